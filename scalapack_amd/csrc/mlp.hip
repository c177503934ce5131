// mlp.hip -- the precision selector's network: 12 -> 64 -> 64 -> 7, tanh
// hidden layers, softmax over the seven choices (FP8, BF16x3 or BF16x6
// factors, each with classical IR or GMRES-IR; plain FP64 LU), FP32 weights.  Forward, backward
// (softmax cross-entropy) and Adam run as HIP kernels; every reduction has a
// fixed order (no atomics), so training is reproducible bit for bit.
//
// Parameter vector (floats): W1 [64][12], b1 [64], W2 [64][64], b2 [64],
// W3 [7][64], b3 [7], then the input normalisation mu [12], sigma [12]
// (x_in = (f - mu) / sigma).  Samples are rows of X [B][12].
#include "sla_common.h"
#include "sla_handle.h"
#include "sla_mlp.h"

#include <cmath>
#include <stdexcept>
#include <vector>

namespace sla {

namespace {

constexpr int H = MLP_HIDDEN, IN = MLP_IN, OUT = MLP_OUT;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + H * IN, OFF_W2 = OFF_B1 + H, OFF_B2 = OFF_W2 + H * H,
              OFF_W3 = OFF_B2 + H, OFF_B3 = OFF_W3 + OUT * H, OFF_MU = OFF_B3 + OUT,
              OFF_SD = OFF_MU + IN;
static_assert(OFF_SD + IN == MLP_NPARAM, "parameter layout");
constexpr int NTRAIN = OFF_MU;  // trained parameters (the normalisation is fixed)

// Y[s][o] = act(b[o] + sum_k W[o][k] X[s][k]); one thread per (s, o).
// norm: X is raw features, normalised with mu / sd on the fly.
template <int NI, int NO, bool TANH, bool NORM>
__global__ void dense_fwd_kernel(int B, const float *__restrict__ X, const float *__restrict__ P,
                                 int offw, int offb, float *__restrict__ Y)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= B * NO) return;
    const int s = idx / NO, o = idx % NO;
    const float *w = P + offw + o * NI;
    float acc = P[offb + o];
    for (int k = 0; k < NI; k++) {
        float x = X[s * NI + k];
        if (NORM) x = (x - P[OFF_MU + k]) / P[OFF_SD + k];
        acc = fmaf(w[k], x, acc);
    }
    Y[idx] = TANH ? tanhf(acc) : acc;
}

// Softmax cross-entropy with per-class costs: dZ[s] = cost[y_s] (p - e_y) / B,
// loss[s] = cost[y_s] * -log p[y_s]; probabilities into Pr.
__global__ void softmax_xent_kernel(int B, const float *__restrict__ Z, const int *__restrict__ y,
                                    const float *__restrict__ cost, float *__restrict__ Pr,
                                    float *__restrict__ dZ, float *__restrict__ loss)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= B) return;
    float z[OUT], m = -INFINITY;
    for (int c = 0; c < OUT; c++) {
        z[c] = Z[s * OUT + c];
        m = fmaxf(m, z[c]);
    }
    float sum = 0.f;
    for (int c = 0; c < OUT; c++) {
        z[c] = expf(z[c] - m);
        sum += z[c];
    }
    const int t = y ? y[s] : 0;
    const float w = cost ? cost[t] : 1.f;
    for (int c = 0; c < OUT; c++) {
        const float p = z[c] / sum;
        Pr[s * OUT + c] = p;
        if (dZ) dZ[s * OUT + c] = w * (p - (c == t ? 1.f : 0.f)) / (float)B;
    }
    if (loss) loss[s] = -w * logf(fmaxf(z[t] / sum, 1e-30f));
}

// dX[s][k] = (sum_o dY[s][o] W[o][k]) * (1 - X[s][k]^2) (the tanh of the
// layer below); one thread per (s, k).
template <int NI, int NO>
__global__ void dense_bwd_input_kernel(int B, const float *__restrict__ dY, const float *__restrict__ P,
                                       int offw, const float *__restrict__ Xact, float *__restrict__ dX)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= B * NI) return;
    const int s = idx / NI, k = idx % NI;
    float acc = 0.f;
    for (int o = 0; o < NO; o++) acc = fmaf(dY[s * NO + o], P[offw + o * NI + k], acc);
    const float a = Xact[idx];
    dX[idx] = acc * (1.f - a * a);
}

// G[offw + o*NI + k] = sum_s dY[s][o] X[s][k]; G[offb + o] = sum_s dY[s][o];
// one thread per weight (k == NI: the bias), samples summed in order.
template <int NI, int NO, bool NORM>
__global__ void dense_bwd_weights_kernel(int B, const float *__restrict__ dY, const float *__restrict__ X,
                                         const float *__restrict__ P, int offw, int offb,
                                         float *__restrict__ G)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= NO * (NI + 1)) return;
    const int o = idx / (NI + 1), k = idx % (NI + 1);
    float acc = 0.f;
    if (k == NI) {
        for (int s = 0; s < B; s++) acc += dY[s * NO + o];
        G[offb + o] = acc;
    } else {
        const float mu = NORM ? P[OFF_MU + k] : 0.f, sd = NORM ? P[OFF_SD + k] : 1.f;
        for (int s = 0; s < B; s++) {
            float x = X[s * NI + k];
            if (NORM) x = (x - mu) / sd;
            acc = fmaf(dY[s * NO + o], x, acc);
        }
        G[offw + o * NI + k] = acc;
    }
}

// Fixed-order sum of the per-sample losses (one block).
__global__ void sum_kernel(int B, const float *__restrict__ v, float *out)
{
    __shared__ float part[256];
    float a = 0.f;
    for (int i = threadIdx.x; i < B; i += 256) a += v[i];
    part[threadIdx.x] = a;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) part[threadIdx.x] += part[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = part[0];
}

// Adam on the trained parameters (bias-corrected, step t >= 1).
__global__ void adam_kernel(int n, float *__restrict__ P, const float *__restrict__ G,
                            float *__restrict__ M, float *__restrict__ V, float lr, float b1,
                            float b2, float eps, float c1, float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = G[i];
    const float m = b1 * M[i] + (1.f - b1) * g;
    const float v = b2 * V[i] + (1.f - b2) * g * g;
    M[i] = m;
    V[i] = v;
    P[i] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

__global__ void argmax_kernel(int B, const float *__restrict__ Pr, int *__restrict__ pred)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= B) return;
    int best = 0;
    for (int c = 1; c < OUT; c++)
        if (Pr[s * OUT + c] > Pr[s * OUT + best]) best = c;
    pred[s] = best;
}

unsigned grid(int n) { return (unsigned)std::max(1, (n + 255) / 256); }

}  // namespace

// Workspace: activations and gradients of a batch of B samples.
struct MlpWork {
    float *h1, *h2, *z, *pr, *dz, *d2, *d1, *loss, *lsum, *g;
    explicit MlpWork(Handle &h, int B)
    {
        const size_t n = (size_t)B * (2 * H + 3 * OUT + 2 * H + 1) + 1 + MLP_NPARAM + 64;
        float *p = static_cast<float *>(h.buffer(BUF_MLP, n * sizeof(float)));
        h1 = p;
        h2 = h1 + (size_t)B * H;
        z = h2 + (size_t)B * H;
        pr = z + (size_t)B * OUT;
        dz = pr + (size_t)B * OUT;
        d2 = dz + (size_t)B * OUT;
        d1 = d2 + (size_t)B * H;
        loss = d1 + (size_t)B * H;
        lsum = loss + B;
        g = lsum + 1;
    }
};

namespace {

void forward(hipStream_t s, int B, const float *X, const float *P, MlpWork &w)
{
    hipLaunchKernelGGL((dense_fwd_kernel<IN, H, true, true>), dim3(grid(B * H)), dim3(256), 0, s, B, X, P,
                       OFF_W1, OFF_B1, w.h1);
    hipLaunchKernelGGL((dense_fwd_kernel<H, H, true, false>), dim3(grid(B * H)), dim3(256), 0, s, B, w.h1,
                       P, OFF_W2, OFF_B2, w.h2);
    hipLaunchKernelGGL((dense_fwd_kernel<H, OUT, false, false>), dim3(grid(B * OUT)), dim3(256), 0, s, B,
                       w.h2, P, OFF_W3, OFF_B3, w.z);
    SLA_LAUNCH_CHECK();
}

// Gradient of the mean (cost-weighted) cross-entropy into w.g; loss into w.lsum.
void backward(hipStream_t s, int B, const float *X, const int *y, const float *cost, const float *P,
              MlpWork &w)
{
    hipLaunchKernelGGL(softmax_xent_kernel, dim3(grid(B)), dim3(256), 0, s, B, w.z, y, cost, w.pr, w.dz,
                       w.loss);
    hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(256), 0, s, B, w.loss, w.lsum);
    hipLaunchKernelGGL((dense_bwd_weights_kernel<H, OUT, false>), dim3(grid(OUT * (H + 1))), dim3(256), 0,
                       s, B, w.dz, w.h2, P, OFF_W3, OFF_B3, w.g);
    hipLaunchKernelGGL((dense_bwd_input_kernel<H, OUT>), dim3(grid(B * H)), dim3(256), 0, s, B, w.dz, P,
                       OFF_W3, w.h2, w.d2);
    hipLaunchKernelGGL((dense_bwd_weights_kernel<H, H, false>), dim3(grid(H * (H + 1))), dim3(256), 0, s,
                       B, w.d2, w.h1, P, OFF_W2, OFF_B2, w.g);
    hipLaunchKernelGGL((dense_bwd_input_kernel<H, H>), dim3(grid(B * H)), dim3(256), 0, s, B, w.d2, P,
                       OFF_W2, w.h1, w.d1);
    hipLaunchKernelGGL((dense_bwd_weights_kernel<IN, H, true>), dim3(grid(H * (IN + 1))), dim3(256), 0, s,
                       B, w.d1, X, P, OFF_W1, OFF_B1, w.g);
    SLA_LAUNCH_CHECK();
}

}  // namespace

float mlp_grad_device(Handle &h, int B, const float *dX, const int *dy, const float *dcost,
                      const float *dP, float *dG)
{
    if (B < 1) throw std::invalid_argument("mlp: at least one sample");
    MlpWork w(h, B);
    forward(h.stream, B, dX, dP, w);
    backward(h.stream, B, dX, dy, dcost, dP, w);
    SLA_HIP(hipMemcpyAsync(dG, w.g, sizeof(float) * NTRAIN, hipMemcpyDeviceToDevice, h.stream));
    float loss = 0.f;
    SLA_HIP(hipMemcpyAsync(&loss, w.lsum, sizeof(float), hipMemcpyDeviceToHost, h.stream));
    SLA_HIP(hipStreamSynchronize(h.stream));
    return loss / (float)B;
}

void mlp_train_device(Handle &h, int B, const float *dX, const int *dy, const float *dcost,
                      float *dP, int epochs, float lr, float *loss_history)
{
    if (B < 1 || epochs < 0) throw std::invalid_argument("mlp: at least one sample");
    MlpWork w(h, B);
    float *MV = nullptr;
    SLA_HIP(hipMalloc(&MV, sizeof(float) * 2 * NTRAIN));
    SLA_HIP(hipMemsetAsync(MV, 0, sizeof(float) * 2 * NTRAIN, h.stream));
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    try {
        for (int e = 0; e < epochs; e++) {
            forward(h.stream, B, dX, dP, w);
            backward(h.stream, B, dX, dy, dcost, dP, w);
            const float c1 = 1.f - std::pow(b1, (float)(e + 1)), c2 = 1.f - std::pow(b2, (float)(e + 1));
            hipLaunchKernelGGL(adam_kernel, dim3(grid(NTRAIN)), dim3(256), 0, h.stream, NTRAIN, dP, w.g, MV,
                               MV + NTRAIN, lr, b1, b2, eps, c1, c2);
            SLA_LAUNCH_CHECK();
            if (loss_history) {
                float l = 0.f;
                SLA_HIP(hipMemcpyAsync(&l, w.lsum, sizeof(float), hipMemcpyDeviceToHost, h.stream));
                SLA_HIP(hipStreamSynchronize(h.stream));
                loss_history[e] = l / (float)B;
            }
        }
        SLA_HIP(hipStreamSynchronize(h.stream));
    } catch (...) {
        (void)hipFree(MV);
        throw;
    }
    SLA_HIP(hipFree(MV));
}

void mlp_predict_device(Handle &h, int B, const float *dX, const float *dP, int *dpred, float *dprob)
{
    if (B < 1) return;
    MlpWork w(h, B);
    forward(h.stream, B, dX, dP, w);
    hipLaunchKernelGGL(softmax_xent_kernel, dim3(grid(B)), dim3(256), 0, h.stream, B, w.z, nullptr,
                       nullptr, w.pr, nullptr, nullptr);
    if (dpred) hipLaunchKernelGGL(argmax_kernel, dim3(grid(B)), dim3(256), 0, h.stream, B, w.pr, dpred);
    if (dprob)
        SLA_HIP(hipMemcpyAsync(dprob, w.pr, sizeof(float) * B * OUT, hipMemcpyDeviceToDevice, h.stream));
    SLA_LAUNCH_CHECK();
    SLA_HIP(hipStreamSynchronize(h.stream));
}

}  // namespace sla
