// capi.hip -- the C-ABI of libscalapack_amd.so (include/scalapack_amd.h).
//
// The Fortran drop-in entry points validate their arguments with the
// reference's rules and INFO codes (TOOLS/chk1mat.f:1-100; SRC/pdgetrf.f:
// 173-199; SRC/pdgetrs.f:188-237; SRC/pdgesv.f:178-225; SRC/pdgetf2.f:170-193),
// print the PXERBLA message (PBLAS/SRC/PTZBLAS/pxerbla.f:55-58) for an illegal
// argument, and run the factorization / solve on the GPU.  On a grid other
// than 1 x 1 the call is forwarded to the next definition of the same symbol
// in the process (a real ScaLAPACK linked after this library); without one it
// reports the illegal context.
#include "sla_common.h"
#include "sla_handle.h"
#include "sla_kernels.h"
#include "sla_mixed.h"
#include "sla_aux.h"
#include "sla_batched.h"
#include "sla_mlp.h"

#include "scalapack_amd.h"
#include "sla_blacs1x1.h"

#include <dlfcn.h>

#include <algorithm>
#include <cctype>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

namespace {

thread_local std::string g_last_error;

// Fortran descriptor entries (1-based, TOOLS/descinit.f).
enum { DTYPE_ = 1, CTXT_ = 2, M_ = 3, N_ = 4, MB_ = 5, NB_ = 6, RSRC_ = 7, CSRC_ = 8, LLD_ = 9 };
inline int D(const int *desc, int e) { return desc[e - 1]; }

struct Grid {
    int nprow, npcol, myrow, mycol;
};
Grid gridinfo(int ctxt)
{
    Grid g;
    blacs_gridinfo_(&ctxt, &g.nprow, &g.npcol, &g.myrow, &g.mycol);
    return g;
}

int numroc(int n, int nb, int iproc, int isrc, int nprocs)
{
    return numroc_(&n, &nb, &iproc, &isrc, &nprocs);
}

// CHK1MAT (TOOLS/chk1mat.f), with its DESCMULT encoding of descriptor entries.
void chk1mat(int ma, int mapos0, int na, int napos0, int ia, int ja, const int *desca,
             int descapos0, int &info)
{
    // The grid is that of this descriptor's own context (chk1mat.f).
    const Grid g = gridinfo(D(desca, CTXT_));
    const int DESCMULT = 100, BIGNUM = DESCMULT * DESCMULT;
    if (info >= 0) info = BIGNUM;
    else if (info < -DESCMULT) info = -info;
    else info = -info * DESCMULT;
    const int mapos = mapos0 * DESCMULT, napos = napos0 * DESCMULT;
    const int iapos = (descapos0 - 2) * DESCMULT, japos = (descapos0 - 1) * DESCMULT;
    const int descapos = descapos0 * DESCMULT;
    if (D(desca, DTYPE_) != 1) info = std::min(info, descapos + DTYPE_);
    else if (ma < 0) info = std::min(info, mapos);
    else if (na < 0) info = std::min(info, napos);
    else if (ia < 1) info = std::min(info, iapos);
    else if (ja < 1) info = std::min(info, japos);
    else if (D(desca, MB_) < 1) info = std::min(info, descapos + MB_);
    else if (D(desca, NB_) < 1) info = std::min(info, descapos + NB_);
    else if (D(desca, RSRC_) < 0 || D(desca, RSRC_) >= g.nprow) info = std::min(info, descapos + RSRC_);
    else if (D(desca, CSRC_) < 0 || D(desca, CSRC_) >= g.npcol) info = std::min(info, descapos + CSRC_);
    else if (D(desca, LLD_) < 1) info = std::min(info, descapos + LLD_);
    else if (D(desca, LLD_) < numroc(D(desca, M_), D(desca, MB_), g.myrow, D(desca, RSRC_), g.nprow)) {
        if (numroc(D(desca, N_), D(desca, NB_), g.mycol, D(desca, CSRC_), g.npcol) > 0)
            info = std::min(info, descapos + LLD_);
    }
    if (ma == 0 || na == 0) {
        if (D(desca, M_) < 0) info = std::min(info, descapos + M_);
        if (D(desca, N_) < 0) info = std::min(info, descapos + N_);
    } else {
        if (D(desca, M_) < 1) info = std::min(info, descapos + M_);
        else if (D(desca, N_) < 1) info = std::min(info, descapos + N_);
        else {
            if (ia > D(desca, M_)) info = std::min(info, iapos);
            else if (ja > D(desca, N_)) info = std::min(info, japos);
            else {
                if (ia + ma - 1 > D(desca, M_)) info = std::min(info, mapos);
                if (ja + na - 1 > D(desca, N_)) info = std::min(info, napos);
            }
        }
    }
    if (info == BIGNUM) info = 0;
    else if (info % DESCMULT == 0) info = -info / DESCMULT;
    else info = -info;
}

void pxerbla(const Grid &g, const char *name, int pos)
{
    std::printf("{%5d,%5d}:  On entry to %s parameter number %4d had an illegal value\n",
                g.myrow, g.mycol, name, pos);
    std::fflush(stdout);
}

int indxg2p(int g, int nb, int isrc, int nprocs) { return (isrc + (g - 1) / nb) % nprocs; }

// --- process-wide default handle per device ---------------------------------
std::mutex g_mu;
std::map<int, std::unique_ptr<sla::Handle>> g_handles;

sla::Handle &default_handle()
{
    int dev = 0;
    SLA_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_mu);
    auto &p = g_handles[dev];
    if (!p) p.reset(new sla::Handle());
    return *p;
}

// The handle caches the row permutation of its last factorization keyed on
// the device addresses; the Fortran entry points free their device copies on
// return, so they drop that association when they finish.
struct PermKeyReset {
    sla::Handle &h;
    ~PermKeyReset()
    {
        h.perm_key_a = nullptr;
        h.perm_key_ipiv = nullptr;
        h.perm_m = 0;
    }
};

inline bool lsame(char a, char b) { return std::toupper((unsigned char)a) == b; }

// Device copies of host arrays for the Fortran entry points.
struct DevBuf {
    void *p = nullptr;
    explicit DevBuf(size_t bytes) { SLA_HIP(hipMalloc(&p, bytes ? bytes : 8)); }
    ~DevBuf() { if (p) (void)hipFree(p); }
    template <typename T> T *as() { return static_cast<T *>(p); }
};

void to_device(double *dst, int ldd, const double *src, int lds, int m, int n)
{
    SLA_HIP(hipMemcpy2D(dst, (size_t)ldd * 8, src, (size_t)lds * 8, (size_t)m * 8, n,
                        hipMemcpyHostToDevice));
}
void to_host(double *dst, int ldd, const double *src, int lds, int m, int n)
{
    SLA_HIP(hipMemcpy2D(dst, (size_t)ldd * 8, src, (size_t)lds * 8, (size_t)m * 8, n,
                        hipMemcpyDeviceToHost));
}
int pad_ld(int m) { return std::max(64, (m + 63) / 64 * 64); }

template <typename Fn>
Fn next_symbol(const char *name)
{
    return reinterpret_cast<Fn>(dlsym(RTLD_NEXT, name));
}

[[noreturn]] void fatal(const char *routine, const std::exception &e)
{
    std::fprintf(stderr, "%s: GPU execution failed: %s\n", routine, e.what());
    std::fflush(stderr);
    std::abort();
}

// Factor sub(A) (host, local array a with LLD) on the device.  Leaves the
// device copy and pivots in dA/dipiv for a following solve.
int getrf_host(sla::Handle &h, int m, int n, double *a, int ia, int ja, int lld, int *ipiv,
               DevBuf &dA, int ldd, DevBuf &dipiv)
{
    double *asub = a + (size_t)(ia - 1) + (size_t)(ja - 1) * lld;
    to_device(dA.as<double>(), ldd, asub, lld, m, n);
    const int info = sla::dgetrf_device(h, m, n, dA.as<double>(), ldd, dipiv.as<int>());
    to_host(asub, lld, dA.as<double>(), ldd, m, n);
    const int mn = std::min(m, n);
    std::vector<int> piv(mn);
    SLA_HIP(hipMemcpy(piv.data(), dipiv.p, sizeof(int) * mn, hipMemcpyDeviceToHost));
    // IPIV holds global rows, indexed by the local row (SRC/pdgetrf.f:118-121).
    for (int i = 0; i < mn; i++) ipiv[(ia - 1) + i] = piv[i] + (ia - 1);
    return info;
}

}  // namespace

// ============================================================================
// Part 1: Fortran drop-in
// ============================================================================
extern "C" {

void pdgetrf_(const int *m, const int *n, double *a, const int *ia, const int *ja,
              const int *desca, int *ipiv, int *info)
{
    const int ictxt = D(desca, CTXT_);
    const Grid g = gridinfo(ictxt);
    *info = 0;
    if (g.nprow == -1) {
        *info = -(600 + CTXT_);
    } else {
        chk1mat(*m, 1, *n, 2, *ia, *ja, desca, 6, *info);
        if (*info == 0) {
            const int iroff = (*ia - 1) % D(desca, MB_), icoff = (*ja - 1) % D(desca, NB_);
            if (iroff != 0) *info = -4;
            else if (icoff != 0) *info = -5;
            else if (D(desca, MB_) != D(desca, NB_)) *info = -(600 + NB_);
        }
    }
    if (*info != 0) {
        pxerbla(g, "PDGETRF", -*info);
        return;
    }
    if (D(desca, M_) == 1) {  // pdgetrf.f:201-203
        ipiv[0] = 1;
        return;
    }
    if (*m == 0 || *n == 0) return;
    if (g.nprow * g.npcol != 1) {
        auto next = next_symbol<void (*)(const int *, const int *, double *, const int *,
                                         const int *, const int *, int *, int *)>("pdgetrf_");
        if (next) return next(m, n, a, ia, ja, desca, ipiv, info);
        *info = -(600 + CTXT_);
        pxerbla(g, "PDGETRF", -*info);
        return;
    }
    try {
        sla::Handle &h = default_handle();
        PermKeyReset reset{h};
        const int ldd = pad_ld(*m);
        DevBuf dA((size_t)ldd * *n * 8), dp((size_t)std::min(*m, *n) * 4);
        *info = getrf_host(h, *m, *n, a, *ia, *ja, D(desca, LLD_), ipiv, dA, ldd, dp);
    } catch (const std::exception &e) {
        fatal("PDGETRF", e);
    }
}

void pdgetf2_(const int *m, const int *n, double *a, const int *ia, const int *ja,
              const int *desca, int *ipiv, int *info)
{
    const int ictxt = D(desca, CTXT_);
    const Grid g = gridinfo(ictxt);
    *info = 0;
    if (g.nprow == -1) {
        *info = -(600 + CTXT_);
    } else {
        chk1mat(*m, 1, *n, 2, *ia, *ja, desca, 6, *info);
        if (*info == 0) {
            const int iroff = (*ia - 1) % D(desca, MB_), icoff = (*ja - 1) % D(desca, NB_);
            if (*n + icoff > D(desca, NB_)) *info = -2;
            else if (iroff != 0) *info = -4;
            else if (icoff != 0) *info = -5;
            else if (D(desca, MB_) != D(desca, NB_)) *info = -(600 + NB_);
        }
    }
    if (*info != 0) {
        pxerbla(g, "PDGETF2", -*info);
        int one = 1;
        blacs_abort_(&ictxt, &one);  // pdgetf2.f:190-193
        return;
    }
    if (*m == 0 || *n == 0) return;
    if (g.nprow * g.npcol != 1) {
        auto next = next_symbol<void (*)(const int *, const int *, double *, const int *,
                                         const int *, const int *, int *, int *)>("pdgetf2_");
        if (next) return next(m, n, a, ia, ja, desca, ipiv, info);
        *info = -(600 + CTXT_);
        pxerbla(g, "PDGETF2", -*info);
        return;
    }
    try {
        sla::Handle &h = default_handle();
        PermKeyReset reset{h};
        const int ldd = pad_ld(*m);
        DevBuf dA((size_t)ldd * *n * 8), dp((size_t)std::min(*m, *n) * 4);
        *info = getrf_host(h, *m, *n, a, *ia, *ja, D(desca, LLD_), ipiv, dA, ldd, dp);
    } catch (const std::exception &e) {
        fatal("PDGETF2", e);
    }
}

static int getrs_checks(char trans, int n, int nrhs, int ia, int ja, const int *desca, int ib,
                        int jb, const int *descb, const Grid &g)
{
    int info = 0;
    if (g.nprow == -1) return -(700 + CTXT_);
    const bool notran = (trans == 'N' || trans == 'n');
    chk1mat(n, 2, n, 2, ia, ja, desca, 7, info);
    chk1mat(n, 2, nrhs, 3, ib, jb, descb, 12, info);
    if (info == 0) {
        const int iarow = indxg2p(ia, D(desca, MB_), D(desca, RSRC_), g.nprow);
        const int ibrow = indxg2p(ib, D(descb, MB_), D(descb, RSRC_), g.nprow);
        const int iroffa = (ia - 1) % D(desca, MB_), icoffa = (ja - 1) % D(desca, NB_);
        const int iroffb = (ib - 1) % D(descb, MB_);
        if (!notran && !(trans == 'T' || trans == 't') && !(trans == 'C' || trans == 'c')) info = -1;
        else if (iroffa != 0) info = -5;
        else if (icoffa != 0) info = -6;
        else if (D(desca, MB_) != D(desca, NB_)) info = -(700 + NB_);
        else if (iroffb != 0 || ibrow != iarow) info = -10;
        else if (D(descb, MB_) != D(desca, NB_)) info = -(1200 + NB_);
        else if (D(desca, CTXT_) != D(descb, CTXT_)) info = -(1200 + CTXT_);
    }
    return info;
}

void pdgetrs_(const char *trans, const int *n, const int *nrhs, const double *a, const int *ia,
              const int *ja, const int *desca, const int *ipiv, double *b, const int *ib,
              const int *jb, const int *descb, int *info, size_t)
{
    const Grid g = gridinfo(D(desca, CTXT_));
    *info = getrs_checks(*trans, *n, *nrhs, *ia, *ja, desca, *ib, *jb, descb, g);
    if (*info != 0) {
        pxerbla(g, "PDGETRS", -*info);
        return;
    }
    if (*n == 0 || *nrhs == 0) return;
    if (g.nprow * g.npcol != 1) {
        auto next = next_symbol<void (*)(const char *, const int *, const int *, const double *,
                                         const int *, const int *, const int *, const int *,
                                         double *, const int *, const int *, const int *, int *,
                                         size_t)>("pdgetrs_");
        if (next) return next(trans, n, nrhs, a, ia, ja, desca, ipiv, b, ib, jb, descb, info, 1);
        *info = -(700 + CTXT_);
        pxerbla(g, "PDGETRS", -*info);
        return;
    }
    const bool notran = lsame(*trans, 'N');
    try {
        sla::Handle &h = default_handle();
        PermKeyReset reset{h};
        const int N = *n, NR = *nrhs, ldd = pad_ld(N), lldb = D(descb, LLD_);
        DevBuf dA((size_t)ldd * N * 8), dB((size_t)ldd * NR * 8), dp((size_t)N * 4);
        to_device(dA.as<double>(), ldd, a + (size_t)(*ia - 1) + (size_t)(*ja - 1) * D(desca, LLD_),
                  D(desca, LLD_), N, N);
        double *bsub = b + (size_t)(*ib - 1) + (size_t)(*jb - 1) * lldb;
        to_device(dB.as<double>(), ldd, bsub, lldb, N, NR);
        std::vector<int> piv(N);
        for (int i = 0; i < N; i++) piv[i] = ipiv[(*ia - 1) + i] - (*ia - 1);  // pdlapv2.f:229
        SLA_HIP(hipMemcpy(dp.p, piv.data(), sizeof(int) * N, hipMemcpyHostToDevice));
        sla::dgetrs_device(h, N, NR, dA.as<double>(), ldd, dp.as<int>(), dB.as<double>(), ldd,
                           !notran);
        to_host(bsub, lldb, dB.as<double>(), ldd, N, NR);
    } catch (const std::exception &e) {
        fatal("PDGETRS", e);
    }
}

void pdgesv_(const int *n, const int *nrhs, double *a, const int *ia, const int *ja,
             const int *desca, int *ipiv, double *b, const int *ib, const int *jb,
             const int *descb, int *info)
{
    const Grid g = gridinfo(D(desca, CTXT_));
    *info = 0;
    if (g.nprow == -1) {
        *info = -(600 + CTXT_);
    } else {
        chk1mat(*n, 1, *n, 1, *ia, *ja, desca, 6, *info);
        chk1mat(*n, 1, *nrhs, 2, *ib, *jb, descb, 11, *info);
        if (*info == 0) {
            const int iarow = indxg2p(*ia, D(desca, MB_), D(desca, RSRC_), g.nprow);
            const int ibrow = indxg2p(*ib, D(descb, MB_), D(descb, RSRC_), g.nprow);
            const int iroffa = (*ia - 1) % D(desca, MB_), icoffa = (*ja - 1) % D(desca, NB_);
            const int iroffb = (*ib - 1) % D(descb, MB_);
            if (iroffa != 0) *info = -4;
            else if (icoffa != 0) *info = -5;
            else if (D(desca, MB_) != D(desca, NB_)) *info = -(600 + NB_);
            else if (ibrow != iarow || icoffa != iroffb) *info = -9;
            else if (D(descb, MB_) != D(desca, NB_)) *info = -(1100 + NB_);
            else if (D(desca, CTXT_) != D(descb, CTXT_)) *info = -(1100 + CTXT_);
        }
    }
    if (*info != 0) {
        pxerbla(g, "PDGESV", -*info);
        return;
    }
    if (g.nprow * g.npcol != 1) {
        auto next = next_symbol<void (*)(const int *, const int *, double *, const int *,
                                         const int *, const int *, int *, double *, const int *,
                                         const int *, const int *, int *)>("pdgesv_");
        if (next) return next(n, nrhs, a, ia, ja, desca, ipiv, b, ib, jb, descb, info);
        *info = -(600 + CTXT_);
        pxerbla(g, "PDGESV", -*info);
        return;
    }
    if (*n == 0) return;
    if (D(desca, M_) == 1) {  // PDGETRF's quick return (pdgetrf.f:201-203)
        ipiv[0] = 1;
    }
    try {
        sla::Handle &h = default_handle();
        PermKeyReset reset{h};
        const int N = *n, NR = *nrhs, ldd = pad_ld(N), lldb = D(descb, LLD_);
        DevBuf dA((size_t)ldd * N * 8), dp((size_t)N * 4), dB((size_t)ldd * std::max(NR, 1) * 8);
        if (D(desca, M_) != 1)
            *info = getrf_host(h, N, N, a, *ia, *ja, D(desca, LLD_), ipiv, dA, ldd, dp);
        else
            to_device(dA.as<double>(), ldd, a + (size_t)(*ia - 1) + (size_t)(*ja - 1) * D(desca, LLD_),
                      D(desca, LLD_), N, N);
        if (*info == 0 && NR > 0) {  // pdgesv.f:231
            if (D(desca, M_) == 1) {
                int one = 1;
                SLA_HIP(hipMemcpy(dp.p, &one, sizeof one, hipMemcpyHostToDevice));
            }
            double *bsub = b + (size_t)(*ib - 1) + (size_t)(*jb - 1) * lldb;
            to_device(dB.as<double>(), ldd, bsub, lldb, N, NR);
            sla::dgetrs_device(h, N, NR, dA.as<double>(), ldd, dp.as<int>(), dB.as<double>(), ldd);
            to_host(bsub, lldb, dB.as<double>(), ldd, N, NR);
        }
    } catch (const std::exception &e) {
        fatal("PDGESV", e);
    }
}

void pdlaswp_(const char *direc, const char *rowcol, const int *n, double *a, const int *ia,
              const int *ja, const int *desca, const int *k1, const int *k2, const int *ipiv,
              size_t, size_t)
{
    if (*n == 0) return;
    const Grid g = gridinfo(D(desca, CTXT_));
    if (g.nprow * g.npcol != 1) {
        auto next = next_symbol<void (*)(const char *, const char *, const int *, double *,
                                         const int *, const int *, const int *, const int *,
                                         const int *, const int *, size_t, size_t)>("pdlaswp_");
        if (next) return next(direc, rowcol, n, a, ia, ja, desca, k1, k2, ipiv, 1, 1);
        pxerbla(g, "PDLASWP", 7);
        return;
    }
    const bool rows = (*rowcol == 'R' || *rowcol == 'r');
    const bool fwd = (*direc == 'F' || *direc == 'f');
    const int lld = D(desca, LLD_);
    try {
        sla::Handle &h = default_handle();
        if (rows) {
            // Rows K1..K2 interchange with rows IPIV(K1..K2) of columns JA..JA+N-1
            // (pdlaswp.f:163-185); the touched rows are 1..max(K2, max IPIV).
            int hi = *k2;
            for (int i = *k1; i <= *k2; i++) {
                if (ipiv[i - 1] < 1) {
                    pxerbla(g, "PDLASWP", 10);
                    return;
                }
                hi = std::max(hi, ipiv[i - 1]);
            }
            DevBuf dA((size_t)hi * *n * 8), dp((size_t)hi * 4);
            double *acol = a + (size_t)(*ja - 1) * lld;
            to_device(dA.as<double>(), hi, acol, lld, hi, *n);
            SLA_HIP(hipMemcpy(dp.p, ipiv, sizeof(int) * *k2, hipMemcpyHostToDevice));
            sla::laswp_rows_device(h, dA.as<double>(), hi, *n, *k1 - 1, *k2, dp.as<int>(), fwd);
            to_host(acol, lld, dA.as<double>(), hi, hi, *n);
        } else {
            // Columns K1..K2 interchange with columns IPIV(K1..K2) over rows
            // IA..IA+N-1 (pdlaswp.f:186-203).
            int hi = *k2;
            for (int i = *k1; i <= *k2; i++) {
                if (ipiv[i - 1] < 1) {
                    pxerbla(g, "PDLASWP", 10);
                    return;
                }
                hi = std::max(hi, ipiv[i - 1]);
            }
            DevBuf dA((size_t)*n * hi * 8), dp((size_t)hi * 4);
            double *arow = a + (size_t)(*ia - 1);
            to_device(dA.as<double>(), *n, arow, lld, *n, hi);
            SLA_HIP(hipMemcpy(dp.p, ipiv, sizeof(int) * *k2, hipMemcpyHostToDevice));
            sla::laswp_cols_device(h, dA.as<double>(), *n, *n, *k1 - 1, *k2, dp.as<int>(), fwd);
            to_host(arow, lld, dA.as<double>(), *n, *n, hi);
        }
    } catch (const std::exception &e) {
        fatal("PDLASWP", e);
    }
}

// PDLANGE (SRC/pdlange.f:1-2): a function with no argument checks; WORK is
// not used on the device path.
double pdlange_(const char *norm, const int *m, const int *n, const double *a, const int *ia,
                const int *ja, const int *desca, double *work, size_t)
{
    (void)work;
    const Grid g = gridinfo(D(desca, CTXT_));
    if (std::min(*m, *n) <= 0) return 0.0;
    if (g.nprow * g.npcol != 1) {
        auto next = next_symbol<double (*)(const char *, const int *, const int *, const double *,
                                           const int *, const int *, const int *, double *,
                                           size_t)>("pdlange_");
        if (next) return next(norm, m, n, a, ia, ja, desca, work, 1);
        pxerbla(g, "PDLANGE", 7);
        return 0.0;
    }
    const char c = (char)std::toupper((unsigned char)*norm);
    if (!(c == 'M' || c == '1' || c == 'O' || c == 'I' || c == 'F' || c == 'E')) return 0.0;
    try {
        sla::Handle &h = default_handle();
        const int M = *m, N = *n, lld = D(desca, LLD_), ldd = pad_ld(M);
        DevBuf dA((size_t)ldd * N * 8);
        to_device(dA.as<double>(), ldd, a + (size_t)(*ia - 1) + (size_t)(*ja - 1) * lld, lld, M, N);
        return sla::dlange_device(h, c, M, N, dA.as<double>(), ldd);
    } catch (const std::exception &e) {
        fatal("PDLANGE", e);
    }
}

// PDGECON (SRC/pdgecon.f:1-3); argument checks of pdgecon.f:215-285.
void pdgecon_(const char *norm, const int *n, const double *a, const int *ia, const int *ja,
              const int *desca, const double *anorm, double *rcond, double *work, const int *lwork,
              int *iwork, const int *liwork, int *info, size_t)
{
    const Grid g = gridinfo(D(desca, CTXT_));
    *info = 0;
    int lwmin = 0, liwmin = 0;
    bool lquery = false;
    const bool onenrm = (*norm == '1' || lsame(*norm, 'O'));
    if (g.nprow == -1) {
        *info = -(600 + CTXT_);
    } else {
        chk1mat(*n, 2, *n, 2, *ia, *ja, desca, 6, *info);
        if (*info == 0) {
            const int mb = D(desca, MB_), nb = D(desca, NB_);
            const int iarow = indxg2p(*ia, mb, D(desca, RSRC_), g.nprow);
            const int iacol = indxg2p(*ja, nb, D(desca, CSRC_), g.npcol);
            const int npmod = numroc(*n + (*ia - 1) % mb, mb, g.myrow, iarow, g.nprow);
            const int nqmod = numroc(*n + (*ja - 1) % nb, nb, g.mycol, iacol, g.npcol);
            auto iceil = [](int x, int y) { return (x + y - 1) / y; };
            lwmin = 2 * npmod + 2 * nqmod +
                    std::max(2, std::max(nb * std::max(1, iceil(g.nprow - 1, g.npcol)),
                                         nqmod + nb * std::max(1, iceil(g.npcol - 1, g.nprow))));
            work[0] = (double)lwmin;
            liwmin = std::max(1, npmod);
            iwork[0] = liwmin;
            lquery = (*lwork == -1 || *liwork == -1);
            if (!onenrm && !lsame(*norm, 'I')) *info = -1;
            else if (*anorm < 0.0) *info = -7;
            else if (*lwork < lwmin && !lquery) *info = -10;
            else if (*liwork < liwmin && !lquery) *info = -12;
        }
    }
    if (*info != 0) {
        pxerbla(g, "PDGECON", -*info);
        return;
    }
    if (lquery) return;
    if (g.nprow * g.npcol != 1) {
        auto next = next_symbol<void (*)(const char *, const int *, const double *, const int *,
                                         const int *, const int *, const double *, double *,
                                         double *, const int *, int *, const int *, int *,
                                         size_t)>("pdgecon_");
        if (next) return next(norm, n, a, ia, ja, desca, anorm, rcond, work, lwork, iwork, liwork,
                              info, 1);
        *info = -(600 + CTXT_);
        pxerbla(g, "PDGECON", -*info);
        return;
    }
    // quick returns, pdgecon.f:287-297
    *rcond = 0.0;
    if (*n == 0) {
        *rcond = 1.0;
        return;
    }
    if (*anorm == 0.0) return;
    if (*n == 1) {
        *rcond = 1.0;
        return;
    }
    try {
        sla::Handle &h = default_handle();
        const int N = *n, lld = D(desca, LLD_), ldd = pad_ld(N);
        DevBuf dA((size_t)ldd * N * 8);
        to_device(dA.as<double>(), ldd, a + (size_t)(*ia - 1) + (size_t)(*ja - 1) * lld, lld, N, N);
        *rcond = sla::dgecon_device(h, onenrm ? '1' : 'I', N, dA.as<double>(), ldd, *anorm);
    } catch (const std::exception &e) {
        fatal("PDGECON", e);
    }
}

// PDGERFS (SRC/pdgerfs.f:1-4); argument checks of pdgerfs.f:308-404.
void pdgerfs_(const char *trans, const int *n, const int *nrhs, const double *a, const int *ia,
              const int *ja, const int *desca, const double *af, const int *iaf, const int *jaf,
              const int *descaf, const int *ipiv, const double *b, const int *ib, const int *jb,
              const int *descb, double *x, const int *ix, const int *jx, const int *descx,
              double *ferr, double *berr, double *work, const int *lwork, int *iwork,
              const int *liwork, int *info, size_t)
{
    const Grid g = gridinfo(D(desca, CTXT_));
    const bool notran = lsame(*trans, 'N');
    *info = 0;
    bool lquery = false;
    if (g.nprow == -1) {
        *info = -(700 + CTXT_);
    } else {
        chk1mat(*n, 2, *n, 2, *ia, *ja, desca, 7, *info);
        chk1mat(*n, 2, *n, 2, *iaf, *jaf, descaf, 11, *info);
        chk1mat(*n, 2, *nrhs, 3, *ib, *jb, descb, 16, *info);
        chk1mat(*n, 2, *nrhs, 3, *ix, *jx, descx, 20, *info);
        if (*info == 0) {
            const int iroffa = (*ia - 1) % D(desca, MB_), icoffa = (*ja - 1) % D(desca, NB_);
            const int iroffaf = (*iaf - 1) % D(descaf, MB_), icoffaf = (*jaf - 1) % D(descaf, NB_);
            const int iroffb = (*ib - 1) % D(descb, MB_), icoffb = (*jb - 1) % D(descb, NB_);
            const int iroffx = (*ix - 1) % D(descx, MB_), icoffx = (*jx - 1) % D(descx, NB_);
            const int iarow = indxg2p(*ia, D(desca, MB_), D(desca, RSRC_), g.nprow);
            const int iafcol = indxg2p(*jaf, D(descaf, NB_), D(descaf, CSRC_), g.npcol);
            const int iafrow = indxg2p(*iaf, D(descaf, MB_), D(descaf, RSRC_), g.nprow);
            const int iacol = indxg2p(*ja, D(desca, NB_), D(desca, CSRC_), g.npcol);
            const int ixbrow = indxg2p(*ib, D(descb, MB_), D(descb, RSRC_), g.nprow);
            const int ixbcol = indxg2p(*jb, D(descb, NB_), D(descb, CSRC_), g.npcol);
            const int ixrow = indxg2p(*ix, D(descx, MB_), D(descx, RSRC_), g.nprow);
            const int ixcol = indxg2p(*jx, D(descx, NB_), D(descx, CSRC_), g.npcol);
            const int npmod = numroc(*n + iroffa, D(desca, MB_), g.myrow, iarow, g.nprow);
            const int lwmin = 3 * npmod, liwmin = npmod;
            work[0] = (double)lwmin;
            iwork[0] = liwmin;
            lquery = (*lwork == -1 || *liwork == -1);
            const int ictxt = D(desca, CTXT_);
            if (!notran && !lsame(*trans, 'T') && !lsame(*trans, 'C')) *info = -1;
            else if (*n < 0) *info = -2;
            else if (*nrhs < 0) *info = -3;
            else if (iroffa != 0) *info = -5;
            else if (icoffa != 0) *info = -6;
            else if (D(desca, MB_) != D(desca, NB_)) *info = -(700 + NB_);
            else if (D(desca, MB_) != D(descaf, MB_)) *info = -(1100 + MB_);
            else if (iroffaf != 0 || iarow != iafrow) *info = -9;
            else if (D(desca, NB_) != D(descaf, NB_)) *info = -(1100 + NB_);
            else if (icoffaf != 0 || iacol != iafcol) *info = -10;
            else if (ictxt != D(descaf, CTXT_)) *info = -(1100 + CTXT_);
            else if (iroffa != iroffb || iarow != ixbrow) *info = -14;
            else if (D(desca, MB_) != D(descb, MB_)) *info = -(1600 + MB_);
            else if (ictxt != D(descb, CTXT_)) *info = -(1600 + CTXT_);
            else if (D(descb, MB_) != D(descx, MB_)) *info = -(2000 + MB_);
            else if (iroffx != 0 || ixbrow != ixrow) *info = -18;
            else if (D(descb, NB_) != D(descx, NB_)) *info = -(2000 + NB_);
            else if (icoffb != icoffx || ixbcol != ixcol) *info = -19;
            else if (ictxt != D(descx, CTXT_)) *info = -(2000 + CTXT_);
            else if (*lwork < lwmin && !lquery) *info = -24;
            else if (*liwork < liwmin && !lquery) *info = -26;
        }
    }
    if (*info != 0) {
        pxerbla(g, "PDGERFS", -*info);
        return;
    }
    if (lquery) return;
    if (g.nprow * g.npcol != 1) {
        auto next = next_symbol<void (*)(const char *, const int *, const int *, const double *,
                                         const int *, const int *, const int *, const double *,
                                         const int *, const int *, const int *, const int *,
                                         const double *, const int *, const int *, const int *,
                                         double *, const int *, const int *, const int *, double *,
                                         double *, double *, const int *, int *, const int *,
                                         int *, size_t)>("pdgerfs_");
        if (next) return next(trans, n, nrhs, a, ia, ja, desca, af, iaf, jaf, descaf, ipiv, b, ib,
                              jb, descb, x, ix, jx, descx, ferr, berr, work, lwork, iwork, liwork,
                              info, 1);
        *info = -(700 + CTXT_);
        pxerbla(g, "PDGERFS", -*info);
        return;
    }
    const int N = *n, NR = *nrhs;
    if (N <= 1 || NR == 0) {  // pdgerfs.f:440-446
        for (int k = 0; k < NR; k++) ferr[(*jb - 1) + k] = berr[(*jb - 1) + k] = 0.0;
        return;
    }
    try {
        sla::Handle &h = default_handle();
        PermKeyReset reset{h};
        const int ldd = pad_ld(N);
        const int llda = D(desca, LLD_), lldaf = D(descaf, LLD_), lldb = D(descb, LLD_),
                  lldx = D(descx, LLD_);
        DevBuf dA((size_t)ldd * N * 8), dAF((size_t)ldd * N * 8), dB((size_t)ldd * NR * 8),
            dX((size_t)ldd * NR * 8), dp((size_t)N * 4);
        to_device(dA.as<double>(), ldd, a + (size_t)(*ia - 1) + (size_t)(*ja - 1) * llda, llda, N, N);
        to_device(dAF.as<double>(), ldd, af + (size_t)(*iaf - 1) + (size_t)(*jaf - 1) * lldaf, lldaf,
                  N, N);
        to_device(dB.as<double>(), ldd, b + (size_t)(*ib - 1) + (size_t)(*jb - 1) * lldb, lldb, N, NR);
        double *xsub = x + (size_t)(*ix - 1) + (size_t)(*jx - 1) * lldx;
        to_device(dX.as<double>(), ldd, xsub, lldx, N, NR);
        std::vector<int> piv(N);
        for (int i = 0; i < N; i++) piv[i] = ipiv[(*iaf - 1) + i] - (*iaf - 1);  // as pdlapv2.f:229
        SLA_HIP(hipMemcpy(dp.p, piv.data(), sizeof(int) * N, hipMemcpyHostToDevice));
        // FERR / BERR are indexed by the local column of B (pdgerfs.f:437, JJFBE = JJXB)
        std::vector<double> fe(NR), be(NR);
        for (int k = 0; k < NR; k++) fe[k] = ferr[(*jb - 1) + k];
        sla::dgerfs_device(h, !notran, N, NR, dA.as<double>(), ldd, dAF.as<double>(), ldd,
                           dp.as<int>(), dB.as<double>(), ldd, dX.as<double>(), ldd, fe.data(),
                           be.data());
        to_host(xsub, lldx, dX.as<double>(), ldd, N, NR);
        for (int k = 0; k < NR; k++) {
            ferr[(*jb - 1) + k] = fe[k];
            berr[(*jb - 1) + k] = be[k];
        }
    } catch (const std::exception &e) {
        fatal("PDGERFS", e);
    }
}

// PDGEEQU (SRC/pdgeequ.f:1-2); checks as CHK1MAT (pdgeequ.f:190-198).
void pdgeequ_(const int *m, const int *n, const double *a, const int *ia, const int *ja,
              const int *desca, double *r, double *c, double *rowcnd, double *colcnd, double *amax,
              int *info)
{
    const Grid g = gridinfo(D(desca, CTXT_));
    *info = 0;
    if (g.nprow == -1) *info = -(600 + CTXT_);
    else chk1mat(*m, 1, *n, 2, *ia, *ja, desca, 6, *info);
    if (*info != 0) {
        pxerbla(g, "PDGEEQU", -*info);
        return;
    }
    if (*m == 0 || *n == 0) {
        *rowcnd = 1.0;
        *colcnd = 1.0;
        *amax = 0.0;
        return;
    }
    if (g.nprow * g.npcol != 1) {
        auto next = next_symbol<void (*)(const int *, const int *, const double *, const int *,
                                         const int *, const int *, double *, double *, double *,
                                         double *, double *, int *)>("pdgeequ_");
        if (next) return next(m, n, a, ia, ja, desca, r, c, rowcnd, colcnd, amax, info);
        *info = -(600 + CTXT_);
        pxerbla(g, "PDGEEQU", -*info);
        return;
    }
    try {
        sla::Handle &h = default_handle();
        const int M = *m, N = *n, lld = D(desca, LLD_), ldd = pad_ld(M);
        DevBuf dA((size_t)ldd * N * 8), dR((size_t)M * 8), dC((size_t)N * 8);
        to_device(dA.as<double>(), ldd, a + (size_t)(*ia - 1) + (size_t)(*ja - 1) * lld, lld, M, N);
        *info = sla::dgeequ_device(h, M, N, dA.as<double>(), ldd, dR.as<double>(), dC.as<double>(),
                                   rowcnd, colcnd, amax);
        // R(IA..), C(JA..) as far as the reference computes them before a
        // zero row (R holds the row maxima) or column returns
        SLA_HIP(hipMemcpy(r + (*ia - 1), dR.p, (size_t)M * 8, hipMemcpyDeviceToHost));
        if (*info == 0 || *info > M)
            SLA_HIP(hipMemcpy(c + (*ja - 1), dC.p, (size_t)N * 8, hipMemcpyDeviceToHost));
    } catch (const std::exception &e) {
        fatal("PDGEEQU", e);
    }
}

// PDLAQGE (SRC/pdlaqge.f:1-2): no argument checks in the reference.
void pdlaqge_(const int *m, const int *n, double *a, const int *ia, const int *ja, const int *desca,
              const double *r, const double *c, const double *rowcnd, const double *colcnd,
              const double *amax, char *equed, size_t)
{
    if (*m <= 0 || *n <= 0) {
        *equed = 'N';
        return;
    }
    const Grid g = gridinfo(D(desca, CTXT_));
    if (g.nprow * g.npcol != 1) {
        auto next = next_symbol<void (*)(const int *, const int *, double *, const int *, const int *,
                                         const int *, const double *, const double *, const double *,
                                         const double *, const double *, char *, size_t)>("pdlaqge_");
        if (next) return next(m, n, a, ia, ja, desca, r, c, rowcnd, colcnd, amax, equed, 1);
        pxerbla(g, "PDLAQGE", 6);
        return;
    }
    try {
        sla::Handle &h = default_handle();
        const int M = *m, N = *n, lld = D(desca, LLD_), ldd = pad_ld(M);
        DevBuf dA((size_t)ldd * N * 8), dR((size_t)M * 8), dC((size_t)N * 8);
        double *asub = a + (size_t)(*ia - 1) + (size_t)(*ja - 1) * lld;
        to_device(dA.as<double>(), ldd, asub, lld, M, N);
        SLA_HIP(hipMemcpy(dR.p, r + (*ia - 1), (size_t)M * 8, hipMemcpyHostToDevice));
        SLA_HIP(hipMemcpy(dC.p, c + (*ja - 1), (size_t)N * 8, hipMemcpyHostToDevice));
        *equed = sla::dlaqge_device(h, M, N, dA.as<double>(), ldd, dR.as<double>(), dC.as<double>(),
                                    *rowcnd, *colcnd, *amax);
        if (*equed != 'N') to_host(asub, lld, dA.as<double>(), ldd, M, N);
    } catch (const std::exception &e) {
        fatal("PDLAQGE", e);
    }
}

// PDGESVX (SRC/pdgesvx.f:1-5); argument checks of pdgesvx.f:452-620.
void pdgesvx_(const char *fact, const char *trans, const int *n, const int *nrhs, double *a,
              const int *ia, const int *ja, const int *desca, double *af, const int *iaf,
              const int *jaf, const int *descaf, int *ipiv, char *equed, double *r, double *c,
              double *b, const int *ib, const int *jb, const int *descb, double *x, const int *ix,
              const int *jx, const int *descx, double *rcond, double *ferr, double *berr,
              double *work, const int *lwork, int *iwork, const int *liwork, int *info, size_t,
              size_t, size_t)
{
    const Grid g = gridinfo(D(desca, CTXT_));
    const bool nofact = lsame(*fact, 'N'), equil = lsame(*fact, 'E'), notran = lsame(*trans, 'N');
    const bool factored = lsame(*fact, 'F');
    bool rowequ = false, colequ = false, lquery = false;
    const int N = *n;
    *info = 0;
    int lwmin = 0, liwmin = 0;
    if (g.nprow == -1) {
        *info = -(800 + CTXT_);
    } else {
        chk1mat(N, 3, N, 3, *ia, *ja, desca, 8, *info);
        if (factored) chk1mat(N, 3, N, 3, *iaf, *jaf, descaf, 12, *info);
        chk1mat(N, 3, *nrhs, 4, *ib, *jb, descb, 20, *info);
        chk1mat(N, 3, *nrhs, 4, *ix, *jx, descx, 24, *info);
        if (!(nofact || equil)) {
            rowequ = lsame(*equed, 'R') || lsame(*equed, 'B');
            colequ = lsame(*equed, 'C') || lsame(*equed, 'B');
        }
        if (*info == 0) {
            const int nb = D(desca, NB_);
            const int iarow = indxg2p(*ia, D(desca, MB_), D(desca, RSRC_), g.nprow);
            const int iafrow = indxg2p(*iaf, D(descaf, MB_), D(descaf, RSRC_), g.nprow);
            const int ibrow = indxg2p(*ib, D(descb, MB_), D(descb, RSRC_), g.nprow);
            const int ixrow = indxg2p(*ix, D(descx, MB_), D(descx, RSRC_), g.nprow);
            const int iroffa = (*ia - 1) % D(desca, MB_), iroffaf = (*iaf - 1) % D(descaf, MB_);
            const int icoffa = (*ja - 1) % nb;
            // 1 x 1: NP = NQ = N; LCM = 1 (pdgesvx.f:498-517)
            const int conwrk = 2 * N + 2 * N + std::max(2, std::max(nb, N + nb));
            int rfswrk = 3 * N + (notran ? N + N + ((N + iroffa + nb - 1) / nb) * nb : N + N);
            lwmin = std::max(conwrk, rfswrk);
            liwmin = N;
            if (!nofact && !equil && !factored) *info = -1;
            else if (!notran && !lsame(*trans, 'T') && !lsame(*trans, 'C')) *info = -2;
            else if (iroffa != 0) *info = -6;
            else if (icoffa != 0 || iroffa != icoffa) *info = -7;
            else if (D(desca, MB_) != D(desca, NB_)) *info = -(800 + NB_);
            else if (iafrow != iarow) *info = -10;
            else if (iroffaf != 0) *info = -10;
            else if (D(desca, CTXT_) != D(descaf, CTXT_)) *info = -(1200 + CTXT_);
            else if (factored && !(rowequ || colequ || lsame(*equed, 'N'))) *info = -13;
            else {
                if (rowequ) {
                    double mn = 1.0 / DBL_MIN;
                    for (int i = 0; i < N; i++) mn = std::min(mn, r[(*ia - 1) + i]);
                    if (mn <= 0.0) *info = -14;
                }
                if (colequ && *info == 0) {
                    double mn = 1.0 / DBL_MIN;
                    for (int j = 0; j < N; j++) mn = std::min(mn, c[(*ja - 1) + j]);
                    if (mn <= 0.0) *info = -15;
                }
            }
            work[0] = (double)lwmin;
            iwork[0] = liwmin;
            lquery = (*lwork == -1 || *liwork == -1);
            if (*info == 0) {
                if (ibrow != iarow) *info = -18;
                else if (ixrow != ibrow) *info = -22;
                else if (D(descb, MB_) != D(desca, NB_)) *info = -(2000 + NB_);
                else if (D(descb, CTXT_) != D(desca, CTXT_)) *info = -(2000 + CTXT_);
                else if (D(descx, MB_) != D(desca, NB_)) *info = -(2400 + NB_);
                else if (D(descx, CTXT_) != D(desca, CTXT_)) *info = -(2400 + CTXT_);
                else if (*lwork < lwmin && !lquery) *info = -29;
                else if (*liwork < liwmin && !lquery) *info = -31;
            }
        }
    }
    if (*info != 0) {
        pxerbla(g, "PDGESVX", -*info);
        return;
    }
    if (lquery) return;
    if (g.nprow * g.npcol != 1) {
        auto next = next_symbol<void (*)(const char *, const char *, const int *, const int *, double *,
                                         const int *, const int *, const int *, double *, const int *,
                                         const int *, const int *, int *, char *, double *, double *,
                                         double *, const int *, const int *, const int *, double *,
                                         const int *, const int *, const int *, double *, double *,
                                         double *, double *, const int *, int *, const int *, int *,
                                         size_t, size_t, size_t)>("pdgesvx_");
        if (next) return next(fact, trans, n, nrhs, a, ia, ja, desca, af, iaf, jaf, descaf, ipiv,
                              equed, r, c, b, ib, jb, descb, x, ix, jx, descx, rcond, ferr, berr,
                              work, lwork, iwork, liwork, info, 1, 1, 1);
        *info = -(800 + CTXT_);
        pxerbla(g, "PDGESVX", -*info);
        return;
    }
    if (N == 0) return;
    try {
        sla::Handle &h = default_handle();
        PermKeyReset reset{h};
        const int NR = *nrhs, ldd = pad_ld(N);
        const int llda = D(desca, LLD_), lldaf = D(descaf, LLD_), lldb = D(descb, LLD_),
                  lldx = D(descx, LLD_);
        DevBuf dA((size_t)ldd * N * 8), dAF((size_t)ldd * N * 8), dB((size_t)ldd * std::max(NR, 1) * 8),
            dX((size_t)ldd * std::max(NR, 1) * 8), dp((size_t)N * 4), dR((size_t)N * 8),
            dC((size_t)N * 8);
        double *asub = a + (size_t)(*ia - 1) + (size_t)(*ja - 1) * llda;
        double *afsub = af + (size_t)(*iaf - 1) + (size_t)(*jaf - 1) * lldaf;
        double *bsub = b + (size_t)(*ib - 1) + (size_t)(*jb - 1) * lldb;
        double *xsub = x + (size_t)(*ix - 1) + (size_t)(*jx - 1) * lldx;
        to_device(dA.as<double>(), ldd, asub, llda, N, N);
        if (NR > 0) to_device(dB.as<double>(), ldd, bsub, lldb, N, NR);
        if (factored) {
            to_device(dAF.as<double>(), ldd, afsub, lldaf, N, N);
            std::vector<int> piv(N);
            for (int i = 0; i < N; i++) piv[i] = ipiv[(*iaf - 1) + i] - (*iaf - 1);
            SLA_HIP(hipMemcpy(dp.p, piv.data(), sizeof(int) * N, hipMemcpyHostToDevice));
            SLA_HIP(hipMemcpy(dR.p, r + (*ia - 1), (size_t)N * 8, hipMemcpyHostToDevice));
            SLA_HIP(hipMemcpy(dC.p, c + (*ja - 1), (size_t)N * 8, hipMemcpyHostToDevice));
        }
        std::vector<double> fe(std::max(NR, 1)), be(std::max(NR, 1));
        for (int k = 0; k < NR; k++) fe[k] = ferr[(*jb - 1) + k];
        char eq = *equed;
        const int dinfo = sla::dgesvx_device(h, *fact, !notran, N, NR, dA.as<double>(), ldd,
                                             dAF.as<double>(), ldd, dp.as<int>(), &eq,
                                             dR.as<double>(), dC.as<double>(), dB.as<double>(), ldd,
                                             dX.as<double>(), ldd, rcond, fe.data(), be.data());
        *equed = eq;
        // outputs as the reference leaves them: the equilibrated A, the
        // factors and pivots, the scales, the scaled B, X and the bounds
        if (eq != 'N') to_host(asub, llda, dA.as<double>(), ldd, N, N);
        if (!factored) {
            to_host(afsub, lldaf, dAF.as<double>(), ldd, N, N);
            std::vector<int> piv(N);
            SLA_HIP(hipMemcpy(piv.data(), dp.p, sizeof(int) * N, hipMemcpyDeviceToHost));
            for (int i = 0; i < N; i++) ipiv[(*iaf - 1) + i] = piv[i] + (*iaf - 1);
        }
        if (lsame(*fact, 'E')) {
            SLA_HIP(hipMemcpy(r + (*ia - 1), dR.p, (size_t)N * 8, hipMemcpyDeviceToHost));
            SLA_HIP(hipMemcpy(c + (*ja - 1), dC.p, (size_t)N * 8, hipMemcpyDeviceToHost));
        }
        if (NR > 0) to_host(bsub, lldb, dB.as<double>(), ldd, N, NR);
        if (dinfo == 0 && NR > 0) {
            to_host(xsub, lldx, dX.as<double>(), ldd, N, NR);
            for (int k = 0; k < NR; k++) {
                ferr[(*jb - 1) + k] = fe[k];
                berr[(*jb - 1) + k] = be[k];
            }
        }
        *info = dinfo == N + 1 ? *ia + N : dinfo;
        work[0] = (double)lwmin;
        iwork[0] = liwmin;
    } catch (const std::exception &e) {
        fatal("PDGESVX", e);
    }
}

// ============================================================================
// Part 2: device-resident interface
// ============================================================================
const char *sla_last_error(void) { return g_last_error.c_str(); }

#define SLA_NEED_HANDLE(h)                                   \
    do {                                                     \
        if (!(h)) {                                          \
            g_last_error = "sla: null handle";               \
            return SLA_ERR_ARG;                              \
        }                                                    \
    } while (0)

#define SLA_TRY(body)                                  \
    try {                                              \
        body;                                          \
        return SLA_OK;                                 \
    } catch (const std::exception &e) {                \
        g_last_error = e.what();                       \
        return SLA_ERR_RUNTIME;                        \
    }

int sla_create(sla_handle_t *h)
{
    SLA_TRY(*h = reinterpret_cast<sla_handle_t>(new sla::Handle()))
}

int sla_destroy(sla_handle_t h)
{
    SLA_NEED_HANDLE(h);
    SLA_TRY(delete reinterpret_cast<sla::Handle *>(h))
}

int sla_set_stream(sla_handle_t h, void *stream)
{
    SLA_NEED_HANDLE(h);
    auto *H = reinterpret_cast<sla::Handle *>(h);
    SLA_TRY(H->stream = static_cast<hipStream_t>(stream))
}

int sla_set_param(sla_handle_t h, const char *name, int value)
{
    auto *H = reinterpret_cast<sla::Handle *>(h);
    const std::string k(name ? name : "");
    static const char *known[] = {"nb",      "leaf_w",  "panel_wgs",        "profile",
                                  "ir_max_iter", "gmres_m", "ir_gmres_on_stall", "gmres_tol_exp",
                                  "pivoting", "lookahead", "update_cus", "lacon_lapack",
                                  "batched_w", "batched_wgs_per_cu", "batched_graph",
                                  "update_lds_kb", "ir_accept_stall", "trsv_blockinv",
                                  "ir_fallback", "trsm_mfma"};
    if (std::find_if(std::begin(known), std::end(known), [&](const char *x) { return k == x; }) ==
        std::end(known)) {
        g_last_error = "sla_set_param: unknown parameter " + k;
        return SLA_ERR_ARG;
    }
    SLA_NEED_HANDLE(h);
    // ranges are what the kernels accept: the factorization's block columns
    // are at most 512 wide, a panel leaf launches at most 128 workgroups
    if (k == "nb" && value >= 1 && value <= 512) H->nb = value;
    else if (k == "leaf_w" && value >= 1 && value <= 32) H->leaf_w = value;
    else if (k == "panel_wgs" && value >= 1 && value <= 128) H->panel_wgs = value;
    else if (k == "profile") H->profile = value != 0;
    else if (k == "ir_max_iter" && value >= 3 && value <= 1000) H->ir_max_iter = value;
    else if (k == "gmres_m" && value >= 1 && value <= 1000) H->gmres_m = value;
    else if (k == "ir_gmres_on_stall") H->ir_gmres_on_stall = value != 0;
    else if (k == "ir_accept_stall") H->ir_accept_stall = value != 0;
    else if (k == "trsv_blockinv" && (value == 0 || value == 1)) H->trsv_blockinv = value;
    else if (k == "trsm_mfma" && (value == 0 || value == 1)) H->trsm_mfma = value;
    else if (k == "ir_fallback" && (value == 0 || value == 1)) H->ir_fallback = value;
    else if (k == "pivoting" && (value == 0 || value == 1)) H->pivoting = value;
    else if (k == "lookahead" && (value == 0 || value == 1)) H->lookahead = value;
    else if (k == "update_cus" && value >= 0 && value <= H->num_cus) H->update_cus = value;
    else if (k == "lacon_lapack" && (value == 0 || value == 1)) H->lacon_lapack = value;
    else if (k == "batched_w" && (value == 8 || value == 16 || value == 32)) H->batched_w = value;
    else if (k == "batched_wgs_per_cu" && value >= 1 && value <= 8) H->batched_wgs_per_cu = value;
    else if (k == "batched_graph" && (value == 0 || value == 1)) H->batched_graph = value;
    else if (k == "update_lds_kb" && value >= 0 && value <= 160) H->update_lds_kb = value;
    else if (k == "gmres_tol_exp" && value >= 1 && value <= 15) H->gmres_tol = std::pow(10.0, -value);
    else {
        g_last_error = "sla_set_param: value out of range for " + k;
        return SLA_ERR_ARG;
    }
    return SLA_OK;
}

int sla_get_param(sla_handle_t h, const char *name, int *value)
{
    SLA_NEED_HANDLE(h);
    auto *H = reinterpret_cast<sla::Handle *>(h);
    const std::string k(name ? name : "");
    if (k == "nb") *value = H->nb;
    else if (k == "leaf_w") *value = H->leaf_w;
    else if (k == "panel_wgs") *value = H->panel_wgs;
    else if (k == "profile") *value = H->profile ? 1 : 0;
    else if (k == "ir_max_iter") *value = H->ir_max_iter;
    else if (k == "gmres_m") *value = H->gmres_m;
    else if (k == "ir_gmres_on_stall") *value = H->ir_gmres_on_stall;
    else if (k == "ir_accept_stall") *value = H->ir_accept_stall;
    else if (k == "trsv_blockinv") *value = H->trsv_blockinv;
    else if (k == "trsm_mfma") *value = H->trsm_mfma;
    else if (k == "ir_fallback") *value = H->ir_fallback;
    else if (k == "pivoting") *value = H->pivoting;
    else if (k == "lookahead") *value = H->lookahead;
    else if (k == "update_cus") *value = H->update_cus;
    else if (k == "lacon_lapack") *value = H->lacon_lapack;
    else if (k == "batched_w") *value = H->batched_w;
    else if (k == "batched_wgs_per_cu") *value = H->batched_wgs_per_cu;
    else if (k == "batched_graph") *value = H->batched_graph;
    else if (k == "update_lds_kb") *value = H->update_lds_kb;
    else if (k == "gmres_tol_exp") *value = (int)std::lround(-std::log10(H->gmres_tol));
    else {
        g_last_error = "sla_get_param: unknown parameter " + k;
        return SLA_ERR_ARG;
    }
    return SLA_OK;
}

int sla_get_stats(sla_handle_t h, int cls, double *ms, double *work, long long *launches)
{
    SLA_NEED_HANDLE(h);
    auto *H = reinterpret_cast<sla::Handle *>(h);
    if (cls < 0 || cls >= sla::CAT_COUNT) {
        g_last_error = "sla_get_stats: class out of range";
        return SLA_ERR_ARG;
    }
    *ms = H->stats[cls].ms;
    *work = H->stats[cls].work;
    *launches = H->stats[cls].launches;
    return SLA_OK;
}

int sla_get_update_exec(sla_handle_t h, double *ms, double *work, long long *launches)
{
    SLA_NEED_HANDLE(h);
    if (!ms || !work || !launches) {
        g_last_error = "sla_get_update_exec: null output";
        return SLA_ERR_ARG;
    }
    auto *H = reinterpret_cast<sla::Handle *>(h);
    *ms = H->exec.ms;
    *work = H->exec.work;
    *launches = H->exec.launches;
    return SLA_OK;
}

int sla_reset_stats(sla_handle_t h)
{
    SLA_NEED_HANDLE(h);
    SLA_TRY(reinterpret_cast<sla::Handle *>(h)->reset_stats())
}

int sla_dgetrf(sla_handle_t h, int m, int n, double *dA, int lda, int *dipiv, int *info)
{
    SLA_NEED_HANDLE(h);
    if (m < 0 || n < 0 || lda < std::max(1, m)) {
        g_last_error = "sla_dgetrf: illegal m, n or lda";
        return SLA_ERR_ARG;
    }
    SLA_TRY(*info = sla::dgetrf_device(*reinterpret_cast<sla::Handle *>(h), m, n, dA, lda, dipiv))
}

int sla_dgetrs(sla_handle_t h, int n, int nrhs, const double *dA, int lda, const int *dipiv,
               double *dB, int ldb)
{
    SLA_NEED_HANDLE(h);
    if (n < 0 || nrhs < 0 || lda < std::max(1, n) || ldb < std::max(1, n)) {
        g_last_error = "sla_dgetrs: illegal n, nrhs, lda or ldb";
        return SLA_ERR_ARG;
    }
    SLA_TRY(sla::dgetrs_device(*reinterpret_cast<sla::Handle *>(h), n, nrhs, dA, lda, dipiv, dB, ldb))
}

int sla_dgetrs_trans(sla_handle_t h, char trans, int n, int nrhs, const double *dA, int lda,
                     const int *dipiv, double *dB, int ldb)
{
    SLA_NEED_HANDLE(h);
    const bool notran = lsame(trans, 'N');
    if (!notran && !lsame(trans, 'T') && !lsame(trans, 'C')) {
        g_last_error = "sla_dgetrs_trans: TRANS must be 'N', 'T' or 'C'";
        return SLA_ERR_ARG;
    }
    if (n < 0 || nrhs < 0 || lda < std::max(1, n) || ldb < std::max(1, n)) {
        g_last_error = "sla_dgetrs_trans: illegal n, nrhs, lda or ldb";
        return SLA_ERR_ARG;
    }
    SLA_TRY(sla::dgetrs_device(*reinterpret_cast<sla::Handle *>(h), n, nrhs, dA, lda, dipiv, dB, ldb,
                               !notran))
}

int sla_gesv_batched(sla_handle_t h, int prec, int batch, const int *n, const int *nrhs,
                     void *const *dA, const int *lda, int *const *dipiv, void *const *dB,
                     const int *ldb, int *dinfo)
{
    SLA_NEED_HANDLE(h);
    if (prec != SLA_PREC_FP64 && prec != SLA_PREC_FP32) {
        g_last_error = "sla_gesv_batched: prec must be SLA_PREC_FP64 or SLA_PREC_FP32";
        return SLA_ERR_ARG;
    }
    if (batch < 0 || (batch > 0 && (!n || !nrhs || !lda || !ldb || !dA || !dipiv || !dB || !dinfo))) {
        g_last_error = "sla_gesv_batched: illegal batch or null array";
        return SLA_ERR_ARG;
    }
    for (int s = 0; s < batch; s++)
        if (n[s] < 0 || n[s] > sla::BATCHED_NMAX || nrhs[s] < 0 || nrhs[s] > sla::BATCHED_NRHS_MAX ||
            lda[s] < std::max(1, n[s]) || ldb[s] < std::max(1, n[s])) {
            g_last_error = "sla_gesv_batched: system " + std::to_string(s) +
                           ": n (<= 1024), nrhs (<= 16), lda or ldb out of range";
            return SLA_ERR_ARG;
        }
    auto &H = *reinterpret_cast<sla::Handle *>(h);
    if (prec == SLA_PREC_FP64)
        SLA_TRY(sla::gesv_batched_device<double>(H, batch, n, nrhs,
                                                 reinterpret_cast<double *const *>(dA), lda, dipiv,
                                                 reinterpret_cast<double *const *>(dB), ldb, dinfo))
    SLA_TRY(sla::gesv_batched_device<float>(H, batch, n, nrhs, reinterpret_cast<float *const *>(dA),
                                            lda, dipiv, reinterpret_cast<float *const *>(dB), ldb,
                                            dinfo))
}

static_assert(SLA_BATCHED_NOALLOC_MAX == sla::BATCHED_NOALLOC_MAX,
              "systems per call of the allocation-free batched entry");

int sla_gesv_batched_noalloc(sla_handle_t h, int prec, int batch, const int *dn,
                             const int *dnrhs, void *const *dA, const int *dlda,
                             int *const *dipiv, void *const *dB, const int *dldb, int *dinfo)
{
    SLA_NEED_HANDLE(h);
    if (prec != SLA_PREC_FP64 && prec != SLA_PREC_FP32) {
        g_last_error = "sla_gesv_batched_noalloc: prec must be SLA_PREC_FP64 or SLA_PREC_FP32";
        return SLA_ERR_ARG;
    }
    if (batch < 0 || batch > SLA_BATCHED_NOALLOC_MAX ||
        (batch > 0 && (!dn || !dnrhs || !dlda || !dldb || !dA || !dipiv || !dB || !dinfo))) {
        g_last_error = "sla_gesv_batched_noalloc: batch outside 0.." +
                       std::to_string(SLA_BATCHED_NOALLOC_MAX) + " or null array";
        return SLA_ERR_ARG;
    }
    auto &H = *reinterpret_cast<sla::Handle *>(h);
    if (prec == SLA_PREC_FP64)
        SLA_TRY(sla::gesv_batched_noalloc_device<double>(
            H, batch, dn, dnrhs, reinterpret_cast<double *const *>(dA), dlda, dipiv,
            reinterpret_cast<double *const *>(dB), dldb, dinfo))
    SLA_TRY(sla::gesv_batched_noalloc_device<float>(
        H, batch, dn, dnrhs, reinterpret_cast<float *const *>(dA), dlda, dipiv,
        reinterpret_cast<float *const *>(dB), dldb, dinfo))
}

int sla_gesv_batched_tune(sla_handle_t h, int prec, int batch, const int *n, const int *nrhs,
                          int *w, int *wgs_per_cu, double *ms)
{
    SLA_NEED_HANDLE(h);
    if ((prec != SLA_PREC_FP64 && prec != SLA_PREC_FP32) || batch < 1 || !n || !nrhs) {
        g_last_error = "sla_gesv_batched_tune: illegal prec, batch or sizes";
        return SLA_ERR_ARG;
    }
    for (int s = 0; s < batch; s++)
        if (n[s] < 0 || n[s] > sla::BATCHED_NMAX || nrhs[s] < 0 || nrhs[s] > sla::BATCHED_NRHS_MAX) {
            g_last_error = "sla_gesv_batched_tune: size out of range";
            return SLA_ERR_ARG;
        }
    auto &H = *reinterpret_cast<sla::Handle *>(h);
    try {
        const sla::BatchedTune t = prec == SLA_PREC_FP64
                                       ? sla::tune_batched_device<double>(H, batch, n, nrhs, 3)
                                       : sla::tune_batched_device<float>(H, batch, n, nrhs, 3);
        if (w) *w = t.w;
        if (wgs_per_cu) *wgs_per_cu = t.wgs_per_cu;
        if (ms) *ms = t.ms;
        return SLA_OK;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return SLA_ERR_RUNTIME;
    }
}

int sla_dgeequ(sla_handle_t h, int m, int n, const double *dA, int lda, double *dR, double *dC,
               double *rowcnd, double *colcnd, double *amax, int *info)
{
    SLA_NEED_HANDLE(h);
    if (m < 0 || n < 0 || lda < std::max(1, m) || !rowcnd || !colcnd || !amax || !info) {
        g_last_error = "sla_dgeequ: illegal m, n, lda or output";
        return SLA_ERR_ARG;
    }
    SLA_TRY(*info = sla::dgeequ_device(*reinterpret_cast<sla::Handle *>(h), m, n, dA, lda, dR, dC,
                                       rowcnd, colcnd, amax))
}

int sla_dlaqge(sla_handle_t h, int m, int n, double *dA, int lda, const double *dR,
               const double *dC, double rowcnd, double colcnd, double amax, char *equed)
{
    SLA_NEED_HANDLE(h);
    if (m < 0 || n < 0 || lda < std::max(1, m) || !equed) {
        g_last_error = "sla_dlaqge: illegal m, n, lda or output";
        return SLA_ERR_ARG;
    }
    SLA_TRY(*equed = sla::dlaqge_device(*reinterpret_cast<sla::Handle *>(h), m, n, dA, lda, dR, dC,
                                        rowcnd, colcnd, amax))
}

int sla_dgesvx(sla_handle_t h, char fact, char trans, int n, int nrhs, double *dA, int lda,
               double *dAF, int ldaf, int *dipiv, char *equed, double *dR, double *dC, double *dB,
               int ldb, double *dX, int ldx, double *rcond, double *ferr, double *berr, int *info)
{
    SLA_NEED_HANDLE(h);
    const bool fok = lsame(fact, 'N') || lsame(fact, 'E') || lsame(fact, 'F');
    const bool tok = lsame(trans, 'N') || lsame(trans, 'T') || lsame(trans, 'C');
    const int ld1 = std::max(1, n);
    if (!fok || !tok || n < 0 || nrhs < 0 || lda < ld1 || ldaf < ld1 || ldb < ld1 || ldx < ld1 ||
        !equed || !rcond || !info || (nrhs > 0 && (!ferr || !berr))) {
        g_last_error = "sla_dgesvx: illegal FACT, TRANS, sizes or output";
        return SLA_ERR_ARG;
    }
    SLA_TRY(*info = sla::dgesvx_device(*reinterpret_cast<sla::Handle *>(h), fact, !lsame(trans, 'N'),
                                       n, nrhs, dA, lda, dAF, ldaf, dipiv, equed, dR, dC, dB, ldb,
                                       dX, ldx, rcond, ferr, berr))
}

int sla_dlange(sla_handle_t h, char norm, int m, int n, const double *dA, int lda, double *value)
{
    SLA_NEED_HANDLE(h);
    const char c = (char)std::toupper((unsigned char)norm);
    if (!(c == 'M' || c == '1' || c == 'O' || c == 'I' || c == 'F' || c == 'E')) {
        g_last_error = "sla_dlange: NORM must be one of M 1 O I F E";
        return SLA_ERR_ARG;
    }
    if (m < 0 || n < 0 || lda < std::max(1, m) || !value) {
        g_last_error = "sla_dlange: illegal m, n, lda or value";
        return SLA_ERR_ARG;
    }
    SLA_TRY(*value = sla::dlange_device(*reinterpret_cast<sla::Handle *>(h), c, m, n, dA, lda))
}

int sla_dgecon(sla_handle_t h, char norm, int n, const double *dLU, int lda, double anorm,
               double *rcond)
{
    SLA_NEED_HANDLE(h);
    const char c = (char)std::toupper((unsigned char)norm);
    if (!(c == '1' || c == 'O' || c == 'I')) {
        g_last_error = "sla_dgecon: NORM must be '1', 'O' or 'I'";
        return SLA_ERR_ARG;
    }
    if (n < 0 || lda < std::max(1, n) || !(anorm >= 0.0) || !rcond) {
        g_last_error = "sla_dgecon: illegal n, lda, anorm or rcond";
        return SLA_ERR_ARG;
    }
    SLA_TRY(*rcond = sla::dgecon_device(*reinterpret_cast<sla::Handle *>(h), c, n, dLU, lda, anorm))
}

int sla_dgerfs(sla_handle_t h, char trans, int n, int nrhs, const double *dA, int lda,
               const double *dAF, int ldaf, const int *dipiv, const double *dB, int ldb, double *dX,
               int ldx, double *ferr, double *berr)
{
    SLA_NEED_HANDLE(h);
    const bool notran = lsame(trans, 'N');
    if (!notran && !lsame(trans, 'T') && !lsame(trans, 'C')) {
        g_last_error = "sla_dgerfs: TRANS must be 'N', 'T' or 'C'";
        return SLA_ERR_ARG;
    }
    const int ld1 = std::max(1, n);
    if (n < 0 || nrhs < 0 || lda < ld1 || ldaf < ld1 || ldb < ld1 || ldx < ld1 ||
        (nrhs > 0 && (!ferr || !berr))) {
        g_last_error = "sla_dgerfs: illegal n, nrhs, leading dimension or output pointer";
        return SLA_ERR_ARG;
    }
    SLA_TRY(sla::dgerfs_device(*reinterpret_cast<sla::Handle *>(h), !notran, n, nrhs, dA, lda, dAF,
                               ldaf, dipiv, dB, ldb, dX, ldx, ferr, berr))
}

int sla_dgemm_sub(sla_handle_t h, int m, int n, int k, const double *dA, int lda,
                  const double *dB, int ldb, double *dC, int ldc)
{
    SLA_NEED_HANDLE(h);
    if (m < 0 || n < 0 || k < 0 || lda < std::max(1, m) || ldb < std::max(1, k) ||
        ldc < std::max(1, m)) {
        g_last_error = "sla_dgemm_sub: illegal dimensions";
        return SLA_ERR_ARG;
    }
    SLA_TRY({
        auto &H = *reinterpret_cast<sla::Handle *>(h);
        H.timing_begin(sla::CAT_GEMM);
        sla::dgemm_nn_sub(m, n, k, dA, lda, dB, ldb, dC, ldc, H.stream);
        H.timing_end(2.0 * m * n * k);
        SLA_HIP(hipStreamSynchronize(H.stream));
        H.timing_collect();
    })
}

int sla_dgesv(sla_handle_t h, int n, int nrhs, double *dA, int lda, int *dipiv, double *dB,
              int ldb, int *info)
{
    SLA_NEED_HANDLE(h);
    if (n < 0 || nrhs < 0 || lda < std::max(1, n) || ldb < std::max(1, n)) {
        g_last_error = "sla_dgesv: illegal n, nrhs, lda or ldb";
        return SLA_ERR_ARG;
    }
    SLA_TRY({
        auto &H = *reinterpret_cast<sla::Handle *>(h);
        *info = sla::dgetrf_device(H, n, n, dA, lda, dipiv);
        if (*info == 0) sla::dgetrs_device(H, n, nrhs, dA, lda, dipiv, dB, ldb);
    })
}

int sla_dxgesv(sla_handle_t h, int prec, int refine, int n, int nrhs, double *dA, int lda,
               int *dipiv, const double *dB, int ldb, double *dX, int ldx, int *iter, int *info)
{
    SLA_NEED_HANDLE(h);
    if (prec < SLA_PREC_FP64 || prec > SLA_PREC_FP16X3 || refine < SLA_REFINE_IR ||
        refine > SLA_REFINE_GMRES) {
        g_last_error = "sla_dxgesv: unknown precision or refinement";
        return SLA_ERR_ARG;
    }
    if (n < 0 || nrhs < 0 || nrhs > 16 || lda < std::max(1, n) || ldb < std::max(1, n) ||
        ldx < std::max(1, n)) {
        g_last_error = "sla_dxgesv: illegal n, nrhs (<= 16), lda, ldb or ldx";
        return SLA_ERR_ARG;
    }
    SLA_TRY({
        auto &H = *reinterpret_cast<sla::Handle *>(h);
        sla::MixedInfo mi;
        *info = sla::dxgesv_device(H, prec, refine, n, nrhs, dA, lda, dipiv, dB, ldb, dX, ldx, mi);
        *iter = mi.iter;
        H.last_mixed = mi;
    })
}

// ---------------------------------------------------------------- selector
static_assert(sla::MLP_NPARAM == SLA_MLP_NPARAM, "parameter count of the selector network");

// The selector's class -> factorization precision and refinement of
// sla_dxgesv: {FP8, BF16x3, BF16x6} x {classical IR, GMRES-IR}, then FP64.
static_assert(sla::MLP_OUT == SLA_SELECT_CLASSES, "classes of the selector network");
static int selector_precision(int choice)
{
    static const int prec[SLA_SELECT_CLASSES] = {SLA_PREC_FP8,    SLA_PREC_FP8,    SLA_PREC_BF16X3,
                                                 SLA_PREC_BF16X3, SLA_PREC_BF16X6, SLA_PREC_BF16X6,
                                                 SLA_PREC_FP64};
    return (choice >= 0 && choice < SLA_SELECT_CLASSES) ? prec[choice] : SLA_PREC_FP64;
}
static int selector_refinement(int choice)
{
    return (choice >= 0 && choice < SLA_SELECT_CLASSES - 1 && choice % 2 == 1) ? SLA_REFINE_GMRES
                                                                               : SLA_REFINE_IR;
}

int sla_selector_features(sla_handle_t h, int n, const double *dA, int lda, float *feat)
{
    SLA_NEED_HANDLE(h);
    if (n < 1 || lda < n || !dA || !feat) {
        g_last_error = "sla_selector_features: illegal n, lda or pointer";
        return SLA_ERR_ARG;
    }
    SLA_TRY(sla::selector_features_device(*reinterpret_cast<sla::Handle *>(h), n, dA, lda, feat))
}

int sla_mlp_grad(sla_handle_t h, int batch, const float *dX, const int *dy, const float *dcost,
                 const float *dparams, float *dgrad, float *loss)
{
    SLA_NEED_HANDLE(h);
    if (batch < 1 || !dX || !dy || !dparams || !dgrad || !loss) {
        g_last_error = "sla_mlp_grad: illegal batch or pointer";
        return SLA_ERR_ARG;
    }
    SLA_TRY(*loss = sla::mlp_grad_device(*reinterpret_cast<sla::Handle *>(h), batch, dX, dy, dcost,
                                         dparams, dgrad))
}

int sla_mlp_train(sla_handle_t h, int batch, const float *dX, const int *dy, const float *dcost,
                  float *dparams, int epochs, float lr, float *loss_history)
{
    SLA_NEED_HANDLE(h);
    if (batch < 1 || epochs < 0 || !(lr > 0.f) || !dX || !dy || !dparams) {
        g_last_error = "sla_mlp_train: illegal batch, epochs, lr or pointer";
        return SLA_ERR_ARG;
    }
    SLA_TRY(sla::mlp_train_device(*reinterpret_cast<sla::Handle *>(h), batch, dX, dy, dcost, dparams,
                                  epochs, lr, loss_history))
}

int sla_mlp_predict(sla_handle_t h, int batch, const float *dX, const float *dparams, int *dpred,
                    float *dprob)
{
    SLA_NEED_HANDLE(h);
    if (batch < 0 || (batch > 0 && (!dX || !dparams))) {
        g_last_error = "sla_mlp_predict: illegal batch or pointer";
        return SLA_ERR_ARG;
    }
    SLA_TRY(sla::mlp_predict_device(*reinterpret_cast<sla::Handle *>(h), batch, dX, dparams, dpred,
                                    dprob))
}

int sla_selector_set_model(sla_handle_t h, const float *params)
{
    SLA_NEED_HANDLE(h);
    if (!params) {
        g_last_error = "sla_selector_set_model: null parameters";
        return SLA_ERR_ARG;
    }
    SLA_TRY({
        auto &H = *reinterpret_cast<sla::Handle *>(h);
        for (int i = 0; i < sla::MLP_IN; i++)
            if (!(params[sla::MLP_NPARAM - sla::MLP_IN + i] > 0.f))
                throw std::invalid_argument("sla_selector_set_model: normalisation sigma must be > 0");
        float *d = static_cast<float *>(H.buffer(sla::BUF_MODEL, 4 * (sla::MLP_NPARAM + 64)));
        SLA_HIP(hipMemcpyAsync(d, params, 4 * sla::MLP_NPARAM, hipMemcpyHostToDevice, H.stream));
        SLA_HIP(hipStreamSynchronize(H.stream));
        H.has_model = true;
    })
}

int sla_dxgesv_auto(sla_handle_t h, int n, int nrhs, double *dA, int lda, int *dipiv,
                    const double *dB, int ldb, double *dX, int ldx, int *choice, int *iter,
                    int *info)
{
    SLA_NEED_HANDLE(h);
    if (n < 1 || nrhs < 0 || nrhs > 16 || lda < n || ldb < n || ldx < n || !choice || !iter ||
        !info) {
        g_last_error = "sla_dxgesv_auto: illegal n, nrhs (<= 16), lda, ldb, ldx or output";
        return SLA_ERR_ARG;
    }
    auto &H = *reinterpret_cast<sla::Handle *>(h);
    if (!H.has_model) {
        g_last_error = "sla_dxgesv_auto: no selector model (sla_selector_set_model)";
        return SLA_ERR_ARG;
    }
    SLA_TRY({
        float feat[sla::MLP_IN];
        sla::selector_features_device(H, n, dA, lda, feat);
        float *d = static_cast<float *>(H.buffer(sla::BUF_MODEL, 4 * (sla::MLP_NPARAM + 64)));
        float *df = d + sla::MLP_NPARAM;
        int *dc = reinterpret_cast<int *>(df + 16);
        SLA_HIP(hipMemcpyAsync(df, feat, sizeof feat, hipMemcpyHostToDevice, H.stream));
        sla::mlp_predict_device(H, 1, df, d, dc, nullptr);
        SLA_HIP(hipMemcpy(choice, dc, sizeof(int), hipMemcpyDeviceToHost));
        sla::MixedInfo mi;
        *info = sla::dxgesv_device(H, selector_precision(*choice), selector_refinement(*choice), n,
                                   nrhs, dA, lda, dipiv, dB, ldb, dX, ldx, mi);
        *iter = mi.iter;
        H.last_mixed = mi;
    })
}

int sla_last_solve_info(sla_handle_t h, sla_solve_info_t *out)
{
    SLA_NEED_HANDLE(h);
    if (!out) {
        g_last_error = "sla_last_solve_info: null output";
        return SLA_ERR_ARG;
    }
    const sla::MixedInfo &mi = reinterpret_cast<sla::Handle *>(h)->last_mixed;
    out->precision = mi.precision;
    out->refinement = mi.refinement;
    out->iter = mi.iter;
    out->outer = mi.outer;
    out->inner = mi.inner;
    out->fell_back = mi.fell_back;
    out->anorm = mi.anorm;
    out->bwd_err = mi.bwd_err;
    out->sresid = mi.sresid;
    for (int i = 0; i < 16; i++) out->history[i] = mi.history[i];
    out->stop_rule = mi.stop_rule;
    return SLA_OK;
}

static_assert(SLA_PREC_FP16X3 + 1 == sla::PREC_COUNT, "precision codes of the C-ABI");

static int schur_of(int prec) { return sla::schur_of_prec(prec); }

int sla_sgetrf(sla_handle_t h, int prec, int m, int n, float *dA, int lda, int *dipiv, int *info)
{
    SLA_NEED_HANDLE(h);
    if (prec < SLA_PREC_FP32 || prec > SLA_PREC_FP16X3) {
        g_last_error = "sla_sgetrf: precision must be a low-precision code (FP32 .. FP16X3)";
        return SLA_ERR_ARG;
    }
    if (m < 0 || n < 0 || lda < std::max(1, m)) {
        g_last_error = "sla_sgetrf: illegal m, n or lda";
        return SLA_ERR_ARG;
    }
    SLA_TRY(*info = sla::sgetrf_device(*reinterpret_cast<sla::Handle *>(h), m, n, dA, lda, dipiv,
                                       schur_of(prec)))
}

int sla_sgetrs(sla_handle_t h, int n, int nrhs, const float *dA, int lda, const int *dipiv,
               float *dB, int ldb)
{
    SLA_NEED_HANDLE(h);
    if (n < 0 || nrhs < 0 || lda < std::max(1, n) || ldb < std::max(1, n)) {
        g_last_error = "sla_sgetrs: illegal n, nrhs, lda or ldb";
        return SLA_ERR_ARG;
    }
    SLA_TRY(sla::sgetrs_device(*reinterpret_cast<sla::Handle *>(h), n, nrhs, dA, lda, dipiv, dB,
                               ldb))
}

int sla_dresidual(sla_handle_t h, int n, int nrhs, const double *dA, int lda, const double *dX,
                  int ldx, const double *dB, int ldb, double *dR, int ldr, double *rnorm)
{
    SLA_NEED_HANDLE(h);
    if (n < 0 || nrhs < 0 || lda < std::max(1, n) || ldx < std::max(1, n) ||
        ldb < std::max(1, n) || ldr < std::max(1, n) || (nrhs > 0 && !rnorm)) {
        g_last_error = "sla_dresidual: illegal dimensions or null rnorm";
        return SLA_ERR_ARG;
    }
    SLA_TRY(sla::dresidual_device(*reinterpret_cast<sla::Handle *>(h), n, nrhs, dA, lda, dX, ldx,
                                  dB, ldb, dR, ldr, rnorm))
}

int sla_gemm_lowp_sub(sla_handle_t h, int prec, int m, int n, int k, const float *dA, int lda,
                      const float *dB, int ldb, float *dC, int ldc)
{
    SLA_NEED_HANDLE(h);
    if (prec < SLA_PREC_FP32 || prec > SLA_PREC_FP16X3) {
        g_last_error = "sla_gemm_lowp_sub: precision must be a low-precision code (FP32 .. FP16X3)";
        return SLA_ERR_ARG;
    }
    if (m < 0 || n < 0 || k < 0 || lda < std::max(1, m) || ldb < std::max(1, k) ||
        ldc < std::max(1, m)) {
        g_last_error = "sla_gemm_lowp_sub: illegal dimensions";
        return SLA_ERR_ARG;
    }
    SLA_TRY({
        auto &H = *reinterpret_cast<sla::Handle *>(h);
        hipStream_t s = H.stream;
        if (m > 0 && n > 0 && k > 0) {
            H.timing_begin(sla::CAT_GEMM);
            if (prec == SLA_PREC_FP32) {
                sla::sgemm_nn_sub(m, n, k, dA, lda, dB, ldb, dC, ldc, s);
            } else if (const int split = sla::split_of_schur(sla::schur_of_prec(prec))) {
                const int S = sla::split_slices(split);
                const int kpad = (k + 31) / 32 * 32;
                const size_t lplane = (size_t)m * kpad;
                const size_t uplane = (size_t)n * kpad;
                void *lp = H.buffer(sla::BUF_PACK_L, lplane * 2 * S + 256);
                void *up = H.buffer(sla::BUF_PACK_U, uplane * 2 * S + 256);
                float *sc = static_cast<float *>(H.buffer(sla::BUF_SCALES, 64 * sizeof(float)));
                sla::fill_u32(reinterpret_cast<unsigned *>(sc), 2, 0u, s);
                const bool scaled = split == sla::SPLIT_FP16X3;
                if (scaled) {
                    sla::absmax_scale(m, k, dA, lda, sc, s);
                    sla::absmax_scale(k, n, dB, ldb, sc + 1, s);
                }
                sla::pack_split(split, true, m, k, kpad, dA, lda, scaled ? sc : nullptr, lp, lplane, s);
                sla::pack_split(split, false, n, k, kpad, dB, ldb, scaled ? sc + 1 : nullptr, up,
                                uplane, s);
                sla::gemm_split_nn_sub(split, m, n, kpad, lp, kpad, lplane, up, kpad, uplane, sc, dC,
                                       ldc, s);
            } else {
                const int kq = prec == SLA_PREC_BF16 ? 32 : 128;
                const int kpad = (k + kq - 1) / kq * kq;
                const size_t esz = prec == SLA_PREC_BF16 ? 2 : 1;
                void *lp = H.buffer(sla::BUF_PACK_L, (size_t)m * kpad * esz);
                void *up = H.buffer(sla::BUF_PACK_U, (size_t)n * kpad * esz);
                float *sc = static_cast<float *>(H.buffer(sla::BUF_SCALES, 64 * sizeof(float)));
                if (prec == SLA_PREC_BF16) {
                    sla::pack_rows_bf16(m, k, kpad, dA, lda, static_cast<__bf16 *>(lp), s);
                    sla::pack_cols_bf16(k, kpad, n, dB, ldb, static_cast<__bf16 *>(up), s);
                    sla::gemm_bf16_nn_sub(m, n, kpad, static_cast<__bf16 *>(lp), kpad,
                                          static_cast<__bf16 *>(up), kpad, dC, ldc, s);
                } else {
                    sla::fill_u32(reinterpret_cast<unsigned *>(sc), 2, 0u, s);
                    sla::absmax_scale(m, k, dA, lda, sc, s);
                    sla::absmax_scale(k, n, dB, ldb, sc + 1, s);
                    sla::pack_rows_fp8(m, k, kpad, dA, lda, sc, static_cast<uint8_t *>(lp), s);
                    sla::pack_cols_fp8(k, kpad, n, dB, ldb, sc + 1, static_cast<uint8_t *>(up), s);
                    sla::gemm_fp8_nn_sub(m, n, kpad, static_cast<uint8_t *>(lp), kpad,
                                         static_cast<uint8_t *>(up), kpad, sc, dC, ldc, s);
                }
            }
            H.timing_end(2.0 * m * n * k);
        }
        SLA_HIP(hipStreamSynchronize(s));
        H.timing_collect();
    })
}

}  // extern "C"
