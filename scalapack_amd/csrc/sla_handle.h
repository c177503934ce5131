// sla_handle.h -- per-device state of the library: streams, the small device
// control block used by the device-wide barriers and dependency waits, the
// panel's exchange slots, reusable work buffers, tuning knobs and optional
// per-kernel timing.
#pragma once

#include "sla_common.h"
#include "sla_kernels.h"

#include <array>
#include <vector>

namespace sla {

// CAT_GEMM: the outer trailing updates; CAT_GEMM_PANEL: the GEMMs inside the
// recursive panel and TRSM.
enum Category {
    CAT_PANEL = 0,
    CAT_LASWP,
    CAT_TRSM,
    CAT_GEMM,
    CAT_TRSV,
    CAT_OTHER,
    CAT_GEMM_PANEL,
    CAT_COUNT
};

// Precision of the outer trailing update of an FP32-storage factorization:
// FP32 MFMA, BF16 or FP8 operands, or split 16-bit products (Split).
enum Schur {
    SCHUR_NATIVE = 0,
    SCHUR_BF16 = 1,
    SCHUR_FP8 = 2,
    SCHUR_BF16X3 = 3,
    SCHUR_BF16X6 = 4,
    SCHUR_FP16X3 = 5
};

enum Precision {
    PREC_FP64 = 0,
    PREC_FP32 = 1,
    PREC_BF16 = 2,
    PREC_FP8 = 3,
    PREC_BF16X3 = 4,
    PREC_BF16X6 = 5,
    PREC_FP16X3 = 6,
    PREC_COUNT
};

// The trailing-update precision of a low-precision factorization.
inline int schur_of_prec(int prec)
{
    switch (prec) {
    case PREC_BF16: return SCHUR_BF16;
    case PREC_FP8: return SCHUR_FP8;
    case PREC_BF16X3: return SCHUR_BF16X3;
    case PREC_BF16X6: return SCHUR_BF16X6;
    case PREC_FP16X3: return SCHUR_FP16X3;
    default: return SCHUR_NATIVE;
    }
}
inline int split_of_schur(int schur)
{
    return schur == SCHUR_BF16X3 ? SPLIT_BF16X3
         : schur == SCHUR_BF16X6 ? SPLIT_BF16X6
         : schur == SCHUR_FP16X3 ? SPLIT_FP16X3 : 0;
}
enum Refinement { REFINE_IR = 0, REFINE_GMRES = 1 };

struct MixedInfo {
    int precision = 0;   // factorization precision asked for (Precision)
    int refinement = 0;  // refinement that produced the result (Refinement)
    int iter = 0;        // DSGESV's ITER: refinement steps, < 0 after the FP64 fallback
    int outer = 0;       // FP64 residuals computed after the first solve
    int inner = 0;       // GMRES steps in total
    int fell_back = 0;   // 1 if the FP64 factorization produced X
    double anorm = 0.0;  // ||A||_inf
    double bwd_err = 0.0;  // max_j ||r_j||_inf / (||A||_inf ||x_j||_inf) of the last residual
    double sresid = 0.0;   // the same divided by N eps (the testers' scaled residual)
    double history[16] = {};  // bwd_err after the first solve and each refinement step
    int stop_rule = 0;        // StopRule: which test ended the refinement
};

// Why the refinement of a mixed-precision solve stopped.
enum StopRule {
    STOP_CONVERGED = 0,       // DSGESV's test: ||r|| <= ||x|| ||A|| sqrt(N) eps
    STOP_STALL_ACCEPTED = 1,  // a stall (PDGERFS's rule) with scaled residual < 16
    STOP_FALLBACK = 2,        // X came from the FP64 factorization
    STOP_FAILED = 3           // refinement failed, FP64 fallback disabled: X is the last iterate
};

// Reusable device buffers, grown on demand.
enum BufId {
    BUF_PACK_L = 0,  // packed low-precision L21 of one block column
    BUF_PACK_U,      // packed low-precision U12 of one block row
    BUF_SCALES,      // FP8 block scales
    BUF_SA,          // FP32 copy of the matrix (mixed-precision solve)
    BUF_IPIV_S,      // its pivots
    BUF_VEC,         // refinement vectors (FP64 / FP32)
    BUF_KRYLOV,      // GMRES bases
    BUF_PARTIAL,     // partial sums of the residual / reductions
    BUF_TSLU,        // candidate exchange of the tournament-pivoting leaf
    BUF_AUX,         // vectors and partial sums of the norms / estimators / PDGERFS
    BUF_MLP,         // activations and gradients of the selector network
    BUF_SELECT,      // sampled submatrices of the selector's features
    BUF_MODEL,       // the selector's parameters (sla_selector_set_model)
    BUF_TRSV_MAIL,   // epoch-tagged hand-off words of the persistent TRSV
    BUF_DINV,        // inverses of the diagonal blocks of the low-precision factors
    BUF_TRSM_INV,    // inverses of the 32 x 32 diagonal blocks of L11 (two streams)
    BUF_COUNT
};

struct Stat {
    double ms = 0.0;      // summed kernel time (HIP events around each launch)
    double work = 0.0;    // summed algorithmic flops (GEMM/TRSM/PANEL) or bytes (LASWP)
    long long launches = 0;
};

struct BatchedPlan;
void batched_plan_free(BatchedPlan *p);

struct Handle {
    Handle();
    ~Handle();
    Handle(const Handle &) = delete;
    Handle &operator=(const Handle &) = delete;

    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;   // main stream (own_stream, or the caller's; may be
                                    // HIP's default stream, nullptr)
    hipStream_t stream2 = nullptr;  // look-ahead panel stream (default priority)
    hipEvent_t ev_next = nullptr, ev_panel = nullptr, ev_join = nullptr;
    uint64_t spin_limit = 0;

    // device control block: [0] barrier counter, [1] abort flag,
    // [2] first zero-pivot column (INT_MAX: none), [3] ticket, [4] pair count,
    // [5] pivot-row-out-of-range flag
    int *ctl = nullptr;
    unsigned *ctr = nullptr;
    int *abort_flag = nullptr, *zero_col = nullptr, *ticket = nullptr, *pair_count = nullptr;
    int *bad_pivot = nullptr;
    unsigned ctr_base = 0;

    uint64_t *panel_mail = nullptr;     // the panel leaf's mailbox (PANEL_MAIL_WORDS)
    PanelMailState panel_mail_state;    // its last launch

    int *pairs = nullptr;        // two pair lists of 2048 ints
    int *pair_counts = nullptr;  // their counts
    int trsv_epoch = 0;

    int *perm = nullptr;
    int perm_cap = 0;
    const void *perm_key_a = nullptr;
    const int *perm_key_ipiv = nullptr;
    int perm_m = 0;

    char *work = nullptr;
    size_t work_cap = 0;

    struct Buf {
        void *p = nullptr;
        size_t cap = 0;
    };
    std::array<Buf, BUF_COUNT> bufs{};
    // Device buffer `id` of at least `bytes` (the streams are synchronised
    // before it grows).
    void *buffer(int id, size_t bytes);

    // tuning knobs
    int nb = 256;          // block column width of the outer factorization
    int leaf_w = 32;       // panel leaf width (columns per device-wide-barrier kernel)
    int panel_wgs = 32;    // preferred workgroups per leaf launch
    int pivoting = 1;      // FP32-storage factorization: 0 partial, 1 tournament
    int lookahead = 1;     // factor block column p+1 on a second stream during p's update
    int update_cus = 0;    // > 0: the trailing update runs on at most this many CUs' worth
                           // of workgroups (the rest stay free for the panel)
    int num_cus = 256;     // CUs of the device
    int update_lds_kb = 112;  // LDS (KiB) reserved per workgroup of the FP32 trailing update:
                              // one of its workgroups per CU, leaving registers to the panel
                              // leaf that runs beside it (0: no reservation)
    int lacon_lapack = 0;  // condition / error estimates: 0 the reference's PDLACON result,
                           // 1 LAPACK DLACON's iteration (estimate kept across steps)
    // batched solves
    int batched_w = 16;           // panel width of the per-system factorization (8, 16, 32)
    int batched_wgs_per_cu = 2;   // persistent workgroups per CU
    int batched_graph = 0;        // 1: launch through a captured, replayed hipGraph
    BatchedPlan *batched = nullptr;
    bool has_model = false;       // a selector model is in BUF_MODEL
    // mixed-precision refinement
    int ir_max_iter = 30;        // refinement steps before the FP64 fallback (DSGESV's ITERMAX)
    int ir_gmres_on_stall = 1;   // classical refinement that stalls continues as GMRES-IR
    int ir_fallback = 1;         // a failed refinement is redone by the FP64 factorization
                                 // (0: report the failure, ITER < 0, and leave A as it is)
    int ir_accept_stall = 1;     // a stalled refinement below scaled residual 16 is a result
                                 // (0: only DSGESV's convergence test ends it)
    int trsv_blockinv = 1;       // solves with low-precision factors use the inverses of
                                 // their 128 x 128 diagonal blocks (0: substitution)
    int trsm_mfma = 1;           // U12 solves: one blocked MFMA launch per block (0: strips)
    int gmres_m = 50;            // GMRES steps per correction
    double gmres_tol = 1e-6;     // relative residual reduction per correction

    MixedInfo last_mixed;  // details of the last mixed-precision solve

    // timing
    bool profile = false;
    std::array<Stat, CAT_COUNT> stats{};
    struct Pending {
        int cat;
        double work;
        hipEvent_t e0, e1;
        hipStream_t s;
        bool closed = false;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> ev_pool, ev_free;
    // execution intervals of the outer trailing-update launches, measured
    // inside the kernels while profiling: device pairs (first workgroup
    // start, last workgroup end) in wall-clock ticks, summed into `exec`
    static constexpr int KSTAMP_CAP = 4096;
    unsigned long long *kstamps = nullptr;
    int kstamp_used = 0;
    std::vector<double> kstamp_work;  // algorithmic flops of each stamped launch
    double wall_khz = 100000.0;
    Stat exec{};
    // the next stamp pair for a trailing-update launch of `work` flops, or
    // nullptr when not profiling (or the pairs of this collection are used up)
    unsigned long long *next_kstamp(double work);
    void timing_begin(int cat, hipStream_t s = nullptr);
    void timing_end(double work_units, hipStream_t s = nullptr);
    void timing_collect();
    void reset_stats();

    template <typename T>
    PanelWork<T> panel_work();
    TrsvWork trsv_work(int n);  // mailbox sized for triangles of order n
    void ensure_perm(int m);
    void ensure_work(size_t bytes);
};

// Device-resident FP64 factorization / solve (pointers are device memory).
// Returns INFO > 0 for an exactly zero pivot U(INFO, INFO); throws on a
// runtime failure.
int dgetrf_device(Handle &h, int m, int n, double *dA, int lda, int *dipiv);
// The same on FP32 storage; `schur` (enum Schur) selects the precision of the
// outer trailing update.
int sgetrf_device(Handle &h, int m, int n, float *dA, int lda, int *dipiv, int schur);
// trans: solve A^T X = B (PDGETRS TRANS = 'T' or 'C') instead of A X = B.
int dgetrs_device(Handle &h, int n, int nrhs, const double *dA, int lda, const int *dipiv,
                  double *dB, int ldb, bool trans = false);
// The factorization's net row permutation rebuilt from IPIV alone (h.perm;
// the cached association with a factorization is dropped).
void perm_from_ipiv(Handle &h, int n, const int *dipiv);
// Solve with the permutation already in h.perm (no wait for completion;
// check_solve_flags reports what the kernels flagged).
void getrs_perm_device(Handle &h, bool trans, int n, int nrhs, const double *dA, int lda,
                       double *dB, int ldb);
void check_solve_flags(Handle &h, const char *routine);
// PDLASWP on a device matrix: row interchanges k1 .. k2-1 (0-based) with
// rows ipiv[k]-1 over ncols columns, forward or backward; and the column form.
void laswp_rows_device(Handle &h, double *dA, int lda, int ncols, int k1, int k2,
                       const int *dipiv, bool forward);
void laswp_cols_device(Handle &h, double *dA, int lda, int nrows, int k1, int k2,
                       const int *dipiv, bool forward);

}  // namespace sla
