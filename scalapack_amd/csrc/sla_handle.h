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

enum Category { CAT_PANEL = 0, CAT_LASWP, CAT_TRSM, CAT_GEMM, CAT_TRSV, CAT_OTHER, CAT_COUNT };

struct Stat {
    double ms = 0.0;      // summed kernel time (HIP events around each launch)
    double work = 0.0;    // summed algorithmic flops (GEMM/TRSM/PANEL) or bytes (LASWP)
    long long launches = 0;
};

struct Handle {
    Handle();
    ~Handle();
    Handle(const Handle &) = delete;
    Handle &operator=(const Handle &) = delete;

    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;   // main stream (the caller's, or own_stream)
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

    double *slot_cand_v = nullptr;
    int *slot_cand_r = nullptr;
    double *slot_rows = nullptr, *slot_rowj = nullptr;  // sized for doubles, reused for floats
    int slot_gmax = 0;

    int *pairs = nullptr;        // two pair lists of 2048 ints
    int *pair_counts = nullptr;  // their counts
    int *trsv_flags = nullptr;
    int trsv_max_blocks = 0, trsv_epoch = 0;

    int *perm = nullptr;
    int perm_cap = 0;
    const void *perm_key_a = nullptr;
    const int *perm_key_ipiv = nullptr;
    int perm_m = 0;

    char *work = nullptr;
    size_t work_cap = 0;

    // tuning knobs
    int nb = 256;          // block column width of the outer factorization
    int leaf_w = 32;       // panel leaf width (columns per device-wide-barrier kernel)
    int panel_wgs = 32;    // preferred workgroups per leaf launch

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
    void timing_begin(int cat, hipStream_t s = nullptr);
    void timing_end(double work_units, hipStream_t s = nullptr);
    void timing_collect();
    void reset_stats();

    template <typename T>
    PanelWork<T> panel_work();
    TrsvWork trsv_work();
    void ensure_perm(int m);
    void ensure_work(size_t bytes);
};

// Device-resident FP64 factorization / solve (pointers are device memory).
// Returns INFO > 0 for an exactly zero pivot U(INFO, INFO); throws on a
// runtime failure.
int dgetrf_device(Handle &h, int m, int n, double *dA, int lda, int *dipiv);
int dgetrs_device(Handle &h, int n, int nrhs, const double *dA, int lda, const int *dipiv,
                  double *dB, int ldb);
// PDLASWP on a device matrix: row interchanges k1 .. k2-1 (0-based) with
// rows ipiv[k]-1 over ncols columns, forward or backward; and the column form.
void laswp_rows_device(Handle &h, double *dA, int lda, int ncols, int k1, int k2,
                       const int *dipiv, bool forward);
void laswp_cols_device(Handle &h, double *dA, int lda, int nrows, int k1, int k2,
                       const int *dipiv, bool forward);

}  // namespace sla
