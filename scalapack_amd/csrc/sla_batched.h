// sla_batched.h -- many small independent systems factored and solved in one
// persistent launch (batched.hip).
#pragma once

#include "sla_handle.h"

#include <vector>

namespace sla {

constexpr int BATCHED_NMAX = 1024, BATCHED_NRHS_MAX = 16;
// Systems per call of the allocation-free entry (its fixed scratch array).
constexpr int BATCHED_NOALLOC_MAX = 65536;

struct BatchedTune {
    struct Entry {
        int w, wgs_per_cu;
        double ms;
    };
    int w = 16, wgs_per_cu = 2;
    double ms = 0.0;
    std::vector<Entry> tried;
};

// A_s X_s = B_s for s < batch: n, nrhs, lda, ldb are host arrays; dA, dipiv,
// dB are device arrays of device pointers; dinfo a device int[batch]
// (PDGESV's INFO per system).  Uses the handle's panel width / grid
// (batched_w, batched_wgs_per_cu) and, with batched_graph, a captured
// hipGraph replayed while the pointer arrays and sizes stay the same.
template <typename T>
void gesv_batched_device(Handle &h, int batch, const int *n, const int *nrhs, T *const *dA,
                         const int *lda, int *const *dipiv, T *const *dB, const int *ldb,
                         int *dinfo);
// The same with the sizes in device arrays and no device memory allocated:
// the queue, the order and the descriptors live in a fixed device array of
// this library (at most BATCHED_NOALLOC_MAX systems).  A system whose sizes
// are out of range gets INFO = -1 and no work.  Only enqueues work on the
// handle's stream (no synchronisation, so a caller may capture it in a
// graph); calls that overlap on one GPU share the array and race.
template <typename T>
void gesv_batched_noalloc_device(Handle &h, int batch, const int *dn, const int *dnrhs,
                                 T *const *dA, const int *dlda, int *const *dipiv, T *const *dB,
                                 const int *dldb, int *dinfo);
// Times every (panel width, workgroups per CU) on synthetic systems of these
// sizes, stores the fastest in the handle and returns the table.
template <typename T>
BatchedTune tune_batched_device(Handle &h, int batch, const int *n, const int *nrhs, int reps);

}  // namespace sla
