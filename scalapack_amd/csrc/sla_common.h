// sla_common.h -- shared definitions for the MI355X (gfx950) LU solve library.
//
// Layout conventions used by every kernel in this directory:
//   * matrices are column-major with a leading dimension (the LLD of the
//     ScaLAPACK descriptor on a 1 x 1 grid), element (i, j) at a[i + j*ld];
//   * pivot vectors hold 1-based global row numbers, as PDGETRF returns them
//     (SRC/pdgetrf.f:118-121);
//   * every kernel runs on the stream it is given; nothing synchronises the
//     device except the host entry points that must return data.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace sla {

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

[[noreturn]] inline void hip_fail(hipError_t e, const char *what, const char *file, int line)
{
    char buf[512];
    std::snprintf(buf, sizeof buf, "%s failed at %s:%d: %s", what, file, line,
                  hipGetErrorString(e));
    throw HipError(buf);
}

#define SLA_HIP(call)                                                          \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess) ::sla::hip_fail(e_, #call, __FILE__, __LINE__);  \
    } while (0)

#define SLA_LAUNCH_CHECK() SLA_HIP(hipGetLastError())

__host__ __device__ inline int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

// ---------------------------------------------------------------------------
// Device-side helpers
// ---------------------------------------------------------------------------

// Loads and stores of data that another workgroup of the same launch reads
// or writes.  Agent-scope relaxed atomics keep them out of the per-XCD
// non-coherent cache path; ordering comes from grid_sync's fences.
template <typename T>
__device__ __forceinline__ T coherent_load(const T *p)
{
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "word-sized only");
    if constexpr (sizeof(T) == 8) {
        uint64_t v = __hip_atomic_load(reinterpret_cast<const uint64_t *>(p),
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return __builtin_bit_cast(T, v);
    } else {
        uint32_t v = __hip_atomic_load(reinterpret_cast<const uint32_t *>(p),
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return __builtin_bit_cast(T, v);
    }
}

template <typename T>
__device__ __forceinline__ void coherent_store(T *p, T v)
{
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "word-sized only");
    if constexpr (sizeof(T) == 8) {
        __hip_atomic_store(reinterpret_cast<uint64_t *>(p), __builtin_bit_cast(uint64_t, v),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        __hip_atomic_store(reinterpret_cast<uint32_t *>(p), __builtin_bit_cast(uint32_t, v),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Device-wide barrier for a launch whose workgroups are all resident.
// `ctr` counts arrivals monotonically; the caller passes the count that marks
// this barrier (base + k * gridDim.x).  The wait is bounded: after
// `spin_limit` wall-clock ticks (100 MHz on gfx9) a workgroup raises *abort
// and every waiter leaves, so a launch whose workgroups cannot all be resident
// finishes with an error flag instead of hanging the device.  Returns false
// once the launch has been aborted.
//
// Wait until every vector-memory access this thread issued has completed
// (s_waitcnt with all counters at zero).  A coherent_store is visible
// device-wide once complete, so draining before a workgroup barrier and only
// then raising a device-wide flag publishes the stores of every wave.  A
// workgroup-scope fence alone does not wait for other waves' global stores.
__device__ __forceinline__ void drain_memory() { __builtin_amdgcn_s_waitcnt(0); }

// Memory ordering: everything the workgroups exchange across a barrier goes
// through coherent_load / coherent_store (agent-scope atomics, coherent at
// the device level without touching the per-XCD L2's other lines).  Every
// thread drains its stores before the workgroup barrier, so the arrival
// increment cannot overtake them; the fences only order this workgroup's
// accesses around the counter -- workgroup scope.  An agent-scope fence would
// write back and invalidate the whole L2 of the XCD at every barrier,
// stalling a concurrently running GEMM.
__device__ __forceinline__ bool grid_sync(unsigned *ctr, unsigned target, int *abort_flag,
                                          uint64_t spin_limit)
{
    __shared__ int s_ok;
    drain_memory();
    __syncthreads();
    if (threadIdx.x == 0) {
        int ok = 1;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");   // publish this workgroup's stores
        __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t t0 = wall_clock64();
        while (true) {
            unsigned v = __hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if ((int)(v - target) >= 0) break;
            if (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                ok = 0;
                break;
            }
            if (wall_clock64() - t0 > spin_limit) {
                __hip_atomic_store(abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");   // see the other workgroups' stores
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

// Execution interval of one launch measured inside it: the earliest
// workgroup start (stamp[0], atomic min) and the latest workgroup end
// (stamp[1], atomic max) on the device's constant-rate wall clock.  Unlike an
// event pair on the stream, it excludes the time the launch waits for CUs
// held by kernels of another stream.  stamp == nullptr: nothing is recorded.
__device__ __forceinline__ void launch_stamp_begin(unsigned long long *stamp)
{
    if (stamp && threadIdx.x == 0) atomicMin(stamp, (unsigned long long)wall_clock64());
}
__device__ __forceinline__ void launch_stamp_end(unsigned long long *stamp)
{
    if (stamp) {
        __syncthreads();
        if (threadIdx.x == 0) atomicMax(stamp + 1, (unsigned long long)wall_clock64());
    }
}

// (|v|, row) ordering of the pivot search: larger magnitude wins, the lower
// row wins a tie -- IDAMAX's "first index of the maximum" on the current row
// order (PBLAS/SRC/pdamax_.c:421-425).
// Value of v in lane `lane` (uniform index) of the wave.
template <typename T>
__device__ __forceinline__ T lane_read(T v, int lane)
{
    if constexpr (sizeof(T) == 8) {
        const uint64_t u = __builtin_bit_cast(uint64_t, v);
        const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)u, lane);
        const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(u >> 32), lane);
        return __builtin_bit_cast(T, ((uint64_t)hi << 32) | lo);
    } else {
        return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(uint32_t, v), lane));
    }
}

__device__ __forceinline__ bool pivot_better(double v1, int r1, double v2, int r2)
{
    return v1 > v2 || (v1 == v2 && r1 < r2);
}

}  // namespace sla
