// device.h -- device-side helpers shared by the sweep kernels.
//
// Cross-workgroup publication uses agent-scope relaxed atomics, which gfx950
// compiles to write-through (sc1) stores and L2-bypassing (sc1) loads, and an
// explicit s_waitcnt between a row's value stores and its ready flag.  Every
// value another workgroup may read is stored this way, so no L2 writeback
// (buffer_wbl2) is needed on the release side; the reader issues its value
// loads only after it has observed the flag.
#pragma once
#include <hip/hip_runtime.h>

#include "internal.h"

namespace baas {

template <class T>
__device__ __forceinline__ T ld_agent(const T *p) {
    return __hip_atomic_load(const_cast<T *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void st_agent(T *p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// wait until every store this thread issued has been acknowledged
__device__ __forceinline__ void wait_stores() { __builtin_amdgcn_s_waitcnt(0); }

__device__ __forceinline__ void raise_status(DevStatus *st, int code, int detail) {
    if (atomicCAS(&st->code, 0, code) == 0) st->detail = detail;
}

// Pause between polls: short at first, longer once a wait has lasted, so
// thousands of waiting lanes do not flood the memory system with polls.
__device__ __forceinline__ void poll_pause(long long spins) {
    if (spins < 32)
        __builtin_amdgcn_s_sleep(1);
    else
        __builtin_amdgcn_s_sleep(8);
}

// Spin until flags[c] == epoch; a watchdog turns a missing producer into
// BAAS_TIMEOUT instead of a hang.
__device__ __forceinline__ void wait_ready(const int *flags, int c, int epoch, long long limit,
                                           DevStatus *st) {
    long long spins = 0;
    while (ld_agent(&flags[c]) != epoch) {
        if (++spins > limit) {
            raise_status(st, BAAS_TIMEOUT, c);
            break;
        }
        poll_pause(spins);
    }
}

template <int KEYMODE>
__device__ __forceinline__ int key_of(int r, int n, const int *key) {
    if (KEYMODE == 0) return r;
    if (KEYMODE == 1) return n - 1 - r;
    return key[r];
}

// Claim the next work unit (monotone ticket, so every unit's producers were
// claimed earlier by running workgroups).
__device__ __forceinline__ int claim_unit(int *ticket, int *slot) {
    if (threadIdx.x == 0) *slot = atomicAdd(ticket, 1);
    __syncthreads();
    int u = *slot;
    __syncthreads();
    return u;
}

// Launch context: ticket, exit counter and ready-flag epoch live in device
// memory (handle->counters[0..2]) so graph replays see fresh values.
struct Ctx {
    int *ticket;
    int *done;
    int *epoch;
};

__device__ __forceinline__ int read_epoch(const Ctx &c) { return ld_agent(c.epoch); }

// The last workgroup out resets the ticket and the exit counter and advances
// the epoch, so the next launch (or graph replay) needs no memset and its
// ready flags cannot match stale ones.  Every workgroup has read the epoch
// before it counts itself out.
__device__ __forceinline__ void retire_block(const Ctx &c, int epoch) {
    if (threadIdx.x == 0) {
        __threadfence();
        int d = atomicAdd(c.done, 1);
        if (d == (int)gridDim.x - 1) {
            atomicExch(c.ticket, 0);
            atomicExch(c.epoch, epoch == 0x7fffffff ? 1 : epoch + 1);
            __threadfence();
            atomicExch(c.done, 0);
        }
    }
}

// Device-wide barrier for persistent grids whose workgroups are all
// resident: each workgroup counts itself in on one of eight partial
// counters (by workgroup id, each counter on its own cache line, so the
// arrivals do not all serialise on one address) and waits (bounded) until
// the eight sum to target.  Data crossing the barrier must be written with
// st_agent and read with ld_agent: the barrier only waits for this
// workgroup's stores to complete before counting in, so no cache write-back
// or invalidation (which would cost the whole L2) is needed.
__device__ __forceinline__ int barrier_count(int *bar) {
    int v[8];
#pragma unroll
    for (int p = 0; p < 8; p++) v[p] = ld_agent(&bar[32 * p]);
    int s = 0;
#pragma unroll
    for (int p = 0; p < 8; p++) s += v[p];
    return s;
}

__device__ __forceinline__ void grid_barrier(int *bar, int target, DevStatus *status, long long limit) {
    wait_stores();
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd(&bar[32 * (blockIdx.x & 7)], 1);
        long long spins = 0;
        while (barrier_count(bar) < target) {
            if (++spins > limit) {
                raise_status(status, BAAS_TIMEOUT, -1);
                break;
            }
            poll_pause(spins);
        }
    }
    __syncthreads();
}

// Reset the barrier's partial counters (the last workgroup out does this).
__device__ __forceinline__ void reset_barrier(int *bar) {
#pragma unroll
    for (int p = 0; p < 8; p++) atomicExch(&bar[32 * p], 0);
}

// Value-as-flag publication: a signalling-NaN bit pattern that arithmetic
// never produces marks a slot whose value has not arrived yet.
template <class T> struct Sentinel;
template <> struct Sentinel<double> {
    static __device__ __forceinline__ double value() { return __longlong_as_double(0x7ff4deadbeef0001LL); }
    static __device__ __forceinline__ bool is(double v) { return __double_as_longlong(v) == 0x7ff4deadbeef0001LL; }
};
template <> struct Sentinel<float> {
    static __device__ __forceinline__ float value() { return __int_as_float(0x7fa0dead); }
    static __device__ __forceinline__ bool is(float v) { return __float_as_int(v) == 0x7fa0dead; }
};

// LDS accesses that other wavefronts of the workgroup poll
template <class T>
__device__ __forceinline__ T ld_lds(const T *p) {
    return __hip_atomic_load(const_cast<T *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <class T>
__device__ __forceinline__ void st_lds(T *p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

constexpr int DF_PF = 8;  // entries an exchange lane polls at once

// Exchange wavefront, incoming side: copy published values from other
// workgroups into LDS slots as they appear.  With nr values per source,
// item i is source cols[i / nr], value i % nr, read at cols[.] * nr + i % nr
// and staged at slot base * nr + i.  Each lane walks its own stride of the
// (first-use ordered) list, DF_PF items at a time: it polls all of them with
// one load each per pass and stages each as soon as it is published.  A
// wait beyond the spin limit reports BAAS_TIMEOUT and stages zeros, so the
// workgroup still finishes.
template <class T>
__device__ __forceinline__ void stage_published(const T *pub, T *sl, const int *cols, int base, int ne, int nr,
                                                int lane, long long limit, DevStatus *status) {
    const int items = ne * nr;
    for (int e = lane; e < items; e += DF_PF * 64) {
        int c[DF_PF];
#pragma unroll
        for (int k = 0; k < DF_PF; k++) {
            const int ek = e + k * 64 < items ? e + k * 64 : items - 1;
            c[k] = nr == 1 ? cols[ek] : cols[ek / nr] * nr + ek % nr;
        }
        unsigned todo = 0;
#pragma unroll
        for (int k = 0; k < DF_PF; k++)
            if (e + k * 64 < items) todo |= 1u << k;
        long long spins = 0;
        while (todo) {
            T v[DF_PF];
#pragma unroll
            for (int k = 0; k < DF_PF; k++) v[k] = ld_agent(&pub[c[k]]);
#pragma unroll
            for (int k = 0; k < DF_PF; k++) {
                if (((todo >> k) & 1u) && !Sentinel<T>::is(v[k])) {
                    st_lds(&sl[base * nr + e + k * 64], v[k]);
                    todo &= ~(1u << k);
                }
            }
            if (todo) {
                if (++spins > limit) {
                    raise_status(status, BAAS_TIMEOUT, c[__builtin_ctz(todo)]);
#pragma unroll
                    for (int k = 0; k < DF_PF; k++)
                        if ((todo >> k) & 1u) st_lds(&sl[base * nr + e + k * 64], T(0));
                    todo = 0;
                } else {
                    __builtin_amdgcn_s_sleep(1);
                }
            }
        }
    }
}

}  // namespace baas
