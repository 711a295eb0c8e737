// blas.h -- reductions and host entry points shared by blas.hip and pcg.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "internal.h"

namespace baas {

// Sum over a 64-wide wavefront with DPP/shuffle moves (fixed tree order).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Sum over the workgroup; the result is valid in thread 0.  `red` holds one
// double per wave.  Every thread of the block must call it.
__device__ __forceinline__ double block_sum(double v, double *red) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) red[wv] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        for (int k = 0; k < nw; k++) s += red[k];
    }
    __syncthreads();
    return s;
}

int spmv_dev(int m, const int *rp, const int *ci, const double *va, const double *x, double *y, double alpha,
             double beta, const double *w, double *dotp, const int *skip, hipStream_t st);
int spmv_blocks(int m);
int dot_dev(baas_handle *h, int n, const double *x, const double *y, double *result, int do_sqrt, hipStream_t st);

// fixed_xp: device address of a caller-chosen publication area of
// fixed_bytes (re-armed by this call) instead of the plan-owned one, which
// the plan allocates on first use; nullptr = the plan-owned area.
// unit_skip: optional per-unit flags of the plan in use (nonzero = skip;
// honoured by the multi-RHS lane kernel only).
template <class T>
int sweep(baas_handle *h, baas_schedule *s, int uplo, const T *vals, const T *b, T *x, int nrhs, hipStream_t st,
          const int *skip, void *fixed_xp = nullptr, size_t fixed_bytes = 0, const int *unit_skip = nullptr);
// Look up this device's copy of the allocation-free entry point's scratch
// array (once per device; the device must be current).
int noalloc_scratch_init(int device);
int ic0_split_dev(baas_handle *h, baas_schedule *s, const double *lu, double *lv, double *uv, hipStream_t st);
int ic0_split_dev32(baas_handle *h, baas_schedule *s, const double *lu, float *lv, float *uv, hipStream_t st);

}  // namespace baas
