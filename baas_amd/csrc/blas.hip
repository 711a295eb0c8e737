// blas.hip -- SpMV and BLAS-1 on gfx950 (blas/spmv.cpp, blas/blasone.cpp).
//
// SpMV keeps the reference's per-row ascending FMA sum (bit-identical to
// spmv_omp / dcsrmv) but streams the CSR through LDS: a workgroup of 256
// rows loads its whole nonzero range coalesced, then each lane reduces its
// own row from LDS in order.  Dots reduce per workgroup (wave64 shuffles)
// into partials that a fixed-order second pass adds, so results are
// deterministic run to run.
#include "blas.h"
#include "device.h"
#include "internal.h"

namespace baas {

constexpr int SPMV_ROWS = 256;
constexpr int SPMV_LDS_NNZ = 4096;  // 4096 x 12 B = 48 KB

// One row's ascending FMA sum, its x gathers issued eight at a time (slots
// past the row's end re-read its last entry and are not summed), so a row
// waits for one gather latency per eight entries instead of one per entry.
constexpr int SPMV_CH = 8;
__device__ __forceinline__ double spmv_row_from(const int *ci, const double *va, int b, int e, const double *x) {
    double s = 0.0;
    for (int j0 = b; j0 < e; j0 += SPMV_CH) {
        double xv[SPMV_CH], av[SPMV_CH];
#pragma unroll
        for (int t = 0; t < SPMV_CH; t++) {
            const int j = j0 + t < e ? j0 + t : e - 1;
            av[t] = va[j];
            xv[t] = x[ci[j]];
        }
#pragma unroll
        for (int t = 0; t < SPMV_CH; t++)
            if (j0 + t < e) s = fma(av[t], xv[t], s);
    }
    return s;
}

// y = alpha*(A x) + beta*y; with dotp != nullptr also per-block partials of
// dot(w, y_new) (used by PCG for p.Ap).  beta == 0 does not read y.
__global__ __launch_bounds__(SPMV_ROWS) void k_spmv(int m, const int *__restrict__ rowptr,
                                                    const int *__restrict__ colidx, const double *__restrict__ vals,
                                                    const double *__restrict__ x, double *y, double alpha,
                                                    double beta, const double *__restrict__ w, double *dotp,
                                                    const int *skip) {
    __shared__ int s_ci[SPMV_LDS_NNZ];
    __shared__ double s_va[SPMV_LDS_NNZ];
    __shared__ double s_red[SPMV_ROWS / 64];
    if (skip && ld_agent(skip)) return;
    const int r0 = blockIdx.x * SPMV_ROWS;
    const int r1 = min(m, r0 + SPMV_ROWS);
    const int i = r0 + threadIdx.x;
    const int b0 = rowptr[r0], b1 = rowptr[r1];
    const bool staged = (b1 - b0) <= SPMV_LDS_NNZ;
    if (staged) {
        for (int j = b0 + threadIdx.x; j < b1; j += SPMV_ROWS) {
            s_ci[j - b0] = colidx[j];
            s_va[j - b0] = vals[j];
        }
        __syncthreads();
    }
    double out = 0.0, prod = 0.0;
    if (i < r1) {
        const int rb = rowptr[i], re = rowptr[i + 1];
        double s;
        if (staged) s = spmv_row_from(s_ci, s_va, rb - b0, re - b0, x);
        else s = spmv_row_from(colidx, vals, rb, re, x);
        out = (beta == 0.0) ? alpha * s : fma(alpha, s, beta * y[i]);
        y[i] = out;
        if (dotp) prod = w[i] * out;
    }
    if (dotp) {
        double v = block_sum(prod, s_red);
        if (threadIdx.x == 0) dotp[blockIdx.x] = v;
    }
}

__global__ __launch_bounds__(256) void k_dot_partial(int n, const double *__restrict__ x,
                                                     const double *__restrict__ y, double *part) {
    __shared__ double s_red[4];
    double s = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) s = fma(x[i], y[i], s);
    s = block_sum(s, s_red);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void k_sum_partials(int np, const double *part, double *out, int do_sqrt) {
    __shared__ double s_red[4];
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += blockDim.x) s += part[i];
    s = block_sum(s, s_red);
    if (threadIdx.x == 0) *out = do_sqrt ? sqrt(s) : s;
}

// w = alpha x + beta y with the special cases of blas/blasone.cpp:24-96
__global__ void k_axpby(int n, double *w, double alpha, const double *x, double beta, const double *y) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double r;
    if (alpha == 1.0) {
        if (beta == 1.0) r = x[i] + y[i];
        else if (beta == -1.0) r = x[i] - y[i];
        else r = fma(beta, y[i], x[i]);
    } else if (beta == 1.0) {
        r = fma(alpha, x[i], y[i]);
    } else if (alpha == -1.0) {
        r = (beta == 0.0) ? -x[i] : fma(beta, y[i], -x[i]);
    } else if (beta == 0.0) {
        r = alpha * x[i];
    } else {
        r = fma(alpha, x[i], beta * y[i]);
    }
    w[i] = r;
}

static inline int nblk(long long n, int b = 256) { return (int)((n + b - 1) / b); }

int spmv_dev(int m, const int *rp, const int *ci, const double *va, const double *x, double *y, double alpha,
             double beta, const double *w, double *dotp, const int *skip, hipStream_t st) {
    hipLaunchKernelGGL(k_spmv, dim3(nblk(m, SPMV_ROWS)), dim3(SPMV_ROWS), 0, st, m, rp, ci, va, x, y, alpha, beta, w,
                       dotp, skip);
    BAAS_HIP_TRY(hipGetLastError());
    return BAAS_OK;
}

int spmv_blocks(int m) { return nblk(m, SPMV_ROWS); }

int dot_dev(baas_handle *h, int n, const double *x, const double *y, double *result, int do_sqrt, hipStream_t st) {
    const int np = 512;
    if (h->dscratch_n < (size_t)np) {
        if (h->dscratch) (void)hipFree(h->dscratch);
        h->dscratch = nullptr;
        h->dscratch_n = 0;
        BAAS_HIP_TRY(hipMalloc(&h->dscratch, np * sizeof(double)));
        h->dscratch_n = np;
    }
    int grid = nblk(n);
    if (grid > np) grid = np;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(k_dot_partial, dim3(grid), dim3(256), 0, st, n, x, y, h->dscratch);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, grid, h->dscratch, result, do_sqrt);
    BAAS_HIP_TRY(hipGetLastError());
    return BAAS_OK;
}

}  // namespace baas

using namespace baas;

extern "C" int baas_spmv_f64(baas_handle *h, int m, const int *rp, const int *ci, const double *va,
                             const double *x, double *y, double alpha, double beta) {
    if (!h || m <= 0 || !rp || !ci || !va || !x || !y) return BAAS_INVALID_ARG;
    return spmv_dev(m, rp, ci, va, x, y, alpha, beta, nullptr, nullptr, nullptr, h->stream);
}

extern "C" int baas_dot_f64(baas_handle *h, int n, const double *x, const double *y, double *result) {
    if (!h || n <= 0 || !x || !y || !result) return BAAS_INVALID_ARG;
    return dot_dev(h, n, x, y, result, 0, h->stream);
}

extern "C" int baas_nrm2_f64(baas_handle *h, int n, const double *x, double *result) {
    if (!h || n <= 0 || !x || !result) return BAAS_INVALID_ARG;
    return dot_dev(h, n, x, x, result, 1, h->stream);
}

extern "C" int baas_axpby_f64(baas_handle *h, int n, double *w, double alpha, const double *x, double beta,
                              const double *y) {
    if (!h || n <= 0 || !w || !x || !y) return BAAS_INVALID_ARG;
    hipLaunchKernelGGL(k_axpby, dim3(nblk(n)), dim3(256), 0, h->stream, n, w, alpha, x, beta, y);
    BAAS_HIP_TRY(hipGetLastError());
    return BAAS_OK;
}
