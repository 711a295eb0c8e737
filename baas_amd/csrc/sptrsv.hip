// sptrsv.hip -- triangular solves on gfx950.
//
// Row arithmetic follows the reference exactly: the lower solve sums
// L_ij * x_j in ascending column order with fused multiply-adds and divides
// by the diagonal (merge/SpTRSV.cpp:24-39, the reference build contracts the
// sum to FMA); the upper solve on U = L^T sums in descending column order as
// the backward P2P kernel does (merge/SpTRSV.cpp:643-652).  Every schedule
// therefore produces the serial solve's bits.
//
// k_sweep is the persistent, sync-free replacement of the P2P group-merged
// kernel (merge/SpTRSV.cpp:421-478): workgroups claim work units in order;
// inside a unit, sub-levels are separated by barriers and solved values are
// staged in LDS; parents in earlier units are waited for on per-row ready
// flags (epoch-stamped, so no reset between calls).
#include <algorithm>

#include "device.h"
#include "internal.h"

namespace baas {

struct SweepArgs {
    int n, nunits, nrhs;
    const int *rowptr, *colidx;
    const int *key_begin, *sub_begin, *sub_ptr, *rows, *key;
    int *flags;
    Ctx ctx;
    const int *skip;  // optional device flag: nonzero = return at once (converged PCG)
    DevStatus *status;
    long long spin_limit;
};

template <class T, int KEYMODE, bool UPPER, bool LDS>
__global__ __launch_bounds__(1024) void k_sweep(SweepArgs a, const T *__restrict__ vals,
                                               const T *__restrict__ b, T *x) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *xs = reinterpret_cast<T *>(smem_raw);
    __shared__ int s_unit;
    if (a.skip && ld_agent(a.skip)) return;
    const int nrhs = a.nrhs;
    const int epoch = read_epoch(a.ctx);
    for (;;) {
        const int u = claim_unit(a.ctx.ticket, &s_unit);
        if (u >= a.nunits) break;
        const int kb = a.key_begin[u];
        const int s0 = a.sub_begin[u], s1 = a.sub_begin[u + 1];
        for (int si = s0; si < s1; si++) {
            const int qb = a.sub_ptr[si], qe = a.sub_ptr[si + 1];
            const int items = (qe - qb) * nrhs;
            for (int t = threadIdx.x; t < items; t += blockDim.x) {
                const int q = qb + (nrhs == 1 ? t : t / nrhs);
                const int r = nrhs == 1 ? 0 : t - (q - qb) * nrhs;
                const int rw = a.rows[q];
                const int row = rw & 0x7fffffff;
                const int rb = a.rowptr[row], re = a.rowptr[row + 1];
                const int ob = UPPER ? rb + 1 : rb, oe = UPPER ? re : re - 1;  // off-diagonals
                const int dg = UPPER ? rb : re - 1;
                // wait for parents produced by other units
                for (int j = ob; j < oe; j++) {
                    const int c = a.colidx[j];
                    if (key_of<KEYMODE>(c, a.n, a.key) < kb) wait_ready(a.flags, c, epoch, a.spin_limit, a.status);
                }
                T s = T(0);
                if (UPPER) {
                    for (int j = oe - 1; j >= ob; j--) {
                        const int c = a.colidx[j];
                        const int kc = key_of<KEYMODE>(c, a.n, a.key);
                        T xv;
                        if (LDS && kc >= kb) xv = xs[(size_t)(kc - kb) * nrhs + r];
                        else xv = ld_agent(&x[(size_t)c * nrhs + r]);
                        s = fma(xv, vals[j], s);
                    }
                } else {
                    for (int j = ob; j < oe; j++) {
                        const int c = a.colidx[j];
                        const int kc = key_of<KEYMODE>(c, a.n, a.key);
                        T xv;
                        if (LDS && kc >= kb) xv = xs[(size_t)(kc - kb) * nrhs + r];
                        else xv = ld_agent(&x[(size_t)c * nrhs + r]);
                        s = fma(vals[j], xv, s);
                    }
                }
                const T xr = (b[(size_t)row * nrhs + r] - s) / vals[dg];
                st_agent(&x[(size_t)row * nrhs + r], xr);
                if (LDS) xs[(size_t)(key_of<KEYMODE>(row, a.n, a.key) - kb) * nrhs + r] = xr;
                if (nrhs == 1 && rw < 0) {
                    wait_stores();
                    st_agent(&a.flags[row], epoch);
                }
            }
            if (nrhs > 1) {
                wait_stores();
                __syncthreads();
                for (int t = threadIdx.x; t < qe - qb; t += blockDim.x) {
                    const int rw = a.rows[qb + t];
                    if (rw < 0) st_agent(&a.flags[rw & 0x7fffffff], epoch);
                }
            } else if (!LDS) {
                wait_stores();
                __syncthreads();
            } else {
                __syncthreads();
            }
        }
    }
    retire_block(a.ctx, epoch);
}

// ---------------------------------------------------------------------------
// Dataflow sweep.  One lane owns one row; the unit's solved values live in
// LDS slots pre-filled with a signalling-NaN sentinel, so a lane reads its
// parent's slot until the sentinel is gone -- no sub-level barriers, every
// row waits exactly for its own parents.  Parents in earlier units are staged
// into LDS slots (after the unit's own keys) by dedicated prefetch waves that
// poll the per-row ready flags, so the flag round trips overlap with compute.
// Each lane retries in one loop that consumes every ready entry in ascending
// (descending for U) column order, so lanes of one wavefront that depend on
// each other still make progress.
template <class T> struct Sentinel;
template <> struct Sentinel<double> {
    static __device__ __forceinline__ double value() { return __longlong_as_double(0x7ff4deadbeef0001LL); }
    static __device__ __forceinline__ bool is(double v) { return __double_as_longlong(v) == 0x7ff4deadbeef0001LL; }
};
template <> struct Sentinel<float> {
    static __device__ __forceinline__ float value() { return __int_as_float(0x7fa0dead); }
    static __device__ __forceinline__ bool is(float v) { return __float_as_int(v) == 0x7fa0dead; }
};

template <class T>
__device__ __forceinline__ T ld_lds(const T *p) {
    return __hip_atomic_load(const_cast<T *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <class T>
__device__ __forceinline__ void st_lds(T *p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

struct DfArgs {
    int n, nunits, npf;
    const int *rowptr, *scol;
    const int *key_begin, *unit_ptr, *rows, *key, *ext_ptr, *ext_cols;
    int *flags;
    Ctx ctx;
    const int *skip;
    DevStatus *status;
    long long spin_limit;
};

constexpr int DF_CHUNK = 8;

template <class T, int KEYMODE, bool UPPER>
__global__ __launch_bounds__(1024) void k_sweep_df(DfArgs a, const T *__restrict__ vals, const T *__restrict__ b,
                                                  T *x) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *sl = reinterpret_cast<T *>(smem_raw);
    __shared__ int s_unit;
    if (a.skip && ld_agent(a.skip)) return;
    const int epoch = read_epoch(a.ctx);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int npf = a.npf;
    for (;;) {
        const int u = claim_unit(a.ctx.ticket, &s_unit);
        if (u >= a.nunits) break;
        const int kb = a.key_begin[u], nk = a.key_begin[u + 1] - kb;
        const int e0 = a.ext_ptr[u], e1 = a.ext_ptr[u + 1];
        const int nslots = nk + (e1 - e0);
        for (int i = threadIdx.x; i < nslots; i += blockDim.x) st_lds(&sl[i], Sentinel<T>::value());
        __syncthreads();
        if (wv < npf) {
            // prefetch waves: stage external parents in order of first use
            for (int e = e0 + wv * 64 + lane; e < e1; e += npf * 64) {
                const int c = a.ext_cols[e];
                wait_ready(a.flags, c, epoch, a.spin_limit, a.status);
                st_lds(&sl[nk + (e - e0)], ld_agent(&x[c]));
            }
        } else {
            const int q0 = a.unit_ptr[u], q1 = a.unit_ptr[u + 1];
            const int C = nw - npf, cw = wv - npf;
            for (int base = q0 + cw * 64; base < q1; base += C * 64) {
                const int q = base + lane;
                bool pending = q < q1;
                int rw = 0, row = 0, dg = 0, left = 0, jn = 0;
                T bv = T(0), s = T(0);
                if (pending) {
                    rw = a.rows[q];
                    row = rw & 0x7fffffff;
                    const int rb = a.rowptr[row], re = a.rowptr[row + 1];
                    dg = UPPER ? rb : re - 1;
                    left = re - rb - 1;               // off-diagonal entries still to consume
                    jn = UPPER ? re - 1 : rb;         // next entry to load
                    bv = b[row];
                }
                int cs[DF_CHUNK];
                T cv[DF_CHUNK];
                int nc = 0, pc = 0;
                long long spins = 0;
                while (pending) {
                    if (pc == nc && left > 0) {  // load the next chunk of entries
                        nc = left < DF_CHUNK ? left : DF_CHUNK;
#pragma unroll
                        for (int t = 0; t < DF_CHUNK; t++) {
                            if (t < nc) {
                                const int j = UPPER ? jn - t : jn + t;
                                cs[t] = a.scol[j];
                                cv[t] = vals[j];
                            }
                        }
                        jn = UPPER ? jn - nc : jn + nc;
                        left -= nc;
                        pc = 0;
                    }
                    if (pc < nc) {
#pragma unroll
                        for (int t = 0; t < DF_CHUNK; t++) {
                            if (t == pc && t < nc) {
                                const T v = ld_lds(&sl[cs[t]]);
                                if (!Sentinel<T>::is(v)) {
                                    s = UPPER ? fma(v, cv[t], s) : fma(cv[t], v, s);
                                    pc++;
                                }
                            }
                        }
                        if (pc < nc && ++spins > a.spin_limit) {
                            raise_status(a.status, BAAS_TIMEOUT, row);
                            pc = nc;  // give up on this row's remaining parents
                            left = 0;
                        }
                    } else {
                        const T xr = (bv - s) / vals[dg];
                        st_lds(&sl[key_of<KEYMODE>(row, a.n, a.key) - kb], xr);
                        st_agent(&x[row], xr);
                        if (rw < 0) {
                            wait_stores();
                            st_agent(&a.flags[row], epoch);
                        }
                        pending = false;
                    }
                }
            }
        }
        __syncthreads();
    }
    retire_block(a.ctx, epoch);
}

// Level-synchronous solve: one launch per row level (merge/SpTRSV.cpp:41-80).
template <class T>
__global__ void k_level_rows(const int *__restrict__ perm, int qb, int qe, const int *__restrict__ rowptr,
                             const int *__restrict__ colidx, const T *__restrict__ vals, const T *__restrict__ b,
                             T *x) {
    int q = qb + blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= qe) return;
    int row = perm[q];
    int rb = rowptr[row], re = rowptr[row + 1] - 1;
    T s = T(0);
    for (int j = rb; j < re; j++) s = fma(vals[j], x[colidx[j]], s);
    x[row] = (b[row] - s) / vals[re];
}

template <class T, class S>
__global__ void k_gather(int n, const int *__restrict__ src, const S *__restrict__ in, T *__restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (T)in[src[i]];
}

template <class T, class S>
__global__ void k_cast(int n, const S *__restrict__ in, T *__restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (T)in[i];
}

// dL_ij = -sum_r lambda_{i,r} x_{j,r}
__global__ void k_grad_vals(int nnz, int nrhs, const int *__restrict__ row_of, const int *__restrict__ colidx,
                            const double *__restrict__ lam, const double *__restrict__ x, double *__restrict__ dv) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    int i = row_of[e], j = colidx[e];
    double s = 0.0;
    for (int r = 0; r < nrhs; r++) s = fma(lam[(size_t)i * nrhs + r], x[(size_t)j * nrhs + r], s);
    dv[e] = -s;
}

// dA_ij = -sum_r lambda_{i,r} x_{j,r} over any CSR pattern, one lane per row
__global__ void k_csr_outer_grad(int m, int nrhs, const int *__restrict__ rp, const int *__restrict__ ci,
                                 const double *__restrict__ lam, const double *__restrict__ x, double *__restrict__ dv) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    for (int e = rp[i]; e < rp[i + 1]; e++) {
        const int j = ci[e];
        double s = 0.0;
        for (int r = 0; r < nrhs; r++) s = fma(lam[(size_t)i * nrhs + r], x[(size_t)j * nrhs + r], s);
        dv[e] = -s;
    }
}

static inline int nblk(long long n, int b = 256) { return (int)((n + b - 1) / b); }

template <class T, int KM, bool UP, bool LDS>
static int launch_sweep_t(baas_handle *h, const Plan &p, const SweepArgs &a, const T *vals, const T *b, T *x,
                          size_t lds, hipStream_t stream) {
    auto kern = k_sweep<T, KM, UP, LDS>;
    int per_cu = 0;
    BAAS_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, p.block, lds));
    if (per_cu <= 0) return BAAS_UNSUPPORTED;
    long long grid = (long long)per_cu * h->num_cus;
    if (grid > p.nunits) grid = p.nunits;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(p.block), lds, stream, a, vals, b, x);
    BAAS_HIP_TRY(hipGetLastError());
    return BAAS_OK;
}

template <class T, bool UP, bool LDS>
static int launch_sweep_k(baas_handle *h, const Plan &p, const SweepArgs &a, const T *vals, const T *b, T *x,
                          size_t lds, hipStream_t st) {
    switch (p.keymode) {
        case 0: return launch_sweep_t<T, 0, UP, LDS>(h, p, a, vals, b, x, lds, st);
        case 1: return launch_sweep_t<T, 1, UP, LDS>(h, p, a, vals, b, x, lds, st);
        default: return launch_sweep_t<T, 2, UP, LDS>(h, p, a, vals, b, x, lds, st);
    }
}

size_t lds_limit(baas_handle *h);

template <class T, int KM, bool UP>
static int launch_df_t(baas_handle *h, const Plan &p, const DfArgs &a, const T *vals, const T *b, T *x, size_t lds,
                       hipStream_t stream) {
    auto kern = k_sweep_df<T, KM, UP>;
    int per_cu = 0;
    BAAS_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, p.block, lds));
    if (per_cu <= 0) return BAAS_UNSUPPORTED;
    long long grid = (long long)per_cu * h->num_cus;
    if (grid > p.nunits) grid = p.nunits;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(p.block), lds, stream, a, vals, b, x);
    BAAS_HIP_TRY(hipGetLastError());
    return BAAS_OK;
}

template <class T, bool UP>
static int launch_df(baas_handle *h, const Plan &p, const DfArgs &a, const T *vals, const T *b, T *x, size_t lds,
                     hipStream_t st) {
    switch (p.keymode) {
        case 0: return launch_df_t<T, 0, UP>(h, p, a, vals, b, x, lds, st);
        case 1: return launch_df_t<T, 1, UP>(h, p, a, vals, b, x, lds, st);
        default: return launch_df_t<T, 2, UP>(h, p, a, vals, b, x, lds, st);
    }
}

template <class T>
int sweep(baas_handle *h, baas_schedule *s, int uplo, const T *vals, const T *b, T *x, int nrhs,
          hipStream_t stream, const int *skip) {
    if (!h || !s || !vals || !b || !x || nrhs < 1 || (uplo != 0 && uplo != 1)) return BAAS_INVALID_ARG;
    if ((const void *)b == (const void *)x) return BAAS_INVALID_ARG;
    const Plan &p = uplo ? s->bwd : s->fwd;
    if (nrhs == 1 && p.dataflow && p.block >= 128 && (size_t)p.max_slots * sizeof(T) <= lds_limit(h)) {
        DfArgs d;
        d.n = s->n;
        d.nunits = p.nunits;
        d.npf = p.block >= 512 ? 2 : 1;
        d.rowptr = uplo ? s->u_rowptr.p : s->l_rowptr.p;
        d.scol = p.scol.p;
        d.key_begin = p.key_begin.p;
        d.unit_ptr = p.unit_ptr.p;
        d.rows = p.rows.p;
        d.key = p.key.p;
        d.ext_ptr = p.ext_ptr.p;
        d.ext_cols = p.ext_cols.p;
        d.flags = s->flags.p;
        d.ctx = Ctx{h->counters, h->counters + 1, h->counters + 2};
        d.skip = skip;
        d.status = h->status;
        d.spin_limit = h->spin_limit;
        size_t lds = (size_t)p.max_slots * sizeof(T);
        return uplo ? launch_df<T, true>(h, p, d, vals, b, x, lds, stream)
                    : launch_df<T, false>(h, p, d, vals, b, x, lds, stream);
    }
    SweepArgs a;
    a.n = s->n;
    a.nunits = p.nunits;
    a.nrhs = nrhs;
    a.rowptr = uplo ? s->u_rowptr.p : s->l_rowptr.p;
    a.colidx = uplo ? s->u_colidx.p : s->l_colidx.p;
    a.key_begin = p.key_begin.p;
    a.sub_begin = p.sub_begin.p;
    a.sub_ptr = p.sub_ptr.p;
    a.rows = p.rows.p;
    a.key = p.key.p;
    a.flags = s->flags.p;
    a.ctx = Ctx{h->counters, h->counters + 1, h->counters + 2};
    a.skip = skip;
    a.status = h->status;
    a.spin_limit = h->spin_limit;
    size_t lds = (size_t)p.max_unit_keys * nrhs * sizeof(T);
    bool use_lds = lds <= lds_limit(h);
    if (uplo)
        return use_lds ? launch_sweep_k<T, true, true>(h, p, a, vals, b, x, lds, stream)
                       : launch_sweep_k<T, true, false>(h, p, a, vals, b, x, 0, stream);
    return use_lds ? launch_sweep_k<T, false, true>(h, p, a, vals, b, x, lds, stream)
                   : launch_sweep_k<T, false, false>(h, p, a, vals, b, x, 0, stream);
}

template int sweep<double>(baas_handle *, baas_schedule *, int, const double *, const double *, double *, int,
                           hipStream_t, const int *);
template int sweep<float>(baas_handle *, baas_schedule *, int, const float *, const float *, float *, int,
                          hipStream_t, const int *);

}  // namespace baas

using namespace baas;

extern "C" int baas_sptrsv_f64(baas_handle *h, baas_schedule *s, int uplo, const double *vals, const double *b,
                               double *x, int nrhs) {
    return sweep<double>(h, s, uplo, vals, b, x, nrhs, h ? h->stream : nullptr, nullptr);
}

extern "C" int baas_sptrsv_f32(baas_handle *h, baas_schedule *s, int uplo, const float *vals, const float *b,
                               float *x, int nrhs) {
    return sweep<float>(h, s, uplo, vals, b, x, nrhs, h ? h->stream : nullptr, nullptr);
}

extern "C" int baas_sptrsv_levelsync_f64(baas_handle *h, baas_schedule *s, const double *vals, const double *b,
                                         double *x) {
    if (!h || !s || !vals || !b || !x) return BAAS_INVALID_ARG;
    const auto &lp = s->row_level_ptr;
    for (int l = 0; l + 1 < (int)lp.size(); l++) {
        int cnt = lp[l + 1] - lp[l];
        hipLaunchKernelGGL(k_level_rows<double>, dim3(nblk(cnt)), dim3(256), 0, h->stream, s->row_perm_d.p, lp[l],
                           lp[l + 1], s->l_rowptr.p, s->l_colidx.p, vals, b, x);
    }
    BAAS_HIP_TRY(hipGetLastError());
    return BAAS_OK;
}

extern "C" int baas_transpose_values_f64(baas_handle *h, baas_schedule *s, const double *lv, double *uv) {
    if (!h || !s || !lv || !uv) return BAAS_INVALID_ARG;
    int nnz = s->nnz_lower;
    hipLaunchKernelGGL((k_gather<double, double>), dim3(nblk(nnz)), dim3(256), 0, h->stream, nnz, s->u_src.p, lv, uv);
    BAAS_HIP_TRY(hipGetLastError());
    return BAAS_OK;
}

extern "C" int baas_cast_f64_f32(baas_handle *h, int n, const double *in, float *out) {
    if (!h || n < 0 || (n && (!in || !out))) return BAAS_INVALID_ARG;
    if (n) hipLaunchKernelGGL((k_cast<float, double>), dim3(nblk(n)), dim3(256), 0, h->stream, n, in, out);
    BAAS_HIP_TRY(hipGetLastError());
    return BAAS_OK;
}

extern "C" int baas_cast_f32_f64(baas_handle *h, int n, const float *in, double *out) {
    if (!h || n < 0 || (n && (!in || !out))) return BAAS_INVALID_ARG;
    if (n) hipLaunchKernelGGL((k_cast<double, float>), dim3(nblk(n)), dim3(256), 0, h->stream, n, in, out);
    BAAS_HIP_TRY(hipGetLastError());
    return BAAS_OK;
}

extern "C" int baas_sptrsv_grad_vals_f64(baas_handle *h, baas_schedule *s, const double *lam, const double *x,
                                         double *dv, int nrhs) {
    if (!h || !s || !lam || !x || !dv || nrhs < 1) return BAAS_INVALID_ARG;
    int nnz = s->nnz_lower;
    hipLaunchKernelGGL(k_grad_vals, dim3(nblk(nnz)), dim3(256), 0, h->stream, nnz, nrhs, s->l_row_of.p,
                       s->l_colidx.p, lam, x, dv);
    BAAS_HIP_TRY(hipGetLastError());
    return BAAS_OK;
}

extern "C" int baas_csr_outer_grad_f64(baas_handle *h, int m, const int *rp, const int *ci, const double *lam,
                                       const double *x, double *dv, int nrhs) {
    if (!h || m <= 0 || !rp || !ci || !lam || !x || !dv || nrhs < 1) return BAAS_INVALID_ARG;
    hipLaunchKernelGGL(k_csr_outer_grad, dim3(nblk(m)), dim3(256), 0, h->stream, m, nrhs, rp, ci, lam, x, dv);
    BAAS_HIP_TRY(hipGetLastError());
    return BAAS_OK;
}

extern "C" int baas_plan_dataflow(baas_schedule *s, int enable) {
    if (!s) return BAAS_INVALID_ARG;
    s->fwd.dataflow = s->bwd.dataflow = enable ? 1 : 0;
    return BAAS_OK;
}
