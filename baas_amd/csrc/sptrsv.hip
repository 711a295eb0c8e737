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
    const int *udep_ptr, *udeps;  // unit-level dependencies (UNITDEP kernels)
    int *uflags;
};

template <class T, int KEYMODE, bool UPPER, bool LDS, bool UNITDEP>
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
        if (UNITDEP) {
            // one poll per producer unit, then every external parent is final
            for (int d = a.udep_ptr[u] + threadIdx.x; d < a.udep_ptr[u + 1]; d += blockDim.x)
                wait_ready(a.uflags, a.udeps[d], epoch, a.spin_limit, a.status);
            __syncthreads();
        }
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
                if (!UNITDEP)
                    for (int j = ob; j < oe; j++) {
                        const int c = a.colidx[j];
                        if (key_of<KEYMODE>(c, a.n, a.key) < kb)
                            wait_ready(a.flags, c, epoch, a.spin_limit, a.status);
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
                if (!UNITDEP && nrhs == 1 && rw < 0) {
                    wait_stores();
                    st_agent(&a.flags[row], epoch);
                }
            }
            if (UNITDEP) {
                wait_stores();
                __syncthreads();
            } else if (nrhs > 1) {
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
        if (UNITDEP && threadIdx.x == 0) st_agent(&a.uflags[u], epoch);
    }
    retire_block(a.ctx, epoch);
}

// ---------------------------------------------------------------------------
// Dataflow sweep.  One lane owns one row; the unit's solved values live in
// LDS slots pre-filled with a signalling-NaN sentinel, so a lane reads its
// parent's slot until the sentinel is gone -- no sub-level barriers, every
// row waits exactly for its own parents.  Each lane retries in one loop that
// consumes every ready entry in ascending (descending for U) column order,
// so lanes of one wavefront that depend on each other still make progress.
//
// Rows read by later units publish their value itself: one atomic 64-bit
// (32-bit for FP32) store into a plan-owned array of two halves pre-filled
// with the sentinel.  A sweep publishes into the half selected by the plan's
// sweep counter and re-arms the same rows in the other half for the next
// sweep, so a consumer needs one load per parent (the value is the flag) and
// a producer never waits for its stores.  Dedicated prefetch waves stage
// those parents into LDS while the compute waves work, and every compute
// lane loads its next row's data before it starts waiting on the current one.
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
    int *pcount;  // the plan's sweep counter (parity selects the publication half)
    Ctx ctx;
    const int *skip;
    DevStatus *status;
    long long spin_limit;
};

constexpr int DF_CHUNK = 8;

// one row's data, loaded ahead of its turn
template <class T>
struct DfRow {
    int rw, row, dg, left, jn, nc;
    T bv;
    int cs[DF_CHUNK];
    T cv[DF_CHUNK];
};

template <class T, bool UPPER>
__device__ __forceinline__ void df_load_row(const DfArgs &a, const T *__restrict__ vals, const T *__restrict__ b,
                                            int q, int q1, DfRow<T> &R) {
    R.nc = 0;
    R.left = 0;
    R.rw = -1;
    if (q >= q1) return;
    R.rw = a.rows[q];
    R.row = R.rw & 0x7fffffff;
    const int rb = a.rowptr[R.row], re = a.rowptr[R.row + 1];
    R.dg = UPPER ? rb : re - 1;
    R.left = re - rb - 1;
    R.jn = UPPER ? re - 1 : rb;
    R.bv = b[R.row];
    R.nc = R.left < DF_CHUNK ? R.left : DF_CHUNK;
#pragma unroll
    for (int t = 0; t < DF_CHUNK; t++) {
        if (t < R.nc) {
            const int j = UPPER ? R.jn - t : R.jn + t;
            R.cs[t] = a.scol[j];
            R.cv[t] = vals[j];
        }
    }
    R.jn = UPPER ? R.jn - R.nc : R.jn + R.nc;
    R.left -= R.nc;
}

template <class T, int KEYMODE, bool UPPER>
__global__ __launch_bounds__(1024) void k_sweep_df(DfArgs a, const T *__restrict__ vals, const T *__restrict__ b,
                                                  T *x, T *xp) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *sl = reinterpret_cast<T *>(smem_raw);
    __shared__ int s_unit;
    if (a.skip && ld_agent(a.skip)) return;
    const int epoch = read_epoch(a.ctx);
    const int par = ld_agent(a.pcount) & 1;
    T *pub = xp + (size_t)par * a.n;
    T *nxt = xp + (size_t)(par ^ 1) * a.n;
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
            // prefetch waves: stage parents of other units in order of first use
            for (int e = e0 + wv * 64 + lane; e < e1; e += npf * 64) {
                const int c = a.ext_cols[e];
                T v = ld_agent(&pub[c]);
                long long spins = 0;
                while (Sentinel<T>::is(v)) {
                    if (++spins > a.spin_limit) {
                        raise_status(a.status, BAAS_TIMEOUT, c);
                        v = T(0);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    v = ld_agent(&pub[c]);
                }
                st_lds(&sl[nk + (e - e0)], v);
            }
        } else {
            const int q0 = a.unit_ptr[u], q1 = a.unit_ptr[u + 1];
            const int C = nw - npf, cw = wv - npf;
            DfRow<T> cur, nxr;
            df_load_row<T, UPPER>(a, vals, b, q0 + cw * 64 + lane, q1, cur);
            for (int base = q0 + cw * 64; base < q1; base += C * 64) {
                df_load_row<T, UPPER>(a, vals, b, base + C * 64 + lane, q1, nxr);
                bool pending = base + lane < q1;
                T s = T(0);
                int pc = 0;
                long long spins = 0;
                while (pending) {
                    if (pc == cur.nc && cur.left > 0) {  // rows longer than one chunk
                        cur.nc = cur.left < DF_CHUNK ? cur.left : DF_CHUNK;
#pragma unroll
                        for (int t = 0; t < DF_CHUNK; t++) {
                            if (t < cur.nc) {
                                const int j = UPPER ? cur.jn - t : cur.jn + t;
                                cur.cs[t] = a.scol[j];
                                cur.cv[t] = vals[j];
                            }
                        }
                        cur.jn = UPPER ? cur.jn - cur.nc : cur.jn + cur.nc;
                        cur.left -= cur.nc;
                        pc = 0;
                    }
                    if (pc < cur.nc) {
#pragma unroll
                        for (int t = 0; t < DF_CHUNK; t++) {
                            if (t == pc && t < cur.nc) {
                                const T v = ld_lds(&sl[cur.cs[t]]);
                                if (!Sentinel<T>::is(v)) {
                                    s = UPPER ? fma(v, cur.cv[t], s) : fma(cur.cv[t], v, s);
                                    pc++;
                                }
                            }
                        }
                        if (pc < cur.nc && ++spins > a.spin_limit) {
                            raise_status(a.status, BAAS_TIMEOUT, cur.row);
                            pc = cur.nc;  // give up on this row's remaining parents
                            cur.left = 0;
                        }
                    } else {
                        const T xr = (cur.bv - s) / vals[cur.dg];
                        st_lds(&sl[key_of<KEYMODE>(cur.row, a.n, a.key) - kb], xr);
                        x[cur.row] = xr;
                        if (cur.rw < 0) {
                            st_agent(&pub[cur.row], xr);
                            nxt[cur.row] = Sentinel<T>::value();
                        }
                        pending = false;
                    }
                }
                cur = nxr;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        __threadfence();
        const int d = atomicAdd(a.ctx.done, 1);
        if (d == (int)gridDim.x - 1) {
            atomicExch(a.ctx.ticket, 0);
            atomicExch(a.ctx.epoch, epoch == 0x7fffffff ? 1 : epoch + 1);
            atomicAdd(a.pcount, 1);
            __threadfence();
            atomicExch(a.ctx.done, 0);
        }
    }
}

template <class T>
__global__ void k_fill_sentinel(size_t n, T *p) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = Sentinel<T>::value();
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
    auto kern = p.unitdep ? k_sweep<T, KM, UP, LDS, true> : k_sweep<T, KM, UP, LDS, false>;
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
static int launch_df_t(baas_handle *h, const Plan &p, const DfArgs &a, const T *vals, const T *b, T *x, T *xp,
                       size_t lds, hipStream_t stream) {
    auto kern = k_sweep_df<T, KM, UP>;
    int per_cu = 0;
    BAAS_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, p.block, lds));
    if (per_cu <= 0) return BAAS_UNSUPPORTED;
    long long grid = (long long)per_cu * h->num_cus;
    if (grid > p.nunits) grid = p.nunits;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(p.block), lds, stream, a, vals, b, x, xp);
    BAAS_HIP_TRY(hipGetLastError());
    return BAAS_OK;
}

template <class T, bool UP>
static int launch_df(baas_handle *h, const Plan &p, const DfArgs &a, const T *vals, const T *b, T *x, T *xp,
                     size_t lds, hipStream_t st) {
    switch (p.keymode) {
        case 0: return launch_df_t<T, 0, UP>(h, p, a, vals, b, x, xp, lds, st);
        case 1: return launch_df_t<T, 1, UP>(h, p, a, vals, b, x, xp, lds, st);
        default: return launch_df_t<T, 2, UP>(h, p, a, vals, b, x, xp, lds, st);
    }
}

template <class T>
int sweep(baas_handle *h, baas_schedule *s, int uplo, const T *vals, const T *b, T *x, int nrhs,
          hipStream_t stream, const int *skip) {
    if (!h || !s || !vals || !b || !x || nrhs < 1 || (uplo != 0 && uplo != 1)) return BAAS_INVALID_ARG;
    if ((const void *)b == (const void *)x) return BAAS_INVALID_ARG;
    Plan &p = uplo ? s->bwd : s->fwd;
    if (nrhs == 1 && p.dataflow && p.block >= 128 && (size_t)p.max_slots * sizeof(T) <= lds_limit(h)) {
        // publication halves, armed with this value type's sentinel
        if (p.xp_type != (int)sizeof(T)) {
            const size_t cnt = 2 * (size_t)s->n;
            BAAS_TRY(p.xp_raw.alloc(cnt * sizeof(T)));
            hipLaunchKernelGGL(k_fill_sentinel<T>, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, stream, cnt,
                               reinterpret_cast<T *>(p.xp_raw.p));
            BAAS_HIP_TRY(hipGetLastError());
            p.xp_type = (int)sizeof(T);
        }
        DfArgs d;
        d.pcount = p.pcount.p;
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
        d.ctx = Ctx{h->counters, h->counters + 1, h->counters + 2};
        d.skip = skip;
        d.status = h->status;
        d.spin_limit = h->spin_limit;
        size_t lds = (size_t)p.max_slots * sizeof(T);
        T *xp = reinterpret_cast<T *>(p.xp_raw.p);
        return uplo ? launch_df<T, true>(h, p, d, vals, b, x, xp, lds, stream)
                    : launch_df<T, false>(h, p, d, vals, b, x, xp, lds, stream);
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
    a.udep_ptr = p.udep_ptr.p;
    a.udeps = p.udeps.p;
    a.uflags = p.uflags.p;
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
