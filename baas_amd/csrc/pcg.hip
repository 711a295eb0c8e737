// pcg.hip -- IC0-preconditioned CG on gfx950 (pcg_opt / pcg_ref,
// test/iccg_level_merge.cpp:380-533 and :698-807).
//
// Order of operations per iteration follows the reference loop:
//   Ap = A p; alpha = rz / (p.Ap); x += alpha p; r -= alpha Ap;
//   rel = ||r|| / ||r0||  (stop at the first rel < tol);
//   z = U^-1 L^-1 r; rz' = r.z; beta = rz'/rz; p = z + beta p.
// All scalars stay on the device.  precision 1 stores the IC0 factor in FP32
// and runs both triangular solves in FP32 while the residual, dots, x/r/p
// updates and the convergence test stay FP64 (FP64 CG acts as refinement).
// `chunk` iterations are captured once in a hipGraph and replayed; after the
// stop condition every remaining kernel of the chunk returns at once.
#include <chrono>
#include <cmath>
#include <cstring>

#include "blas.h"
#include "device.h"
#include "internal.h"

namespace baas {

// state words (int) and scalars (double)
enum { ST_DONE = 0, ST_ITER = 1, ST_FIRST = 2, ST_MAXIT = 3 };
enum { SC_NORM0 = 0, SC_TOL = 1, SC_REL = 2, SC_PAP = 3, SC_RZ0 = 4, SC_RZ1 = 5, SC_RR = 6 };

__global__ void k_reduce(int np, const double *__restrict__ part, double *out, const int *skip) {
    __shared__ double red[4];
    if (skip && ld_agent(skip)) return;
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += blockDim.x) s += part[i];
    s = block_sum(s, red);
    if (threadIdx.x == 0) *out = s;
}

// Sum of np partials in the order k_reduce uses (256 threads, strided, then
// the block sum), computed by every workgroup of a consumer kernel so the
// value needs no launch of its own; every workgroup gets the same bits.
__device__ __forceinline__ double sum_partials(int np, const double *part, double *red, double *bcast) {
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += blockDim.x) s += ld_agent(&part[i]);
    s = block_sum(s, red);
    if (threadIdx.x == 0) *bcast = s;
    __syncthreads();
    return *bcast;
}

// x += alpha p; r -= alpha Ap; partials of r.r; r32 = (float) r
// (alpha = rz / pAp, pAp summed here from the SpMV's partials)
__global__ __launch_bounds__(256) void k_update_xr(int n, double *x, double *r, const double *__restrict__ p,
                                                   const double *__restrict__ Ap, double *sc, int cur,
                                                   int np_pap, const double *part_pap, double *part_rr,
                                                   float *r32, const int *st) {
    __shared__ double red[4];
    __shared__ double bc;
    if (ld_agent(&st[ST_DONE])) return;
    const double pap = sum_partials(np_pap, part_pap, red, &bc);
    if (blockIdx.x == 0 && threadIdx.x == 0) sc[SC_PAP] = pap;
    const double alpha = sc[SC_RZ0 + cur] / pap;
    double acc = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        x[i] = fma(alpha, p[i], x[i]);
        const double ri = fma(-alpha, Ap[i], r[i]);
        r[i] = ri;
        if (r32) r32[i] = (float)ri;
        acc = fma(ri, ri, acc);
    }
    acc = block_sum(acc, red);
    if (threadIdx.x == 0) part_rr[blockIdx.x] = acc;
}

// rel = ||r|| / ||r0||; stop at the first rel < tol or at maxiter
__global__ void k_conv(int np, const double *__restrict__ part_rr, double *sc, int *st) {
    __shared__ double red[4];
    if (ld_agent(&st[ST_DONE])) return;
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += blockDim.x) s += part_rr[i];
    s = block_sum(s, red);
    if (threadIdx.x == 0) {
        const int it = st[ST_ITER] + 1;
        st[ST_ITER] = it;
        const double rel = sqrt(s) / sc[SC_NORM0];
        sc[SC_REL] = rel;
        if (rel < sc[SC_TOL]) {
            st[ST_FIRST] = it;
            st_agent(&st[ST_DONE], 1);
        } else if (it >= st[ST_MAXIT]) {
            st_agent(&st[ST_DONE], 2);
        }
    }
}

// z64 = (double) z32 (mixed) and partials of r.z
__global__ __launch_bounds__(256) void k_rz(int n, const double *__restrict__ r, double *z,
                                            const float *__restrict__ z32, double *part, const int *st) {
    __shared__ double red[4];
    if (st && ld_agent(&st[ST_DONE])) return;
    double acc = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        double zi;
        if (z32) {
            zi = (double)z32[i];
            z[i] = zi;
        } else {
            zi = z[i];
        }
        acc = fma(r[i], zi, acc);
    }
    acc = block_sum(acc, red);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// p = z + beta p, beta = rz_new / rz_old (rz_new summed here from the
// partials of k_rz; workgroup 0 keeps it for the next iteration)
__global__ __launch_bounds__(256) void k_update_p(int n, double *p, const double *__restrict__ z, double *sc,
                                                  int cur, int np_rz, const double *part_rz, const int *st) {
    __shared__ double red[4];
    __shared__ double bc;
    if (st && ld_agent(&st[ST_DONE])) return;
    const double rz_new = sum_partials(np_rz, part_rz, red, &bc);
    if (blockIdx.x == 0 && threadIdx.x == 0) sc[SC_RZ0 + (cur ^ 1)] = rz_new;
    const double beta = rz_new / sc[SC_RZ0 + cur];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        p[i] = fma(beta, p[i], z[i]);
}

__global__ void k_copy(int n, const double *__restrict__ a, double *__restrict__ b) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

__global__ void k_cast32(int n, const double *__restrict__ a, float *__restrict__ b) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

__global__ void k_init_state(int *st, double *sc, double tol, int maxiter, const double *norm0_src) {
    st[ST_DONE] = 0;
    st[ST_ITER] = 0;
    st[ST_FIRST] = -1;
    st[ST_MAXIT] = maxiter;
    sc[SC_TOL] = tol;
    sc[SC_NORM0] = *norm0_src;
    if (!(sc[SC_NORM0] > 0.0)) st[ST_DONE] = 1, st[ST_FIRST] = 0, sc[SC_REL] = 0.0;
}

static inline int nblk(long long n, int b = 256) { return (int)((n + b - 1) / b); }

}  // namespace baas

using namespace baas;

struct baas_pcg {
    baas_handle *h = nullptr;
    baas_schedule *s = nullptr;
    const double *a_vals = nullptr;
    int precision = 0, n = 0, chunk = 8;
    int nb_spmv = 0, nb_vec = 0;
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    DBuf<double> lu, lv, uv, x, r, z, y, p, Ap, part_pap, part_rr, part_rz, sc, norm0;
    DBuf<float> lv32, uv32, r32, y32, z32;
    DBuf<int> state;
    int *h_state = nullptr;  // pinned
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool factored = false;
    int cap_gen = -1;  // schedule plan generation the graph was captured with
    double t_ic = 0, t_iter = 0;
};

namespace baas {

static int precond(baas_pcg *P, const int *skip) {
    baas_schedule *s = P->s;
    if (P->precision == 1) {
        BAAS_TRY(sweep<float>(P->h, s, 0, P->lv32.p, P->r32.p, P->y32.p, 1, P->st, skip));
        BAAS_TRY(sweep<float>(P->h, s, 1, P->uv32.p, P->y32.p, P->z32.p, 1, P->st, skip));
    } else {
        BAAS_TRY(sweep<double>(P->h, s, 0, P->lv.p, P->r.p, P->y.p, 1, P->st, skip));
        BAAS_TRY(sweep<double>(P->h, s, 1, P->uv.p, P->y.p, P->z.p, 1, P->st, skip));
    }
    return BAAS_OK;
}

// one CG iteration (parity `cur` selects which rz slot is the old one)
static int iteration(baas_pcg *P, int cur) {
    const int n = P->n;
    baas_schedule *s = P->s;
    const int *done = P->state.p + ST_DONE;
    BAAS_TRY(spmv_dev(n, s->a_rowptr.p, s->a_colidx.p, P->a_vals, P->p.p, P->Ap.p, 1.0, 0.0, P->p.p, P->part_pap.p,
                      done, P->st));
    hipLaunchKernelGGL(k_update_xr, dim3(P->nb_vec), dim3(256), 0, P->st, n, P->x.p, P->r.p, P->p.p, P->Ap.p,
                       P->sc.p, cur, P->nb_spmv, P->part_pap.p, P->part_rr.p,
                       P->precision == 1 ? P->r32.p : nullptr, P->state.p);
    hipLaunchKernelGGL(k_conv, dim3(1), dim3(256), 0, P->st, P->nb_vec, P->part_rr.p, P->sc.p, P->state.p);
    BAAS_TRY(precond(P, done));
    hipLaunchKernelGGL(k_rz, dim3(P->nb_vec), dim3(256), 0, P->st, n, P->r.p, P->z.p,
                       P->precision == 1 ? P->z32.p : nullptr, P->part_rz.p, P->state.p);
    hipLaunchKernelGGL(k_update_p, dim3(P->nb_vec), dim3(256), 0, P->st, n, P->p.p, P->z.p, P->sc.p, cur,
                       P->nb_vec, P->part_rz.p, P->state.p);
    BAAS_HIP_TRY(hipGetLastError());
    return BAAS_OK;
}

static int capture(baas_pcg *P) {
    if (P->exec && P->cap_gen == P->s->plan_gen) return BAAS_OK;
    // the graph holds the plans' kernels and buffers: re-capture after any
    // change of plan
    if (P->exec) (void)hipGraphExecDestroy(P->exec);
    if (P->graph) (void)hipGraphDestroy(P->graph);
    P->exec = nullptr;
    P->graph = nullptr;
    P->cap_gen = P->s->plan_gen;
    BAAS_HIP_TRY(hipStreamBeginCapture(P->st, hipStreamCaptureModeThreadLocal));
    int rc = BAAS_OK;
    for (int t = 0; t < P->chunk && rc == BAAS_OK; t++) rc = iteration(P, t & 1);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(P->st, &g);
    if (rc != BAAS_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return BAAS_HIP_ERROR;
    P->graph = g;
    BAAS_HIP_TRY(hipGraphInstantiate(&P->exec, g, nullptr, nullptr, 0));
    return BAAS_OK;
}

static int factorize(baas_pcg *P) {
    baas_handle *h = P->h;
    hipStream_t keep = h->stream;
    h->stream = P->st;
    int rc = baas_ic0_f64(h, P->s, P->a_vals, P->lu.p);
    h->stream = keep;
    BAAS_TRY(rc);
    BAAS_TRY(ic0_split_dev(h, P->s, P->lu.p, P->lv.p, P->uv.p, P->st));
    if (P->precision == 1) BAAS_TRY(ic0_split_dev32(h, P->s, P->lu.p, P->lv32.p, P->uv32.p, P->st));
    P->factored = true;
    return BAAS_OK;
}

}  // namespace baas

extern "C" int baas_pcg_create(baas_handle *h, baas_schedule *s, const double *a_vals, int precision,
                               baas_pcg **out) {
    if (!h || !s || !a_vals || !out || (precision != 0 && precision != 1)) return BAAS_INVALID_ARG;
    if (s->kind != BAAS_KIND_IC0) return BAAS_INVALID_ARG;
    *out = nullptr;
    baas_pcg *P = new baas_pcg();
    P->h = h;
    P->s = s;
    P->a_vals = a_vals;
    P->precision = precision;
    const int n = s->n, nl = s->nnz_lower;
    P->n = n;
    P->nb_spmv = spmv_blocks(n);
    P->nb_vec = nblk(n);
    if (P->nb_vec > 1024) P->nb_vec = 1024;
    int rc = BAAS_OK;
    auto A = [&](int r) {
        if (rc == BAAS_OK) rc = r;
    };
    A(P->lu.alloc(s->nnz_full));
    A(P->lv.alloc(nl));
    A(P->uv.alloc(nl));
    A(P->x.alloc(n));
    A(P->r.alloc(n));
    A(P->z.alloc(n));
    A(P->y.alloc(n));
    A(P->p.alloc(n));
    A(P->Ap.alloc(n));
    A(P->part_pap.alloc(P->nb_spmv));
    A(P->part_rr.alloc(P->nb_vec));
    A(P->part_rz.alloc(P->nb_vec));
    A(P->sc.alloc(16));
    A(P->norm0.alloc(1));
    A(P->state.alloc(8));
    if (precision == 1) {
        A(P->lv32.alloc(nl));
        A(P->uv32.alloc(nl));
        A(P->r32.alloc(n));
        A(P->y32.alloc(n));
        A(P->z32.alloc(n));
    }
    if (rc == BAAS_OK && hipStreamCreateWithFlags(&P->st, hipStreamNonBlocking) != hipSuccess) rc = BAAS_HIP_ERROR;
    if (rc == BAAS_OK && hipEventCreateWithFlags(&P->ev, hipEventDisableTiming) != hipSuccess) rc = BAAS_HIP_ERROR;
    if (rc == BAAS_OK && hipHostMalloc((void **)&P->h_state, 8 * sizeof(int), hipHostMallocDefault) != hipSuccess)
        rc = BAAS_HIP_ERROR;
    if (rc != BAAS_OK) {
        baas_pcg_destroy(P);
        return rc;
    }
    *out = P;
    return BAAS_OK;
}

extern "C" int baas_pcg_destroy(baas_pcg *P) {
    if (!P) return BAAS_INVALID_ARG;
    if (P->st) (void)hipStreamSynchronize(P->st);
    if (P->exec) (void)hipGraphExecDestroy(P->exec);
    if (P->graph) (void)hipGraphDestroy(P->graph);
    if (P->ev) (void)hipEventDestroy(P->ev);
    if (P->st) (void)hipStreamDestroy(P->st);
    if (P->h_state) (void)hipHostFree(P->h_state);
    delete P;
    return BAAS_OK;
}

extern "C" int baas_pcg_solve(baas_pcg *P, const double *b, double *x, double tol, int maxiter, int refactor,
                              int *iters, double *relres) {
    if (!P || !b || !x || !(tol > 0) || maxiter < 1) return BAAS_INVALID_ARG;
    baas_handle *h = P->h;
    const int n = P->n;
    baas_schedule *s = P->s;
    // order after the caller's stream
    BAAS_HIP_TRY(hipEventRecord(P->ev, h->stream));
    BAAS_HIP_TRY(hipStreamWaitEvent(P->st, P->ev, 0));
    auto t0 = std::chrono::steady_clock::now();
    if (!P->factored || refactor) BAAS_TRY(factorize(P));
    BAAS_HIP_TRY(hipStreamSynchronize(P->st));
    auto t1 = std::chrono::steady_clock::now();
    // r = b - A x ; norm0 ; z = M^-1 r ; p = z ; rz
    hipLaunchKernelGGL(k_copy, dim3(nblk(n)), dim3(256), 0, P->st, n, x, P->x.p);
    hipLaunchKernelGGL(k_copy, dim3(nblk(n)), dim3(256), 0, P->st, n, b, P->r.p);
    BAAS_TRY(spmv_dev(n, s->a_rowptr.p, s->a_colidx.p, P->a_vals, P->x.p, P->r.p, -1.0, 1.0, nullptr, nullptr,
                      nullptr, P->st));
    BAAS_TRY(dot_dev(h, n, P->r.p, P->r.p, P->norm0.p, 1, P->st));
    hipLaunchKernelGGL(k_init_state, dim3(1), dim3(1), 0, P->st, P->state.p, P->sc.p, tol, maxiter, P->norm0.p);
    if (P->precision == 1) hipLaunchKernelGGL(k_cast32, dim3(nblk(n)), dim3(256), 0, P->st, n, P->r.p, P->r32.p);
    BAAS_TRY(precond(P, nullptr));
    hipLaunchKernelGGL(k_rz, dim3(P->nb_vec), dim3(256), 0, P->st, n, P->r.p, P->z.p,
                       P->precision == 1 ? P->z32.p : nullptr, P->part_rz.p, nullptr);
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(256), 0, P->st, P->nb_vec, P->part_rz.p, P->sc.p + SC_RZ0, nullptr);
    hipLaunchKernelGGL(k_copy, dim3(nblk(n)), dim3(256), 0, P->st, n, P->z.p, P->p.p);
    BAAS_HIP_TRY(hipGetLastError());
    BAAS_TRY(capture(P));
    for (;;) {
        BAAS_HIP_TRY(hipGraphLaunch(P->exec, P->st));
        BAAS_HIP_TRY(hipMemcpyAsync(P->h_state, P->state.p, 4 * sizeof(int), hipMemcpyDeviceToHost, P->st));
        BAAS_HIP_TRY(hipStreamSynchronize(P->st));
        if (P->h_state[ST_DONE]) break;
    }
    auto t2 = std::chrono::steady_clock::now();
    double rel = 0;
    BAAS_HIP_TRY(hipMemcpy(&rel, P->sc.p + SC_REL, sizeof(double), hipMemcpyDeviceToHost));
    hipLaunchKernelGGL(k_copy, dim3(nblk(n)), dim3(256), 0, P->st, n, P->x.p, x);
    BAAS_HIP_TRY(hipEventRecord(P->ev, P->st));
    BAAS_HIP_TRY(hipStreamWaitEvent(h->stream, P->ev, 0));
    P->t_ic = std::chrono::duration<double>(t1 - t0).count();
    P->t_iter = std::chrono::duration<double>(t2 - t1).count();
    if (iters) *iters = P->h_state[ST_DONE] == 1 ? P->h_state[ST_FIRST] : P->h_state[ST_ITER];
    if (relres) *relres = rel;
    return P->h_state[ST_DONE] == 1 ? BAAS_OK : BAAS_NOT_CONVERGED;
}

extern "C" int baas_pcg_autotune(baas_pcg *P, int *chosen) {
    if (!P) return BAAS_INVALID_ARG;
    baas_handle *h = P->h;
    BAAS_HIP_TRY(hipEventRecord(P->ev, h->stream));
    BAAS_HIP_TRY(hipStreamWaitEvent(P->st, P->ev, 0));
    if (!P->factored) BAAS_TRY(factorize(P));
    hipStream_t keep = h->stream;
    h->stream = P->st;
    const int rc = P->precision == 1 ? autotune<float>(h, P->s, P->lv32.p, 0, 1, chosen)
                                     : autotune<double>(h, P->s, P->lv.p, 0, 1, chosen);
    h->stream = keep;
    return rc;
}

extern "C" int baas_pcg_times(const baas_pcg *P, double *ic_s, double *iter_s) {
    if (!P) return BAAS_INVALID_ARG;
    if (ic_s) *ic_s = P->t_ic;
    if (iter_s) *iter_s = P->t_iter;
    return BAAS_OK;
}
