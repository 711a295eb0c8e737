// batch.hip -- many independent SPD systems, each with several right-hand
// sides, solved by IC0-preconditioned CG in one set of launches.
//
// The systems are packed back to back as one block-diagonal CSR.  One
// schedule over it drives both the IC0 factorisation (FP64) and the FP32
// triangular sweeps of all systems and all right-hand sides at once (one
// persistent launch per sweep; rows are row-major N x nrhs).  CG runs in
// FP64 with one (alpha, beta, ||r||) per system and right-hand side; a
// column stops updating at its own first ||r||/||r0|| < tol, the whole batch
// stops when every column has.  `chunk` iterations are captured in a hipGraph.
// The execution plan's unit sizes are chosen per system-size bucket by the
// autotuner (timed FP32 sweeps of the actual batch).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "blas.h"
#include "device.h"
#include "internal.h"

namespace baas {
int build_plans_mode(baas_schedule *s, int mode, int rows_per_unit);

int make_plan_custom(baas_schedule *s, const std::vector<int> &unit_ptr, int block);  // plan.cpp

struct Chunk {
    int sys, r0, r1;
};

__device__ __forceinline__ bool stopped(const int *alldone) { return alldone && ld_agent(alldone); }

// y[i,r] = (A x)[i,r] for all right-hand sides (ascending FMA sum per row),
// the x gathers of up to CH entries issued at once
template <int CH>
__global__ void k_spmv_multi(int N, int k, const int *__restrict__ rp, const int *__restrict__ ci,
                             const double *__restrict__ va, const double *__restrict__ x, double *__restrict__ y,
                             const int *alldone) {
    if (stopped(alldone)) return;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)N * k) return;
    int i = (int)(t / k), r = (int)(t - (long long)i * k);
    const int b = rp[i], e = rp[i + 1];
    double s = 0.0;
    for (int j0 = b; j0 < e; j0 += CH) {
        double xv[CH], av[CH];
#pragma unroll
        for (int q = 0; q < CH; q++) {
            const int j = j0 + q < e ? j0 + q : e - 1;
            av[q] = va[j];
            xv[q] = x[(size_t)ci[j] * k + r];
        }
#pragma unroll
        for (int q = 0; q < CH; q++)
            if (j0 + q < e) s = fma(av[q], xv[q], s);
    }
    y[t] = s;
}

// the symmetric pattern's longest row has 2 * (longest lower row) + 1
// entries: 5 for a 5-point stencil, 7 for a 7-point one
static void launch_spmv_multi(int max_lower, hipStream_t st, int N, int k, const int *rp, const int *ci,
                              const double *va, const double *x, double *y, const int *alldone) {
    const int w = 2 * max_lower + 1;
    const long long M = (long long)N * k;
    const dim3 grid((unsigned)((M + 255) / 256)), block(256);
    if (w <= 5)
        hipLaunchKernelGGL(k_spmv_multi<5>, grid, block, 0, st, N, k, rp, ci, va, x, y, alldone);
    else if (w <= 7)
        hipLaunchKernelGGL(k_spmv_multi<7>, grid, block, 0, st, N, k, rp, ci, va, x, y, alldone);
    else
        hipLaunchKernelGGL(k_spmv_multi<8>, grid, block, 0, st, N, k, rp, ci, va, x, y, alldone);
}

// per chunk and right-hand side: sum over the chunk's rows of u*v (fixed order)
__global__ void k_chunk_dot(int k, const Chunk *__restrict__ ch, const double *__restrict__ u,
                            const double *__restrict__ v, double *part, const int *alldone) {
    extern __shared__ double red[];
    if (stopped(alldone)) return;
    const Chunk c = ch[blockIdx.x];
    const size_t e0 = (size_t)c.r0 * k, e1 = (size_t)c.r1 * k;
    // blockDim is a multiple of k: thread t always sees column t % k
    double acc = 0.0;
    for (size_t e = e0 + threadIdx.x; e < e1; e += blockDim.x) acc = fma(u[e], v[e], acc);
    red[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x < (unsigned)k) {
        double s = 0.0;
        for (int q = threadIdx.x; q < (int)blockDim.x; q += k) s += red[q];
        part[(size_t)blockIdx.x * k + threadIdx.x] = s;
    }
}

// per system and column: sum the chunk partials of the system in order
__global__ void k_sys_sum(int nsys, int k, const int *__restrict__ sys_chunk, const double *__restrict__ part,
                          double *out, int do_sqrt, const int *alldone) {
    if (stopped(alldone)) return;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nsys * k) return;
    int s = t / k, r = t - s * k;
    double acc = 0.0;
    for (int c = sys_chunk[s]; c < sys_chunk[s + 1]; c++) acc += part[(size_t)c * k + r];
    out[t] = do_sqrt ? sqrt(acc) : acc;
}

// r = b - Ap (initial residual)
__global__ void k_bres(long long M, const double *__restrict__ b, const double *__restrict__ Ap, double *r,
                       float *r32) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= M) return;
    double v = b[t] - Ap[t];
    r[t] = v;
    r32[t] = (float)v;
}

// x += alpha p, r -= alpha Ap for columns still running; r32 = (float) r
__global__ void k_bupdate_xr(int N, int k, const int *__restrict__ sys_of, double *x, double *r,
                             const double *__restrict__ p, const double *__restrict__ Ap, const double *__restrict__ rz,
                             const double *__restrict__ pap, const int *__restrict__ coldone, float *r32,
                             const int *alldone) {
    if (stopped(alldone)) return;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)N * k) return;
    int i = (int)(t / k), c = sys_of[i] * k + (int)(t - (long long)i * k);
    double ri = r[t];
    if (!coldone[c]) {
        const double alpha = rz[c] / pap[c];
        x[t] = fma(alpha, p[t], x[t]);
        ri = fma(-alpha, Ap[t], ri);
        r[t] = ri;
    }
    r32[t] = (float)ri;
}

// convergence per column (rel = ||r|| / ||r0|| < tol), count of running columns
__global__ void k_bconv(int ncol, const double *__restrict__ rr, const double *__restrict__ norm0,
                        const double *__restrict__ tolp, int *coldone, int *coliter, int *ctl) {
    if (stopped(ctl)) return;
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncol || coldone[c]) return;
    const int it = ctl[1] + 1;
    const double rel = sqrt(rr[c]) / norm0[c];
    if (rel < *tolp) {
        coldone[c] = 1;
        coliter[c] = it;
    } else {
        atomicAdd(&ctl[2], 1);
    }
}

// ctl = {alldone, iteration, running, maxiter}
__global__ void k_bconv_finish(int *ctl) {
    if (stopped(ctl)) return;
    const int it = ctl[1] + 1;
    ctl[1] = it;
    if (ctl[2] == 0 || it >= ctl[3]) st_agent(&ctl[0], 1);
    ctl[2] = 0;
}

__global__ void k_cast_back(long long M, const float *__restrict__ z32, double *__restrict__ z, const int *alldone) {
    if (stopped(alldone)) return;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < M) z[t] = (double)z32[t];
}

// p = z + (rz_new / rz_old) p for running columns
__global__ void k_bupdate_p(int N, int k, const int *__restrict__ sys_of, double *p, const double *__restrict__ z,
                            const double *__restrict__ rz_new, const double *__restrict__ rz_old,
                            const int *__restrict__ coldone, const int *alldone) {
    if (stopped(alldone)) return;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)N * k) return;
    int i = (int)(t / k), c = sys_of[i] * k + (int)(t - (long long)i * k);
    if (coldone[c]) return;
    p[t] = fma(rz_new[c] / rz_old[c], p[t], z[t]);
}

__global__ void k_bcopy(long long M, const double *__restrict__ a, double *__restrict__ b) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < M) b[t] = a[t];
}

__global__ void k_binit(int ncol, int maxiter, int *coldone, int *coliter, int *ctl, const double *norm0) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < ncol) {
        coldone[c] = norm0[c] > 0.0 ? 0 : 1;
        coliter[c] = norm0[c] > 0.0 ? -1 : 0;
    }
    if (c == 0) {
        ctl[0] = 0;
        ctl[1] = 0;
        ctl[2] = 0;
        ctl[3] = maxiter;
    }
}

static inline int nblk(long long n, int b = 256) { return (int)((n + b - 1) / b); }

}  // namespace baas

using namespace baas;

struct baas_batch {
    baas_handle *h = nullptr;
    baas_schedule *s = nullptr;
    int nsys = 0, N = 0, k = 0, nchunks = 0, chunk_iters = 8;
    const double *a_vals = nullptr;
    const int *rowptr = nullptr, *colidx = nullptr;
    std::vector<int> sys_ptr;
    DBuf<int> sys_of, sys_chunk;
    DBuf<Chunk> chunks;
    DBuf<double> lu, x, r, z, p, Ap, part, rz0, rz1, pap, rr, norm0, tol;
    DBuf<float> lv32, uv32, r32, y32, z32;
    DBuf<int> coldone, coliter, ctl;
    int *h_ctl = nullptr;  // pinned
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    std::vector<int> tuned_rpu;
    bool tuned = false;
    int cap_gen = -1;  // schedule plan generation the graph was captured with
    double t_tune = 0, t_ic = 0, t_iter = 0;
};

namespace baas {

static int bucket_of(int n) { return n <= 4096 ? 0 : (n <= 16384 ? 1 : 2); }

// Unit boundaries never cross systems; each system is cut into units of its
// size bucket's rows-per-unit.
static std::vector<int> batch_units(const baas_batch *B, const std::vector<int> &rpu) {
    std::vector<int> up;
    for (int s = 0; s < B->nsys; s++) {
        int b = B->sys_ptr[s], e = B->sys_ptr[s + 1];
        int R = rpu[bucket_of(e - b)];
        for (int q = b; q < e; q += R) up.push_back(q);
    }
    up.push_back(B->N);
    return up;
}

static int batch_precond(baas_batch *B, const int *skip) {
    BAAS_TRY(sweep<float>(B->h, B->s, 0, B->lv32.p, B->r32.p, B->y32.p, B->k, B->st, skip));
    BAAS_TRY(sweep<float>(B->h, B->s, 1, B->uv32.p, B->y32.p, B->z32.p, B->k, B->st, skip));
    return BAAS_OK;
}

static float time_sweeps(baas_batch *B, int reps) {
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1e30f;
    float best = 1e30f;
    for (int t = 0; t < reps; t++) {
        (void)hipEventRecord(e0, B->st);
        if (batch_precond(B, nullptr) != BAAS_OK) break;
        (void)hipEventRecord(e1, B->st);
        (void)hipEventSynchronize(e1);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (t > 0) best = std::min(best, ms);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return best;
}

// Coordinate search over rows-per-unit, one size bucket at a time, timing
// the FP32 forward + backward sweeps of the whole batch with all RHS; then
// the level-set plan (all systems' rows of one level at a time), which is
// kept if faster.
static int autotune(baas_batch *B) {
    auto t0 = std::chrono::steady_clock::now();
    const int cand[] = {512, 1024, 2048, 4096, 8192};
    std::vector<int> rpu = {2048, 2048, 2048};
    const int block = B->s->opts.block_threads;
    BAAS_TRY(make_plan_custom(B->s, batch_units(B, rpu), block));
    float best = time_sweeps(B, 3);
    for (int bk = 0; bk < 3; bk++) {
        for (int c : cand) {
            if (c == rpu[bk]) continue;
            std::vector<int> trial = rpu;
            trial[bk] = c;
            BAAS_TRY(make_plan_custom(B->s, batch_units(B, trial), block));
            float t = time_sweeps(B, 3);
            if (t < best) best = t, rpu = trial;
        }
    }
    int lev_block = 0;
    for (int lb : {256, 512, 1024}) {
        B->s->opts.block_threads = lb;
        BAAS_TRY(build_plans_mode(B->s, BAAS_PLAN_LEVELS, 0));
        const float t_lev = time_sweeps(B, 3);
        if (t_lev < best) best = t_lev, lev_block = lb;
    }
    B->s->opts.block_threads = block;
    if (lev_block) {
        B->s->opts.block_threads = lev_block;
        BAAS_TRY(build_plans_mode(B->s, BAAS_PLAN_LEVELS, 0));
        B->s->opts.block_threads = block;
        B->tuned_rpu = {0, 0, lev_block};
        B->tuned = true;
        B->t_tune = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return BAAS_OK;
    }
    BAAS_TRY(make_plan_custom(B->s, batch_units(B, rpu), block));
    B->tuned_rpu = rpu;
    B->tuned = true;
    B->t_tune = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return BAAS_OK;
}

static int chunk_dot_sys(baas_batch *B, const double *u, const double *v, double *out, int do_sqrt,
                         const int *skip) {
    hipLaunchKernelGGL(k_chunk_dot, dim3(B->nchunks), dim3(256), 256 * sizeof(double), B->st, B->k, B->chunks.p, u, v,
                       B->part.p, skip);
    hipLaunchKernelGGL(k_sys_sum, dim3(nblk((long long)B->nsys * B->k)), dim3(256), 0, B->st, B->nsys, B->k,
                       B->sys_chunk.p, B->part.p, out, do_sqrt, skip);
    BAAS_HIP_TRY(hipGetLastError());
    return BAAS_OK;
}

static int batch_iteration(baas_batch *B, int cur) {
    const int N = B->N, k = B->k, ncol = B->nsys * B->k;
    const long long M = (long long)N * k;
    int *ctl = B->ctl.p;
    double *rz_old = cur ? B->rz1.p : B->rz0.p, *rz_new = cur ? B->rz0.p : B->rz1.p;
    launch_spmv_multi(B->s->fwd.max_deg, B->st, N, k, B->rowptr, B->colidx, B->a_vals, B->p.p, B->Ap.p, ctl);
    BAAS_TRY(chunk_dot_sys(B, B->p.p, B->Ap.p, B->pap.p, 0, ctl));
    hipLaunchKernelGGL(k_bupdate_xr, dim3(nblk(M)), dim3(256), 0, B->st, N, k, B->sys_of.p, B->x.p, B->r.p, B->p.p,
                       B->Ap.p, rz_old, B->pap.p, B->coldone.p, B->r32.p, ctl);
    BAAS_TRY(chunk_dot_sys(B, B->r.p, B->r.p, B->rr.p, 0, ctl));
    hipLaunchKernelGGL(k_bconv, dim3(nblk(ncol)), dim3(256), 0, B->st, ncol, B->rr.p, B->norm0.p, B->tol.p,
                       B->coldone.p, B->coliter.p, ctl);
    hipLaunchKernelGGL(k_bconv_finish, dim3(1), dim3(1), 0, B->st, ctl);
    BAAS_TRY(batch_precond(B, ctl));
    hipLaunchKernelGGL(k_cast_back, dim3(nblk(M)), dim3(256), 0, B->st, M, B->z32.p, B->z.p, ctl);
    BAAS_TRY(chunk_dot_sys(B, B->r.p, B->z.p, rz_new, 0, ctl));
    hipLaunchKernelGGL(k_bupdate_p, dim3(nblk(M)), dim3(256), 0, B->st, N, k, B->sys_of.p, B->p.p, B->z.p, rz_new,
                       rz_old, B->coldone.p, ctl);
    BAAS_HIP_TRY(hipGetLastError());
    return BAAS_OK;
}

static int batch_capture(baas_batch *B) {
    if (B->exec && B->cap_gen == B->s->plan_gen) return BAAS_OK;
    if (B->exec) (void)hipGraphExecDestroy(B->exec);
    if (B->graph) (void)hipGraphDestroy(B->graph);
    B->exec = nullptr;
    B->graph = nullptr;
    B->cap_gen = B->s->plan_gen;
    BAAS_HIP_TRY(hipStreamBeginCapture(B->st, hipStreamCaptureModeThreadLocal));
    int rc = BAAS_OK;
    for (int t = 0; t < B->chunk_iters && rc == BAAS_OK; t++) rc = batch_iteration(B, t & 1);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(B->st, &g);
    if (rc != BAAS_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return BAAS_HIP_ERROR;
    B->graph = g;
    BAAS_HIP_TRY(hipGraphInstantiate(&B->exec, g, nullptr, nullptr, 0));
    return BAAS_OK;
}

}  // namespace baas

extern "C" int baas_batch_create(baas_handle *h, int nsys, const int *sys_ptr_host, int N, const int *rowptr,
                                 const int *colidx, const double *a_vals, int nrhs, const baas_opts *opts,
                                 baas_batch **out) {
    if (!h || nsys < 1 || !sys_ptr_host || N < 1 || !rowptr || !colidx || !a_vals || !out || nrhs < 1 ||
        nrhs > 256 || (256 % nrhs) != 0)
        return BAAS_INVALID_ARG;
    if (sys_ptr_host[0] != 0 || sys_ptr_host[nsys] != N) return BAAS_INVALID_ARG;
    for (int s = 0; s < nsys; s++)
        if (sys_ptr_host[s + 1] <= sys_ptr_host[s]) return BAAS_INVALID_ARG;
    *out = nullptr;
    baas_batch *B = new baas_batch();
    B->h = h;
    B->nsys = nsys;
    B->N = N;
    B->k = nrhs;
    B->a_vals = a_vals;
    B->rowptr = rowptr;
    B->colidx = colidx;
    B->sys_ptr.assign(sys_ptr_host, sys_ptr_host + nsys + 1);
    baas_opts o{};
    if (opts) o = *opts;
    o.plan_mode = BAAS_PLAN_RANGE;
    int rc = baas_analyze(h, N, rowptr, colidx, BAAS_KIND_IC0, &o, &B->s);
    // independent systems: the level-set plan runs every system's level at
    // once (the first factorisation uses it; the autotuner may keep it)
    if (rc == BAAS_OK) rc = build_plans_mode(B->s, BAAS_PLAN_LEVELS, 0);
    std::vector<int> sys_of(N), sys_chunk(nsys + 1);
    std::vector<Chunk> ch;
    for (int s = 0; s < nsys && rc == BAAS_OK; s++) {
        sys_chunk[s] = (int)ch.size();
        for (int q = sys_ptr_host[s]; q < sys_ptr_host[s + 1]; q++) sys_of[q] = s;
        for (int q = sys_ptr_host[s]; q < sys_ptr_host[s + 1]; q += 256)
            ch.push_back(Chunk{s, q, std::min(q + 256, sys_ptr_host[s + 1])});
    }
    sys_chunk[nsys] = (int)ch.size();
    B->nchunks = (int)ch.size();
    const long long M = (long long)N * nrhs;
    const int nl = rc == BAAS_OK ? B->s->nnz_lower : 0, ncol = nsys * nrhs;
    auto A = [&](int r) {
        if (rc == BAAS_OK) rc = r;
    };
    if (rc == BAAS_OK) {
        A(B->sys_of.alloc(N));
        A(B->sys_chunk.alloc(nsys + 1));
        A(B->chunks.alloc(ch.size()));
        A(B->lu.alloc(B->s->nnz_full));
        A(B->lv32.alloc(nl));
        A(B->uv32.alloc(nl));
        for (auto *d : {&B->x, &B->r, &B->z, &B->p, &B->Ap}) A(d->alloc(M));
        for (auto *f : {&B->r32, &B->y32, &B->z32}) A(f->alloc(M));
        A(B->part.alloc((size_t)B->nchunks * nrhs));
        for (auto *d : {&B->rz0, &B->rz1, &B->pap, &B->rr, &B->norm0}) A(d->alloc(ncol));
        A(B->tol.alloc(1));
        A(B->coldone.alloc(ncol));
        A(B->coliter.alloc(ncol));
        A(B->ctl.alloc(4));
    }
    if (rc == BAAS_OK && hipStreamCreateWithFlags(&B->st, hipStreamNonBlocking) != hipSuccess) rc = BAAS_HIP_ERROR;
    if (rc == BAAS_OK && hipEventCreateWithFlags(&B->ev, hipEventDisableTiming) != hipSuccess) rc = BAAS_HIP_ERROR;
    if (rc == BAAS_OK && hipHostMalloc((void **)&B->h_ctl, 4 * sizeof(int), hipHostMallocDefault) != hipSuccess)
        rc = BAAS_HIP_ERROR;
    if (rc == BAAS_OK) {
        if (hipMemcpy(B->sys_of.p, sys_of.data(), N * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(B->sys_chunk.p, sys_chunk.data(), (nsys + 1) * sizeof(int), hipMemcpyHostToDevice) !=
                hipSuccess ||
            hipMemcpy(B->chunks.p, ch.data(), ch.size() * sizeof(Chunk), hipMemcpyHostToDevice) != hipSuccess)
            rc = BAAS_HIP_ERROR;
    }
    if (rc != BAAS_OK) {
        baas_batch_destroy(B);
        return rc;
    }
    *out = B;
    return BAAS_OK;
}

extern "C" int baas_batch_destroy(baas_batch *B) {
    if (!B) return BAAS_INVALID_ARG;
    if (B->st) (void)hipStreamSynchronize(B->st);
    if (B->exec) (void)hipGraphExecDestroy(B->exec);
    if (B->graph) (void)hipGraphDestroy(B->graph);
    if (B->ev) (void)hipEventDestroy(B->ev);
    if (B->st) (void)hipStreamDestroy(B->st);
    if (B->h_ctl) (void)hipHostFree(B->h_ctl);
    if (B->s) baas_schedule_destroy(B->s);
    delete B;
    return BAAS_OK;
}

extern "C" int baas_batch_pcg(baas_batch *B, const double *Bm, double *X, double tol, int maxiter, int *iters) {
    if (!B || !Bm || !X || !(tol > 0) || maxiter < 1) return BAAS_INVALID_ARG;
    baas_handle *h = B->h;
    const int N = B->N, k = B->k, ncol = B->nsys * B->k;
    const long long M = (long long)N * k;
    BAAS_HIP_TRY(hipEventRecord(B->ev, h->stream));
    BAAS_HIP_TRY(hipStreamWaitEvent(B->st, B->ev, 0));
    auto t0 = std::chrono::steady_clock::now();
    {  // IC0 (FP64) on all systems in one launch, FP32 split into L and U = L^T
        hipStream_t keep = h->stream;
        h->stream = B->st;
        int rc = baas_ic0_f64(h, B->s, B->a_vals, B->lu.p);
        h->stream = keep;
        BAAS_TRY(rc);
        BAAS_TRY(ic0_split_dev32(h, B->s, B->lu.p, B->lv32.p, B->uv32.p, B->st));
    }
    BAAS_HIP_TRY(hipMemcpyAsync(B->tol.p, &tol, sizeof(double), hipMemcpyHostToDevice, B->st));
    // r = b - A x, per-column ||r0||
    hipLaunchKernelGGL(k_bcopy, dim3(nblk(M)), dim3(256), 0, B->st, M, X, B->x.p);
    launch_spmv_multi(B->s->fwd.max_deg, B->st, N, k, B->rowptr, B->colidx, B->a_vals, B->x.p, B->Ap.p, nullptr);
    hipLaunchKernelGGL(k_bres, dim3(nblk(M)), dim3(256), 0, B->st, M, Bm, B->Ap.p, B->r.p, B->r32.p);
    BAAS_TRY(chunk_dot_sys(B, B->r.p, B->r.p, B->norm0.p, 1, nullptr));
    hipLaunchKernelGGL(k_binit, dim3(nblk(ncol)), dim3(256), 0, B->st, ncol, maxiter, B->coldone.p, B->coliter.p,
                       B->ctl.p, B->norm0.p);
    BAAS_HIP_TRY(hipStreamSynchronize(B->st));
    auto t1 = std::chrono::steady_clock::now();
    if (!B->tuned) BAAS_TRY(autotune(B));  // times sweeps on r32 only; z32/y32 are recomputed below
    auto t2 = std::chrono::steady_clock::now();
    // z = M^-1 r, p = z, rz
    BAAS_TRY(batch_precond(B, nullptr));
    hipLaunchKernelGGL(k_cast_back, dim3(nblk(M)), dim3(256), 0, B->st, M, B->z32.p, B->z.p, nullptr);
    BAAS_TRY(chunk_dot_sys(B, B->r.p, B->z.p, B->rz0.p, 0, nullptr));
    hipLaunchKernelGGL(k_bcopy, dim3(nblk(M)), dim3(256), 0, B->st, M, B->z.p, B->p.p);
    BAAS_HIP_TRY(hipGetLastError());
    BAAS_TRY(batch_capture(B));
    for (;;) {
        BAAS_HIP_TRY(hipGraphLaunch(B->exec, B->st));
        BAAS_HIP_TRY(hipMemcpyAsync(B->h_ctl, B->ctl.p, 4 * sizeof(int), hipMemcpyDeviceToHost, B->st));
        BAAS_HIP_TRY(hipStreamSynchronize(B->st));
        if (B->h_ctl[0]) break;
    }
    auto t3 = std::chrono::steady_clock::now();
    hipLaunchKernelGGL(k_bcopy, dim3(nblk(M)), dim3(256), 0, B->st, M, B->x.p, X);
    if (iters) BAAS_HIP_TRY(hipMemcpyAsync(iters, B->coliter.p, ncol * sizeof(int), hipMemcpyDeviceToHost, B->st));
    BAAS_HIP_TRY(hipStreamSynchronize(B->st));
    BAAS_HIP_TRY(hipEventRecord(B->ev, B->st));
    BAAS_HIP_TRY(hipStreamWaitEvent(h->stream, B->ev, 0));
    B->t_ic = std::chrono::duration<double>(t1 - t0).count();
    B->t_iter = std::chrono::duration<double>((t3 - t2)).count();
    bool all = true;
    if (iters)
        for (int c = 0; c < ncol; c++) all = all && iters[c] >= 0;
    return all ? BAAS_OK : BAAS_NOT_CONVERGED;
}

extern "C" int baas_batch_info(const baas_batch *B, double *times, int *rpu) {
    if (!B) return BAAS_INVALID_ARG;
    if (times) {
        times[0] = B->t_ic;
        times[1] = B->t_iter;
        times[2] = B->t_tune;
    }
    if (rpu)
        for (int q = 0; q < 3; q++) rpu[q] = B->tuned ? B->tuned_rpu[q] : 0;
    return BAAS_OK;
}
