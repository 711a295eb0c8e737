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
// The execution plan -- the level-set plan, boxes of each system's own
// stencil grid on the multi-RHS lane kernel, or row ranges aligned to
// systems -- is chosen by timing FP32 sweeps of the actual batch.
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
size_t lds_limit(baas_handle *h);

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
    const int sh = __builtin_ctz(k);  // k is a power of two
    int i = (int)(t >> sh), r = (int)(t & (k - 1));
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

// The iteration's fused kernels: one workgroup per chunk (256 rows of one
// system, every right-hand side), each element computed where it is first
// needed and its dot-product term accumulated in the same pass, in the
// order k_chunk_dot uses (thread t walks elements t, t + 256, ... of the
// chunk: always column t % k), so the partial sums are the same bits.
__device__ __forceinline__ void chunk_reduce(int k, double acc, double *red, double *part) {
    red[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x < (unsigned)k) {
        double s = 0.0;
        for (int q = threadIdx.x; q < (int)blockDim.x; q += k) s += red[q];
        part[(size_t)blockIdx.x * k + threadIdx.x] = s;
    }
}

// Ap = A p and the chunk's p . Ap partials
template <int CH>
__global__ void k_spmv_pap(int k, const Chunk *__restrict__ ch, const int *__restrict__ rp, const int *__restrict__ ci,
                           const double *__restrict__ va, const double *__restrict__ p, double *__restrict__ Ap,
                           double *part, const int *alldone, const int *__restrict__ coldone) {
    extern __shared__ double red[];
    if (stopped(alldone)) return;
    const Chunk c = ch[blockIdx.x];
    const int sh = __builtin_ctz(k), col = (int)(threadIdx.x & (unsigned)(k - 1));
    double acc = 0.0;
    if (!coldone[c.sys * k + col]) {
        for (size_t e = (size_t)c.r0 * k + threadIdx.x; e < (size_t)c.r1 * k; e += blockDim.x) {
            const int i = (int)(e >> sh);
            const int b = rp[i], en = rp[i + 1];
            double s = 0.0;
            for (int j0 = b; j0 < en; j0 += CH) {
                double xv[CH], av[CH];
#pragma unroll
                for (int q = 0; q < CH; q++) {
                    const int j = j0 + q < en ? j0 + q : en - 1;
                    av[q] = va[j];
                    xv[q] = p[(size_t)ci[j] * k + col];
                }
#pragma unroll
                for (int q = 0; q < CH; q++)
                    if (j0 + q < en) s = fma(av[q], xv[q], s);
            }
            Ap[e] = s;
            acc = fma(p[e], s, acc);
        }
    }
    chunk_reduce(k, acc, red, part);
}

// The same with two columns a thread (k even): a row's indices and values
// are read once for both, and p and Ap move in 16-byte pairs.  The chunk's
// p . Ap partials sum each column's per-thread partials in thread order.
template <int CH>
__global__ void k_spmv_pap2(int k, const Chunk *__restrict__ ch, const int *__restrict__ rp,
                            const int *__restrict__ ci, const double *__restrict__ va, const double *__restrict__ p,
                            double *__restrict__ Ap, double *part, const int *alldone, const int *__restrict__ coldone) {
    extern __shared__ double red[];  // 2 blockDim.x values
    if (stopped(alldone)) return;
    const Chunk c = ch[blockIdx.x];
    const int kp = k >> 1, sh = __builtin_ctz(kp), cp = (int)(threadIdx.x & (unsigned)(kp - 1));
    const bool run0 = !coldone[c.sys * k + 2 * cp], run1 = !coldone[c.sys * k + 2 * cp + 1];
    double acc0 = 0.0, acc1 = 0.0;
    if (run0 | run1) {
        for (size_t e = (size_t)c.r0 * kp + threadIdx.x; e < (size_t)c.r1 * kp; e += blockDim.x) {
            const int i = (int)(e >> sh);
            const int b = rp[i], en = rp[i + 1];
            double s0 = 0.0, s1 = 0.0;
            for (int j0 = b; j0 < en; j0 += CH) {
                double2 xv[CH];
                double av[CH];
#pragma unroll
                for (int q = 0; q < CH; q++) {
                    const int j = j0 + q < en ? j0 + q : en - 1;
                    av[q] = va[j];
                    xv[q] = *reinterpret_cast<const double2 *>(p + (size_t)ci[j] * k + 2 * cp);
                }
#pragma unroll
                for (int q = 0; q < CH; q++)
                    if (j0 + q < en) {
                        s0 = fma(av[q], xv[q].x, s0);
                        s1 = fma(av[q], xv[q].y, s1);
                    }
            }
            reinterpret_cast<double2 *>(Ap)[e] = make_double2(s0, s1);
            const double2 pe = reinterpret_cast<const double2 *>(p)[e];
            acc0 = fma(pe.x, s0, acc0);
            acc1 = fma(pe.y, s1, acc1);
        }
    }
    red[2 * threadIdx.x] = run0 ? acc0 : 0.0;
    red[2 * threadIdx.x + 1] = run1 ? acc1 : 0.0;
    __syncthreads();
    if (threadIdx.x < (unsigned)k) {
        const int pc = (int)threadIdx.x >> 1, hf = (int)threadIdx.x & 1;
        double s = 0.0;
        for (int q = pc; q < (int)blockDim.x; q += kp) s += red[2 * q + hf];
        part[(size_t)blockIdx.x * k + threadIdx.x] = s;
    }
}

// x += alpha p, r -= alpha Ap, r32 = (float) r and the chunk's r . r partials
__global__ void k_update_xr_rr(int k, const Chunk *__restrict__ ch, double *x, double *r,
                               const double *__restrict__ p, const double *__restrict__ Ap,
                               const double *__restrict__ rz, const double *__restrict__ pap,
                               const int *__restrict__ coldone, float *r32, double *part, const int *alldone) {
    extern __shared__ double red[];
    if (stopped(alldone)) return;
    const Chunk c = ch[blockIdx.x];
    const int col = (int)(threadIdx.x & (unsigned)(k - 1)), cc = c.sys * k + col;
    double acc = 0.0;
    if (!coldone[cc]) {
        const double alpha = rz[cc] / pap[cc];
        for (size_t e = (size_t)c.r0 * k + threadIdx.x; e < (size_t)c.r1 * k; e += blockDim.x) {
            x[e] = fma(alpha, p[e], x[e]);
            const double ri = fma(-alpha, Ap[e], r[e]);
            r[e] = ri;
            r32[e] = (float)ri;
            acc = fma(ri, ri, acc);
        }
    }
    chunk_reduce(k, acc, red, part);
}

// the chunk's r . z partials with z = (double) z32 (exact)
__global__ void k_rz32(int k, const Chunk *__restrict__ ch, const double *__restrict__ r,
                       const float *__restrict__ z32, double *part, const int *alldone,
                       const int *__restrict__ coldone) {
    extern __shared__ double red[];
    if (stopped(alldone)) return;
    const Chunk c = ch[blockIdx.x];
    const int col = (int)(threadIdx.x & (unsigned)(k - 1));
    double acc = 0.0;
    if (!(coldone && coldone[c.sys * k + col]))
        for (size_t e = (size_t)c.r0 * k + threadIdx.x; e < (size_t)c.r1 * k; e += blockDim.x)
            acc = fma(r[e], (double)z32[e], acc);
    chunk_reduce(k, acc, red, part);
}

// p = z + (rz_new / rz_old) p for running columns, z = (double) z32; with
// rz_old == nullptr p = z (the first direction)
__global__ void k_update_p32(int k, const Chunk *__restrict__ ch, double *p, const float *__restrict__ z32,
                             const double *__restrict__ rz_new, const double *__restrict__ rz_old,
                             const int *__restrict__ coldone, const int *alldone) {
    if (stopped(alldone)) return;
    const Chunk c = ch[blockIdx.x];
    const int col = (int)(threadIdx.x & (unsigned)(k - 1)), cc = c.sys * k + col;
    if (coldone && coldone[cc]) return;
    if (!rz_old) {
        for (size_t e = (size_t)c.r0 * k + threadIdx.x; e < (size_t)c.r1 * k; e += blockDim.x) p[e] = (double)z32[e];
        return;
    }
    const double beta = rz_new[cc] / rz_old[cc];
    for (size_t e = (size_t)c.r0 * k + threadIdx.x; e < (size_t)c.r1 * k; e += blockDim.x)
        p[e] = fma(beta, p[e], (double)z32[e]);
}

// per work unit of a plan: skip it when every column of its system is done
__global__ void k_unit_skip(int nunits, int k, const int *__restrict__ usys, const int *__restrict__ coldone,
                            int *usk, const int *alldone) {
    if (stopped(alldone)) return;
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nunits) return;
    const int *c = coldone + (size_t)usys[u] * k;
    int all = 1;
    for (int r = 0; r < k; r++) all &= c[r] != 0;
    usk[u] = all;
}

// per system and column: sum the chunk partials of the system in order
__global__ void k_sys_sum(int nsys, int k, const int *__restrict__ sys_chunk, const double *__restrict__ part,
                          double *out, int do_sqrt, const int *alldone) {
    if (stopped(alldone)) return;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nsys * k) return;
    const int sh = __builtin_ctz(k);
    int s = t >> sh, r = t & (k - 1);
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
    DBuf<int> sys_chunk;
    DBuf<Chunk> chunks;
    DBuf<double> lu, x, r, p, Ap, part, rz0, rz1, pap, rr, norm0, tol;
    DBuf<float> lv32, uv32, r32, y32, z32;
    DBuf<int> coldone, coliter, ctl;
    // per unit of the forward / backward plan: its system, and the skip
    // flags the iteration sets for units of finished systems (box plans)
    DBuf<int> usys_f, usys_b, usk_f, usk_b;
    int nu_f = 0, nu_b = 0;
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

static int batch_precond(baas_batch *B, const int *skip, bool skip_done = false) {
    const int *uf = skip_done && B->nu_f ? B->usk_f.p : nullptr, *ub = skip_done && B->nu_b ? B->usk_b.p : nullptr;
    BAAS_TRY(sweep<float>(B->h, B->s, 0, B->lv32.p, B->r32.p, B->y32.p, B->k, B->st, skip, nullptr, 0, uf));
    BAAS_TRY(sweep<float>(B->h, B->s, 1, B->uv32.p, B->y32.p, B->z32.p, B->k, B->st, skip, nullptr, 0, ub));
    return BAAS_OK;
}

// The system of every unit of the plans in use (each unit of a box plan
// lies in one system), for skipping finished systems' units; other plans
// get none.
static int build_unit_sys(baas_batch *B) {
    B->nu_f = B->nu_b = 0;
    if (B->s->fwd.mode != BAAS_PLAN_BOXES || B->s->fwd.dataflow < 3) return BAAS_OK;
    std::vector<int> sys_of(B->N);
    for (int s = 0; s < B->nsys; s++)
        for (int q = B->sys_ptr[s]; q < B->sys_ptr[s + 1]; q++) sys_of[q] = s;
    int nus[2] = {0, 0};
    for (int dir = 0; dir < 2; dir++) {
        const Plan &p = dir ? B->s->bwd : B->s->fwd;
        const int nu = p.nunits;
        if ((int)p.h_unit_ptr.size() != nu + 1 || (int)p.h_rows.size() != B->N) return BAAS_OK;
        std::vector<int> us(nu);
        for (int u = 0; u < nu; u++) {
            us[u] = sys_of[p.h_rows[p.h_unit_ptr[u]]];
            for (int q = p.h_unit_ptr[u]; q < p.h_unit_ptr[u + 1]; q++)
                if (sys_of[p.h_rows[q]] != us[u]) return BAAS_OK;  // a unit across systems: no skipping
        }
        DBuf<int> &d = dir ? B->usys_b : B->usys_f;
        DBuf<int> &f = dir ? B->usk_b : B->usk_f;
        BAAS_TRY(d.alloc(nu));
        BAAS_TRY(f.alloc(nu));
        BAAS_HIP_TRY(hipMemcpy(d.p, us.data(), sizeof(int) * nu, hipMemcpyHostToDevice));
        nus[dir] = nu;
    }
    B->nu_f = nus[0];
    B->nu_b = nus[1];
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

// Plan search for the batch, timing the FP32 forward + backward sweeps of
// the whole batch with all RHS: the level-set plan, per-system boxes on the
// multi-RHS lane kernel, and (where boxes do not apply) row ranges aligned
// to systems; the fastest is kept.
static int autotune(baas_batch *B) {
    auto t0 = std::chrono::steady_clock::now();
    const int block = B->s->opts.block_threads;
    auto done = [&](std::vector<int> rpu) {
        B->s->opts.block_threads = block;
        B->tuned_rpu = rpu;
        B->tuned = true;
        B->t_tune = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return BAAS_OK;
    };
    // the level-set plan (all systems' rows of one level at a time)
    float best = 1e30f;
    int lev_block = 0;
    for (int lb : {256, 512, 1024}) {
        B->s->opts.block_threads = lb;
        BAAS_TRY(build_plans_mode(B->s, BAAS_PLAN_LEVELS, 0));
        const float t_lev = time_sweeps(B, 3);
        if (t_lev < best) best = t_lev, lev_block = lb;
    }
    // every system a stencil grid: boxes of each system's own grid on the
    // multi-RHS lane kernel (b staged in LDS or read from memory), no
    // device-wide barrier; a system that is not a stencil makes the plan
    // unsupported
    int box_rows = 0, box_block = 0, box_df = 0;
    bool boxes = true;
    for (int R : {64, 144, 256, 400, 576}) {
        for (int blk : {512, 1024}) {
            B->s->opts.block_threads = blk;
            if (build_plans_mode(B->s, BAAS_PLAN_BOXES, R) != BAAS_OK) {
                boxes = false;
                break;
            }
            for (int df : {3, 4}) {
                B->s->fwd.dataflow = B->s->bwd.dataflow = df;
                if (B->s->fwd.mr_lds((int)sizeof(float), B->k, false) > lds_limit(B->h) ||
                    B->s->bwd.mr_lds((int)sizeof(float), B->k, false) > lds_limit(B->h))
                    continue;
                const float t = time_sweeps(B, 3);
                if (t < best) best = t, box_rows = R, box_block = blk, box_df = df;
            }
        }
        if (!boxes) break;
    }
    if (box_rows) {
        B->s->opts.block_threads = box_block;
        BAAS_TRY(build_plans_mode(B->s, BAAS_PLAN_BOXES, box_rows));
        B->s->fwd.dataflow = B->s->bwd.dataflow = box_df;
        return done({box_rows, box_rows, box_rows});
    }
    // otherwise row ranges aligned to systems: a coordinate search over rows
    // per unit, one size bucket at a time (only where boxes do not apply)
    std::vector<int> rpu = {2048, 2048, 2048};
    bool range_won = false;
    if (!boxes) {
        const int cand[] = {512, 1024, 2048, 4096, 8192};
        for (int bk = 0; bk < 3; bk++) {
            for (int c : cand) {
                std::vector<int> trial = rpu;
                trial[bk] = c;
                if (c == rpu[bk] && bk > 0) continue;
                BAAS_TRY(make_plan_custom(B->s, batch_units(B, trial), block));
                const float t = time_sweeps(B, 3);
                if (t < best) best = t, rpu = trial, range_won = true;
            }
        }
    }
    if (range_won) {
        BAAS_TRY(make_plan_custom(B->s, batch_units(B, rpu), block));
        return done(rpu);
    }
    B->s->opts.block_threads = lev_block;
    BAAS_TRY(build_plans_mode(B->s, BAAS_PLAN_LEVELS, 0));
    return done({0, 0, lev_block});
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

static int sys_sum(baas_batch *B, double *out, int do_sqrt, const int *skip) {
    hipLaunchKernelGGL(k_sys_sum, dim3(nblk((long long)B->nsys * B->k)), dim3(256), 0, B->st, B->nsys, B->k,
                       B->sys_chunk.p, B->part.p, out, do_sqrt, skip);
    BAAS_HIP_TRY(hipGetLastError());
    return BAAS_OK;
}

static int batch_iteration(baas_batch *B, int cur) {
    const int k = B->k, ncol = B->nsys * B->k;
    int *ctl = B->ctl.p;
    const int *cd = B->coldone.p;
    double *rz_old = cur ? B->rz1.p : B->rz0.p, *rz_new = cur ? B->rz0.p : B->rz1.p;
    const dim3 grid(B->nchunks), block(256);
    const size_t red = 256 * sizeof(double);
    const int w = 2 * B->s->fwd.max_deg + 1;  // the symmetric pattern's longest row
    if (k >= 2 && w <= 5)  // two columns a thread (k is a power of two)
        hipLaunchKernelGGL(k_spmv_pap2<5>, grid, block, 2 * red, B->st, k, B->chunks.p, B->rowptr, B->colidx,
                           B->a_vals, B->p.p, B->Ap.p, B->part.p, ctl, cd);
    else if (k >= 2 && w <= 7)
        hipLaunchKernelGGL(k_spmv_pap2<7>, grid, block, 2 * red, B->st, k, B->chunks.p, B->rowptr, B->colidx,
                           B->a_vals, B->p.p, B->Ap.p, B->part.p, ctl, cd);
    else if (w <= 5)
        hipLaunchKernelGGL(k_spmv_pap<5>, grid, block, red, B->st, k, B->chunks.p, B->rowptr, B->colidx, B->a_vals,
                           B->p.p, B->Ap.p, B->part.p, ctl, cd);
    else if (w <= 7)
        hipLaunchKernelGGL(k_spmv_pap<7>, grid, block, red, B->st, k, B->chunks.p, B->rowptr, B->colidx, B->a_vals,
                           B->p.p, B->Ap.p, B->part.p, ctl, cd);
    else
        hipLaunchKernelGGL(k_spmv_pap<8>, grid, block, red, B->st, k, B->chunks.p, B->rowptr, B->colidx, B->a_vals,
                           B->p.p, B->Ap.p, B->part.p, ctl, cd);
    BAAS_TRY(sys_sum(B, B->pap.p, 0, ctl));
    hipLaunchKernelGGL(k_update_xr_rr, grid, block, red, B->st, k, B->chunks.p, B->x.p, B->r.p, B->p.p, B->Ap.p,
                       rz_old, B->pap.p, cd, B->r32.p, B->part.p, ctl);
    BAAS_TRY(sys_sum(B, B->rr.p, 0, ctl));
    hipLaunchKernelGGL(k_bconv, dim3(nblk(ncol)), dim3(256), 0, B->st, ncol, B->rr.p, B->norm0.p, B->tol.p,
                       B->coldone.p, B->coliter.p, ctl);
    hipLaunchKernelGGL(k_bconv_finish, dim3(1), dim3(1), 0, B->st, ctl);
    if (B->nu_f && B->nu_b) {
        hipLaunchKernelGGL(k_unit_skip, dim3(nblk(B->nu_f)), dim3(256), 0, B->st, B->nu_f, k, B->usys_f.p, cd,
                           B->usk_f.p, ctl);
        hipLaunchKernelGGL(k_unit_skip, dim3(nblk(B->nu_b)), dim3(256), 0, B->st, B->nu_b, k, B->usys_b.p, cd,
                           B->usk_b.p, ctl);
    }
    BAAS_TRY(batch_precond(B, ctl, B->nu_f && B->nu_b));
    hipLaunchKernelGGL(k_rz32, grid, block, red, B->st, k, B->chunks.p, B->r.p, B->z32.p, B->part.p, ctl, cd);
    BAAS_TRY(sys_sum(B, rz_new, 0, ctl));
    hipLaunchKernelGGL(k_update_p32, grid, block, 0, B->st, k, B->chunks.p, B->p.p, B->z32.p, rz_new, rz_old, cd,
                       ctl);
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
    if (rc == BAAS_OK) B->s->block_ptr = B->sys_ptr;  // BOXES plans cut each system's grid
    // independent systems: the level-set plan runs every system's level at
    // once (the first factorisation uses it; the autotuner may keep it)
    if (rc == BAAS_OK) rc = build_plans_mode(B->s, BAAS_PLAN_LEVELS, 0);
    std::vector<int> sys_chunk(nsys + 1);
    std::vector<Chunk> ch;
    for (int s = 0; s < nsys && rc == BAAS_OK; s++) {
        sys_chunk[s] = (int)ch.size();
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
        A(B->sys_chunk.alloc(nsys + 1));
        A(B->chunks.alloc(ch.size()));
        A(B->lu.alloc(B->s->nnz_full));
        A(B->lv32.alloc(nl));
        A(B->uv32.alloc(nl));
        for (auto *d : {&B->x, &B->r, &B->p, &B->Ap}) A(d->alloc(M));
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
        if (hipMemcpy(B->sys_chunk.p, sys_chunk.data(), (nsys + 1) * sizeof(int), hipMemcpyHostToDevice) !=
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
    if (!B->tuned) {
        BAAS_TRY(autotune(B));  // times sweeps on r32 only; z32/y32 are recomputed below
        BAAS_TRY(build_unit_sys(B));
    }
    auto t2 = std::chrono::steady_clock::now();
    // z = M^-1 r, p = z, rz
    BAAS_TRY(batch_precond(B, nullptr));
    hipLaunchKernelGGL(k_rz32, dim3(B->nchunks), dim3(256), 256 * sizeof(double), B->st, k, B->chunks.p, B->r.p,
                       B->z32.p, B->part.p, (const int *)nullptr, (const int *)nullptr);
    BAAS_TRY(sys_sum(B, B->rz0.p, 0, nullptr));
    hipLaunchKernelGGL(k_update_p32, dim3(B->nchunks), dim3(256), 0, B->st, k, B->chunks.p, B->p.p, B->z32.p,
                       (const double *)nullptr, (const double *)nullptr, (const int *)nullptr, (const int *)nullptr);
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
