// tune.hip -- the execution-plan autotuner.
//
// The reference fixes its level-merge and threads-per-level thresholds at
// compile time (merge/LevelMerge.cpp:118-161, merge/include/MergeMarco.hpp).
// Here the packing is chosen per matrix by timing real sweeps: RANGE plans
// over several rows-per-unit and workgroup sizes, with and without the
// dataflow kernel, SCHEDULE plans (units cut from the merged levels) over
// several rows-per-unit, and LEVELS plans (level sets with device-wide
// barriers).  The fastest plan stays on the schedule.  Values need not be
// the final ones: any non-zero diagonal gives the same timing.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "blas.h"

#include "internal.h"

namespace baas {
int build_plans_mode(baas_schedule *s, int mode, int rows_per_unit);
int build_factor_lane(baas_schedule *s);
size_t lds_limit(baas_handle *h);
size_t pipe_lds(const Plan &p, int ts);
}  // namespace baas

using namespace baas;

namespace {

struct Cand {
    int mode, rpu, block, df, band;
};

// a candidate replaces the fastest so far only below this fraction of its time
constexpr float TUNE_MARGIN = 0.97f;

// BAAS_TUNE_LOG set (to anything but 0) prints every timed candidate to stderr
bool tune_log() {
    static const bool on = [] {
        const char *e = std::getenv("BAAS_TUNE_LOG");
        return e && *e && std::strcmp(e, "0") != 0;
    }();
    return on;
}

void log_cand(const char *what, const Cand &c, bool ok, float ms) {
    if (tune_log())
        std::fprintf(stderr, "[baas tune %s] mode %d rows %d block %d dataflow %d band %d: %s%.1f us\n", what, c.mode,
                     c.rpu, c.block, c.df, c.band, ok ? "" : "failed ", ok ? ms * 1000.0f : 0.0f);
}

// TILES candidates: row ranges by level bands
constexpr int TILE_RANGES[] = {8192, 16384};
constexpr int TILE_BANDS[] = {16, 64};

template <class T>
__global__ void k_fill(int n, T *p, T v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace

namespace baas {

// Time every candidate plan for value type T and keep the fastest.  FP32
// sweeps need half the LDS per row, so their dataflow units may be larger.
template <class T>
int autotune(baas_handle *h, baas_schedule *s, const T *vals, int uplo, int nrhs, int *chosen) {
    if (!h || !s || !vals || nrhs < 1 || (uplo != 0 && uplo != 1)) return BAAS_INVALID_ARG;
    const int n = s->n;
    DBuf<T> b, x;
    BAAS_TRY(b.alloc((size_t)n * nrhs));
    BAAS_TRY(x.alloc((size_t)n * nrhs));
    hipLaunchKernelGGL(k_fill<T>, dim3((n * nrhs + 255) / 256), dim3(256), 0, h->stream, n * nrhs, b.p, T(1));
    // candidates sharing (mode, rows per unit, band) are adjacent: their
    // plan is built once and only the workgroup size changes
    std::vector<Cand> cands;
    std::vector<int> rpus = {1024, 2048, 4096, 8192, 12288, 16384};
    if (sizeof(T) == 4) rpus.insert(rpus.end(), {24576, 32768});
    for (int rpu : rpus)
        for (int blk : {256, 384, 512, 1024}) {
            const bool big = blk == 1024 && rpu < 4096;
            if (rpu <= 16384 && !big) cands.push_back({BAAS_PLAN_RANGE, rpu, blk, 0, 0});
            if (nrhs == 1 && blk <= 512) cands.push_back({BAAS_PLAN_RANGE, rpu, blk, 1, 0});
        }
    // lane kernel: units small enough for their rows (times the right-hand
    // sides) to live in LDS
    if (nrhs == 1)
        for (int rpu : {384, 512, 768, 1024, 1536})
            for (int blk : {512, 768, 1024}) cands.push_back({BAAS_PLAN_RANGE, rpu, blk, 2, 0});
    for (int rpu : {256, 1024, 4096})
        for (int blk : {256, 1024}) cands.push_back({BAAS_PLAN_SCHEDULE, rpu, blk, 0, 0});
    // small units of the merged levels on the lane kernel: every row starts
    // as soon as its own parents are published, with no level barrier
    if (nrhs == 1) {
        for (int rpu : {16, 32, 40, 64, 96, 128})
            for (int blk : {512, 1024}) cands.push_back({BAAS_PLAN_SCHEDULE, rpu, blk, 2, 0});
    } else {
        for (int rpu : {16, 32}) cands.push_back({BAAS_PLAN_SCHEDULE, rpu, 1024, 2, 0});
    }
    for (int blk : {64, 128, 256, 512, 1024}) cands.push_back({BAAS_PLAN_LEVELS, 0, blk, 0, 0});
    // grid boxes (matrices without stencil-grid structure fail to build and
    // are skipped); 900 rows (10 x 10 x 9 on a 3D grid) leave no lane of a
    // 1024-thread workgroup with a second row
    for (int R : {512, 900, 1000, 1331}) {
        if (nrhs == 1) {
            cands.push_back({BAAS_PLAN_BOXES, R, 1024, 2, 0});
            cands.push_back({BAAS_PLAN_BOXES, R, 512, 2, 0});
            cands.push_back({BAAS_PLAN_BOXES, R, 384, 1, 0});
        }
        cands.push_back({BAAS_PLAN_BOXES, R, 256, 0, 0});
    }
    // the wave-pipelined grid sweep: boxes of 8 x 8 (64 x 1 on a 2D grid)
    // rows across by R / 64 planes; several right-hand sides run their
    // pipelines side by side in one launch
    for (int R : {512, 1024, 2048}) cands.push_back({BAAS_PLAN_PIPE, R, 128, 5, 0});
    if (nrhs > 1) {
        for (int R : {64, 125, 216}) cands.push_back({BAAS_PLAN_BOXES, R, 1024, 2, 0});
        // the multi-RHS lane kernel (parents of other units read directly,
        // b staged in LDS where it fits or read from memory) on small
        // boxes, small units of the merged levels and thin level bands
        for (int R : {64, 125, 216, 512})
            for (int df : {3, 4}) cands.push_back({BAAS_PLAN_BOXES, R, 1024, df, 0});
        for (int rpu : {16, 32, 64}) cands.push_back({BAAS_PLAN_SCHEDULE, rpu, 1024, 4, 0});
    }
    if (nrhs > 1)
        for (int S : {1024, 2048})
            for (int hb : {2, 4}) cands.push_back({BAAS_PLAN_TILES, S, 1024, 4, hb});
    for (int S : TILE_RANGES) {
        if (S >= 2 * n && S != TILE_RANGES[0]) continue;
        for (int hb : TILE_BANDS)
            for (int blk : {256, 512, 1024}) {
                if (blk != 512) cands.push_back({BAAS_PLAN_TILES, S, blk, 0, hb});
                if (nrhs == 1 && blk <= 512) cands.push_back({BAAS_PLAN_TILES, S, blk, 1, hb});
                if (nrhs == 1 && blk >= 512) cands.push_back({BAAS_PLAN_TILES, S, blk, 2, hb});
            }
    }
    hipEvent_t e0, e1;
    BAAS_HIP_TRY(hipEventCreate(&e0));
    BAAS_HIP_TRY(hipEventCreate(&e1));
    float best = 1e30f;
    Cand win = cands[0];
    const int band0 = s->tile_band;
    Cand built = {-1, -1, 0, 0, -1};
    for (const Cand &c : cands) {
        if (c.rpu >= 2 * n && c.mode == BAAS_PLAN_RANGE && c.rpu != 1024) continue;
        s->opts.block_threads = c.block;
        if (c.mode != built.mode || c.rpu != built.rpu || c.band != built.band ||
            (c.mode == BAAS_PLAN_SCHEDULE && c.rpu == 0 && c.block != built.block)) {
            if (c.mode == BAAS_PLAN_TILES) s->tile_band = c.band;
            built = {-1, -1, 0, 0, -1};
            if (build_plans_mode(s, c.mode, c.rpu) != BAAS_OK) continue;
            built = c;
        }
        s->fwd.block = s->bwd.block = c.block;
        Plan &p = uplo ? s->bwd : s->fwd;
        if (c.df == 1 && p.df_lds[sizeof(T) == 8] > lds_limit(h)) continue;
        if (c.df == 2 && p.lane_lds((int)sizeof(T), nrhs) > lds_limit(h)) continue;
        if ((c.df == 3 || c.df == 4) && p.mr_lds((int)sizeof(T), nrhs, false) > lds_limit(h)) continue;
        if (c.df == 5 && (!p.pipe || pipe_lds(p, (int)sizeof(T)) > lds_limit(h))) continue;
        s->fwd.dataflow = s->bwd.dataflow = c.df;
        float t_best = 1e30f;
        bool ok = true;
        for (int rep = 0; rep < 4 && ok; rep++) {
            (void)hipEventRecord(e0, h->stream);
            ok = sweep<T>(h, s, uplo, vals, b.p, x.p, nrhs, h->stream, nullptr) == BAAS_OK;
            (void)hipEventRecord(e1, h->stream);
            (void)hipEventSynchronize(e1);
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (rep > 0 && ms < t_best) t_best = ms;
        }
        log_cand("sweep", c, ok, t_best);
        // a later candidate must be clearly faster to replace the incumbent,
        // so near-ties resolve the same way from run to run
        if (ok && t_best < best * TUNE_MARGIN) best = t_best, win = c;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    s->opts.block_threads = win.block;
    s->opts.plan_mode = win.mode;
    s->opts.rows_per_unit = win.rpu;
    s->tile_band = win.mode == BAAS_PLAN_TILES ? win.band : band0;
    BAAS_TRY(build_plans_mode(s, win.mode, win.rpu));
    s->fwd.dataflow = s->bwd.dataflow = win.df;
    // a factorisation on this schedule will use the lane kernel too: build
    // its records now, as part of the analysis, not on the first call
    if (win.df == 2 && s->kind != BAAS_KIND_TRSV) (void)build_factor_lane(s);
    if (chosen) {
        chosen[0] = win.mode;
        chosen[1] = win.rpu;
        chosen[2] = win.block;
        chosen[3] = win.df;
        chosen[4] = (int)(best * 1000.0f);  // microseconds of the winning sweep
    }
    return baas_check(h) == BAAS_OK ? BAAS_OK : BAAS_HIP_ERROR;
}

template int autotune<double>(baas_handle *, baas_schedule *, const double *, int, int, int *);
template int autotune<float>(baas_handle *, baas_schedule *, const float *, int, int, int *);

}  // namespace baas

// The same search for the numeric factorisation of an IC0 or ILU0
// schedule: unit plans, schedule plans and level-set plans, timed on the
// caller's values (into a scratch factor).
static int autotune_factor(baas_handle *h, baas_schedule *s, const double *a_vals, int *chosen) {
    if (!h || !s || !a_vals || (s->kind != BAAS_KIND_IC0 && s->kind != BAAS_KIND_ILU0)) return BAAS_INVALID_ARG;
    DBuf<double> lu;
    BAAS_TRY(lu.alloc(s->nnz_full));
    std::vector<Cand> cands;
    for (int rpu : {1024, 2048, 4096, 8192})
        for (int blk : {256, 512}) cands.push_back({BAAS_PLAN_RANGE, rpu, blk, 0, 0});
    for (int rpu : {0, 1024})
        for (int blk : {256, 1024}) cands.push_back({BAAS_PLAN_SCHEDULE, rpu, blk, 0, 0});
    for (int blk : {64, 128, 256, 512}) cands.push_back({BAAS_PLAN_LEVELS, 0, blk, 0, 0});
    for (int hb : {16, 64})
        for (int blk : {256, 512}) cands.push_back({BAAS_PLAN_TILES, 16384, blk, 0, hb});
    for (int R : {512, 1024})
        for (int blk : {256, 512}) cands.push_back({BAAS_PLAN_BOXES, R, blk, 0, 0});
    {  // IC0 and ILU0 on the lane kernel
        for (int rpu : {256, 512, 1024})
            for (int blk : {512, 1024}) cands.push_back({BAAS_PLAN_RANGE, rpu, blk, 2, 0});
        for (int rpu : {16, 32, 64})
            for (int blk : {512, 1024}) cands.push_back({BAAS_PLAN_SCHEDULE, rpu, blk, 2, 0});
        for (int R : {512, 576, 900, 1000}) cands.push_back({BAAS_PLAN_BOXES, R, 1024, 2, 0});
    }
    // the pipelined factorisation on full stencils (other patterns fall
    // back to the lane kernel on the same plan)
    for (int R : {512, 1024, 2048}) cands.push_back({BAAS_PLAN_PIPE, R, 128, 5, 0});
    hipEvent_t e0, e1;
    BAAS_HIP_TRY(hipEventCreate(&e0));
    BAAS_HIP_TRY(hipEventCreate(&e1));
    float best = 1e30f;
    Cand win = cands[0];
    const int band0 = s->tile_band;
    Cand built = {-1, -1, 0, 0, -1};
    for (const Cand &c : cands) {
        if (c.mode == BAAS_PLAN_RANGE && c.rpu >= 2 * s->n && c.rpu != 1024) continue;
        s->opts.block_threads = c.block;
        if (c.mode != built.mode || c.rpu != built.rpu || c.band != built.band ||
            (c.mode == BAAS_PLAN_SCHEDULE && c.rpu == 0 && c.block != built.block)) {
            if (c.mode == BAAS_PLAN_TILES) s->tile_band = c.band;
            built = {-1, -1, 0, 0, -1};
            if (build_plans_mode(s, c.mode, c.rpu) != BAAS_OK) continue;
            built = c;
        }
        s->fwd.block = s->bwd.block = c.block;
        s->fwd.dataflow = s->bwd.dataflow = c.df;
        float t_best = 1e30f;
        bool ok = true;
        for (int rep = 0; rep < 3 && ok; rep++) {
            (void)hipEventRecord(e0, h->stream);
            ok = (s->kind == BAAS_KIND_IC0 ? baas_ic0_f64(h, s, a_vals, lu.p) : baas_ilu0_f64(h, s, a_vals, lu.p)) ==
                 BAAS_OK;
            (void)hipEventRecord(e1, h->stream);
            (void)hipEventSynchronize(e1);
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (rep > 0 && ms < t_best) t_best = ms;
        }
        log_cand("factor", c, ok, t_best);
        // a later candidate must be clearly faster to replace the incumbent,
        // so near-ties resolve the same way from run to run
        if (ok && t_best < best * TUNE_MARGIN) best = t_best, win = c;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    s->opts.block_threads = win.block;
    s->opts.plan_mode = win.mode;
    s->opts.rows_per_unit = win.rpu;
    s->tile_band = win.mode == BAAS_PLAN_TILES ? win.band : band0;
    BAAS_TRY(build_plans_mode(s, win.mode, win.rpu));
    s->fwd.dataflow = s->bwd.dataflow = win.df;
    if (win.df == 2) (void)build_factor_lane(s);  // the records, now rather than on the first call
    if (chosen) {
        chosen[0] = win.mode;
        chosen[1] = win.rpu;
        chosen[2] = win.block;
        chosen[3] = win.df;
        chosen[4] = (int)(best * 1000.0f);
    }
    // pivot reports of the trial runs are the caller's to see on the real call
    const int st = baas_check(h);
    return st == BAAS_OK || st == BAAS_NONPOSITIVE_PIVOT || st == BAAS_ZERO_PIVOT ? BAAS_OK : st;
}

extern "C" int baas_autotune_factor(baas_handle *h, baas_schedule *s, const double *a_vals, int *chosen) {
    return autotune_factor(h, s, a_vals, chosen);
}

extern "C" int baas_autotune(baas_handle *h, baas_schedule *s, const double *vals, int uplo, int nrhs, int *chosen) {
    return baas::autotune<double>(h, s, vals, uplo, nrhs, chosen);
}

extern "C" int baas_autotune_f32(baas_handle *h, baas_schedule *s, const float *vals, int uplo, int nrhs,
                                 int *chosen) {
    return baas::autotune<float>(h, s, vals, uplo, nrhs, chosen);
}
