// pipe.hip -- the wave-pipelined sweep of stencil-grid triangular matrices
// (PIPE plans, dataflow 5).
//
// Same arithmetic as the reference's solves (merge/SpTRSV.cpp:24-39 forward,
// :601-660 backward with descending sums): x_r = (b_r - sum_j L_rj x_j) /
// L_rr, the sum in ascending column order for L (descending for U = L^T),
// one fma per term, IEEE division.  On a 7-point (5-point) grid in natural
// order the parents of row (x, y, z) are (x, y, z-1), (x, y-1, z) and
// (x-1, y, z) in that summation order, so the rows of level l are those
// with x + y + z = l.
//
// A work unit is a box of TX x TY rows across (x, y) by kz planes along z,
// swept by ONE wavefront: lane (i, j) owns the column (x0+i, y0+j) and at
// step t solves its row of plane k = t - i - j.  Its three parents were all
// solved at step t-1: the z parent by the lane itself, the y parent by lane
// (i, j-1) and the x parent by lane (i-1, j), moved by DPP shifts or the LDS
// crossbar.  Inside a box a dependency therefore costs one step -- a lane
// move, three fma and the division -- instead of an LDS hand-off and a
// poll.  Parents in other boxes (the plane below, the x and y faces) are
// staged into LDS by the workgroup's other wavefront as their producers
// publish them (value as flag, into the double-buffered publication array
// of the dataflow kernels); a box publishes its own x face, y face and top
// plane for the boxes after it.  Boxes are claimed from a ticket in
// wavefront order, so every box's producers are claimed before it.
//
// Plans exist only for full stencils (every row has exactly the parents its
// position implies), so the step needs no per-row masks: a parent that
// does not exist (a row on the grid's low faces) has coefficient -0.0 and
// value +0.0, and fma(-0.0, +0.0, s) == s for every s (the exact product is
// -0, and s + (-0) == s in round-to-nearest, s = -0 included): the sum is
// bit for bit the reference's sum over the entries present.
//
// (Precomputing the denominator-only half of the division's expansion per
// row was measured slower than the compiler's division in the step.)
#include <hip/hip_runtime.h>


#include "device.h"
#include "internal.h"

namespace baas {

struct PipeArgs {
    int n, nunits, nrhs, rhs;
    int nx, ny, nz, sy, sz;  // grid; sy < 0: a 2D grid (no y direction)
    int TX, TY, K;           // box shape
    const int4 *box;         // per unit: x0, y0, z0, tx | ty << 8 | kz << 16
    const int *rowptr, *colidx;
    int *pcount;
    Ctx ctx;
    const int *skip;
    DevStatus *status;
    long long spin_limit;
};

constexpr int PIPE_MAX_EXT = 1024;  // parents in other boxes per box (the plan keeps boxes within)
constexpr int PIPE_POLL_AHEAD = 16;  // steps ahead of the progress word the exchange polls
#ifndef PIPE_POLL_SLEEP
#define PIPE_POLL_SLEEP 1  // s_sleep units (64 clocks) between the exchange's polling passes
#endif

// Cross-lane moves.  shr1: the value of lane L - 1 (a wavefront-wide DPP
// shift; lane 0 keeps `old`).  lane_read: the value of the lane at byte
// address addr (lane * 4), through the LDS crossbar.
__device__ __forceinline__ int shr1_i(int old, int v) {
    return __builtin_amdgcn_update_dpp(old, v, 0x138, 0xf, 0xf, false);
}
__device__ __forceinline__ double shr1(double old, double v) {
    const long long b = __double_as_longlong(v), o = __double_as_longlong(old);
    const int lo = shr1_i((int)(o & 0xffffffff), (int)(b & 0xffffffff));
    const int hi = shr1_i((int)(o >> 32), (int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ float shr1(float old, float v) {
    return __int_as_float(shr1_i(__float_as_int(old), __float_as_int(v)));
}
__device__ __forceinline__ double lane_read(double v, int addr) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_ds_bpermute(addr, (int)(b & 0xffffffff));
    const int hi = __builtin_amdgcn_ds_bpermute(addr, (int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ float lane_read(float v, int addr) {
    return __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(v)));
}

// Grid coordinates -> row of this sweep (the backward sweep's coordinates
// count from the last row down).
template <bool UPPER>
__device__ __forceinline__ int pipe_row(const PipeArgs &a, int x, int y, int z) {
    const int key = x + (a.sy > 0 ? a.sy * y : 0) + a.sz * z;
    return UPPER ? a.n - 1 - key : key;
}

// Minimum over the wavefront, in every lane: DPP row shifts fold each row
// of 16 lanes into its last lane, row broadcasts fold the rows into lane
// 63, which is then read out (no LDS round trips).
__device__ __forceinline__ int wave_min(int v) {
    constexpr int BIG = 0x7fffffff;
    v = min(v, __builtin_amdgcn_update_dpp(BIG, v, 0x111, 0xf, 0xf, false));  // row_shr:1
    v = min(v, __builtin_amdgcn_update_dpp(BIG, v, 0x112, 0xf, 0xf, false));  // row_shr:2
    v = min(v, __builtin_amdgcn_update_dpp(BIG, v, 0x114, 0xf, 0xf, false));  // row_shr:4
    v = min(v, __builtin_amdgcn_update_dpp(BIG, v, 0x118, 0xf, 0xf, false));  // row_shr:8
    v = min(v, __builtin_amdgcn_update_dpp(BIG, v, 0x142, 0xa, 0xf, false));  // row_bcast:15
    v = min(v, __builtin_amdgcn_update_dpp(BIG, v, 0x143, 0xc, 0xf, false));  // row_bcast:31
    return __builtin_amdgcn_readlane(v, 63);
}

// The exchange wavefront: copy the box's parents in other boxes (item e
// of the list: row er[e], slot ev[e]) from the publication half into LDS as
// they are published, and keep *ready at the first pipeline step that
// needs an item not yet staged (nsteps once all are): the plane below's item
// (i, j) is needed at step i + j, an x-face item (j, plane k) at k + j, a
// y-face item (i, plane k) at k + i.  Each lane owns items lane, lane + 64,
// ..., at most R of them, and polls, in every pass, its missing ones needed
// within PIPE_POLL_AHEAD steps of the progress word, with a short sleep
// between passes (more frequent polling -- two passes in flight, or none
// of the sleep -- measured slower: the polls load the memory system the
// publications travel through).  A pass has no branches: items not polled
// load a fixed harmless address, and values not found are written to a
// dump slot.  A wait beyond the spin limit reports BAAS_TIMEOUT and stages
// zeros.
template <class T, int R>
__device__ __forceinline__ void pipe_stage(const T *pub, T *ev, const int *er, int ne, int nzf, int nxk, int nyk,
                                           int tx, int nsteps, int *ready, int lane, long long limit,
                                           DevStatus *status, int dump) {
    constexpr int BIG = 0x7fffffff;
    const int fs = nxk + nyk;
    int need[R], rows[R];
    unsigned todo = 0;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int e = lane + 64 * r;
        const bool have = e < ne;
        todo |= (have ? 1u : 0u) << r;
        rows[r] = have ? er[e] : 0;
        int nd = BIG;
        if (have) {
            if (e < nzf) {
                nd = e % tx + e / tx;
            } else {
                const int f = e - nzf, k = f / fs, w = f - k * fs;
                nd = k + (w < nxk ? w : w - nxk);
            }
        }
        need[r] = nd;
    }
    int last = -1;  // the progress word last written
    long long spins = 0;
    for (;;) {
        T v[R];
        unsigned pm = 0;
#pragma unroll
        for (int r = 0; r < R; r++) {
            const bool p = ((todo >> r) & 1u) & (need[r] < last + PIPE_POLL_AHEAD);
            pm |= (p ? 1u : 0u) << r;
            v[r] = ld_agent(&pub[p ? rows[r] : 0]);
        }
        int first = BIG;
#pragma unroll
        for (int r = 0; r < R; r++) {
            const bool found = ((pm >> r) & 1u) & !Sentinel<T>::is(v[r]);
            st_lds(&ev[found ? lane + 64 * r : dump], v[r]);
            todo &= ~((found ? 1u : 0u) << r);
            first = min(first, ((todo >> r) & 1u) ? need[r] : BIG);
        }
        first = min(wave_min(first), nsteps);
        // release: the staged values are in LDS before ready moves past them
        if (first != last && lane == 0)
            __hip_atomic_store(ready, first, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        last = first;
        if (first >= nsteps) break;
        if (++spins > limit) {
            raise_status(status, BAAS_TIMEOUT, last);
#pragma unroll
            for (int r = 0; r < R; r++)
                if ((todo >> r) & 1u) st_lds(&ev[lane + 64 * r], T(0));
            if (lane == 0) __hip_atomic_store(ready, nsteps, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
            break;
        }
        __builtin_amdgcn_s_sleep(PIPE_POLL_SLEEP);
    }
}

// The compute wavefront reads the progress word relaxed, then the staged
// slots: a wavefront's LDS operations complete in program order, and the
// exchange wavefront's release store orders its slot writes before the
// word, so slots read after a word past step t hold step t's parents.  The
// empty asm keeps the compiler from moving the slot reads above the word.
__device__ __forceinline__ int ld_ready(const int *ready) {
    const int r = __hip_atomic_load(const_cast<int *>(ready), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __asm__ volatile("" ::: "memory");
    return r;
}

// LDS of one box of at most RMX rows (compile-time, so every record read is
// one address register plus an immediate offset): per row the three parent
// coefficients (z, y, x), the diagonal, b and the solution, each in an
// array of RMX + 1 values (index RMX is the
// spare row idle lanes read and write); then the external parents' slots
// (PIPE_MAX_EXT values, a zero slot and a dump slot), their rows and the
// progress word.
template <class T, int RMX>
struct PipeLds {
    static constexpr int NS = RMX + 1;
    static constexpr size_t C0 = 0, C1 = C0 + NS * sizeof(T), C2 = C1 + NS * sizeof(T), DG = C2 + NS * sizeof(T),
                            BV = DG + NS * sizeof(T), XV = BV + NS * sizeof(T),
                            EV = XV + NS * sizeof(T), ER = EV + (PIPE_MAX_EXT + 2) * sizeof(T),
                            RDY = ER + PIPE_MAX_EXT * sizeof(int), TOTAL = RDY + 16;
    static constexpr int ZERO = PIPE_MAX_EXT;      // the zero slot of ev
    static constexpr int DUMP = PIPE_MAX_EXT + 1;  // where the exchange writes values it did not find
};

// One step's inputs of a lane, read a step ahead so that their LDS reads
// complete while the previous step computes: the progress word, the row's
// coefficients, diagonal and b, and the staged values of the
// parents it takes from other boxes (valid when the progress word read
// before them is past the step).
template <class T>
struct PipeRec {
    T cz, cy, cx, dd, bb, ez, ey, ex;
    int rdy;
};

// The compute wavefront.  Lane (i, j) = (lane / TY, lane % TY); its y parent
// is lane - 1 (shr1), its x parent lane - TY (shr1 when TY = 1, XS = 1: lane
// 0 then takes the x-face value as the shift's `old`; the LDS crossbar
// otherwise).  A lane without a row this step (before its first plane,
// after its last, outside a partial box) reads and writes the spare row,
// whose coefficients are -0, diagonal 1 and b +0, and so computes +0.
template <class T, bool UPPER, int XS, int RMX>
__device__ __forceinline__ void pipe_wave(const PipeArgs &a, int lane, int x0, int y0, int z0, int tx, int ty,
                                          int kz, int nzf, int nxk, int nyk, unsigned char *smem, const int *ready,
                                          T *pub, T *nxt) {
    using Lay = PipeLds<T, RMX>;
    const T *c0 = reinterpret_cast<const T *>(smem + Lay::C0);
    const T *c1 = reinterpret_cast<const T *>(smem + Lay::C1);
    const T *c2 = reinterpret_cast<const T *>(smem + Lay::C2);
    const T *dg = reinterpret_cast<const T *>(smem + Lay::DG);
    const T *bv = reinterpret_cast<const T *>(smem + Lay::BV);
    T *xv = reinterpret_cast<T *>(smem + Lay::XV);
    const T *ev = reinterpret_cast<const T *>(smem + Lay::EV);
    const int TY = a.TY;
    const int i = lane / TY, j = lane - (lane / TY) * TY;
    const bool lok = i < tx && j < ty;
    const int nsteps = kz + tx + ty - 2;
    const int txy = tx * ty;
    const bool pubx = x0 + tx < a.nx, puby = y0 + ty < a.ny, pubz = z0 + kz < a.nz;
    const bool face_xy = lok && ((i == tx - 1 && pubx) || (j == ty - 1 && puby));
    const int kpub = pubz ? kz - 1 : -1;
    const int xaddr = ((lane + 64 - TY) & 63) << 2;
    const int fs = nxk + nyk;
    // exchange slots: the plane below, and per plane the x face (lanes with
    // i = 0) or the y face (lanes with j = 0); the zero slot where a box
    // has no such face (a parent that does not exist has value +0)
    const bool fx = lok && i == 0, fy = lok && j == 0;
    const int zslot = lok && nzf > 0 ? i + tx * j : Lay::ZERO;
    const int xs0 = fx && nxk > 0 ? nzf + j : Lay::ZERO, xst = fx && nxk > 0 ? fs : 0;
    const int ys0 = fy && nyk > 0 ? nzf + nxk + i : Lay::ZERO, yst = fy && nyk > 0 ? fs : 0;
    const int qb = j * tx + i;
    auto slot = [&](int k) { return lok & ((unsigned)k < (unsigned)kz) ? qb + k * txy : RMX; };
    auto load_ext = [&](int k, PipeRec<T> &R) {
        const int kc = min(max(k, 0), kz - 1);
        R.ez = ev[zslot];
        if (XS != 1) R.ey = ev[ys0 + kc * yst];
        R.ex = ev[xs0 + kc * xst];
    };
    auto load = [&](int k, PipeRec<T> &R) {
        R.rdy = ld_ready(ready);
        const int q = slot(k);
        R.cz = c0[q];
        if (XS != 1) R.cy = c1[q];
        R.cx = c2[q];
        R.dd = dg[q];
        R.bb = bv[q];
        load_ext(k, R);
    };
    T xl = T(0);  // the value this lane solved at the previous step
    auto step = [&](int t, PipeRec<T> &C, PipeRec<T> &N) {
        const int k = t - i - j;
        if (__builtin_expect(t >= C.rdy, 0)) {
            // a parent from another box this step is not staged yet: wait
            // for the exchange wavefront, then read the staged values again
            long long spins = 0;
            int r = ld_ready(ready);
            while (r <= t) {
                __builtin_amdgcn_s_sleep(1);
                if (++spins > a.spin_limit) {
                    raise_status(a.status, BAAS_TIMEOUT, -1);
                    break;  // the staged slots then hold sentinels or zeros
                }
                r = ld_ready(ready);
            }
            load_ext(k, C);
        }
        // the next step's inputs, in flight while this step computes (after
        // the wait above, so the common path never waits on these reads)
        load(k + 1, N);
        // the parents: z from this lane, y from lane - 1, x from lane - TY
        // (staged values on the box's faces)
        // The lane moves run on every lane before any select: a cross-lane
        // read from a lane that is off in EXEC returns zero, so they must
        // not be sunk into the branch a select may become (the empty asm
        // pins them here).
        T vx = XS == 1 ? shr1(C.ex, xl) : lane_read(xl, xaddr);
        T vy = XS == 1 ? T(0) : shr1(T(0), xl);
        __asm__ volatile("" : "+v"(vx), "+v"(vy));
        const T pz = k == 0 ? C.ez : xl;
        const T px = XS == 1 ? vx : (fx ? C.ex : vx);
        // z, y, x: ascending columns for L, descending for U (a 2D grid,
        // XS = 1, has no y term)
        T s = UPPER ? fma(pz, C.cz, T(0)) : fma(C.cz, pz, T(0));
        if (XS != 1) {
            const T py = fy ? C.ey : vy;
            s = UPPER ? fma(py, C.cy, s) : fma(C.cy, py, s);
        }
        s = UPPER ? fma(px, C.cx, s) : fma(C.cx, px, s);
        xl = (C.bb - s) / C.dd;
        const int q = slot(k);
        xv[q] = xl;
        if (q != RMX && (face_xy || k == kpub)) {
            const int row = pipe_row<UPPER>(a, x0 + i, y0 + j, z0 + k);
            st_agent(&pub[row], xl);
            nxt[row] = Sentinel<T>::value();
        }
    };
    // two register sets used in turn, so no step copies one into the other
    PipeRec<T> A, B;
    load(-i - j, A);
    for (int t = 0; t < nsteps; t += 2) {
        step(t, A, B);
        if (t + 1 < nsteps) step(t + 1, B, A);
    }
}

template <class T, bool UPPER, int XS, int RMX, int R>
__global__ __launch_bounds__(128) void k_sweep_pipe(PipeArgs a, const T *__restrict__ vals, const T *__restrict__ b,
                                                    T *x, T *xp) {
    using Lay = PipeLds<T, RMX>;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ int s_unit;
    if (a.skip && ld_agent(a.skip)) return;
    const int epoch = read_epoch(a.ctx);
    const int par = ld_agent(a.pcount) & 1;
    T *pub = xp + (size_t)par * a.n;
    T *nxt = xp + (size_t)(par ^ 1) * a.n;
    T *c0 = reinterpret_cast<T *>(smem_raw + Lay::C0);
    T *c1 = reinterpret_cast<T *>(smem_raw + Lay::C1);
    T *c2 = reinterpret_cast<T *>(smem_raw + Lay::C2);
    T *dg = reinterpret_cast<T *>(smem_raw + Lay::DG);
    T *bv = reinterpret_cast<T *>(smem_raw + Lay::BV);
    const T *xv = reinterpret_cast<const T *>(smem_raw + Lay::XV);
    T *ev = reinterpret_cast<T *>(smem_raw + Lay::EV);
    int *er = reinterpret_cast<int *>(smem_raw + Lay::ER);
    int *ready = reinterpret_cast<int *>(smem_raw + Lay::RDY);
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nt = blockDim.x;
    // the spare row (idle lanes compute +0 on it) and the zero slot
    if (tid == 0) {
        c0[RMX] = c1[RMX] = c2[RMX] = -T(0);
        dg[RMX] = T(1);
        bv[RMX] = T(0);
        ev[Lay::ZERO] = T(0);
    }
    for (;;) {
        const int u = claim_unit(a.ctx.ticket, &s_unit);
        if (u >= a.nunits) break;
        const int4 bx = a.box[u];
        const int x0 = bx.x, y0 = bx.y, z0 = bx.z;
        const int tx = bx.w & 255, ty = (bx.w >> 8) & 255, kz = (bx.w >> 16) & 255;
        const int nrow = tx * ty * kz, txy = tx * ty, nsteps = kz + tx + ty - 2;
        // parents in other boxes: the plane below, then per plane the x face
        // (ty rows) and the y face (tx rows)
        const int nzf = z0 > 0 ? txy : 0, nxk = x0 > 0 ? ty : 0, nyk = (a.sy > 0 && y0 > 0) ? tx : 0;
        const int ne = nzf + kz * (nxk + nyk);
        for (int e = tid; e < ne; e += nt) {
            int xx, yy, zz;
            if (e < nzf) {
                xx = x0 + e % tx;
                yy = y0 + e / tx;
                zz = z0 - 1;
            } else {
                const int f = e - nzf, k = f / (nxk + nyk), w = f - k * (nxk + nyk);
                zz = z0 + k;
                if (w < nxk) {
                    xx = x0 - 1;
                    yy = y0 + w;
                } else {
                    xx = x0 + (w - nxk);
                    yy = y0 - 1;
                }
            }
            er[e] = pipe_row<UPPER>(a, xx, yy, zz);
            st_lds(&ev[e], Sentinel<T>::value());
        }
        if (tid == 0) *ready = ne > 0 ? 0 : nsteps;
        // the box's rows: eight per thread per pass, every load of a pass in
        // flight together (row pointers and b, then the entries); a parent
        // the row lacks gets coefficient -0
        constexpr int PB = 8;
        for (int q0 = tid; q0 < nrow; q0 += PB * nt) {
            int row[PB], p0[PB], p1[PB];
            T bb[PB];
#pragma unroll
            for (int v = 0; v < PB; v++) {
                const int q = min(q0 + v * nt, nrow - 1);
                const int ii = q % tx, r2 = q / tx, jj = r2 % ty, kk = r2 / ty;
                row[v] = pipe_row<UPPER>(a, x0 + ii, y0 + jj, z0 + kk);
                p0[v] = a.rowptr[row[v]];
                p1[v] = a.rowptr[row[v] + 1];
                bb[v] = b[(size_t)row[v] * a.nrhs + a.rhs];
            }
            int cc[PB][4];
            T cv[PB][4];
#pragma unroll
            for (int v = 0; v < PB; v++)
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const int idx = min(p0[v] + e, p1[v] - 1);
                    cc[v][e] = a.colidx[idx];
                    cv[v][e] = vals[idx];
                }
#pragma unroll
            for (int v = 0; v < PB; v++) {
                const int q = q0 + v * nt;
                if (q >= nrow) continue;
                const int cnt = p1[v] - p0[v];
                T vz = -T(0), vy = -T(0), vx = -T(0), vd = T(1);
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    if (e >= cnt) continue;
                    const int d = UPPER ? cc[v][e] - row[v] : row[v] - cc[v][e];
                    if (d == 0)
                        vd = cv[v][e];
                    else if (d == a.sz)
                        vz = cv[v][e];
                    else if (d == a.sy)
                        vy = cv[v][e];
                    else if (d == 1)
                        vx = cv[v][e];
                }
                c0[q] = vz;
                c1[q] = vy;
                c2[q] = vx;
                dg[q] = vd;
                bv[q] = bb[v];
            }
        }
        __syncthreads();
        if (wv == 0)
            pipe_stage<T, R>(pub, ev, er, ne, nzf, nxk, nyk, tx, nsteps, ready, lane, a.spin_limit, a.status,
                             Lay::DUMP);
        else if (wv == 1)
            pipe_wave<T, UPPER, XS, RMX>(a, lane, x0, y0, z0, tx, ty, kz, nzf, nxk, nyk, smem_raw, ready, pub, nxt);
        __syncthreads();
        for (int q = tid; q < nrow; q += nt) {
            const int ii = q % tx, r2 = q / tx, jj = r2 % ty, kk = r2 / ty;
            const int row = pipe_row<UPPER>(a, x0 + ii, y0 + jj, z0 + kk);
            x[(size_t)row * a.nrhs + a.rhs] = xv[q];
        }
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

// Rows per box the kernel instances hold (the plan keeps TX TY K within).
static int pipe_rmx(const Plan &p) { return p.pg[5] * p.pg[6] * p.pg[7] <= 1024 ? 1024 : 2048; }

// LDS bytes the pipelined sweep needs per workgroup on plan p.
size_t pipe_lds(const Plan &p, int ts) {
    if (p.pg[5] * p.pg[6] * p.pg[7] > 2048) return (size_t)1 << 40;  // no instance holds such boxes
    const bool r1 = pipe_rmx(p) == 1024;
    return ts == 8 ? (r1 ? PipeLds<double, 1024>::TOTAL : PipeLds<double, 2048>::TOTAL)
                   : (r1 ? PipeLds<float, 1024>::TOTAL : PipeLds<float, 2048>::TOTAL);
}

template <class T, int RMX, int R>
static int launch_pipe_r(baas_handle *h, const PipeArgs &a0, int uplo, const T *vals, const T *b, T *x, T *xp,
                         int nunits, int nrhs, hipStream_t stream) {
    // the x parent sits TY lanes down: a DPP shift for TY = 1 (2D grids),
    // the LDS crossbar otherwise
    using K = void (*)(PipeArgs, const T *, const T *, T *, T *);
    K kern = a0.TY == 1 ? (uplo ? k_sweep_pipe<T, true, 1, RMX, R> : k_sweep_pipe<T, false, 1, RMX, R>)
                        : (uplo ? k_sweep_pipe<T, true, 0, RMX, R> : k_sweep_pipe<T, false, 0, RMX, R>);
    const size_t lds = PipeLds<T, RMX>::TOTAL;
    int per_cu = 0;
    BAAS_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 128, lds));
    if (per_cu <= 0) return BAAS_UNSUPPORTED;
    long long grid = std::min<long long>((long long)per_cu * h->num_cus, nunits);
    if (grid < 1) grid = 1;
    PipeArgs a = a0;
    for (int r = 0; r < nrhs; r++) {
        a.rhs = r;
        hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(128), lds, stream, a, vals, b, x, xp);
        BAAS_HIP_TRY(hipGetLastError());
    }
    return BAAS_OK;
}

// One launch per right-hand side (b and x are n x nrhs, row-major).
template <class T>
int launch_pipe(baas_handle *h, baas_schedule *s, const Plan &p, int uplo, const T *vals, const T *b, T *x, T *xp,
                int nrhs, hipStream_t stream, const int *skip) {
    PipeArgs a;
    a.n = s->n;
    a.nunits = p.nunits;
    a.nrhs = nrhs;
    a.rhs = 0;
    a.nx = p.pg[0];
    a.ny = p.pg[1];
    a.nz = p.pg[2];
    a.sy = p.pg[3];
    a.sz = p.pg[4];
    a.TX = p.pg[5];
    a.TY = p.pg[6];
    a.K = p.pg[7];
    a.box = reinterpret_cast<const int4 *>(p.pipe_box.p);
    a.rowptr = uplo ? s->u_rowptr.p : s->l_rowptr.p;
    a.colidx = uplo ? s->u_colidx.p : s->l_colidx.p;
    a.pcount = p.pcount.p;
    a.ctx = Ctx{h->counters, h->counters + 1, h->counters + 2};
    a.skip = skip;
    a.status = h->status;
    a.spin_limit = h->spin_limit;
    if (a.TX * a.TY * a.K > 2048) return BAAS_UNSUPPORTED;
    // items per exchange lane: the largest box's parents in other boxes / 64
    const int nem = (int)pipe_ext_max(a.TX, a.TY, a.K, a.sy > 0);
    if (nem > PIPE_MAX_EXT) return BAAS_UNSUPPORTED;
    const bool r1 = pipe_rmx(p) == 1024;
    if (nem <= 128)
        return r1 ? launch_pipe_r<T, 1024, 2>(h, a, uplo, vals, b, x, xp, p.nunits, nrhs, stream)
                  : launch_pipe_r<T, 2048, 2>(h, a, uplo, vals, b, x, xp, p.nunits, nrhs, stream);
    if (nem <= 256)
        return r1 ? launch_pipe_r<T, 1024, 4>(h, a, uplo, vals, b, x, xp, p.nunits, nrhs, stream)
                  : launch_pipe_r<T, 2048, 4>(h, a, uplo, vals, b, x, xp, p.nunits, nrhs, stream);
    if (nem <= 512)
        return r1 ? launch_pipe_r<T, 1024, 8>(h, a, uplo, vals, b, x, xp, p.nunits, nrhs, stream)
                  : launch_pipe_r<T, 2048, 8>(h, a, uplo, vals, b, x, xp, p.nunits, nrhs, stream);
    return r1 ? launch_pipe_r<T, 1024, 16>(h, a, uplo, vals, b, x, xp, p.nunits, nrhs, stream)
              : launch_pipe_r<T, 2048, 16>(h, a, uplo, vals, b, x, xp, p.nunits, nrhs, stream);
}

template int launch_pipe<double>(baas_handle *, baas_schedule *, const Plan &, int, const double *, const double *,
                                 double *, double *, int, hipStream_t, const int *);
template int launch_pipe<float>(baas_handle *, baas_schedule *, const Plan &, int, const float *, const float *,
                                float *, float *, int, hipStream_t, const int *);

}  // namespace baas
