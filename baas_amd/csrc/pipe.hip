// pipe.hip -- the wave-pipelined sweep of stencil-grid triangular matrices
// (PIPE plans, dataflow 5).
//
// Same arithmetic as the reference's solves (merge/SpTRSV.cpp:24-39 forward,
// :601-660 backward with descending sums): x_r = (b_r - sum_j L_rj x_j) /
// L_rr, the sum in ascending column order for L (descending for U = L^T),
// one fma per term.  On a 7-point (5-point) grid in natural order the
// parents of row (x, y, z) are (x-1, y, z), (x, y-1, z) and (x, y, z-1),
// so the rows of level l are those with x + y + z = l.
//
// A work unit is a box of TX x TY rows across (x, y) by kz planes along z,
// swept by ONE wavefront: lane (i, j) owns the column (x0+i, y0+j) and at
// step t solves its row of plane k = t - i - j.  Its three parents were all
// solved at step t-1: the x parent by lane (i-1, j), the y parent by lane
// (i, j-1) -- both read with a cross-lane permute -- and the z parent by the
// lane itself.  Inside a box a dependency therefore costs one step (a
// permute, three fma and the IEEE division) instead of an LDS hand-off and
// a poll, and every lane is busy once the pipeline has filled.  Parents in
// other boxes (the plane below, the x and y faces) are staged into LDS by
// the workgroup's second wavefront as their producers publish them (value
// as flag, as in the lane kernel); a box publishes its own x face, y face
// and top plane for the boxes after it.  Boxes are claimed from a ticket in
// wavefront order (every box's producers are claimed before it).
//
// LDS per box (pipe_layout): the rows' coefficients, b and mask, preloaded
// when the box is claimed, and the solution, written out at the end.
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

// Cross-lane moves of the pipeline.  lane_read: the value of the lane at
// byte address addr (lane * 4), through the LDS crossbar.  shr1: the value
// of lane L - 1 (one DPP move per 32 bits, a wavefront-wide shift).  Lanes
// whose source lies below lane 0 get values they do not use.
__device__ __forceinline__ int lane_read_i(int v, int addr) { return __builtin_amdgcn_ds_bpermute(addr, v); }
__device__ __forceinline__ int shr1_i(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, false); }

// The x parent's move: lane L - 1 (XS = 1) or the lane at addr (XS = 0).
template <int XS>
__device__ __forceinline__ int xshift_i(int v, int addr) {
    return XS == 1 ? shr1_i(v) : lane_read_i(v, addr);
}
template <int XS>
__device__ __forceinline__ double xshift(double v, int addr) {
    const long long b = __double_as_longlong(v);
    const int lo = xshift_i<XS>((int)(b & 0xffffffff), addr);
    const int hi = xshift_i<XS>((int)(b >> 32), addr);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <int XS>
__device__ __forceinline__ float xshift(float v, int addr) {
    return __int_as_float(xshift_i<XS>(__float_as_int(v), addr));
}
__device__ __forceinline__ double shr1(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = shr1_i((int)(b & 0xffffffff)), hi = shr1_i((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ float shr1(float v) { return __int_as_float(shr1_i(__float_as_int(v))); }

// Grid key of box-local (i, j, k) -> row of this sweep (the backward sweep's
// keys count from the last row down).
template <bool UPPER>
__device__ __forceinline__ int pipe_row(const PipeArgs &a, int x, int y, int z) {
    const int key = x + (a.sy > 0 ? a.sy * y : 0) + a.sz * z;
    return UPPER ? a.n - 1 - key : key;
}

// The exchange wavefront: copy the box's parents in other boxes (item e
// of the list: row er[e], slot ev[e]) from the publication half into LDS as
// they are published, and keep *ready at the first pipeline step that
// needs an item not yet staged (nsteps once all are): the plane below's item
// (i, j) is needed at step i + j, an x-face item (j, plane k) at k + j, a
// y-face item (i, plane k) at k + i.  Each lane owns items lane, lane + 64,
// ... and polls all of its missing ones in every pass.  A wait beyond the
// spin limit reports BAAS_TIMEOUT and stages zeros.
constexpr int PIPE_MAX_EXT = 1024;  // items per box (the plan keeps K small enough)
template <class T>
__device__ __forceinline__ void pipe_stage(const T *pub, T *ev, const int *er, int ne, int nzf, int nxk, int nyk,
                                           int tx, int nsteps, int *ready, int lane, long long limit,
                                           DevStatus *status) {
    constexpr int R = PIPE_MAX_EXT / 64;
    const int fs = nxk + nyk;
    int need[R], rows[R];
    unsigned todo = 0;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int e = lane + 64 * r;
        if (e < ne) {
            todo |= 1u << r;
            rows[r] = er[e];
            if (e < nzf) {
                need[r] = e % tx + e / tx;
            } else {
                const int f = e - nzf, k = f / fs, w = f - k * fs;
                need[r] = k + (w < nxk ? w : w - nxk);
            }
        } else {
            rows[r] = 0;
            need[r] = nsteps;
        }
    }
    long long spins = 0;
    for (;;) {
        T v[R];
#pragma unroll
        for (int r = 0; r < R; r++)
            if ((todo >> r) & 1u) v[r] = ld_agent(&pub[rows[r]]);
        int first = nsteps;
#pragma unroll
        for (int r = 0; r < R; r++) {
            if ((todo >> r) & 1u) {
                if (!Sentinel<T>::is(v[r])) {
                    st_lds(&ev[lane + 64 * r], v[r]);
                    todo &= ~(1u << r);
                } else {
                    first = min(first, need[r]);
                }
            }
        }
        // the wavefront's earliest step still waiting
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) first = min(first, __shfl_xor(first, d));
        // release: the staged values are in LDS before ready moves past them
        if (lane == 0) __hip_atomic_store(ready, first, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (first >= nsteps) break;
        if (++spins > limit) {
            raise_status(status, BAAS_TIMEOUT, first);
#pragma unroll
            for (int r = 0; r < R; r++)
                if ((todo >> r) & 1u) st_lds(&ev[lane + 64 * r], T(0));
            if (lane == 0) __hip_atomic_store(ready, nsteps, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

__device__ __forceinline__ int ld_ready(const int *ready) {
    return __hip_atomic_load(const_cast<int *>(ready), __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// One step's inputs of a lane, read a step ahead so that their LDS reads
// complete while the previous step computes: the exchange progress word,
// then its row's coefficients, b and parent mask, and the staged values of
// the parents it may take from other boxes (valid when the progress word
// read before them is past the step).
template <class T>
struct PipeRec {
    T cz, cy, cx, dd, bb, ez, ey, ex;
    int mq, rdy;
};

// The compute wavefront: sweep the box as a diagonal pipeline.  Lane
// (i, j) = (lane / TY, lane % TY) solves plane k = t - i - j at step t; its
// y parent is lane - 1 (a DPP shift), its x parent lane - TY (a DPP shift
// when TY = 1, XS = 1; the LDS crossbar otherwise), its z parent its own
// value of the previous step.  A lane without a row this step (before its
// first plane, after its last, or outside a partial box) computes on
// harmless data and stores to the spare slot: its value is never read,
// since its neighbours have no row at the next step either.
template <class T, bool UPPER, int XS>
__device__ __forceinline__ void pipe_wave(const PipeArgs &a, int lane, int x0, int y0, int z0, int tx, int ty,
                                          int kz, int nzf, int nxk, int nyk, const PipeLayout &lay,
                                          unsigned char *smem, const int *ready, T *pub, T *nxt) {
    const T *c0 = reinterpret_cast<const T *>(smem + lay.c0);
    const T *c1 = reinterpret_cast<const T *>(smem + lay.c1);
    const T *c2 = reinterpret_cast<const T *>(smem + lay.c2);
    const T *c3 = reinterpret_cast<const T *>(smem + lay.c3);
    const T *bv = reinterpret_cast<const T *>(smem + lay.bv);
    T *xv = reinterpret_cast<T *>(smem + lay.xv);
    const int *msk = reinterpret_cast<const int *>(smem + lay.msk);
    const T *ev = reinterpret_cast<const T *>(smem + lay.ev);
    const int TY = a.TY;
    const int i = lane / TY, j = lane - (lane / TY) * TY;
    const bool lok = i < tx && j < ty;
    const int nsteps = kz + tx + ty - 2;
    const int nrow = tx * ty * kz, txy = tx * ty;
    const bool pubx = x0 + tx < a.nx, puby = y0 + ty < a.ny, pubz = z0 + kz < a.nz;
    const bool face_xy = lok && ((i == tx - 1 && pubx) || (j == ty - 1 && puby));
    const int kpub = pubz ? kz - 1 : -1;
    const int xaddr = ((lane + 64 - TY) & 63) << 2;
    const int fstride = nxk + nyk;
    // exchange slots: the plane below (z), per plane the x face (lanes with
    // i = 0) and the y face (lanes with j = 0); slot 0 for lanes off them
    const int zslot = lok ? i + tx * j : 0;
    const bool onx = lok && i == 0 && nxk > 0, ony = lok && j == 0 && nyk > 0;
    const int xs0 = onx ? nzf + j : 0, xst = onx ? fstride : 0;
    const int ys0 = ony ? nzf + nxk + i : 0, yst = ony ? fstride : 0;
    const int qb = j * tx + i;
    auto load_ext = [&](int k, PipeRec<T> &R) {
        const bool act = lok & ((unsigned)k < (unsigned)kz);
        const int kc = act ? k : 0;
        R.ez = ev[kc == 0 ? zslot : 0];
        R.ey = ev[ys0 + kc * yst];
        R.ex = ev[xs0 + kc * xst];
    };
    auto load = [&](int k, PipeRec<T> &R) {
        R.rdy = ld_ready(ready);
        const bool act = lok & ((unsigned)k < (unsigned)kz);
        const int p = pipe_slot(act ? qb + k * txy : nrow);
        R.cz = c0[p];
        R.cy = c1[p];
        R.cx = c2[p];
        R.dd = c3[p];
        R.bb = bv[p];
        R.mq = msk[p];
        load_ext(k, R);
    };
    T xl = T(0);  // the value this lane solved at the previous step
    auto step = [&](int t, PipeRec<T> &C, PipeRec<T> &N) {
        const int k = t - i - j;
        load(k + 1, N);
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
        const bool act = lok & ((unsigned)k < (unsigned)kz);
        const int p = pipe_slot(act ? qb + k * txy : nrow);
        const int m = act ? C.mq : 0;
        const bool wz = ((m & 1) != 0) & (k == 0), wy = ((m & 2) != 0) & ony, wx = ((m & 4) != 0) & onx;
        // the neighbours' values of the previous step: y and x parents
        const T vy = shr1(xl);
        const T vx = xshift<XS>(xl, xaddr);
        const T pz = wz ? C.ez : xl, py = wy ? C.ey : vy, px = wx ? C.ex : vx;
        // z, y, x: ascending columns for L, descending for U
        T s = T(0);
        s = (m & 1) ? (UPPER ? fma(pz, C.cz, s) : fma(C.cz, pz, s)) : s;
        s = (m & 2) ? (UPPER ? fma(py, C.cy, s) : fma(C.cy, py, s)) : s;
        s = (m & 4) ? (UPPER ? fma(px, C.cx, s) : fma(C.cx, px, s)) : s;
        xl = (C.bb - s) / C.dd;
        xv[p] = xl;
        const bool pubme = act & (face_xy | (k == kpub));
        if (__ballot(pubme) != 0 && pubme) {
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

template <class T, bool UPPER, int XS>
__global__ __launch_bounds__(128) void k_sweep_pipe(PipeArgs a, const T *__restrict__ vals, const T *__restrict__ b,
                                                    T *x, T *xp) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ int s_unit;
    if (a.skip && ld_agent(a.skip)) return;
    const int epoch = read_epoch(a.ctx);
    const int par = ld_agent(a.pcount) & 1;
    T *pub = xp + (size_t)par * a.n;
    T *nxt = xp + (size_t)(par ^ 1) * a.n;
    const PipeLayout lay = pipe_layout(a.TX, a.TY, a.K, sizeof(T), a.sy > 0);
    T *c0 = reinterpret_cast<T *>(smem_raw + lay.c0);
    T *c1 = reinterpret_cast<T *>(smem_raw + lay.c1);
    T *c2 = reinterpret_cast<T *>(smem_raw + lay.c2);
    T *c3 = reinterpret_cast<T *>(smem_raw + lay.c3);
    T *bv = reinterpret_cast<T *>(smem_raw + lay.bv);
    const T *xv = reinterpret_cast<const T *>(smem_raw + lay.xv);
    int *msk = reinterpret_cast<int *>(smem_raw + lay.msk);
    T *ev = reinterpret_cast<T *>(smem_raw + lay.ev);
    int *er = reinterpret_cast<int *>(smem_raw + lay.er);
    __shared__ int s_ready;  // exchange progress: the first step still waiting for a staged parent
    int *ready = &s_ready;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nt = blockDim.x;
    const int syk = a.sy > 0 ? a.sy : 0;
    for (;;) {
        const int u = claim_unit(a.ctx.ticket, &s_unit);
        if (u >= a.nunits) break;
        const int4 bx = a.box[u];
        const int x0 = bx.x, y0 = bx.y, z0 = bx.z;
        const int tx = bx.w & 255, ty = (bx.w >> 8) & 255, kz = (bx.w >> 16) & 255;
        const int nrow = tx * ty * kz, txy = tx * ty;
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
        if (tid == 0) *ready = ne > 0 ? 0 : kz + tx + ty - 2;
        // the box's rows: four per thread per pass, every load of a pass in
        // flight together (row pointers and b, then the entries)
        constexpr int PB = 8;
        for (int q0 = tid; q0 < nrow; q0 += PB * nt) {
            int row[PB], p0[PB], p1[PB];
            T bb[PB];
#pragma unroll
            for (int v = 0; v < PB; v++) {
                const int q = min(q0 + v * nt, nrow - 1);
                const int ii = q % tx, r2 = q / tx, jj = r2 % ty, kk = r2 / ty;
                row[v] = UPPER ? a.n - 1 - (x0 + ii + syk * (y0 + jj) + a.sz * (z0 + kk))
                               : x0 + ii + syk * (y0 + jj) + a.sz * (z0 + kk);
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
                T z0c = T(0), y0c = T(0), x0c = T(0), dg = T(1);
                int m = 0;
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    if (e >= cnt) continue;
                    const int d = UPPER ? cc[v][e] - row[v] : row[v] - cc[v][e];
                    if (d == 0) {
                        dg = cv[v][e];
                    } else if (d == a.sz) {
                        z0c = cv[v][e];
                        m |= 1;
                    } else if (d == a.sy) {
                        y0c = cv[v][e];
                        m |= 2;
                    } else if (d == 1) {
                        x0c = cv[v][e];
                        m |= 4;
                    }
                }
                const int ps = pipe_slot(q);
                c0[ps] = z0c;
                c1[ps] = y0c;
                c2[ps] = x0c;
                c3[ps] = dg;
                bv[ps] = bb[v];
                msk[ps] = m;
            }
        }
        __syncthreads();
        if (wv == 0)
            pipe_stage<T>(pub, ev, er, ne, nzf, nxk, nyk, tx, kz + tx + ty - 2, ready, lane, a.spin_limit,
                          a.status);
        else if (wv == 1)
            pipe_wave<T, UPPER, XS>(a, lane, x0, y0, z0, tx, ty, kz, nzf, nxk, nyk, lay, smem_raw, ready, pub,
                                    nxt);
        __syncthreads();
        for (int q = tid; q < nrow; q += nt) {
            const int ii = q % tx, r2 = q / tx, jj = r2 % ty, kk = r2 / ty;
            const int row = pipe_row<UPPER>(a, x0 + ii, y0 + jj, z0 + kk);
            x[(size_t)row * a.nrhs + a.rhs] = xv[pipe_slot(q)];
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

// LDS bytes the pipelined sweep needs per workgroup on plan p.
size_t pipe_lds(const Plan &p, int ts) { return pipe_layout(p.pg[5], p.pg[6], p.pg[7], ts, p.pg[3] > 0).total; }

// One launch per right-hand side (b and x are n x nrhs, row-major).
template <class T>
int launch_pipe(baas_handle *h, baas_schedule *s, const Plan &p, int uplo, const T *vals, const T *b, T *x, T *xp,
                int nrhs, hipStream_t stream, const int *skip) {
    PipeArgs a;
    a.n = s->n;
    a.nunits = p.nunits;
    a.nrhs = nrhs;
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
    const size_t lds = pipe_lds(p, (int)sizeof(T));
    // the x parent sits TY lanes down: a DPP shift for TY = 1 (2D grids),
    // the LDS crossbar otherwise (measured faster than the row rotation and
    // swaps that also reach lane - 8)
    using K = void (*)(PipeArgs, const T *, const T *, T *, T *);
    K kern = a.TY == 1 ? (uplo ? k_sweep_pipe<T, true, 1> : k_sweep_pipe<T, false, 1>)
                       : (uplo ? k_sweep_pipe<T, true, 0> : k_sweep_pipe<T, false, 0>);
    int per_cu = 0;
    BAAS_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 128, lds));
    if (per_cu <= 0) return BAAS_UNSUPPORTED;
    long long grid = std::min<long long>((long long)per_cu * h->num_cus, p.nunits);
    if (grid < 1) grid = 1;
    for (int r = 0; r < nrhs; r++) {
        a.rhs = r;
        hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(128), lds, stream, a, vals, b, x, xp);
        BAAS_HIP_TRY(hipGetLastError());
    }
    return BAAS_OK;
}

template int launch_pipe<double>(baas_handle *, baas_schedule *, const Plan &, int, const double *, const double *,
                                 double *, double *, int, hipStream_t, const int *);
template int launch_pipe<float>(baas_handle *, baas_schedule *, const Plan &, int, const float *, const float *,
                                float *, float *, int, hipStream_t, const int *);

}  // namespace baas
