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
// publish them.  A box publishes its x face, y face and top plane for the
// boxes after it (value as flag, into a double-buffered publication array)
// in the order its consumer needs them: a face by diagonals, so that one
// step of the producer writes, and one step of the consumer needs, one run
// of consecutive values.  Boxes are claimed from a ticket in wavefront
// order, so every box's producers are claimed before it.
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

size_t lds_limit(baas_handle *h);

struct PipeArgs {
    int n, nunits, nrhs;  // b and x are n x nrhs, row-major
    int nx, ny, nz, sy, sz;  // grid; sy < 0: a 2D grid (no y direction)
    int TX, TY, K;           // box shape
    long long half;          // values per publication half: nrhs regions of rstride
    long long rstride;       // one right-hand side's box faces
    int nq;                  // ticket queues: 1 (the context's ticket) or one per XCD
    int *qticket;            // queue q's ticket at qticket[32 q]
    // per unit two records: x0, y0, z0, tx | ty << 8 | kz << 16; then the
    // offsets of the box's own faces, of the x face before it, of the y
    // face before it and of the top plane below it (-1: none)
    const int4 *box;
    int nnz;  // entries of L (and of U = L^T)
    int *pcount;
    Ctx ctx;
    const int *skip;
    DevStatus *status;
    long long spin_limit;
};

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
// The same shift where lane 0's result does not matter (it reads 0).
__device__ __forceinline__ int shr1z_i(int v) { return __builtin_amdgcn_mov_dpp(v, 0x138, 0xf, 0xf, true); }
__device__ __forceinline__ double shr1z(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = shr1z_i((int)(b & 0xffffffff)), hi = shr1z_i((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ float shr1z(float v) { return __int_as_float(shr1z_i(__float_as_int(v))); }
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

// The exchange wavefront: copy the faces the box reads from the boxes
// before it (ex: the x face, ey: the y face, ez: the plane below; px, py,
// pz in the publication half, null where there is none) into LDS as they
// are published, and keep *ready at the first pipeline step that needs a
// value not yet staged (nsteps once all are).  The faces come in the order
// they are needed: x-face value c at step c / ty, y-face value c at c / tx;
// of the plane below, value i + tx j at step i + j.  For each face a cursor
// marks the first value not yet seen; every pass loads the 64 values from
// the cursor (rounded down to a multiple of 8) in one coalesced load, and
// the whole plane below (one value a lane) until it is complete; values
// found are written to their slots again, which changes nothing.  A pass
// has no branches: lanes with nothing to load read `harmless`, and values
// not found are written to the dump slot.  A short sleep separates passes
// (more frequent polling loads the memory system the publications travel
// through).  A wait beyond the spin limit reports BAAS_TIMEOUT and stages
// zeros.  Once every value is staged (or the wait gave up), the faces are
// re-armed with the sentinel in the other publication half (`other`
// values away), so the producers' steps carry no second store.  With NR
// right-hand sides a face value is NR consecutive values, in the
// publication half and in LDS alike, and it is found when all NR are.
template <class T, int NR = 1>
__device__ __forceinline__ void pipe_stage(const T *px, int nxf, int ty, const T *py, int nyf, int tx, const T *pz,
                                           int nzf, T *ex, T *ey, T *ez, T *dump, const T *harmless, int nsteps,
                                           int *ready, int lane, long long limit, DevStatus *status, long long other) {
    constexpr int BIG = 0x7fffffff;
    if (!px) nxf = 0;
    if (!py) nyf = 0;
    bool zt = pz && lane < nzf;  // this lane's value of the plane below is still missing
    const int zneed = lane % tx + lane / tx;
    int cx = 0, cy = 0;
    int last = -1;  // the progress word last written
    long long spins = 0;
    for (;;) {
        const int wx = cx & ~7, wy = cy & ~7;
        const bool qx = (cx < nxf) & (wx + lane < nxf), qy = (cy < nyf) & (wy + lane < nyf);
        T vx[NR], vy[NR], vz[NR];
        bool fx = qx, fy = qy, fz = zt;
#pragma unroll
        for (int r = 0; r < NR; r++) {
            vx[r] = ld_agent(qx ? px + (wx + lane) * NR + r : harmless);
            vy[r] = ld_agent(qy ? py + (wy + lane) * NR + r : harmless);
            vz[r] = ld_agent(zt ? pz + lane * NR + r : harmless);
        }
#pragma unroll
        for (int r = 0; r < NR; r++) {
            fx = fx & !Sentinel<T>::is(vx[r]);
            fy = fy & !Sentinel<T>::is(vy[r]);
            fz = fz & !Sentinel<T>::is(vz[r]);
        }
#pragma unroll
        for (int r = 0; r < NR; r++) {
            st_lds(fx ? ex + (wx + lane) * NR + r : dump, vx[r]);
            st_lds(fy ? ey + (wy + lane) * NR + r : dump, vy[r]);
            st_lds(fz ? ez + lane * NR + r : dump, vz[r]);
        }
        zt = zt & !fz;
        const unsigned long long mx = __ballot(qx & !fx), my = __ballot(qy & !fy);
        cx = mx ? wx + (int)__builtin_ctzll(mx) : min(wx + 64, nxf);
        cy = my ? wy + (int)__builtin_ctzll(my) : min(wy + 64, nyf);
        int first = nsteps;
        if (cx < nxf) first = min(first, cx / ty);
        if (cy < nyf) first = min(first, cy / tx);
        if (__ballot(zt)) first = min(first, wave_min(zt ? zneed : BIG));
        // release: the staged values are in LDS before ready moves past them
        if (first != last && lane == 0)
            __hip_atomic_store(ready, first, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        last = first;
        if (first >= nsteps) break;
        if (++spins > limit) {
            raise_status(status, BAAS_TIMEOUT, last);
            for (int c = cx * NR + lane; c < nxf * NR; c += 64) st_lds(ex + c, T(0));
            for (int c = cy * NR + lane; c < nyf * NR; c += 64) st_lds(ey + c, T(0));
            if (zt)
                for (int r = 0; r < NR; r++) st_lds(ez + lane * NR + r, T(0));
            if (lane == 0) __hip_atomic_store(ready, nsteps, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
            break;
        }
        __builtin_amdgcn_s_sleep(PIPE_POLL_SLEEP);
    }
    // every value of these faces is read by this box alone: re-arm them in
    // the other half, which the next sweep publishes into
    for (int c = lane; c < nxf * NR; c += 64) const_cast<T *>(px)[other + c] = Sentinel<T>::value();
    for (int c = lane; c < nyf * NR; c += 64) const_cast<T *>(py)[other + c] = Sentinel<T>::value();
    if (pz)
        for (int c = lane; c < nzf * NR; c += 64) const_cast<T *>(pz)[other + c] = Sentinel<T>::value();
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

// What the pipelined sweep computes: a triangular solve, or the stencil
// IC0 or ILU0 factorisation (the published value is then the row's
// diagonal of L, or of U).
enum { PIPE_SOLVE = 0, PIPE_IC0 = 1, PIPE_ILU0 = 2 };

// LDS of one box of at most RMX rows (compile-time, so every record read is
// one address register plus an immediate offset), in arrays of RMX + 1
// values (index RMX is the spare row idle lanes read and write): per row
// the three parent coefficients (z, y, x) and the diagonal; for a solve b,
// whose slot then takes the solution (a row's b is read a step before its
// value is written; idle lanes write the spare row's b, 0, back); for ILU0
// the parents' entries in this row's column (z, y, x).  A factorisation
// writes its results over the coefficients and the diagonal.  Then the
// staged faces (EX: the x face, EY: the y face, each in the order of its
// publication and of at most RMX / 2 values, EZ: the plane below), the
// slot parents outside the grid read (1) and a dump slot; then the
// progress word.  A solve of NR right-hand sides at once keeps NR values
// per row in b's array and per slot in the faces (slot s at s NR .. s NR +
// NR - 1).
template <class T, int RMX, int MODE = PIPE_SOLVE, int NR = 1>
struct PipeLds {
    static constexpr int NS = RMX + 1, FMX = RMX / 2;
    static constexpr int NA = MODE == PIPE_SOLVE ? 4 + NR : (MODE == PIPE_IC0 ? 4 : 7);
    static constexpr int EX = 0, EY = FMX, EZ = 2 * FMX, NONE = EZ + 64, DUMP = NONE + 1, NE = DUMP + 1;
    static constexpr size_t A = NS * sizeof(T);
    static constexpr size_t C0 = 0, C1 = A, C2 = 2 * A, DG = 3 * A, BV = 4 * A, XV = BV, U0 = 4 * A, U1 = 5 * A,
                            U2 = 6 * A, EV = NA * A, RDY = EV + NE * NR * sizeof(T), TOTAL = RDY + 16;
};

// One step's inputs of a lane, read a step ahead so that their LDS reads
// complete while the previous step computes: the progress word, the row's
// coefficients, diagonal and b, and the staged values of the
// parents it takes from other boxes (valid when the progress word read
// before them is past the step).
template <class T, int NR = 1>
struct PipeRec {
    T cz, cy, cx, dd, uz, uy, ux;
    T bb[NR], ez[NR], ey[NR], ex[NR];  // per right-hand side
    int rdy, q, ax, ay;                // progress word; record slot; staged x- and y-face slots read
};

// The compute wavefront.  Lane (i, j) = (lane / TY, lane % TY); its y parent
// is lane - 1 (shr1), its x parent lane - TY (shr1 when TY = 1, XS = 1: lane
// 0 then takes the x-face value as the shift's `old`; the LDS crossbar
// otherwise).  A lane without a row this step (before its first plane,
// after its last, outside a partial box) reads and writes the spare row,
// and what it computes is read by no lane with a row.  A parent outside
// the grid has coefficient -0 and value 1 (the NONE slot), which adds
// nothing to a sum and divides nothing by zero.  hx, hy, hz: the box has
// staged x, y and z faces; opub: its own faces in the publication half.
// The x-face lanes (i = tx - 1) write diagonal d = t - (tx - 1) of the x
// face at step t, all of it; the y face likewise; the top plane row by
// row.  A face value's offset is linear in t, and whether a step writes a
// diagonal depends on t alone.  With NR right-hand sides (solves only) a
// lane carries NR values through the same steps, on the same coefficients.
template <class T, bool UPPER, int XS, int RMX, int MODE = PIPE_SOLVE, int NR = 1>
__device__ __forceinline__ void pipe_wave(const PipeArgs &a, int lane, int x0, int y0, int z0, int tx, int ty,
                                          int kz, bool hx, bool hy, bool hz, int3 fsz, unsigned char *smem,
                                          const int *ready, T *opub) {
    static_assert(NR == 1 || MODE == PIPE_SOLVE, "factorisations carry one value a row");
    using Lay = PipeLds<T, RMX, MODE, NR>;
    const T *c0 = reinterpret_cast<const T *>(smem + Lay::C0);
    const T *c1 = reinterpret_cast<const T *>(smem + Lay::C1);
    const T *c2 = reinterpret_cast<const T *>(smem + Lay::C2);
    const T *dg = reinterpret_cast<const T *>(smem + Lay::DG);
    const T *bv = reinterpret_cast<const T *>(smem + Lay::BV);
    T *xv = reinterpret_cast<T *>(smem + Lay::XV);
    const T *u0 = reinterpret_cast<const T *>(smem + Lay::U0);
    const T *u1 = reinterpret_cast<const T *>(smem + Lay::U1);
    const T *u2 = reinterpret_cast<const T *>(smem + Lay::U2);
    const T *ev = reinterpret_cast<const T *>(smem + Lay::EV);
    const int TY = a.TY;
    const int i = lane / TY, j = lane - (lane / TY) * TY;
    const bool lok = i < tx && j < ty;
    const int nsteps = kz + tx + ty - 2;
    const int txy = tx * ty;
    const bool bx = x0 + tx < a.nx, by = XS != 1 && a.sy > 0 && y0 + ty < a.ny;  // consumers exist
    const bool oz = lok && z0 + kz < a.nz;
    // The helper lane's x move reads the corner lane (tx - 1, ty - 1)
    // instead of its x parent: see the y face below.  The helper is lane 0,
    // whose own x parent is staged, or, in a box one lane wide in x or y
    // (where lane 0 is itself on a face), lane 63, which then lies outside
    // the box.
    const int corner = (tx - 1) * TY + (ty - 1);
    const int helper = tx == 1 || ty == 1 ? 63 : 0;
    const int xaddr = (XS != 1 && lane == helper ? corner : (lane + 64 - TY) & 63) << 2;
    // staged slots: the plane below, and per plane the x face (lanes with
    // i = 0) or the y face (lanes with j = 0); the NONE slot where a box has
    // no such face (a parent outside the grid)
    const bool fx = lok && i == 0, fy = lok && j == 0;
    const int zslot = lok && hz ? Lay::EZ + i + tx * j : Lay::NONE;
    const int xs0 = fx && hx ? Lay::EX + j * (ty + 1) : Lay::NONE, xst = fx && hx ? ty : 0;
    const int ys0 = fy && hy ? Lay::EY + i * (tx + 1) : Lay::NONE, yst = fy && hy ? tx : 0;
    const int qb = j * tx + i;
    // Publication, one store of every lane per step (lanes with nothing to
    // publish write the box's dump slot): at step t a lane writes slot
    // t sA + cA while t - aA is in [0, nA).  x-face lanes (i = tx - 1) write
    // diagonal t - (tx - 1) of the x face; y-face lanes (j = ty - 1) that of
    // the y face, except the corner, which is on the x face too: its y-face
    // value reaches the helper lane through the helper's x move one step
    // later, and the helper writes it (a last such store follows the final
    // step).  The top plane is a second store, in the steps from kz - 1 on:
    // the row of plane kz - 1 at step tz, else the dump slot.  (Sending a
    // top-plane row with the first store where that is free measured
    // slower.)
    const int dmp = fsz.x + fsz.y + fsz.z;
    int sA = 0, cA = 0, aA = 0, nA = 0;
    const bool cy_help = XS != 1 && lane == helper && by && bx;
    if (lok && i == tx - 1 && bx) {
        sA = ty, cA = j - (tx - 1) * ty, aA = tx - 1, nA = kz + ty - 1;
    } else if (lok && j == ty - 1 && by) {
        sA = tx, cA = fsz.x + i - (ty - 1) * tx, aA = ty - 1, nA = kz + tx - 1;
    } else if (cy_help) {
        sA = tx, cA = fsz.x + (tx - 1) - ty * tx, aA = ty, nA = kz + tx - 1;
    }
    const int oz0 = oz ? fsz.x + fsz.y + i + tx * j : dmp;
    const int tz = kz - 1 + i + j;
    // Offsets in bytes from opub (a 32-bit offset from the box's uniform base
    // address), advanced a step at a time rather than multiplied out: pob is
    // (t sA + cA) face values at step t, taken modulo 2^32, exact at the
    // steps that publish; pt is t - aA.
    constexpr unsigned VB = NR * sizeof(T);  // one face value
    const unsigned dmpb = (unsigned)dmp * VB, oz0b = (unsigned)oz0 * VB, psb = (unsigned)sA * VB;
    unsigned pob = (unsigned)cA * VB;
    int pt = -aA;
    auto st_pub = [&](unsigned boff, T v) {
        st_agent(reinterpret_cast<T *>(reinterpret_cast<char *>(opub) + boff), v);
    };
    auto publish_a = [&](const T *v) {  // this step's store, then on to the next step
        const unsigned ob = (unsigned)pt < (unsigned)nA ? pob : dmpb;
#pragma unroll
        for (int r = 0; r < NR; r++) st_pub(ob + r * (unsigned)sizeof(T), v[r]);
        pob += psb;
        pt++;
    };
    // the record slot of the row the next load reads (plane kk, slot qk),
    // advanced a plane at a time
    int kk = -i - j, qk = qb + (-i - j) * txy;
    auto slot_next = [&]() {
        const int q = lok & ((unsigned)kk < (unsigned)kz) ? qk : RMX;
        kk++;
        qk += txy;
        return q;
    };
    // Staged face slots advance by xst (yst) a step.  A lane reads them at
    // every step but uses them only at the steps it has a row (i = 0 or
    // j = 0, 0 <= k < kz); at the others the slot is still inside the staged
    // faces' LDS, and the value read goes to an idle computation.
    int ax = xs0 + (-i - j) * xst, ay = ys0 + (-i - j) * yst;
    auto load_ext = [&](PipeRec<T, NR> &R) {
#pragma unroll
        for (int r = 0; r < NR; r++) {
            R.ez[r] = ev[zslot * NR + r];
            if (XS != 1) R.ey[r] = ev[R.ay * NR + r];
            R.ex[r] = ev[R.ax * NR + r];
        }
    };
    auto load = [&](PipeRec<T, NR> &R) {  // the rows of planes -i-j, -i-j+1, ... in turn
        R.rdy = ld_ready(ready);
        const int q = slot_next();
        R.q = q;
        R.cz = c0[q];
        if (XS != 1) R.cy = c1[q];
        R.cx = c2[q];
        R.dd = dg[q];
        if (MODE == PIPE_SOLVE)
#pragma unroll
            for (int r = 0; r < NR; r++) R.bb[r] = bv[q * NR + r];
        if (MODE == PIPE_ILU0) {
            R.uz = u0[q];
            if (XS != 1) R.uy = u1[q];
            R.ux = u2[q];
        }
        R.ax = ax;
        R.ay = ay;
        ax += xst;
        ay += yst;
        load_ext(R);
    };
    T xl[NR];  // the values this lane solved at the previous step
#pragma unroll
    for (int r = 0; r < NR; r++) xl[r] = T(0);
    auto step = [&](int t, PipeRec<T, NR> &C, PipeRec<T, NR> &N) {
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
            load_ext(C);
        }
        // the next step's inputs, in flight while this step computes (after
        // the wait above, so the common path never waits on these reads)
        load(N);
        // the parents: z from this lane, y from lane - 1, x from lane - TY
        // (staged values on the box's faces)
        // The lane moves run on every lane before any select: a cross-lane
        // read from a lane that is off in EXEC returns zero, so they must
        // not be sunk into the branch a select may become (the empty asm
        // pins them here).
        T vx[NR], vy[NR], pzv[NR], pxv[NR], pyv[NR];
#pragma unroll
        for (int r = 0; r < NR; r++) {
            vx[r] = XS == 1 ? shr1(C.ex[r], xl[r]) : lane_read(xl[r], xaddr);
            vy[r] = XS == 1 ? T(0) : shr1z(xl[r]);
            __asm__ volatile("" : "+v"(vx[r]), "+v"(vy[r]));
        }
#pragma unroll
        for (int r = 0; r < NR; r++) {
            pzv[r] = k == 0 ? C.ez[r] : xl[r];
            pxv[r] = XS == 1 ? vx[r] : (fx ? C.ex[r] : vx[r]);
            pyv[r] = XS == 1 ? T(0) : (fy ? C.ey[r] : vy[r]);
        }
        if (MODE == PIPE_SOLVE) {
            // z, y, x: ascending columns for L, descending for U (a 2D grid,
            // XS = 1, has no y term)
#pragma unroll
            for (int r = 0; r < NR; r++) {
                const T pz = pzv[r], px = pxv[r], py = pyv[r];
                T s = UPPER ? fma(pz, C.cz, T(0)) : fma(C.cz, pz, T(0));
                if (XS != 1) s = UPPER ? fma(py, C.cy, s) : fma(C.cy, py, s);
                s = UPPER ? fma(px, C.cx, s) : fma(C.cx, px, s);
                xl[r] = (C.bb[r] - s) / C.dd;
                xv[C.q * NR + r] = xl[r];
            }
        } else {
            const T pz = pzv[0], px = pxv[0], py = pyv[0];
            // the row's entries over its parents' diagonals (the parents'
            // rows share no earlier column with it on a 7- or 5-point stencil,
            // so the reference's dot products are zero: A_ik - 0 = A_ik)
            const T lz = C.cz / pz, ly = XS == 1 ? T(0) : C.cy / py, lx = C.cx / px;
            T d;
            if (MODE == PIPE_IC0) {
                // L_ii = sqrt(A_ii - sum L_ik^2) over the row's entries in
                // ascending order; a non-positive pivot keeps A_ii
                T dv = fma(lz, lz, T(0));
                if (XS != 1) dv = fma(ly, ly, dv);
                dv = fma(lx, lx, dv);
                const T e = C.dd - dv;
                d = e > T(0) ? sqrt(e) : C.dd;
                if (!(e > T(0)) && C.q != RMX)
                    raise_status(a.status, BAAS_NONPOSITIVE_PIVOT, pipe_row<false>(a, x0 + i, y0 + j, z0 + k));
            } else {
                // U_ii = A_ii - sum L_ik U_ki, the parents' U_ki being A_ki
                d = fma(-lz, C.uz, C.dd);
                if (XS != 1) d = fma(-ly, C.uy, d);
                d = fma(-lx, C.ux, d);
                // a pivot the reference asserts against (parents outside the
                // grid divide by 1)
                const bool small = !(fabs(pz) > 1.0e-16) | !(fabs(px) > 1.0e-16) | (XS != 1 && !(fabs(py) > 1.0e-16));
                if (small && C.q != RMX)
                    raise_status(a.status, BAAS_ZERO_PIVOT, pipe_row<false>(a, x0 + i, y0 + j, z0 + k));
            }
            T *w0 = reinterpret_cast<T *>(smem + Lay::C0), *w1 = reinterpret_cast<T *>(smem + Lay::C1);
            T *w2 = reinterpret_cast<T *>(smem + Lay::C2), *wd = reinterpret_cast<T *>(smem + Lay::DG);
            w0[C.q] = lz;
            if (XS != 1) w1[C.q] = ly;
            w2[C.q] = lx;
            wd[C.q] = d;
            xl[0] = d;
        }
        // (a lane without a row this step publishes a value no consumer
        // reads: idle lanes take parents only from idle lanes, and their
        // arithmetic never yields the signalling-NaN sentinel)
        T pv[NR];
#pragma unroll
        for (int r = 0; r < NR; r++) pv[r] = cy_help ? vx[r] : xl[r];
        publish_a(pv);
        if (t >= kz - 1)
#pragma unroll
            for (int r = 0; r < NR; r++) st_pub((t == tz ? oz0b : dmpb) + r * (unsigned)sizeof(T), xl[r]);
    };
    // two register sets used in turn, so no step copies one into the other
    PipeRec<T, NR> A, B;
    load(A);
    for (int t = 0; t < nsteps; t += 2) {
        step(t, A, B);
        if (t + 1 < nsteps) step(t + 1, B, A);
    }
    if (XS != 1) {
        // the corner's last y-face value (its final step's row)
        T vc[NR];
#pragma unroll
        for (int r = 0; r < NR; r++) {
            vc[r] = lane_read(xl[r], xaddr);
            __asm__ volatile("" : "+v"(vc[r]));
        }
        if (cy_help) publish_a(vc);  // step nsteps
    }
}

template <class T, bool UPPER, int XS, int RMX, int NR>
__global__ __launch_bounds__(128) void k_sweep_pipe(PipeArgs a, const T *__restrict__ vals, const T *__restrict__ b,
                                                    T *x, T *xp) {
    using Lay = PipeLds<T, RMX, PIPE_SOLVE, NR>;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ int s_unit;
    if (a.skip && ld_agent(a.skip)) return;
    const int epoch = read_epoch(a.ctx);
    // wave-uniform values are made so explicitly (readfirstlane), so that
    // the box's fields and the publication pointers live in scalar
    // registers and stores address them as scalar base + 32-bit offset
    const int par = __builtin_amdgcn_readfirstlane(ld_agent(a.pcount)) & 1;
    T *pub = xp + (size_t)par * a.half;
    T *nxt = xp + (size_t)(par ^ 1) * a.half;
    T *c0 = reinterpret_cast<T *>(smem_raw + Lay::C0);
    T *c1 = reinterpret_cast<T *>(smem_raw + Lay::C1);
    T *c2 = reinterpret_cast<T *>(smem_raw + Lay::C2);
    T *dg = reinterpret_cast<T *>(smem_raw + Lay::DG);
    T *bv = reinterpret_cast<T *>(smem_raw + Lay::BV);
    const T *xv = reinterpret_cast<const T *>(smem_raw + Lay::XV);
    T *ev = reinterpret_cast<T *>(smem_raw + Lay::EV);
    int *ready = reinterpret_cast<int *>(smem_raw + Lay::RDY);
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nt = blockDim.x;
    // the spare row idle lanes read and the slot of parents outside the grid
    if (tid == 0) {
        c0[RMX] = c1[RMX] = c2[RMX] = -T(0);
        dg[RMX] = T(1);
        for (int r = 0; r < NR; r++) {
            bv[RMX * NR + r] = T(0);
            ev[Lay::NONE * NR + r] = T(1);
        }
    }
    const int qid = (int)(blockIdx.x % (unsigned)a.nq);
    int *tkt = a.nq > 1 ? a.qticket + 32 * qid : a.ctx.ticket;
    const int qboxes = (a.nunits - qid + a.nq - 1) / a.nq;
    for (;;) {
        // tickets name (box, group of NR right-hand sides) pairs, box-major:
        // the boxes a box waits for (same group, earlier in wavefront order)
        // hold earlier tickets.  With several groups, workgroup w takes
        // tickets from queue w mod nq, which holds the boxes u = q mod nq:
        // every group of a box then runs on one XCD (workgroups are
        // dealt to the XCDs in turn) and reads the box's coefficients
        // through one L2.  Every wait still ends: each queue is in box
        // order and has a workgroup, so the smallest box any workgroup waits
        // for has its producers finished or next in their queues.
        const int ng = a.nrhs / NR;
        const int tk = __builtin_amdgcn_readfirstlane(claim_unit(tkt, &s_unit));
        if (tk >= qboxes * ng) break;
        const int u = qid + a.nq * (tk / ng), grp = tk - (tk / ng) * ng, rhs = grp * NR;
        T *rpub = pub + (long long)grp * a.rstride * NR;  // this group's faces
        const int4 bx = a.box[2 * u], bf = a.box[2 * u + 1];
        const int x0 = bx.x, y0 = bx.y, z0 = bx.z;
        const int tx = bx.w & 255, ty = (bx.w >> 8) & 255, kz = (bx.w >> 16) & 255;
        const int nrow = tx * ty * kz, nsteps = kz + tx + ty - 2;
        // the faces this box reads: the x face and y face of the boxes
        // before it and the top plane of the box below (none at the grid's
        // low faces), each as long as this box's own
        const int3 fsz = pipe_face_sizes(tx, ty, kz, a.sy > 0);
        const bool hx = bf.y >= 0, hy = bf.z >= 0, hz = bf.w >= 0;
        if (tid == 0) *ready = hx | hy | hz ? 0 : nsteps;
        // the box's rows: eight per thread per pass, every load of a pass in
        // flight together.  A row's entries sit where the full stencil puts
        // them (pipe_lp): for L the z, y and x parents in that order, then
        // the diagonal; for U = L^T (the backward sweep, whose flipped
        // coordinates make the same count of parents) the diagonal first.
        // A parent the row lacks reads the diagonal and gets coefficient -0.
        constexpr int PB = NR == 1 ? 8 : 4;
        for (int q0 = tid; q0 < nrow; q0 += PB * nt) {
            bool hxr[PB], hyr[PB], hzr[PB];
            T cz[PB], cy[PB], cx[PB], cd[PB], bb[PB][NR];
#pragma unroll
            for (int v = 0; v < PB; v++) {
                const int q = min(q0 + v * nt, nrow - 1);
                const int ii = q % tx, r2 = q / tx, jj = r2 % ty, kk = r2 / ty;
                const int xg = x0 + ii, yg = y0 + jj, zg = z0 + kk;
                const int ex = xg > 0, ey = a.sy > 0 && yg > 0, ez = zg > 0;
                const int lp = pipe_lp(a.nx, a.ny, a.sy, a.sz, xg, yg, zg);
                int pz, py, px, pd;
                if (!UPPER) {
                    pz = lp, py = pz + ez, px = py + ey, pd = px + ex;
                } else {
                    pd = a.nnz - lp - (1 + ex + ey + ez), px = pd + 1, py = px + ex, pz = py + ey;
                }
                hxr[v] = ex, hyr[v] = ey, hzr[v] = ez;
                cz[v] = vals[ez ? pz : pd];
                cy[v] = vals[ey ? py : pd];
                cx[v] = vals[ex ? px : pd];
                cd[v] = vals[pd];
                const size_t rb = (size_t)pipe_row<UPPER>(a, xg, yg, zg) * a.nrhs + rhs;
#pragma unroll
                for (int r = 0; r < NR; r++) bb[v][r] = b[rb + r];
            }
#pragma unroll
            for (int v = 0; v < PB; v++) {
                const int q = q0 + v * nt;
                if (q >= nrow) continue;
                c0[q] = hzr[v] ? cz[v] : -T(0);
                c1[q] = hyr[v] ? cy[v] : -T(0);
                c2[q] = hxr[v] ? cx[v] : -T(0);
                dg[q] = cd[v];
#pragma unroll
                for (int r = 0; r < NR; r++) bv[q * NR + r] = bb[v][r];
            }
        }
        __syncthreads();
        if (wv == 0)
            pipe_stage<T, NR>(hx ? rpub + bf.y * NR : nullptr, fsz.x, ty, hy ? rpub + bf.z * NR : nullptr, fsz.y, tx,
                              hz ? rpub + bf.w * NR : nullptr, fsz.z, ev + Lay::EX * NR, ev + Lay::EY * NR,
                              ev + Lay::EZ * NR, ev + Lay::DUMP * NR, rpub, nsteps, ready, lane, a.spin_limit,
                              a.status, (long long)(nxt - pub));
        else if (wv == 1)
            pipe_wave<T, UPPER, XS, RMX, PIPE_SOLVE, NR>(a, lane, x0, y0, z0, tx, ty, kz, hx, hy, hz, fsz, smem_raw,
                                                         ready, rpub + bf.x * NR);
        __syncthreads();
        for (int q = tid; q < nrow; q += nt) {
            const int ii = q % tx, r2 = q / tx, jj = r2 % ty, kk = r2 / ty;
            const int row = pipe_row<UPPER>(a, x0 + ii, y0 + jj, z0 + kk);
#pragma unroll
            for (int r = 0; r < NR; r++) x[(size_t)row * a.nrhs + rhs + r] = xv[q * NR + r];
        }
    }
    if (threadIdx.x == 0) {
        __threadfence();
        const int d = atomicAdd(a.ctx.done, 1);
        if (d == (int)gridDim.x - 1) {
            atomicExch(a.ctx.ticket, 0);
            for (int q = 0; q < a.nq && a.nq > 1; q++) atomicExch(a.qticket + 32 * q, 0);
            atomicExch(a.ctx.epoch, epoch == 0x7fffffff ? 1 : epoch + 1);
            atomicAdd(a.pcount, 1);
            __threadfence();
            atomicExch(a.ctx.done, 0);
        }
    }
}

// Where the entries of row (x, y, z) of the full stencil matrix A sit for
// the factorisations: its parents' columns (z, y, x; pos of the diagonal
// when absent) and its diagonal, and (ILU0) its parents' entries in this
// row's column -- the last entry of the z parent's row, the entry after
// the y parent's diagonal and its x child, the entry after the x parent's
// diagonal.
struct PipeApos {
    int z, y, x, d, uz, uy, ux;
    bool ez, ey, ex;
};
__device__ __forceinline__ PipeApos pipe_apos(const PipeArgs &a, int xg, int yg, int zg) {
    PipeApos o;
    const bool three = a.sy > 0;
    o.ex = xg > 0, o.ey = three && yg > 0, o.ez = zg > 0;
    const int pa = pipe_pa(a.nx, a.ny, a.nz, a.sy, a.sz, xg, yg, zg);
    o.z = pa;
    o.y = o.z + o.ez;
    o.x = o.y + o.ey;
    o.d = o.x + o.ex;
    const int hasx1 = xg < a.nx - 1, hasy1 = three && yg < a.ny - 1;
    // z parent (xg, yg, zg - 1): its row ends with column r
    o.uz = o.d;
    if (o.ez) {
        const int low = (xg > 0) + (three && yg > 0) + (zg - 1 > 0);
        o.uz = pipe_pa(a.nx, a.ny, a.nz, a.sy, a.sz, xg, yg, zg - 1) + low + hasx1 + hasy1 + 1;
    }
    o.uy = o.d;
    if (o.ey) {
        const int low = (xg > 0) + (yg - 1 > 0) + (zg > 0);
        o.uy = pipe_pa(a.nx, a.ny, a.nz, a.sy, a.sz, xg, yg - 1, zg) + low + 1 + hasx1;
    }
    o.ux = o.d;
    if (o.ex) {
        const int low = (xg - 1 > 0) + (three && yg > 0) + (zg > 0);
        o.ux = pipe_pa(a.nx, a.ny, a.nz, a.sy, a.sz, xg - 1, yg, zg) + low + 1;
    }
    return o;
}

// IC0 (MODE PIPE_IC0) or ILU0 (PIPE_ILU0) of a full 7- or 5-point stencil
// matrix in natural order on PIPE plans (merge/IC.cpp:32-89 up-looking IC0,
// merge/ILU.cpp:350-386 up-looking ILU0), in place of lu (which holds A's
// values): the box machinery of k_sweep_pipe, with each row's value the
// diagonal its children divide by.
template <int MODE, int XS, int RMX>
__global__ __launch_bounds__(128) void k_factor_pipe(PipeArgs a, const double *__restrict__ av, double *lu,
                                                     double *xp) {
    using T = double;
    using Lay = PipeLds<T, RMX, MODE>;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ int s_unit;
    const int epoch = read_epoch(a.ctx);
    const int par = __builtin_amdgcn_readfirstlane(ld_agent(a.pcount)) & 1;
    T *pub = xp + (size_t)par * a.half;
    T *nxt = xp + (size_t)(par ^ 1) * a.half;
    T *c0 = reinterpret_cast<T *>(smem_raw + Lay::C0);
    T *c1 = reinterpret_cast<T *>(smem_raw + Lay::C1);
    T *c2 = reinterpret_cast<T *>(smem_raw + Lay::C2);
    T *dg = reinterpret_cast<T *>(smem_raw + Lay::DG);
    T *u0 = reinterpret_cast<T *>(smem_raw + Lay::U0);
    T *u1 = reinterpret_cast<T *>(smem_raw + Lay::U1);
    T *u2 = reinterpret_cast<T *>(smem_raw + Lay::U2);
    T *ev = reinterpret_cast<T *>(smem_raw + Lay::EV);
    int *ready = reinterpret_cast<int *>(smem_raw + Lay::RDY);
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nt = blockDim.x;
    if (tid == 0) {
        c0[RMX] = c1[RMX] = c2[RMX] = -T(0);
        dg[RMX] = T(1);
        if (MODE == PIPE_ILU0) u0[RMX] = u1[RMX] = u2[RMX] = T(0);
        ev[Lay::NONE] = T(1);
    }
    for (;;) {
        const int u = __builtin_amdgcn_readfirstlane(claim_unit(a.ctx.ticket, &s_unit));
        if (u >= a.nunits) break;
        const int4 bx = a.box[2 * u], bf = a.box[2 * u + 1];
        const int x0 = bx.x, y0 = bx.y, z0 = bx.z;
        const int tx = bx.w & 255, ty = (bx.w >> 8) & 255, kz = (bx.w >> 16) & 255;
        const int nrow = tx * ty * kz, nsteps = kz + tx + ty - 2;
        const int3 fsz = pipe_face_sizes(tx, ty, kz, a.sy > 0);
        const bool hx = bf.y >= 0, hy = bf.z >= 0, hz = bf.w >= 0;
        if (tid == 0) *ready = hx | hy | hz ? 0 : nsteps;
        // the box's rows of A, eight per thread per pass with every load of a
        // pass in flight together; a parent the row lacks gets -0 (and, for
        // ILU0, an entry of 0 in this row's column)
        constexpr int PB = 8;
        for (int q0 = tid; q0 < nrow; q0 += PB * nt) {
            bool hxr[PB], hyr[PB], hzr[PB];
            T cz[PB], cy[PB], cx[PB], cd[PB], vz[PB], vy[PB], vx[PB];
#pragma unroll
            for (int v = 0; v < PB; v++) {
                const int q = min(q0 + v * nt, nrow - 1);
                const int ii = q % tx, r2 = q / tx, jj = r2 % ty, kk = r2 / ty;
                const PipeApos o = pipe_apos(a, x0 + ii, y0 + jj, z0 + kk);
                hxr[v] = o.ex, hyr[v] = o.ey, hzr[v] = o.ez;
                cz[v] = av[o.z];
                cy[v] = av[o.y];
                cx[v] = av[o.x];
                cd[v] = av[o.d];
                if (MODE == PIPE_ILU0) {
                    vz[v] = av[o.uz];
                    vy[v] = av[o.uy];
                    vx[v] = av[o.ux];
                }
            }
#pragma unroll
            for (int v = 0; v < PB; v++) {
                const int q = q0 + v * nt;
                if (q >= nrow) continue;
                c0[q] = hzr[v] ? cz[v] : -T(0);
                c1[q] = hyr[v] ? cy[v] : -T(0);
                c2[q] = hxr[v] ? cx[v] : -T(0);
                dg[q] = cd[v];
                if (MODE == PIPE_ILU0) {
                    u0[q] = hzr[v] ? vz[v] : T(0);
                    u1[q] = hyr[v] ? vy[v] : T(0);
                    u2[q] = hxr[v] ? vx[v] : T(0);
                }
            }
        }
        __syncthreads();
        if (wv == 0)
            pipe_stage<T>(hx ? pub + bf.y : nullptr, fsz.x, ty, hy ? pub + bf.z : nullptr, fsz.y, tx,
                          hz ? pub + bf.w : nullptr, fsz.z, ev + Lay::EX, ev + Lay::EY, ev + Lay::EZ, ev + Lay::DUMP,
                          pub, nsteps, ready, lane, a.spin_limit, a.status, (long long)(nxt - pub));
        else if (wv == 1)
            pipe_wave<T, false, XS, RMX, MODE>(a, lane, x0, y0, z0, tx, ty, kz, hx, hy, hz, fsz, smem_raw, ready,
                                               pub + bf.x);
        __syncthreads();
        // the box's L (IC0) or L multipliers and U diagonal (ILU0) into lu,
        // and, when lu is not A itself, the rows' upper entries (A's values)
        for (int q = tid; q < nrow; q += nt) {
            const int ii = q % tx, r2 = q / tx, jj = r2 % ty, kk = r2 / ty;
            const int xg = x0 + ii, yg = y0 + jj, zg = z0 + kk;
            const PipeApos o = pipe_apos(a, xg, yg, zg);
            if (o.ez) lu[o.z] = c0[q];
            if (o.ey) lu[o.y] = c1[q];
            if (o.ex) lu[o.x] = c2[q];
            lu[o.d] = dg[q];
            if (av != lu) {
                const int nup = (xg < a.nx - 1) + (a.sy > 0 && yg < a.ny - 1) + (zg < a.nz - 1);
                for (int e = 1; e <= nup; e++) lu[o.d + e] = av[o.d + e];
            }
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

// The kernel instance for plan p: the smallest of 512, 1024 and 2048 rows
// that holds the box's rows and, in RMX / 2 slots each, its x and y faces
// (0: none does; the plan builder keeps boxes within 2048 rows and faces
// within PIPE_FACE_MAX = 1024 values, so one always does).  Smaller
// instances take less LDS, and more boxes share a CU.
static int pipe_rmx(const Plan &p) {
    const int rows = p.pg[5] * p.pg[6] * p.pg[7];
    const int3 f = pipe_face_sizes(p.pg[5], p.pg[6], p.pg[7], p.pg[3] > 0);
    if (f.z > 64) return 0;
    for (int r : {512, 1024, 2048})
        if (rows <= r && f.x <= r / 2 && f.y <= r / 2) return r;
    return 0;
}

// LDS bytes the pipelined sweep needs per workgroup on plan p.
size_t pipe_lds(const Plan &p, int ts) {
    switch (pipe_rmx(p)) {
    case 512: return ts == 8 ? PipeLds<double, 512>::TOTAL : PipeLds<float, 512>::TOTAL;
    case 1024: return ts == 8 ? PipeLds<double, 1024>::TOTAL : PipeLds<float, 1024>::TOTAL;
    case 2048: return ts == 8 ? PipeLds<double, 2048>::TOTAL : PipeLds<float, 2048>::TOTAL;
    }
    return (size_t)1 << 40;  // no instance holds such boxes
}

template <class T, int RMX, int NR>
static int launch_pipe_r(baas_handle *h, const PipeArgs &a0, int uplo, const T *vals, const T *b, T *x, T *xp,
                         int nunits, int nrhs, hipStream_t stream) {
    // the x parent sits TY lanes down: a DPP shift for TY = 1 (2D grids),
    // the LDS crossbar otherwise
    using K = void (*)(PipeArgs, const T *, const T *, T *, T *);
    K kern = a0.TY == 1 ? (uplo ? k_sweep_pipe<T, true, 1, RMX, NR> : k_sweep_pipe<T, false, 1, RMX, NR>)
                        : (uplo ? k_sweep_pipe<T, true, 0, RMX, NR> : k_sweep_pipe<T, false, 0, RMX, NR>);
    const size_t lds = PipeLds<T, RMX, PIPE_SOLVE, NR>::TOTAL;
    int per_cu = 0;
    BAAS_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 128, lds));
    if (per_cu <= 0) return BAAS_UNSUPPORTED;
    long long grid = std::min<long long>((long long)per_cu * h->num_cus, (long long)nunits * (nrhs / NR));
    if (grid < 1) grid = 1;
    PipeArgs a = a0;
    // several groups: one ticket queue per XCD (fewer when the grid or the
    // boxes are fewer, so that every queue has a workgroup and a box)
    a.nq = nrhs / NR > 1 ? (int)std::min<long long>({8, grid, (long long)nunits}) : 1;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(128), lds, stream, a, vals, b, x, xp);
    BAAS_HIP_TRY(hipGetLastError());
    return BAAS_OK;
}

// Right-hand sides per lane: two where the count is even (each lane then
// solves two columns on the same coefficients, half the tickets, and b and
// x move in 16-byte pairs), else one.  On the 64^3 Laplacian with 32
// right-hand sides, two a lane took 0.17 ms per launch against 0.20 ms for
// one and 0.20 ms for four (boxes of 1024 rows, per-XCD ticket queues;
// 512- and 2048-row boxes ranked the same).
template <class T, int RMX>
static int launch_pipe_nr(baas_handle *h, const PipeArgs &a, int uplo, const T *vals, const T *b, T *x, T *xp,
                          int nunits, int nrhs, hipStream_t stream) {
    if (nrhs % 2 == 0 && PipeLds<T, RMX, PIPE_SOLVE, 2>::TOTAL <= lds_limit(h))
        return launch_pipe_r<T, RMX, 2>(h, a, uplo, vals, b, x, xp, nunits, nrhs, stream);
    return launch_pipe_r<T, RMX, 1>(h, a, uplo, vals, b, x, xp, nunits, nrhs, stream);
}

// All right-hand sides in one launch (b and x are n x nrhs, row-major):
// each has its own copy of the box faces in each publication half, and the
// workgroups take (box, right-hand side) pairs from one ticket.
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
    a.half = p.pipe_half * nrhs;
    a.rstride = p.pipe_half;
    a.nq = 1;
    a.qticket = h->counters + BAAS_QTICKET_BASE;
    a.box = reinterpret_cast<const int4 *>(p.pipe_box.p);
    a.nnz = s->nnz_lower;
    a.pcount = p.pcount.p;
    a.ctx = Ctx{h->counters, h->counters + 1, h->counters + 2};
    a.skip = skip;
    a.status = h->status;
    a.spin_limit = h->spin_limit;
    // the kernel's LDS holds boxes of up to 2048 rows and faces of up to
    // PIPE_FACE_MAX values (the plan builder keeps within both)
    if (a.TX * a.TY * a.K > 2048 || !pipe_faces_fit(a.TX, a.TY, a.K, a.sy > 0) || a.half < 1) return BAAS_UNSUPPORTED;
    if ((long long)p.nunits * nrhs > 0x7fffffff) return BAAS_TOO_LARGE;  // tickets are 32-bit
    switch (pipe_rmx(p)) {
    case 512: return launch_pipe_nr<T, 512>(h, a, uplo, vals, b, x, xp, p.nunits, nrhs, stream);
    case 1024: return launch_pipe_nr<T, 1024>(h, a, uplo, vals, b, x, xp, p.nunits, nrhs, stream);
    case 2048: return launch_pipe_nr<T, 2048>(h, a, uplo, vals, b, x, xp, p.nunits, nrhs, stream);
    }
    return BAAS_UNSUPPORTED;
}

__global__ void k_pipe_arm(size_t n, double *p) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = Sentinel<double>::value();
}

// LDS bytes of the factorisation on plan p.
static size_t pipe_factor_lds(const Plan &p, int kind) {
    const bool ic = kind == BAAS_KIND_IC0;
    switch (pipe_rmx(p)) {
    case 512: return ic ? PipeLds<double, 512, PIPE_IC0>::TOTAL : PipeLds<double, 512, PIPE_ILU0>::TOTAL;
    case 1024: return ic ? PipeLds<double, 1024, PIPE_IC0>::TOTAL : PipeLds<double, 1024, PIPE_ILU0>::TOTAL;
    case 2048: return ic ? PipeLds<double, 2048, PIPE_IC0>::TOTAL : PipeLds<double, 2048, PIPE_ILU0>::TOTAL;
    }
    return (size_t)1 << 40;
}

template <int MODE, int RMX>
static int launch_factor_r(baas_handle *h, const PipeArgs &a, const double *av, double *lu, double *xp, int nunits) {
    using K = void (*)(PipeArgs, const double *, double *, double *);
    K kern = a.TY == 1 ? k_factor_pipe<MODE, 1, RMX> : k_factor_pipe<MODE, 0, RMX>;
    const size_t lds = PipeLds<double, RMX, MODE>::TOTAL;
    int per_cu = 0;
    BAAS_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 128, lds));
    if (per_cu <= 0) return BAAS_UNSUPPORTED;
    long long grid = std::min<long long>((long long)per_cu * h->num_cus, nunits);
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(128), lds, h->stream, a, av, lu, xp);
    BAAS_HIP_TRY(hipGetLastError());
    return BAAS_OK;
}

// IC0 / ILU0 on the forward plan when it is a PIPE plan over a full stencil
// (p.pipe_fact), from A's values in a_vals into lu (which may be a_vals);
// BAAS_UNSUPPORTED otherwise, with lu not touched.
int launch_pipe_factor(baas_handle *h, baas_schedule *s, Plan &p, int kind, const double *a_vals, double *lu) {
    if (!p.pipe || !p.pipe_fact || p.dataflow != 5) return BAAS_UNSUPPORTED;
    if (pipe_factor_lds(p, kind) > lds_limit(h)) return BAAS_UNSUPPORTED;
    if (p.pg[5] * p.pg[6] * p.pg[7] > 2048 || !pipe_faces_fit(p.pg[5], p.pg[6], p.pg[7], p.pg[3] > 0) ||
        p.pipe_half < 1)
        return BAAS_UNSUPPORTED;
    // the publication halves, armed (shared with FP64 solves on this plan)
    if (p.xp_type != (int)sizeof(double) || p.xp_nrhs != -1) {
        const size_t cnt = 2 * (size_t)p.pipe_half;
        BAAS_TRY(p.xp_raw.alloc(cnt * sizeof(double)));
        hipLaunchKernelGGL(k_pipe_arm, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, h->stream, cnt,
                           reinterpret_cast<double *>(p.xp_raw.p));
        BAAS_HIP_TRY(hipGetLastError());
        p.xp_type = (int)sizeof(double);
        p.xp_nrhs = -1;
    }
    PipeArgs a;
    a.n = s->n;
    a.nunits = p.nunits;
    a.nrhs = 1;
    a.nx = p.pg[0];
    a.ny = p.pg[1];
    a.nz = p.pg[2];
    a.sy = p.pg[3];
    a.sz = p.pg[4];
    a.TX = p.pg[5];
    a.TY = p.pg[6];
    a.K = p.pg[7];
    a.half = p.pipe_half;
    a.rstride = p.pipe_half;
    a.nq = 1;
    a.qticket = nullptr;
    a.box = reinterpret_cast<const int4 *>(p.pipe_box.p);
    a.nnz = s->nnz_lower;
    a.pcount = p.pcount.p;
    a.ctx = Ctx{h->counters, h->counters + 1, h->counters + 2};
    a.skip = nullptr;
    a.status = h->status;
    a.spin_limit = h->spin_limit;
    double *xp = reinterpret_cast<double *>(p.xp_raw.p);
    const bool ic = kind == BAAS_KIND_IC0;
    switch (pipe_rmx(p)) {
    case 512:
        return ic ? launch_factor_r<PIPE_IC0, 512>(h, a, a_vals, lu, xp, p.nunits)
                  : launch_factor_r<PIPE_ILU0, 512>(h, a, a_vals, lu, xp, p.nunits);
    case 1024:
        return ic ? launch_factor_r<PIPE_IC0, 1024>(h, a, a_vals, lu, xp, p.nunits)
                  : launch_factor_r<PIPE_ILU0, 1024>(h, a, a_vals, lu, xp, p.nunits);
    case 2048:
        return ic ? launch_factor_r<PIPE_IC0, 2048>(h, a, a_vals, lu, xp, p.nunits)
                  : launch_factor_r<PIPE_ILU0, 2048>(h, a, a_vals, lu, xp, p.nunits);
    }
    return BAAS_UNSUPPORTED;
}

template int launch_pipe<double>(baas_handle *, baas_schedule *, const Plan &, int, const double *, const double *,
                                 double *, double *, int, hipStream_t, const int *);
template int launch_pipe<float>(baas_handle *, baas_schedule *, const Plan &, int, const float *, const float *,
                                float *, float *, int, hipStream_t, const int *);

}  // namespace baas
