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
#include <mutex>
#include <type_traits>

#include "blas.h"
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
// row waits exactly for its own parents.  Each lane retries in one loop that
// consumes every ready entry in ascending (descending for U) column order,
// so lanes of one wavefront that depend on each other still make progress.
//
// Rows read by later units publish their value itself: one atomic 64-bit
// (32-bit for FP32) store into a plan-owned array of two halves pre-filled
// with the sentinel.  A sweep publishes into the half selected by the plan's
// sweep counter and re-arms the same rows in the other half for the next
// sweep, so a consumer needs one load per parent (the value is the flag) and
// a producer never waits for its stores.  Two exchange waves per workgroup
// move values between LDS and the publication array (one stages parents from
// earlier units, one publishes rows later units read), so the compute waves
// touch global memory only to load matrix data, a round ahead of its use.
struct DfArgs {
    int n, nunits, keymode;
    int nrhs;  // right-hand sides per row (1 for k_sweep_df)
    const int4 *meta;  // per position: row | ext flag, first off-diagonal entry, count, LDS slot
    const int *scol;
    const int *unit_ptr, *key_begin, *ext_ptr, *ext_cols;
    const int *out_ptr, *out_slot;  // slots of rows later units read
    const int *key_row;             // keymode 2: row of each key
    int *pcount;  // the plan's sweep counter (parity selects the publication half)
    Ctx ctx;
    const int *skip;
    // per work unit, nonzero = skip the unit (multi-RHS lane kernel only):
    // its rows' values are not needed and no unit that reads them runs (a
    // batch's converged systems); the unit only re-arms its published rows
    const int *unit_skip;
    DevStatus *status;
    long long spin_limit;
};


// one row's data, loaded a round ahead of its turn; CH parents per chunk
template <class T, int CH>
struct DfRow {
    int rw, slot, left, jn, nc;
    T bv, dv;
    int cs[CH];
    T cv[CH];
};

// Row record of position q; positions past the unit's end read the unit's
// last record (so every load stays unconditional and in bounds) and are
// marked empty by the caller.
__device__ __forceinline__ int4 df_meta(const DfArgs &a, int q, int q1) {
    return a.meta[q < q1 ? q : q1 - 1];
}

// Issue the loads of one row (b, diagonal, first chunk of parents); nothing
// here waits for them.  The loads are unconditional (a short row re-reads its
// last parent), which keeps the compiler from merging divergent paths into a
// wait for every load in flight.
template <class T, bool UPPER, int CH>
__device__ __forceinline__ void df_load_row(const DfArgs &a, const T *__restrict__ vals, const T *__restrict__ b,
                                            const int4 m, bool valid, DfRow<T, CH> &R) {
    const int cnt = m.z;
    R.slot = valid ? m.w : -1;
    R.rw = m.x;
    R.bv = b[m.x & 0x7fffffff];
    R.dv = vals[UPPER ? m.y - cnt : m.y + cnt];
    R.nc = valid ? (cnt < CH ? cnt : CH) : 0;
    const int last = cnt > 0 ? (cnt < CH ? cnt : CH) - 1 : 0;
#pragma unroll
    for (int t = 0; t < CH; t++) {
        const int o = t < last ? t : last;
        const int j = UPPER ? m.y - o : m.y + o;
        R.cs[t] = a.scol[j];  // no arithmetic on loaded values here: that would wait for them
        R.cv[t] = vals[j];
    }
    R.jn = UPPER ? m.y - R.nc : m.y + R.nc;
    R.left = valid ? cnt - R.nc : 0;
}

// Exchange wave, incoming side: stage the unit's parents in other units into
// LDS slots [nk * nrhs, ...) (stage_published).
template <class T>
__device__ __forceinline__ void df_incoming(const DfArgs &a, const T *pub, T *sl, const int *cols, int nk, int ne,
                                            int lane) {
    stage_published<T>(pub, sl, cols, nk, ne, a.nrhs, lane, a.spin_limit, a.status);
}

// Exchange wave, outgoing side: publish the unit's rows that later units read
// (their LDS slots in execution order, outs) as soon as they are solved (the
// value is the flag), and re-arm the same rows in the other half for the
// next sweep.  Compute waves therefore issue no global stores.
template <class T>
__device__ __forceinline__ int df_row_of(const DfArgs &a, int kb, int slot) {
    const int key = kb + slot;
    return a.keymode == 0 ? key : a.keymode == 1 ? a.n - 1 - key : a.key_row[key];
}

template <class T>
__device__ __forceinline__ void df_outgoing(const DfArgs &a, T *pub, T *nxt, const T *sl, const int *outs, int no,
                                            int kb, int lane) {
    bool gave_up = false;
    for (int o = lane; o < no; o += DF_PF * 64) {
        int sk[DF_PF];
#pragma unroll
        for (int k = 0; k < DF_PF; k++) {
            const int ok = o + k * 64;
            sk[k] = outs[ok < no ? ok : no - 1];
        }
        unsigned todo = 0;
#pragma unroll
        for (int k = 0; k < DF_PF; k++)
            if (o + k * 64 < no) todo |= 1u << k;
        long long spins = 0;
        while (todo && !gave_up) {
            T v[DF_PF];
#pragma unroll
            for (int k = 0; k < DF_PF; k++) v[k] = ld_lds(&sl[sk[k]]);
            unsigned got = 0;
#pragma unroll
            for (int k = 0; k < DF_PF; k++) {
                if (((todo >> k) & 1u) && !Sentinel<T>::is(v[k])) {
                    const int row = df_row_of<T>(a, kb, sk[k]);
                    st_agent(&pub[row], v[k]);
                    nxt[row] = Sentinel<T>::value();
                    got |= 1u << k;
                }
            }
            todo &= ~got;
            if (todo && !got) {
                if (++spins > a.spin_limit) {
                    gave_up = true;  // the unit's own rows are stuck: its compute waves report it
                } else {
                    __builtin_amdgcn_s_sleep(1);
                }
            }
        }
        // After giving up, still re-arm the other half for every row left,
        // so the next sweep starts from sentinels (its own half must hold
        // no value from two sweeps back).
        if (gave_up) {
#pragma unroll
            for (int k = 0; k < DF_PF; k++)
                if ((todo >> k) & 1u) nxt[df_row_of<T>(a, kb, sk[k])] = Sentinel<T>::value();
        }
    }
}

// Solve the row in R (one lane's row of a round): read its parents' LDS
// slots until every one is solved, consuming the ready prefix in order, then
// write the result to the row's slot.  The wavefront leaves when all its
// lanes are done.
template <class T, bool UPPER, int CH>
__device__ __forceinline__ void df_solve(const DfArgs &a, const T *__restrict__ vals, T *sl, DfRow<T, CH> &R) {
    bool pending = R.slot >= 0;
    T s = T(0);
    int pc = 0;
    long long spins = 0;
    while (pending) {
        if (pc == R.nc && R.left > 0) {  // rows longer than one chunk
            R.nc = R.left < CH ? R.left : CH;
            const int last = R.nc - 1;
#pragma unroll
            for (int t = 0; t < CH; t++) {
                const int o = t < last ? t : last;
                const int j = UPPER ? R.jn - o : R.jn + o;
                R.cs[t] = a.scol[j];
                R.cv[t] = vals[j];
            }
            R.jn = UPPER ? R.jn - R.nc : R.jn + R.nc;
            R.left -= R.nc;
            pc = 0;
            // complete these loads here (this also waits for the loads of
            // later rounds; rows this long are rare)
            __builtin_amdgcn_s_waitcnt(0);
        }
        if (pc < R.nc) {
            // read every parent slot of the chunk at once (slots past the
            // row's count repeat its last parent), then consume the ready
            // prefix in order
            T v[CH];
#pragma unroll
            for (int t = 0; t < CH; t++) v[t] = ld_lds(&sl[R.cs[t]]);
#pragma unroll
            for (int t = 0; t < CH; t++) {
                if (t == pc && t < R.nc && !Sentinel<T>::is(v[t])) {
                    s = UPPER ? fma(v[t], R.cv[t], s) : fma(R.cv[t], v[t], s);
                    pc++;
                }
            }
            if (pc < R.nc && ++spins > a.spin_limit) {
                raise_status(a.status, BAAS_TIMEOUT, R.rw & 0x7fffffff);
                pc = R.nc;  // give up on this row's remaining parents
                R.left = 0;
            }
        } else {
            st_lds(&sl[R.slot], (R.bv - s) / R.dv);
            pending = false;
        }
    }
}

// One round of a compute wave on buffer R (loaded DF_DEPTH rounds ago), then
// refill R with the round DF_DEPTH ahead (position qn) from its record m, and
// load the record of the round after that (position qm) into m.
template <class T, bool UPPER, int CH>
__device__ __forceinline__ void df_round(const DfArgs &a, const T *__restrict__ vals, const T *__restrict__ b,
                                         T *sl, DfRow<T, CH> &R, int4 &m, int qn, int qm, int q1) {
    df_solve<T, UPPER, CH>(a, vals, sl, R);
    df_load_row<T, UPPER, CH>(a, vals, b, m, qn < q1, R);
    m = df_meta(a, qm, q1);
}

template <class T, bool UPPER, int CH>
__global__ __launch_bounds__(512) void k_sweep_df(DfArgs a, const T *__restrict__ vals, const T *__restrict__ b,
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
    for (;;) {
        const int u = claim_unit(a.ctx.ticket, &s_unit);
        if (u >= a.nunits) break;
        const int q0 = a.unit_ptr[u], q1 = a.unit_ptr[u + 1], nk = q1 - q0;
        const int e0 = a.ext_ptr[u], ne = a.ext_ptr[u + 1] - e0;
        const int o0 = a.out_ptr[u], no = a.out_ptr[u + 1] - o0;
        const int kb = a.key_begin[u];
        // LDS: one value slot per row of the unit, then one per external parent
        const int nslots = nk + ne;
        for (int i = threadIdx.x; i < nslots; i += blockDim.x) st_lds(&sl[i], Sentinel<T>::value());
        __syncthreads();
        if (wv == 0) {
            df_incoming<T>(a, pub, sl, a.ext_cols + e0, nk, ne, lane);
        } else if (wv == 1) {
            df_outgoing<T>(a, pub, nxt, sl, a.out_slot + o0, no, kb, lane);
        } else {
            // Compute waves: round r of wave cw covers positions
            // q0 + (r * C + cw) * 64 + lane.  Each wave keeps DF_DEPTH rounds
            // of row data in flight in fixed buffers (the round loop is
            // unrolled DF_DEPTH times, so no buffer is ever copied while its
            // loads are outstanding) and the row records DF_DEPTH rounds
            // further ahead, so a round waits for memory only when it ran
            // faster than a load's latency divided by DF_DEPTH.  Results go to
            // LDS only; the unit's x is written out once it is complete.
            const int C = nw - 2, cw = wv - 2, stride = C * 64;
            const int qa = q0 + cw * 64 + lane;
            DfRow<T, CH> r0, r1, r2;
            int4 m0 = df_meta(a, qa, q1), m1 = df_meta(a, qa + stride, q1), m2 = df_meta(a, qa + 2 * stride, q1);
            df_load_row<T, UPPER, CH>(a, vals, b, m0, qa < q1, r0);
            df_load_row<T, UPPER, CH>(a, vals, b, m1, qa + stride < q1, r1);
            df_load_row<T, UPPER, CH>(a, vals, b, m2, qa + 2 * stride < q1, r2);
            m0 = df_meta(a, qa + 3 * stride, q1);
            m1 = df_meta(a, qa + 4 * stride, q1);
            m2 = df_meta(a, qa + 5 * stride, q1);
            for (int base = q0 + cw * 64; base < q1; base += 3 * stride) {
                const int q = base + lane;
                df_round<T, UPPER, CH>(a, vals, b, sl, r0, m0, q + 3 * stride, q + 6 * stride, q1);
                if (base + stride >= q1) break;
                df_round<T, UPPER, CH>(a, vals, b, sl, r1, m1, q + 4 * stride, q + 7 * stride, q1);
                if (base + 2 * stride >= q1) break;
                df_round<T, UPPER, CH>(a, vals, b, sl, r2, m2, q + 5 * stride, q + 8 * stride, q1);
            }
        }
        __syncthreads();
        // the unit's solution, slot i holding the row with key kb + i
        for (int i = threadIdx.x; i < nk; i += blockDim.x) {
            const int key = kb + i;
            const int row = a.keymode == 0 ? key : a.keymode == 1 ? a.n - 1 - key : a.key_row[key];
            x[row] = sl[i];
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

// ---------------------------------------------------------------------------
// Lane kernel (dataflow plans, variant 2).  Same protocol and the same
// summation order as k_sweep_df, with three differences:
//  * the unit's row records and off-diagonal entries are copied into LDS
//    when the unit starts, so the solving loop issues no global loads and
//    never waits on the memory counter;
//  * every lane walks its own positions (position p, p + stride, ...) in one
//    flat loop, so it moves to its next row as soon as its current row is
//    solved instead of waiting for the other lanes of its wavefront;
//  * a lane publishes its own row for later units (the value is the flag,
//    as in the outgoing exchange), so only wavefront 0 exchanges (incoming).
// Load a chunk of a row's parents: their value slots (times nr, plus the
// lane's right-hand side r) and coefficients; slots past the chunk repeat its
// last entry, an empty chunk points at slot 0 with coefficient 0.
template <class T, bool UPPER, int CH>
__device__ __forceinline__ void lane_chunk(const T *lev, const int *les, int eo, int k, int nc, int nr, int r,
                                           int (&cs)[CH], T (&cv)[CH]) {
    const int last = nc > 0 ? nc - 1 : 0;
#pragma unroll
    for (int t = 0; t < CH; t++) {
        const int i = eo + k + (t < last ? t : last);
        cs[t] = nc > 0 ? les[i] * nr + r : 0;
        cv[t] = nc > 0 ? lev[i] : T(0);
    }
}

// One lane's items: item it is position it / nr of the unit, right-hand side
// it % nr; the lane takes it, it + stride, ... in one flat loop.  ONE: every
// row's parents fit in one chunk (the multi-chunk steps are compiled out).
template <class T, bool UPPER, int CH, bool ONE>
__device__ __forceinline__ void lane_solve(const DfArgs &a, T *sl, const T *lb, const T *lg, const int2 *lrec,
                                           const int *lslot, const int *lrow, const T *lev, const int *les, T *pub,
                                           T *nxt, int it, int stride, int nk) {
    const int nr = a.nrhs, items = nk * nr;
    bool have = it < items;
    int eo = 0, cnt = 0, k = 0, nc = 0, pc = 0, vslot = 0, vrow = 0, r = 0;
    bool ext = false;
    T s = T(0), bv = T(0), dv = T(1);
    int cs[CH];
    T cv[CH];
    long long spins = 0;
    // start the item `it`: its row record, b and first chunk (LDS only)
    auto start = [&]() {
        const int p = nr == 1 ? it : it / nr;
        r = it - p * nr;
        const int2 rc = lrec[p];
        eo = rc.x;
        cnt = rc.y & 0x7fffffff;
        ext = rc.y < 0;
        vslot = lslot[p] * nr + r;
        vrow = lrow[p] * nr + r;
        bv = lb[p * nr + r];
        dv = lg[p];
        k = 0;
        pc = 0;
        s = T(0);
        spins = 0;
        nc = cnt < CH ? cnt : CH;
        lane_chunk<T, UPPER, CH>(lev, les, eo, 0, nc, nr, r, cs, cv);
    };
    if (have) start();
    // One loop whose exit is uniform across the wavefront and whose body is
    // a single divergent step, so no lane waits at an inner loop's exit for
    // the others: a lane whose row is solved starts its next item in the
    // same iteration in which the others keep polling.
    while (__ballot(have) != 0) {
        if (!have) continue;
        {
            T v[CH];
#pragma unroll
            for (int t = 0; t < CH; t++) v[t] = ld_lds(&sl[cs[t]]);
#pragma unroll
            for (int t = 0; t < CH; t++) {
                if (t == pc && t < nc && !Sentinel<T>::is(v[t])) {
                    s = UPPER ? fma(v[t], cv[t], s) : fma(cv[t], v[t], s);
                    pc++;
                }
            }
        }
        if (pc == nc) {
            k += nc;
            if (!ONE && k < cnt) {  // rows longer than one chunk
                nc = cnt - k < CH ? cnt - k : CH;
                pc = 0;
                lane_chunk<T, UPPER, CH>(lev, les, eo, k, nc, nr, r, cs, cv);
            } else {
                const T xv = (bv - s) / dv;
                st_lds(&sl[vslot], xv);
                if (ext) {
                    st_agent(&pub[vrow], xv);
                    nxt[vrow] = Sentinel<T>::value();
                }
                it += stride;
                have = it < items;
                if (have) start();
            }
        } else if (++spins > a.spin_limit) {
            raise_status(a.status, BAAS_TIMEOUT, vrow);
            pc = nc;  // give up on this row's remaining parents
            cnt = k + nc;
        }
    }
}

template <class T, bool UPPER, int CH, bool ONE>
__global__ __launch_bounds__(1024) void k_sweep_lane(DfArgs a, const int *__restrict__ eoff,
                                                    const int *__restrict__ eptr, const T *__restrict__ vals,
                                                    const T *__restrict__ b, T *x, T *xp) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ int s_unit;
    if (a.skip && ld_agent(a.skip)) return;
    const int epoch = read_epoch(a.ctx);
    const int par = ld_agent(a.pcount) & 1;
    const int nr = a.nrhs;
    T *pub = xp + (size_t)par * a.n * nr;
    T *nxt = xp + (size_t)(par ^ 1) * a.n * nr;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (;;) {
        const int u = claim_unit(a.ctx.ticket, &s_unit);
        if (u >= a.nunits) break;
        const int q0 = a.unit_ptr[u], q1 = a.unit_ptr[u + 1], nk = q1 - q0;
        const int e0 = a.ext_ptr[u], ne = a.ext_ptr[u + 1] - e0;
        const int kb = a.key_begin[u];
        const LaneLayout lay = lane_layout(nk, ne, eptr[u + 1] - eptr[u], sizeof(T), nr);
        T *sl = reinterpret_cast<T *>(smem_raw + lay.sl);
        T *lb = reinterpret_cast<T *>(smem_raw + lay.bv);
        T *lg = reinterpret_cast<T *>(smem_raw + lay.dv);
        int2 *lrec = reinterpret_cast<int2 *>(smem_raw + lay.rec);
        int *lslot = reinterpret_cast<int *>(smem_raw + lay.slot);
        int *lrow = reinterpret_cast<int *>(smem_raw + lay.row);
        T *lev = reinterpret_cast<T *>(smem_raw + lay.ev);
        int *les = reinterpret_cast<int *>(smem_raw + lay.es);
        for (int i = threadIdx.x; i < (nk + ne) * nr; i += blockDim.x) st_lds(&sl[i], Sentinel<T>::value());
        for (int q = threadIdx.x; q < nk; q += blockDim.x) {
            const int4 m = a.meta[q0 + q];
            const int cnt = m.z, eo = eoff[q0 + q], row = m.x & 0x7fffffff;
            if (nr == 1) lb[q] = b[row];
            lg[q] = vals[UPPER ? m.y - cnt : m.y + cnt];
            lrec[q] = make_int2(eo, cnt | (m.x & (int)0x80000000));
            lslot[q] = m.w;
            lrow[q] = row;
            for (int e = 0; e < cnt; e++) {
                const int j = UPPER ? m.y - e : m.y + e;
                lev[eo + e] = vals[j];
                les[eo + e] = a.scol[j];
            }
        }
        if (nr > 1) {  // b of every (position, right-hand side), spread over all threads
            __syncthreads();
            for (int i = threadIdx.x; i < nk * nr; i += blockDim.x) {
                const int q = i / nr;
                lb[i] = b[(size_t)lrow[q] * nr + (i - q * nr)];
            }
        }
        __syncthreads();
        if (wv == 0) {
            df_incoming<T>(a, pub, sl, a.ext_cols + e0, nk, ne, lane);
        } else {
            lane_solve<T, UPPER, CH, ONE>(a, sl, lb, lg, lrec, lslot, lrow, lev, les, pub, nxt, (wv - 1) * 64 + lane,
                                     (nw - 1) * 64, nk);
        }
        __syncthreads();
        // the unit's solution, slot i holding the row with key kb + i
        for (int i = threadIdx.x; i < nk * nr; i += blockDim.x) {
            const int key = kb + (nr == 1 ? i : i / nr);
            const int row = a.keymode == 0 ? key : a.keymode == 1 ? a.n - 1 - key : a.key_row[key];
            x[(size_t)row * nr + (nr == 1 ? 0 : i % nr)] = sl[i];
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

// ---------------------------------------------------------------------------
// Multi-RHS lane kernel (dataflow plans, variant 3).  The lane kernel's
// (row, right-hand side) items, with the exchange moved into the compute
// lanes: a lane reads a parent in another unit straight from the
// publication half (the value is the flag, one load per poll) instead of
// waiting for an exchange wavefront to stage it, reads b when its item
// starts and writes x when the item is solved.  So LDS holds only the
// unit's own values (nk x nrhs), every wavefront computes, and with k
// right-hand sides the k lanes of a row load and store its k contiguous
// values together.  Same summation order and division as every other sweep.
template <class T, bool UPPER, int CH>
__device__ __forceinline__ void mr_chunk(const T *lev, const int *les, int eo, int k, int nc, int nr, int r,
                                         int (&cs)[CH], T (&cv)[CH]) {
    const int last = nc > 0 ? nc - 1 : 0;
#pragma unroll
    for (int t = 0; t < CH; t++) {
        const int i = eo + k + (t < last ? t : last);
        const int e = les[i];
        // LDS slot x nr + r for the unit's rows; ~(row x nr + r) for others'
        cs[t] = nc > 0 ? (e >= 0 ? e * nr + r : ~((~e) * nr + r)) : 0;
        cv[t] = nc > 0 ? lev[i] : T(0);
    }
}

// V right-hand sides per lane: a lane's item is position p with the V
// contiguous right-hand sides r .. r + V - 1 (V divides nrhs), so a row's
// control flow is shared by V values.
template <class T, int V>
struct MrVec {
    T e[V];
};

template <class T, int V>
__device__ __forceinline__ MrVec<T, V> ld_lds_vec(const T *p) {
    MrVec<T, V> w;
    if constexpr (V == 1) {
        w.e[0] = ld_lds(p);
    } else {
        // one LDS read of V values (polled: volatile, never hoisted); a torn
        // read only shows a value as not yet there
        typedef T vt __attribute__((ext_vector_type(V)));
        const vt u = *reinterpret_cast<const volatile vt *>(p);
#pragma unroll
        for (int i = 0; i < V; i++) w.e[i] = u[i];
    }
    return w;
}

template <class T, bool UPPER, int CH, bool ONE, bool SB, int V>
__device__ __forceinline__ void mr_solve(const DfArgs &a, T *sl, const T *lb, const T *lg, const int2 *lrec,
                                         const int *lslot, const int *lrow, const T *lev, const int *les,
                                         const T *__restrict__ b, T *x, const T *pub, T *nxt, int it,
                                         int stride, int nk) {
    const int nr = a.nrhs, ng = nr / V, items = nk * ng;
    bool have = it < items;
    int eo = 0, cnt = 0, k = 0, nc = 0, pc = 0, vslot = 0, vrow = 0, r = 0;
    bool ext = false;
    T s[V], bv[V], dv = T(1);
    int cs[CH];
    T cv[CH];
    long long spins = 0;
    auto start = [&]() {
        const int p = ng == 1 ? it : it / ng;
        r = (it - p * ng) * V;
        const int2 rc = lrec[p];
        eo = rc.x;
        cnt = rc.y & 0x7fffffff;
        ext = rc.y < 0;
        vslot = lslot[p] * nr + r;
        vrow = lrow[p] * nr + r;
#pragma unroll
        for (int i = 0; i < V; i++) {
            bv[i] = SB ? lb[p * nr + r + i] : b[vrow + i];
            s[i] = T(0);
        }
        dv = lg[p];
        k = 0;
        pc = 0;
        spins = 0;
        nc = cnt < CH ? cnt : CH;
        mr_chunk<T, UPPER, CH>(lev, les, eo, 0, nc, nr, r, cs, cv);
    };
    if (have) start();
    // The body runs in every lane of a wavefront that still has work, the
    // lanes without any masked by predicates rather than branches: every
    // divergent branch costs scalar instructions (exec-mask saves and
    // restores), and the CU's one scalar unit issues them for all its
    // wavefronts.
    while (__ballot(have) != 0) {
        const int pc0 = pc;
        {
            // LDS slots for every parent (other units' rows read slot 0 and
            // are replaced below), then, only in lanes that still wait for a
            // parent of another unit, those values from memory.  The memory
            // wait stays inside that branch: a wait after it would make
            // every wavefront of the unit wait for its lanes' outstanding
            // stores (loads and stores share one counter) on every poll.
            MrVec<T, V> v[CH];
            bool need = false;
#pragma unroll
            for (int t = 0; t < CH; t++) {
                v[t] = ld_lds_vec<T, V>(&sl[cs[t] >= 0 ? cs[t] : 0]);
                need |= have && cs[t] < 0 && t >= pc && t < nc;
            }
            if (need) {
                MrVec<T, V> gv[CH];
#pragma unroll
                for (int t = 0; t < CH; t++)
#pragma unroll
                    for (int i = 0; i < V; i++)
                        gv[t].e[i] = (cs[t] < 0 && t >= pc && t < nc) ? ld_agent(&pub[~cs[t] + i])
                                                                      : Sentinel<T>::value();
                __builtin_amdgcn_s_waitcnt(0);
#pragma unroll
                for (int t = 0; t < CH; t++)
                    if (cs[t] < 0) v[t] = gv[t];
            }
            // consume the ready prefix in order, by selects
#pragma unroll
            for (int t = 0; t < CH; t++) {
                bool take = have && t == pc && t < nc;
#pragma unroll
                for (int i = 0; i < V; i++) take = take && !Sentinel<T>::is(v[t].e[i]);
#pragma unroll
                for (int i = 0; i < V; i++) {
                    const T f = UPPER ? fma(v[t].e[i], cv[t], s[i]) : fma(cv[t], v[t].e[i], s[i]);
                    s[i] = take ? f : s[i];
                }
                pc += take ? 1 : 0;
            }
        }
        if (have && pc == nc) {
            k += nc;
            if (!ONE && k < cnt) {
                nc = cnt - k < CH ? cnt - k : CH;
                pc = 0;
                mr_chunk<T, UPPER, CH>(lev, les, eo, k, nc, nr, r, cs, cv);
            } else {
                T xv[V];
#pragma unroll
                for (int i = 0; i < V; i++) {
                    xv[i] = (bv[i] - s[i]) / dv;
                    st_lds(&sl[vslot + i], xv[i]);
                    x[vrow + i] = xv[i];
                }
                if (ext) {
#pragma unroll
                    for (int i = 0; i < V; i++) {
                        st_agent(const_cast<T *>(&pub[vrow + i]), xv[i]);
                        nxt[vrow + i] = Sentinel<T>::value();
                    }
                }
                it += stride;
                have = it < items;
                if (have) start();
            }
        } else {
            spins += have ? 1 : 0;
            // a wavefront none of whose lanes moved yields its issue slots
            // to the others of its SIMD for a moment
            if (__ballot(pc != pc0) == 0) __builtin_amdgcn_s_sleep(1);
            if (spins > a.spin_limit) {
                raise_status(a.status, BAAS_TIMEOUT, vrow);
                pc = nc;  // give up on this row's remaining parents
                cnt = k + nc;
            }
        }
    }
}

template <class T, bool UPPER, int CH, bool ONE, bool SB, int V>
__global__ __launch_bounds__(1024) void k_sweep_mr(DfArgs a, const int *__restrict__ eoff,
                                                  const int *__restrict__ eptr, const T *__restrict__ vals,
                                                  const T *__restrict__ b, T *x, T *xp) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ int s_unit;
    if (a.skip && ld_agent(a.skip)) return;
    const int epoch = read_epoch(a.ctx);
    const int par = ld_agent(a.pcount) & 1;
    const int nr = a.nrhs;
    T *pub = xp + (size_t)par * a.n * nr;
    T *nxt = xp + (size_t)(par ^ 1) * a.n * nr;
    for (;;) {
        const int u = claim_unit(a.ctx.ticket, &s_unit);
        if (u >= a.nunits) break;
        const int q0 = a.unit_ptr[u], q1 = a.unit_ptr[u + 1], nk = q1 - q0;
        if (a.unit_skip && a.unit_skip[u]) {
            // the other half of every published row must be the sentinel
            // when the next sweep starts, as after a solved unit
            for (int i = threadIdx.x; i < nk * nr; i += blockDim.x) {
                const int q = i / nr;
                const int mx = a.meta[q0 + q].x;
                if (mx < 0) nxt[(size_t)(mx & 0x7fffffff) * nr + (i - q * nr)] = Sentinel<T>::value();
            }
            continue;
        }
        const int e0 = a.ext_ptr[u];
        const MrLayout lay = mr_layout(nk, eptr[u + 1] - eptr[u], sizeof(T), nr, SB);
        T *sl = reinterpret_cast<T *>(smem_raw + lay.sl);
        T *lb = reinterpret_cast<T *>(smem_raw + lay.bv);
        T *lg = reinterpret_cast<T *>(smem_raw + lay.dv);
        int2 *lrec = reinterpret_cast<int2 *>(smem_raw + lay.rec);
        int *lslot = reinterpret_cast<int *>(smem_raw + lay.slot);
        int *lrow = reinterpret_cast<int *>(smem_raw + lay.row);
        T *lev = reinterpret_cast<T *>(smem_raw + lay.ev);
        int *les = reinterpret_cast<int *>(smem_raw + lay.es);
        for (int i = threadIdx.x; i < nk * nr; i += blockDim.x) st_lds(&sl[i], Sentinel<T>::value());
        for (int q = threadIdx.x; q < nk; q += blockDim.x) {
            const int4 m = a.meta[q0 + q];
            const int cnt = m.z, eo = eoff[q0 + q];
            lg[q] = vals[UPPER ? m.y - cnt : m.y + cnt];
            lrec[q] = make_int2(eo, cnt | (m.x & (int)0x80000000));
            lslot[q] = m.w;
            lrow[q] = m.x & 0x7fffffff;
            for (int e = 0; e < cnt; e++) {
                const int j = UPPER ? m.y - e : m.y + e;
                const int sc = a.scol[j];
                lev[eo + e] = vals[j];
                les[eo + e] = sc < nk ? sc : ~a.ext_cols[e0 + sc - nk];
            }
        }
        if (SB) {  // b of every (position, right-hand side), contiguous per row
            __syncthreads();
            for (int i = threadIdx.x; i < nk * nr; i += blockDim.x) {
                const int q = nr == 1 ? i : i / nr;
                lb[i] = b[(size_t)lrow[q] * nr + (i - q * nr)];
            }
        }
        __syncthreads();
        mr_solve<T, UPPER, CH, ONE, SB, V>(a, sl, lb, lg, lrec, lslot, lrow, lev, les, b, x, pub, nxt, threadIdx.x,
                                           blockDim.x, nk);
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
size_t pipe_lds(const Plan &p, int ts);
template <class T>
int launch_pipe(baas_handle *h, baas_schedule *s, const Plan &p, int uplo, const T *vals, const T *b, T *x, T *xp,
                int nrhs, hipStream_t stream, const int *skip);

template <class T, bool UP>
static int launch_df(baas_handle *h, const Plan &p, const DfArgs &a, const T *vals, const T *b, T *x, T *xp,
                     size_t lds, hipStream_t stream) {
    // four parents per chunk cover stencil-like rows; wider rows take eight
    auto kern = p.max_deg <= 4 ? k_sweep_df<T, UP, 4> : k_sweep_df<T, UP, 8>;
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
static int launch_lane(baas_handle *h, const Plan &p, const DfArgs &a, const T *vals, const T *b, T *x, T *xp,
                       size_t lds, hipStream_t stream) {
    // Chunk width: one chunk per row, as narrow as the longest row allows,
    // while rows have at most eight parents (every polled slot costs each
    // step); two for matrices with longer rows, whose lanes then make more
    // but cheaper steps (measured faster than four, eight or twelve, and
    // than one, on rows of up to 20-40 parents).
    auto kern = p.max_deg <= 2   ? k_sweep_lane<T, UP, 2, true>
                : p.max_deg <= 3 ? k_sweep_lane<T, UP, 3, true>
                : p.max_deg <= 4 ? k_sweep_lane<T, UP, 4, true>
                : p.max_deg <= 6 ? k_sweep_lane<T, UP, 6, true>
                : p.max_deg <= 8 ? k_sweep_lane<T, UP, 8, true>
                                 : k_sweep_lane<T, UP, 2, false>;
    int per_cu = 0;
    BAAS_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, p.block, lds));
    if (per_cu <= 0) return BAAS_UNSUPPORTED;
    // at most 16 wavefronts per CU: more resident units share the SIMDs'
    // issue slots and slow every poll loop (whatever the register count
    // would allow)
    per_cu = std::min(per_cu, std::max(1, 1024 / p.block));
    long long grid = (long long)per_cu * h->num_cus;
    if (grid > p.nunits) grid = p.nunits;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(p.block), lds, stream, a, p.lane_eoff.p, p.lane_eptr.p, vals,
                       b, x, xp);
    BAAS_HIP_TRY(hipGetLastError());
    return BAAS_OK;
}

template <class T, bool UP>
static int launch_mr(baas_handle *h, const Plan &p, const DfArgs &a, const T *vals, const T *b, T *x, T *xp,
                     size_t lds, bool sb, hipStream_t stream) {
    using K = void (*)(DfArgs, const int *, const int *, const T *, const T *, T *, T *);
    // chunk width as in the lane kernel (rows of up to four parents, else
    // eight or two-wide chunks).  One value per lane: several right-hand
    // sides per lane chain their IEEE divisions (they share VCC) and
    // measured slower on every case.
    auto pick = [&](auto sbtag) -> K {
        constexpr bool S = decltype(sbtag)::value;
        return p.max_deg <= 2   ? k_sweep_mr<T, UP, 2, true, S, 1>
               : p.max_deg <= 3 ? k_sweep_mr<T, UP, 3, true, S, 1>
               : p.max_deg <= 4 ? k_sweep_mr<T, UP, 4, true, S, 1>
               : p.max_deg <= 8 ? k_sweep_mr<T, UP, 8, true, S, 1>
                                : k_sweep_mr<T, UP, 2, false, S, 1>;
    };
    K kern = sb ? pick(std::true_type{}) : pick(std::false_type{});
    int per_cu = 0;
    BAAS_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, p.block, lds));
    if (per_cu <= 0) return BAAS_UNSUPPORTED;
    per_cu = std::min(per_cu, std::max(1, 1024 / p.block));
    long long grid = (long long)per_cu * h->num_cus;
    if (grid > p.nunits) grid = p.nunits;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(p.block), lds, stream, a, p.lane_eoff.p, p.lane_eptr.p, vals,
                       b, x, xp);
    BAAS_HIP_TRY(hipGetLastError());
    return BAAS_OK;
}

// ---------------------------------------------------------------------------
// Level-set sweep (LEVELS plans).  One persistent launch walks the row level
// sets; every workgroup takes a grid-stride share of each level's (row, rhs)
// items, and a device-wide barrier separates levels (merge/SpTRSV.cpp:41-80
// with the OpenMP barrier replaced).  Rows of one level are independent, so
// no per-row flag is needed; x is written through and read at device scope,
// so the previous levels' values are visible once the barrier is passed.  The barrier counter lives in the handle's
// counter block and is reset by the last workgroup out.
struct LvArgs {
    int n, nlev, nrhs;
    const int *lptr;
    const int4 *meta;  // per position: row, first off-diagonal entry, count
    const int *colidx;
    Ctx ctx;
    int *bar;
    const int *skip;
    DevStatus *status;
    long long spin_limit;
};

// One (row, rhs) item of a level with its first CH parents' columns and
// values, the right-hand side and the diagonal: everything but x, so it can
// be loaded before the barrier that makes the parents' x final.
template <class T, int CH>
struct LvItem {
    int row, r, j0, cnt;
    int c[CH];
    T w[CH];
    T bv, dv;
};

template <class T, bool UPPER, int CH>
__device__ __forceinline__ void lv_load(const LvArgs &a, const T *__restrict__ vals, const T *__restrict__ b, int q,
                                        int r, LvItem<T, CH> &I) {
    const int4 m = a.meta[q];
    I.row = m.x;
    I.j0 = m.y;
    I.cnt = m.z;
    I.r = r;
    const int nc = m.z < CH ? m.z : CH;
#pragma unroll
    for (int t = 0; t < CH; t++) {
        const int o = nc > 0 ? (t < nc ? t : nc - 1) : 0;
        const int j = UPPER ? m.y - o : m.y + o;
        I.c[t] = a.colidx[j];
        I.w[t] = vals[j];
    }
    I.bv = b[(size_t)m.x * a.nrhs + r];
    I.dv = vals[UPPER ? m.y - m.z : m.y + m.z];
}

// Solve one item: parents CH at a time (the first chunk already loaded);
// every x load of a chunk is in flight together and the sum keeps the
// reference order (slots past the count repeat the last entry, not summed).
template <class T, bool UPPER, int CH>
__device__ __forceinline__ void lv_solve(const LvArgs &a, const T *__restrict__ vals, T *x, LvItem<T, CH> &I) {
    const int nrhs = a.nrhs;
    T s = T(0);
    for (int k0 = 0; k0 < I.cnt; k0 += CH) {
        const int nc = I.cnt - k0 < CH ? I.cnt - k0 : CH;
        if (k0 > 0) {
#pragma unroll
            for (int t = 0; t < CH; t++) {
                const int o = k0 + (t < nc ? t : nc - 1);
                const int j = UPPER ? I.j0 - o : I.j0 + o;
                I.c[t] = a.colidx[j];
                I.w[t] = vals[j];
            }
        }
        T xv[CH];
#pragma unroll
        for (int t = 0; t < CH; t++) xv[t] = ld_agent(&x[(size_t)I.c[t] * nrhs + I.r]);
#pragma unroll
        for (int t = 0; t < CH; t++)
            if (t < nc) s = UPPER ? fma(xv[t], I.w[t], s) : fma(I.w[t], xv[t], s);
    }
    st_agent(&x[(size_t)I.row * nrhs + I.r], (I.bv - s) / I.dv);
}

template <class T, bool UPPER, int CH>
__global__ __launch_bounds__(1024) void k_sweep_levels(LvArgs a, const T *__restrict__ vals, const T *__restrict__ b,
                                                      T *x) {
    if (a.skip && ld_agent(a.skip)) return;
    const int epoch = read_epoch(a.ctx);
    const int nrhs = a.nrhs;
    const long long nth = (long long)gridDim.x * blockDim.x;
    const long long gt = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    // this thread's first item of each level is loaded before the barrier
    // that precedes the level
    LvItem<T, CH> first;
    bool have = false;
    {
        const int l = UPPER ? a.nlev - 1 : 0;
        const int qb = a.lptr[l];
        if (gt < (long long)(a.lptr[l + 1] - qb) * nrhs) {
            lv_load<T, UPPER, CH>(a, vals, b, qb + (int)(gt / nrhs), (int)(gt % nrhs), first);
            have = true;
        }
    }
    for (int li = 0; li < a.nlev; li++) {
        const int l = UPPER ? a.nlev - 1 - li : li;
        const int qb = a.lptr[l], qe = a.lptr[l + 1];
        const long long items = (long long)(qe - qb) * nrhs;
        if (have) lv_solve<T, UPPER, CH>(a, vals, x, first);
        for (long long t = gt + nth; t < items; t += nth) {
            LvItem<T, CH> it;
            lv_load<T, UPPER, CH>(a, vals, b, qb + (int)(t / nrhs), (int)(t % nrhs), it);
            lv_solve<T, UPPER, CH>(a, vals, x, it);
        }
        if (li + 1 < a.nlev) {
            const int ln = UPPER ? l - 1 : l + 1;
            const int nb = a.lptr[ln];
            have = gt < (long long)(a.lptr[ln + 1] - nb) * nrhs;
            if (have) lv_load<T, UPPER, CH>(a, vals, b, nb + (int)(gt / nrhs), (int)(gt % nrhs), first);
            if (gridDim.x == 1) {
                wait_stores();  // a single workgroup needs only its own barrier
                __syncthreads();
            } else {
                grid_barrier(a.bar, (li + 1) * (int)gridDim.x, a.status, a.spin_limit);
            }
        }
    }
    if (threadIdx.x == 0) {
        __threadfence();
        const int d = atomicAdd(a.ctx.done, 1);
        if (d == (int)gridDim.x - 1) {
            reset_barrier(a.bar);
            atomicExch(a.ctx.ticket, 0);
            atomicExch(a.ctx.epoch, epoch == 0x7fffffff ? 1 : epoch + 1);
            __threadfence();
            atomicExch(a.ctx.done, 0);
        }
    }
}

template <class T>
static int launch_levels(baas_handle *h, baas_schedule *s, const Plan &p, int uplo, const T *vals, const T *b, T *x,
                         int nrhs, hipStream_t stream, const int *skip) {
    LvArgs a;
    a.n = s->n;
    a.nlev = p.nlev;
    a.nrhs = nrhs;
    a.lptr = p.lptr.p;
    a.meta = reinterpret_cast<const int4 *>(p.meta.p);
    a.colidx = uplo ? s->u_colidx.p : s->l_colidx.p;
    a.ctx = Ctx{h->counters, h->counters + 1, h->counters + 2};
    a.bar = h->counters + BAAS_BARRIER_BASE;
    a.skip = skip;
    a.status = h->status;
    a.spin_limit = h->spin_limit;
    // parents gathered per chunk: as many as the longest row has, up to
    // four (padded slots would be loaded for nothing), else eight at a time
    auto kern = p.max_deg <= 2   ? (uplo ? k_sweep_levels<T, true, 2> : k_sweep_levels<T, false, 2>)
                : p.max_deg <= 3 ? (uplo ? k_sweep_levels<T, true, 3> : k_sweep_levels<T, false, 3>)
                : p.max_deg <= 4 ? (uplo ? k_sweep_levels<T, true, 4> : k_sweep_levels<T, false, 4>)
                                 : (uplo ? k_sweep_levels<T, true, 8> : k_sweep_levels<T, false, 8>);
    int per_cu = 0;
    BAAS_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, p.block, 0));
    if (per_cu <= 0) return BAAS_UNSUPPORTED;
    // every workgroup must be resident for the barrier; size the grid to
    // the widest level
    int widest = 1;
    for (int l = 0; l + 1 < (int)s->row_level_ptr.size(); l++)
        widest = std::max(widest, s->row_level_ptr[l + 1] - s->row_level_ptr[l]);
    // at most two workgroups per CU: every workgroup arrives at every
    // barrier, so fewer, fuller workgroups make the barriers cheaper
    long long need = ((long long)widest * nrhs + p.block - 1) / p.block;
    long long grid = std::min<long long>((long long)std::min(per_cu, 2) * h->num_cus, need);
    if (p.lv_cap > 0 && grid > p.lv_cap) grid = p.lv_cap;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(p.block), 0, stream, a, vals, b, x);
    BAAS_HIP_TRY(hipGetLastError());
    return BAAS_OK;
}

template <class T>
int sweep(baas_handle *h, baas_schedule *s, int uplo, const T *vals, const T *b, T *x, int nrhs,
          hipStream_t stream, const int *skip, void *fixed_xp, size_t fixed_bytes, const int *unit_skip) {
    if (!h || !s || !vals || !b || !x || nrhs < 1 || (uplo != 0 && uplo != 1)) return BAAS_INVALID_ARG;
    if ((const void *)b == (const void *)x) return BAAS_INVALID_ARG;
    // a fixed publication area must hold both halves at any plan's layout
    if (fixed_xp && 2 * (size_t)s->n * (size_t)nrhs * sizeof(T) > fixed_bytes) return BAAS_TOO_LARGE;
    Plan &p = uplo ? s->bwd : s->fwd;
    if (p.mode == BAAS_PLAN_LEVELS) return launch_levels<T>(h, s, p, uplo, vals, b, x, nrhs, stream, skip);
    // multi-RHS lane kernel: b staged in LDS where it fits (dataflow 3), or
    // always read from memory (dataflow 4)
    const bool mr = (p.dataflow == 3 || p.dataflow == 4) && p.block >= 64 &&
                    p.mr_lds((int)sizeof(T), nrhs, false) <= lds_limit(h);
    const bool mr_sb = mr && p.dataflow == 3 && p.mr_lds((int)sizeof(T), nrhs, true) <= lds_limit(h);
    const bool lane = mr || (p.dataflow == 2 && p.block >= 128 && p.lane_lds((int)sizeof(T), nrhs) <= lds_limit(h));
    // the pipelined grid sweep (PIPE plans): one value per row published
    const bool pipe = !lane && p.pipe && p.dataflow == 5 && pipe_lds(p, (int)sizeof(T)) <= lds_limit(h);
    if (lane || pipe ||
        (nrhs == 1 && p.dataflow == 1 && p.block >= 192 && p.block <= 512 &&
         p.df_lds[sizeof(T) == 8] <= lds_limit(h))) {
        // publication halves (one value per row and right-hand side, or a
        // PIPE plan's box faces per right-hand side), armed with this value
        // type's sentinel; any change of layout re-arms them
        const int xr = lane ? nrhs : 1;
        const size_t half = pipe ? (size_t)p.pipe_half * nrhs : (size_t)s->n * xr;
        const int xkey = pipe ? -nrhs : xr;
        T *xp = reinterpret_cast<T *>(p.xp_raw.p);
        if (fixed_xp) {
            // shared by every call on this GPU and left holding the last
            // call's values: arm both halves (a fill kernel, no memset, so
            // the call can be captured in a graph)
            const size_t cnt = 2 * half;
            if (cnt * sizeof(T) > fixed_bytes) return BAAS_TOO_LARGE;
            xp = reinterpret_cast<T *>(fixed_xp);
            hipLaunchKernelGGL(k_fill_sentinel<T>, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, stream, cnt, xp);
            BAAS_HIP_TRY(hipGetLastError());
        } else if (p.xp_type != (int)sizeof(T) || p.xp_nrhs != xkey) {
            const size_t cnt = 2 * half;
            BAAS_TRY(p.xp_raw.alloc(cnt * sizeof(T)));
            hipLaunchKernelGGL(k_fill_sentinel<T>, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, stream, cnt,
                               reinterpret_cast<T *>(p.xp_raw.p));
            BAAS_HIP_TRY(hipGetLastError());
            p.xp_type = (int)sizeof(T);
            p.xp_nrhs = xkey;
            xp = reinterpret_cast<T *>(p.xp_raw.p);
        }
        DfArgs d;
        d.nrhs = xr;
        d.pcount = p.pcount.p;
        d.n = s->n;
        d.nunits = p.nunits;
        d.keymode = p.keymode;
        d.meta = reinterpret_cast<const int4 *>(p.meta.p);
        d.scol = p.scol.p;
        d.unit_ptr = p.unit_ptr.p;
        d.key_begin = p.key_begin.p;
        d.out_ptr = p.out_ptr.p;
        d.out_slot = p.out_slot.p;
        d.key_row = p.key_row.p;
        d.ext_ptr = p.ext_ptr.p;
        d.ext_cols = p.ext_cols.p;
        d.ctx = Ctx{h->counters, h->counters + 1, h->counters + 2};
        d.skip = skip;
        d.unit_skip = mr ? unit_skip : nullptr;
        d.status = h->status;
        d.spin_limit = h->spin_limit;
        if (pipe) return launch_pipe<T>(h, s, p, uplo, vals, b, x, xp, nrhs, stream, skip);
        if (mr) {
            const size_t lds = p.mr_lds((int)sizeof(T), nrhs, mr_sb);
            return uplo ? launch_mr<T, true>(h, p, d, vals, b, x, xp, lds, mr_sb, stream)
                        : launch_mr<T, false>(h, p, d, vals, b, x, xp, lds, mr_sb, stream);
        }
        if (lane) {
            const size_t lds = p.lane_lds((int)sizeof(T), nrhs);
            return uplo ? launch_lane<T, true>(h, p, d, vals, b, x, xp, lds, stream)
                        : launch_lane<T, false>(h, p, d, vals, b, x, xp, lds, stream);
        }
        const size_t lds = p.df_lds[sizeof(T) == 8];
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
    size_t lds = (size_t)p.max_unit_keys * nrhs * sizeof(T);
    bool use_lds = lds <= lds_limit(h);
    if (uplo)
        return use_lds ? launch_sweep_k<T, true, true>(h, p, a, vals, b, x, lds, stream)
                       : launch_sweep_k<T, true, false>(h, p, a, vals, b, x, 0, stream);
    return use_lds ? launch_sweep_k<T, false, true>(h, p, a, vals, b, x, lds, stream)
                   : launch_sweep_k<T, false, false>(h, p, a, vals, b, x, 0, stream);
}

template int sweep<double>(baas_handle *, baas_schedule *, int, const double *, const double *, double *, int,
                           hipStream_t, const int *, void *, size_t, const int *);
template int sweep<float>(baas_handle *, baas_schedule *, int, const float *, const float *, float *, int,
                          hipStream_t, const int *, void *, size_t, const int *);

// ---------------------------------------------------------------------------
// Scratch of the allocation-free entry point (baas_sptrsv_f64_noalloc): the
// two publication halves of the dataflow sweeps, 2 x n x nrhs doubles, in a
// fixed array that each GPU holds from the library's load until the process
// ends.  In host code `scratch` names a host-side stand-in, not device
// memory: kernels get the device address found by hipGetSymbolAddress, one
// per device.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)BAAS_NOALLOC_MAX_VALUES * sizeof(double);
static __device__ __align__(256) double scratch[BAAS_NOALLOC_MAX_VALUES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static void *g_scratch_dev[NOALLOC_MAX_DEVICES];
static std::mutex g_scratch_mu;

int noalloc_scratch_init(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return BAAS_UNSUPPORTED;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_dev[device]) return BAAS_OK;
    void *ptr = nullptr;
    BAAS_HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return BAAS_HIP_ERROR;
    g_scratch_dev[device] = ptr;
    return BAAS_OK;
}

static void *noalloc_scratch(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_dev[device];
}

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

extern "C" int baas_sptrsv_f64_noalloc(baas_handle *h, baas_schedule *s, int uplo, const double *vals,
                                       const double *b, double *x, int nrhs) {
    if (!h) return BAAS_INVALID_ARG;
    void *ptr = noalloc_scratch(h->device);  // set up by baas_create on this device
    if (!ptr) return BAAS_HIP_ERROR;
    return sweep<double>(h, s, uplo, vals, b, x, nrhs, h->stream, nullptr, ptr, NOALLOC_SCRATCH_BYTES);
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
    if (enable < 0 || enable > 5) return BAAS_INVALID_ARG;
    s->fwd.dataflow = s->bwd.dataflow = enable;
    s->plan_gen++;
    return BAAS_OK;
}
