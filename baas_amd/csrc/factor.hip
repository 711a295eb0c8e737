// factor.hip -- IC0 and ILU0 numeric factorisation on gfx950.
//
// Both run on the forward execution plan of the lower pattern of A (a row
// depends exactly on the rows of its strictly lower entries) with the same
// unit / sub-level / ready-flag protocol as the triangular solve.  One lane
// factors one row with the reference's arithmetic:
//   IC0  (merge/IC.cpp:32-89 and :296-393): up-looking, L_ij = (A_ij - <L_i,
//        L_j>)/L_jj over sorted two-pointer dots with FMA; the diagonal is
//        sqrt(A_ii - sum L_ik^2) only if positive, else A_ii is kept.  The
//        self dot follows the reference build's vectorised form (blocks of
//        four rounded squares added in order, then FMA for the tail).
//   ILU0 (merge/ILU.cpp:350-386 and :638-720): IKJ, l_ij = a_ij / u_jj then
//        a_ik -= l_ij u_jk over the matching upper entries of row j (FMA).
// Values another row reads are stored with agent-scope write-through stores
// and read with agent-scope loads, so rows of the same unit (after the
// sub-level barrier) and of earlier units (after the ready flag) see them.
#include <algorithm>

#include "device.h"
#include "internal.h"

namespace baas {
int build_ic0_plans(baas_schedule *s);
int launch_pipe_factor(baas_handle *h, baas_schedule *s, Plan &p, int kind, const double *a_vals, double *lu);
int build_factor_lane(baas_schedule *s);
size_t lds_limit(baas_handle *h);

struct FactorArgs {
    int n, nunits;
    const int *rowptr, *colidx, *diagptr;
    const int *ic_djpos, *ic_pptr;  // IC0 dot plans (may be null)
    const int2 *ic_pairs;
    const int *key_begin, *sub_begin, *sub_ptr, *rows, *key;
    int *flags;
    Ctx ctx;
    DevStatus *status;
    long long spin_limit;
};

__device__ __forceinline__ double sparse_dot_dev(int l1, int u1, int l2, int u2, const int *__restrict__ idx,
                                                 const double *lu) {
    double r = 0.0;
    while (l1 < u1 && l2 < u2) {
        const int a = idx[l1], b = idx[l2];
        if (a == b) {
            r = fma(ld_agent(&lu[l1]), ld_agent(&lu[l2]), r);
            l1++;
            l2++;
        } else if (a < b) {
            l1++;
        } else {
            l2++;
        }
    }
    return r;
}

__device__ __forceinline__ double self_dot_dev(int l, int u, const double *lu) {
    const int k = u - l;
    double r = 0.0;
    int i = l;
    if (k > 3) {
        const int stop = l + (k & ~3);
        for (; i < stop; i++) {
            const double v = ld_agent(&lu[i]);
            const double p = v * v;
            r = r + p;
        }
    }
    for (; i < u; i++) {
        const double v = ld_agent(&lu[i]);
        r = fma(v, v, r);
    }
    return r;
}

template <int KIND>
__device__ __forceinline__ void factor_row(const FactorArgs &a, int i, double *lu) {
    const int rb = a.rowptr[i], dg = a.diagptr[i];
    if (KIND == BAAS_KIND_IC0) {
        for (int k = rb; k < dg; k++) {
            const int j = a.colidx[k];
            const double dp = sparse_dot_dev(rb, dg, a.rowptr[j], a.diagptr[j], a.colidx, lu);
            const double aij = ld_agent(&lu[k]);
            const double v = (aij - dp) / ld_agent(&lu[a.diagptr[j]]);
            st_agent(&lu[k], v);
            wait_stores();
        }
        const double dv = self_dot_dev(rb, dg, lu);
        const double aii = ld_agent(&lu[dg]);
        if (aii - dv > 0) {
            st_agent(&lu[dg], sqrt(aii - dv));
        } else {
            raise_status(a.status, BAAS_NONPOSITIVE_PIVOT, i);
        }
    } else {
        const int re = a.rowptr[i + 1];
        for (int j = rb; j < dg; j++) {
            const int col = a.colidx[j];
            const double pivot = ld_agent(&lu[a.diagptr[col]]);
            if (!(fabs(pivot) > 1.0e-16)) raise_status(a.status, BAAS_ZERO_PIVOT, col);
            const double tmp = ld_agent(&lu[j]) / pivot;
            st_agent(&lu[j], tmp);
            int k1 = j + 1, k2 = a.diagptr[col] + 1;
            const int e2 = a.rowptr[col + 1];
            while (k1 < re && k2 < e2) {
                const int c1 = a.colidx[k1], c2 = a.colidx[k2];
                if (c1 < c2) {
                    k1++;
                } else if (c1 > c2) {
                    k2++;
                } else {
                    const double v = fma(-tmp, ld_agent(&lu[k2]), ld_agent(&lu[k1]));
                    st_agent(&lu[k1], v);
                    wait_stores();
                    k1++;
                    k2++;
                }
            }
            wait_stores();
        }
    }
}

// IC0 row from its dot plan, for rows of at most IC_MAXE lower entries and
// IC_MAXT pairs: every load the row needs (its A values, the diagonals of
// its columns, the other rows' entries of every dot) is issued at once, then
// the entries are computed in order with the row's own values in LDS (own,
// stride apart) -- the same products in the same order as the two-pointer
// dots.  Returns false for rows outside these bounds.
constexpr int IC_MAXE = 8, IC_MAXT = 24;

__device__ __forceinline__ bool ic0_row_planned(const FactorArgs &a, int i, double *lu, double *own, int stride) {
    const int rb = a.rowptr[i], dg = a.diagptr[i], ne = dg - rb;
    const int p0 = a.ic_pptr[i], np = a.ic_pptr[i + 1] - p0;
    if (ne > IC_MAXE || np > IC_MAXT) return false;
    double av[IC_MAXE], dj[IC_MAXE], pv[IC_MAXT];
    int djp[IC_MAXE];
    int2 pr[IC_MAXT];
#pragma unroll
    for (int e = 0; e < IC_MAXE; e++) {
        const int k = rb + (e < ne ? e : (ne > 0 ? ne - 1 : 0));
        av[e] = ld_agent(&lu[k]);
        djp[e] = a.ic_djpos[k];
    }
    // a row without dot pairs reads no pair record at all (its slots point at
    // its own diagonal); otherwise slots past the count repeat the last pair
#pragma unroll
    for (int t = 0; t < IC_MAXT; t++) pr[t] = np > 0 ? a.ic_pairs[p0 + (t < np ? t : np - 1)] : make_int2(dg, 0);
    const double aii = ld_agent(&lu[dg]);
#pragma unroll
    for (int e = 0; e < IC_MAXE; e++) dj[e] = ld_agent(&lu[djp[e]]);
#pragma unroll
    for (int t = 0; t < IC_MAXT; t++) pv[t] = ld_agent(&lu[pr[t].x]);
#pragma unroll
    for (int e = 0; e < IC_MAXE; e++) {
        if (e < ne) {
            double dp = 0.0;
#pragma unroll
            for (int t = 0; t < IC_MAXT; t++)
                if (t < np && (pr[t].y >> 8) == e) dp = fma(own[(pr[t].y & 255) * stride], pv[t], dp);
            const double v = (av[e] - dp) / dj[e];
            own[e * stride] = v;
            st_agent(&lu[rb + e], v);
        }
    }
    // self dot in the reference build's order (blocks of four rounded
    // squares added in order when there are more than three, FMA tail)
    double r = 0.0;
    const int stop = ne > 3 ? (ne & ~3) : 0;
#pragma unroll
    for (int e = 0; e < IC_MAXE; e++) {
        if (e < ne) {
            const double v = own[e * stride];
            if (e < stop) {
                const double p = v * v;
                r = r + p;
            } else {
                r = fma(v, v, r);
            }
        }
    }
    if (aii - r > 0) {
        st_agent(&lu[dg], sqrt(aii - r));
    } else {
        raise_status(a.status, BAAS_NONPOSITIVE_PIVOT, i);
    }
    return true;
}

// One row of the factorisation: the planned IC0 row when it applies, the
// general two-pointer form otherwise.
template <int KIND>
__device__ __forceinline__ void factor_any(const FactorArgs &a, int i, double *lu, double *own, int stride) {
    if (KIND == BAAS_KIND_IC0 && a.ic_pptr && ic0_row_planned(a, i, lu, own, stride)) return;
    factor_row<KIND>(a, i, lu);
}

template <int KIND, int KEYMODE>
__global__ __launch_bounds__(1024) void k_factor(FactorArgs a, double *lu) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double *own = reinterpret_cast<double *>(smem_raw) + threadIdx.x;
    __shared__ int s_unit;
    const int epoch = read_epoch(a.ctx);
    for (;;) {
        const int u = claim_unit(a.ctx.ticket, &s_unit);
        if (u >= a.nunits) break;
        const int kb = a.key_begin[u];
        for (int si = a.sub_begin[u]; si < a.sub_begin[u + 1]; si++) {
            const int qb = a.sub_ptr[si], qe = a.sub_ptr[si + 1];
            for (int q = qb + threadIdx.x; q < qe; q += blockDim.x) {
                const int rw = a.rows[q];
                const int i = rw & 0x7fffffff;
                for (int k = a.rowptr[i]; k < a.diagptr[i]; k++) {
                    const int c = a.colidx[k];
                    if (key_of<KEYMODE>(c, a.n, a.key) < kb) wait_ready(a.flags, c, epoch, a.spin_limit, a.status);
                }
                factor_any<KIND>(a, i, lu, own, blockDim.x);
                wait_stores();
                if (rw < 0) st_agent(&a.flags[i], epoch);
            }
            __syncthreads();
        }
    }
    retire_block(a.ctx, epoch);
}

// Level-set factorisation (LEVELS plans): all workgroups share each row
// level, a device-wide barrier separates levels (the reference's level-set
// kernels, merge/IC.cpp:209-281 and merge/ILU.cpp:578-625).
template <int KIND>
__global__ __launch_bounds__(1024) void k_factor_levels(FactorArgs a, int nlev, const int *__restrict__ lptr,
                                                       const int4 *__restrict__ meta, int *bar, double *lu) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double *own = reinterpret_cast<double *>(smem_raw) + threadIdx.x;
    const int epoch = read_epoch(a.ctx);
    const int nth = (int)(gridDim.x * blockDim.x);
    const int gt = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    for (int l = 0; l < nlev; l++) {
        for (int q = lptr[l] + gt; q < lptr[l + 1]; q += nth) factor_any<KIND>(a, meta[q].x, lu, own, blockDim.x);
        if (l + 1 < nlev) {
            if (gridDim.x == 1) {
                wait_stores();  // a single workgroup needs only its own barrier
                __syncthreads();
            } else {
                grid_barrier(bar, (l + 1) * (int)gridDim.x, a.status, a.spin_limit);
            }
        }
    }
    if (threadIdx.x == 0) {
        __threadfence();
        const int d = atomicAdd(a.ctx.done, 1);
        if (d == (int)gridDim.x - 1) {
            reset_barrier(bar);
            atomicExch(a.ctx.ticket, 0);
            atomicExch(a.ctx.epoch, epoch == 0x7fffffff ? 1 : epoch + 1);
            __threadfence();
            atomicExch(a.ctx.done, 0);
        }
    }
}

// ---------------------------------------------------------------------------
// IC0 and ILU0 on the lane kernel (dataflow plans).  Every entry that
// another unit reads is published into a double-buffered array indexed by
// its lu position, the value being its own flag (as the sweep publishes x).
// A unit stages its rows' records, entries (A values, divisor slots, pair
// lists) and pairs in LDS; wavefront 0 stages the external entries the unit
// reads as they are published; every other lane factors its own rows, entry
// by entry, over the precomputed pairs in the reference's order:
//  IC0 (merge/IC.cpp:6-89): L_ik = (A_ik - sum_c L_ic L_jc) / L_jj with the
//  products ascending in c, then the diagonal with the self dot in the
//  reference build's order;
//  ILU0 (merge/ILU.cpp:350-386): lu_ik1 = A_ik1 updated by
//  fma(-L_ij, U_{col(j),c1}, .) for the row's lower entries j in ascending
//  order, then divided by U_{c1,c1} when k1 is a lower entry.
// No level barrier and no flag wait: an entry starts as soon as its own
// inputs are in LDS.
struct IclArgs {
    int nunits;
    size_t nnz;
    const int *unit_ptr;  // the forward plan's units (positions)
    const int4 *rec, *ent, *unit;
    const int *pair, *ext;
    const int *diagptr;
    int *pcount;
    Ctx ctx;
    DevStatus *status;
    long long spin_limit;
};

constexpr int ICL_CH = 4;  // pairs read per poll

// ALLF: every row with lower entries is marked (a stencil).  IC0 then has no
// entry-by-entry path; ILU0 uses it for diagonals only (no pivot, at most FW
// updates, one per lower entry), so its pair chunk is FW wide.
template <int KIND, int FW, bool ALLF>
__device__ __forceinline__ void icl_factor_rows(const IclArgs &a, double *sl, const int4 *lrec, const double *laii,
                                                const int *ldgp, const int4 *lent, const double *leav,
                                                const int *lpair, double *lu, double *pub, double *nxt, int p,
                                                int stride, int nk) {
    constexpr bool DIAG_ONLY = KIND == BAAS_KIND_ILU0 && ALLF;
    constexpr int PCH = DIAG_ONLY ? FW : ICL_CH;
    bool have = p < nk;
    int row = 0, eo = 0, ne = 0, base = 0, dgp = 0, e = 0, k = 0, djs = 0, ps = 0, np = 0, pc = 0;
    bool dpub = false, epub = false, force = false, fast = false;
    double dp = 0.0, aii = 0.0;
    long long spins = 0;
    // rows marked by the host (stencil rows): their nfl lower entries (at
    // most FW) take no updates and are factored together; an ILU0 row's
    // upper entries take none either, so they are A's values from the start
    int nfl = 0, fk[FW], fdj[FW];
    bool fpub[FW];
    double fav[FW];
    // DIAG_ONLY: the diagonal entry of a marked row (its lu position, A's
    // value, publication mark) and its pairs (the LDS slot of the parent's
    // upper value, the index of the row's own lower entry), at most FW
    int dk = 0, dnp = 0, uslot[FW], uo[FW];
    bool dpubd = false;
    double dav = 0.0;
    auto load_entry = [&]() {
        dp = 0.0;
        if (e < ne) {
            const int4 q = lent[eo + e];
            k = q.x;
            djs = q.y;
            ps = q.z;
            np = q.w & 0x7fffffff;
            epub = q.w < 0;
            if (KIND == BAAS_KIND_ILU0) dp = leav[eo + e];  // ILU0 updates A's value in place
        }
        pc = 0;
        spins = 0;
        force = false;
    };
    auto start = [&]() {
        const int4 r = lrec[p];
        row = r.x & 0x3fffffff;
        fast = (r.x & (1 << 30)) != 0;
        eo = r.y;
        ne = r.z & 0x7fffffff;
        dpub = r.z < 0;
        base = r.w;
        aii = laii[p];
        dgp = ldgp[p];
        e = 0;
        load_entry();
        if (fast) {
            // IC0: every entry is lower; ILU0: the lower entries precede the
            // diagonal, whose lu position is dgp (the first entry's is A's row start)
            nfl = KIND == BAAS_KIND_IC0 ? ne : dgp - lent[eo].x;
            const int last = nfl > 0 ? nfl - 1 : 0;  // a row has at least one entry
#pragma unroll
            for (int t = 0; t < FW; t++) {
                const int4 q = lent[eo + (t < nfl ? t : last)];
                fk[t] = q.x;
                fdj[t] = q.y;
                fpub[t] = q.w < 0;
                fav[t] = leav[eo + (t < nfl ? t : last)];
            }
            if (DIAG_ONLY) {
                const int4 q = lent[eo + nfl];
                dk = q.x;
                dnp = q.w & 0x7fffffff;
                dpubd = q.w < 0;
                dav = leav[eo + nfl];
#pragma unroll
                for (int t = 0; t < FW; t++) {
                    const int w = t < dnp ? lpair[q.z + t] : 0;
                    uslot[t] = t < dnp ? w >> 8 : base;  // any slot of the unit when unused
                    uo[t] = w & 255;
                }
            }
            if (KIND == BAAS_KIND_ILU0) {
                for (int t = nfl + 1; t < ne; t++) {
                    const int4 q = lent[eo + t];
                    const double L = leav[eo + t];
                    st_lds(&sl[base + t], L);
                    lu[q.x] = L;
                    if (q.w < 0) {
                        st_agent(&pub[q.x], L);
                        nxt[q.x] = Sentinel<double>::value();
                    }
                }
            }
        }
    };
    // One entry's step: its pairs, then its pivot (an ILU0 row whose last
    // entry this completes moves on to its next row in the same step).
    auto pair_step = [&]() {
        if (pc < np) {
            double v[PCH], o[PCH];
#pragma unroll
            for (int t = 0; t < PCH; t++) {
                const int i = ps + (pc + t < np ? pc + t : np - 1);
                const int w = lpair[i];
                v[t] = ld_lds(&sl[w >> 8]);
                o[t] = ld_lds(&sl[base + (w & 255)]);
            }
            int took = 0;
#pragma unroll
            for (int t = 0; t < PCH; t++) {
                if (t == took && pc + t < np && (force || !Sentinel<double>::is(v[t]))) {
                    const double pv = Sentinel<double>::is(v[t]) ? 0.0 : v[t];
                    dp = KIND == BAAS_KIND_IC0 ? fma(o[t], pv, dp) : fma(-o[t], pv, dp);
                    took++;
                }
            }
            pc += took;
        }
        bool done_entry = false;
        if (pc == np) {
            const double dj = !DIAG_ONLY && djs >= 0 ? ld_lds(&sl[djs]) : 1.0;
            if (force || !Sentinel<double>::is(dj)) {
                const double d = Sentinel<double>::is(dj) ? 1.0 : dj;
                double L;
                if (KIND == BAAS_KIND_IC0) {
                    L = (leav[eo + e] - dp) / d;
                } else {
                    if (!DIAG_ONLY && djs >= 0 && !(fabs(d) > 1.0e-16))
                        raise_status(a.status, BAAS_ZERO_PIVOT, row);
                    L = !DIAG_ONLY && djs >= 0 ? dp / d : dp;
                }
                st_lds(&sl[base + e], L);
                lu[k] = L;
                if (epub) {
                    st_agent(&pub[k], L);
                    nxt[k] = Sentinel<double>::value();
                }
                // a marked ILU0 row comes here for its diagonal only
                e = KIND == BAAS_KIND_ILU0 && fast ? ne : e + 1;
                load_entry();
                done_entry = true;
                if (KIND == BAAS_KIND_ILU0 && e >= ne) {
                    p += stride;  // every entry of the row, diagonal included, is done
                    have = p < nk;
                    if (have) start();
                }
            }
        }
        if (!done_entry && !force && ++spins > a.spin_limit) {
            raise_status(a.status, BAAS_TIMEOUT, row);
            force = true;  // finish this entry with what has arrived
        }
    };
    if (have) start();
    // one loop with a wavefront-uniform exit and a single divergent step;
    // every path of a step ends at the one back edge (a second one lets the
    // compiler nest loops, and lanes parked outside an inner spin loop can
    // starve the lanes spinning in it)
    while (__ballot(have) != 0) {
        if (!have) continue;
        if (fast && e < nfl) {
            // every lower entry needs only its pivot: read them all, and when
            // all are there factor the entries in order
            double d[FW], uv[FW];
#pragma unroll
            for (int t = 0; t < FW; t++) d[t] = ld_lds(&sl[fdj[t]]);
            bool ready = true, uready = true;
#pragma unroll
            for (int t = 0; t < FW; t++)
                if (t < nfl && Sentinel<double>::is(d[t])) ready = false;
            if (DIAG_ONLY) {
                // the parents' upper values for the diagonal, published when
                // those rows started: normally there before their pivots
#pragma unroll
                for (int t = 0; t < FW; t++) uv[t] = ld_lds(&sl[uslot[t]]);
#pragma unroll
                for (int t = 0; t < FW; t++)
                    if (t < dnp && Sentinel<double>::is(uv[t])) uready = false;
            }
            if (ready || force) {
                double Lv[FW];
#pragma unroll
                for (int t = 0; t < FW; t++) {
                    Lv[t] = 0.0;
                    if (t < nfl) {
                        const double dd = Sentinel<double>::is(d[t]) ? 1.0 : d[t];
                        if (KIND == BAAS_KIND_ILU0 && !(fabs(dd) > 1.0e-16))
                            raise_status(a.status, BAAS_ZERO_PIVOT, row);
                        // IC0 subtracts a zero dot, ILU0 divides A's value
                        const double L = (KIND == BAAS_KIND_IC0 ? fav[t] - 0.0 : fav[t]) / dd;
                        Lv[t] = L;
                        st_lds(&sl[base + t], L);
                        lu[fk[t]] = L;
                        if (fpub[t]) {
                            st_agent(&pub[fk[t]], L);
                            nxt[fk[t]] = Sentinel<double>::value();
                        }
                    }
                }
                if (KIND == BAAS_KIND_IC0) {
                    // the diagonal in the same step, from the entries just
                    // computed: at most FW <= 3 of them, so the self dot is
                    // its FMA tail alone, in order (as in the step below),
                    // then the pivot rule of merge/IC.cpp:82-87
                    double r = 0.0;
#pragma unroll
                    for (int t = 0; t < FW; t++)
                        if (t < nfl) r = fma(Lv[t], Lv[t], r);
                    double dg = aii;
                    if (aii - r > 0) {
                        dg = sqrt(aii - r);
                    } else {
                        raise_status(a.status, BAAS_NONPOSITIVE_PIVOT, row);
                    }
                    st_lds(&sl[base + ne], dg);
                    lu[dgp] = dg;
                    if (dpub) {
                        st_agent(&pub[dgp], dg);
                        nxt[dgp] = Sentinel<double>::value();
                    }
                    p += stride;
                    have = p < nk;
                    if (have) start();
                } else if (DIAG_ONLY && (uready || force)) {
                    // the diagonal in the same step: A's value less the
                    // pairs' products in order (as in pair_step), no division
                    double dd = dav;
#pragma unroll
                    for (int t = 0; t < FW; t++) {
                        if (t < dnp) {
                            const double o = uo[t] == 0 ? Lv[0] : (FW > 1 && uo[t] == 1) ? Lv[1 < FW ? 1 : 0]
                                                                                        : Lv[FW - 1];
                            const double pv = Sentinel<double>::is(uv[t]) ? 0.0 : uv[t];
                            dd = fma(-o, pv, dd);
                        }
                    }
                    st_lds(&sl[base + nfl], dd);
                    lu[dk] = dd;
                    if (dpubd) {
                        st_agent(&pub[dk], dd);
                        nxt[dk] = Sentinel<double>::value();
                    }
                    p += stride;
                    have = p < nk;
                    if (have) start();
                } else {
                    e = nfl;  // the diagonal follows in the next step
                    load_entry();
                }
            } else if (!force && ++spins > a.spin_limit) {
                raise_status(a.status, BAAS_TIMEOUT, row);
                force = true;
            }
        } else if ((!ALLF || KIND == BAAS_KIND_ILU0) && e < ne) {
            pair_step();
        } else if (KIND == BAAS_KIND_ILU0) {
            p += stride;  // every entry of the row, diagonal included, is done
            have = p < nk;
            if (have) start();
        } else {
            // diagonal: the self dot in the reference build's order (blocks
            // of four rounded squares added in order when there are more
            // than three, FMA tail), then the pivot rule of merge/IC.cpp:82-87
            double r = 0.0;
            const int stop = ne > 3 ? (ne & ~3) : 0;
            for (int t0 = 0; t0 < ne; t0 += 4) {
                double v[4];
#pragma unroll
                for (int t = 0; t < 4; t++) v[t] = ld_lds(&sl[base + (t0 + t < ne ? t0 + t : ne - 1)]);
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    if (t0 + t < ne) {
                        if (t0 + t < stop) {
                            const double sq = v[t] * v[t];
                            r = r + sq;
                        } else {
                            r = fma(v[t], v[t], r);
                        }
                    }
                }
            }
            double d = aii;
            if (aii - r > 0) {
                d = sqrt(aii - r);
            } else {
                raise_status(a.status, BAAS_NONPOSITIVE_PIVOT, row);
            }
            st_lds(&sl[base + ne], d);
            lu[dgp] = d;
            if (dpub) {
                st_agent(&pub[dgp], d);
                nxt[dgp] = Sentinel<double>::value();
            }
            p += stride;
            have = p < nk;
            if (have) start();
        }
    }
}

template <int KIND, int FW, bool ALLF>
__global__ __launch_bounds__(1024) void k_factor_lane(IclArgs a, double *lu, double *xp) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ int s_unit;
    const int epoch = read_epoch(a.ctx);
    const int par = ld_agent(a.pcount) & 1;
    double *pub = xp + (size_t)par * a.nnz;
    double *nxt = xp + (size_t)(par ^ 1) * a.nnz;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (;;) {
        const int u = claim_unit(a.ctx.ticket, &s_unit);
        if (u >= a.nunits) break;
        const int q0 = a.unit_ptr[u], nk = a.unit_ptr[u + 1] - q0;
        const int4 U0 = a.unit[u], U1 = a.unit[u + 1];
        const int nown = U0.x, nent = U1.y - U0.y, npair = U1.z - U0.z, next = U1.w - U0.w;
        const IclLayout lay = icl_layout(nown + next, nk, nent, npair);
        double *sl = reinterpret_cast<double *>(smem_raw + lay.vals);
        int4 *lrec = reinterpret_cast<int4 *>(smem_raw + lay.rec);
        double *laii = reinterpret_cast<double *>(smem_raw + lay.aii);
        int *ldgp = reinterpret_cast<int *>(smem_raw + lay.dgp);
        int4 *lent = reinterpret_cast<int4 *>(smem_raw + lay.ent);
        double *leav = reinterpret_cast<double *>(smem_raw + lay.eav);
        int *lpair = reinterpret_cast<int *>(smem_raw + lay.pair);
        for (int i = threadIdx.x; i < nown + next; i += blockDim.x) st_lds(&sl[i], Sentinel<double>::value());
        for (int q = threadIdx.x; q < nk; q += blockDim.x) {
            const int4 r = a.rec[q0 + q];
            const int dg = a.diagptr[r.x & 0x3fffffff];  // bit 30 marks rows without dot pairs
            lrec[q] = r;
            ldgp[q] = dg;
            laii[q] = lu[dg];
        }
        for (int i = threadIdx.x; i < nent; i += blockDim.x) {
            const int4 q = a.ent[U0.y + i];
            lent[i] = q;
            leav[i] = lu[q.x];
        }
        for (int i = threadIdx.x; i < npair; i += blockDim.x) lpair[i] = a.pair[U0.z + i];
        __syncthreads();
        if (wv == 0) {
            stage_published<double>(pub, sl, a.ext + U0.w, nown, next, 1, lane, a.spin_limit, a.status);
        } else {
            icl_factor_rows<KIND, FW, ALLF>(a, sl, lrec, laii, ldgp, lent, leav, lpair, lu, pub, nxt, (wv - 1) * 64 + lane,
                                  (nw - 1) * 64, nk);
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

__global__ void k_fill_sentinel_f64(size_t n, double *p) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = Sentinel<double>::value();
}

__global__ void k_copy_f64(int n, const double *__restrict__ in, double *__restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i];
}

// L values (lower CSR) and U = L^T values from an IC0 result in A's layout.
template <class T>
__global__ void k_split(int nnz, const int *__restrict__ l_src, const int *__restrict__ u_src,
                        const double *__restrict__ lu, T *__restrict__ lv, T *__restrict__ uv) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    lv[e] = (T)lu[l_src[e]];
    uv[e] = (T)lu[l_src[u_src[e]]];
}

static inline int nblk(long long n, int b = 256) { return (int)((n + b - 1) / b); }

template <int KIND, int KM>
static int launch_factor(baas_handle *h, const Plan &p, const FactorArgs &a, double *lu) {
    auto kern = k_factor<KIND, KM>;
    int per_cu = 0;
    const size_t lds = KIND == BAAS_KIND_IC0 ? (size_t)IC_MAXE * p.block * sizeof(double) : 0;
    BAAS_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, p.block, lds));
    if (per_cu <= 0) return BAAS_UNSUPPORTED;
    long long grid = (long long)per_cu * h->num_cus;
    if (grid > p.nunits) grid = p.nunits;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(p.block), lds, h->stream, a, lu);
    BAAS_HIP_TRY(hipGetLastError());
    return BAAS_OK;
}

template <int KIND>
static int factor(baas_handle *h, baas_schedule *s, const double *a_vals, double *lu) {
    if (!h || !s || !a_vals || !lu) return BAAS_INVALID_ARG;
    if (s->kind != BAAS_KIND_IC0 && s->kind != BAAS_KIND_ILU0) return BAAS_INVALID_ARG;
    // PIPE plans over a full stencil: the pipelined factorisation, which
    // writes every entry of lu itself (A's values where the factor leaves
    // them).  In place it reads each row's entries when its box starts, and
    // only that box writes them.
    {
        const int rc = launch_pipe_factor(h, s, s->fwd, KIND, a_vals, lu);
        if (rc != BAAS_UNSUPPORTED) return rc;
    }
    if (a_vals != lu)
        hipLaunchKernelGGL(k_copy_f64, dim3(nblk(s->nnz_full)), dim3(256), 0, h->stream, s->nnz_full, a_vals, lu);
    const Plan &p = s->fwd;
    FactorArgs a;
    a.n = s->n;
    a.nunits = p.nunits;
    a.rowptr = s->a_rowptr.p;
    a.colidx = s->a_colidx.p;
    a.diagptr = s->a_diagptr.p;
    a.ic_djpos = a.ic_pptr = nullptr;
    a.ic_pairs = nullptr;
    if (KIND == BAAS_KIND_IC0) {
        BAAS_TRY(build_ic0_plans(s));
        a.ic_djpos = s->ic_djpos.p;
        a.ic_pptr = s->ic_pptr.p;
        a.ic_pairs = reinterpret_cast<const int2 *>(s->ic_pairs.p);
    }
    a.key_begin = p.key_begin.p;
    a.sub_begin = p.sub_begin.p;
    a.sub_ptr = p.sub_ptr.p;
    a.rows = p.rows.p;
    a.key = p.key.p;
    a.flags = s->flags.p;
    a.ctx = Ctx{h->counters, h->counters + 1, h->counters + 2};
    a.status = h->status;
    a.spin_limit = h->spin_limit;
    if ((p.dataflow == 2 || p.dataflow == 5) && p.block >= 128) {
        const int rc = build_factor_lane(s);
        if (rc != BAAS_OK && rc != BAAS_UNSUPPORTED) return rc;
        if (rc == BAAS_OK && s->icl_lds <= lds_limit(h)) {
            const size_t nnz = s->nnz_full;
            if (!s->icl_armed) {
                hipLaunchKernelGGL(k_fill_sentinel_f64, dim3((unsigned)((2 * nnz + 255) / 256)), dim3(256), 0,
                                   h->stream, 2 * nnz, reinterpret_cast<double *>(s->icl_xp.p));
                BAAS_HIP_TRY(hipGetLastError());
                s->icl_armed = true;
            }
            IclArgs c;
            c.nunits = p.nunits;
            c.nnz = nnz;
            c.unit_ptr = p.unit_ptr.p;
            c.rec = reinterpret_cast<const int4 *>(s->icl_rec.p);
            c.ent = reinterpret_cast<const int4 *>(s->icl_ent.p);
            c.unit = reinterpret_cast<const int4 *>(s->icl_unit.p);
            c.pair = s->icl_pair.p;
            c.ext = s->icl_ext.p;
            c.diagptr = s->a_diagptr.p;
            c.pcount = s->icl_pcount.p;
            c.ctx = Ctx{h->counters, h->counters + 1, h->counters + 2};
            c.status = h->status;
            c.spin_limit = h->spin_limit;
            // marked rows have at most icl_fw lower entries (2 or 3)
            auto kern = s->icl_fw == 2 ? k_factor_lane<KIND, 2, false> : k_factor_lane<KIND, 3, false>;
            if (s->icl_allfast) kern = s->icl_fw == 2 ? k_factor_lane<KIND, 2, true> : k_factor_lane<KIND, 3, true>;
            int per_cu = 0;
            BAAS_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, p.block, s->icl_lds));
            // at most 16 wavefronts per CU, as for the lane sweep
            per_cu = std::min(per_cu, std::max(1, 1024 / p.block));
            if (per_cu > 0) {
                long long grid = std::min<long long>((long long)per_cu * h->num_cus, p.nunits);
                if (grid < 1) grid = 1;
                hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(p.block), s->icl_lds, h->stream, c,
                                   lu, reinterpret_cast<double *>(s->icl_xp.p));
                BAAS_HIP_TRY(hipGetLastError());
                return BAAS_OK;
            }
        }
    }
    if (p.mode == BAAS_PLAN_LEVELS) {
        auto kern = k_factor_levels<KIND>;
        const size_t lds = KIND == BAAS_KIND_IC0 ? (size_t)IC_MAXE * p.block * sizeof(double) : 0;
        int per_cu = 0;
        BAAS_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, p.block, lds));
        if (per_cu <= 0) return BAAS_UNSUPPORTED;
        int widest = 1;
        for (int l = 0; l + 1 < (int)s->row_level_ptr.size(); l++)
            widest = std::max(widest, s->row_level_ptr[l + 1] - s->row_level_ptr[l]);
        long long grid =
            std::min<long long>((long long)std::min(per_cu, 2) * h->num_cus, (widest + p.block - 1) / p.block);
        if (p.lv_cap > 0 && grid > p.lv_cap) grid = p.lv_cap;
        if (grid < 1) grid = 1;
        hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(p.block), lds, h->stream, a, p.nlev, p.lptr.p,
                           reinterpret_cast<const int4 *>(p.meta.p), h->counters + BAAS_BARRIER_BASE, lu);
        BAAS_HIP_TRY(hipGetLastError());
        return BAAS_OK;
    }
    switch (p.keymode) {
        case 0: return launch_factor<KIND, 0>(h, p, a, lu);
        case 1: return launch_factor<KIND, 1>(h, p, a, lu);
        default: return launch_factor<KIND, 2>(h, p, a, lu);
    }
}

int ic0_split_dev(baas_handle *h, baas_schedule *s, const double *lu, double *lv, double *uv, hipStream_t st) {
    int nnz = s->nnz_lower;
    hipLaunchKernelGGL(k_split<double>, dim3(nblk(nnz)), dim3(256), 0, st, nnz, s->l_src.p, s->u_src.p, lu, lv, uv);
    BAAS_HIP_TRY(hipGetLastError());
    return BAAS_OK;
}

int ic0_split_dev32(baas_handle *h, baas_schedule *s, const double *lu, float *lv, float *uv, hipStream_t st) {
    int nnz = s->nnz_lower;
    hipLaunchKernelGGL(k_split<float>, dim3(nblk(nnz)), dim3(256), 0, st, nnz, s->l_src.p, s->u_src.p, lu, lv, uv);
    BAAS_HIP_TRY(hipGetLastError());
    return BAAS_OK;
}

}  // namespace baas

using namespace baas;

extern "C" int baas_ic0_f64(baas_handle *h, baas_schedule *s, const double *a_vals, double *lu) {
    return factor<BAAS_KIND_IC0>(h, s, a_vals, lu);
}

extern "C" int baas_ilu0_f64(baas_handle *h, baas_schedule *s, const double *a_vals, double *lu) {
    return factor<BAAS_KIND_ILU0>(h, s, a_vals, lu);
}

extern "C" int baas_ic0_split_f64(baas_handle *h, baas_schedule *s, const double *lu, double *lv, double *uv) {
    if (!h || !s || !lu || !lv || !uv || s->kind == BAAS_KIND_TRSV) return BAAS_INVALID_ARG;
    return ic0_split_dev(h, s, lu, lv, uv, h->stream);
}

extern "C" int baas_ic0_split_f32(baas_handle *h, baas_schedule *s, const double *lu, float *lv, float *uv) {
    if (!h || !s || !lu || !lv || !uv || s->kind == BAAS_KIND_TRSV) return BAAS_INVALID_ARG;
    return ic0_split_dev32(h, s, lu, lv, uv, h->stream);
}
