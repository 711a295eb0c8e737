// plan.cpp -- execution plans: packing a schedule into GPU work units.
//
// The reference maps groups of each (merged) level to OpenMP threads with its
// affinity rule and builds a task DAG of point-to-point waits
// (merge/LevelTask.cpp:2295-2684, :305-485, :3096-3184).  On the GPU a work
// unit is a contiguous run of an execution order processed by one workgroup
// sub-level by sub-level (barrier in between, solved values staged in LDS);
// rows that another unit needs publish a per-row ready flag.  Two packings:
//
//   RANGE     execution order = row order (reversed for the backward sweep),
//             units = consecutive row ranges.  Dependencies of banded and
//             naturally ordered matrices stay inside a unit, so most of the
//             critical path runs on LDS barriers instead of HBM flags.
//   SCHEDULE  execution order = the BAAS schedule: merged levels, groups in
//             level order, nodes ascending inside a group; units = chunks of
//             at most rows_per_unit rows of one merged level.
//
// Built on the host once per analysis (like the reference's task build).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.h"

namespace baas {

namespace {

struct HostCsr {
    std::vector<int> ptr, idx;
};

int fetch(HostCsr &c, const int *ptr, const int *idx, int n, hipStream_t st) {
    c.ptr.resize(n + 1);
    if (hipMemcpyAsync(c.ptr.data(), ptr, sizeof(int) * (n + 1), hipMemcpyDeviceToHost, st) != hipSuccess)
        return BAAS_HIP_ERROR;
    if (hipStreamSynchronize(st) != hipSuccess) return BAAS_HIP_ERROR;
    c.idx.resize(c.ptr[n]);
    if (c.ptr[n] &&
        hipMemcpyAsync(c.idx.data(), idx, sizeof(int) * c.ptr[n], hipMemcpyDeviceToHost, st) != hipSuccess)
        return BAAS_HIP_ERROR;
    if (hipStreamSynchronize(st) != hipSuccess) return BAAS_HIP_ERROR;
    return BAAS_OK;
}

template <class T>
int upload(DBuf<T> &d, const std::vector<T> &h, hipStream_t st) {
    BAAS_TRY(d.alloc(h.empty() ? 1 : h.size()));
    if (!h.empty() && hipMemcpyAsync(d.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, st) != hipSuccess)
        return BAAS_HIP_ERROR;
    return BAAS_OK;
}

// Build one plan.
//   order[n]       execution order (topological for the sweep)
//   unit_ptr[]     unit boundaries in `order`
//   deps           parents of each row for this sweep (excluding the diagonal)
//   kids           children of each row for this sweep
//   keymode        0 key = row, 1 key = n-1-row, 2 key = position in order
int make_plan(Plan &p, int n, const std::vector<int> &order, const std::vector<int> &unit_ptr,
              const std::vector<int> &dptr, const std::vector<int> &didx, const std::vector<int> &kptr,
              const std::vector<int> &kidx, int keymode, int mode, int block, hipStream_t st,
              const HostCsr &full, bool upper) {
    int nunits = (int)unit_ptr.size() - 1;
    std::vector<int> pos(n), unit_of(n), key(n);
    for (int q = 0; q < n; q++) pos[order[q]] = q;
    for (int u = 0; u < nunits; u++)
        for (int q = unit_ptr[u]; q < unit_ptr[u + 1]; q++) unit_of[order[q]] = u;
    for (int r = 0; r < n; r++) key[r] = keymode == 0 ? r : keymode == 1 ? n - 1 - r : pos[r];
    std::vector<int> key_begin(nunits + 1), sub_begin(nunits + 1), sub_ptr, rows(n);
    std::vector<int> sub(n, 0);
    int max_keys = 0;
    sub_ptr.push_back(0);
    for (int u = 0; u < nunits; u++) {
        int b = unit_ptr[u], e = unit_ptr[u + 1];
        // key range of the unit (contiguous by construction)
        int kmin = 1 << 30, kmax = -1;
        for (int q = b; q < e; q++) {
            kmin = std::min(kmin, key[order[q]]);
            kmax = std::max(kmax, key[order[q]]);
        }
        if (kmax - kmin + 1 != e - b) return BAAS_INVALID_ARG;
        key_begin[u] = kmin;
        max_keys = std::max(max_keys, e - b);
        int maxsub = 0;
        for (int q = b; q < e; q++) {
            int r = order[q], s = 0;
            for (int j = dptr[r]; j < dptr[r + 1]; j++) {
                int c = didx[j];
                if (unit_of[c] == u) s = std::max(s, sub[c] + 1);
                else if (unit_of[c] > u) return BAAS_INVALID_ARG;  // not topological
            }
            sub[r] = s;
            maxsub = std::max(maxsub, s);
        }
        // counting sort of the unit's rows by sub-level (stable)
        std::vector<int> cnt(maxsub + 2, 0);
        for (int q = b; q < e; q++) cnt[sub[order[q]] + 1]++;
        for (int s = 0; s <= maxsub; s++) cnt[s + 1] += cnt[s];
        sub_begin[u] = (int)sub_ptr.size() - 1;
        for (int s = 0; s <= maxsub; s++) sub_ptr.push_back(b + cnt[s + 1]);
        for (int q = b; q < e; q++) {
            int r = order[q];
            int ext = 0;
            for (int j = kptr[r]; j < kptr[r + 1]; j++)
                if (unit_of[kidx[j]] != u) { ext = 1; break; }
            rows[b + cnt[sub[r]]++] = r | (ext ? (int)0x80000000u : 0);
        }
    }
    key_begin[nunits] = n;
    sub_begin[nunits] = (int)sub_ptr.size() - 1;
    // dataflow slots: intra-unit parents live at key - key_begin, external
    // parents after the unit's keys in order of first use
    // (the diagonal's entry holds slot 0: a row without parents may load it,
    // never use it)
    std::vector<int> scol(full.idx.size(), 0), ext_ptr(nunits + 1, 0), ext_cols;
    std::vector<int> ext_slot(n, -1), ext_touched;
    int max_slots = 0;
    for (int u = 0; u < nunits; u++) {
        const int kb = key_begin[u], nk = unit_ptr[u + 1] - unit_ptr[u];
        ext_ptr[u] = (int)ext_cols.size();
        for (int q = unit_ptr[u]; q < unit_ptr[u + 1]; q++) {
            const int r = rows[q] & 0x7fffffff;
            const int rb = full.ptr[r], re = full.ptr[r + 1];
            const int dg = upper ? rb : re - 1;
            for (int j = rb; j < re; j++) {
                if (j == dg) continue;
                const int c = full.idx[j];
                if (unit_of[c] == u) {
                    scol[j] = key[c] - kb;
                } else {
                    if (ext_slot[c] < 0) {
                        ext_slot[c] = nk + (int)(ext_cols.size() - ext_ptr[u]);
                        ext_cols.push_back(c);
                        ext_touched.push_back(c);
                    }
                    scol[j] = ext_slot[c];
                }
            }
        }
        for (int c : ext_touched) ext_slot[c] = -1;
        ext_touched.clear();
        max_slots = std::max(max_slots, nk + (int)ext_cols.size() - ext_ptr[u]);
    }
    ext_ptr[nunits] = (int)ext_cols.size();
    // rows each unit publishes for later units, in execution order, as LDS
    // slots; and the dataflow kernel's LDS footprint per unit (one value per
    // row of the unit and per external parent)
    std::vector<int> out_ptr(nunits + 1, 0), out_slot;
    size_t df_lds8 = 0, df_lds4 = 0;
    for (int u = 0; u < nunits; u++) {
        out_ptr[u] = (int)out_slot.size();
        for (int q = unit_ptr[u]; q < unit_ptr[u + 1]; q++)
            if (rows[q] < 0) out_slot.push_back(key[rows[q] & 0x7fffffff] - key_begin[u]);
        const size_t nk = unit_ptr[u + 1] - unit_ptr[u], ne = ext_ptr[u + 1] - ext_ptr[u];
        df_lds8 = std::max(df_lds8, (nk + ne) * 8);
        df_lds4 = std::max(df_lds4, (nk + ne) * 4);
    }
    out_ptr[nunits] = (int)out_slot.size();
    // per-position row records for the dataflow kernel: one 16-byte load
    // gives a lane everything it needs to fetch its row
    std::vector<int> meta(4 * (size_t)n);
    p.max_deg = 0;
    for (int u = 0; u < nunits; u++)
        for (int q = unit_ptr[u]; q < unit_ptr[u + 1]; q++) {
            const int r = rows[q] & 0x7fffffff;
            const int rb = full.ptr[r], re = full.ptr[r + 1];
            meta[4 * (size_t)q + 0] = rows[q];
            meta[4 * (size_t)q + 1] = upper ? re - 1 : rb;
            meta[4 * (size_t)q + 2] = re - rb - 1;
            p.max_deg = std::max(p.max_deg, re - rb - 1);
            meta[4 * (size_t)q + 3] = key[r] - key_begin[u];
        }
    // lane kernel: unit-relative entry offsets and the LDS footprint
    std::vector<int> lane_eoff(n), lane_eptr(nunits + 1, 0);
    p.h_lane_units.assign(3 * (size_t)nunits, 0);
    for (int u = 0; u < nunits; u++) {
        int acc = 0;
        for (int q = unit_ptr[u]; q < unit_ptr[u + 1]; q++) {
            lane_eoff[q] = acc;
            acc += meta[4 * (size_t)q + 2];
        }
        lane_eptr[u + 1] = lane_eptr[u] + acc;
        p.h_lane_units[3 * (size_t)u + 0] = unit_ptr[u + 1] - unit_ptr[u];
        p.h_lane_units[3 * (size_t)u + 1] = ext_ptr[u + 1] - ext_ptr[u];
        p.h_lane_units[3 * (size_t)u + 2] = acc;
    }
    BAAS_TRY(upload(p.lane_eoff, lane_eoff, st));
    BAAS_TRY(upload(p.lane_eptr, lane_eptr, st));
    BAAS_TRY(p.pcount.alloc(1));
    if (hipMemsetAsync(p.pcount.p, 0, sizeof(int), st) != hipSuccess) return BAAS_HIP_ERROR;
    p.xp_raw.free();
    p.xp_type = 0;
    p.xp_nrhs = 0;
    p.lane_lds_key = -1;
    p.mr_lds_key = -1;
    p.pipe = 0;  // a PIPE plan's boxes are set by its builder after this
    p.pipe_half = 0;
    p.pipe_fact = 0;
    BAAS_TRY(upload(p.scol, scol, st));
    BAAS_TRY(upload(p.ext_ptr, ext_ptr, st));
    BAAS_TRY(upload(p.ext_cols, ext_cols, st));
    BAAS_TRY(upload(p.meta, meta, st));
    BAAS_TRY(upload(p.out_ptr, out_ptr, st));
    BAAS_TRY(upload(p.out_slot, out_slot, st));
    p.df_lds[0] = df_lds4;
    p.df_lds[1] = df_lds8;
    BAAS_TRY(upload(p.unit_ptr, unit_ptr, st));
    p.max_slots = max_slots;
    p.n = n;
    p.nunits = nunits;
    p.nsub = (int)sub_ptr.size() - 1;
    p.keymode = keymode;
    p.max_unit_keys = max_keys;
    p.block = block;
    p.mode = mode;
    BAAS_TRY(upload(p.key_begin, key_begin, st));
    BAAS_TRY(upload(p.sub_begin, sub_begin, st));
    BAAS_TRY(upload(p.sub_ptr, sub_ptr, st));
    BAAS_TRY(upload(p.rows, rows, st));
    if (keymode == 2) {
        BAAS_TRY(upload(p.key, key, st));
        BAAS_TRY(upload(p.key_row, order, st));  // key = position in order
    }
    p.h_unit_ptr = unit_ptr;
    p.h_rows.resize(n);
    for (int q = 0; q < n; q++) p.h_rows[q] = rows[q] & 0x7fffffff;
    if (hipStreamSynchronize(st) != hipSuccess) return BAAS_HIP_ERROR;
    return BAAS_OK;
}

}  // namespace

size_t Plan::mr_lds(int ts, int nrhs, bool stage_b) {
    const int key = (nrhs * 16 + ts) * 2 + (stage_b ? 1 : 0);
    if (key == mr_lds_key) return mr_lds_val;
    size_t m = 0;
    for (size_t u = 0; u + 2 < h_lane_units.size(); u += 3)
        m = std::max(m, mr_layout(h_lane_units[u], h_lane_units[u + 2], ts, nrhs, stage_b).total);
    mr_lds_key = key;
    mr_lds_val = m;
    return m;
}

size_t Plan::lane_lds(int ts, int nrhs) {
    const int key = nrhs * 16 + ts;
    if (key == lane_lds_key) return lane_lds_val;
    size_t m = 0;
    for (size_t u = 0; u + 2 < h_lane_units.size(); u += 3)
        m = std::max(m, lane_layout(h_lane_units[u], h_lane_units[u + 1], h_lane_units[u + 2], ts, nrhs).total);
    lane_lds_key = key;
    lane_lds_val = m;
    return m;
}

// Strip the diagonal: parents of row r in L (lower CSR, diag last) and in
// U = L^T (upper CSR, diag first).
// LEVELS plan: rows in (level, id) order with a 16-byte record each (row,
// first off-diagonal entry, off-diagonal count), and the level pointer.  The
// backward plan walks the same levels in reverse over U.
static int make_levels_plan(Plan &p, const baas_schedule *s, const HostCsr &M, bool upper, int block, int cap,
                            hipStream_t st) {
    const int n = s->n;
    const auto &perm = s->row_perm, &lp = s->row_level_ptr;
    if ((int)perm.size() != n || lp.empty() || lp.back() != n) return BAAS_INVALID_ARG;
    std::vector<int> meta(4 * (size_t)n);
    int max_deg = 0;
    for (int q = 0; q < n; q++) {
        const int r = perm[q], rb = M.ptr[r], re = M.ptr[r + 1];
        meta[4 * (size_t)q + 0] = r;
        meta[4 * (size_t)q + 1] = upper ? re - 1 : rb;
        meta[4 * (size_t)q + 2] = re - rb - 1;
        meta[4 * (size_t)q + 3] = 0;
        max_deg = std::max(max_deg, re - rb - 1);
    }
    BAAS_TRY(upload(p.meta, meta, st));
    BAAS_TRY(upload(p.lptr, lp, st));
    p.nlev = (int)lp.size() - 1;
    p.n = n;
    p.nunits = p.nlev;
    p.max_deg = max_deg;
    p.mode = BAAS_PLAN_LEVELS;
    p.block = block;
    p.lv_cap = cap;
    p.keymode = 0;
    p.max_slots = 0;
    p.df_lds[0] = p.df_lds[1] = 0;
    p.max_unit_keys = 0;
    if (hipStreamSynchronize(st) != hipSuccess) return BAAS_HIP_ERROR;
    return BAAS_OK;
}

// SCHEDULE units: merged levels -> groups in level order -> nodes ascending,
// each merged level cut into units of at most R rows.
static int schedule_units(const baas_schedule *s, int R, std::vector<int> &order, std::vector<int> &fu) {
    const int n = s->n;
    const auto &gp = s->group_ptr, &gs = s->group_set, &pto = s->perm_to_orig, &mp = s->merged_ptr;
    int q = 0;
    for (int m = 0; m + 1 < (int)mp.size(); m++) {
        int start = q;
        for (int t = mp[m]; t < mp[m + 1]; t++) {
            int g = pto[t];
            for (int k = gp[g]; k < gp[g + 1]; k++) order[q++] = gs[k];
        }
        for (int b = start; b < q; b += R) fu.push_back(b);
    }
    fu.push_back(n);
    return q == n ? BAAS_OK : BAAS_INVALID_ARG;
}

// TILES units: the (row level, row id) plane cut into tiles of hb
// consecutive levels by S consecutive row ids, taken band by band and, in a
// band, by ascending row range; rows inside a tile ascending.  A parent has a
// lower level and a lower row id than its child, so it lies in the same
// tile or in an earlier one: the order is topological.  On matrices whose
// row numbering follows the geometry (natural orderings of stencils, banded
// and RCM-ordered matrices) a tile is a compact block of the domain, so a
// dependency chain crosses few tiles.  Tiles above BAAS_TILE_MAX_ROWS are cut
// into pieces of consecutive rows (each piece's parents lie in earlier
// pieces or tiles).
static int tile_units(const baas_schedule *s, int S, int hb, std::vector<int> &order, std::vector<int> &fu) {
    const int n = s->n;
    const auto &lev = s->row_level;
    if ((int)lev.size() != n || S < 1 || hb < 1) return BAAS_INVALID_ARG;
    int maxlev = 1;
    for (int r = 0; r < n; r++) maxlev = std::max(maxlev, lev[r]);
    const long long nr = (n + (long long)S - 1) / S;
    const long long ntiles = ((long long)(maxlev - 1) / hb + 1) * nr;
    if (ntiles > (1LL << 28)) return BAAS_INVALID_ARG;
    std::vector<int> start((size_t)ntiles + 1, 0);
    auto tile_of = [&](int r) { return (long long)((lev[r] - 1) / hb) * nr + r / S; };
    for (int r = 0; r < n; r++) start[tile_of(r) + 1]++;
    for (long long t = 0; t < ntiles; t++) start[t + 1] += start[t];
    std::vector<int> next(start.begin(), start.end() - 1);
    for (int r = 0; r < n; r++) order[next[tile_of(r)]++] = r;
    for (long long t = 0; t < ntiles; t++) {
        const int b = start[t], sz = start[t + 1] - b;
        if (sz == 0) continue;
        const int pieces = (sz + BAAS_TILE_MAX_ROWS - 1) / BAAS_TILE_MAX_ROWS, per = (sz + pieces - 1) / pieces;
        for (int q = b; q < b + sz; q += per) fu.push_back(q);
    }
    fu.push_back(n);
    return BAAS_OK;
}

// BOXES units: rows of a structured grid numbered in natural (lexicographic)
// order, recognised from the lower triangle itself.  When every
// off-diagonal entry's offset (row - column) is one of at most three values
// 1 = s1 < s2 < s3 with s2 dividing s3, row r sits at (x, y, z) =
// (r mod s2, (r div s2) mod (s3 / s2), r div s3) -- the 5- and 7-point
// stencil orderings -- and units are boxes of about R rows, as cubic as the
// grid allows, taken in (z, y, x) box order with rows ascending inside a
// box.  A parent then lies in the same box or an earlier one (the plan
// builder checks this and rejects the order otherwise), and a dependency
// chain crosses a unit once per box edge in every direction instead of at
// every plane.  Matrices without this structure get BAAS_UNSUPPORTED.
static int box_units(const baas_schedule *s, const HostCsr &L, int R, std::vector<int> &order, std::vector<int> &fu) {
    const int n = s->n;
    if (R < 1) return BAAS_INVALID_ARG;
    std::vector<long long> offs;
    for (int r = 0; r < n; r++)
        for (int j = L.ptr[r]; j < L.ptr[r + 1] - 1; j++) {
            const long long d = (long long)r - L.idx[j];
            if (std::find(offs.begin(), offs.end(), d) == offs.end()) {
                offs.push_back(d);
                if (offs.size() > 3) return BAAS_UNSUPPORTED;
            }
        }
    std::sort(offs.begin(), offs.end());
    if (offs.empty() || offs[0] != 1) return BAAS_UNSUPPORTED;
    const long long sy = offs.size() >= 2 ? offs[1] : 0, sz = offs.size() == 3 ? offs[2] : 0;
    if (sz && sz % sy != 0) return BAAS_UNSUPPORTED;
    const long long nx = sy ? sy : n;
    const long long ny = sz ? sz / sy : (sy ? (n + sy - 1) / sy : 1);
    const long long nz = sz ? (n + sz - 1) / sz : 1;
    const int dims = (nx > 1) + (ny > 1) + (nz > 1);
    const double edge = std::pow((double)R, 1.0 / std::max(dims, 1));
    auto clampl = [](long long v, long long hi) { return std::max(1LL, std::min(v, hi)); };
    const long long bx = clampl(std::llround(edge), nx);
    const long long by = ny > 1 ? clampl(std::llround(edge), ny) : 1;
    const long long bz = nz > 1 ? clampl(R / (bx * by), nz) : 1;
    const long long gx = (nx + bx - 1) / bx, gy = (ny + by - 1) / by, gz = (nz + bz - 1) / bz;
    const long long nbox = gx * gy * gz;
    if (nbox > (1LL << 26)) return BAAS_INVALID_ARG;
    // boxes numbered in wavefront order: by the sum of their box
    // coordinates (every parent box has a smaller sum), then z, y, x; the
    // persistent grid claims units in this order, so the boxes it holds at
    // any moment are those whose parents are done or running
    std::vector<long long> rank((size_t)nbox);
    {
        std::vector<long long> ids((size_t)nbox);
        for (long long t = 0; t < nbox; t++) ids[t] = t;
        auto key = [&](long long t) {
            const long long tx = t % gx, ty = (t / gx) % gy, tz = t / (gx * gy);
            return std::make_pair(tx + ty + tz, t);
        };
        std::sort(ids.begin(), ids.end(), [&](long long a, long long b) { return key(a) < key(b); });
        for (long long q = 0; q < nbox; q++) rank[ids[q]] = q;
    }
    auto box_of = [&](int r) {
        const long long x = r % nx, y = (r / nx) % ny, z = r / (nx * ny);
        return rank[((z / bz) * gy + y / by) * gx + x / bx];
    };
    std::vector<int> start((size_t)nbox + 1, 0);
    for (int r = 0; r < n; r++) start[box_of(r) + 1]++;
    for (long long t = 0; t < nbox; t++) start[t + 1] += start[t];
    std::vector<int> next(start.begin(), start.end() - 1);
    for (int r = 0; r < n; r++) order[next[box_of(r)]++] = r;
    for (long long t = 0; t < nbox; t++)
        if (start[t + 1] > start[t]) fu.push_back(start[t]);
    fu.push_back(n);
    return BAAS_OK;
}

// PIPE units: boxes of a stencil grid shaped for one wavefront each, swept
// by the pipelined grid kernel (pipe.hip).  The grid is recognised as for
// BOXES (offsets 1 < sy < sz, sy dividing sz, natural order), and must also
// fill whole planes (n a multiple of sz; of sy on a 2D grid) with a full
// stencil: a row has an entry at offset 1 exactly when x > 0, at sy
// exactly when y > 0, at sz exactly when z > 0.  Boxes hold at most 2048
// rows.  A 2D grid
// (offsets 1 < sy) is taken as nx = sy, ny = 1, nz = n / sy, its rows
// becoming the pipeline's planes.  Boxes are TX x TY rows across (8 x 8 on
// a 3D grid, 64 x 1 on a 2D grid, narrower where the grid is) by
// K = R / (TX TY) planes, numbered in wavefront order like BOXES units.
// The backward sweep takes the same boxes in reverse order, each in flipped
// coordinates (x -> nx - 1 - x, ...).  Per unit and sweep, box[8u..8u+7] =
// origin x, y, z and tx | ty << 8 | kz << 16 in that sweep's coordinates,
// then where faces live in a publication half: the box's own faces (see
// pipe_face_sizes; then PIPE_DUMP values its lanes with nothing to publish
// in a step write instead), and the x face of the box before it in x, the y face
// of the box before it in y and the top plane of the box below it (-1
// where there is none).  half = the values of all boxes' faces.
struct PipeGeom {
    int nx = 0, ny = 1, nz = 1, sy = -1, sz = 0, TX = 1, TY = 1, K = 1;
};
// A's rows are the grid's full stencil: row r holds r - sz, r - sy, r - 1,
// r, r + 1, r + sy, r + sz (those inside the grid, ascending) from
// position pipe_pa.
static bool pipe_stencil_rows(const PipeGeom &g, const HostCsr &A) {
    const int n = (int)A.ptr.size() - 1;
    const bool three = g.sy > 0;
    int want[7];
    for (int r = 0; r <= n; r++) {
        const int x = r % g.nx, y = three ? (r / g.nx) % g.ny : 0, z = r / g.sz;
        if (A.ptr[r] != pipe_pa(g.nx, g.ny, g.nz, g.sy, g.sz, x, y, z)) return false;
        if (r == n) break;
        int c = 0;
        if (z > 0) want[c++] = r - g.sz;
        if (three && y > 0) want[c++] = r - g.sy;
        if (x > 0) want[c++] = r - 1;
        want[c++] = r;
        if (x < g.nx - 1) want[c++] = r + 1;
        if (three && y < g.ny - 1) want[c++] = r + g.sy;
        if (z < g.nz - 1) want[c++] = r + g.sz;
        if (A.ptr[r + 1] - A.ptr[r] != c) return false;
        for (int e = 0; e < c; e++)
            if (A.idx[A.ptr[r] + e] != want[e]) return false;
    }
    return true;
}
static int pipe_units(const baas_schedule *s, const HostCsr &L, int R, std::vector<int> &order, std::vector<int> &fu,
                      PipeGeom &g, std::vector<int> &fbox, std::vector<int> &bbox, long long &half) {
    const int n = s->n;
    if (R < 64 || !s->block_ptr.empty()) return R < 64 ? BAAS_INVALID_ARG : BAAS_UNSUPPORTED;
    std::vector<long long> offs;
    for (int r = 0; r < n; r++) {
        // the diagonal last, at most three parents in ascending columns
        const int cnt = L.ptr[r + 1] - L.ptr[r];
        if (cnt < 1 || cnt > 4 || L.idx[L.ptr[r + 1] - 1] != r) return BAAS_UNSUPPORTED;
        for (int j = L.ptr[r] + 1; j < L.ptr[r + 1]; j++)
            if (L.idx[j] <= L.idx[j - 1]) return BAAS_UNSUPPORTED;
        for (int j = L.ptr[r]; j < L.ptr[r + 1] - 1; j++) {
            const long long d = (long long)r - L.idx[j];
            if (std::find(offs.begin(), offs.end(), d) == offs.end()) {
                offs.push_back(d);
                if (offs.size() > 3) return BAAS_UNSUPPORTED;
            }
        }
    }
    std::sort(offs.begin(), offs.end());
    if (offs.size() < 2 || offs[0] != 1) return BAAS_UNSUPPORTED;
    const bool three = offs.size() == 3;
    const long long s1 = offs[1], s2 = three ? offs[2] : offs[1];
    if (s2 % s1 != 0 || n % s2 != 0 || s2 > (1LL << 30)) return BAAS_UNSUPPORTED;
    g.nx = (int)s1;
    g.ny = three ? (int)(s2 / s1) : 1;
    g.nz = (int)(n / s2);
    g.sy = three ? (int)s1 : -1;
    g.sz = (int)s2;
    // a full stencil: every row has exactly the parents its position
    // implies, each in the direction its offset names
    for (int r = 0; r < n; r++) {
        const int x = r % g.nx, y = (r / g.nx) % g.ny, z = r / g.sz;
        const int expect = (x > 0) + (three && y > 0) + (z > 0);
        if (L.ptr[r + 1] - L.ptr[r] - 1 != expect) return BAAS_UNSUPPORTED;
        for (int j = L.ptr[r]; j < L.ptr[r + 1] - 1; j++) {
            const int d = r - L.idx[j];
            if ((d == 1 && x == 0) || (three && d == g.sy && y == 0) || (d == g.sz && z == 0))
                return BAAS_UNSUPPORTED;
        }
    }
    // the entries of every row start where the stencil puts them (the kernel
    // computes these positions instead of reading the row pointers)
    for (int r = 0; r <= n; r++) {
        const int x = r % g.nx, y = (r / g.nx) % g.ny, z = r / g.sz;
        if (L.ptr[r] != pipe_lp(g.nx, g.ny, g.sy, g.sz, x, y, z)) return BAAS_UNSUPPORTED;
    }
    g.TX = three ? std::min(8, g.nx) : std::min(64, g.nx);
    g.TY = three ? std::min(64 / g.TX, g.ny) : 1;
    g.K = std::max(1, std::min({R / (g.TX * g.TY), g.nz, 255, 2048 / (g.TX * g.TY)}));
    // a box's staged faces have room for PIPE_FACE_MAX values each
    while (g.K > 1 && !pipe_faces_fit(g.TX, g.TY, g.K, three)) g.K--;
    if (!pipe_faces_fit(g.TX, g.TY, g.K, three)) return BAAS_UNSUPPORTED;
    const long long gx = (g.nx + g.TX - 1) / g.TX, gy = (g.ny + g.TY - 1) / g.TY, gz = (g.nz + g.K - 1) / g.K;
    const long long nbox = gx * gy * gz;
    if (nbox > (1LL << 26)) return BAAS_INVALID_ARG;
    std::vector<long long> ids((size_t)nbox);
    for (long long t = 0; t < nbox; t++) ids[t] = t;
    auto dsum = [&](long long t) { return t % gx + (t / gx) % gy + t / (gx * gy); };
    std::stable_sort(ids.begin(), ids.end(), [&](long long a, long long b) { return dsum(a) < dsum(b); });
    std::vector<long long> rank((size_t)nbox);
    for (long long q = 0; q < nbox; q++) rank[ids[q]] = q;
    std::vector<int> start((size_t)nbox + 1, 0);
    auto box_of = [&](int r) {
        const int x = r % g.nx, y = (r / g.nx) % g.ny, z = r / g.sz;
        return rank[((z / g.K) * gy + y / g.TY) * gx + x / g.TX];
    };
    for (int r = 0; r < n; r++) start[box_of(r) + 1]++;
    for (long long t = 0; t < nbox; t++) start[t + 1] += start[t];
    std::vector<int> next(start.begin(), start.end() - 1);
    for (int r = 0; r < n; r++) order[next[box_of(r)]++] = r;
    // forward unit of the box at grid position (bx, by, bz); -1 outside
    auto unit_at = [&](long long bx, long long by, long long bz) -> long long {
        if (bx < 0 || by < 0 || bz < 0 || bx >= gx || by >= gy || bz >= gz) return -1;
        return rank[(bz * gy + by) * gx + bx];
    };
    std::vector<int> dims((size_t)nbox);
    std::vector<long long> fsz((size_t)nbox);  // face values of forward unit q
    for (long long q = 0; q < nbox; q++) {
        const long long t = ids[q];
        const int x0 = (int)(t % gx) * g.TX, y0 = (int)((t / gx) % gy) * g.TY, z0 = (int)(t / (gx * gy)) * g.K;
        const int tx = std::min(g.TX, g.nx - x0), ty = std::min(g.TY, g.ny - y0), kz = std::min(g.K, g.nz - z0);
        dims[q] = tx | (ty << 8) | (kz << 16);
        const int3 f = pipe_face_sizes(tx, ty, kz, three);
        fsz[q] = (long long)f.x + f.y + f.z + PIPE_DUMP;
    }
    // the faces of forward unit q start at fb[q] in the forward sweep and at
    // bb[q] in the backward one (which takes the units in reverse order)
    std::vector<long long> fb((size_t)nbox), bb((size_t)nbox);
    half = 0;
    for (long long q = 0; q < nbox; q++) fb[q] = half, half += fsz[q];
    long long acc = 0;
    for (long long q = nbox - 1; q >= 0; q--) bb[q] = acc, acc += fsz[q];
    if (half >= (1LL << 31)) return BAAS_UNSUPPORTED;
    // offset of part 0 (x face), 1 (y face) or 2 (top plane) of unit q's faces
    auto part = [&](const std::vector<long long> &base, long long q, int which) -> int {
        if (q < 0) return -1;
        const int3 f = pipe_face_sizes(dims[q] & 255, (dims[q] >> 8) & 255, (dims[q] >> 16) & 255, three);
        return (int)(base[q] + (which > 0 ? f.x : 0) + (which > 1 ? f.y : 0));
    };
    fbox.assign(8 * (size_t)nbox, -1);
    bbox.assign(8 * (size_t)nbox, -1);
    for (long long q = 0; q < nbox; q++) {
        const long long t = ids[q];
        const int bx = (int)(t % gx), by = (int)((t / gx) % gy), bz = (int)(t / (gx * gy));
        const int x0 = bx * g.TX, y0 = by * g.TY, z0 = bz * g.K;
        const int tx = dims[q] & 255, ty = (dims[q] >> 8) & 255, kz = (dims[q] >> 16) & 255;
        int *f = &fbox[8 * (size_t)q];
        f[0] = x0, f[1] = y0, f[2] = z0, f[3] = dims[q];
        f[4] = (int)fb[q];
        f[5] = part(fb, unit_at(bx - 1, by, bz), 0);
        f[6] = three ? part(fb, unit_at(bx, by - 1, bz), 1) : -1;
        f[7] = part(fb, unit_at(bx, by, bz - 1), 2);
        // backward: unit nbox-1-q, flipped; its producers are the forward
        // neighbours after it
        int *r = &bbox[8 * (size_t)(nbox - 1 - q)];
        r[0] = g.nx - (x0 + tx), r[1] = g.ny - (y0 + ty), r[2] = g.nz - (z0 + kz), r[3] = dims[q];
        r[4] = (int)bb[q];
        r[5] = part(bb, unit_at(bx + 1, by, bz), 0);
        r[6] = three ? part(bb, unit_at(bx, by + 1, bz), 1) : -1;
        r[7] = part(bb, unit_at(bx, by, bz + 1), 2);
        fu.push_back(start[q]);
    }
    fu.push_back(n);
    return BAAS_OK;
}

// Grid geometry of one stencil block from its lower pattern: at most three
// distinct offsets 1 < sy < sz, sy dividing sz (rows r0 .. r1 of L); the box
// edge for about R rows per box.
struct BoxGeom {
    long long nx = 0, ny = 1, nz = 1, bx = 1, by = 1, bz = 1, gx = 1, gy = 1, gz = 1;
};
static int box_geometry(const HostCsr &L, int r0, int r1, int R, BoxGeom &g) {
    std::vector<long long> offs;
    for (int r = r0; r < r1; r++)
        for (int j = L.ptr[r]; j < L.ptr[r + 1] - 1; j++) {
            const long long d = (long long)r - L.idx[j];
            if (L.idx[j] < r0) return BAAS_UNSUPPORTED;  // a parent outside the block
            if (std::find(offs.begin(), offs.end(), d) == offs.end()) {
                offs.push_back(d);
                if (offs.size() > 3) return BAAS_UNSUPPORTED;
            }
        }
    std::sort(offs.begin(), offs.end());
    const long long n = r1 - r0;
    if (!offs.empty() && offs[0] != 1) return BAAS_UNSUPPORTED;
    const long long sy = offs.size() >= 2 ? offs[1] : 0, sz = offs.size() == 3 ? offs[2] : 0;
    if (sz && sz % sy != 0) return BAAS_UNSUPPORTED;
    g.nx = sy ? sy : n;
    g.ny = sz ? sz / sy : (sy ? (n + sy - 1) / sy : 1);
    g.nz = sz ? (n + sz - 1) / sz : 1;
    const int dims = (g.nx > 1) + (g.ny > 1) + (g.nz > 1);
    const double edge = std::pow((double)R, 1.0 / std::max(dims, 1));
    auto clampl = [](long long v, long long hi) { return std::max(1LL, std::min(v, hi)); };
    g.bx = clampl(std::llround(edge), g.nx);
    g.by = g.ny > 1 ? clampl(std::llround(edge), g.ny) : 1;
    g.bz = g.nz > 1 ? clampl(R / (g.bx * g.by), g.nz) : 1;
    g.gx = (g.nx + g.bx - 1) / g.bx;
    g.gy = (g.ny + g.by - 1) / g.by;
    g.gz = (g.nz + g.bz - 1) / g.bz;
    return BAAS_OK;
}

// BOXES units of a block-diagonal matrix: each block is a stencil grid of
// its own size, cut into boxes of about R rows; the boxes of all blocks are
// claimed in one wavefront order (box-coordinate sum, then block, then box),
// so every block's first boxes start at once and each box's parents (same
// block, smaller sum) come earlier.
static int block_box_units(const baas_schedule *s, const HostCsr &L, int R, std::vector<int> &order,
                           std::vector<int> &fu) {
    const std::vector<int> &bp = s->block_ptr;
    const int nb = (int)bp.size() - 1;
    if (R < 1 || nb < 1 || bp[0] != 0 || bp[nb] != s->n) return BAAS_INVALID_ARG;
    std::vector<BoxGeom> geo(nb);
    std::vector<long long> box0(nb + 1, 0);
    for (int b = 0; b < nb; b++) {
        BAAS_TRY(box_geometry(L, bp[b], bp[b + 1], R, geo[b]));
        box0[b + 1] = box0[b] + geo[b].gx * geo[b].gy * geo[b].gz;
    }
    const long long nbox = box0[nb];
    if (nbox > (1LL << 26)) return BAAS_INVALID_ARG;
    struct Key {
        long long d, b, t;
    };
    std::vector<Key> keys((size_t)nbox);
    for (int b = 0; b < nb; b++) {
        const BoxGeom &g = geo[b];
        for (long long t = 0; t < box0[b + 1] - box0[b]; t++) {
            const long long tx = t % g.gx, ty = (t / g.gx) % g.gy, tz = t / (g.gx * g.gy);
            keys[(size_t)(box0[b] + t)] = Key{tx + ty + tz, b, t};
        }
    }
    std::vector<long long> ids((size_t)nbox), rank((size_t)nbox);
    for (long long q = 0; q < nbox; q++) ids[q] = q;
    std::sort(ids.begin(), ids.end(), [&](long long a, long long c) {
        const Key &x = keys[a], &y = keys[c];
        return x.d != y.d ? x.d < y.d : (x.b != y.b ? x.b < y.b : x.t < y.t);
    });
    for (long long q = 0; q < nbox; q++) rank[ids[q]] = q;
    const int n = s->n;
    std::vector<long long> box_of(n);
    for (int b = 0; b < nb; b++) {
        const BoxGeom &g = geo[b];
        for (int r = bp[b]; r < bp[b + 1]; r++) {
            const long long q = r - bp[b];
            const long long x = q % g.nx, y = (q / g.nx) % g.ny, z = q / (g.nx * g.ny);
            box_of[r] = rank[box0[b] + ((z / g.bz) * g.gy + y / g.by) * g.gx + x / g.bx];
        }
    }
    std::vector<int> start((size_t)nbox + 1, 0);
    for (int r = 0; r < n; r++) start[box_of[r] + 1]++;
    for (long long t = 0; t < nbox; t++) start[t + 1] += start[t];
    std::vector<int> next(start.begin(), start.end() - 1);
    for (int r = 0; r < n; r++) order[next[box_of[r]]++] = r;
    for (long long t = 0; t < nbox; t++)
        if (start[t + 1] > start[t]) fu.push_back(start[t]);
    fu.push_back(n);
    return BAAS_OK;
}

int build_plans_mode(baas_schedule *s, int mode, int rows_per_unit) {
    s->plan_gen++;
    hipStream_t st = s->h->stream;
    const int n = s->n;
    HostCsr L, U;
    BAAS_TRY(fetch(L, s->l_rowptr.p, s->l_colidx.p, n, st));
    BAAS_TRY(fetch(U, s->u_rowptr.p, s->u_colidx.p, n, st));
    if (mode == BAAS_PLAN_LEVELS) {
        const int block = s->opts.block_threads > 0 ? s->opts.block_threads : 256;
        BAAS_TRY(make_levels_plan(s->fwd, s, L, false, block, rows_per_unit, st));
        BAAS_TRY(make_levels_plan(s->bwd, s, U, true, block, rows_per_unit, st));
        return BAAS_OK;
    }
    std::vector<int> lp(n + 1), li, up(n + 1), ui;
    li.reserve(L.idx.size());
    ui.reserve(U.idx.size());
    lp[0] = up[0] = 0;
    for (int r = 0; r < n; r++) {
        for (int j = L.ptr[r]; j < L.ptr[r + 1] - 1; j++) li.push_back(L.idx[j]);
        lp[r + 1] = (int)li.size();
        for (int j = U.ptr[r] + 1; j < U.ptr[r + 1]; j++) ui.push_back(U.idx[j]);
        up[r + 1] = (int)ui.size();
    }
    int block = s->opts.block_threads > 0 ? s->opts.block_threads : 256;
    std::vector<int> fwd_order(n), bwd_order(n), fu, bu;
    PipeGeom pg;
    std::vector<int> pbox, bbox;
    long long phalf = 0;
    if (mode == BAAS_PLAN_RANGE) {
        int R = rows_per_unit > 0 ? rows_per_unit : 4096;
        for (int r = 0; r < n; r++) fwd_order[r] = r, bwd_order[r] = n - 1 - r;
        for (int b = 0; b < n; b += R) fu.push_back(b);
        fu.push_back(n);
        bu = fu;
        // forward: parents = L off-diagonals, children = U off-diagonals
        BAAS_TRY(make_plan(s->fwd, n, fwd_order, fu, lp, li, up, ui, 0, mode, block, st, L, false));
        BAAS_TRY(make_plan(s->bwd, n, bwd_order, bu, up, ui, lp, li, 1, mode, block, st, U, true));
        return BAAS_OK;
    }
    if (mode == BAAS_PLAN_TILES) {
        BAAS_TRY(tile_units(s, rows_per_unit > 0 ? rows_per_unit : 16384, s->tile_band, fwd_order, fu));
    } else if (mode == BAAS_PLAN_PIPE) {
        BAAS_TRY(pipe_units(s, L, rows_per_unit > 0 ? rows_per_unit : 1024, fwd_order, fu, pg, pbox, bbox, phalf));
    } else if (mode == BAAS_PLAN_BOXES) {
        if (!s->block_ptr.empty())
            BAAS_TRY(block_box_units(s, L, rows_per_unit > 0 ? rows_per_unit : 512, fwd_order, fu));
        else
            BAAS_TRY(box_units(s, L, rows_per_unit > 0 ? rows_per_unit : 512, fwd_order, fu));
    } else {
        BAAS_TRY(schedule_units(s, rows_per_unit > 0 ? rows_per_unit : block, fwd_order, fu));
    }
    for (int i = 0; i < n; i++) bwd_order[i] = fwd_order[n - 1 - i];
    const int nu = (int)fu.size() - 1;
    bu.resize(nu + 1);
    for (int u = 0; u <= nu; u++) bu[u] = n - fu[nu - u];
    BAAS_TRY(make_plan(s->fwd, n, fwd_order, fu, lp, li, up, ui, 2, mode, block, st, L, false));
    BAAS_TRY(make_plan(s->bwd, n, bwd_order, bu, up, ui, lp, li, 2, mode, block, st, U, true));
    if (mode == BAAS_PLAN_PIPE) {
        const int g8[8] = {pg.nx, pg.ny, pg.nz, pg.sy, pg.sz, pg.TX, pg.TY, pg.K};
        BAAS_TRY(upload(s->fwd.pipe_box, pbox, st));
        BAAS_TRY(upload(s->bwd.pipe_box, bbox, st));
        if (hipStreamSynchronize(st) != hipSuccess) return BAAS_HIP_ERROR;
        // the factorisations take the pipelined path when A's rows hold
        // exactly the stencil's entries where pipe_pa puts them
        int fact = 0;
        if (s->kind != BAAS_KIND_TRSV) {
            HostCsr A;
            BAAS_TRY(fetch(A, s->a_rowptr.p, s->a_colidx.p, n, st));
            fact = pipe_stencil_rows(pg, A) ? 1 : 0;
        }
        for (Plan *p : {&s->fwd, &s->bwd}) {
            std::copy(g8, g8 + 8, p->pg);
            p->pipe_half = phalf;
            p->pipe_fact = fact;
            p->pipe = 1;
        }
    }
    return BAAS_OK;
}

// RANGE plans over caller-chosen unit boundaries (row space, ascending);
// the backward plan uses the same units in reverse order.
int make_plan_custom(baas_schedule *s, const std::vector<int> &fu, int block) {
    s->plan_gen++;
    hipStream_t st = s->h->stream;
    const int n = s->n;
    if (fu.empty() || fu.front() != 0 || fu.back() != n) return BAAS_INVALID_ARG;
    HostCsr L, U;
    BAAS_TRY(fetch(L, s->l_rowptr.p, s->l_colidx.p, n, st));
    BAAS_TRY(fetch(U, s->u_rowptr.p, s->u_colidx.p, n, st));
    std::vector<int> lp(n + 1), li, up(n + 1), ui;
    lp[0] = up[0] = 0;
    for (int r = 0; r < n; r++) {
        for (int j = L.ptr[r]; j < L.ptr[r + 1] - 1; j++) li.push_back(L.idx[j]);
        lp[r + 1] = (int)li.size();
        for (int j = U.ptr[r] + 1; j < U.ptr[r + 1]; j++) ui.push_back(U.idx[j]);
        up[r + 1] = (int)ui.size();
    }
    std::vector<int> fo(n), bo(n);
    for (int r = 0; r < n; r++) fo[r] = r, bo[r] = n - 1 - r;
    int nu = (int)fu.size() - 1;
    std::vector<int> bu(nu + 1);
    for (int u = 0; u <= nu; u++) bu[u] = n - fu[nu - u];
    if (block <= 0) block = 256;
    BAAS_TRY(make_plan(s->fwd, n, fo, fu, lp, li, up, ui, 0, BAAS_PLAN_RANGE, block, st, L, false));
    BAAS_TRY(make_plan(s->bwd, n, bo, bu, up, ui, lp, li, 1, BAAS_PLAN_RANGE, block, st, U, true));
    return BAAS_OK;
}

// Default packing before any autotuning: RANGE when a row-range unit keeps
// most dependencies inside the unit (the matrix is banded), SCHEDULE else.
int build_plans(baas_schedule *s) {
    int mode = s->opts.plan_mode;
    if (mode == BAAS_PLAN_AUTO) {
        // fraction of off-diagonal entries whose column lies within 4096 rows
        HostCsr L;
        BAAS_TRY(fetch(L, s->l_rowptr.p, s->l_colidx.p, s->n, s->h->stream));
        long long near = 0, tot = 0;
        for (int r = 0; r < s->n; r++)
            for (int j = L.ptr[r]; j < L.ptr[r + 1] - 1; j++) {
                tot++;
                if (r - L.idx[j] <= 4096) near++;
            }
        mode = (tot == 0 || near * 10 >= tot * 9) ? BAAS_PLAN_RANGE : BAAS_PLAN_SCHEDULE;
    }
    return build_plans_mode(s, mode, s->opts.rows_per_unit);
}

// IC0 dot plans: for row i and its lower entry e (column j), the matching
// columns of row i's lower part and row j's lower part in ascending order --
// exactly the pairs the up-looking sparse dot multiplies
// (merge/IC.cpp:6-29) -- as (position in row j, e << 8 | position in row i
// relative to its first entry).
int build_ic0_plans(baas_schedule *s) {
    if (s->ic_plans) return BAAS_OK;
    hipStream_t st = s->h->stream;
    const int n = s->n;
    HostCsr A;
    BAAS_TRY(fetch(A, s->a_rowptr.p, s->a_colidx.p, n, st));
    std::vector<int> dg(n);
    if (hipMemcpy(dg.data(), s->a_diagptr.p, sizeof(int) * n, hipMemcpyDeviceToHost) != hipSuccess)
        return BAAS_HIP_ERROR;
    std::vector<int> djpos(A.idx.size(), 0), pptr(n + 1, 0), pairs;
    for (int i = 0; i < n; i++) {
        pptr[i] = (int)(pairs.size() / 2);
        const int rb = A.ptr[i], di = dg[i];
        for (int k = rb; k < di; k++) {
            const int j = A.idx[k];
            djpos[k] = dg[j];
            int l1 = rb, l2 = A.ptr[j];
            const int u1 = di, u2 = dg[j];
            while (l1 < u1 && l2 < u2) {
                const int c1 = A.idx[l1], c2 = A.idx[l2];
                if (c1 == c2) {
                    pairs.push_back(l2);
                    pairs.push_back(((k - rb) << 8) | (l1 - rb));
                    l1++;
                    l2++;
                } else if (c1 < c2) {
                    l1++;
                } else {
                    l2++;
                }
            }
        }
    }
    pptr[n] = (int)(pairs.size() / 2);
    if (pairs.empty()) pairs = {0, 0};  // keep one valid record (never read: every row has np = 0)
    BAAS_TRY(upload(s->ic_djpos, djpos, st));
    BAAS_TRY(upload(s->ic_pptr, pptr, st));
    BAAS_TRY(upload(s->ic_pairs, pairs, st));
    if (hipStreamSynchronize(st) != hipSuccess) return BAAS_HIP_ERROR;
    s->ic_plans = true;
    return BAAS_OK;
}

// IC0 and ILU0 on the lane kernel (dataflow plans): for the forward plan's
// units, the value slots of every entry a unit computes or reads, the entry
// and pair records in the reference's order, and the entries other units
// read, which are published.
//  IC0 (merge/IC.cpp:6-89): a row's entries are its lower part; entry k
//  (column j) has the pairs (L_jc, own L_ic) of the up-looking sparse dot,
//  c ascending, and divides by L_jj; the diagonal follows (self dot).
//  ILU0 (merge/ILU.cpp:350-386): a row's entries are all of it; entry k1
//  (column c1) is updated by every lower entry j < k1 whose row col(j) has
//  c1 in its upper part, in ascending j (pairs (U_{col(j),c1}, own L_ij)),
//  and a lower entry then divides by U_{c1,c1}.
// The own-entry index is a byte: rows with more than 255 lower entries get
// BAAS_UNSUPPORTED, and the factorisation keeps the unit kernels.
int build_factor_lane(baas_schedule *s) {
    if (s->icl_gen == s->plan_gen) return s->icl_status;
    s->icl_gen = s->plan_gen;
    s->icl_status = BAAS_UNSUPPORTED;
    const bool ilu = s->kind == BAAS_KIND_ILU0;
    const Plan &p = s->fwd;
    if (p.mode == BAAS_PLAN_LEVELS || (int)p.h_rows.size() != s->n) return s->icl_status;
    hipStream_t st = s->h->stream;
    const int n = s->n;
    HostCsr A;
    BAAS_TRY(fetch(A, s->a_rowptr.p, s->a_colidx.p, n, st));
    std::vector<int> dg(n);
    if (hipMemcpy(dg.data(), s->a_diagptr.p, sizeof(int) * n, hipMemcpyDeviceToHost) != hipSuccess)
        return BAAS_HIP_ERROR;
    const size_t nnz = A.idx.size();
    const int nunits = p.nunits;
    // a row's entries: IC0 its lower part and diagonal, ILU0 all of it
    auto row_end = [&](int r) { return ilu ? A.ptr[r + 1] : dg[r] + 1; };
    std::vector<int> unit_of(n), row_of_lu(nnz, -1);
    for (int u = 0; u < nunits; u++)
        for (int q = p.h_unit_ptr[u]; q < p.h_unit_ptr[u + 1]; q++) unit_of[p.h_rows[q]] = u;
    if (n >= (1 << 30)) return s->icl_status;  // bit 30 of the row word is a flag
    for (int r = 0; r < n; r++) {
        if (dg[r] - A.ptr[r] > 255) return s->icl_status;
        for (int k = A.ptr[r]; k < row_end(r); k++) row_of_lu[k] = r;
    }
    std::vector<int> rec(4 * (size_t)n), ent, pair, unit(4 * ((size_t)nunits + 1), 0), ext;
    std::vector<char> pub(nnz, 0);
    std::vector<int> own_base(n, 0), ext_slot(nnz, -1), touched;
    std::vector<std::pair<int, int>> upd;  // ILU0: (k1, pair word), j ascending
    size_t lds = 0;
    for (int u = 0; u < nunits; u++) {
        const int q0 = p.h_unit_ptr[u], q1 = p.h_unit_ptr[u + 1];
        int nown = 0;
        for (int q = q0; q < q1; q++) {
            const int i = p.h_rows[q];
            own_base[i] = nown;
            nown += row_end(i) - A.ptr[i];
        }
        const int ent0 = (int)ent.size() / 4, pair0 = (int)pair.size(), ext0 = (int)ext.size();
        unit[4 * (size_t)u + 0] = nown;
        unit[4 * (size_t)u + 1] = ent0;
        unit[4 * (size_t)u + 2] = pair0;
        unit[4 * (size_t)u + 3] = ext0;
        // the value slot of lu position P for unit u
        auto slot_of = [&](int P) {
            const int r = row_of_lu[P];
            if (unit_of[r] == u) return own_base[r] + (P - A.ptr[r]);
            if (ext_slot[P] < 0) {
                ext_slot[P] = nown + (int)(ext.size() - ext0);
                ext.push_back(P);
                touched.push_back(P);
                pub[P] = 1;
            }
            return ext_slot[P];
        };
        for (int q = q0; q < q1; q++) {
            const int i = p.h_rows[q], rb = A.ptr[i], di = dg[i], ne = di - rb;
            rec[4 * (size_t)q + 0] = i;
            rec[4 * (size_t)q + 1] = (int)ent.size() / 4 - ent0;
            rec[4 * (size_t)q + 2] = ilu ? A.ptr[i + 1] - rb : ne;
            rec[4 * (size_t)q + 3] = own_base[i];
            if (!ilu) {
                for (int k = rb; k < di; k++) {
                    const int j = A.idx[k];
                    const int ps = (int)pair.size() - pair0;
                    int l1 = rb, l2 = A.ptr[j];
                    const int u1 = di, u2 = dg[j];
                    while (l1 < u1 && l2 < u2) {
                        const int c1 = A.idx[l1], c2 = A.idx[l2];
                        if (c1 == c2) {
                            pair.push_back((slot_of(l2) << 8) | (l1 - rb));
                            l1++;
                            l2++;
                        } else if (c1 < c2) {
                            l1++;
                        } else {
                            l2++;
                        }
                    }
                    ent.push_back(k);
                    ent.push_back(slot_of(dg[j]));
                    ent.push_back(ps);
                    ent.push_back((int)pair.size() - pair0 - ps);
                }
            } else {
                // the reference's update loop, recorded per target entry k1
                upd.clear();
                const int re = A.ptr[i + 1];
                for (int j = rb; j < di; j++) {
                    const int col = A.idx[j];
                    int k1 = j + 1, k2 = dg[col] + 1;
                    const int e2 = A.ptr[col + 1];
                    while (k1 < re && k2 < e2) {
                        const int c1 = A.idx[k1], c2 = A.idx[k2];
                        if (c1 < c2) {
                            k1++;
                        } else if (c1 > c2) {
                            k2++;
                        } else {
                            upd.push_back({k1, (slot_of(k2) << 8) | (j - rb)});
                            k1++;
                            k2++;
                        }
                    }
                }
                std::stable_sort(upd.begin(), upd.end(),
                                 [](const std::pair<int, int> &x, const std::pair<int, int> &y) {
                                     return x.first < y.first;
                                 });
                size_t t = 0;
                for (int k1 = rb; k1 < re; k1++) {
                    const int ps = (int)pair.size() - pair0;
                    for (; t < upd.size() && upd[t].first == k1; t++) pair.push_back(upd[t].second);
                    ent.push_back(k1);
                    ent.push_back(k1 < di ? slot_of(dg[A.idx[k1]]) : -1);  // lower entries divide by U_cc
                    ent.push_back(ps);
                    ent.push_back((int)pair.size() - pair0 - ps);
                }
            }
        }
        const size_t nslots = nown + ext.size() - ext0;
        if (nslots >= (1u << 23)) return s->icl_status;
        lds = std::max(lds, icl_layout(nslots, q1 - q0, ent.size() / 4 - ent0, pair.size() - pair0).total);
        for (int P : touched) ext_slot[P] = -1;
        touched.clear();
    }
    unit[4 * (size_t)nunits + 0] = 0;
    unit[4 * (size_t)nunits + 1] = (int)ent.size() / 4;
    unit[4 * (size_t)nunits + 2] = (int)pair.size();
    unit[4 * (size_t)nunits + 3] = (int)ext.size();
    // publish flags: bit 31 of the count words (IC0: the diagonal's on the row record)
    for (size_t e = 0; e < ent.size(); e += 4)
        if (pub[ent[e]]) ent[e + 3] |= (int)0x80000000;
    if (!ilu)
        for (int q = 0; q < n; q++)
            if (pub[dg[rec[4 * (size_t)q]]]) rec[4 * (size_t)q + 2] |= (int)0x80000000;
    // rows of at most three lower entries whose entries other than the
    // diagonal take no pairs (stencil rows) are marked in bit 30 of the row
    // word: their lower entries need only their pivots and are factored
    // together (IC0 also needs at least one entry: its row is all lower).
    // The kernel reads icl_fw pivots per step: 2 when no marked row has
    // three lower entries (5-point stencils), else 3.
    std::vector<char> mark(n, 0);
    int fw = 2;
    bool allfast = true;  // IC0: unmarked rows have no lower entries; ILU0: none unmarked
    for (int u = 0; u < nunits; u++)
        for (int q = p.h_unit_ptr[u]; q < p.h_unit_ptr[u + 1]; q++) {
            const int i = rec[4 * (size_t)q + 0], ne = rec[4 * (size_t)q + 2] & 0x7fffffff;
            const int nl = dg[i] - A.ptr[i];
            const int e0 = unit[4 * (size_t)u + 1] + rec[4 * (size_t)q + 1];
            bool nopairs = nl <= 3 && (ilu || ne >= 1);
            for (int e = 0; e < ne && nopairs; e++)
                if (!ilu || e != nl) nopairs = (ent[4 * ((size_t)e0 + e) + 3] & 0x7fffffff) == 0;
            if (nopairs) {
                mark[q] = 1;
                fw = std::max(fw, nl);
            } else if (ne > 0) {
                allfast = false;
            }
        }
    for (int q = 0; q < n; q++)
        if (mark[q]) rec[4 * (size_t)q + 0] |= 1 << 30;
    s->icl_fw = fw;
    s->icl_allfast = allfast;
    if (ent.empty()) ent.assign(4, 0);
    if (pair.empty()) pair.push_back(0);
    if (ext.empty()) ext.push_back(0);
    BAAS_TRY(upload(s->icl_rec, rec, st));
    BAAS_TRY(upload(s->icl_ent, ent, st));
    BAAS_TRY(upload(s->icl_pair, pair, st));
    BAAS_TRY(upload(s->icl_unit, unit, st));
    BAAS_TRY(upload(s->icl_ext, ext, st));
    BAAS_TRY(s->icl_pcount.alloc(1));
    if (hipMemsetAsync(s->icl_pcount.p, 0, sizeof(int), st) != hipSuccess) return BAAS_HIP_ERROR;
    BAAS_TRY(s->icl_xp.alloc(2 * nnz * sizeof(double)));
    s->icl_armed = false;
    if (hipStreamSynchronize(st) != hipSuccess) return BAAS_HIP_ERROR;
    s->icl_lds = lds;
    s->icl_status = BAAS_OK;
    return BAAS_OK;
}

}  // namespace baas

extern "C" int baas_replan(baas_schedule *s, int mode, int rows_per_unit, int block_threads) {
    if (!s || mode < 1 || mode > 6 || rows_per_unit < 0 || block_threads < 0 || block_threads > 1024 ||
        (block_threads % 64) != 0)
        return BAAS_INVALID_ARG;
    const baas_opts old = s->opts;
    if (block_threads) s->opts.block_threads = block_threads;
    s->opts.plan_mode = mode;
    s->opts.rows_per_unit = rows_per_unit;
    const int rc = baas::build_plans_mode(s, mode, rows_per_unit);
    if (rc != BAAS_OK) s->opts = old;  // the plans are rebuilt only on success
    return rc;
}

extern "C" int baas_replan_tiles(baas_schedule *s, int row_range, int band_levels, int block_threads) {
    if (!s || row_range < 1 || band_levels < 1) return BAAS_INVALID_ARG;
    const int old = s->tile_band;
    s->tile_band = band_levels;
    const int rc = baas_replan(s, BAAS_PLAN_TILES, row_range, block_threads);
    if (rc != BAAS_OK) s->tile_band = old;
    return rc;
}

extern "C" int baas_schedule_pipe(const baas_schedule *s, int *info) {
    if (!s || !info) return BAAS_INVALID_ARG;
    info[0] = s->fwd.pipe;
    info[1] = s->fwd.pipe && s->fwd.pipe_fact;
    return BAAS_OK;
}

extern "C" int baas_schedule_plan(const baas_schedule *s, int *plan) {
    if (!s || !plan) return BAAS_INVALID_ARG;
    plan[0] = s->fwd.mode;
    plan[1] = s->opts.rows_per_unit;
    plan[2] = s->fwd.block;
    plan[3] = s->fwd.dataflow;
    plan[4] = s->fwd.mode == BAAS_PLAN_TILES ? s->tile_band : 0;
    return BAAS_OK;
}
