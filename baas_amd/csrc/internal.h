// internal.h -- shared structures and device helpers of libbaas_hip.so.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <vector>

#include "../../include/baas_hip.h"

#define BAAS_HIP_TRY(expr)                                                        \
    do {                                                                          \
        hipError_t _e = (expr);                                                   \
        if (_e != hipSuccess) {                                                   \
            fprintf(stderr, "baas: %s failed: %s (%s:%d)\n", #expr,               \
                    hipGetErrorString(_e), __FILE__, __LINE__);                   \
            return BAAS_HIP_ERROR;                                                \
        }                                                                         \
    } while (0)

#define BAAS_TRY(expr)                    \
    do {                                  \
        int _s = (expr);                  \
        if (_s != BAAS_OK) return _s;     \
    } while (0)

namespace baas {

// Device-side error word (first error wins) and watchdog accounting.
struct DevStatus {
    int code;
    int detail;
};

// A device buffer owned by the library.
template <class T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    int alloc(size_t count) {
        free();
        n = count;
        if (count == 0) return BAAS_OK;
        if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) {
            p = nullptr;
            n = 0;
            return BAAS_ALLOC;
        }
        return BAAS_OK;
    }
    void free() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DBuf() { free(); }
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
};

// Execution plan: how a triangular sweep (solve or factorisation) is cut
// into work units.  A unit is a contiguous run of the execution order; its
// rows are stored sorted by their sub-level inside the unit (longest path
// restricted to the unit), so a workgroup runs a unit sub-level by sub-level
// with a barrier in between and keeps the unit's solved x in LDS.
// keymode: 0 key(row) = row, 1 key(row) = n-1-row, 2 key(row) = key[row].
// Parents of a row in unit u have key < key_begin[u] (another unit, wait on
// the ready flag) or key in [key_begin[u], key_begin[u+1]) (same unit, LDS).
struct Plan {
    int n = 0;
    int nunits = 0;
    int nsub = 0;
    int keymode = 0;
    int max_unit_keys = 0;
    int block = 256;
    int mode = 0;
    DBuf<int> key_begin;  // [nunits+1]
    DBuf<int> sub_begin;  // [nunits+1] index into sub_ptr
    DBuf<int> sub_ptr;    // [nsub+1]   offsets into rows
    DBuf<int> rows;       // [n]        row ids; bit 31 set = row has a child in a later unit
    DBuf<int> key;        // [n]        keymode 2 only
    // dataflow form (one lane per row, LDS slots instead of sub-level barriers)
    DBuf<int> unit_ptr;   // [nunits+1] unit boundaries in rows
    DBuf<int> scol;       // [nnz]      LDS slot of each entry's parent (0 on the diagonal)
    DBuf<int> ext_ptr;    // [nunits+1] external parents of each unit ...
    DBuf<int> ext_cols;   // [...]      ... in order of first use
    DBuf<int> meta;       // [4n] per position: row | ext flag, first off-diagonal entry, count, LDS slot
    DBuf<int> out_ptr;    // [nunits+1] rows of each unit read by later units ...
    DBuf<int> out_slot;   // [...]      ... as LDS slots, in execution order
    size_t df_lds[2] = {0, 0};  // dataflow kernel LDS bytes per workgroup (FP32, FP64)
    DBuf<int> key_row;    // [n]        row of each key (keymode 2 only)
    int max_slots = 0;    // max over units of keys + external parents
    int max_deg = 0;      // max off-diagonal entries of a row
    // LEVELS plans: row level sets in execution order (meta holds the rows)
    DBuf<int> lptr;       // [nlev+1]
    int nlev = 0;
    int lv_cap = 0;       // at most this many workgroups (0: as the widest level needs)
    int dataflow = 0;     // single-RHS sweeps: 1 dataflow kernel, 2 lane kernel
    // lane kernel: per position, the offset of its entries in its unit's LDS
    // entry area; per unit, the running entry count; LDS bytes (FP32, FP64)
    DBuf<int> lane_eoff;  // [n]
    DBuf<int> lane_eptr;  // [nunits+1]
    std::vector<int> h_lane_units;  // per unit: rows, external parents, entries
    size_t lane_lds(int ts, int nrhs);  // LDS bytes per workgroup (cached per ts, nrhs)
    int lane_lds_key = -1;
    size_t lane_lds_val = 0;
    size_t mr_lds(int ts, int nrhs, bool stage_b);  // the same for the multi-RHS lane kernel
    int mr_lds_key = -1;
    size_t mr_lds_val = 0;
    DBuf<unsigned char> xp_raw;  // 2 x n x nrhs published values (dataflow kernels)
    int xp_type = 0;             // sizeof the value type the halves are armed for
    int xp_nrhs = 0;             // ... and their right-hand sides per row (-k: a PIPE plan's faces for k)
    DBuf<int> pcount;            // [1] sweeps run on this plan
    // PIPE plans: the grid (nx, ny, nz, sy (-1 on a 2D grid), sz) and box
    // shape (TX, TY, K) in pg; per unit the box's origin x, y, z and
    // extents tx | ty << 8 | kz << 16 in this sweep's coordinates (the
    // backward plan's flipped), then the offsets of its own faces and of
    // the three faces it reads in a publication half (-1: none); each
    // half holds pipe_half values; pipe = 1 once they are built
    int pipe = 0;
    int pg[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    DBuf<int> pipe_box;  // [8 nunits]
    long long pipe_half = 0;
    int pipe_fact = 0;  // A (IC0/ILU0 schedules) is the full stencil pipe_pa describes
    std::vector<int> h_unit_ptr;  // host copy: unit boundaries in rows
    std::vector<int> h_rows;      // host copy of rows (without flag bit)
};

template <class T>
int autotune(baas_handle *h, baas_schedule *s, const T *vals, int uplo, int nrhs, int *chosen);

// LDS layout of one unit in the lane kernel (byte offsets, 8-byte aligned
// sections): value slots (keys, then external parents), per position b, the
// diagonal, {entry offset, count | external-child flag}, LDS slot and row,
// then the unit's off-diagonal values and their slots.
// With k right-hand sides the value slots and b hold k values per row.
struct LaneLayout {
    size_t sl, bv, dv, rec, slot, row, ev, es, total;
};
__host__ __device__ inline size_t lane_align8(size_t v) { return (v + 7) & ~size_t(7); }
// Multi-RHS lane kernel (dataflow 3): value slots of the unit's own rows
// only (parents in other units are read from the publication half), no
// staged b; per row the diagonal, record, LDS slot and row id; per entry the
// coefficient and its parent (LDS slot, or ~row for another unit's row).
// With stage_b the unit's b values are copied into LDS too (nk x nrhs).
struct MrLayout {
    size_t sl, bv, dv, rec, slot, row, ev, es, total;
};
__host__ __device__ inline MrLayout mr_layout(size_t nk, size_t E, size_t ts, size_t nrhs, bool stage_b = false) {
    MrLayout l;
    l.sl = 0;
    l.bv = lane_align8(nk * nrhs * ts);
    l.dv = l.bv + (stage_b ? lane_align8(nk * nrhs * ts) : 0);
    l.rec = l.dv + lane_align8(nk * ts);
    l.slot = l.rec + nk * 8;
    l.row = l.slot + lane_align8(nk * 4);
    l.ev = l.row + lane_align8(nk * 4);
    l.es = l.ev + lane_align8(E * ts);
    l.total = l.es + lane_align8(E * 4);
    return l;
}
// PIPE plans: what a box of tx x ty rows by kz planes publishes for the
// boxes after it, in the order they consume it -- x: the x face by
// diagonals d = k + j (ty values each, (kz + ty - 1) ty in all), y: the y
// face by d = k + i ((kz + tx - 1) tx; none on a 2D grid), z: its top
// plane (tx ty).  A box stages at most PIPE_FACE_MAX values of each face.
// Where the entries of row (x, y, z) start in the lower triangle of a full
// 7-point (5-point: sy < 0, a line of nx rows is a plane) stencil grid in
// natural order: one entry per earlier row, plus one per earlier row with
// an x parent, a y parent and a z parent.  The PIPE plan builder checks it
// against the matrix's row pointers.
__host__ __device__ inline int pipe_lp(int nx, int ny, int sy, int sz, int x, int y, int z) {
    const int key = x + (sy > 0 ? sy * y : 0) + sz * z;
    const int cz = key - sz > 0 ? key - sz : 0;
    const int w = (y - 1) * nx + x;
    const int cy = sy > 0 ? z * (sz - nx) + (w > 0 ? w : 0) : 0;
    const int lines = sy > 0 ? z * ny + y : z;
    const int cx = lines * (nx - 1) + (x > 1 ? x - 1 : 0);
    return key + cx + cy + cz;
}
// The same for the full matrix A = L + L^T pattern (IC0 and ILU0 schedules):
// also one entry per earlier row with an x, a y and a z child.
__host__ __device__ inline int pipe_pa(int nx, int ny, int nz, int sy, int sz, int x, int y, int z) {
    const int key = x + (sy > 0 ? sy * y : 0) + sz * z;
    const int lines = sy > 0 ? z * ny + y : z;
    int up = lines * (nx - 1) + x;
    if (sy > 0) {
        const int w = y * nx + x, m = sz - nx;
        up += z * m + (w < m ? w : m);
    }
    const int nm = sz * (nz - 1);
    up += key < nm ? key : nm;
    return pipe_lp(nx, ny, sy, sz, x, y, z) + up;
}
constexpr int PIPE_FACE_MAX = 1024;
constexpr int PIPE_DUMP = 8;  // values after a box's faces that its idle publishing lanes write
__host__ __device__ inline int3 pipe_face_sizes(int tx, int ty, int kz, bool has_y) {
    return make_int3((kz + ty - 1) * ty, has_y ? (kz + tx - 1) * tx : 0, tx * ty);
}
__host__ __device__ inline bool pipe_faces_fit(int TX, int TY, int K, bool has_y) {
    const int3 f = pipe_face_sizes(TX, TY, K, has_y);
    return f.x <= PIPE_FACE_MAX && f.y <= PIPE_FACE_MAX && f.z <= 64;
}
__host__ __device__ inline LaneLayout lane_layout(size_t nk, size_t ne, size_t E, size_t ts, size_t nrhs = 1) {
    LaneLayout l;
    l.sl = 0;
    l.bv = lane_align8((nk + ne) * nrhs * ts);
    l.dv = l.bv + lane_align8(nk * nrhs * ts);
    l.rec = l.dv + lane_align8(nk * ts);
    l.slot = l.rec + nk * 8;
    l.row = l.slot + lane_align8(nk * 4);
    l.ev = l.row + lane_align8(nk * 4);
    l.es = l.ev + lane_align8(E * ts);
    l.total = l.es + lane_align8(E * 4);
    return l;
}
}  // namespace baas

constexpr int BAAS_BARRIER_BASE = 64;
// per-XCD ticket queues of multi-RHS PIPE sweeps, one 128-byte line each
constexpr int BAAS_QTICKET_BASE = BAAS_BARRIER_BASE + 8 * 32;
constexpr int BAAS_COUNTER_INTS = BAAS_QTICKET_BASE + 8 * 32;
// TILES plans: largest work unit (rows); its values fit the dataflow kernel's LDS
constexpr int BAAS_TILE_MAX_ROWS = 6144;

// IC0 lane kernel: LDS layout of one unit (byte offsets, 16-byte aligned
// sections): value slots (the unit's lower entries and diagonals, then the
// external entries it reads), per position record, A's diagonal value and
// its lu position, per entry record and A value, per pair word.
struct IclLayout {
    size_t vals, rec, aii, dgp, ent, eav, pair, total;
};
__host__ __device__ inline size_t icl_align16(size_t v) { return (v + 15) & ~size_t(15); }
__host__ __device__ inline IclLayout icl_layout(size_t nslots, size_t npos, size_t nent, size_t npair) {
    IclLayout l;
    l.vals = 0;
    l.rec = icl_align16(nslots * 8);
    l.aii = l.rec + npos * 16;
    l.dgp = l.aii + icl_align16(npos * 8);
    l.ent = l.dgp + icl_align16(npos * 4);
    l.eav = l.ent + nent * 16;
    l.pair = l.eav + nent * 8;
    l.total = l.pair + icl_align16(npair * 4);
    return l;
}

struct baas_handle {
    hipStream_t stream = nullptr;
    int device = 0;
    int num_cus = 0;
    baas::DevStatus *status = nullptr;  // device
    int *counters = nullptr;            // device counters [BAAS_COUNTER_INTS]: 0 ticket, 1 done,
                                        // 2 epoch; the grid barrier's 8 partial counts at
                                        // BAAS_BARRIER_BASE + 32 p (one 128-byte line each)
    double *dscratch = nullptr;         // device reduction scratch
    size_t dscratch_n = 0;
    int epoch = 0;                      // ready-flag epoch
    long long spin_limit = 1LL << 22;
};

struct baas_schedule {
    baas_handle *h = nullptr;
    int kind = 0;
    int n = 0;
    int nnz_full = 0;   // A (IC0/ILU0) or L (TRSV)
    int nnz_lower = 0;
    // patterns (device)
    baas::DBuf<int> a_rowptr, a_colidx, a_diagptr;  // full A (IC0/ILU0)
    baas::DBuf<int> l_rowptr, l_colidx;              // lower part incl. diagonal
    baas::DBuf<int> l_src;                           // position in A of each L entry
    baas::DBuf<int> u_rowptr, u_colidx;              // U = L^T (diag first)
    baas::DBuf<int> u_src;                           // position in L of each U entry
    baas::DBuf<int> l_row_of;                        // row of each L nonzero
    // schedule exports (host copies)
    std::vector<int> group_ptr, group_set, level_ptr, perm_to_orig, node_to_level, nthreads,
        merged_ptr, row_level_ptr, row_perm, row_level, reduced_ptr, reduced_set, group_cost;
    int ngroups = 0, nlevels = 0, nmerged = 0, reduced_edges = 0, nrowlevels = 0;
    // plans (forward over L, backward over U) and ready flags
    baas::Plan fwd, bwd;
    int plan_gen = 0;       // bumped whenever the plans change (captured graphs re-capture)
    int tile_band = 16;     // TILES plans: row levels per band
    baas::DBuf<int> flags;  // [n] epoch of the last sweep that finished the row
    baas::DBuf<int> row_level_ptr_d, row_perm_d;  // for the level-synchronous solve
    // IC0 dot plans (built at the first factorisation): per lower entry the
    // position of its column's diagonal; per row the (other row's entry,
    // entry << 8 | own entry) pairs of every sparse dot, in the reference's
    // summation order
    baas::DBuf<int> ic_djpos, ic_pptr, ic_pairs;
    bool ic_plans = false;
    // IC0 on the lane kernel, built for the forward plan of generation
    // icl_gen: per position {row, entry offset, count | publish flag, own
    // slot base}; per entry {lu position, diagonal slot, pair offset, pair
    // count | publish flag}; per pair slot << 8 | own index; per unit {own
    // slots, entry start, pair start, external start}; external lu positions
    baas::DBuf<int> icl_rec, icl_ent, icl_pair, icl_unit, icl_ext;
    baas::DBuf<unsigned char> icl_xp;  // 2 x nnz(A) published values
    baas::DBuf<int> icl_pcount;
    int icl_gen = -1, icl_status = 0;
    bool icl_armed = false;  // icl_xp holds sentinels
    size_t icl_lds = 0;
    int icl_fw = 3;  // lower entries of the rows marked for the one-step path (2 or 3)
    bool icl_allfast = false;  // every row with lower entries is marked
    double t_gpu = 0, t_host = 0;
    baas_opts opts{};
    // block-diagonal matrices (a batch of independent systems): the first
    // row of each block and n; BOXES plans then cut each block's own grid
    std::vector<int> block_ptr;
};
