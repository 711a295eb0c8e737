// analysis.hip -- the BAAS schedule analysis (the runtimes' buildset()).
//
// GPU integer kernels build every parallel stage: the CSC transpose of the
// lower pattern (spm/CSR.cpp:914-950), the two-hop transitive reduction
// (spm/DAG.cpp:295-344), the group DAGs (merge/Coarsen.cpp:748-801), the
// group remap (merge/Coarsen.cpp:896-928), the level set (spm/DAG.cpp:572-629,
// computed as longest-path levels by frontier expansion and ordered by a
// stable radix sort on the level, which is the reference's (level, id)
// order), threads per level (merge/LevelMerge.cpp:164-262) and the 1-thread
// level merge (merge/LevelMerge.cpp:331-355).
//
// The two tree coarsenings (merge/Coarsen.cpp:73-229 and :234-403) are FIFO
// walks whose group membership depends on visit order; they run on the host
// over the reduced DAG that the GPU produced, exactly as the reference runs
// them serially.  Everything else stays on the device.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <vector>

#include "internal.h"

namespace baas {

// ---------------------------------------------------------------- kernels
__global__ void k_fill_int(int *p, int v, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_iota(int *p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

// count entries per column (CSR -> CSC)
__global__ void k_col_count(const int *idx, int nnz, int *cnt) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < nnz) atomicAdd(&cnt[idx[e] + 1], 1);
}

// scatter row ids into columns (unsorted inside a column), remember source
__global__ void k_scatter_t(int m, const int *ptr, const int *idx, int *fill, int *oidx, int *osrc) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    for (int e = ptr[i]; e < ptr[i + 1]; e++) {
        int p = atomicAdd(&fill[idx[e]], 1);
        oidx[p] = i;
        osrc[p] = e;
    }
}

// two-hop reduction marks: drop child j of i if j is a child of a child of i
__global__ void k_two_hop_mark(int n, const int *ptr, const int *set, unsigned char *keep) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int cb = ptr[i] + 1, ce = ptr[i + 1];
    keep[ptr[i]] = 1;
    for (int q = cb; q < ce; q++) keep[q] = 1;
    for (int kp = cb; kp < ce; kp++) {
        int k = set[kp];
        for (int jp = ptr[k] + 1; jp < ptr[k + 1]; jp++) {
            int j = set[jp];
            int lo = cb, hi = ce;
            while (lo < hi) {
                int mid = (lo + hi) >> 1;
                if (set[mid] < j) lo = mid + 1; else hi = mid;
            }
            if (lo < ce && set[lo] == j) keep[lo] = 0;
        }
    }
}

__global__ void k_u8_to_int(const unsigned char *a, int *b, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

// compact kept edges: pos = exclusive scan of keep
__global__ void k_compact_edges(int n, const int *ptr, const int *set, const unsigned char *keep,
                                const int *pos, int *optr, int *oset, int total) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    if (i == n) { optr[n] = total; return; }
    optr[i] = pos[ptr[i]];
    for (int e = ptr[i]; e < ptr[i + 1]; e++)
        if (keep[e]) oset[pos[e]] = set[e];
}

// in-degree including the self entry (spm/DAG.cpp:578-586)
__global__ void k_indeg(const int *set, int nnz, int *indeg) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < nnz) atomicAdd(&indeg[set[e]], 1);
}

__global__ void k_sources(int n, const int *indeg, int *level, int *front, int *cnt) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    if (indeg[v] == 1) {
        level[v] = 1;
        front[atomicAdd(cnt, 1)] = v;
    }
}

// expand one frontier: a child whose last parent finishes joins the next level
__global__ void k_expand(const int *ptr, const int *set, const int *cur, int ncur, int lvl,
                         int *indeg, int *level, int *next, int *cnt) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncur) return;
    int v = cur[t];
    for (int e = ptr[v]; e < ptr[v + 1]; e++) {
        int c = set[e];
        if (c == v) continue;
        if (atomicSub(&indeg[c], 1) == 2) {
            level[c] = lvl + 1;
            next[atomicAdd(cnt, 1)] = c;
        }
    }
}

// level_ptr from levels sorted ascending (1-based levels)
__global__ void k_level_ptr(int n, const int *sorted_lvl, int *level_ptr, int nlevels) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    if (i == n) { level_ptr[nlevels] = n; return; }
    if (i == 0 || sorted_lvl[i] != sorted_lvl[i - 1]) level_ptr[sorted_lvl[i] - 1] = i;
}

// node -> group
__global__ void k_group_inv(int ng, const int *gptr, const int *gset, int *ginv) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    for (int q = gptr[g]; q < gptr[g + 1]; q++) ginv[gset[q]] = g;
}

// total edge count per group (before dedup)
__global__ void k_group_edge_count(int ng, const int *gptr, const int *gset, const int *ptr, int *cnt) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    int c = 0;
    for (int q = gptr[g]; q < gptr[g + 1]; q++) {
        int v = gset[q];
        c += ptr[v + 1] - ptr[v];
    }
    cnt[g + 1] = c;
}

__global__ void k_group_edge_fill(int ng, const int *gptr, const int *gset, const int *ptr,
                                  const int *set, const int *ginv, const int *off, int *buf) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    int o = off[g];
    for (int q = gptr[g]; q < gptr[g + 1]; q++) {
        int v = gset[q];
        for (int e = ptr[v]; e < ptr[v + 1]; e++) buf[o++] = ginv[set[e]];
    }
}

// first-of-run flags inside each segment of sorted keys
__global__ void k_unique_flags(int ng, const int *off, const int *keys, int *flag) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    for (int q = off[g]; q < off[g + 1]; q++) flag[q] = (q == off[g] || keys[q] != keys[q - 1]) ? 1 : 0;
}

__global__ void k_unique_compact(int ng, const int *off, const int *keys, const int *flag,
                                 const int *pos, int *optr, int *oset, int total) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g > ng) return;
    if (g == ng) { optr[ng] = total; return; }
    optr[g] = pos[off[g]];
    for (int q = off[g]; q < off[g + 1]; q++)
        if (flag[q]) oset[pos[q]] = keys[q];
}

// group remap: size of each forward group in nodes
__global__ void k_remap_size(int ngf, const int *fptr, const int *fset, const int *gptr, int *sz) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= ngf) return;
    int s = 0;
    for (int q = fptr[f]; q < fptr[f + 1]; q++) {
        int g = fset[q];
        s += gptr[g + 1] - gptr[g];
    }
    sz[f + 1] = s;
}

__global__ void k_remap_copy(int ngf, const int *fptr, const int *fset, const int *gptr,
                             const int *gset, const int *nptr, int *nset) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= ngf) return;
    int o = nptr[f];
    for (int q = fptr[f]; q < fptr[f + 1]; q++) {
        int g = fset[q];
        for (int r = gptr[g]; r < gptr[g + 1]; r++) nset[o++] = gset[r];
    }
}

// group id of each entry of group_set (for the in-group node sort)
__global__ void k_group_of_slot(int ng, const int *gptr, int *gid) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    for (int q = gptr[g]; q < gptr[g + 1]; q++) gid[q] = g;
}

// threads per level, merge/LevelMerge.cpp:118-262: one workgroup per level
// sums its groups' nodes (and, for TRSV, their rows' nonzeros), then applies
// the rule
constexpr int NT_BLOCK = 256;
__global__ __launch_bounds__(NT_BLOCK) void k_nthreads(int kind, int nlevels, const int *lptr, const int *pto,
                                                       const int *gptr, const int *gset, const int *rowptr,
                                                       int ncore, int *nt) {
    const int l = blockIdx.x;
    if (l >= nlevels) return;
    const int groups = lptr[l + 1] - lptr[l];
    if (nlevels == 1) {
        if (threadIdx.x == 0) nt[0] = groups < ncore ? groups : ncore;
        return;
    }
    long long nodes = 0, nnz = 0;
    for (int q = lptr[l] + threadIdx.x; q < lptr[l + 1]; q += NT_BLOCK) {
        const int g = pto[q];
        nodes += gptr[g + 1] - gptr[g];
        if (kind == BAAS_KIND_TRSV)
            for (int k = gptr[g]; k < gptr[g + 1]; k++) {
                const int v = gset[k];
                nnz += rowptr[v + 1] - rowptr[v];
            }
    }
    __shared__ long long s_nodes[NT_BLOCK], s_nnz[NT_BLOCK];
    s_nodes[threadIdx.x] = nodes;
    s_nnz[threadIdx.x] = nnz;
    __syncthreads();
    for (int w = NT_BLOCK / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s_nodes[threadIdx.x] += s_nodes[threadIdx.x + w];
            s_nnz[threadIdx.x] += s_nnz[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    nodes = s_nodes[0];
    nnz = s_nnz[0];
    int t;
    if (kind == BAAS_KIND_TRSV) {
        if (nodes < 16 || groups <= 4 || nnz < 32) {
            t = 1;
        } else {
            int avg = (int)ceil(1.0 * (double)nodes / 8);
            int mx = (int)(nnz / 32) > avg ? (int)(nnz / 32) : avg;
            t = mx < groups ? mx : groups;
            t = t < ncore ? t : ncore;
        }
    } else {
        if (nodes < 4 || groups <= 2) {
            t = 1;
        } else {
            int avg = (int)ceil(1.0 * (double)nodes / 2);
            t = avg < groups ? avg : groups;
            t = t < ncore ? t : ncore;
        }
    }
    nt[l] = t;
}

// Merge of runs of 1-thread levels (merge/LevelMerge.cpp:236-262).  Level i
// >= 1 starts a merged level unless it and level i-1 both have one thread
// (a 1-thread level is absorbed exactly when the last kept level has one
// thread, which is the case exactly when level i-1 has one); a scan of
// these marks gives each kept level its slot.
__global__ void k_merge_marks(int nlevels, const int *nt, int *mark) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nlevels) return;
    mark[i] = i == 0 ? 0 : (nt[i] == 1 && nt[i - 1] == 1 ? 0 : 1);
}
// pos = inclusive scan of the marks: kept level i goes to slot pos[i]; the
// merged count is pos[nlevels-1] + 1 and its end is lptr[nlevels]
__global__ void k_merge_scatter(int nlevels, const int *lptr, const int *nt, const int *pos, int *mptr,
                                int *nmerged) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nlevels) return;
    if (i == 0) mptr[0] = 0;
    const bool kept = i > 0 && !(nt[i] == 1 && nt[i - 1] == 1);
    if (kept) mptr[pos[i]] = lptr[i];
    if (i == nlevels - 1) {
        mptr[pos[i] + 1] = lptr[nlevels];
        *nmerged = pos[i] + 1;
    }
}

// costComputation (merge/Coarsen.cpp:854-891, the SpTRSV_LL form every
// runtime calls): a group's cost is the nonzeros of its rows in the lower
// CSR, summed over the group's nodes
__global__ void k_group_cost(int ngroups, const int *gptr, const int *gset, const int *rowptr, int *cost) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngroups) return;
    int c = 0;
    for (int k = gptr[g]; k < gptr[g + 1]; k++) {
        const int v = gset[k];
        c += rowptr[v + 1] - rowptr[v];
    }
    cost[g] = c;
}

// lower part of a full matrix: diagptr, lengths, copy (test/ICRuntime.hpp:416-435)
__global__ void k_diagptr(int n, const int *rowptr, const int *colidx, int *diagptr, int *llen) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int d = rowptr[i + 1];
    for (int j = rowptr[i]; j < rowptr[i + 1]; j++)
        if (colidx[j] >= i) { d = j; break; }
    diagptr[i] = d;
    llen[i + 1] = d - rowptr[i] + 1;
}

__global__ void k_extract_lower(int n, const int *rowptr, const int *colidx, const int *lrowptr,
                                int *lcolidx, int *lsrc) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int o = lrowptr[i];
    int len = lrowptr[i + 1] - o;
    for (int q = 0; q < len; q++) {
        lcolidx[o + q] = colidx[rowptr[i] + q];
        lsrc[o + q] = rowptr[i] + q;
    }
}

// validation of a lower pattern: sorted strictly increasing columns, last = diagonal
__global__ void k_validate_lower(int n, const int *rowptr, const int *colidx, int *bad) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int b = rowptr[i], e = rowptr[i + 1];
    if (e <= b || colidx[e - 1] != i) { atomicOr(bad, 1); return; }
    for (int j = b; j < e - 1; j++)
        if (colidx[j] < 0 || colidx[j] >= colidx[j + 1]) { atomicOr(bad, 1); return; }
}

__global__ void k_row_of(int n, const int *rowptr, int *row_of) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int e = rowptr[i]; e < rowptr[i + 1]; e++) row_of[e] = i;
}

// ---------------------------------------------------------------- helpers
static inline int nb(long long n, int b = 256) { return (int)((n + b - 1) / b); }

struct Scratch {
    void *p = nullptr;
    size_t n = 0;
    int need(size_t bytes) {
        if (bytes <= n) return BAAS_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        if (hipMalloc(&p, bytes) != hipSuccess) return BAAS_ALLOC;
        n = bytes;
        return BAAS_OK;
    }
    ~Scratch() {
        if (p) (void)hipFree(p);
    }
};

static int exclusive_scan(Scratch &tmp, const int *in, int *out, int n, hipStream_t st) {
    size_t bytes = 0;
    BAAS_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n, st));
    BAAS_TRY(tmp.need(bytes));
    BAAS_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(tmp.p, bytes, in, out, n, st));
    return BAAS_OK;
}

static int inclusive_scan_inplace(Scratch &tmp, int *io, int n, hipStream_t st) {
    size_t bytes = 0;
    BAAS_HIP_TRY(hipcub::DeviceScan::InclusiveSum(nullptr, bytes, io, io, n, st));
    BAAS_TRY(tmp.need(bytes));
    BAAS_HIP_TRY(hipcub::DeviceScan::InclusiveSum(tmp.p, bytes, io, io, n, st));
    return BAAS_OK;
}

static int read_int(const int *d, int *h, hipStream_t st) {
    BAAS_HIP_TRY(hipMemcpyAsync(h, d, sizeof(int), hipMemcpyDeviceToHost, st));
    BAAS_HIP_TRY(hipStreamSynchronize(st));
    return BAAS_OK;
}

// CSR (m rows, n cols, nnz) -> CSC with ascending row ids per column.
// optr[n+1], oidx[nnz], osrc[nnz] (position of each entry in the input).
int transpose_pattern(int m, int n, int nnz, const int *ptr, const int *idx, int *optr, int *oidx,
                      int *osrc, hipStream_t st) {
    Scratch tmp;
    DBuf<int> fill, k2, v2;
    BAAS_TRY(fill.alloc(n + 1));
    hipLaunchKernelGGL(k_fill_int, dim3(nb(n + 1)), dim3(256), 0, st, optr, 0, n + 1);
    if (nnz > 0) hipLaunchKernelGGL(k_col_count, dim3(nb(nnz)), dim3(256), 0, st, idx, nnz, optr);
    BAAS_TRY(inclusive_scan_inplace(tmp, optr, n + 1, st));
    BAAS_HIP_TRY(hipMemcpyAsync(fill.p, optr, sizeof(int) * n, hipMemcpyDeviceToDevice, st));
    if (nnz == 0) return BAAS_OK;
    BAAS_TRY(k2.alloc(nnz));
    BAAS_TRY(v2.alloc(nnz));
    hipLaunchKernelGGL(k_scatter_t, dim3(nb(m)), dim3(256), 0, st, m, ptr, idx, fill.p, k2.p, v2.p);
    // sort row ids inside each column; carry the source position
    size_t bytes = 0;
    int bits = 1;
    while ((1LL << bits) < (long long)m) bits++;
    BAAS_HIP_TRY(hipcub::DeviceSegmentedRadixSort::SortPairs(nullptr, bytes, k2.p, oidx, v2.p, osrc, nnz,
                                                             n, optr, optr + 1, 0, bits, st));
    BAAS_TRY(tmp.need(bytes));
    BAAS_HIP_TRY(hipcub::DeviceSegmentedRadixSort::SortPairs(tmp.p, bytes, k2.p, oidx, v2.p, osrc, nnz, n,
                                                             optr, optr + 1, 0, bits, st));
    return BAAS_OK;
}

// two-hop reduced DAG of a CSC DAG [self, children].  Returns edges in *edges.
static int two_hop_reduce(int n, const int *ptr, const int *set, int nnz, DBuf<int> &optr,
                          DBuf<int> &oset, int *edges, hipStream_t st) {
    Scratch tmp;
    DBuf<unsigned char> keep;
    DBuf<int> kint, pos;
    BAAS_TRY(keep.alloc(nnz));
    BAAS_TRY(kint.alloc(nnz + 1));
    BAAS_TRY(pos.alloc(nnz + 1));
    hipLaunchKernelGGL(k_two_hop_mark, dim3(nb(n)), dim3(256), 0, st, n, ptr, set, keep.p);
    hipLaunchKernelGGL(k_u8_to_int, dim3(nb(nnz)), dim3(256), 0, st, keep.p, kint.p, nnz);
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(1), 0, st, kint.p + nnz, 0, 1);
    BAAS_TRY(exclusive_scan(tmp, kint.p, pos.p, nnz + 1, st));
    int total = 0;
    BAAS_TRY(read_int(pos.p + nnz, &total, st));
    BAAS_TRY(optr.alloc(n + 1));
    BAAS_TRY(oset.alloc(total > 0 ? total : 1));
    hipLaunchKernelGGL(k_compact_edges, dim3(nb(n + 1)), dim3(256), 0, st, n, ptr, set, keep.p, pos.p,
                       optr.p, oset.p, total);
    *edges = total;
    return BAAS_OK;
}

// Level set of a CSC DAG: level_ptr[nl+1], perm_to_orig[n], node_to_level[n] (1-based).
int level_set(int n, const int *ptr, const int *set, int nnz, DBuf<int> &level_ptr,
              DBuf<int> &perm_to_orig, DBuf<int> &node_to_level, int *nlevels, hipStream_t st) {
    Scratch tmp;
    DBuf<int> indeg, fa, fb, cnt, ids, sorted_lvl;
    BAAS_TRY(indeg.alloc(n));
    BAAS_TRY(fa.alloc(n));
    BAAS_TRY(fb.alloc(n));
    BAAS_TRY(cnt.alloc(2));
    BAAS_TRY(node_to_level.alloc(n));
    BAAS_TRY(perm_to_orig.alloc(n));
    hipLaunchKernelGGL(k_fill_int, dim3(nb(n)), dim3(256), 0, st, indeg.p, 0, n);
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(2), 0, st, cnt.p, 0, 2);
    hipLaunchKernelGGL(k_indeg, dim3(nb(nnz)), dim3(256), 0, st, set, nnz, indeg.p);
    hipLaunchKernelGGL(k_sources, dim3(nb(n)), dim3(256), 0, st, n, indeg.p, node_to_level.p, fa.p, cnt.p);
    int ncur = 0, lvl = 1, done = 0;
    BAAS_TRY(read_int(cnt.p, &ncur, st));
    done = ncur;
    int *cur = fa.p, *nxt = fb.p;
    while (ncur > 0) {
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(1), 0, st, cnt.p + 1, 0, 1);
        hipLaunchKernelGGL(k_expand, dim3(nb(ncur)), dim3(256), 0, st, ptr, set, cur, ncur, lvl, indeg.p,
                           node_to_level.p, nxt, cnt.p + 1);
        int nn = 0;
        BAAS_TRY(read_int(cnt.p + 1, &nn, st));
        if (nn == 0) break;
        lvl++;
        done += nn;
        ncur = nn;
        std::swap(cur, nxt);
    }
    if (done != n) return BAAS_INVALID_ARG;  // a cycle: not a DAG
    // stable sort of ids by level == (level, id) order
    BAAS_TRY(ids.alloc(n));
    BAAS_TRY(sorted_lvl.alloc(n));
    hipLaunchKernelGGL(k_iota, dim3(nb(n)), dim3(256), 0, st, ids.p, n);
    int bits = 1;
    while ((1LL << bits) <= (long long)lvl) bits++;
    size_t bytes = 0;
    BAAS_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, node_to_level.p, sorted_lvl.p, ids.p,
                                                    perm_to_orig.p, n, 0, bits, st));
    BAAS_TRY(tmp.need(bytes));
    BAAS_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(tmp.p, bytes, node_to_level.p, sorted_lvl.p, ids.p,
                                                    perm_to_orig.p, n, 0, bits, st));
    BAAS_TRY(level_ptr.alloc(lvl + 1));
    hipLaunchKernelGGL(k_level_ptr, dim3(nb(n + 1)), dim3(256), 0, st, n, sorted_lvl.p, level_ptr.p, lvl);
    BAAS_HIP_TRY(hipStreamSynchronize(st));
    *nlevels = lvl;
    return BAAS_OK;
}

// Group DAG (CSC, per group sorted unique child groups, self first).
static int group_dag(int n, int ng, const int *gptr, const int *gset, const int *ptr, const int *set,
                     DBuf<int> &optr, DBuf<int> &oset, hipStream_t st) {
    Scratch tmp;
    DBuf<int> ginv, off, buf, sorted, flag, pos;
    BAAS_TRY(ginv.alloc(n));
    BAAS_TRY(off.alloc(ng + 1));
    hipLaunchKernelGGL(k_group_inv, dim3(nb(ng)), dim3(256), 0, st, ng, gptr, gset, ginv.p);
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(1), 0, st, off.p, 0, 1);
    hipLaunchKernelGGL(k_group_edge_count, dim3(nb(ng)), dim3(256), 0, st, ng, gptr, gset, ptr, off.p);
    BAAS_TRY(inclusive_scan_inplace(tmp, off.p, ng + 1, st));
    int total = 0;
    BAAS_TRY(read_int(off.p + ng, &total, st));
    BAAS_TRY(buf.alloc(total + 1));
    BAAS_TRY(sorted.alloc(total + 1));
    BAAS_TRY(flag.alloc(total + 1));
    BAAS_TRY(pos.alloc(total + 1));
    hipLaunchKernelGGL(k_group_edge_fill, dim3(nb(ng)), dim3(256), 0, st, ng, gptr, gset, ptr, set, ginv.p,
                       off.p, buf.p);
    int bits = 1;
    while ((1LL << bits) < (long long)ng) bits++;
    size_t bytes = 0;
    BAAS_HIP_TRY(hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, bytes, buf.p, sorted.p, total, ng, off.p,
                                                            off.p + 1, 0, bits, st));
    BAAS_TRY(tmp.need(bytes));
    BAAS_HIP_TRY(hipcub::DeviceSegmentedRadixSort::SortKeys(tmp.p, bytes, buf.p, sorted.p, total, ng, off.p,
                                                            off.p + 1, 0, bits, st));
    hipLaunchKernelGGL(k_unique_flags, dim3(nb(ng)), dim3(256), 0, st, ng, off.p, sorted.p, flag.p);
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(1), 0, st, flag.p + total, 0, 1);
    BAAS_TRY(exclusive_scan(tmp, flag.p, pos.p, total + 1, st));
    int uniq = 0;
    BAAS_TRY(read_int(pos.p + total, &uniq, st));
    BAAS_TRY(optr.alloc(ng + 1));
    BAAS_TRY(oset.alloc(uniq > 0 ? uniq : 1));
    hipLaunchKernelGGL(k_unique_compact, dim3(nb(ng + 1)), dim3(256), 0, st, ng, off.p, sorted.p, flag.p,
                       pos.p, optr.p, oset.p, uniq);
    BAAS_HIP_TRY(hipStreamSynchronize(st));
    return BAAS_OK;
}

// ---------------------------------------------------------------- host FIFO walks
// Name groups by their root (max node for reverse trees, min node for
// forward trees) and counting-sort them by that name keeping BFS order
// inside a group: merge/Coarsen.cpp:181-222 and :353-398.
static int name_and_sort(int n, int ng, std::vector<int> &gp, std::vector<int> &gs, bool by_max) {
    std::vector<int> gid(n), gsize(n + 1, 0);
    for (int g = 0; g < ng; g++) {
        int id = by_max ? 0 : n;
        for (int i = gp[g]; i < gp[g + 1]; i++) {
            int v = gs[i];
            if (by_max ? (v > id) : (v < id)) id = v;
        }
        gsize[id + 1] = gp[g + 1] - gp[g];
        for (int i = gp[g]; i < gp[g + 1]; i++) gid[gs[i]] = id;
    }
    for (int i = 1; i < n; i++) gsize[i + 1] += gsize[i];
    int m = 0;
    gp.assign(1, 0);
    for (int i = 0; i < n; i++)
        if (gsize[i + 1] != gsize[i]) gp.push_back(gsize[i + 1]), m++;
    std::vector<int> tmp(n);
    for (int i = 0; i < n; i++) tmp[gsize[gid[gs[i]]]++] = gs[i];
    gs.swap(tmp);
    return m;
}

// Tree coarsening walk (reverse: merge/Coarsen.cpp:234-403 with
// isSingleReverseTree :43-71; forward: :73-229 with isSingleForwardTree
// :12-40).  A FIFO of roots; each root's group grows breadth first; a head's
// unvisited neighbours on the walk side are marked visited and join the group
// only if each has exactly one edge back (a tree), otherwise all become roots.
static int tree_coarsen(int n, const std::vector<int> &ptr, const std::vector<int> &set,
                        const std::vector<int> &iptr, const std::vector<int> &iset, bool reverse,
                        std::vector<int> &gp, std::vector<int> &gs) {
    std::vector<int> roots;
    roots.reserve(n);
    std::vector<int> gq(n), nb_(n);
    std::vector<char> vis(n, 0);
    for (int i = 0; i < n; i++) {
        int deg = reverse ? (ptr[i + 1] - ptr[i]) : (iptr[i + 1] - iptr[i]);
        if (deg == 1) roots.push_back(i);
    }
    gp.clear();
    gs.assign(n, 0);
    size_t rf = 0;
    int cnt = 0, ng = 0;
    while (rf < roots.size()) {
        int head = roots[rf++];
        vis[head] = 1;
        int gf = 0, gr = 0;
        gq[gr++] = head;
        gp.push_back(cnt);
        ng++;
        while (gf < gr) {
            head = gq[gf++];
            gs[cnt++] = head;
            int k = 0;
            bool tree = true;
            if (reverse) {
                for (int e = iptr[head]; e < iptr[head + 1] - 1; e++) {
                    int p = iset[e];
                    if (!vis[p]) {
                        vis[p] = 1;
                        nb_[k++] = p;
                        if (ptr[p + 1] - ptr[p] != 2) tree = false;
                    }
                }
            } else {
                for (int e = ptr[head] + 1; e < ptr[head + 1]; e++) {
                    int c = set[e];
                    if (!vis[c]) {
                        vis[c] = 1;
                        nb_[k++] = c;
                        if (iptr[c + 1] - iptr[c] != 2) tree = false;
                    }
                }
            }
            if (tree)
                for (int q = 0; q < k; q++) gq[gr++] = nb_[q];
            else
                for (int q = 0; q < k; q++) roots.push_back(nb_[q]);
        }
    }
    gp.push_back(cnt);
    if (cnt != n) return -1;
    return name_and_sort(n, ng, gp, gs, reverse);
}

template <class T>
static int to_host(std::vector<T> &h, const T *d, size_t n, hipStream_t st) {
    h.resize(n);
    if (n) BAAS_HIP_TRY(hipMemcpyAsync(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost, st));
    BAAS_HIP_TRY(hipStreamSynchronize(st));
    return BAAS_OK;
}
template <class T>
static int to_dev(DBuf<T> &d, const std::vector<T> &h, hipStream_t st) {
    BAAS_TRY(d.alloc(h.size() ? h.size() : 1));
    if (h.size()) BAAS_HIP_TRY(hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st));
    return BAAS_OK;
}

int build_plans(baas_schedule *s);  // plan.cpp
int build_plans_mode(baas_schedule *s, int mode, int rows_per_unit);

// The whole analysis.  l_rowptr/l_colidx (device) hold the lower pattern.
// The arrays every schedule derives from its lower pattern alone: the
// pattern's validation, each entry's row, U = L^T's pattern and the
// row-level level set (also what a schedule restored from saved bytes
// recomputes).
int derive_patterns(baas_schedule *s) {
    hipStream_t st = s->h->stream;
    const int n = s->n, nnzL = s->nnz_lower;
    // validate the lower pattern
    {
        DBuf<int> bad;
        BAAS_TRY(bad.alloc(1));
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(1), 0, st, bad.p, 0, 1);
        hipLaunchKernelGGL(k_validate_lower, dim3(nb(n)), dim3(256), 0, st, n, s->l_rowptr.p, s->l_colidx.p,
                           bad.p);
        int hb = 0;
        BAAS_TRY(read_int(bad.p, &hb, st));
        if (hb) return BAAS_INVALID_ARG;
    }
    // row of each L entry (used by the gradient and split kernels)
    BAAS_TRY(s->l_row_of.alloc(nnzL));
    hipLaunchKernelGGL(k_row_of, dim3(nb(n)), dim3(256), 0, st, n, s->l_rowptr.p, s->l_row_of.p);
    // CSC of L == CSR of U = L^T (diag first): the DAG [self, children]
    BAAS_TRY(s->u_rowptr.alloc(n + 1));
    BAAS_TRY(s->u_colidx.alloc(nnzL));
    BAAS_TRY(s->u_src.alloc(nnzL));
    BAAS_TRY(transpose_pattern(n, n, nnzL, s->l_rowptr.p, s->l_colidx.p, s->u_rowptr.p, s->u_colidx.p,
                               s->u_src.p, st));
    // row-level level set (for the level-synchronous solve and the exports)
    {
        DBuf<int> lp, pto, ntl;
        int nl = 0;
        BAAS_TRY(level_set(n, s->u_rowptr.p, s->u_colidx.p, nnzL, lp, pto, ntl, &nl, st));
        s->nrowlevels = nl;
        BAAS_TRY(to_host(s->row_level_ptr, lp.p, nl + 1, st));
        BAAS_TRY(to_host(s->row_perm, pto.p, n, st));
        BAAS_TRY(to_host(s->row_level, ntl.p, n, st));
        BAAS_TRY(s->row_level_ptr_d.alloc(nl + 1));
        BAAS_TRY(s->row_perm_d.alloc(n));
        BAAS_HIP_TRY(hipMemcpyAsync(s->row_level_ptr_d.p, lp.p, sizeof(int) * (nl + 1), hipMemcpyDeviceToDevice, st));
        BAAS_HIP_TRY(hipMemcpyAsync(s->row_perm_d.p, pto.p, sizeof(int) * n, hipMemcpyDeviceToDevice, st));
    }
    return BAAS_OK;
}

int run_analysis(baas_schedule *s, const int *thr_rowptr) {
    hipStream_t st = s->h->stream;
    const int n = s->n, nnzL = s->nnz_lower;
    double host_s = 0;
    auto t0 = std::chrono::steady_clock::now();
    BAAS_TRY(derive_patterns(s));
    // two-hop reduction
    DBuf<int> dptr, dset;
    int red = 0;
    BAAS_TRY(two_hop_reduce(n, s->u_rowptr.p, s->u_colidx.p, nnzL, dptr, dset, &red, st));
    s->reduced_edges = red;
    // inverse of the reduced DAG (CSR [parents, self])
    DBuf<int> iptr, iset, isrc;
    BAAS_TRY(iptr.alloc(n + 1));
    BAAS_TRY(iset.alloc(red));
    BAAS_TRY(isrc.alloc(red));
    BAAS_TRY(transpose_pattern(n, n, red, dptr.p, dset.p, iptr.p, iset.p, isrc.p, st));
    std::vector<int> h_ptr, h_set, h_iptr, h_iset;
    BAAS_TRY(to_host(h_ptr, dptr.p, n + 1, st));
    BAAS_TRY(to_host(h_set, dset.p, red, st));
    BAAS_TRY(to_host(h_iptr, iptr.p, n + 1, st));
    BAAS_TRY(to_host(h_iset, iset.p, red, st));
    s->reduced_ptr = h_ptr;
    s->reduced_set = h_set;
    // reverse coarsening (host FIFO walk)
    auto th = std::chrono::steady_clock::now();
    std::vector<int> gp, gs;
    int ng = tree_coarsen(n, h_ptr, h_set, h_iptr, h_iset, true, gp, gs);
    host_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - th).count();
    if (ng < 0) return BAAS_INVALID_ARG;
    DBuf<int> gp_d, gs_d;
    BAAS_TRY(to_dev(gp_d, gp, st));
    BAAS_TRY(to_dev(gs_d, gs, st));
    // group DAG after reverse coarsening
    DBuf<int> g1p, g1s;
    BAAS_TRY(group_dag(n, ng, gp_d.p, gs_d.p, dptr.p, dset.p, g1p, g1s, st));
    int g1e = 0;
    BAAS_TRY(read_int(g1p.p + ng, &g1e, st));
    DBuf<int> g1ip, g1is, g1isrc;
    BAAS_TRY(g1ip.alloc(ng + 1));
    BAAS_TRY(g1is.alloc(g1e));
    BAAS_TRY(g1isrc.alloc(g1e));
    BAAS_TRY(transpose_pattern(ng, ng, g1e, g1p.p, g1s.p, g1ip.p, g1is.p, g1isrc.p, st));
    std::vector<int> h_g1p, h_g1s, h_g1ip, h_g1is;
    BAAS_TRY(to_host(h_g1p, g1p.p, ng + 1, st));
    BAAS_TRY(to_host(h_g1s, g1s.p, g1e, st));
    BAAS_TRY(to_host(h_g1ip, g1ip.p, ng + 1, st));
    BAAS_TRY(to_host(h_g1is, g1is.p, g1e, st));
    // forward coarsening on the group DAG (host FIFO walk)
    th = std::chrono::steady_clock::now();
    std::vector<int> fp, fs;
    int ngf = tree_coarsen(ng, h_g1p, h_g1s, h_g1ip, h_g1is, false, fp, fs);
    host_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - th).count();
    if (ngf < 0) return BAAS_INVALID_ARG;
    DBuf<int> fp_d, fs_d;
    BAAS_TRY(to_dev(fp_d, fp, st));
    BAAS_TRY(to_dev(fs_d, fs, st));
    // remap to node groups
    Scratch tmp;
    DBuf<int> np_d, ns_d;
    BAAS_TRY(np_d.alloc(ngf + 1));
    BAAS_TRY(ns_d.alloc(n));
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(1), 0, st, np_d.p, 0, 1);
    hipLaunchKernelGGL(k_remap_size, dim3(nb(ngf)), dim3(256), 0, st, ngf, fp_d.p, fs_d.p, gp_d.p, np_d.p);
    BAAS_TRY(inclusive_scan_inplace(tmp, np_d.p, ngf + 1, st));
    hipLaunchKernelGGL(k_remap_copy, dim3(nb(ngf)), dim3(256), 0, st, ngf, fp_d.p, fs_d.p, gp_d.p, gs_d.p,
                       np_d.p, ns_d.p);
    // final group DAG and its level set
    DBuf<int> g2p, g2s;
    BAAS_TRY(group_dag(ng, ngf, fp_d.p, fs_d.p, g1p.p, g1s.p, g2p, g2s, st));
    int g2e = 0;
    BAAS_TRY(read_int(g2p.p + ngf, &g2e, st));
    DBuf<int> lp, pto, ntl;
    int nl = 0;
    BAAS_TRY(level_set(ngf, g2p.p, g2s.p, g2e, lp, pto, ntl, &nl, st));
    // threads per level (on group_set before the in-group sort, as the reference)
    DBuf<int> nt, pos, mp, nm, cost;
    BAAS_TRY(nt.alloc(nl));
    BAAS_TRY(pos.alloc(nl));
    BAAS_TRY(mp.alloc(nl + 1));
    BAAS_TRY(nm.alloc(1));
    hipLaunchKernelGGL(k_nthreads, dim3(nl), dim3(NT_BLOCK), 0, st, s->kind, nl, lp.p, pto.p, np_d.p, ns_d.p,
                       thr_rowptr, s->opts.ncore, nt.p);
    hipLaunchKernelGGL(k_merge_marks, dim3(nb(nl)), dim3(256), 0, st, nl, nt.p, pos.p);
    BAAS_TRY(inclusive_scan_inplace(tmp, pos.p, nl, st));
    hipLaunchKernelGGL(k_merge_scatter, dim3(nb(nl)), dim3(256), 0, st, nl, lp.p, nt.p, pos.p, mp.p, nm.p);
    // group costs on the lower pattern (the in-group order does not change them)
    BAAS_TRY(cost.alloc(ngf));
    hipLaunchKernelGGL(k_group_cost, dim3(nb(ngf)), dim3(256), 0, st, ngf, np_d.p, ns_d.p, s->l_rowptr.p, cost.p);
    // sort nodes inside each group (orderingGroupSet, test/SpTRSVRuntime.hpp:633-640)
    DBuf<int> gid, gid2, ns2;
    BAAS_TRY(gid.alloc(n));
    BAAS_TRY(gid2.alloc(n));
    BAAS_TRY(ns2.alloc(n));
    {
        int bits = 1;
        while ((1LL << bits) < (long long)n) bits++;
        size_t bytes = 0;
        BAAS_HIP_TRY(hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, bytes, ns_d.p, ns2.p, n, ngf, np_d.p,
                                                                np_d.p + 1, 0, bits, st));
        BAAS_TRY(tmp.need(bytes));
        BAAS_HIP_TRY(hipcub::DeviceSegmentedRadixSort::SortKeys(tmp.p, bytes, ns_d.p, ns2.p, n, ngf, np_d.p,
                                                                np_d.p + 1, 0, bits, st));
    }
    int nmerged = 0;
    BAAS_TRY(read_int(nm.p, &nmerged, st));
    s->ngroups = ngf;
    s->nlevels = nl;
    s->nmerged = nmerged;
    BAAS_TRY(to_host(s->group_ptr, np_d.p, ngf + 1, st));
    BAAS_TRY(to_host(s->group_set, ns2.p, n, st));
    BAAS_TRY(to_host(s->level_ptr, lp.p, nl + 1, st));
    BAAS_TRY(to_host(s->perm_to_orig, pto.p, ngf, st));
    BAAS_TRY(to_host(s->node_to_level, ntl.p, ngf, st));
    BAAS_TRY(to_host(s->nthreads, nt.p, nl, st));
    BAAS_TRY(to_host(s->merged_ptr, mp.p, nmerged + 1, st));
    BAAS_TRY(to_host(s->group_cost, cost.p, ngf, st));
    double total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    // execution plans
    th = std::chrono::steady_clock::now();
    BAAS_TRY(build_plans(s));
    double plan_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - th).count();
    s->t_host = host_s + plan_s;
    s->t_gpu = total - host_s;
    return BAAS_OK;
}

int prepare_lower(baas_schedule *s, const int *rowptr, const int *colidx, int nnz_full) {
    hipStream_t st = s->h->stream;
    const int n = s->n;
    if (s->kind == BAAS_KIND_TRSV) {
        BAAS_TRY(s->l_rowptr.alloc(n + 1));
        BAAS_TRY(s->l_colidx.alloc(nnz_full));
        BAAS_HIP_TRY(hipMemcpyAsync(s->l_rowptr.p, rowptr, sizeof(int) * (n + 1), hipMemcpyDeviceToDevice, st));
        BAAS_HIP_TRY(hipMemcpyAsync(s->l_colidx.p, colidx, sizeof(int) * nnz_full, hipMemcpyDeviceToDevice, st));
        s->nnz_lower = nnz_full;
        return BAAS_OK;
    }
    Scratch tmp;
    BAAS_TRY(s->a_rowptr.alloc(n + 1));
    BAAS_TRY(s->a_colidx.alloc(nnz_full));
    BAAS_TRY(s->a_diagptr.alloc(n));
    BAAS_HIP_TRY(hipMemcpyAsync(s->a_rowptr.p, rowptr, sizeof(int) * (n + 1), hipMemcpyDeviceToDevice, st));
    BAAS_HIP_TRY(hipMemcpyAsync(s->a_colidx.p, colidx, sizeof(int) * nnz_full, hipMemcpyDeviceToDevice, st));
    BAAS_TRY(s->l_rowptr.alloc(n + 1));
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(1), 0, st, s->l_rowptr.p, 0, 1);
    hipLaunchKernelGGL(k_diagptr, dim3(nb(n)), dim3(256), 0, st, n, s->a_rowptr.p, s->a_colidx.p,
                       s->a_diagptr.p, s->l_rowptr.p);
    BAAS_TRY(inclusive_scan_inplace(tmp, s->l_rowptr.p, n + 1, st));
    int nl = 0;
    BAAS_TRY(read_int(s->l_rowptr.p + n, &nl, st));
    s->nnz_lower = nl;
    BAAS_TRY(s->l_colidx.alloc(nl));
    BAAS_TRY(s->l_src.alloc(nl));
    hipLaunchKernelGGL(k_extract_lower, dim3(nb(n)), dim3(256), 0, st, n, s->a_rowptr.p, s->a_colidx.p,
                       s->l_rowptr.p, s->l_colidx.p, s->l_src.p);
    BAAS_HIP_TRY(hipStreamSynchronize(st));
    return BAAS_OK;
}

}  // namespace baas

using namespace baas;

extern "C" int baas_analyze(baas_handle *h, int n, const int *rowptr, const int *colidx, int kind,
                            const baas_opts *opts, baas_schedule **out) {
    if (!h || !out || n <= 0 || !rowptr || !colidx || kind < 0 || kind > 2) return BAAS_INVALID_ARG;
    *out = nullptr;
    int nnz = 0;
    BAAS_HIP_TRY(hipSetDevice(h->device));
    BAAS_HIP_TRY(hipMemcpyAsync(&nnz, rowptr + n, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    BAAS_HIP_TRY(hipStreamSynchronize(h->stream));
    if (nnz <= 0) return BAAS_INVALID_ARG;
    baas_schedule *s = new baas_schedule();
    s->h = h;
    s->kind = kind;
    s->n = n;
    s->nnz_full = nnz;
    if (opts) s->opts = *opts;
    if (s->opts.ncore <= 0) s->opts.ncore = 8;
    if (s->opts.block_threads <= 0) s->opts.block_threads = 256;
    int st = prepare_lower(s, rowptr, colidx, nnz);
    if (st == BAAS_OK) st = run_analysis(s, kind == BAAS_KIND_TRSV ? s->l_rowptr.p : s->a_rowptr.p);
    if (st == BAAS_OK) st = s->flags.alloc(n);
    if (st == BAAS_OK) {
        hipLaunchKernelGGL(k_fill_int, dim3(nb(n)), dim3(256), 0, h->stream, s->flags.p, 0, n);
        if (hipStreamSynchronize(h->stream) != hipSuccess) st = BAAS_HIP_ERROR;
    }
    if (st != BAAS_OK) {
        delete s;
        return st;
    }
    *out = s;
    return BAAS_OK;
}

// ---------------------------------------------------------------- persistence
// Saved schedule: a header of int32 words, the pattern's digest, then the
// analysis arrays as (int64 length, int32 values); little-endian, as the
// reference's CSR::storeBin writes its matrices (spm/CSR.cpp:399-435).
namespace {
constexpr char SAVE_MAGIC[8] = {'B', 'A', 'A', 'S', 'S', 'C', 'H', 'D'};
constexpr int SAVE_VERSION = 1;
constexpr int SAVE_HDR = 18;  // int32 words after the magic

// FNV-1a over the pattern's bytes: a saved schedule only loads onto the
// pattern it was made for
uint64_t pattern_digest(const std::vector<int> &rp, const std::vector<int> &ci) {
    uint64_t h = 1469598103934665603ULL;
    auto mix = [&](const std::vector<int> &v) {
        const unsigned char *b = reinterpret_cast<const unsigned char *>(v.data());
        for (size_t i = 0; i < v.size() * sizeof(int); i++) h = (h ^ b[i]) * 1099511628211ULL;
    };
    mix(rp);
    mix(ci);
    return h;
}

int pattern_digest_dev(int n, const int *rowptr, const int *colidx, hipStream_t st, uint64_t *out, int *nnz) {
    std::vector<int> rp(n + 1);
    BAAS_HIP_TRY(hipMemcpyAsync(rp.data(), rowptr, sizeof(int) * (n + 1), hipMemcpyDeviceToHost, st));
    BAAS_HIP_TRY(hipStreamSynchronize(st));
    if (rp[n] <= 0) return BAAS_INVALID_ARG;
    std::vector<int> ci(rp[n]);
    BAAS_HIP_TRY(hipMemcpyAsync(ci.data(), colidx, sizeof(int) * rp[n], hipMemcpyDeviceToHost, st));
    BAAS_HIP_TRY(hipStreamSynchronize(st));
    *out = pattern_digest(rp, ci);
    *nnz = rp[n];
    return BAAS_OK;
}

std::vector<std::vector<int> *> saved_arrays(baas_schedule *s) {
    return {&s->group_ptr, &s->group_set, &s->level_ptr, &s->perm_to_orig, &s->node_to_level, &s->nthreads,
            &s->merged_ptr, &s->reduced_ptr, &s->reduced_set, &s->group_cost};
}
}  // namespace

extern "C" int baas_schedule_save(const baas_schedule *cs, void *buf, size_t *len) {
    if (!cs || !len) return BAAS_INVALID_ARG;
    baas_schedule *s = const_cast<baas_schedule *>(cs);
    const int *rp = s->kind == BAAS_KIND_TRSV ? s->l_rowptr.p : s->a_rowptr.p;
    const int *ci = s->kind == BAAS_KIND_TRSV ? s->l_colidx.p : s->a_colidx.p;
    uint64_t dig = 0;
    int nnz = 0;
    BAAS_HIP_TRY(hipSetDevice(s->h->device));
    BAAS_TRY(pattern_digest_dev(s->n, rp, ci, s->h->stream, &dig, &nnz));
    int plan[5];
    BAAS_TRY(baas_schedule_plan(s, plan));
    const int hdr[SAVE_HDR] = {SAVE_VERSION, s->kind, s->n, s->nnz_full, s->nnz_lower, s->opts.ncore,
                               s->opts.merge_threshold, s->ngroups, s->nlevels, s->nmerged, s->reduced_edges,
                               plan[0], plan[1], plan[2], plan[3], s->tile_band, (int)(dig & 0xffffffffu),
                               (int)(dig >> 32)};
    size_t need = sizeof(SAVE_MAGIC) + sizeof(hdr);
    for (auto *v : saved_arrays(s)) need += sizeof(int64_t) + v->size() * sizeof(int);
    if (!buf) {
        *len = need;
        return BAAS_OK;
    }
    if (*len < need) {
        *len = need;
        return BAAS_INVALID_ARG;
    }
    unsigned char *o = static_cast<unsigned char *>(buf);
    auto put = [&](const void *p, size_t b) {
        std::memcpy(o, p, b);
        o += b;
    };
    put(SAVE_MAGIC, sizeof(SAVE_MAGIC));
    put(hdr, sizeof(hdr));
    for (auto *v : saved_arrays(s)) {
        const int64_t k = (int64_t)v->size();
        put(&k, sizeof(k));
        put(v->data(), v->size() * sizeof(int));
    }
    *len = need;
    return BAAS_OK;
}

extern "C" int baas_schedule_load(baas_handle *h, int n, const int *rowptr, const int *colidx, int kind,
                                  const void *buf, size_t len, baas_schedule **out) {
    if (!h || !out || !buf || n <= 0 || !rowptr || !colidx || kind < 0 || kind > 2) return BAAS_INVALID_ARG;
    *out = nullptr;
    const unsigned char *in = static_cast<const unsigned char *>(buf), *end = in + len;
    int hdr[SAVE_HDR];
    if (len < sizeof(SAVE_MAGIC) + sizeof(hdr) || std::memcmp(in, SAVE_MAGIC, sizeof(SAVE_MAGIC)) != 0)
        return BAAS_INVALID_ARG;
    std::memcpy(hdr, in + sizeof(SAVE_MAGIC), sizeof(hdr));
    in += sizeof(SAVE_MAGIC) + sizeof(hdr);
    if (hdr[0] != SAVE_VERSION || hdr[1] != kind || hdr[2] != n) return BAAS_INVALID_ARG;
    BAAS_HIP_TRY(hipSetDevice(h->device));
    uint64_t dig = 0;
    int nnz = 0;
    BAAS_TRY(pattern_digest_dev(n, rowptr, colidx, h->stream, &dig, &nnz));
    if (nnz != hdr[3] || (int)(dig & 0xffffffffu) != hdr[16] || (int)(dig >> 32) != hdr[17]) return BAAS_INVALID_ARG;
    baas_schedule *s = new baas_schedule();
    s->h = h;
    s->kind = kind;
    s->n = n;
    s->nnz_full = nnz;
    s->opts.ncore = hdr[5];
    s->opts.merge_threshold = hdr[6];
    int st = BAAS_OK;
    for (auto *v : saved_arrays(s)) {
        int64_t k = 0;
        if (end - in < (ptrdiff_t)sizeof(k)) { st = BAAS_INVALID_ARG; break; }
        std::memcpy(&k, in, sizeof(k));
        in += sizeof(k);
        if (k < 0 || (end - in) / (ptrdiff_t)sizeof(int) < k) { st = BAAS_INVALID_ARG; break; }
        v->resize((size_t)k);
        std::memcpy(v->data(), in, (size_t)k * sizeof(int));
        in += (size_t)k * sizeof(int);
    }
    if (st == BAAS_OK) {
        s->ngroups = hdr[7];
        s->nlevels = hdr[8];
        s->nmerged = hdr[9];
        s->reduced_edges = hdr[10];
        if ((int)s->group_ptr.size() != s->ngroups + 1 || (int)s->group_set.size() != n ||
            (int)s->level_ptr.size() != s->nlevels + 1 || (int)s->perm_to_orig.size() != s->ngroups ||
            (int)s->merged_ptr.size() != s->nmerged + 1)
            st = BAAS_INVALID_ARG;
    }
    auto t0 = std::chrono::steady_clock::now();
    if (st == BAAS_OK) st = prepare_lower(s, rowptr, colidx, nnz);
    if (st == BAAS_OK && s->nnz_lower != hdr[4]) st = BAAS_INVALID_ARG;
    if (st == BAAS_OK) st = derive_patterns(s);
    double gpu_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    // the saved plan: mode, rows per unit, threads, dataflow kernel, band
    if (st == BAAS_OK) {
        auto t1 = std::chrono::steady_clock::now();
        s->opts.plan_mode = hdr[11];
        s->opts.rows_per_unit = hdr[12];
        s->opts.block_threads = hdr[13];
        s->tile_band = hdr[15];
        st = build_plans_mode(s, hdr[11], hdr[12]);
        if (st == BAAS_OK) st = baas_plan_dataflow(s, hdr[14]);
        s->t_host = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
        s->t_gpu = gpu_s;
    }
    if (st == BAAS_OK) st = s->flags.alloc(n);
    if (st == BAAS_OK) {
        hipLaunchKernelGGL(k_fill_int, dim3(nb(n)), dim3(256), 0, h->stream, s->flags.p, 0, n);
        if (hipStreamSynchronize(h->stream) != hipSuccess) st = BAAS_HIP_ERROR;
    }
    if (st != BAAS_OK) {
        delete s;
        return st;
    }
    *out = s;
    return BAAS_OK;
}

extern "C" int baas_schedule_destroy(baas_schedule *s) {
    delete s;
    return BAAS_OK;
}

extern "C" int baas_schedule_info(const baas_schedule *s, int *c) {
    if (!s || !c) return BAAS_INVALID_ARG;
    c[0] = s->ngroups;
    c[1] = s->nlevels;
    c[2] = s->nmerged;
    c[3] = s->reduced_edges;
    c[4] = s->nrowlevels;
    c[5] = s->fwd.nunits;
    c[6] = s->nnz_lower;
    c[7] = s->fwd.mode;
    return BAAS_OK;
}

extern "C" int baas_schedule_times(const baas_schedule *s, double *g, double *hs) {
    if (!s) return BAAS_INVALID_ARG;
    if (g) *g = s->t_gpu;
    if (hs) *hs = s->t_host;
    return BAAS_OK;
}

extern "C" int baas_schedule_export(const baas_schedule *s, int which, int *out, long long *len) {
    if (!s || !len) return BAAS_INVALID_ARG;
    const std::vector<int> *v = nullptr;
    switch (which) {
        case BAAS_EXPORT_GROUP_PTR: v = &s->group_ptr; break;
        case BAAS_EXPORT_GROUP_SET: v = &s->group_set; break;
        case BAAS_EXPORT_LEVEL_PTR: v = &s->level_ptr; break;
        case BAAS_EXPORT_PERM_TO_ORIG: v = &s->perm_to_orig; break;
        case BAAS_EXPORT_NODE_TO_LEVEL: v = &s->node_to_level; break;
        case BAAS_EXPORT_NTHREADS: v = &s->nthreads; break;
        case BAAS_EXPORT_MERGED_PTR: v = &s->merged_ptr; break;
        case BAAS_EXPORT_ROW_LEVEL_PTR: v = &s->row_level_ptr; break;
        case BAAS_EXPORT_ROW_PERM: v = &s->row_perm; break;
        case BAAS_EXPORT_ROW_LEVEL: v = &s->row_level; break;
        case BAAS_EXPORT_REDUCED_PTR: v = &s->reduced_ptr; break;
        case BAAS_EXPORT_REDUCED_SET: v = &s->reduced_set; break;
        case BAAS_EXPORT_UNIT_PTR: v = &s->fwd.h_unit_ptr; break;
        case BAAS_EXPORT_UNIT_ROWS: v = &s->fwd.h_rows; break;
        case BAAS_EXPORT_GROUP_COST: v = &s->group_cost; break;
        default: return BAAS_INVALID_ARG;
    }
    *len = (long long)v->size();
    if (out && !v->empty()) std::memcpy(out, v->data(), v->size() * sizeof(int));
    return BAAS_OK;
}

extern "C" int baas_schedule_upper_pattern(const baas_schedule *s, const int **rp, const int **ci) {
    if (!s) return BAAS_INVALID_ARG;
    if (rp) *rp = s->u_rowptr.p;
    if (ci) *ci = s->u_colidx.p;
    return BAAS_OK;
}
