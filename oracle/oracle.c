/*
 * oracle.c -- CPU restatement of the BAAS sparse triangular pipeline.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Never linked into the product.
 *
 * Every function cites the reference file:line it restates and says how it is
 * pinned (tests/test_oracle.py holds the checks):
 *   G  = golden vectors produced by the compiled reference (oracle/_ref) and
 *        committed under tests/golden/
 *   R  = compared live against oracle/_ref in this container
 *   M  = pinned by mathematics (closed form, invariant, brute force, library)
 * Parity unpinned: none; orc_sptrsv_* have no buildable reference
 * (merge/SpTRSV.cpp needs papi.h) and are pinned by M only (manufactured x=1,
 * scipy spsolve_triangular, transposed-solve identity).
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int *ialloc(long n) { return (int *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int)); }
static int cmp_int(const void *a, const void *b) {
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

/* ------------------------------------------------------------------------ */
/* Lower-part extraction used by every runtime that analyses a full matrix:
 * test/ICRuntime.hpp:416-435, test/Schedule.hpp:134-154.  Pinned: R, G. */
int orc_lower_nnz(int n, const int *rowptr, const int *diagptr) {
    int s = 0;
    for (int i = 0; i < n; i++) s += diagptr[i] - rowptr[i] + 1;
    return s;
}
void orc_extract_lower(int n, const int *rowptr, const int *colidx, const int *diagptr,
                       int *lrowptr, int *lcolidx) {
    lrowptr[0] = 0;
    for (int i = 0; i < n; i++) {
        int len = diagptr[i] - rowptr[i] + 1;
        memcpy(lcolidx + lrowptr[i], colidx + rowptr[i], sizeof(int) * (size_t)len);
        lrowptr[i + 1] = lrowptr[i] + len;
    }
}

/* First column >= i in each row: spm/CSR.cpp:853-871.  Pinned: M (definition). */
void orc_diagptr(int n, const int *rowptr, const int *colidx, int *diagptr) {
    for (int i = 0; i < n; i++) {
        diagptr[i] = rowptr[i + 1];
        for (int j = rowptr[i]; j < rowptr[i + 1]; j++)
            if (colidx[j] >= i) { diagptr[i] = j; break; }
    }
}

/* CSR -> CSC with ascending row indices per column: spm/CSR.cpp:914-950
 * (counting scatter + per-column sort; a stable scatter is already sorted).
 * Pinned: M (scipy tocsc), R. */
void orc_csr_to_csc(int m, int n, const int *rowptr, const int *colidx, const double *val,
                    int *colptr, int *rowidx, double *cval) {
    memset(colptr, 0, sizeof(int) * (size_t)(n + 1));
    for (int j = 0; j < rowptr[m]; j++) colptr[colidx[j] + 1]++;
    for (int c = 0; c < n; c++) colptr[c + 1] += colptr[c];
    int *fill = ialloc(n);
    memcpy(fill, colptr, sizeof(int) * (size_t)n);
    for (int i = 0; i < m; i++)
        for (int j = rowptr[i]; j < rowptr[i + 1]; j++) {
            int p = fill[colidx[j]]++;
            rowidx[p] = i;
            if (val && cval) cval[p] = val[j];
        }
    free(fill);
}

/* Two-hop transitive reduction, non-clique branch: spm/DAG.cpp:295-344.
 * Edge i->j (j a child of i, self entry first and kept) is deleted when j is
 * also a child of some child k of i.  Pinned: M (brute force), R, G. */
int orc_two_hop_reduce(int n, const int *ptr, const int *set, int *out_ptr, int *out_set) {
    int nnz = ptr[n];
    char *del = (char *)calloc((size_t)(nnz > 0 ? nnz : 1), 1);
    for (int i = 0; i < n; i++) {
        int cb = ptr[i] + 1, ce = ptr[i + 1];
        for (int kp = cb; kp < ce; kp++) {
            int k = set[kp];
            for (int jp = ptr[k] + 1; jp < ptr[k + 1]; jp++) {
                int j = set[jp];
                /* lower_bound in i's (sorted) child list */
                int lo = cb, hi = ce;
                while (lo < hi) { int mid = (lo + hi) >> 1; if (set[mid] < j) lo = mid + 1; else hi = mid; }
                if (lo < ce && set[lo] == j) del[lo] = 1;
            }
        }
    }
    int e = 0;
    out_ptr[0] = 0;
    for (int i = 0; i < n; i++) {
        for (int kp = ptr[i]; kp < ptr[i + 1]; kp++)
            if (!del[kp]) out_set[e++] = set[kp];
        out_ptr[i + 1] = e;
    }
    free(del);
    return e;
}

/* DAG transpose, sorted: spm/DAG.cpp:481-564.  CSC [self, children] becomes
 * CSR [parents, self].  Pinned: M (it is a transpose), R. */
void orc_inverse_dag(int n, const int *ptr, const int *set, int *inv_ptr, int *inv_set) {
    orc_csr_to_csc(n, n, ptr, set, NULL, inv_ptr, inv_set, NULL);
}

/* Level set by in-degree BFS, levels sorted ascending: spm/DAG.cpp:572-629
 * (DAG_levelSet_CSC), wrapped by findLevelsPostOrder spm/DAG.cpp:348-392.
 * Closed form: level(v) = 1 + max level(parents), sources at 1.
 * node_to_level is 1-based, level_ptr[0] = 0.  Returns nlevels.
 * Pinned: M (longest path), R, G. */
int orc_levelset_csc(int n, const int *ptr, const int *set,
                     int *level_ptr, int *node_to_level, int *perm_to_orig) {
    int *indeg = (int *)calloc((size_t)(n > 0 ? n : 1), sizeof(int));
    for (int e = 0; e < ptr[n]; e++) indeg[set[e]]++;
    int front = 0, rear = 0, lvl = 0;
    level_ptr[0] = 0;
    for (int i = 0; i < n; i++) if (indeg[i] == 1) perm_to_orig[rear++] = i;
    while (front < rear) {
        int stop = rear;
        lvl++;
        while (front < stop) {
            int cur = perm_to_orig[front++];
            node_to_level[cur] = lvl;
            for (int e = ptr[cur]; e < ptr[cur + 1]; e++) {
                int c = set[e];
                if (c != cur && --indeg[c] == 1) perm_to_orig[rear++] = c;
            }
        }
        level_ptr[lvl] = front;
    }
    for (int l = 0; l < lvl; l++)
        qsort(perm_to_orig + level_ptr[l], (size_t)(level_ptr[l + 1] - level_ptr[l]), sizeof(int), cmp_int);
    free(indeg);
    return lvl;
}

/* Shared tail of both coarsenings: name each BFS group by its root (the max
 * node for reverse trees, the min node for forward trees) and counting-sort
 * the groups by that name, keeping BFS order inside a group.
 * merge/Coarsen.cpp:181-222 (forward) and :353-398 (reverse). */
static int name_and_sort_groups(int n, int ngroups, int *group_ptr, int *group_set, int by_max) {
    int *gid = ialloc(n), *gsize = (int *)calloc((size_t)n + 1, sizeof(int));
    for (int g = 0; g < ngroups; g++) {
        int id = by_max ? 0 : n;
        for (int i = group_ptr[g]; i < group_ptr[g + 1]; i++) {
            int v = group_set[i];
            if (by_max ? (v > id) : (v < id)) id = v;
        }
        gsize[id + 1] = group_ptr[g + 1] - group_ptr[g];
        for (int i = group_ptr[g]; i < group_ptr[g + 1]; i++) gid[group_set[i]] = id;
    }
    for (int i = 1; i < n; i++) gsize[i + 1] += gsize[i];
    int ng = 0;
    group_ptr[0] = 0;
    for (int i = 0; i < n; i++)
        if (gsize[i + 1] != gsize[i]) group_ptr[++ng] = gsize[i + 1];
    int *tmp = ialloc(n);
    for (int i = 0; i < n; i++) tmp[gsize[gid[group_set[i]]]++] = group_set[i];
    memcpy(group_set, tmp, sizeof(int) * (size_t)n);
    free(tmp); free(gid); free(gsize);
    return ng;
}

/* Reverse-tree (convergent) BFS coarsening: merge/Coarsen.cpp:234-403 with
 * isSingleReverseTree :43-71.  Roots are sinks (CSC column of length 1);
 * a FIFO of roots; each group is expanded breadth-first; for a head, its
 * unvisited parents are marked visited and join the group only if every one
 * of them has exactly one child, otherwise they all become new roots.
 * Order dependent -- pinned: R, G (no closed form). */
int orc_reverse_coarsen(int n, const int *ptr, const int *set, int *group_ptr, int *group_set) {
    int *iptr = ialloc(n + 1), *iset = ialloc(ptr[n]);
    orc_inverse_dag(n, ptr, set, iptr, iset);
    int *roots = ialloc(n), *gq = ialloc(n), *par = ialloc(n);
    char *vis = (char *)calloc((size_t)(n > 0 ? n : 1), 1);
    int rf = 0, rr = 0, cnt = 0, ng = 0;
    for (int i = 0; i < n; i++) if (ptr[i + 1] - ptr[i] == 1) roots[rr++] = i;
    while (rf < rr) {
        int head = roots[rf++];
        vis[head] = 1;
        int gf = 0, gr = 0;
        gq[gr++] = head;
        group_ptr[ng++] = cnt;
        while (gf < gr) {
            head = gq[gf++];
            group_set[cnt++] = head;
            int np = 0, tree = 1;
            for (int e = iptr[head]; e < iptr[head + 1] - 1; e++) {
                int p = iset[e];
                if (!vis[p]) {
                    vis[p] = 1;
                    par[np++] = p;
                    if (ptr[p + 1] - ptr[p] != 2) tree = 0;
                }
            }
            if (tree) for (int k = 0; k < np; k++) gq[gr++] = par[k];
            else      for (int k = 0; k < np; k++) roots[rr++] = par[k];
        }
    }
    group_ptr[ng] = cnt;
    ng = name_and_sort_groups(n, ng, group_ptr, group_set, 1);
    free(iptr); free(iset); free(roots); free(gq); free(par); free(vis);
    return ng;
}

/* Forward-tree (divergent) BFS coarsening: merge/Coarsen.cpp:73-229 with
 * isSingleForwardTree :12-40.  Roots are sources (CSR row of the inverse DAG
 * of length 1); unvisited children join only if each has exactly one parent.
 * Pinned: R, G. */
int orc_forward_coarsen(int n, const int *ptr, const int *set, int *group_ptr, int *group_set) {
    int *iptr = ialloc(n + 1), *iset = ialloc(ptr[n]);
    orc_inverse_dag(n, ptr, set, iptr, iset);
    int *roots = ialloc(n), *gq = ialloc(n), *ch = ialloc(n);
    char *vis = (char *)calloc((size_t)(n > 0 ? n : 1), 1);
    int rf = 0, rr = 0, cnt = 0, ng = 0;
    for (int i = 0; i < n; i++) if (iptr[i + 1] - iptr[i] == 1) roots[rr++] = i;
    while (rf < rr) {
        int head = roots[rf++];
        vis[head] = 1;
        int gf = 0, gr = 0;
        gq[gr++] = head;
        group_ptr[ng++] = cnt;
        while (gf < gr) {
            head = gq[gf++];
            group_set[cnt++] = head;
            int nc = 0, tree = 1;
            for (int e = ptr[head] + 1; e < ptr[head + 1]; e++) {
                int c = set[e];
                if (!vis[c]) {
                    vis[c] = 1;
                    ch[nc++] = c;
                    if (iptr[c + 1] - iptr[c] != 2) tree = 0;
                }
            }
            if (tree) for (int k = 0; k < nc; k++) gq[gr++] = ch[k];
            else      for (int k = 0; k < nc; k++) roots[rr++] = ch[k];
        }
    }
    group_ptr[ng] = cnt;
    ng = name_and_sort_groups(n, ng, group_ptr, group_set, 0);
    free(iptr); free(iset); free(roots); free(gq); free(ch); free(vis);
    return ng;
}

/* Group-level CSC DAG, per group the sorted unique child groups (self first):
 * merge/Coarsen.cpp:748-801.  Returns the edge count.  Pinned: M, R. */
int orc_build_group_dag(int n, int ngroups, const int *group_ptr, const int *group_set,
                        const int *ptr, const int *set, int *gptr, int *gset) {
    int *ginv = ialloc(n);
    for (int g = 0; g < ngroups; g++)
        for (int i = group_ptr[g]; i < group_ptr[g + 1]; i++) ginv[group_set[i]] = g;
    int *buf = ialloc(ptr[n]);
    int e = 0;
    gptr[0] = 0;
    for (int g = 0; g < ngroups; g++) {
        int cnt = 0;
        for (int i = group_ptr[g]; i < group_ptr[g + 1]; i++) {
            int v = group_set[i];
            for (int c = ptr[v]; c < ptr[v + 1]; c++) buf[cnt++] = ginv[set[c]];
        }
        qsort(buf, (size_t)cnt, sizeof(int), cmp_int);
        for (int k = 0; k < cnt; k++)
            if (k == 0 || buf[k] != buf[k - 1]) gset[e++] = buf[k];
        gptr[g + 1] = e;
    }
    free(ginv); free(buf);
    return e;
}

/* Compose groups-of-groups into node groups: merge/Coarsen.cpp:896-928.
 * Pinned: M (composition), R. */
void orc_group_remap(int ngroups_f, const int *group_ptr, const int *group_set,
                     const int *group_ptr_f, const int *group_set_f,
                     int *new_ptr, int *new_set) {
    int off = 0;
    new_ptr[0] = 0;
    for (int gf = 0; gf < ngroups_f; gf++) {
        for (int k = group_ptr_f[gf]; k < group_ptr_f[gf + 1]; k++) {
            int g = group_set_f[k];
            for (int i = group_ptr[g]; i < group_ptr[g + 1]; i++) new_set[off++] = group_set[i];
        }
        new_ptr[gf + 1] = off;
    }
}

/* Threads per level: merge/LevelMerge.cpp:164-262 with the thresholds of
 * get_nthreads_per_level_group_{TRSV,IC0,ILU0} :118-161.  A single-level DAG
 * gets min(groups, ncore).  Pinned: M (closed form), R. */
void orc_nthreads_per_level(int kind, int nlevels, const int *level_ptr, const int *perm_to_orig,
                            const int *group_ptr, const int *group_set, const int *rowptr,
                            int ncore, int *nthreads) {
    if (nlevels == 1) {
        int g = level_ptr[1] - level_ptr[0];
        nthreads[0] = g < ncore ? g : ncore;
        return;
    }
    for (int l = 0; l < nlevels; l++) {
        int groups = level_ptr[l + 1] - level_ptr[l];
        long nodes = 0, nnz = 0;
        for (int q = level_ptr[l]; q < level_ptr[l + 1]; q++) {
            int g = perm_to_orig[q];
            nodes += group_ptr[g + 1] - group_ptr[g];
            if (kind == ORC_KIND_TRSV)
                for (int k = group_ptr[g]; k < group_ptr[g + 1]; k++) {
                    int v = group_set[k];
                    nnz += rowptr[v + 1] - rowptr[v];
                }
        }
        int t;
        if (kind == ORC_KIND_TRSV) {
            if (nodes < 16 || groups <= 4 || nnz < 32) t = 1;
            else {
                int avg = (int)ceil(1.0 * (double)nodes / 8);
                int mx = (int)(nnz / 32) > avg ? (int)(nnz / 32) : avg;
                t = mx < groups ? mx : groups;
                t = t < ncore ? t : ncore;
            }
        } else {
            if (nodes < 4 || groups <= 2) t = 1;
            else {
                int avg = (int)ceil(1.0 * (double)nodes / 2);
                t = avg < groups ? avg : groups;
                t = t < ncore ? t : ncore;
            }
        }
        nthreads[l] = t;
    }
}

/* Merge runs of consecutive single-thread levels: the two-pointer loop of
 * merge/LevelMerge.cpp:331-355 (level_merge_group_TRSV).  nthreads is
 * compacted in place; returns the merged level count.  Pinned: M, R. */
/* costComputation, merge/Coarsen.cpp:854-891 (kernel SpTRSV_LL, grouped):
 * cost[g] = nonzeros of group g's rows in the lower CSR.
 * Pinned: the compiled reference (ref_group_cost) on the schedule test
 * matrices, and the invariant sum(cost) = nnz(L) (every row in one group). */
void orc_group_cost(int ngroups, const int *group_ptr, const int *group_set, const int *lrowptr, int *cost) {
    for (int g = 0; g < ngroups; g++) {
        int c = 0;
        for (int k = group_ptr[g]; k < group_ptr[g + 1]; k++) c += lrowptr[group_set[k] + 1] - lrowptr[group_set[k]];
        cost[g] = c;
    }
}

int orc_level_merge(int nlevels, const int *level_ptr, int *nthreads, int *merged_ptr) {
    for (int i = 0; i <= nlevels; i++) merged_ptr[i] = 0;
    int m = 0;
    for (int i = 1; i < nlevels; i++) {
        if (nthreads[i] == 1 && nthreads[i] == nthreads[m]) continue;
        m++;
        nthreads[m] = nthreads[i];
        merged_ptr[m] = level_ptr[i];
    }
    m++;
    merged_ptr[m] = level_ptr[nlevels];
    return m;
}

/* The BAAS buildset() analysis: test/SpTRSVRuntime.hpp:675-735 (TRSV),
 * test/ICRuntime.hpp:411-477 (IC0), test/ILURuntime.hpp:265-330 (ILU0):
 * CSC of L -> two-hop reduction -> reverse coarsening -> group DAG ->
 * forward coarsening -> remap -> group DAG -> level set -> threads per
 * level -> sort nodes inside groups (orderingGroupSet :633-640).
 * The merged level_ptr is the 1-thread-run merge that the FourRule mapping
 * applies (merge/LevelTask.cpp:2627-2683), restated as level_merge.
 * Pinned: G (tests/golden), R. */
void orc_baas_schedule(int kind, int n, const int *lrowptr, const int *lcolidx,
                       const int *thr_rowptr, int ncore,
                       int *group_ptr, int *group_set, int *level_ptr, int *perm_to_orig,
                       int *node_to_level, int *nthreads, int *merged_ptr, int *counts) {
    int nnz = lrowptr[n];
    int *cp = ialloc(n + 1), *ri = ialloc(nnz);
    orc_csr_to_csc(n, n, lrowptr, lcolidx, NULL, cp, ri, NULL);
    int *dp = ialloc(n + 1), *ds = ialloc(nnz);
    int red = orc_two_hop_reduce(n, cp, ri, dp, ds);
    free(cp); free(ri);
    int *gp = ialloc(n + 1), *gs = ialloc(n);
    int ng = orc_reverse_coarsen(n, dp, ds, gp, gs);
    int *g1p = ialloc(ng + 1), *g1s = ialloc(red);
    orc_build_group_dag(n, ng, gp, gs, dp, ds, g1p, g1s);
    free(dp); free(ds);
    int *fp = ialloc(ng + 1), *fs = ialloc(ng);
    int ngf = orc_forward_coarsen(ng, g1p, g1s, fp, fs);
    orc_group_remap(ngf, gp, gs, fp, fs, group_ptr, group_set);
    int *g2p = ialloc(ngf + 1), *g2s = ialloc(g1p[ng]);
    orc_build_group_dag(ng, ngf, fp, fs, g1p, g1s, g2p, g2s);
    int nl = orc_levelset_csc(ngf, g2p, g2s, level_ptr, node_to_level, perm_to_orig);
    orc_nthreads_per_level(kind, nl, level_ptr, perm_to_orig, group_ptr, group_set,
                           thr_rowptr, ncore, nthreads);
    for (int g = 0; g < ngf; g++)
        qsort(group_set + group_ptr[g], (size_t)(group_ptr[g + 1] - group_ptr[g]), sizeof(int), cmp_int);
    int *tmp = ialloc(nl);
    memcpy(tmp, nthreads, sizeof(int) * (size_t)nl);
    int nm = orc_level_merge(nl, level_ptr, tmp, merged_ptr);
    free(tmp);
    counts[0] = ngf; counts[1] = nl; counts[2] = nm; counts[3] = red;
    free(gp); free(gs); free(g1p); free(g1s); free(fp); free(fs); free(g2p); free(g2s);
}

/* ------------------------------------------------------------------------ */
/* Lower solve, diagonal last, ascending FMA sum, true division:
 * merge/SpTRSV.cpp:24-39 and :553-571.  Pinned: M (x=1 manufactured solution,
 * scipy spsolve_triangular, <L^-1 b, g> = <b, L^-T g>). */
void orc_sptrsv_lower(int n, const int *rowptr, const int *colidx, const double *val,
                      const double *b, double *x) {
    for (int i = 0; i < n; i++) {
        double s = 0.0;
        for (int j = rowptr[i]; j < rowptr[i + 1] - 1; j++) s = fma(val[j], x[colidx[j]], s);
        x[i] = (b[i] - s) / val[rowptr[i + 1] - 1];
    }
}

/* Upper solve, diagonal first, ascending sum: merge/SpTRSV.cpp:580-599.  Pinned: M. */
void orc_sptrsv_upper(int n, const int *rowptr, const int *colidx, const double *val,
                      const double *b, double *x) {
    for (int i = n - 1; i >= 0; i--) {
        double s = 0.0;
        for (int j = rowptr[i] + 1; j < rowptr[i + 1]; j++) s = fma(val[j], x[colidx[j]], s);
        x[i] = (b[i] - s) / val[rowptr[i]];
    }
}

/* Upper solve with the DESCENDING sum of the BAAS backward P2P kernel:
 * merge/SpTRSV.cpp:643-652.  Pinned: M. */
void orc_sptrsv_upper_desc(int n, const int *rowptr, const int *colidx, const double *val,
                           const double *b, double *x) {
    for (int i = n - 1; i >= 0; i--) {
        double s = 0.0;
        for (int j = rowptr[i + 1] - 1; j > rowptr[i]; j--) s = fma(x[colidx[j]], val[j], s);
        x[i] = (b[i] - s) / val[rowptr[i]];
    }
}

/* FP32 forms used by the mixed-precision preconditioner (no reference
 * counterpart; same operation order as the FP64 forms).  Pinned: bit-equal
 * to an independent numpy float32 sweep with an exactly rounded fma (itself
 * checked against rational arithmetic), Higham's componentwise backward-error
 * bound for substitution, and the x = 1 manufactured solution. */
void orc_sptrsv_lower_f32(int n, const int *rowptr, const int *colidx, const float *val,
                          const float *b, float *x) {
    for (int i = 0; i < n; i++) {
        float s = 0.0f;
        for (int j = rowptr[i]; j < rowptr[i + 1] - 1; j++) s = fmaf(val[j], x[colidx[j]], s);
        x[i] = (b[i] - s) / val[rowptr[i + 1] - 1];
    }
}
void orc_sptrsv_upper_desc_f32(int n, const int *rowptr, const int *colidx, const float *val,
                               const float *b, float *x) {
    for (int i = n - 1; i >= 0; i--) {
        float s = 0.0f;
        for (int j = rowptr[i + 1] - 1; j > rowptr[i]; j--) s = fmaf(x[colidx[j]], val[j], s);
        x[i] = (b[i] - s) / val[rowptr[i]];
    }
}

/* Level-synchronous OpenMP lower solve (the CPU baseline path):
 * merge/SpTRSV.cpp:41-80.  Same per-row arithmetic as orc_sptrsv_lower, so
 * results are bit-identical to it.  Pinned: M (equals the serial solve). */
void orc_sptrsv_lower_levels_omp(int n, const int *rowptr, const int *colidx, const double *val,
                                 const double *b, double *x, int nlevels, const int *level_ptr,
                                 const int *perm_to_orig, int nthreads) {
    (void)n;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel
    for (int l = 0; l < nlevels; l++) {
#pragma omp for schedule(static)
        for (int q = level_ptr[l]; q < level_ptr[l + 1]; q++) {
            int i = perm_to_orig[q];
            double s = 0.0;
            for (int j = rowptr[i]; j < rowptr[i + 1] - 1; j++) s = fma(val[j], x[colidx[j]], s);
            x[i] = (b[i] - s) / val[rowptr[i + 1] - 1];
        }
    }
}

/* OpenMP threads the next parallel region would use (the CPU baseline
 * reports it). */
int omp_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Point-to-point OpenMP lower solve, the synchronisation of the BAAS P2P
 * kernel (merge/SpTRSV.cpp:421-478): each thread walks its own tasks in
 * schedule order, waits on the ready flags of the tasks its rows depend on,
 * solves them and publishes its own flag; there is no barrier.  A task is a
 * thread's contiguous chunk of one row level (the reference's FourRule
 * mapping of coarsened groups to threads, merge/LevelTask.cpp:2295-2684, is
 * not restated); flags[l * nthreads + t] marks task (l, t) done with value
 * epoch, and a thread remembers the highest level it has seen done per
 * thread, so most parents cost no atomic load.  lvl: 0-based row level,
 * own: the thread whose task holds the row.  flags (nlevels * nthreads ints)
 * stay with the caller, who passes a new epoch per call, so nothing is
 * reset.  Same per-row arithmetic as orc_sptrsv_lower: results are
 * bit-identical to it.  Pinned: M (equals the serial solve at every thread
 * count). */
void orc_sptrsv_lower_p2p_omp(int n, const int *rowptr, const int *colidx, const double *val,
                              const double *b, double *x, int nlevels, const int *level_ptr,
                              const int *perm_to_orig, const int *lvl, const int *own, int nthreads,
                              int *flags, int epoch) {
    (void)n;
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
    {
#ifdef _OPENMP
        const int tid = omp_get_thread_num(), nt = omp_get_num_threads();
#else
        const int tid = 0, nt = 1;
#endif
        int known[256];
        for (int t = 0; t < 256; t++) known[t] = -1;
        if (nt <= 256) {
            for (int l = 0; l < nlevels; l++) {
                const int q0 = level_ptr[l], cnt = level_ptr[l + 1] - q0;
                const int chunk = (cnt + nt - 1) / nt;
                const int qa = q0 + tid * chunk < q0 + cnt ? q0 + tid * chunk : q0 + cnt;
                const int qb = qa + chunk < q0 + cnt ? qa + chunk : q0 + cnt;
                for (int q = qa; q < qb; q++) {
                    const int i = perm_to_orig[q];
                    double s = 0.0;
                    for (int j = rowptr[i]; j < rowptr[i + 1] - 1; j++) {
                        const int c = colidx[j], lc = lvl[c], owner = own[c];
                        if (owner != tid && known[owner] < lc) {
                            while (__atomic_load_n(&flags[lc * nt + owner], __ATOMIC_ACQUIRE) != epoch) {
                            }
                            known[owner] = lc;
                        }
                        s = fma(val[j], x[c], s);
                    }
                    x[i] = (b[i] - s) / val[rowptr[i + 1] - 1];
                }
                __atomic_store_n(&flags[l * nt + tid], epoch, __ATOMIC_RELEASE);
            }
        }
    }
}

/* k right-hand sides, row-major b[i*k + r]; each column is the serial solve.
 * Pinned: M (column-by-column equals orc_sptrsv_lower). */
void orc_sptrsv_lower_multi(int n, const int *rowptr, const int *colidx, const double *val,
                            int k, const double *b, double *x) {
    for (int i = 0; i < n; i++)
        for (int r = 0; r < k; r++) {
            double s = 0.0;
            for (int j = rowptr[i]; j < rowptr[i + 1] - 1; j++)
                s = fma(val[j], x[(long)colidx[j] * k + r], s);
            x[(long)i * k + r] = (b[(long)i * k + r] - s) / val[rowptr[i + 1] - 1];
        }
}

/* Sorted two-pointer sparse dot: merge/IC.cpp:6-29. */
static double sparse_dot(int l1, int u1, int l2, int u2, const int *idx, const double *d) {
    double r = 0.0;
    while (l1 < u1 && l2 < u2) {
        if (idx[l1] == idx[l2]) { r = fma(d[l1], d[l2], r); l1++; l2++; }
        else if (idx[l1] < idx[l2]) l1++;
        else l2++;
    }
    return r;
}

/* The diagonal's self dot sum_c L_ic^2.  In the reference's Release build
 * (-O3 -march=native -mfma) GCC vectorises this loop when the two ranges are
 * the same row: whole blocks of 4 squares are rounded products added in order
 * (vmulpd + in-order vaddsd), the remaining 0-3 terms are FMA-contracted, and
 * rows with at most 3 terms are FMA-contracted throughout.  Restated exactly
 * so the oracle (and the GPU kernel) reproduce the reference's bits. */
static double self_dot(int l, int u, const double *d) {
    int k = u - l;
    double r = 0.0;
    int i = l;
    if (k > 3) {
        int stop = l + (k & ~3);
        for (; i < stop; i++) { double p = d[i] * d[i]; r = r + p; }
    }
    for (; i < u; i++) r = fma(d[i], d[i], r);
    return r;
}

/* Up-looking IC0 in place of lu (lu starts as a copy of val; only the lower
 * part incl. diagonal is written): merge/IC.cpp:32-89.  A non-positive pivot
 * leaves the diagonal unchanged.  Pinned: R (spic0_csr_uL_serial), G,
 * M ((L L^T)_ij = A_ij on the lower pattern). */
void orc_ic0(int n, const int *rowptr, const int *colidx, const int *diagptr,
             const double *val, double *lu) {
    for (int i = 0; i < n; i++) {
        for (int k = rowptr[i]; k < diagptr[i]; k++) {
            int j = colidx[k];
            double dp = sparse_dot(rowptr[i], diagptr[i], rowptr[j], diagptr[j], colidx, lu);
            lu[k] = (val[k] - dp) / lu[diagptr[j]];
        }
        double dv = self_dot(rowptr[i], diagptr[i], lu);
        double a = val[diagptr[i]];
        if (a - dv > 0) lu[diagptr[i]] = sqrt(a - dv);
    }
}

/* Up-looking IKJ ILU0 in place: merge/ILU.cpp:350-386.
 * Pinned: R (ilu0csr_uplooking_ref), G, M ((LU)_ij = A_ij on the pattern). */
void orc_ilu0(int n, const int *rowptr, const int *colidx, const int *diagptr, double *lu) {
    for (int i = 0; i < n; i++) {
        for (int j = rowptr[i]; j < diagptr[i]; j++) {
            int col = colidx[j];
            double tmp = lu[j] / lu[diagptr[col]];
            lu[j] = tmp;
            int k1 = j + 1, k2 = diagptr[col] + 1;
            while (k1 < rowptr[i + 1] && k2 < rowptr[col + 1]) {
                if (colidx[k1] < colidx[k2]) k1++;
                else if (colidx[k1] > colidx[k2]) k2++;
                else { lu[k1] = fma(-tmp, lu[k2], lu[k1]); k1++; k2++; }
            }
        }
    }
}

/* Split an IC0 result into L (CSR lower, diag last) and U = L^T (CSR upper,
 * diag first, pattern of A's upper part): test/iccg_level_merge.cpp:809-891.
 * Pinned: M (U equals scipy's transpose of L in pattern and bits; L L^T = A
 * on the pattern). */
void orc_ic0_split(int n, const int *rowptr, const int *colidx, const int *diagptr,
                   const double *lu, int *Lp, int *Li, double *Lx, int *Up, int *Ui, double *Ux) {
    Lp[0] = 0; Up[0] = 0;
    for (int i = 0; i < n; i++) {
        Lp[i + 1] = Lp[i] + diagptr[i] - rowptr[i] + 1;
        Up[i + 1] = Up[i] + rowptr[i + 1] - diagptr[i];
    }
    for (int i = 0; i < n; i++) {
        memcpy(Li + Lp[i], colidx + rowptr[i], sizeof(int) * (size_t)(diagptr[i] - rowptr[i] + 1));
        memcpy(Ui + Up[i], colidx + diagptr[i], sizeof(int) * (size_t)(rowptr[i + 1] - diagptr[i]));
    }
    int *fill = ialloc(n + 1);
    memcpy(fill, Up, sizeof(int) * (size_t)(n + 1));
    for (int i = 0; i < n; i++)
        for (int j = rowptr[i]; j <= diagptr[i]; j++) {
            Lx[Lp[i] + j - rowptr[i]] = lu[j];
            Ux[fill[colidx[j]]++] = lu[j];
        }
    free(fill);
}

/* y = A x, ascending FMA sum: blas/spmv.cpp:8-27.  Pinned: M (scipy). */
void orc_spmv(int m, const int *rowptr, const int *colidx, const double *val,
              const double *x, double *y) {
    for (int i = 0; i < m; i++) {
        double s = 0.0;
        for (int j = rowptr[i]; j < rowptr[i + 1]; j++) s = fma(val[j], x[colidx[j]], s);
        y[i] = s;
    }
}
/* y = alpha*(A x) + beta*y: blas/spmv.cpp:29-45 (contracted as fma(alpha,s,beta*y)).
 * Pinned: G (golden vectors from the compiled reference's dcsrmv), alpha=1,
 * beta=0 equal to SpMV bit for bit, R (live, where oracle/_ref is built). */
void orc_dcsrmv(int m, const int *rowptr, const int *colidx, const double *val,
                const double *x, double *y, double alpha, double beta) {
    for (int i = 0; i < m; i++) {
        double s = 0.0;
        for (int j = rowptr[i]; j < rowptr[i + 1]; j++) s = fma(val[j], x[colidx[j]], s);
        y[i] = fma(alpha, s, beta * y[i]);
    }
}
/* Serial dot (the 1-thread form of blas/blasone.cpp:7-16).  Pinned: M. */
double orc_dot(int n, const double *x, const double *y) {
    double s = 0.0;
    for (int i = 0; i < n; i++) s = fma(x[i], y[i], s);
    return s;
}
double orc_norm(int n, const double *x) { return sqrt(orc_dot(n, x, x)); }
/* w = alpha x + beta y with the special cases of blas/blasone.cpp:24-96.  Pinned: M, R. */
void orc_waxpby(int n, double *w, double alpha, const double *x, double beta, const double *y) {
    for (int i = 0; i < n; i++) {
        double r;
        if (alpha == 1.0) {
            if (beta == 1.0) r = x[i] + y[i];
            else if (beta == -1.0) r = x[i] - y[i];
            else r = fma(beta, y[i], x[i]);
        } else if (beta == 1.0) r = fma(alpha, x[i], y[i]);
        else if (alpha == -1.0) r = (beta == 0.0) ? -x[i] : fma(beta, y[i], -x[i]);
        else if (beta == 0.0) r = alpha * x[i];
        else r = fma(alpha, x[i], beta * y[i]);
        w[i] = r;
    }
}

/* IC0-preconditioned CG, order of operations of pcg_ref
 * (test/iccg_level_merge.cpp:698-807): r = b - A x; z = U^-1 L^-1 r; p = z;
 * loop { Ap; alpha = rz/(p.Ap); x += alpha p; r -= alpha Ap; rel = |r|/|r0|;
 * z = U^-1 L^-1 r; rz' = r.z; beta = rz'/rz; p = z + beta p }.
 * The reference never breaks; here stop_at_tol stops at the first k with
 * rel < tol (the reference's intended stop, :319-323).  precision 1 stores
 * the factor in FP32 and runs both triangular solves in FP32 (no reference
 * counterpart).  The backward solve sums descending like the BAAS kernel
 * (merge/SpTRSV.cpp:643-652).  Pinned: M (converges; FP64 variant equals
 * the reference loop step for step on the same BLAS order). */
int orc_pcg(int n, const int *rowptr, const int *colidx, const int *diagptr, const double *val,
            const double *b, double *x, int maxiter, double tol, int precision,
            double *relres_hist, int stop_at_tol) {
    int nnz = rowptr[n];
    double *lu = (double *)malloc(sizeof(double) * (size_t)nnz);
    memcpy(lu, val, sizeof(double) * (size_t)nnz);
    orc_ic0(n, rowptr, colidx, diagptr, val, lu);
    int nl = orc_lower_nnz(n, rowptr, diagptr);
    int *Lp = ialloc(n + 1), *Li = ialloc(nl), *Up = ialloc(n + 1), *Ui = ialloc(nl);
    double *Lx = (double *)malloc(sizeof(double) * (size_t)nl);
    double *Ux = (double *)malloc(sizeof(double) * (size_t)nl);
    orc_ic0_split(n, rowptr, colidx, diagptr, lu, Lp, Li, Lx, Up, Ui, Ux);
    float *Lf = NULL, *Uf = NULL, *rf = NULL, *yf = NULL, *zf = NULL;
    if (precision == 1) {
        Lf = (float *)malloc(sizeof(float) * (size_t)nl);
        Uf = (float *)malloc(sizeof(float) * (size_t)nl);
        rf = (float *)malloc(sizeof(float) * (size_t)n);
        yf = (float *)malloc(sizeof(float) * (size_t)n);
        zf = (float *)malloc(sizeof(float) * (size_t)n);
        for (int e = 0; e < nl; e++) { Lf[e] = (float)Lx[e]; Uf[e] = (float)Ux[e]; }
    }
    double *r = (double *)malloc(sizeof(double) * (size_t)n);
    double *z = (double *)malloc(sizeof(double) * (size_t)n);
    double *y = (double *)malloc(sizeof(double) * (size_t)n);
    double *p = (double *)malloc(sizeof(double) * (size_t)n);
    double *Ap = (double *)malloc(sizeof(double) * (size_t)n);
    memcpy(r, b, sizeof(double) * (size_t)n);
    orc_dcsrmv(n, rowptr, colidx, val, x, r, -1.0, 1.0);
    double norm0 = orc_norm(n, r);
#define PRECOND()                                                              \
    do {                                                                       \
        if (precision == 1) {                                                  \
            for (int i = 0; i < n; i++) rf[i] = (float)r[i];                   \
            orc_sptrsv_lower_f32(n, Lp, Li, Lf, rf, yf);                       \
            orc_sptrsv_upper_desc_f32(n, Up, Ui, Uf, yf, zf);                  \
            for (int i = 0; i < n; i++) z[i] = (double)zf[i];                  \
        } else {                                                               \
            orc_sptrsv_lower(n, Lp, Li, Lx, r, y);                             \
            orc_sptrsv_upper_desc(n, Up, Ui, Ux, y, z);                        \
        }                                                                      \
    } while (0)
    PRECOND();
    memcpy(p, z, sizeof(double) * (size_t)n);
    double rz = orc_dot(n, r, z);
    if (relres_hist) relres_hist[0] = 1.0;
    int first = -1;
    for (int k = 1; k <= maxiter; k++) {
        double old_rz = rz;
        orc_spmv(n, rowptr, colidx, val, p, Ap);
        double alpha = old_rz / orc_dot(n, p, Ap);
        orc_waxpby(n, x, 1.0, x, alpha, p);
        orc_waxpby(n, r, 1.0, r, -alpha, Ap);
        double rel = orc_norm(n, r) / norm0;
        if (relres_hist) relres_hist[k] = rel;
        if (rel < tol && first < 0) {
            first = k;
            if (stop_at_tol) break;
        }
        PRECOND();
        rz = orc_dot(n, r, z);
        double beta = rz / old_rz;
        orc_waxpby(n, p, 1.0, z, beta, p);
    }
#undef PRECOND
    free(lu); free(Lp); free(Li); free(Up); free(Ui); free(Lx); free(Ux);
    free(Lf); free(Uf); free(rf); free(yf); free(zf);
    free(r); free(z); free(y); free(p); free(Ap);
    return first;
}
