/*
 * oracle.h -- CPU restatement of the BAAS sparse triangular pipeline.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline).  The product (baas_amd/, libbaas_hip.so) never
 * links, loads or calls it.
 *
 * Conventions (shared with the reference, spm/include/CSR.hpp:16-27):
 * 0-based int32 indices, rows sorted ascending by column, double values.
 * Lower-triangular CSR keeps the diagonal LAST in each row, upper-triangular
 * CSR keeps it FIRST; full matrices carry diagptr.
 *
 * Floating-point order follows the reference build (-O3 -march=native -mfma,
 * GCC's default -ffp-contract=fast): every "s += a*b" is an explicit fma(a,b,s)
 * and every "s -= a*b" is fma(-a,b,s); the file is compiled with
 * -ffp-contract=off so nothing else is contracted.
 *
 * Parity pinning of each function is listed at its definition in oracle.c.
 */
#ifndef BAAS_ORACLE_H
#define BAAS_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- structure ---------------------------------------------------------- */
int  orc_lower_nnz(int n, const int *rowptr, const int *diagptr);
void orc_extract_lower(int n, const int *rowptr, const int *colidx, const int *diagptr,
                       int *lrowptr, int *lcolidx);
void orc_diagptr(int n, const int *rowptr, const int *colidx, int *diagptr);
void orc_csr_to_csc(int m, int n, const int *rowptr, const int *colidx, const double *val,
                    int *colptr, int *rowidx, double *cval);
int  orc_two_hop_reduce(int n, const int *ptr, const int *set, int *out_ptr, int *out_set);
void orc_inverse_dag(int n, const int *ptr, const int *set, int *inv_ptr, int *inv_set);
int  orc_levelset_csc(int n, const int *ptr, const int *set,
                      int *level_ptr, int *node_to_level, int *perm_to_orig);
int  orc_reverse_coarsen(int n, const int *ptr, const int *set, int *group_ptr, int *group_set);
int  orc_forward_coarsen(int n, const int *ptr, const int *set, int *group_ptr, int *group_set);
int  orc_build_group_dag(int n, int ngroups, const int *group_ptr, const int *group_set,
                         const int *ptr, const int *set, int *gptr, int *gset);
void orc_group_remap(int ngroups_f, const int *group_ptr, const int *group_set,
                     const int *group_ptr_f, const int *group_set_f,
                     int *new_ptr, int *new_set);

enum { ORC_KIND_TRSV = 0, ORC_KIND_IC0 = 1, ORC_KIND_ILU0 = 2 };
void orc_nthreads_per_level(int kind, int nlevels, const int *level_ptr, const int *perm_to_orig,
                            const int *group_ptr, const int *group_set, const int *rowptr,
                            int ncore, int *nthreads);
int  orc_level_merge(int nlevels, const int *level_ptr, int *nthreads, int *merged_ptr);
void orc_group_cost(int ngroups, const int *group_ptr, const int *group_set, const int *lrowptr, int *cost);

/* The whole BAAS analysis of a lower-triangular pattern (buildset()).
 * Outputs (caller allocates n+1 / n ints each):
 *   group_ptr[ngroups+1], group_set[n]  (groups sorted ascending inside)
 *   level_ptr[nlevels+1], perm_to_orig[ngroups], node_to_level[ngroups] (1-based)
 *   nthreads[nlevels], merged_ptr[nmerged+1]
 * counts[0..3] = {ngroups, nlevels, nmerged, reduced_edges}. */
void orc_baas_schedule(int kind, int n, const int *lrowptr, const int *lcolidx,
                       const int *thr_rowptr, int ncore,
                       int *group_ptr, int *group_set, int *level_ptr, int *perm_to_orig,
                       int *node_to_level, int *nthreads, int *merged_ptr, int *counts);

/* ---- numerics ----------------------------------------------------------- */
void orc_sptrsv_lower(int n, const int *rowptr, const int *colidx, const double *val,
                      const double *b, double *x);
void orc_sptrsv_upper(int n, const int *rowptr, const int *colidx, const double *val,
                      const double *b, double *x);
void orc_sptrsv_upper_desc(int n, const int *rowptr, const int *colidx, const double *val,
                           const double *b, double *x);
void orc_sptrsv_lower_f32(int n, const int *rowptr, const int *colidx, const float *val,
                          const float *b, float *x);
void orc_sptrsv_upper_desc_f32(int n, const int *rowptr, const int *colidx, const float *val,
                               const float *b, float *x);
void orc_sptrsv_lower_levels_omp(int n, const int *rowptr, const int *colidx, const double *val,
                                 const double *b, double *x, int nlevels, const int *level_ptr,
                                 const int *perm_to_orig, int nthreads);
int omp_max_threads(void);
void orc_sptrsv_lower_p2p_omp(int n, const int *rowptr, const int *colidx, const double *val,
                              const double *b, double *x, int nlevels, const int *level_ptr,
                              const int *perm_to_orig, const int *lvl, const int *own, int nthreads,
                              int *flags, int epoch);
void orc_sptrsv_lower_multi(int n, const int *rowptr, const int *colidx, const double *val,
                            int k, const double *b, double *x);
void orc_ic0(int n, const int *rowptr, const int *colidx, const int *diagptr,
             const double *val, double *lu);
void orc_ilu0(int n, const int *rowptr, const int *colidx, const int *diagptr, double *lu);
void orc_ic0_split(int n, const int *rowptr, const int *colidx, const int *diagptr,
                   const double *lu, int *Lp, int *Li, double *Lx, int *Up, int *Ui, double *Ux);
void orc_spmv(int m, const int *rowptr, const int *colidx, const double *val,
              const double *x, double *y);
void orc_dcsrmv(int m, const int *rowptr, const int *colidx, const double *val,
                const double *x, double *y, double alpha, double beta);
double orc_dot(int n, const double *x, const double *y);
double orc_norm(int n, const double *x);
void orc_waxpby(int n, double *w, double alpha, const double *x, double beta, const double *y);

/* PCG as in pcg_ref; precision 0 = all FP64, 1 = IC0 factor stored and applied in FP32.
 * Runs until ||r||/||r0|| < tol or maxiter.  Returns the first iteration k that
 * meets the tolerance (or -1), relres[k] history optional (len maxiter+1). */
int orc_pcg(int n, const int *rowptr, const int *colidx, const int *diagptr, const double *val,
            const double *b, double *x, int maxiter, double tol, int precision,
            double *relres_hist, int stop_at_tol);

#ifdef __cplusplus
}
#endif
#endif
