// ref_harness.cpp -- plain-pointer entry points into the COMPILED reference
// (oracle/_ref/libbaas_ref.so), used only by tests/ and tests/golden/ to pin
// the oracle.  TEST INFRASTRUCTURE ONLY: nothing in the product loads it.
//
// It is linked against the reference's own spm/, merge/ and blas/ sources,
// compiled where they lie under the reference tree by oracle/Makefile.
// merge/SpTRSV.cpp is not part of it: that file includes papi.h, which this
// image does not have, so the reference triangular solves are unbuildable
// here and the oracle's solves are pinned by mathematics instead.
//
// Each entry point repeats the call sequence of a reference runtime; the
// sequence is cited, the work is done by the reference functions.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "SPM.hpp"
#include "Merge.hpp"
#include "blasone.hpp"
#include "spmv.hpp"

using namespace SPM;

namespace {
int *dup_ints(const int *p, int n) {
    int *q = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    std::memcpy(q, p, sizeof(int) * (size_t)n);
    return q;
}
}  // namespace

extern "C" {

// Two-hop reduction of a CSC DAG (spm/DAG.cpp:295-344); returns edge count.
int ref_two_hop(int n, const int *ptr, const int *set, int *out_ptr, int *out_set) {
    std::vector<int> P, S;
    DAG::partialSparsification_CSC(n, ptr[n], ptr, set, P, S, false);
    std::copy(P.begin(), P.end(), out_ptr);
    std::copy(S.begin(), S.end(), out_set);
    return (int)S.size();
}

// Row-level level set of a lower CSR pattern through its CSC DAG
// (spm/DAG.cpp:572-629).  Returns nlevels.
int ref_levelset_rows(int n, const int *lrowptr, const int *lcolidx,
                      int *level_ptr, int *node_to_level, int *perm_to_orig) {
    int nnz = lrowptr[n];
    std::vector<double> v(nnz, 1.0);
    CSR L(n, n, const_cast<int *>(lrowptr), const_cast<int *>(lcolidx), v.data());
    CSC *C = new CSC();
    CSR::transpositionToCSC(&L, C);
    std::memset(level_ptr, 0, sizeof(int) * (size_t)(n + 1));
    int nl = 0;
    DAG_levelSet_CSC(n, C->colptr, C->rowidx, nl, level_ptr, node_to_level, perm_to_orig);
    delete C;
    return nl;
}

// BAAS buildset() analysis, call sequence of test/SpTRSVRuntime.hpp:675-735
// (kind 0), test/ICRuntime.hpp:411-477 (kind 1), test/ILURuntime.hpp:265-330
// (kind 2).  thr_rowptr is the rowptr the runtime hands to the threads-per-
// level rule (L itself for SpTRSV, the full A for IC0/ILU0/ICCG).
// merged_ptr is the 1-thread-run merge of merge/LevelMerge.cpp:313-365
// (level_merge_group_TRSV) for kind 0 and the identical loop of
// merge/LevelTask.cpp:2644-2674 for kinds 1 and 2.
void ref_baas_schedule(int kind, int n, const int *lrowptr, const int *lcolidx,
                       const int *thr_rowptr, int ncore,
                       int *group_ptr_out, int *group_set_out, int *level_ptr_out,
                       int *perm_to_orig_out, int *node_to_level_out, int *nthreads_out,
                       int *merged_ptr_out, int *counts) {
    int nnz = lrowptr[n];
    std::vector<double> v(nnz, 1.0);
    CSR L(n, n, const_cast<int *>(lrowptr), const_cast<int *>(lcolidx), v.data());
    CSC *C = new CSC();
    CSR::transpositionToCSC(&L, C);

    DAG *dag = new DAG(n, nnz);
    dag->DAG_ptr.clear();
    dag->DAG_set.clear();
    DAG::partialSparsification_CSC(n, nnz, C->colptr, C->rowidx, dag->DAG_ptr, dag->DAG_set, false);
    int reduced = (int)dag->DAG_set.size();
    int ngroups = 0;
    std::vector<int> group_ptr, group_set;
    Coarsen::reverseTreeCoarseningBFS_all(n, dag->DAG_ptr, dag->DAG_set, ngroups, group_ptr, group_set);
    DAG *dag_group = new DAG();
    Coarsen::buildGroupDAGParallel(n, ngroups, group_ptr.data(), group_set.data(),
                                   dag->DAG_ptr.data(), dag->DAG_set.data(),
                                   dag_group->DAG_ptr, dag_group->DAG_set);
    dag_group->setNodeNum(ngroups);
    delete dag;

    std::vector<int> group_ptr_f, group_set_f;
    DAG *dag_group_f = new DAG();
    int ngroups_f = 0;
    Coarsen::forwardTreeCoarseningBFS_all(ngroups, dag_group->DAG_ptr, dag_group->DAG_set,
                                          ngroups_f, group_ptr_f, group_set_f, false);
    Coarsen::groupRemapping(group_ptr, group_set, group_ptr_f, group_set_f, ngroups_f);
    Coarsen::buildGroupDAGParallel(ngroups, ngroups_f, group_ptr_f.data(), group_set_f.data(),
                                   dag_group->DAG_ptr.data(), dag_group->DAG_set.data(),
                                   dag_group_f->DAG_ptr, dag_group_f->DAG_set);
    ngroups = ngroups_f;
    dag_group_f->setNodeNum(ngroups);
    dag_group_f->updateEdges();
    LevelSet *levelset = new LevelSet();
    dag_group_f->findLevelsPostOrder(levelset);
    delete dag_group;
    delete dag_group_f;

    int nlevels = levelset->getLevels();
    int nfull = thr_rowptr[n];
    std::vector<int> cols_dummy(nfull, 0);
    std::vector<double> vals_dummy(nfull, 0.0);
    CSR T(n, n, const_cast<int *>(thr_rowptr), cols_dummy.data(), vals_dummy.data());
    int *nthreads = nullptr;
    if (kind == 0)
        Merge::level_nthreads_group_TRSV(levelset, &T, group_ptr.data(), group_set.data(), ncore, nthreads);
    else if (kind == 1)
        Merge::level_nthreads_group_IC0(levelset, &T, group_ptr.data(), group_set.data(), ncore, nthreads);
    else
        Merge::level_nthreads_group_ILU0(levelset, &T, group_ptr.data(), group_set.data(), ncore, nthreads);

    // orderingGroupSet (test/SpTRSVRuntime.hpp:633-640)
    for (int g = 0; g < ngroups; g++)
        std::sort(group_set.begin() + group_ptr[g], group_set.begin() + group_ptr[g + 1]);

    int nmerged = 0;
    if (kind == 0) {
        int *nt_m = nullptr;
        LevelSet *lm = new LevelSet();
        Merge::level_merge_group_TRSV(levelset, &T, nullptr, nt_m, lm, ncore,
                                      group_ptr.data(), group_set.data(), ngroups);
        nmerged = lm->getLevels();
        std::copy(lm->level_ptr, lm->level_ptr + nmerged + 1, merged_ptr_out);
        free(nt_m);
        delete lm;
    } else {
        std::vector<int> tmp(nthreads, nthreads + nlevels);
        std::vector<int> mp(nlevels + 1, 0);
        int m = 0;
        for (int l = 1; l < nlevels;) {
            if (tmp[l] == 1 && tmp[m] == 1) { l++; }
            else { m++; tmp[m] = tmp[l]; mp[m] = levelset->level_ptr[l]; l++; }
        }
        m++;
        mp[m] = levelset->level_ptr[nlevels];
        nmerged = m;
        std::copy(mp.begin(), mp.begin() + m + 1, merged_ptr_out);
    }

    std::copy(group_ptr.begin(), group_ptr.end(), group_ptr_out);
    std::copy(group_set.begin(), group_set.end(), group_set_out);
    std::copy(levelset->level_ptr, levelset->level_ptr + nlevels + 1, level_ptr_out);
    std::copy(levelset->permToOrig, levelset->permToOrig + ngroups, perm_to_orig_out);
    std::copy(levelset->nodeToLevel, levelset->nodeToLevel + ngroups, node_to_level_out);
    std::copy(nthreads, nthreads + nlevels, nthreads_out);
    counts[0] = ngroups; counts[1] = nlevels; counts[2] = nmerged; counts[3] = reduced;
    free(nthreads);
    delete levelset;
    delete C;
}

// costComputation (merge/Coarsen.cpp:854-891) as every runtime calls it:
// SpTRSV_LL on the lower CSR, grouped (test/SpTRSVRuntime.hpp:720,
// test/ICRuntime.hpp:476-477, test/Schedule.hpp:229-230).
void ref_group_cost(int ngroups, const int *group_ptr, const int *group_set, const int *lrowptr,
                    const int *lcolidx, double *cost_out) {
    std::vector<double> cost;
    Coarsen::costComputation(ngroups, nullptr, nullptr, lrowptr, lcolidx, Coarsen::Kernel::SpTRSV_LL,
                             group_ptr, group_set, true, cost);
    std::copy(cost.begin(), cost.end(), cost_out);
}

// IC0, merge/IC.cpp:32-89.  lu must start as a copy of val.
void ref_ic0(int n, const int *rowptr, const int *colidx, const double *val, double *lu) {
    CSR A(n, n, const_cast<int *>(rowptr), const_cast<int *>(colidx), const_cast<double *>(val));
    A.constructDiagPtr();
    spic0_csr_uL_serial(&A, lu);
}

// ILU0, merge/ILU.cpp:350-386.  lu must start as a copy of val.
void ref_ilu0(int n, const int *rowptr, const int *colidx, const double *val, double *lu) {
    CSR A(n, n, const_cast<int *>(rowptr), const_cast<int *>(colidx), const_cast<double *>(val));
    A.constructDiagPtr();
    ilu0csr_uplooking_ref(&A, lu);
}

// blas/spmv.cpp:8-45, blas/blasone.cpp:7-96 (OpenMP; thread count from env).
void ref_spmv(int m, const int *rowptr, const int *colidx, const double *val,
              const double *x, double *y) {
    CSR A(m, m, const_cast<int *>(rowptr), const_cast<int *>(colidx), const_cast<double *>(val));
    double *yy = y;
    spmv_omp(A, x, yy);
}
void ref_dcsrmv(int m, const int *rowptr, const int *colidx, const double *val,
                const double *x, double *y, double alpha, double beta) {
    dcsrmv(m, val, rowptr, colidx, x, y, alpha, beta);
}
double ref_dot(int n, const double *x, const double *y) { return dot(x, y, n); }
void ref_waxpby(int n, double *w, double alpha, const double *x, double beta, const double *y) {
    waxpby(n, w, alpha, x, beta, y);
}

// CSR::storeBin (spm/CSR.cpp:399-435) of a 0-based CSR.
void ref_store_bin(int m, int n, const int *rowptr, const int *colidx, const double *values, const char *path) {
    CSR A(m, n, const_cast<int *>(rowptr), const_cast<int *>(colidx), const_cast<double *>(values));
    A.storeBin(path);
}

// Matrix Market loaders (spm/CSR.cpp:90-125; a name ending in .bin goes
// through CSR::loadBin, spm/CSR.cpp:334-397): lower != 0 uses the "lower"
// loader (strictly lower part + unit diagonal, spm/COO.cpp:554-622).  Call
// with rowptr == nullptr to get m and nnz, then again with buffers.
int ref_load_mtx(const char *path, int lower, int *m_out, int *nnz_out,
                 int *rowptr, int *colidx, double *values) {
    CSR *A = lower ? new CSR(path, true, 0) : new CSR(path, 0, false, 1);
    *m_out = A->m;
    *nnz_out = A->getNnz();
    if (rowptr) {
        std::copy(A->rowptr, A->rowptr + A->m + 1, rowptr);
        std::copy(A->colidx, A->colidx + A->getNnz(), colidx);
        std::copy(A->values, A->values + A->getNnz(), values);
    }
    delete A;
    return 0;
}

}  // extern "C"
