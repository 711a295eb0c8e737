// ICCG driven through Schedule::Scheduler the way the reference's pcg_opt
// and pcg_opt_perm drive it (test/iccg_level_merge.cpp:215-340, :380-533):
// one schedule, IC0 by runIC0, the factor split into L and U = L^T, then per
// CG iteration a forward and a backward solve; then the same on A renumbered
// by the scheduler (permPreprocessing, A->permute, the Perm entry points).
// SpMV and BLAS-1 run on the host here (the reference uses MKL for them).
//   iccg_scheduler <matrix.mtx> <nthread> [tol] [maxiter]
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "baas_runtime.hpp"

using namespace SPM;

static void spmv(const CSR *A, const double *x, double *y) {
    for (int i = 0; i < A->m; i++) {
        double s = 0.0;
        for (int j = A->rowptr[i]; j < A->rowptr[i + 1]; j++) s = std::fma(A->values[j], x[A->colidx[j]], s);
        y[i] = s;
    }
}
static double dot(int n, const double *a, const double *b) {
    double s = 0.0;
    for (int i = 0; i < n; i++) s = std::fma(a[i], b[i], s);
    return s;
}

// L (diagonal last) and U = L^T (diagonal first) from an IC0 factor in A's layout
static void split(const CSR *A, const std::vector<double> &lu, CSR *&L, CSR *&U) {
    const int n = A->m;
    int nl = 0;
    for (int i = 0; i < n; i++) nl += A->diagptr[i] - A->rowptr[i] + 1;
    L = new CSR(n, n, nl);
    U = new CSR(n, n, nl);
    std::vector<int> fill(n + 1, 0);
    for (int i = 0; i < n; i++) {
        L->rowptr[i + 1] = L->rowptr[i] + A->diagptr[i] - A->rowptr[i] + 1;
        for (int j = A->rowptr[i]; j <= A->diagptr[i]; j++) fill[A->colidx[j] + 1]++;
    }
    for (int i = 0; i < n; i++) fill[i + 1] += fill[i];
    std::copy(fill.begin(), fill.end(), U->rowptr);
    for (int i = 0; i < n; i++)
        for (int j = A->rowptr[i]; j <= A->diagptr[i]; j++) {
            const int p = L->rowptr[i] + j - A->rowptr[i];
            L->colidx[p] = A->colidx[j];
            L->values[p] = lu[j];
            const int q = fill[A->colidx[j]]++;
            U->colidx[q] = i;
            U->values[q] = lu[j];
        }
}

// PCG with x0 = 0 and b = A * ones on A (perm: A renumbered by the
// scheduler, the runPerm entry points, as pcg_opt_perm,
// test/iccg_level_merge.cpp:215-340); returns 0 on convergence
static int pcg(Schedule::Scheduler *scheduler, const CSR *A, bool perm, double tol, int maxiter) {
    const int n = A->m, nnz = A->getNnz();
    std::vector<double> lu(nnz);
    if (perm) scheduler->runIC0Perm(A, lu.data());
    else scheduler->runIC0(A, lu.data());
    CSR *L = nullptr, *U = nullptr;
    split(A, lu, L, U);
    auto precond = [&](std::vector<double> &z, std::vector<double> &y, const std::vector<double> &r) {
        if (perm) {
            scheduler->runForwardTRSVPerm(L, y.data(), r.data());
            scheduler->runBackwardTRSVPerm(U, z.data(), y.data());
        } else {
            scheduler->runForwardTRSVNoPerm(L, y.data(), r.data());
            scheduler->runBackwardTRSVNoPerm(U, z.data(), y.data());
        }
    };
    std::vector<double> ones(n, 1.0), b(n), x(n, 0.0), r(n), z(n), y(n), p(n), Ap(n);
    spmv(A, ones.data(), b.data());
    r = b;
    precond(z, y, r);
    p = z;
    double rz = dot(n, r.data(), z.data());
    const double nb = std::sqrt(dot(n, b.data(), b.data()));
    int it = 0;
    double rel = 1.0;
    for (it = 1; it <= maxiter; it++) {
        spmv(A, p.data(), Ap.data());
        const double alpha = rz / dot(n, p.data(), Ap.data());
        for (int i = 0; i < n; i++) x[i] = std::fma(alpha, p[i], x[i]), r[i] = std::fma(-alpha, Ap[i], r[i]);
        rel = std::sqrt(dot(n, r.data(), r.data())) / nb;
        if (rel < tol) break;
        precond(z, y, r);
        const double rz1 = dot(n, r.data(), z.data());
        const double beta = rz1 / rz;
        rz = rz1;
        for (int i = 0; i < n; i++) p[i] = std::fma(beta, p[i], z[i]);
    }
    // the true residual
    spmv(A, x.data(), Ap.data());
    double res = 0.0;
    for (int i = 0; i < n; i++) res += (b[i] - Ap[i]) * (b[i] - Ap[i]);
    res = std::sqrt(res) / nb;
    printf("ICCG%s: %d iterations, recursive relres %.3e, true relres %.3e\n", perm ? " perm" : "", it, rel, res);
    delete L;
    delete U;
    return (rel < tol && res < 10 * tol) ? 0 : 1;
}

int main(int argc, char *argv[]) {
    if (argc < 3) {
        fprintf(stderr, "usage: %s matrix.mtx nthread [tol] [maxiter]\n", argv[0]);
        return 2;
    }
    const int nthread = atoi(argv[2]);
    const double tol = argc > 3 ? atof(argv[3]) : 1e-8;
    const int maxiter = argc > 4 ? atoi(argv[4]) : 5000;
    omp_set_num_threads(nthread);
    CSR *A = new CSR(argv[1]);

    Schedule::Scheduler *scheduler = new Schedule::Scheduler(A, true, true, Schedule::P2P_RULE, nthread);
    scheduler->preprocessing();
    int failures = pcg(scheduler, A, false, tol, maxiter);
    // the renumbered run (test/iccg_level_merge.cpp:229-235)
    scheduler->permPreprocessing();
    const int *inversePerm = scheduler->permToOrigByThread.data();
    const int *forwardPerm = scheduler->origToPermByThread.data();
    CSR *APerm = A->permute(forwardPerm, inversePerm, true);
    failures += pcg(scheduler, APerm, true, tol, maxiter);
    delete APerm;
    delete scheduler;
    delete A;
    return failures ? 1 : 0;
}
