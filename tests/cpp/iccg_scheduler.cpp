// ICCG driven through Schedule::Scheduler the way the reference's pcg_opt
// drives it (test/iccg_level_merge.cpp:380-533): one schedule, IC0 by
// runIC0, the factor split into L and U = L^T, then per CG iteration a
// forward and a backward solve (runForwardTRSVNoPerm / runBackwardTRSVNoPerm).
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
    const int n = A->m, nnz = A->getNnz();

    Schedule::Scheduler *scheduler = new Schedule::Scheduler(A, true, false, Schedule::P2P_RULE, nthread);
    scheduler->preprocessing();
    std::vector<double> lu(nnz);
    scheduler->runIC0(A, lu.data());
    // L (diagonal last) and U = L^T (diagonal first) from the factor
    int nl = 0;
    for (int i = 0; i < n; i++) nl += A->diagptr[i] - A->rowptr[i] + 1;
    CSR *L = new CSR(n, n, nl), *U = new CSR(n, n, nl);
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
    // PCG with x0 = 0 and b = A * ones
    std::vector<double> ones(n, 1.0), b(n), x(n, 0.0), r(n), z(n), y(n), p(n), Ap(n);
    spmv(A, ones.data(), b.data());
    r = b;
    scheduler->runForwardTRSVNoPerm(L, y.data(), r.data());
    scheduler->runBackwardTRSVNoPerm(U, z.data(), y.data());
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
        scheduler->runForwardTRSVNoPerm(L, y.data(), r.data());
        scheduler->runBackwardTRSVNoPerm(U, z.data(), y.data());
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
    printf("ICCG: %d iterations, recursive relres %.3e, true relres %.3e\n", it, rel, res);
    delete L;
    delete U;
    delete scheduler;
    delete A;
    return (rel < tol && res < 10 * tol) ? 0 : 1;
}
