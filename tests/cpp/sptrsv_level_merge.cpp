// A driver shaped like the reference's SpTRSV experiment
// (test/SpTRSV_level_merge.cpp): the lower loader (strictly lower part, unit
// diagonal), x = alpha from b = row sums, every runtime variant, one CSV row
// each, with the MI355X runtime header in place of the reference's headers.
//   sptrsv_level_merge <matrix.mtx> <nthread> <result-csv> <num_test>
#include <cassert>
#include <cstdio>
#include <cstdlib>

#include "baas_runtime.hpp"

using namespace SPM;
using namespace SpTRSVRT;
using namespace IO_Utils;

int main(int argc, char *argv[]) {
    if (argc < 5) {
        fprintf(stderr, "usage: %s matrix.mtx nthread result.csv num_test\n", argv[0]);
        return 2;
    }
    const char *matrix_file = argv[1];
    const int nthread = atoi(argv[2]);
    printf("max num thread: %d, running num thread: %d\n", omp_get_max_threads(), nthread);
    omp_set_num_threads(nthread);
    const std::string res_csv = argv[3];
    const int num_test = atoi(argv[4]);
    const double alpha = 1.0;

    CSR *csrA = new CSR(matrix_file, true, 0);
    const int m = csrA->m, n = csrA->n, nnz = csrA->getNnz();
    assert(m == n);
    double *correct_x = MALLOC(double, n);
    std::fill_n(correct_x, m, alpha);
    CSC *cscA = new CSC(m, n, nnz);
    CSR::transpositionToCSC(csrA, cscA);

    CSVManager csv(res_csv, "some address",
                   {"Matrix_Name", "row", "nnz", "test_turns", "Algorithm", "Core", "Scheduling_Time",
                    "Executor_Runtime", "nlevel"},
                   false);
    int failures = 0;
    auto run = [&](FusionDemo *rt, int nlevel) {
        rt->set_num_test(num_test);
        TimeMeasure t = rt->evaluate();
        failures += rt->passed() ? 0 : 1;
        csv.addElementToRecord(matrix_file, "Matrix_Name");
        csv.addElementToRecord(n, "row");
        csv.addElementToRecord(nnz, "nnz");
        csv.addElementToRecord(num_test, "test_turns");
        csv.addElementToRecord(rt->getName(), "Algorithm");
        csv.addElementToRecord(nthread, "Core");
        csv.addElementToRecord(rt->getAnalysisTime(), "Scheduling_Time");
        csv.addElementToRecord(t.getTime(), "Executor_Runtime");
        csv.addElementToRecord(nlevel, "nlevel");
        csv.addRecord();
        delete rt;
    };

    run(new SpRTSV_Serial(csrA, correct_x, n, nnz, alpha, "Serial"), 1);
    {
        auto *wf = new SpTRSV_Wavefront(csrA, correct_x, n, nnz, alpha, "Wavefront", nthread);
        wf->set_num_test(num_test);
        TimeMeasure t = wf->evaluate();
        failures += wf->passed() ? 0 : 1;
        printf("Wavefront levels %d, %.4f ms\n", wf->getWavefront(), t.getTime());
        delete wf;
    }
    run(new SpTRSV_WF_Merge(csrA, correct_x, n, nnz, alpha, "WF Merge", nthread), 0);
    run(new SpTRSV_DUAL_TREE_BFS_WF_Merge(csrA, cscA, correct_x, n, nnz, alpha, "Dual Group Merge", nthread), 0);
    run(new SpTRSV_DUAL_TREE_BFS_WF_Merge_P2P(csrA, cscA, correct_x, n, nnz, alpha, "Dual Group Merge P2P", nthread,
                                              false),
        0);
    run(new SpTRSV_DUAL_TREE_BFS_WF_Merge_P2P(csrA, cscA, correct_x, n, nnz, alpha, "Dual Group Merge P2P Perm",
                                              nthread, true),
        0);
    run(new SpTRSV_DUAL_TREE_BFS_WF_Merge_P2P_MapFourRule(csrA, cscA, correct_x, n, nnz, alpha,
                                                          "Dual Group Merge P2P MappingFourRule", nthread, false),
        0);
    run(new SpTRSV_DUAL_TREE_BFS_WF_Merge_P2P_MapFourRule(csrA, cscA, correct_x, n, nnz, alpha,
                                                          "Dual Group Merge P2P MappingFourRule Perm", nthread, true),
        0);

    FREE(correct_x);
    delete cscA;
    delete csrA;
    printf("failures: %d\n", failures);
    return failures ? 1 : 0;
}
