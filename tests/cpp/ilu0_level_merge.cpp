// A driver shaped like the reference's ILU0 experiment (test/ilu0_level_merge.cpp)
// with the MI355X runtime header in place of the reference's headers.
//   ilu0_level_merge <matrix.mtx> <nthread> <result-csv> <num_test>
#include <cassert>
#include <cstdio>
#include <cstdlib>

#include "baas_runtime.hpp"

using namespace SPM;
using namespace Merge;
using namespace ILURT;
using namespace IO_Utils;

int main(int argc, char *argv[]) {
    if (argc < 5) {
        fprintf(stderr, "usage: %s matrix.mtx nthread result.csv num_test\n", argv[0]);
        return 2;
    }
    const char *filename = argv[1];
    const int nthread = atoi(argv[2]);
    omp_set_num_threads(nthread);
    const int num_test = atoi(argv[4]);
    CSR *csrA = new CSR(filename);
    const int n = csrA->n, nnz = csrA->getNnz();
    assert(csrA->m == n);
    double *correct_lu = MALLOC(double, nnz);
    std::copy(csrA->values, csrA->values + nnz, correct_lu);
    ilu0csr_uplooking_ref(csrA, correct_lu);

    CSVManager csv(argv[3], "some address",
                   {"Matrix_Name", "row", "nnz", "test_turns", "Algorithm", "Core", "Scheduling_Time",
                    "Executor_Runtime", "nlevel"},
                   false);
    int failures = 0;
    ILUDemo *variants[] = {new ILU_Serial(csrA, correct_lu, n, nnz, "Serial"),
                           new ILU_LevelSet(csrA, correct_lu, n, nnz, "LevelSet", nthread),
                           new ILU_DUAL_TREE_BFS_WF_Merge_P2P_MapFourRule(csrA, correct_lu, n, nnz,
                                                                          "Dual Group Merge P2P MappingFourRule",
                                                                          nthread)};
    for (ILUDemo *rt : variants) {
        rt->set_num_test(num_test);
        TimeMeasure t = rt->evaluate();
        failures += rt->passed() ? 0 : 1;
        csv.addElementToRecord(filename, "Matrix_Name");
        csv.addElementToRecord(n, "row");
        csv.addElementToRecord(nnz, "nnz");
        csv.addElementToRecord(num_test, "test_turns");
        csv.addElementToRecord(rt->getAlgName(), "Algorithm");
        csv.addElementToRecord(nthread, "Core");
        csv.addElementToRecord(rt->getAnalysisTime(), "Scheduling_Time");
        csv.addElementToRecord(t.getTime(), "Executor_Runtime");
        csv.addElementToRecord(0, "nlevel");
        csv.addRecord();
        delete rt;
    }
    FREE(correct_lu);
    delete csrA;
    printf("failures: %d\n", failures);
    return failures ? 1 : 0;
}
