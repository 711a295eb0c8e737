// A driver shaped like the reference's IC0 experiment (test/ic0_level_merge.cpp):
// the same types, runtime classes and calls, with the MI355X runtime header
// in place of the reference's headers.
//   ic0_level_merge <matrix.mtx> <nthread> <result-csv> <num_test>
#include <cassert>
#include <cstdio>
#include <cstdlib>

#include "baas_runtime.hpp"

using namespace SPM;
using namespace Merge;
using namespace ICRT;
using namespace IO_Utils;

int main(int argc, char *argv[]) {
    if (argc < 5) {
        fprintf(stderr, "usage: %s matrix.mtx nthread result.csv num_test\n", argv[0]);
        return 2;
    }
    const char *filename = argv[1];
    const int nthread = atoi(argv[2]);
    omp_set_num_threads(nthread);
    const std::string res_csv = argv[3];
    const int num_test = atoi(argv[4]);

    CSR *csrA = new CSR(filename);  // full SPD matrix, diagptr built by the loader
    const int n = csrA->n, nnz = csrA->getNnz();
    assert(csrA->m == n);

    // the serial factor every runtime is checked against
    double *correct_lu = MALLOC(double, nnz);
    double *serial_lu = MALLOC(double, nnz);
    std::copy(csrA->values, csrA->values + nnz, serial_lu);
    spic0_csr_uL_serial(csrA, serial_lu);
    std::copy(serial_lu, serial_lu + nnz, correct_lu);
    FREE(serial_lu);

    std::vector<std::string> headers = {"Matrix_Name", "row", "nnz", "test_turns", "Algorithm",
                                        "Core", "Scheduling_Time", "Executor_Runtime", "nlevel"};
    CSVManager csv(res_csv, "some address", headers, false);
    int failures = 0;
    auto record = [&](ICDemo *rt, const std::string &alg, int nlevel) {
        rt->set_num_test(num_test);
        TimeMeasure t = rt->evaluate();
        failures += rt->passed() ? 0 : 1;
        csv.addElementToRecord(filename, "Matrix_Name");
        csv.addElementToRecord(n, "row");
        csv.addElementToRecord(nnz, "nnz");
        csv.addElementToRecord(num_test, "test_turns");
        csv.addElementToRecord(alg, "Algorithm");
        csv.addElementToRecord(nthread, "Core");
        csv.addElementToRecord(rt->getAnalysisTime(), "Scheduling_Time");
        csv.addElementToRecord(t.getTime(), "Executor_Runtime");
        csv.addElementToRecord(nlevel, "nlevel");
        csv.addRecord();
    };

    IC_Serial *serial = new IC_Serial(csrA, correct_lu, n, nnz, "Serial");
    record(serial, "Serial", 1);
    delete serial;

    IC_LevelSet *levelset = new IC_LevelSet(csrA, correct_lu, n, nnz, "LevelSet", nthread);
    levelset->set_num_test(num_test);
    record(levelset, "LevelSet", 0);
    printf("LevelSet wavefronts: %d\n", levelset->getWavefront());
    delete levelset;

    IC_DUAL_TREE_BFS_WF_Merge_P2P_MapFourRule *baas = new IC_DUAL_TREE_BFS_WF_Merge_P2P_MapFourRule(
        csrA, correct_lu, n, nnz, "Dual Group Merge P2P MappingFourRule", nthread);
    record(baas, "Dual Group Merge P2P MappingFourRule", 0);
    printf("BAAS work units: %d\n", baas->getWavefront());
    // the split factor: U must be L transposed
    CSR *L = nullptr;
    int nl = 0;
    for (int i = 0; i < n; i++) nl += csrA->diagptr[i] - csrA->rowptr[i] + 1;
    CSR *U = new CSR(n, n, nl);
    baas->getLU(L, U);
    CSC LT;
    CSR::transpositionToCSC(L, &LT);
    bool same = L->getNnz() == nl;
    for (int i = 0; i <= n && same; i++) same = LT.colptr[i] == U->rowptr[i];
    for (int k = 0; k < nl && same; k++) same = LT.rowidx[k] == U->colidx[k] && LT.values[k] == U->values[k];
    printf("%s\n", same ? "the transposed U matrix is equal to L matrix!" : "U is not L transposed");
    failures += same ? 0 : 1;
    delete L;
    delete U;
    delete baas;

    FREE(correct_lu);
    delete csrA;
    printf("failures: %d\n", failures);
    return failures ? 1 : 0;
}
