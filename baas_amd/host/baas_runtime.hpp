// baas_runtime.hpp -- the reference's host interface for the sparse
// triangular pipeline, backed by the MI355X library (include/baas_hip.h).
//
// A caller written against the reference's runtimes includes this header in
// place of SPM.hpp / Merge.hpp / MarcoUtils.hpp / TimeMeasure.hpp /
// csv_utils.hpp / FusionDemo.hpp / SpTRSVRuntime.hpp / ICRuntime.hpp /
// ILURuntime.hpp / Schedule.hpp and keeps its code: the types, namespaces,
// class names, constructor arguments and methods below are the reference's
// (each cites the declaration it replaces).  Every numeric call runs on the
// GPU through the C-ABI; host arrays go in and come out, device copies are
// made in buildset()/setup() and fused_code() times the device work only
// (HIP events), as the reference's fused_code() times its OpenMP kernel.
//
// The reference's OpenMP variants are kept as names and map onto execution
// plans of the one GPU schedule (every plan gives the reference's serial
// bits): *_Serial -> row ranges, *_Wavefront / *_LevelSet -> the row level
// sets with device-wide barriers, *_WF_Merge / *_DUAL_TREE_BFS_WF_Merge ->
// units of the merged BAAS levels, *_P2P -> the same units on the dataflow
// lane kernel, *_MapFourRule -> the autotuned plan (the GPU counterpart of
// the FourRule thread mapping, which ready flags replace).
#pragma once
#include <hip/hip_runtime.h>
#include <omp.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/baas_hip.h"

// utils/MarcoUtils.hpp:19-45
#ifndef MALLOC
#define MALLOC(type, count) ((type *)malloc((size_t)(count) * sizeof(type)))
#endif
#ifndef FREE
#define FREE(x)          \
    {                    \
        if (x) free(x);  \
        x = nullptr;     \
    }
#endif
#ifndef CHECK_POINTER
#define CHECK_POINTER(ptr)                                                             \
    do {                                                                               \
        if ((ptr) == NULL) {                                                           \
            fprintf(stderr, "null pointer at %s:%d\n", __FILE__, __LINE__);            \
            exit(EXIT_FAILURE);                                                        \
        }                                                                              \
    } while (0)
#endif

// utils/TimeMeasure.hpp: wall time in milliseconds (here, for fused_code(),
// the device time of the timed calls).
class TimeMeasure {
    std::chrono::time_point<std::chrono::steady_clock> start_time, end_time;
    bool is_measuring = false;
    bool measure_done = true;

public:
    enum class TimeMetric { SECONDS, MILLISECONDS, MICROSECONDS };
    double elasped_time = 0.0;
    void start_timer();
    void measure_elasped_time();
    double getTime(TimeMetric tm = TimeMetric::MILLISECONDS) const;
    std::string getTimeStr(TimeMetric tm = TimeMetric::MILLISECONDS) const;
};

namespace IO_Utils {

// utils/csv_utils.hpp: one CSV file, a header of column names, records
// filled column by column.  A file that cannot be opened is reported once
// and the records are dropped (the reference does the same).
class CSVManager {
    std::string name_;
    std::vector<std::string> cols_;
    std::vector<std::string> rec_;
    std::ofstream out_;

public:
    CSVManager(std::string output_name, std::string output_address, std::vector<std::string> column_names,
               bool force_creation);
    template <typename T>
    void addElementToRecord(const T &v, const std::string &column) {
        std::ostringstream os;
        os << v;
        set(column, os.str());
    }
    void addElementToRecord(const char *v, const std::string &column) { set(column, v); }
    void addElementToRecord(char *v, const std::string &column) { set(column, v); }
    void addElementToRecord(const std::string &v, const std::string &column) { set(column, v); }
    void addRecord();
    const std::string &fileName() const { return name_; }

private:
    void set(const std::string &column, const std::string &v);
};

}  // namespace IO_Utils

namespace SPM {

class CSR;

// spm/include/CSC.hpp:11-40
class CSC {
public:
    int m = 0, n = 0;
    int *colptr = nullptr, *rowidx = nullptr;
    double *values = nullptr;
    CSC() = default;
    CSC(int m, int n, int nnz);
    ~CSC();
    int getNnz() const { return colptr ? colptr[n] - colptr[0] : 0; }
    CSC(const CSC &) = delete;
    CSC &operator=(const CSC &) = delete;

private:
    bool ownData_ = true;
    friend class CSR;
};

// spm/include/CSR.hpp:16-40: 0-based CSR with rows sorted by column.  The
// file constructors follow spm/CSR.cpp:90-125 (a name ending in .bin is read
// with loadBin, anything else as Matrix Market): the full loader
// expands symmetric files and builds diagptr / diag / idiag (exact diagonal
// match), the lower loader keeps the strictly lower entries and puts 1.0 on
// the diagonal (spm/COO.cpp:554-622).
class CSR {
public:
    int m = 0;
    int n = 0;
    int *rowptr = nullptr;
    int *colidx = nullptr;
    double *values = nullptr;
    double *idiag = nullptr;  // inverse of the diagonal
    double *diag = nullptr;
    int *diagptr = nullptr;
    int *extptr = nullptr;  // non-local columns (MPI builds); always NULL here

    CSR() = default;
    CSR(const char *fileName, int base = 0, bool forceSymmetric = false, int pad = 1);
    CSR(const char *fileName, bool isLower, int base = 0);
    CSR(int m, int n, int nnz);
    CSR(const CSR &A);
    CSR(int m, int n, int *rowptr, int *colidx, double *values);  // non-owning
    ~CSR();
    CSR &operator=(const CSR &) = delete;

    int getNnz() const { return rowptr[m] - rowptr[0]; }
    // spm/Permute.cpp:215-405: row i of the result is row rowInversePerm[i]
    // (the identity if NULL) with column c renamed columnPerm[c]; sort puts
    // each row's entries in ascending column order (the GPU analysis needs
    // it); diagptr / diag / idiag follow the rows when present.
    CSR *permute(const int *columnPerm, const int *rowInversePerm, bool sort = false) const;
    // spm/CSR.cpp:332-435: the binary file {1211216, m, n, nnz, row
    // lengths[m], colidx[nnz], values[nnz]} (native int32 / double);
    // loadBin rebuilds rowptr and diagptr / diag / idiag.  0-based only.
    void loadBin(const char *fileName, int base = 0);
    void storeBin(const char *fileName);
    // first column >= row in each row (spm/CSR.cpp:853-892)
    void constructDiagPtr();
    void computeInverseDiag();
    // column j = ascending rows of column j (spm/CSR.cpp:895-950)
    static void transpositionToCSC(const CSR *A, CSC *C);

private:
    bool ownData_ = false;
    void alloc(int m, int nnz);
};

}  // namespace SPM

namespace Merge {

// The reference's serial kernels (merge/IC.cpp:32-89, merge/ILU.cpp:350-386,
// merge/SpTRSV.cpp:24-39): the same results, computed on the GPU through the
// C-ABI (every execution plan gives the serial kernel's bits).
void spic0_csr_uL_serial(const SPM::CSR *A, double *lu);
void ilu0csr_uplooking_ref(SPM::CSR *A, double *lu);
void sptrsv_serial_csr(const SPM::CSR &A, double x[], const double y[]);

}  // namespace Merge

namespace baas_rt {

// Throws std::runtime_error on any non-OK status.
void check(int status, const char *what);

// Device copies of a CSR (pattern + values).
struct DeviceCsr {
    int n = 0, nnz = 0;
    int *rowptr = nullptr, *colidx = nullptr;
    double *values = nullptr;
    void upload(const SPM::CSR &A);
    void release();
    ~DeviceCsr() { release(); }
};

// A library handle, its stream and two timing events.
class Engine {
public:
    Engine();
    ~Engine();
    baas_handle *h = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    void start();
    double stop_ms();  // waits, returns the device time since start()
    Engine(const Engine &) = delete;
    Engine &operator=(const Engine &) = delete;
};

// One analysed matrix: schedule + plan choice.  plan_mode AUTO autotunes.
struct GpuSchedule {
    Engine eng;
    DeviceCsr dA;
    baas_schedule *sched = nullptr;
    int kind = BAAS_KIND_TRSV;
    void analyze(const SPM::CSR &A, int kind, int ncore, int plan_mode, int rows_per_unit, int block, int dataflow);
    int info(int which) const;  // baas_schedule_info counts[which]
    ~GpuSchedule();
};

}  // namespace baas_rt

namespace SpTRSVRT {

// test/FusionDemo.hpp:14-58: x = L^-1 y for a lower-triangular CSR with
// y = row sums times alpha, so that x = alpha (test/FusionDemo.cpp:47-77).
class FusionDemo {
protected:
    TimeMeasure timeMeasure;
    bool hasPerm = false;
    int n = 0;
    int nnz = 0;
    std::vector<int> permToOrig, origToPerm;
    double *y = nullptr;
    double *x = nullptr;
    double *correct_x = nullptr;
    double alpha = 1.0;
    std::string algName;
    int num_test = 100;
    int num_thread = 1;
    double elasped_time = 0;
    double analysis_time = 0;
    SPM::CSR *csrA = nullptr;
    SPM::CSR *csrA_perm = nullptr;

    // the GPU side
    baas_rt::GpuSchedule gs;
    double *d_b = nullptr, *d_x = nullptr;
    int plan_mode = BAAS_PLAN_AUTO, rows_per_unit = 0, block = 0, dataflow = 0;

    virtual void setup();
    virtual void buildset();
    virtual TimeMeasure fused_code();
    virtual void testing();

public:
    FusionDemo();
    FusionDemo(int n, int nnz, std::string algName);
    FusionDemo(SPM::CSR *csrA, double *correct_x, int n, int nnz, double alpha, std::string algName);
    virtual ~FusionDemo();

    TimeMeasure evaluate();
    double *getSolution() { return x; }
    void set_num_test(int num_test) { this->num_test = num_test; }
    void set_num_thread(int num_thread) { this->num_thread = num_thread; }
    double getAnalysisTime() const { return analysis_time; }
    double getElaspedTime() const { return elasped_time; }
    std::string getName() { return algName; }
    // the last testing() result (the reference only prints it)
    bool passed() const { return passed_; }
    baas_schedule *schedule() const { return gs.sched; }

protected:
    bool passed_ = true;
    void use_plan(int mode, int rpu, int blk, int df) {
        plan_mode = mode, rows_per_unit = rpu, block = blk, dataflow = df;
    }
};

// test/SpTRSVRuntime.hpp:73-96
class SpRTSV_Serial : public FusionDemo {
public:
    SpRTSV_Serial(SPM::CSR *csrA, double *correct_x, int n, int nnz, double alpha, std::string algName);
};

// test/SpTRSVRuntime.hpp:98-145 (row level sets; getWavefront = their count)
class SpTRSV_Wavefront : public FusionDemo {
public:
    SpTRSV_Wavefront(SPM::CSR *csrA, double *correct_x, int n, int nnz, double alpha, std::string algName,
                     int num_thread);
    int getWavefront() const;
};

// test/SpTRSVRuntime.hpp:147-198 (merged levels)
class SpTRSV_WF_Merge : public FusionDemo {
public:
    SpTRSV_WF_Merge(SPM::CSR *csrA, double *correct_x, int n, int nnz, double alpha, std::string algName,
                    int num_thread);
    int getWavefront() const;
};

// test/SpTRSVRuntime.hpp:280-386 (dual tree coarsening + merged levels)
class SpTRSV_DUAL_TREE_BFS_WF_Merge : public FusionDemo {
protected:
    SPM::CSC *cscA = nullptr;

public:
    SpTRSV_DUAL_TREE_BFS_WF_Merge(SPM::CSR *csrA, SPM::CSC *cscA, double *correct_x, int n, int nnz, double alpha,
                                  std::string algName, int num_thread);
    int getWavefront() const;
};

// test/SpTRSVRuntime.hpp:389-614 (point-to-point synchronisation: the
// dataflow lane kernel on units of the merged levels; getWavefront = units)
class SpTRSV_DUAL_TREE_BFS_WF_Merge_P2P : public FusionDemo {
protected:
    SPM::CSC *cscA = nullptr;

public:
    SpTRSV_DUAL_TREE_BFS_WF_Merge_P2P(SPM::CSR *csrA, SPM::CSC *cscA, double *correct_x, int n, int nnz,
                                      double alpha, std::string algName, int num_thread, bool hasPerm);
    int getWavefront() const;
};

// test/SpTRSVRuntime.hpp:617-852, the BAAS runtime (autotuned plan).  With
// hasPerm the reference renumbers rows per thread for locality; the GPU
// plans need no renumbering, so x comes back in the original order either way.
class SpTRSV_DUAL_TREE_BFS_WF_Merge_P2P_MapFourRule : public FusionDemo {
protected:
    SPM::CSC *cscA = nullptr;

public:
    SpTRSV_DUAL_TREE_BFS_WF_Merge_P2P_MapFourRule(SPM::CSR *csrA, SPM::CSC *cscA, double *correct_x, int n,
                                                  int nnz, double alpha, std::string algName, int num_thread,
                                                  bool hasPerm);
    int getWavefront() const;
};

}  // namespace SpTRSVRT

namespace ICRT {

// test/ICRuntime.hpp:18-228: IC0 of an SPD matrix (lu starts as A's
// values), checked against a serial factor on the lower part at 1e-10.
class ICDemo {
protected:
    TimeMeasure timeMeasure;
    int n = 0;
    int lowerNnz = 0;
    int nnz = 0;
    std::vector<int> permToOrig, origToPerm;
    double *lu = nullptr;
    double *correct_lu = nullptr;
    std::string algName;
    int num_test = 10;
    int num_thread = 1;
    double elasped_time = 0;
    double analysis_time = 0;
    SPM::CSR *csrA = nullptr;

    baas_rt::GpuSchedule gs;
    double *d_lu = nullptr;
    int plan_mode = BAAS_PLAN_AUTO, rows_per_unit = 0, block = 0, dataflow = 0;
    bool passed_ = true;

    virtual void setup();
    virtual void buildset();
    virtual TimeMeasure fused_code();
    virtual void testing();

public:
    ICDemo();
    ICDemo(int n, int nnz, std::string algName);
    ICDemo(SPM::CSR *csrA, double *correct_lu, int n, int nnz, std::string algName);
    virtual ~ICDemo();

    TimeMeasure evaluate();
    double *getSolution() { return lu; }
    void set_num_test(int num_test) { this->num_test = num_test; }
    void set_num_thread(int num_thread) { this->num_thread = num_thread; }
    double getAnalysisTime() const { return analysis_time; }
    double getElaspedTime() const { return elasped_time; }
    std::string getAlgName() { return algName; }
    // L (new, lower with the diagonal last) and U = L^T (upper with the
    // diagonal first, into U's arrays when U was allocated as
    // CSR(n, n, nnz of the lower part)) from the factor
    void getLU(SPM::CSR *&L, SPM::CSR *U);
    bool passed() const { return passed_; }
};

// test/ICRuntime.hpp:230-265
class IC_Serial : public ICDemo {
public:
    IC_Serial(SPM::CSR *csrA, double *correct_lu, int n, int nnz, std::string algName);
};

// test/ICRuntime.hpp:267-350
class IC_LevelSet : public ICDemo {
public:
    IC_LevelSet(SPM::CSR *csrA, double *correct_lu, int n, int nnz, std::string algName, int num_thread);
    int getWavefront() const;
};

// test/ICRuntime.hpp:353-583, the BAAS runtime (autotuned plan)
class IC_DUAL_TREE_BFS_WF_Merge_P2P_MapFourRule : public ICDemo {
public:
    IC_DUAL_TREE_BFS_WF_Merge_P2P_MapFourRule(SPM::CSR *csrA, double *correct_lu, int n, int nnz,
                                              std::string algName, int num_thread);
    int getWavefront() const;
};

}  // namespace ICRT

namespace ILURT {

// test/ILUDemo.hpp:12-56: ILU0 (lu starts as A's values), checked against a
// serial factor on every entry at 1e-10 (test/ILUDemo.cpp:40-55).
class ILUDemo {
protected:
    TimeMeasure timeMeasure;
    int n = 0;
    int lowerNnz = 0;
    int nnz = 0;
    std::vector<int> permToOrig, origToPerm;
    double *lu = nullptr;
    double *correct_lu = nullptr;
    std::string algName;
    int num_test = 10;
    int num_thread = 1;
    double elasped_time = 0;
    double analysis_time = 0;
    SPM::CSR *csrA = nullptr;

    baas_rt::GpuSchedule gs;
    double *d_lu = nullptr;
    int plan_mode = BAAS_PLAN_AUTO, rows_per_unit = 0, block = 0, dataflow = 0;
    bool passed_ = true;

    virtual void setup();
    virtual void buildset();
    virtual TimeMeasure fused_code();
    virtual void testing();

public:
    ILUDemo();
    ILUDemo(int n, int nnz, std::string algName);
    ILUDemo(SPM::CSR *csrA, double *correct_lu, int n, int nnz, std::string algName);
    virtual ~ILUDemo();

    TimeMeasure evaluate();
    double *getSolution() { return lu; }
    void set_num_test(int num_test) { this->num_test = num_test; }
    void set_num_thread(int num_thread) { this->num_thread = num_thread; }
    double getAnalysisTime() const { return analysis_time; }
    double getElaspedTime() const { return elasped_time; }
    std::string getAlgName() { return algName; }
    bool passed() const { return passed_; }
};

// test/ILURuntime.hpp:16-51
class ILU_Serial : public ILUDemo {
public:
    ILU_Serial(SPM::CSR *csrA, double *correct_lu, int n, int nnz, std::string algName);
};

// test/ILURuntime.hpp:115-205
class ILU_LevelSet : public ILUDemo {
public:
    ILU_LevelSet(SPM::CSR *csrA, double *correct_lu, int n, int nnz, std::string algName, int num_thread);
    int getWavefront() const;
};

// test/ILURuntime.hpp:207-436, the BAAS runtime (autotuned plan)
class ILU_DUAL_TREE_BFS_WF_Merge_P2P_MapFourRule : public ILUDemo {
public:
    ILU_DUAL_TREE_BFS_WF_Merge_P2P_MapFourRule(SPM::CSR *csrA, double *correct_lu, int n, int nnz,
                                               std::string algName, int num_thread);
    int getWavefront() const;
};

}  // namespace ILURT

namespace Schedule {

// test/Schedule.hpp:31-37.  Every choice maps onto the autotuned GPU plan
// (the reference's variants differ in thread mapping, which ready flags
// replace).
enum ALG_SCHEDULE { P2P, P2P_MAPPING, P2P_TIME, P2P_RULE };

// test/Schedule.hpp:39-353: one schedule for IC0 and the two triangular
// solves of ICCG.  Matrices and vectors passed to run* are host arrays.
// permPreprocessing() (after preprocessing()) sets permToOrigByThread to the
// execution plan's row order -- the GPU counterpart of the reference's
// thread-contiguous order: each work unit's rows become consecutive -- and
// analyses A renumbered by it, A->permute(origToPermByThread.data(),
// permToOrigByThread.data(), true); the Perm entry points take matrices and
// vectors renumbered the same way.
class Scheduler {
    baas_rt::GpuSchedule gs, gsPerm;
    const SPM::CSR *csrA;
    SPM::CSR *csrAPerm = nullptr;
    bool isSymmetric, isTRSVPerm;
    ALG_SCHEDULE alg;
    int num_thread;
    int nnzL = 0;
    double *d_lu = nullptr, *d_lv = nullptr, *d_uv = nullptr, *d_x = nullptr, *d_b = nullptr;
    void ic0_on(baas_rt::GpuSchedule &g, const SPM::CSR *mat, double *lu);
    void trsv_on(baas_rt::GpuSchedule &g, int uplo, const SPM::CSR *T, double *x, const double *b);

public:
    std::vector<int> permToOrigByThread, origToPermByThread;

    Scheduler(const SPM::CSR *csrA, bool isSymmetric, bool isTRSVPerm, ALG_SCHEDULE alg, int num_thread);
    ~Scheduler();
    void preprocessing();
    void permPreprocessing();
    void runIC0(const SPM::CSR *mat, double *lu);  // lu: nnz(A), host
    void runIC0Perm(const SPM::CSR *matPerm, double *lu);
    void runForwardTRSVNoPerm(const SPM::CSR *L, double *x, const double *b);
    void runForwardTRSVPerm(const SPM::CSR *LPerm, double *x, const double *b);
    void runBackwardTRSVNoPerm(const SPM::CSR *U, double *x, const double *b);
    void runBackwardTRSVPerm(const SPM::CSR *UPerm, double *x, const double *b);
    baas_schedule *schedule() const { return gs.sched; }
    baas_schedule *schedulePerm() const { return gsPerm.sched; }
    Scheduler(const Scheduler &) = delete;
    Scheduler &operator=(const Scheduler &) = delete;
};

}  // namespace Schedule
