// baas_runtime.hpp -- C++ host layer with the reference's runtime shapes.
//
// The reference drives every kernel through runtime classes whose
// buildset() runs the analysis and whose fused_code() runs one timed numeric
// call, evaluate() chaining buildset -> setup -> num_test x fused_code ->
// testing (test/FusionDemo.hpp:14-58, test/ICRuntime.hpp:18-228,
// test/ILUDemo.hpp:12-56, test/Schedule.hpp:39-178).  These classes keep
// those shapes and the SPM::CSR / SPM::CSC field names
// (spm/include/CSR.hpp:16-27, spm/include/CSC.hpp:11-40) so callers compile
// unchanged, and do the work on the GPU through include/baas_hip.h.  Host
// arrays in, host arrays out; device mirrors are made once in the
// constructor/setup, and fused_code() times only the device work.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/baas_hip.h"

namespace SPM {

// 0-based CSR with sorted rows; owns its arrays iff ownData.
struct CSR {
    int m = 0, n = 0;
    int *rowptr = nullptr, *colidx = nullptr;
    double *values = nullptr;
    int *diagptr = nullptr;
    bool ownData = false;
    CSR() = default;
    CSR(int m_, int n_, int nnz);                                   // owning, zeroed
    CSR(int m_, int n_, int *rp, int *ci, double *va);             // non-owning
    ~CSR();
    int getNnz() const { return rowptr[m] - rowptr[0]; }
    void constructDiagPtr();                                        // first column >= row
    CSR(const CSR &) = delete;
    CSR &operator=(const CSR &) = delete;
};

struct CSC {
    int m = 0, n = 0;
    int *colptr = nullptr, *rowidx = nullptr;
    double *values = nullptr;
    CSC() = default;
    ~CSC();
    CSC(const CSC &) = delete;
    CSC &operator=(const CSC &) = delete;
};

// CSR -> CSC with ascending row ids per column (spm/CSR.cpp:895-950).
void transpositionToCSC(const CSR *A, CSC *C);

}  // namespace SPM

namespace baas_rt {

// Throws std::runtime_error on any non-OK status.
void check(int status, const char *what);

// One device-resident CSR (pattern + values).
struct DeviceCsr {
    int n = 0, nnz = 0;
    int *rowptr = nullptr, *colidx = nullptr;
    double *values = nullptr;
    void upload(const SPM::CSR &A);
    void release();
    ~DeviceCsr() { release(); }
};

class Engine {  // handle + stream shared by a runtime object
public:
    Engine();
    ~Engine();
    baas_handle *h = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    void start();
    double stop_ms();  // waits, returns elapsed device time
};

}  // namespace baas_rt

namespace SpTRSVRT {

// x = L^-1 b for a lower-triangular CSR; b = row sums so that x = alpha
// (test/FusionDemo.cpp:47-77).
class FusionDemo {
protected:
    SPM::CSR *csrA = nullptr;
    double *x = nullptr, *y = nullptr, *correct_x = nullptr;
    int n = 0, nnz = 0, num_thread = 8, num_test = 100;
    double alpha = 1.0;
    std::string algName;
    double analysis_time = 0.0;
    virtual void buildset() = 0;
    virtual double fused_code() = 0;  // ms of one numeric call
    virtual void setup();
    virtual bool testing();

public:
    FusionDemo(int n, int nnz, std::string name) : n(n), nnz(nnz), algName(std::move(name)) {}
    virtual ~FusionDemo();
    // Runs buildset, setup, num_test solves and testing; returns the mean ms.
    double evaluate(bool *ok = nullptr);
    double getAnalysisTime() const { return analysis_time; }
    const double *solution() const { return x; }
    void setNumTest(int k) { num_test = k; }
};

// The BAAS runtime (SpTRSV_DUAL_TREE_BFS_WF_Merge_P2P_MapFourRule,
// test/SpTRSVRuntime.hpp:617-852) on the GPU.
class SpTRSV_BAAS : public FusionDemo {
    baas_rt::Engine eng;
    baas_rt::DeviceCsr dL;
    baas_schedule *sched = nullptr;
    double *d_b = nullptr, *d_x = nullptr;
    baas_opts opts{};

protected:
    void buildset() override;
    double fused_code() override;
    void setup() override;
    bool testing() override;

public:
    SpTRSV_BAAS(SPM::CSR *csrA, double *correct_x, int n, int nnz, double alpha, std::string algName,
                int num_thread, int plan_mode = BAAS_PLAN_AUTO);
    ~SpTRSV_BAAS() override;
    baas_schedule *schedule() const { return sched; }
};

}  // namespace SpTRSVRT

namespace ICRT {

// IC0 of an SPD matrix; compared to a serial result at 1e-10 on the lower
// part (test/ICRuntime.hpp:95-124).
class IC_BAAS {
    baas_rt::Engine eng;
    baas_rt::DeviceCsr dA;
    baas_schedule *sched = nullptr;
    double *d_lu = nullptr;
    SPM::CSR *csrA;
    double *correct_lu;
    int num_thread;

public:
    std::vector<double> lu;
    double analysis_time = 0.0;
    IC_BAAS(SPM::CSR *csrA, double *correct_lu, int num_thread);
    ~IC_BAAS();
    void buildset();
    double fused_code();
    bool testing();
    double evaluate(int num_test, bool *ok);
};

}  // namespace ICRT

namespace ILURT {

// ILU0; compared to a serial result at 1e-10 on all entries
// (test/ILUDemo.cpp:40-55).
class ILU_BAAS {
    baas_rt::Engine eng;
    baas_rt::DeviceCsr dA;
    baas_schedule *sched = nullptr;
    double *d_lu = nullptr;
    SPM::CSR *csrA;
    double *correct_lu;
    int num_thread;

public:
    std::vector<double> lu;
    double analysis_time = 0.0;
    ILU_BAAS(SPM::CSR *csrA, double *correct_lu, int num_thread);
    ~ILU_BAAS();
    void buildset();
    double fused_code();
    bool testing();
    double evaluate(int num_test, bool *ok);
};

}  // namespace ILURT

namespace Schedule {

// One schedule for IC0 and both triangular solves of ICCG
// (test/Schedule.hpp:39-353).  Pointers passed to run* are host arrays.
class Scheduler {
    baas_rt::Engine eng;
    baas_rt::DeviceCsr dA;
    baas_schedule *sched = nullptr;
    const SPM::CSR *csrA;
    int num_thread;
    int nnzL = 0;
    double *d_lu = nullptr, *d_lv = nullptr, *d_uv = nullptr, *d_x = nullptr, *d_b = nullptr;

public:
    Scheduler(const SPM::CSR *csrA, int num_thread);
    ~Scheduler();
    void preprocessing();
    void runIC0(const SPM::CSR *mat, double *lu);                   // lu: nnz(A), host
    void runForwardTRSVNoPerm(const SPM::CSR *L, double *x, const double *b);
    void runBackwardTRSVNoPerm(const SPM::CSR *U, double *x, const double *b);
    baas_schedule *schedule() const { return sched; }
};

}  // namespace Schedule
