// baas_runtime.cpp -- the reference-shaped host classes on top of the C-ABI.
#include "baas_runtime.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <numeric>
#include <stdexcept>

namespace {

void hip_ok(hipError_t e, const char *what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <class T>
T *dalloc(size_t n) {
    T *p = nullptr;
    hip_ok(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)), "hipMalloc");
    return p;
}

template <class T>
void dfree(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

// Matrix Market coordinate file -> 0-based COO, symmetric and skew-symmetric
// files expanded to both triangles (spm/mmio_highlevel.cpp:748-921).
struct Coo {
    int m = 0, n = 0;
    std::vector<int> r, c;
    std::vector<double> v;
};

Coo read_matrix_market(const char *path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error(std::string("cannot open ") + path);
    std::string line;
    std::getline(f, line);
    std::string lower = line;
    std::transform(lower.begin(), lower.end(), lower.begin(), ::tolower);
    if (lower.rfind("%%matrixmarket", 0) != 0 || lower.find("coordinate") == std::string::npos)
        throw std::runtime_error(std::string(path) + ": only coordinate Matrix Market files are supported");
    if (lower.find("complex") != std::string::npos)
        throw std::runtime_error(std::string(path) + ": complex matrices are not supported");
    const bool pattern = lower.find("pattern") != std::string::npos;
    const bool skew = lower.find("skew-symmetric") != std::string::npos;
    const bool sym = skew || lower.find("symmetric") != std::string::npos || lower.find("hermitian") != std::string::npos;
    while (std::getline(f, line))
        if (!line.empty() && line[0] != '%') break;
    Coo a;
    long long nz = 0;
    {
        std::istringstream hs(line);
        if (!(hs >> a.m >> a.n >> nz)) throw std::runtime_error(std::string(path) + ": bad size line");
    }
    a.r.reserve(sym ? 2 * nz : nz);
    for (long long k = 0; k < nz; k++) {
        long long i, j;
        double v = 1.0;
        if (!(f >> i >> j)) throw std::runtime_error(std::string(path) + ": truncated");
        if (!pattern && !(f >> v)) throw std::runtime_error(std::string(path) + ": truncated");
        const int r = (int)(i - 1), c = (int)(j - 1);
        a.r.push_back(r), a.c.push_back(c), a.v.push_back(v);
        if (sym && r != c) a.r.push_back(c), a.c.push_back(r), a.v.push_back(skew ? -v : v);
    }
    return a;
}

}  // namespace

// ---------------------------------------------------------------- TimeMeasure
void TimeMeasure::start_timer() {
    start_time = std::chrono::steady_clock::now();
    is_measuring = true;
    measure_done = false;
}

void TimeMeasure::measure_elasped_time() {
    if (!is_measuring) throw std::runtime_error("Time measurement is not started. Call start_timer() first.");
    end_time = std::chrono::steady_clock::now();
    elasped_time = std::chrono::duration<double, std::milli>(end_time - start_time).count();
    is_measuring = false;
    measure_done = true;
}

double TimeMeasure::getTime(TimeMetric tm) const {
    if (!measure_done) throw std::runtime_error("No measurement available. Call measure_elasped_time() first.");
    switch (tm) {
        case TimeMetric::SECONDS: return elasped_time / 1000.0;
        case TimeMetric::MILLISECONDS: return elasped_time;
        case TimeMetric::MICROSECONDS: return elasped_time * 1000.0;
    }
    throw std::invalid_argument("Invalid time unit");
}

std::string TimeMeasure::getTimeStr(TimeMetric tm) const {
    std::ostringstream os;
    os << std::fixed << std::setprecision(4) << getTime(tm)
       << (tm == TimeMetric::SECONDS ? " s" : tm == TimeMetric::MILLISECONDS ? " ms" : " us");
    return os.str();
}

// ---------------------------------------------------------------- CSVManager
namespace IO_Utils {

CSVManager::CSVManager(std::string output_name, std::string, std::vector<std::string> column_names,
                       bool force_creation)
    : name_(output_name + ".csv"), cols_(std::move(column_names)), rec_(cols_.size()) {
    bool exists = static_cast<bool>(std::ifstream(name_));
    if (!exists || force_creation) {
        out_.open(name_, std::ios::out | std::ios::trunc);
        if (!out_) {
            std::cout << "Cannot create the file " << name_ << std::endl;
            return;
        }
        for (size_t i = 0; i < cols_.size(); i++) out_ << cols_[i] << (i + 1 < cols_.size() ? "," : "\n");
    } else {
        out_.open(name_, std::ios::out | std::ios::app);
        if (!out_) std::cout << "Cannot open the file " << name_ << std::endl;
    }
}

void CSVManager::set(const std::string &column, const std::string &v) {
    for (size_t i = 0; i < cols_.size(); i++)
        if (cols_[i] == column) {
            rec_[i] = v;
            return;
        }
    throw std::invalid_argument("CSVManager: no column named " + column);
}

void CSVManager::addRecord() {
    if (out_) {
        for (size_t i = 0; i < rec_.size(); i++) out_ << rec_[i] << (i + 1 < rec_.size() ? "," : "\n");
        out_.flush();
    }
    std::fill(rec_.begin(), rec_.end(), std::string());
}

}  // namespace IO_Utils

// ---------------------------------------------------------------- SPM
namespace SPM {

CSC::CSC(int m_, int n_, int nnz) : m(m_), n(n_) {
    colptr = new int[n + 1]();
    rowidx = new int[std::max(nnz, 1)]();
    values = new double[std::max(nnz, 1)]();
}

CSC::~CSC() {
    if (ownData_) {
        delete[] colptr;
        delete[] rowidx;
        delete[] values;
    }
}

void CSR::alloc(int m_, int nnz) {
    m = m_;
    rowptr = new int[m + 1]();
    colidx = new int[std::max(nnz, 1)]();
    values = new double[std::max(nnz, 1)]();
    ownData_ = true;
}

CSR::CSR(int m_, int n_, int nnz) : n(n_) { alloc(m_, nnz); }

CSR::CSR(int m_, int n_, int *rp, int *ci, double *va) : m(m_), n(n_), rowptr(rp), colidx(ci), values(va) {}

CSR::CSR(const CSR &A) : n(A.n) {
    const int nnz = A.getNnz();
    alloc(A.m, nnz);
    std::copy(A.rowptr, A.rowptr + A.m + 1, rowptr);
    std::copy(A.colidx, A.colidx + nnz, colidx);
    std::copy(A.values, A.values + nnz, values);
    if (A.diagptr) {
        diagptr = new int[m];
        std::copy(A.diagptr, A.diagptr + m, diagptr);
    }
    if (A.diag) {
        diag = new double[m];
        std::copy(A.diag, A.diag + m, diag);
    }
    if (A.idiag) {
        idiag = new double[m];
        std::copy(A.idiag, A.idiag + m, idiag);
    }
}

// COO -> CSR with rows sorted by column (spm/COO.cpp:263-320); diagptr,
// diag and idiag from the exact diagonal entries
static void fill_from_coo(CSR &A, const Coo &a, int base) {
    if (base != 0) throw std::invalid_argument("SPM::CSR: only 0-based output is supported");
    const size_t nz = a.r.size();
    std::vector<size_t> ord(nz);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](size_t p, size_t q) {
        return a.r[p] != a.r[q] ? a.r[p] < a.r[q] : a.c[p] < a.c[q];
    });
    for (size_t t = 0; t < nz; t++) {
        A.rowptr[a.r[ord[t]] + 1]++;
        A.colidx[t] = a.c[ord[t]];
        A.values[t] = a.v[ord[t]];
    }
    for (int i = 0; i < A.m; i++) A.rowptr[i + 1] += A.rowptr[i];
    A.diagptr = new int[A.m];
    A.diag = new double[A.m]();
    A.idiag = new double[A.m]();
    for (int i = 0; i < A.m; i++) {
        A.diagptr[i] = A.rowptr[i + 1];
        for (int j = A.rowptr[i]; j < A.rowptr[i + 1]; j++)
            if (A.colidx[j] >= i) {
                A.diagptr[i] = j;
                break;
            }
        for (int j = A.rowptr[i]; j < A.rowptr[i + 1]; j++)
            if (A.colidx[j] == i) {
                A.diagptr[i] = j;
                A.diag[i] = A.values[j];
                A.idiag[i] = 1.0 / A.values[j];
            }
    }
}

CSR::CSR(const char *fileName, int base, bool forceSymmetric, int) {
    const size_t l = strlen(fileName);
    if (l > 4 && !strcmp(fileName + l - 4, ".bin")) {
        if (forceSymmetric) throw std::invalid_argument("SPM::CSR: forceSymmetric with a .bin file");
        loadBin(fileName, base);
        return;
    }
    Coo a = read_matrix_market(fileName);
    n = a.n;
    alloc(a.m, (int)a.r.size());
    fill_from_coo(*this, a, base);
}

CSR::CSR(const char *fileName, bool isLower, int base) {
    Coo a = read_matrix_market(fileName);
    if (a.m != a.n) throw std::invalid_argument("non-square matrix can not split lower matrix");
    Coo lo;
    lo.m = a.m, lo.n = a.n;
    for (size_t k = 0; k < a.r.size(); k++) {
        const bool keep = isLower ? a.c[k] < a.r[k] : a.c[k] > a.r[k];
        if (keep) lo.r.push_back(a.r[k]), lo.c.push_back(a.c[k]), lo.v.push_back(a.v[k]);
    }
    for (int i = 0; i < a.m; i++) lo.r.push_back(i), lo.c.push_back(i), lo.v.push_back(1.0);
    printf("m:%d, n:%d, nnz:%d\n", lo.m, lo.n, (int)lo.r.size());
    n = lo.n;
    alloc(lo.m, (int)lo.r.size());
    fill_from_coo(*this, lo, base);
}

CSR::~CSR() {
    if (ownData_) {
        delete[] rowptr;
        delete[] colidx;
        delete[] values;
    }
    delete[] diagptr;
    delete[] diag;
    delete[] idiag;
}

CSR *CSR::permute(const int *columnPerm, const int *rowInversePerm, bool sort) const {
    const int nnz = getNnz();
    CSR *r = new CSR(m, n, nnz);
    for (int i = 0; i < m; i++) {
        const int row = rowInversePerm ? rowInversePerm[i] : i;
        r->rowptr[i + 1] = r->rowptr[i] + (rowptr[row + 1] - rowptr[row]);
    }
    std::vector<std::pair<int, double>> ent;
    for (int i = 0; i < m; i++) {
        const int row = rowInversePerm ? rowInversePerm[i] : i;
        const int b = rowptr[row], e = rowptr[row + 1], o = r->rowptr[i];
        ent.clear();
        for (int j = b; j < e; j++) ent.emplace_back(columnPerm[colidx[j]], values[j]);
        if (sort)
            std::stable_sort(ent.begin(), ent.end(),
                             [](const std::pair<int, double> &x, const std::pair<int, double> &y) {
                                 return x.first < y.first;
                             });
        for (int k = 0; k < e - b; k++) {
            r->colidx[o + k] = ent[k].first;
            r->values[o + k] = ent[k].second;
        }
        if (diagptr) {
            if (!r->diagptr) r->diagptr = new int[m];
            // the renamed diagonal column of the original row
            const int dcol = diagptr[row] < e && colidx[diagptr[row]] == row ? columnPerm[row] : -1;
            r->diagptr[i] = r->rowptr[i + 1];
            for (int k = 0; k < e - b; k++)
                if (r->colidx[o + k] == dcol) {
                    r->diagptr[i] = o + k;
                    break;
                }
        }
    }
    if (diag) {
        r->diag = new double[m];
        for (int i = 0; i < m; i++) r->diag[i] = diag[rowInversePerm ? rowInversePerm[i] : i];
    }
    if (idiag) {
        r->idiag = new double[m];
        for (int i = 0; i < m; i++) r->idiag[i] = idiag[rowInversePerm ? rowInversePerm[i] : i];
    }
    return r;
}

static const int MAT_FILE_CLASSID = 1211216;

void CSR::loadBin(const char *fileName, int base) {
    if (base != 0) throw std::invalid_argument("SPM::CSR::loadBin: only 0-based matrices");
    FILE *fp = fopen(fileName, "rb");
    if (!fp) throw std::runtime_error(std::string("SPM::CSR::loadBin: cannot open ") + fileName);
    int hdr[4];
    if (fread(hdr, sizeof(int), 4, fp) != 4 || hdr[0] != MAT_FILE_CLASSID || hdr[1] < 0 || hdr[2] < 0 ||
        hdr[3] < 0) {
        fclose(fp);
        throw std::runtime_error(std::string("SPM::CSR::loadBin: not a matrix file: ") + fileName);
    }
    if (ownData_) {
        delete[] rowptr;
        delete[] colidx;
        delete[] values;
    }
    delete[] diagptr;
    delete[] diag;
    delete[] idiag;
    diagptr = nullptr;
    diag = idiag = nullptr;
    n = hdr[2];
    alloc(hdr[1], hdr[3]);
    const bool ok = fread(rowptr + 1, sizeof(int), (size_t)m, fp) == (size_t)m &&
                    fread(colidx, sizeof(int), (size_t)hdr[3], fp) == (size_t)hdr[3] &&
                    fread(values, sizeof(double), (size_t)hdr[3], fp) == (size_t)hdr[3];
    fclose(fp);
    if (!ok) throw std::runtime_error(std::string("SPM::CSR::loadBin: truncated file ") + fileName);
    rowptr[0] = 0;
    for (int i = 0; i < m; i++) rowptr[i + 1] += rowptr[i];
    if (rowptr[m] != hdr[3]) throw std::runtime_error("SPM::CSR::loadBin: row lengths do not add up to nnz");
    // the exact diagonal of each row (rows without one keep diagptr = -1)
    diagptr = new int[m];
    diag = new double[m]();
    idiag = new double[m]();
    for (int i = 0; i < m; i++) {
        diagptr[i] = -1;
        for (int j = rowptr[i]; j < rowptr[i + 1]; j++)
            if (colidx[j] == i) {
                diagptr[i] = j;
                diag[i] = values[j];
                idiag[i] = 1 / values[j];
            }
    }
}

void CSR::storeBin(const char *fileName) {
    FILE *fp = fopen(fileName, "wb");
    if (!fp) throw std::runtime_error(std::string("SPM::CSR::storeBin: cannot open ") + fileName);
    const int nnz = getNnz();
    const int hdr[4] = {MAT_FILE_CLASSID, m, n, nnz};
    std::vector<int> len(m);
    for (int i = 0; i < m; i++) len[i] = rowptr[i + 1] - rowptr[i];
    const bool ok = fwrite(hdr, sizeof(int), 4, fp) == 4 && fwrite(len.data(), sizeof(int), m, fp) == (size_t)m &&
                    fwrite(colidx + rowptr[0], sizeof(int), nnz, fp) == (size_t)nnz &&
                    fwrite(values + rowptr[0], sizeof(double), nnz, fp) == (size_t)nnz;
    if (fclose(fp) != 0 || !ok) throw std::runtime_error(std::string("SPM::CSR::storeBin: write failed ") + fileName);
}

void CSR::constructDiagPtr() {
    if (diagptr) return;
    diagptr = new int[m];
    for (int i = 0; i < m; i++) {
        diagptr[i] = rowptr[i + 1];
        for (int j = rowptr[i]; j < rowptr[i + 1]; j++)
            if (colidx[j] >= i) {
                diagptr[i] = j;
                break;
            }
    }
}

void CSR::computeInverseDiag() {
    constructDiagPtr();
    if (!idiag) idiag = new double[m];
    if (!diag) diag = new double[m];
    for (int i = 0; i < m; i++) {
        diag[i] = values[diagptr[i]];
        idiag[i] = 1.0 / diag[i];
    }
}

void CSR::transpositionToCSC(const CSR *A, CSC *C) {
    const int nnz = A->getNnz();
    if (C->ownData_) {
        delete[] C->colptr;
        delete[] C->rowidx;
        delete[] C->values;
    }
    C->colptr = new int[A->n + 1]();
    C->rowidx = new int[std::max(nnz, 1)];
    C->values = new double[std::max(nnz, 1)];
    C->ownData_ = true;
    C->m = A->m;
    C->n = A->n;
    std::fill(C->colptr, C->colptr + A->n + 1, 0);
    for (int j = 0; j < nnz; j++) C->colptr[A->colidx[j] + 1]++;
    for (int c = 0; c < A->n; c++) C->colptr[c + 1] += C->colptr[c];
    std::vector<int> fill(C->colptr, C->colptr + A->n);
    for (int i = 0; i < A->m; i++)
        for (int j = A->rowptr[i]; j < A->rowptr[i + 1]; j++) {
            const int p = fill[A->colidx[j]]++;
            C->rowidx[p] = i;
            C->values[p] = A->values[j];
        }
}

}  // namespace SPM

// ---------------------------------------------------------------- baas_rt
namespace baas_rt {

void check(int st, const char *what) {
    if (st != BAAS_OK) throw std::runtime_error(std::string(what) + ": " + baas_status_string(st));
}

void DeviceCsr::upload(const SPM::CSR &A) {
    release();
    n = A.m;
    nnz = A.getNnz();
    rowptr = dalloc<int>(n + 1);
    colidx = dalloc<int>(nnz);
    values = dalloc<double>(nnz);
    hip_ok(hipMemcpy(rowptr, A.rowptr, sizeof(int) * (n + 1), hipMemcpyHostToDevice), "upload rowptr");
    hip_ok(hipMemcpy(colidx, A.colidx, sizeof(int) * nnz, hipMemcpyHostToDevice), "upload colidx");
    hip_ok(hipMemcpy(values, A.values, sizeof(double) * nnz, hipMemcpyHostToDevice), "upload values");
}

void DeviceCsr::release() {
    dfree(rowptr);
    dfree(colidx);
    dfree(values);
}

Engine::Engine() {
    hip_ok(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "stream");
    hip_ok(hipEventCreate(&e0), "event");
    hip_ok(hipEventCreate(&e1), "event");
    check(baas_create(&h, stream), "baas_create");
}

Engine::~Engine() {
    if (h) baas_destroy(h);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (stream) (void)hipStreamDestroy(stream);
}

void Engine::start() { hip_ok(hipEventRecord(e0, stream), "event record"); }

double Engine::stop_ms() {
    hip_ok(hipEventRecord(e1, stream), "event record");
    hip_ok(hipEventSynchronize(e1), "event sync");
    float ms = 0;
    hip_ok(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    return ms;
}

void GpuSchedule::analyze(const SPM::CSR &A, int kind_, int ncore, int plan_mode, int rpu, int block, int dataflow) {
    kind = kind_;
    dA.upload(A);
    if (sched) baas_schedule_destroy(sched);
    sched = nullptr;
    baas_opts o{};
    o.ncore = ncore > 0 ? ncore : 1;
    o.plan_mode = plan_mode;
    o.rows_per_unit = rpu;
    o.block_threads = block;
    check(baas_analyze(eng.h, A.m, dA.rowptr, dA.colidx, kind, &o, &sched), "baas_analyze");
    if (plan_mode == BAAS_PLAN_AUTO) {
        // time the execution plans on these values and keep the fastest
        if (kind == BAAS_KIND_TRSV) check(baas_autotune(eng.h, sched, dA.values, 0, 1, nullptr), "baas_autotune");
        else check(baas_autotune_factor(eng.h, sched, dA.values, nullptr), "baas_autotune_factor");
    } else {
        check(baas_plan_dataflow(sched, dataflow), "baas_plan_dataflow");
    }
}

int GpuSchedule::info(int which) const {
    int c[8] = {0};
    if (sched) check(baas_schedule_info(sched, c), "baas_schedule_info");
    return c[which];
}

GpuSchedule::~GpuSchedule() {
    if (sched) baas_schedule_destroy(sched);
}

}  // namespace baas_rt

// ---------------------------------------------------------------- Merge
namespace Merge {

void spic0_csr_uL_serial(const SPM::CSR *A, double *lu) {
    baas_rt::GpuSchedule g;
    g.analyze(*A, BAAS_KIND_IC0, 1, BAAS_PLAN_RANGE, 0, 256, 0);
    double *d_lu = dalloc<double>(g.dA.nnz);
    baas_rt::check(baas_ic0_f64(g.eng.h, g.sched, g.dA.values, d_lu), "baas_ic0_f64");
    const int st = baas_check(g.eng.h);
    if (st != BAAS_OK && st != BAAS_NONPOSITIVE_PIVOT) baas_rt::check(st, "ic0 kernel");
    hip_ok(hipMemcpy(lu, d_lu, sizeof(double) * g.dA.nnz, hipMemcpyDeviceToHost), "download lu");
    dfree(d_lu);
}

void ilu0csr_uplooking_ref(SPM::CSR *A, double *lu) {
    baas_rt::GpuSchedule g;
    g.analyze(*A, BAAS_KIND_ILU0, 1, BAAS_PLAN_RANGE, 0, 256, 0);
    double *d_lu = dalloc<double>(g.dA.nnz);
    baas_rt::check(baas_ilu0_f64(g.eng.h, g.sched, g.dA.values, d_lu), "baas_ilu0_f64");
    baas_rt::check(baas_check(g.eng.h), "ilu0 kernel");
    hip_ok(hipMemcpy(lu, d_lu, sizeof(double) * g.dA.nnz, hipMemcpyDeviceToHost), "download lu");
    dfree(d_lu);
}

void sptrsv_serial_csr(const SPM::CSR &A, double x[], const double y[]) {
    baas_rt::GpuSchedule g;
    g.analyze(A, BAAS_KIND_TRSV, 1, BAAS_PLAN_RANGE, 0, 256, 0);
    double *d_b = dalloc<double>(A.m), *d_x = dalloc<double>(A.m);
    hip_ok(hipMemcpy(d_b, y, sizeof(double) * A.m, hipMemcpyHostToDevice), "upload b");
    baas_rt::check(baas_sptrsv_f64(g.eng.h, g.sched, 0, g.dA.values, d_b, d_x, 1), "baas_sptrsv_f64");
    baas_rt::check(baas_check(g.eng.h), "sptrsv kernel");
    hip_ok(hipMemcpy(x, d_x, sizeof(double) * A.m, hipMemcpyDeviceToHost), "download x");
    dfree(d_b);
    dfree(d_x);
}

}  // namespace Merge

// ---------------------------------------------------------------- SpTRSVRT
namespace SpTRSVRT {

FusionDemo::FusionDemo() {}

FusionDemo::FusionDemo(int n_, int nnz_, std::string name) : n(n_), nnz(nnz_), algName(std::move(name)) {
    x = MALLOC(double, n);
    y = MALLOC(double, n);
    CHECK_POINTER(x);
    CHECK_POINTER(y);
}

FusionDemo::FusionDemo(SPM::CSR *A, double *cx, int n_, int nnz_, double alpha_, std::string name)
    : FusionDemo(n_, nnz_, std::move(name)) {
    csrA = A;
    correct_x = cx;
    alpha = alpha_;
}

FusionDemo::~FusionDemo() {
    delete csrA_perm;
    FREE(x);
    FREE(y);
    dfree(d_b);
    dfree(d_x);
}

// y_i = sum_j L_ij * alpha, x := 0 (test/FusionDemo.cpp:47-65); a Perm
// variant solves on y renumbered like its matrix (test/FusionDemo.cpp:96-112)
void FusionDemo::setup() {
    for (int i = 0; i < n; i++) {
        double s = 0.0;
        for (int j = csrA->rowptr[i]; j < csrA->rowptr[i + 1]; j++) s = std::fma(csrA->values[j], alpha, s);
        y[i] = s;
        x[i] = 0.0;
    }
    if (!d_b) d_b = dalloc<double>(n);
    if (!d_x) d_x = dalloc<double>(n);
    if (hasPerm) {
        std::vector<double> yp(n);
        for (int i = 0; i < n; i++) yp[i] = y[permToOrig[i]];
        hip_ok(hipMemcpy(d_b, yp.data(), sizeof(double) * n, hipMemcpyHostToDevice), "upload b");
    } else {
        hip_ok(hipMemcpy(d_b, y, sizeof(double) * n, hipMemcpyHostToDevice), "upload b");
    }
    hip_ok(hipMemset(d_x, 0, sizeof(double) * n), "clear x");
}

// With hasPerm the matrix is renumbered by the first analysis's execution
// order (every work unit's rows consecutive: the GPU form of the
// reference's thread-contiguous groupPerm, test/SpTRSVRuntime.hpp:751-772)
// and analysed again; both analyses count as analysis time.
void FusionDemo::buildset() {
    gs.analyze(*csrA, BAAS_KIND_TRSV, num_thread, plan_mode, rows_per_unit, block, dataflow);
    if (!hasPerm) return;
    long long len = 0;
    baas_rt::check(baas_schedule_export(gs.sched, BAAS_EXPORT_UNIT_ROWS, nullptr, &len), "export");
    permToOrig.assign((size_t)len, 0);
    baas_rt::check(baas_schedule_export(gs.sched, BAAS_EXPORT_UNIT_ROWS, permToOrig.data(), &len), "export");
    origToPerm.assign((size_t)n, 0);
    for (int i = 0; i < n; i++) origToPerm[permToOrig[i]] = i;
    delete csrA_perm;
    csrA_perm = csrA->permute(origToPerm.data(), permToOrig.data(), true);
    gs.analyze(*csrA_perm, BAAS_KIND_TRSV, num_thread, plan_mode, rows_per_unit, block, dataflow);
}

TimeMeasure FusionDemo::fused_code() {
    TimeMeasure t;
    gs.eng.start();
    baas_rt::check(baas_sptrsv_f64(gs.eng.h, gs.sched, 0, gs.dA.values, d_b, d_x, 1), "baas_sptrsv_f64");
    t.elasped_time = gs.eng.stop_ms();
    return t;
}

// |x - correct_x| <= 1e-10 everywhere (test/FusionDemo.cpp:66-77)
void FusionDemo::testing() {
    baas_rt::check(baas_check(gs.eng.h), "sptrsv kernel");
    hip_ok(hipMemcpy(x, d_x, sizeof(double) * n, hipMemcpyDeviceToHost), "download x");
    if (hasPerm) {  // back to the original numbering (test/FusionDemo.cpp:127-137)
        std::vector<double> xp(x, x + n);
        for (int i = 0; i < n; i++) x[permToOrig[i]] = xp[i];
    }
    passed_ = true;
    if (correct_x == nullptr) return;
    for (int i = 0; i < n; i++)
        if (std::fabs(x[i] - correct_x[i]) > 1e-10) {
            printf("%d, first is %.16f, sencond is %.16f,  error is %.16f\n", i, x[i], correct_x[i],
                   std::fabs(x[i] - correct_x[i]));
            passed_ = false;
            break;
        }
    std::cout << algName + (passed_ ? " code solutions check pass.\n" : " code solutions != reference solutions.\n");
}

TimeMeasure FusionDemo::evaluate() {
    TimeMeasure avg, an;
    an.start_timer();
    buildset();
    an.measure_elasped_time();
    analysis_time = an.getTime();
    setup();
    double total = 0.0;
    for (int i = 0; i < num_test; i++) total += fused_code().elasped_time;
    testing();
    avg.elasped_time = num_test > 0 ? total / num_test : 0.0;
    elasped_time = avg.elasped_time;
    return avg;
}

SpRTSV_Serial::SpRTSV_Serial(SPM::CSR *A, double *cx, int n_, int nnz_, double a, std::string name)
    : FusionDemo(A, cx, n_, nnz_, a, std::move(name)) {
    use_plan(BAAS_PLAN_RANGE, 0, 256, 0);
}

SpTRSV_Wavefront::SpTRSV_Wavefront(SPM::CSR *A, double *cx, int n_, int nnz_, double a, std::string name, int nt)
    : FusionDemo(A, cx, n_, nnz_, a, std::move(name)) {
    num_thread = nt;
    use_plan(BAAS_PLAN_LEVELS, 0, 256, 0);
}
int SpTRSV_Wavefront::getWavefront() const { return gs.info(4); }

SpTRSV_WF_Merge::SpTRSV_WF_Merge(SPM::CSR *A, double *cx, int n_, int nnz_, double a, std::string name, int nt)
    : FusionDemo(A, cx, n_, nnz_, a, std::move(name)) {
    num_thread = nt;
    use_plan(BAAS_PLAN_SCHEDULE, 0, 256, 0);
}
int SpTRSV_WF_Merge::getWavefront() const { return gs.info(2); }

SpTRSV_DUAL_TREE_BFS_WF_Merge::SpTRSV_DUAL_TREE_BFS_WF_Merge(SPM::CSR *A, SPM::CSC *C, double *cx, int n_, int nnz_,
                                                             double a, std::string name, int nt)
    : FusionDemo(A, cx, n_, nnz_, a, std::move(name)), cscA(C) {
    num_thread = nt;
    use_plan(BAAS_PLAN_SCHEDULE, 0, 256, 0);
}
int SpTRSV_DUAL_TREE_BFS_WF_Merge::getWavefront() const { return gs.info(2); }

SpTRSV_DUAL_TREE_BFS_WF_Merge_P2P::SpTRSV_DUAL_TREE_BFS_WF_Merge_P2P(SPM::CSR *A, SPM::CSC *C, double *cx, int n_,
                                                                     int nnz_, double a, std::string name, int nt,
                                                                     bool perm)
    : FusionDemo(A, cx, n_, nnz_, a, std::move(name)), cscA(C) {
    num_thread = nt;
    hasPerm = perm;
    use_plan(BAAS_PLAN_SCHEDULE, 64, 1024, 2);
}
int SpTRSV_DUAL_TREE_BFS_WF_Merge_P2P::getWavefront() const { return gs.info(5); }

SpTRSV_DUAL_TREE_BFS_WF_Merge_P2P_MapFourRule::SpTRSV_DUAL_TREE_BFS_WF_Merge_P2P_MapFourRule(
    SPM::CSR *A, SPM::CSC *C, double *cx, int n_, int nnz_, double a, std::string name, int nt, bool perm)
    : FusionDemo(A, cx, n_, nnz_, a, std::move(name)), cscA(C) {
    num_thread = nt;
    hasPerm = perm;
    use_plan(BAAS_PLAN_AUTO, 0, 0, 0);
}
int SpTRSV_DUAL_TREE_BFS_WF_Merge_P2P_MapFourRule::getWavefront() const { return gs.info(5); }

}  // namespace SpTRSVRT

// ---------------------------------------------------------------- ICRT
namespace ICRT {

ICDemo::ICDemo() {}

ICDemo::ICDemo(int n_, int nnz_, std::string name) : n(n_), nnz(nnz_), algName(std::move(name)) {
    lu = MALLOC(double, nnz);
    CHECK_POINTER(lu);
}

ICDemo::ICDemo(SPM::CSR *A, double *cl, int n_, int nnz_, std::string name) : ICDemo(n_, nnz_, std::move(name)) {
    csrA = A;
    correct_lu = cl;
}

ICDemo::~ICDemo() {
    FREE(lu);
    dfree(d_lu);
}

void ICDemo::setup() {}

void ICDemo::buildset() {
    csrA->constructDiagPtr();
    gs.analyze(*csrA, BAAS_KIND_IC0, num_thread, plan_mode, rows_per_unit, block, dataflow);
    lowerNnz = gs.info(6);
    if (!d_lu) d_lu = dalloc<double>(nnz);
}

// lu := A's values, then IC0 (test/ICRuntime.hpp:533-553)
TimeMeasure ICDemo::fused_code() {
    TimeMeasure t;
    gs.eng.start();
    baas_rt::check(baas_ic0_f64(gs.eng.h, gs.sched, gs.dA.values, d_lu), "baas_ic0_f64");
    t.elasped_time = gs.eng.stop_ms();
    return t;
}

// lower part with the diagonal within 1e-10 (test/ICRuntime.hpp:95-124)
void ICDemo::testing() {
    const int st = baas_check(gs.eng.h);
    if (st != BAAS_OK && st != BAAS_NONPOSITIVE_PIVOT) baas_rt::check(st, "ic0 kernel");
    hip_ok(hipMemcpy(lu, d_lu, sizeof(double) * nnz, hipMemcpyDeviceToHost), "download lu");
    passed_ = true;
    if (correct_lu == nullptr) return;
    std::cout << "=============== result check ===================\n";
    for (int i = 0; i < n && passed_; i++)
        for (int k = csrA->rowptr[i]; k <= csrA->diagptr[i]; k++)
            if (std::fabs(correct_lu[k] - lu[k]) > 1.0e-10) {
                passed_ = false;
                break;
            }
    std::cout << algName + (passed_ ? " code IC0 check pass.\n" : " code IC0 != reference factorization.\n");
    fflush(stdout);
}

TimeMeasure ICDemo::evaluate() {
    std::cout << "################ IC0 running in " << algName << " method #################" << std::endl;
    TimeMeasure avg, an;
    an.start_timer();
    buildset();
    an.measure_elasped_time();
    analysis_time = an.getTime();
    setup();
    double total = 0.0;
    for (int i = 0; i < num_test; i++) total += fused_code().elasped_time;
    testing();
    avg.elasped_time = num_test > 0 ? total / num_test : 0.0;
    elasped_time = avg.elasped_time;
    std::cout << "################ IC0 ending in " << algName << " method #################" << std::endl;
    return avg;
}

void ICDemo::getLU(SPM::CSR *&L, SPM::CSR *U) {
    csrA->constructDiagPtr();
    int nl = 0;
    for (int i = 0; i < n; i++) nl += csrA->diagptr[i] - csrA->rowptr[i] + 1;
    L = new SPM::CSR(n, n, nl);
    std::vector<int> up(n + 1, 0);
    for (int i = 0; i < n; i++) {
        L->rowptr[i + 1] = L->rowptr[i] + csrA->diagptr[i] - csrA->rowptr[i] + 1;
        up[i + 1] = up[i] + csrA->rowptr[i + 1] - csrA->diagptr[i];
    }
    for (int i = 0; i < n; i++)
        for (int j = csrA->rowptr[i]; j <= csrA->diagptr[i]; j++) {
            L->colidx[L->rowptr[i] + j - csrA->rowptr[i]] = csrA->colidx[j];
            L->values[L->rowptr[i] + j - csrA->rowptr[i]] = lu[j];
        }
    if (U == nullptr) return;
    // U = L^T: row c of U holds column c of L, the diagonal first
    std::vector<int> fill(up.begin(), up.end() - 1);
    std::copy(up.begin(), up.end(), U->rowptr);
    for (int i = 0; i < n; i++)
        for (int j = csrA->rowptr[i]; j <= csrA->diagptr[i]; j++) {
            const int c = csrA->colidx[j], p = fill[c]++;
            U->colidx[p] = i;
            U->values[p] = lu[j];
        }
}

IC_Serial::IC_Serial(SPM::CSR *A, double *cl, int n_, int nnz_, std::string name)
    : ICDemo(A, cl, n_, nnz_, std::move(name)) {
    plan_mode = BAAS_PLAN_RANGE, block = 256;
}

IC_LevelSet::IC_LevelSet(SPM::CSR *A, double *cl, int n_, int nnz_, std::string name, int nt)
    : ICDemo(A, cl, n_, nnz_, std::move(name)) {
    num_thread = nt;
    plan_mode = BAAS_PLAN_LEVELS, block = 256;
}
int IC_LevelSet::getWavefront() const { return gs.info(4); }

IC_DUAL_TREE_BFS_WF_Merge_P2P_MapFourRule::IC_DUAL_TREE_BFS_WF_Merge_P2P_MapFourRule(SPM::CSR *A, double *cl, int n_,
                                                                                     int nnz_, std::string name,
                                                                                     int nt)
    : ICDemo(A, cl, n_, nnz_, std::move(name)) {
    num_thread = nt;
    plan_mode = BAAS_PLAN_AUTO;
}
int IC_DUAL_TREE_BFS_WF_Merge_P2P_MapFourRule::getWavefront() const { return gs.info(5); }

}  // namespace ICRT

// ---------------------------------------------------------------- ILURT
namespace ILURT {

ILUDemo::ILUDemo() {}

ILUDemo::ILUDemo(int n_, int nnz_, std::string name) : n(n_), nnz(nnz_), algName(std::move(name)) {
    lu = MALLOC(double, nnz);
    CHECK_POINTER(lu);
}

ILUDemo::ILUDemo(SPM::CSR *A, double *cl, int n_, int nnz_, std::string name) : ILUDemo(n_, nnz_, std::move(name)) {
    csrA = A;
    correct_lu = cl;
}

ILUDemo::~ILUDemo() {
    FREE(lu);
    dfree(d_lu);
}

void ILUDemo::setup() {}

void ILUDemo::buildset() {
    csrA->constructDiagPtr();
    gs.analyze(*csrA, BAAS_KIND_ILU0, num_thread, plan_mode, rows_per_unit, block, dataflow);
    lowerNnz = gs.info(6);
    if (!d_lu) d_lu = dalloc<double>(nnz);
}

TimeMeasure ILUDemo::fused_code() {
    TimeMeasure t;
    gs.eng.start();
    baas_rt::check(baas_ilu0_f64(gs.eng.h, gs.sched, gs.dA.values, d_lu), "baas_ilu0_f64");
    t.elasped_time = gs.eng.stop_ms();
    return t;
}

// every entry within 1e-10 (test/ILUDemo.cpp:40-55)
void ILUDemo::testing() {
    baas_rt::check(baas_check(gs.eng.h), "ilu0 kernel");
    hip_ok(hipMemcpy(lu, d_lu, sizeof(double) * nnz, hipMemcpyDeviceToHost), "download lu");
    passed_ = true;
    if (correct_lu == nullptr) return;
    std::cout << "=============== result check ===================\n";
    for (int k = 0; k < nnz; k++)
        if (std::fabs(correct_lu[k] - lu[k]) > 1.0e-10) {
            printf("%d, first is %.16f, sencond is %.16f,  error is %.16f\n", k, correct_lu[k], lu[k],
                   std::fabs(correct_lu[k] - lu[k]));
            passed_ = false;
            break;
        }
    std::cout << algName + (passed_ ? " code ILU0 check pass.\n" : " code ILU0 != reference factorization.\n");
    fflush(stdout);
}

TimeMeasure ILUDemo::evaluate() {
    std::cout << "################ ILU0 running in " << algName << " method #################" << std::endl;
    TimeMeasure avg, an;
    an.start_timer();
    buildset();
    an.measure_elasped_time();
    analysis_time = an.getTime();
    setup();
    double total = 0.0;
    for (int i = 0; i < num_test; i++) total += fused_code().elasped_time;
    testing();
    avg.elasped_time = num_test > 0 ? total / num_test : 0.0;
    elasped_time = avg.elasped_time;
    std::cout << "################ ILU0 ending in " << algName << " method #################" << std::endl;
    return avg;
}

ILU_Serial::ILU_Serial(SPM::CSR *A, double *cl, int n_, int nnz_, std::string name)
    : ILUDemo(A, cl, n_, nnz_, std::move(name)) {
    plan_mode = BAAS_PLAN_RANGE, block = 256;
}

ILU_LevelSet::ILU_LevelSet(SPM::CSR *A, double *cl, int n_, int nnz_, std::string name, int nt)
    : ILUDemo(A, cl, n_, nnz_, std::move(name)) {
    num_thread = nt;
    plan_mode = BAAS_PLAN_LEVELS, block = 256;
}
int ILU_LevelSet::getWavefront() const { return gs.info(4); }

ILU_DUAL_TREE_BFS_WF_Merge_P2P_MapFourRule::ILU_DUAL_TREE_BFS_WF_Merge_P2P_MapFourRule(SPM::CSR *A, double *cl,
                                                                                       int n_, int nnz_,
                                                                                       std::string name, int nt)
    : ILUDemo(A, cl, n_, nnz_, std::move(name)) {
    num_thread = nt;
    plan_mode = BAAS_PLAN_AUTO;
}
int ILU_DUAL_TREE_BFS_WF_Merge_P2P_MapFourRule::getWavefront() const { return gs.info(5); }

}  // namespace ILURT

// ---------------------------------------------------------------- Schedule
namespace Schedule {

Scheduler::Scheduler(const SPM::CSR *A, bool sym, bool perm, ALG_SCHEDULE a, int nt)
    : csrA(A), isSymmetric(sym), isTRSVPerm(perm), alg(a), num_thread(nt) {
    if (!A->diagptr) throw std::invalid_argument("Scheduler: the matrix needs diagptr (constructDiagPtr())");
}

Scheduler::~Scheduler() {
    for (double **p : {&d_lu, &d_lv, &d_uv, &d_x, &d_b}) dfree(*p);
    delete csrAPerm;
}

// the IC0 schedule of A's lower part serves IC0 and both sweeps
// (test/Schedule.hpp:181-296); every ALG_SCHEDULE takes the autotuned plan
void Scheduler::preprocessing() {
    gs.analyze(*csrA, BAAS_KIND_IC0, num_thread, BAAS_PLAN_AUTO, 0, 0, 0);
    nnzL = gs.info(6);
    d_lu = dalloc<double>(gs.dA.nnz);
    d_lv = dalloc<double>(nnzL);
    d_uv = dalloc<double>(nnzL);
    d_x = dalloc<double>(gs.dA.n);
    d_b = dalloc<double>(gs.dA.n);
    permToOrigByThread.resize(csrA->m);
    std::iota(permToOrigByThread.begin(), permToOrigByThread.end(), 0);
    origToPermByThread = permToOrigByThread;
}

// test/Schedule.hpp:298-314: the renumbering the Perm entry points use,
// here the IC0 plan's row order (parents before children, each unit's rows
// consecutive), and the schedule of A renumbered by it
void Scheduler::permPreprocessing() {
    if (!gs.sched) throw std::logic_error("Scheduler::permPreprocessing: call preprocessing() first");
    long long len = 0;
    baas_rt::check(baas_schedule_export(gs.sched, BAAS_EXPORT_UNIT_ROWS, nullptr, &len), "export");
    permToOrigByThread.assign((size_t)len, 0);
    baas_rt::check(baas_schedule_export(gs.sched, BAAS_EXPORT_UNIT_ROWS, permToOrigByThread.data(), &len),
                   "export");
    origToPermByThread.assign((size_t)csrA->m, 0);
    for (int i = 0; i < csrA->m; i++) origToPermByThread[permToOrigByThread[i]] = i;
    delete csrAPerm;
    csrAPerm = csrA->permute(origToPermByThread.data(), permToOrigByThread.data(), true);
    gsPerm.analyze(*csrAPerm, BAAS_KIND_IC0, num_thread, BAAS_PLAN_AUTO, 0, 0, 0);
}

void Scheduler::ic0_on(baas_rt::GpuSchedule &g, const SPM::CSR *mat, double *lu) {
    if (!g.sched) throw std::logic_error("Scheduler: not preprocessed");
    hip_ok(hipMemcpy(g.dA.values, mat->values, sizeof(double) * g.dA.nnz, hipMemcpyHostToDevice), "upload A");
    baas_rt::check(baas_ic0_f64(g.eng.h, g.sched, g.dA.values, d_lu), "baas_ic0_f64");
    const int st = baas_check(g.eng.h);
    if (st != BAAS_OK && st != BAAS_NONPOSITIVE_PIVOT) baas_rt::check(st, "ic0 kernel");
    hip_ok(hipMemcpy(lu, d_lu, sizeof(double) * g.dA.nnz, hipMemcpyDeviceToHost), "download lu");
    baas_rt::check(baas_ic0_split_f64(g.eng.h, g.sched, d_lu, d_lv, d_uv), "baas_ic0_split_f64");
}

void Scheduler::trsv_on(baas_rt::GpuSchedule &g, int uplo, const SPM::CSR *T, double *x, const double *b) {
    if (!g.sched) throw std::logic_error("Scheduler: not preprocessed");
    double *dv = uplo ? d_uv : d_lv;
    hip_ok(hipMemcpy(dv, T->values, sizeof(double) * nnzL, hipMemcpyHostToDevice), "upload factor");
    hip_ok(hipMemcpy(d_b, b, sizeof(double) * g.dA.n, hipMemcpyHostToDevice), "upload b");
    baas_rt::check(baas_sptrsv_f64(g.eng.h, g.sched, uplo, dv, d_b, d_x, 1), uplo ? "backward" : "forward");
    baas_rt::check(baas_check(g.eng.h), uplo ? "backward kernel" : "forward kernel");
    hip_ok(hipMemcpy(x, d_x, sizeof(double) * g.dA.n, hipMemcpyDeviceToHost), "download x");
}

void Scheduler::runIC0(const SPM::CSR *mat, double *lu) { ic0_on(gs, mat, lu); }
void Scheduler::runIC0Perm(const SPM::CSR *matPerm, double *lu) { ic0_on(gsPerm, matPerm, lu); }
void Scheduler::runForwardTRSVNoPerm(const SPM::CSR *L, double *x, const double *b) { trsv_on(gs, 0, L, x, b); }
void Scheduler::runForwardTRSVPerm(const SPM::CSR *L, double *x, const double *b) { trsv_on(gsPerm, 0, L, x, b); }
void Scheduler::runBackwardTRSVNoPerm(const SPM::CSR *U, double *x, const double *b) { trsv_on(gs, 1, U, x, b); }
void Scheduler::runBackwardTRSVPerm(const SPM::CSR *U, double *x, const double *b) { trsv_on(gsPerm, 1, U, x, b); }

}  // namespace Schedule
