// Loads a Matrix Market file with the runtime's SPM::CSR loaders (full:
// CSR(file); lower with a unit diagonal: CSR(file, true, 0)) and writes the
// arrays as raw little-endian files <prefix>.{full,lower}.{rowptr,colidx,values}
// so a test can compare them with the reference loader's digests; with a
// permutation file it also writes CSR::permute's result; the full matrix is
// also stored with CSR::storeBin as <prefix>.full.bin (a .bin input is read
// with loadBin).  No GPU.
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "baas_runtime.hpp"

static void dump(const SPM::CSR &A, const std::string &p) {
    std::ofstream(p + ".rowptr", std::ios::binary).write((const char *)A.rowptr, sizeof(int) * (A.m + 1));
    std::ofstream(p + ".colidx", std::ios::binary).write((const char *)A.colidx, sizeof(int) * A.getNnz());
    std::ofstream(p + ".values", std::ios::binary).write((const char *)A.values, sizeof(double) * A.getNnz());
}

int main(int argc, char *argv[]) {
    if (argc < 3) {
        fprintf(stderr, "usage: %s matrix.mtx out-prefix [perm.int32]\n", argv[0]);
        return 2;
    }
    SPM::CSR full(argv[1]);
    dump(full, std::string(argv[2]) + ".full");
    full.storeBin((std::string(argv[2]) + ".full.bin").c_str());
    const std::string in(argv[1]);
    if (in.size() > 4 && in.compare(in.size() - 4, 4, ".bin") == 0) {  // a binary input: no lower loader
        printf("full %d x %d, %d nnz\n", full.m, full.n, full.getNnz());
        return 0;
    }
    if (argc > 3) {
        // CSR::permute(origToPerm, permToOrig, sort) with permToOrig read from
        // a raw int32 file; the result's diagptr is dumped as well
        std::vector<int> p2o(full.m), o2p(full.m);
        std::ifstream(argv[3], std::ios::binary).read((char *)p2o.data(), sizeof(int) * full.m);
        for (int i = 0; i < full.m; i++) o2p[p2o[i]] = i;
        full.constructDiagPtr();
        for (int sort = 0; sort < 2; sort++) {
            SPM::CSR *P = full.permute(o2p.data(), p2o.data(), sort == 1);
            const std::string q = std::string(argv[2]) + (sort ? ".perm_sorted" : ".perm");
            dump(*P, q);
            std::ofstream(q + ".diagptr", std::ios::binary).write((const char *)P->diagptr, sizeof(int) * P->m);
            delete P;
        }
    }
    SPM::CSR lower(argv[1], true, 0);
    dump(lower, std::string(argv[2]) + ".lower");
    printf("full %d x %d, %d nnz; lower %d nnz\n", full.m, full.n, full.getNnz(), lower.getNnz());
    return 0;
}
