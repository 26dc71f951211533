// Test driver for the host set-up chain's symmetric eigensolver
// (da-gtm_amd/host/dg_init.cpp): reads n (int32) and the n x n row-major
// matrix (doubles) from argv[1]; writes the convergence flag (one byte), the
// n eigenvalues and, unless argv[3] is "values", the n x n eigenvectors
// (eigenvector i in row i) to argv[2].
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../da-gtm_amd/host/dg_init.hpp"

int main(int argc, char **argv) {
    if (argc < 3) return 2;
    FILE *f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    int n = 0;
    std::vector<double> A;
    bool ok = std::fread(&n, sizeof n, 1, f) == 1 && n >= 0;
    if (ok) {
        A.resize((size_t)n * n);
        ok = std::fread(A.data(), sizeof(double), A.size(), f) == A.size();
    }
    std::fclose(f);
    if (!ok) return 2;
    const bool vectors = !(argc > 3 && std::strcmp(argv[3], "values") == 0);
    std::vector<double> ev, Vt;
    const unsigned char conv = dgh::eig_sym(n, A, ev, vectors ? &Vt : nullptr) ? 1 : 0;
    FILE *g = std::fopen(argv[2], "wb");
    if (!g) return 2;
    std::fwrite(&conv, 1, 1, g);
    std::fwrite(ev.data(), sizeof(double), ev.size(), g);
    if (vectors) std::fwrite(Vt.data(), sizeof(double), Vt.size(), g);
    std::fclose(g);
    return 0;
}
