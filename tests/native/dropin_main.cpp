// A C++ caller of the drop-in par_assemble (include/Cassemble.h, std::string arguments, abort()
// on invalid input), used by tests/test_native_boundary.py.  It reads the mesh and the rules
// from a binary file and calls par_assemble("system", ...) as the reference's Cython module does
// (cython/nlfem.pyx:242-266); the matrix lands in the Armadillo file named on the command line.
// Built twice: as a program,
//
//   dropin_main INPUT OUTPUT KERNEL METHOD_REMOTE METHOD_CLOSE
//
// and, with DROPIN_LIB defined, as a shared library exporting dropin_run (same arguments, same
// result) for callers that must not start a process.
//
// INPUT: int64 [nE, nV, nPx, nPy, outdim, caps, K, nE_Omega, nV_Omega, dim], float64 sqdelta,
// then int64 elements[3 nE], elementLabels[nE], float64 vertices[2 nV], int64 vertexLabels[nV],
// float64 Px[2 nPx], dx[nPx], Py[2 nPy], dy[nPy].
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "Cassemble.h"

template <typename T>
static std::vector<T> take(std::FILE* f, size_t n, bool& ok) {
    std::vector<T> v(n ? n : 1);
    if (n && std::fread(v.data(), sizeof(T), n, f) != n) ok = false;
    return v;
}

static int short_input(std::FILE* f) {
    std::fclose(f);
    std::fprintf(stderr, "dropin_main: short input\n");
    return 2;
}

extern "C" int dropin_run(const char* input, const char* output, const char* kernel, const char* remote,
                          const char* close) {
    std::FILE* f = std::fopen(input, "rb");
    if (!f) return 2;
    bool ok = true;
    const std::vector<long> h = take<long>(f, 10, ok);
    if (!ok) return short_input(f);
    const long nE = h[0], nV = h[1], nPx = h[2], nPy = h[3], outdim = h[4], caps = h[5], K = h[6];
    const long nE_Omega = h[7], nV_Omega = h[8], dim = h[9];
    const double sqdelta = take<double>(f, 1, ok)[0];
    const std::vector<long> el = take<long>(f, 3 * nE, ok), lab = take<long>(f, nE, ok);
    const std::vector<double> verts = take<double>(f, 2 * nV, ok);
    const std::vector<long> vlab = take<long>(f, nV, ok);
    const std::vector<double> Px = take<double>(f, 2 * nPx, ok), dx = take<double>(f, nPx, ok);
    const std::vector<double> Py = take<double>(f, 2 * nPy, ok), dy = take<double>(f, nPy, ok);
    if (!ok) return short_input(f);
    std::fclose(f);
    par_assemble("system", output, "", (int)(nV_Omega * outdim), (int)K, el.data(), lab.data(), verts.data(),
                 vlab.data(), (int)nE, (int)nE_Omega, (int)nV, (int)nV_Omega, Px.data(), (int)nPx, dx.data(),
                 Py.data(), (int)nPy, dy.data(), sqdelta, 0, 0, kernel, "", remote, close, (int)caps, (int)dim,
                 (int)outdim, nullptr, nullptr, 0);
    return 0;
}

#ifndef DROPIN_LIB
int main(int argc, char** argv) {
    if (argc != 6) {
        std::fprintf(stderr, "usage: dropin_main INPUT OUTPUT KERNEL METHOD_REMOTE METHOD_CLOSE\n");
        return 2;
    }
    const int rc = dropin_run(argv[1], argv[2], argv[3], argv[4], argv[5]);
    if (rc == 0) std::printf("par_assemble done\n");
    return rc;
}
#endif
