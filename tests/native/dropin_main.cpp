// A C++ caller of the drop-in par_assemble (include/Cassemble.h, std::string arguments, abort()
// on invalid input), used by tests/test_native_boundary.py.  It reads the mesh and the rules
// from a binary file and calls par_assemble("system", ...) as the reference's Cython module does
// (cython/nlfem.pyx:242-266); the matrix lands in the Armadillo file named on the command line.
//
//   dropin_main INPUT OUTPUT KERNEL METHOD_REMOTE METHOD_CLOSE
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
static std::vector<T> take(std::FILE* f, size_t n) {
    std::vector<T> v(n ? n : 1);
    if (n && std::fread(v.data(), sizeof(T), n, f) != n) {
        std::fprintf(stderr, "dropin_main: short input\n");
        std::exit(2);
    }
    return v;
}

int main(int argc, char** argv) {
    if (argc != 6) {
        std::fprintf(stderr, "usage: dropin_main INPUT OUTPUT KERNEL METHOD_REMOTE METHOD_CLOSE\n");
        return 2;
    }
    std::FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    const std::vector<long> h = take<long>(f, 10);
    const long nE = h[0], nV = h[1], nPx = h[2], nPy = h[3], outdim = h[4], caps = h[5], K = h[6];
    const long nE_Omega = h[7], nV_Omega = h[8], dim = h[9];
    const double sqdelta = take<double>(f, 1)[0];
    const std::vector<long> el = take<long>(f, 3 * nE), lab = take<long>(f, nE);
    const std::vector<double> verts = take<double>(f, 2 * nV);
    const std::vector<long> vlab = take<long>(f, nV);
    const std::vector<double> Px = take<double>(f, 2 * nPx), dx = take<double>(f, nPx);
    const std::vector<double> Py = take<double>(f, 2 * nPy), dy = take<double>(f, nPy);
    std::fclose(f);
    par_assemble("system", argv[2], "", (int)(nV_Omega * outdim), (int)K, el.data(), lab.data(), verts.data(),
                 vlab.data(), (int)nE, (int)nE_Omega, (int)nV, (int)nV_Omega, Px.data(), (int)nPx, dx.data(),
                 Py.data(), (int)nPy, dy.data(), sqdelta, 0, 0, argv[3], "", argv[4], argv[5], (int)caps, (int)dim,
                 (int)outdim, nullptr, nullptr, 0);
    std::printf("par_assemble done\n");
    return 0;
}
