// nlfem_oracle.cpp -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
//
// FP64 CPU restatement (C++17 + OpenMP) of the reference nonlocal stiffness assembly
// (nlfem `par_assemble` -> `par_system`).  It is the checker for the HIP library in
// nlfem_amd/csrc and the timed CPU baseline of bench.py.  Only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg may load it.  Nothing in the product links it.
//
// Every function below names the reference file:line it restates.  Operation order of the
// geometric predicates (method_retriangulate, placePointOnCap, inTriangle, signDet, toPhys,
// vec_sqL2dist) follows the reference exactly so that integer outputs (neighbour lists,
// sparsity pattern) are bit-identical; the file is compiled with -ffp-contract=off so no
// fused multiply-add changes a rounding (the reference's x86-64 GCC build emits none).
//
// Cost structure of the reference is kept on purpose (it is also the CPU baseline):
// per-thread std::map scatter, std::queue BFS with an O(nE) visited reset per outer
// element, every thread scanning all nE elements, serial merge + sort.  METIS is replaced by
// (a) a dual graph built here (elements sharing >= 1 vertex, rows sorted ascending) and
// (b) a contiguous element-block partition instead of METIS_PartMeshDual.  Both only change
// floating-point summation order.
//
// Parity pins (tests/test_oracle*.py): kernels by closed forms, fully-contained pairs by the
// closed form of the 7-point rule, cut pairs of the polynomial kernels (constant, parabola) by
// an independent exact quadrature of the same polygons (tests/test_oracle_cutpairs.py),
// A*1 = 0, symmetry, Duffy tables by measure / polynomial exactness, the close-pair routines
// at s=-1 by an independent polynomial integral, integrate_fractional at s=0.4 on an identical
// pair by its convergence to a semi-analytic value and on vertex- and edge-sharing pairs
// against a direct quadrature, to the quadrature error of ~3e-4 (tests/test_oracle_fractional.py),
// retriangulation by disk-triangle area bounds and hand cases, BFS by brute force.
//
// Both close-pair routines at s=-0.5: identical pair to 1e-12 of the semi-analytic value,
// vertex-sharing pair to 1e-10 and edge-sharing pair to 2e-6 of a direct quadrature.
//
// Parity unpinned: the close-pair routines' touching pairs at other s below the accuracy of
// their direct-quadrature checks, and the cut-pair values of the non-polynomial kernels
// (fractional, microelastic), which share the integration loop that the polynomial kernels
// pin but whose 7-point sums have no exact value to compare with.

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <queue>
#include <string>
#include <vector>
#include <algorithm>
#include <chrono>
#include <omp.h>

namespace {

// ---------------------------------------------------------------- mathhelpers
// src/mathhelpers.cpp:362-368 absolute
inline double absolute(double v) { return v < 0 ? -v : v; }

// src/mathhelpers.cpp:41-71 solve2x2 (column pivot LU).  Returns 0 on success.
int solve2x2(const double* A, const double* b, double* x) {
    int dx0 = 0, dx1 = 1;
    if (absolute(A[0]) < absolute(A[2])) { dx0 = 1; dx1 = 0; }
    if (absolute(A[2 * dx0] - 0) < 1e-15) return -1;
    double l = A[2 * dx1] / A[2 * dx0];
    double u = A[2 * dx1 + 1] - l * A[2 * dx0 + 1];
    if (u == 0) return -1;
    x[1] = (b[dx1] - l * b[dx0]) / u;
    x[0] = (b[dx0] - A[2 * dx0 + 1] * x[1]) / A[2 * dx0];
    return 0;
}

// src/mathhelpers.cpp:74-78 rightNormal
inline void rightNormal(const double* y0, const double* y1, double orientation, double* n) {
    n[0] = y1[1] - y0[1];
    n[1] = y0[0] - y1[0];
    n[0] = n[0] * orientation;
    n[1] = n[1] * orientation;
}

// src/mathhelpers.cpp:101-109 absDet (2-D closed form); the arma::det variant
// (src/mathhelpers.cpp:112-129) evaluates the same 2x2 expression.
inline double absDet(const double* E) {
    double m00 = E[2] - E[0], m01 = E[4] - E[0];
    double m10 = E[3] - E[1], m11 = E[5] - E[1];
    return absolute(m00 * m11 - m01 * m10);
}

// src/mathhelpers.cpp:131-147 signDet
inline double signDet(const double* E) {
    double m00 = E[2] - E[0], m01 = E[4] - E[0];
    double m10 = E[3] - E[1], m11 = E[5] - E[1];
    double det = m00 * m11 - m01 * m10;
    if (det > 0) return 1.;
    if (det < 0) return -1.;
    return 0.0;
}

// src/mathhelpers.cpp:217-225 toPhys(E,p,dim,out): out = (p0*(E1-E0) + p1*(E2-E0)) + E0.
// The 3-argument overload (src/mathhelpers.cpp:210-215) computes the same sum.
inline void toPhys(const double* E, const double* p, double* out) {
    for (int i = 0; i < 2; i++) {
        double o = p[0] * (E[2 + i] - E[i]);
        o += p[1] * (E[4 + i] - E[i]);
        o += E[i];
        out[i] = o;
    }
}

// src/mathhelpers.cpp:227-240 toRef
inline int toRef(const double* E, const double* x, double* p) {
    double M[4] = {E[2] - E[0], E[4] - E[0], E[3] - E[1], E[5] - E[1]};
    double b[2] = {x[0] - E[0], x[1] - E[1]};
    return solve2x2(M, b, p);
}

// src/mathhelpers.cpp:272-279 vec_sqL2dist
inline double sqdist(const double* x, const double* y) {
    double d0 = x[0] - y[0], d1 = x[1] - y[1];
    return d0 * d0 + d1 * d1;
}
// src/mathhelpers.cpp:255-263 vec_dot
inline double dot2(const double* x, const double* y) { return x[0] * y[0] + x[1] * y[1]; }

// src/MeshTypes.cpp:16-24 model_basisFunction
inline void basis(const double* p, double* psi) {
    psi[0] = 1 - p[0] - p[1];
    psi[1] = p[0];
    psi[2] = p[1];
}

// ---------------------------------------------------------------- types
struct Mesh {
    int nE = 0, nV = 0, outdim = 1, K = 0, is_DG = 0;
    const long* tri = nullptr;
    const long* triLabel = nullptr;
    const double* verts = nullptr;
    const long* vertLabel = nullptr;
    double sqdelta = 0, delta = 0, fractional_s = -1;
    std::vector<int> xadj, adjncy;
};

struct Quad {
    const double *Px = nullptr, *dx = nullptr, *Py = nullptr, *dy = nullptr;
    int nPx = 0, nPy = 0;
    std::vector<double> psix, psiy;  // [nP][3]
    const double *Pg = nullptr, *dg = nullptr;
    int degree = 0, nPg = 0;
    // [case][traffo][k][4] and [case][traffo][k]
    std::vector<double> alpha[3], alphaC[3], detFrac[3], detWeak[3];
};

struct Elem {
    double E[6];
    long label;
    double absDet;
    long Tdx;
};

typedef void (*KernelFn)(const double* x, long lx, const double* y, long ly, const Mesh& m, double* out);
typedef int (*MethodFn)(const double* x, const double* TE, double sqdelta, double* list, int caps);

struct Conf {
    KernelFn kernel = nullptr;
    int remote = 0, close = 0;  // 0 retriangulate, 1 baryCenter, 2 fractional, 3 weakSingular
    bool caps = true;
    bool singular = false;
    bool fullCC = false;
};

// Planted mistakes, for tests that must show they catch them (0 in normal use):
// PLANT_BLOCKED_LAYOUT numbers the outdim-2 degrees of freedom component-major
// (comp * nV + node) instead of the reference's interleaved od * node + comp;
// PLANT_FIELD_DIAGONAL exchanges k00 and k11 of the field kernel.
constexpr int PLANT_BLOCKED_LAYOUT = 1, PLANT_FIELD_DIAGONAL = 2;
static int g_plant = 0;
extern "C" void oracle_set_plant(int mode) { g_plant = mode; }

// ---------------------------------------------------------------- kernels (src/model.cpp)
// src/model.cpp:25-28
void kernel_constant(const double*, long, const double*, long, const Mesh& m, double* k) {
    *k = 4 / (M_PI * (m.sqdelta * m.sqdelta));
}
// src/model.cpp:52-60
void kernel_parabola(const double* x, long, const double* y, long, const Mesh& m, double* k) {
    double z[2] = {x[0] - y[0], x[1] - y[1]};
    const double value = m.sqdelta - dot2(z, z);
    const double c = 12. / (M_PI * std::pow(m.sqdelta, 3));
    *k = c * value;
}
// src/model.cpp:116-124
void kernel_linearPrototypeMicroelastic(const double* x, long, const double* y, long, const Mesh& m, double* k) {
    double z[2] = {x[0] - y[0], x[1] - y[1]};
    const double den = 1.0 / std::sqrt(dot2(z, z));
    const double c = 3.0 / (M_PI * std::pow(m.delta, 3));
    *k = c * den;
}
// src/model.cpp:126-138
void kernel_fractional(const double* x, long, const double* y, long, const Mesh& m, double* k) {
    const double exponent = 2 + 2. * m.fractional_s;
    const double cexp = 2. - 2. * m.fractional_s;
    double z[2] = {x[0] - y[0], x[1] - y[1]};
    const double den = 1.0 / std::pow(std::sqrt(dot2(z, z)), exponent);
    const double c = cexp / (M_PI * std::pow(m.delta, cexp));
    *k = c * den;
}
// src/model.cpp:140-153 (row-major [k00, k01, k10, k11])
void kernelField_linearPrototypeMicroelastic(const double* x, long, const double* y, long, const Mesh& m,
                                             double* k) {
    double z[2] = {x[0] - y[0], x[1] - y[1]};
    double den = 1.0 / std::pow(std::sqrt(dot2(z, z)), 3);
    double c = 3.0 / std::pow(m.delta, 3);
    k[0] = c * den * z[0] * z[0];
    k[1] = c * den * z[0] * z[1];
    k[2] = c * den * z[1] * z[0];
    k[3] = c * den * z[1] * z[1];
    if (g_plant & PLANT_FIELD_DIAGONAL) std::swap(k[0], k[3]);
}

// forcing: src/model.cpp:240-242 f_linear, :201-203 f_constant, :204-210 fField_linear
typedef void (*ForceFn)(const double* x, double* out);
void f_linear(const double* x, double* f) { *f = -2. * (x[1] + 1.); }
void f_constant(const double*, double* f) { *f = 1.0; }
void fField_linear(const double* x, double* f) {
    const double c = M_PI / 2.0;
    f[0] = -c * (1.0 + 2 * x[0]);
    f[1] = -c * x[1];
}

// ---------------------------------------------------------------- retriangulation
// src/integration.cpp:1283-1298 inTriangle
inline bool inTriangle(const double* y, const double* p, const double* q, const double* r, const double* na,
                       const double* nb, const double* nc) {
    double v[2];
    v[0] = y[0] - p[0]; v[1] = y[1] - p[1];
    bool a = dot2(na, v) >= 0;
    v[0] = y[0] - q[0]; v[1] = y[1] - q[1];
    bool b = dot2(nb, v) >= 0;
    v[0] = y[0] - r[0]; v[1] = y[1] - r[1];
    bool c = dot2(nc, v) >= 0;
    return a && b && c;
}

// src/integration.cpp:1353-1377 placePointOnCap
inline int placePointOnCap(const double* ypred, const double* ycur, const double* xc, double sqdelta,
                           const double* TE, const double* na, const double* nb, const double* nc,
                           double orientation, int Rdx, double* R) {
    double s[2], ynew[2];
    rightNormal(ycur, ypred, orientation, s);
    double scaling = std::sqrt(sqdelta / dot2(s, s));
    s[0] = s[0] * scaling;
    s[1] = s[1] * scaling;
    ynew[0] = xc[0] + s[0];
    ynew[1] = xc[1] + s[1];
    if (inTriangle(ynew, &TE[0], &TE[2], &TE[4], na, nb, nc)) {
        R[2 * Rdx] = ynew[0];
        R[2 * Rdx + 1] = ynew[1];
        return 1;
    }
    return 0;
}

// src/integration.cpp:1393-1512 (== :1775-1894) method_retriangulate.  Returns the number
// of triangles of the fan (Rdx-2) or 0.
int method_retriangulate(const double* xc, const double* TE, double sqdelta, double* list, int caps) {
    int Rdx = 0;
    double na[2], nb[2], nc[2];
    bool is_onEdge = false, firstOnVertex = true;
    double R[2 * 12];
    std::memset(R, 0, sizeof(R));
    double orientation = -signDet(TE);
    rightNormal(&TE[0], &TE[2], orientation, na);
    rightNormal(&TE[2], &TE[4], orientation, nb);
    rightNormal(&TE[4], &TE[0], orientation, nc);
    for (int k = 0; k < 3; k++) {
        const double* p = &TE[2 * k];
        const double* q = &TE[2 * ((k + 1) % 3)];
        double a[2] = {q[0] - xc[0], q[1] - xc[1]};
        double b[2] = {p[0] - q[0], p[1] - q[1]};
        if (sqdist(p, xc) <= sqdelta) {
            R[2 * Rdx] = p[0];
            R[2 * Rdx + 1] = p[1];
            is_onEdge = false;
            Rdx += 1;
        }
        double ab = dot2(a, b), aa = dot2(a, a), bb = dot2(b, b);
        double v = ab * ab - (aa - sqdelta) * bb;
        if (v >= 0) {
            double term1 = -ab / bb;
            double term2 = std::sqrt(v) / bb;
            double lam1, lam2;
            if (term1 > 0) {
                lam1 = term1 + term2;
                lam2 = 1 / lam1 * (aa - sqdelta) / bb;
            } else {
                lam2 = term1 - term2;
                lam1 = 1 / lam2 * (aa - sqdelta) / bb;
            }
            double y1[2] = {lam1 * b[0] + q[0], lam1 * b[1] + q[1]};
            double y2[2] = {lam2 * b[0] + q[0], lam2 * b[1] + q[1]};
            if ((0 <= lam1) && (lam1 <= 1)) {
                firstOnVertex = firstOnVertex && (bool)Rdx;
                if (is_onEdge && caps)
                    Rdx += placePointOnCap(&R[2 * (Rdx - 1)], y1, xc, sqdelta, TE, na, nb, nc, orientation, Rdx, R);
                R[2 * Rdx] = y1[0];
                R[2 * Rdx + 1] = y1[1];
                is_onEdge = true;
                Rdx += 1;
            }
            double dl = lam1 - lam2;
            if ((0 <= lam2) && (lam2 <= 1) && (dl * dl > 0)) {
                firstOnVertex = firstOnVertex && (bool)Rdx;
                if (is_onEdge && caps)
                    Rdx += placePointOnCap(&R[2 * (Rdx - 1)], y2, xc, sqdelta, TE, na, nb, nc, orientation, Rdx, R);
                R[2 * Rdx] = y2[0];
                R[2 * Rdx + 1] = y2[1];
                is_onEdge = true;
                Rdx += 1;
            }
        }
    }
    if (is_onEdge && (!firstOnVertex && Rdx > 1) && caps)
        Rdx += placePointOnCap(&R[2 * (Rdx - 1)], &R[0], xc, sqdelta, TE, na, nb, nc, orientation, Rdx, R);
    if (Rdx < 3) return 0;
    for (int k = 0; k < Rdx - 2; k++) {
        for (int i = 0; i < 2; i++) {
            list[2 * (3 * k + 0) + i] = R[i];
            list[2 * (3 * k + 1) + i] = R[2 * (k + 1) + i];
            list[2 * (3 * k + 2) + i] = R[2 * (k + 2) + i];
        }
    }
    return Rdx - 2;
}

// src/integration.cpp:1262-1280 method_baryCenter (+ baryCenter src/mathhelpers.cpp:184-196)
int method_baryCenter(const double* xc, const double* TE, double sqdelta, double* list, int) {
    double bary[2];
    for (int j = 0; j < 2; j++) {
        double s = 0;
        for (int i = 0; i < 3; i++) s += TE[2 * i + j];
        bary[j] = s / 3.0;
    }
    double distance = sqdist(xc, bary);
    if (distance > sqdelta) return 0;
    for (int i = 0; i < 6; i++) list[i] = TE[i];
    return -1;
}

// ---------------------------------------------------------------- integration routines
// src/integration.cpp:345-540 integrate_retriangulate
int integrate_retriangulate(const Elem& aT, const Elem& bT, const Quad& q, const Mesh& m, const Conf& c,
                            double* tL, double* tN, double* tLP, double* tNP) {
    const int od = m.outdim, oo = od * od;
    double x[2], innerLocal[4], innerNonloc[12], innerNonlocPrime[12], kv[4], kpv[4];
    double phys[2], ref[2], psi[3];
    double list[9 * 3 * 2];
    int doesInteract = 0;
    for (int k = 0; k < q.nPx; k++) {
        toPhys(aT.E, &q.Px[2 * k], x);
        int Rdx = method_retriangulate(x, bT.E, m.sqdelta, list, c.caps);
        doesInteract += Rdx;
        for (int o = 0; o < oo; o++) innerLocal[o] = 0;
        for (int o = 0; o < 3 * oo; o++) { innerNonloc[o] = 0; innerNonlocPrime[o] = 0; }
        for (int r = 0; r < Rdx; r++) {
            for (int i = 0; i < q.nPy; i++) {
                toPhys(&list[6 * r], &q.Py[2 * i], phys);
                double rTdet = absDet(&list[6 * r]);
                c.kernel(x, aT.label, phys, bT.label, m, kv);
                for (int o = 0; o < oo; o++) innerLocal[o] += kv[o] * q.dy[i] * rTdet;
                if (toRef(bT.E, phys, ref)) { std::fprintf(stderr, "oracle: singular toRef\n"); std::abort(); }
                c.kernel(phys, bT.label, x, aT.label, m, kpv);
                basis(ref, psi);
                for (int b = 0; b < 3 * oo; b++) {
                    innerNonloc[b] += psi[b / oo] * kv[b % oo] * q.dy[i] * rTdet;
                    innerNonlocPrime[b] += psi[b / oo] * kpv[b % oo] * q.dy[i] * rTdet;
                }
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++)
                        for (int o = 0; o < oo; o++)
                            tLP[a * oo * 3 + (o / od) * od * 3 + b * od + o % od] +=
                                psi[a] * psi[b] * kpv[o] * q.dy[i] * q.dx[k] * rTdet * aT.absDet;
            }
        }
        for (int a = 0; a < 3 * od; a++) {
            for (int b = 0; b < 3 * od; b++) {
                tL[3 * od * a + b] += aT.absDet * q.psix[3 * k + a / od] * q.psix[3 * k + b / od] * q.dx[k] *
                                      innerLocal[od * (a % od) + (b % od)];
                tN[3 * od * a + b] += aT.absDet * q.psix[3 * k + a / od] * q.dx[k] *
                                      innerNonloc[(a % od) * od + oo * (b / od) + (b % od)];
                tNP[3 * od * a + b] += aT.absDet * q.psix[3 * k + a / od] * q.dx[k] *
                                       innerNonlocPrime[(a % od) * od + oo * (b / od) + (b % od)];
            }
        }
    }
    return (doesInteract || c.fullCC);
}

// src/integration.cpp:910-1018 integrate_baryCenter
int integrate_baryCenter(const Elem& aT, const Elem& bT, const Quad& q, const Mesh& m, const Conf& c,
                         double* tL, double* tN, double*, double*) {
    const int od = m.outdim, oo = od * od;
    double x[2], innerLocal[4], innerNonloc[12], kv[4], phys[2];
    double list[6];
    int doesInteract = 0;
    for (int k = 0; k < q.nPx; k++) {
        toPhys(aT.E, &q.Px[2 * k], x);
        for (int o = 0; o < oo; o++) innerLocal[o] = 0;
        for (int o = 0; o < 3 * oo; o++) innerNonloc[o] = 0;
        int Rdx = method_baryCenter(x, bT.E, m.sqdelta, list, 0);
        doesInteract += Rdx;
        if (Rdx) {
            for (int i = 0; i < q.nPy; i++) {
                toPhys(bT.E, &q.Py[2 * i], phys);
                double rTdet = absDet(bT.E);
                c.kernel(x, aT.label, phys, bT.label, m, kv);
                for (int o = 0; o < oo; o++) innerLocal[o] += rTdet * kv[o] * q.dy[i];
                c.kernel(phys, bT.label, x, aT.label, m, kv);
                for (int b = 0; b < 3 * oo; b++)
                    innerNonloc[b] += q.psiy[3 * i + b / oo] * kv[b % oo] * q.dy[i] * rTdet;
            }
        }
        for (int a = 0; a < 3 * od; a++) {
            for (int b = 0; b < 3 * od; b++) {
                tL[3 * od * a + b] += 2 * aT.absDet * q.psix[3 * k + a / od] * q.psix[3 * k + b / od] * q.dx[k] *
                                      innerLocal[od * (a % od) + (b % od)];
                tN[3 * od * a + b] += 2 * aT.absDet * q.psix[3 * k + a / od] * q.dx[k] *
                                      innerNonloc[(a % od) * od + oo * (b / od) + (b % od)];
            }
        }
    }
    return (doesInteract || c.fullCC);
}

// src/integration.cpp:1913-1958 join (+ setupElement :1897-1912)
int join(const Mesh& m, const Elem& aT, const Elem& bT, Elem& as, Elem& bs, int* argA, int* argB) {
    int nEqual = 0, AinB[3] = {0, 0, 0}, BinA[3] = {0, 0, 0};
    const long* aV = &m.tri[3 * aT.Tdx];
    const long* bV = &m.tri[3 * bT.Tdx];
    long aS[3], bS[3];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
            if (aV[a] == bV[b]) {
                AinB[a] += 1; BinA[b] += 1;
                aS[nEqual] = aV[a]; argA[nEqual] = a;
                bS[nEqual] = bV[b]; argB[nEqual] = b;
                nEqual += 1;
            }
    int ia = 0, ib = 0;
    for (int i = 0; i < 3; i++) {
        if (!AinB[i]) { aS[nEqual + ia] = aV[i]; argA[nEqual + ia] = i; ia++; }
        if (!BinA[i]) { bS[nEqual + ib] = bV[i]; argB[nEqual + ib] = i; ib++; }
    }
    for (int k = 0; k < 3; k++)
        for (int j = 0; j < 2; j++) {
            as.E[2 * k + j] = m.verts[2 * aS[k] + j];
            bs.E[2 * k + j] = m.verts[2 * bS[k] + j];
        }
    as.absDet = absDet(as.E);
    bs.absDet = absDet(bs.E);
    as.label = aT.label;
    bs.label = bT.label;
    return nEqual;
}

inline int ntraffo(int cs) { return cs == 0 ? 2 : (cs == 1 ? 5 : 6); }

// src/integration.cpp:187-288 integrate_fractional
int integrate_fractional(const Elem& aT, const Elem& bT, const Quad& q, const Mesh& m, const Conf& c,
                         double* tL, double* tN, double* tLP, double* tNP) {
    Elem as, bs;
    int argA[3], argB[3];
    const int nEqual = join(m, aT, bT, as, bs, argA, argB);
    if (nEqual < 1 || nEqual > 3) { std::fprintf(stderr, "oracle: join nEqual=%d\n", nEqual); std::abort(); }
    const int od = m.outdim, cs = nEqual - 1, nT = ntraffo(cs);
    const int nLocal = 6 - nEqual, shift = 3 - nEqual;
    double xc[2], yc[2], kv[4], psix[3], psiy[3], psixy[6];
    for (int t = 0; t < nT; t++) {
        for (int k = 0; k < q.nPg; k++) {
            const double* al = &q.alphaC[cs][(t * q.nPg + k) * 4];
            double det = q.detFrac[cs][t * q.nPg + k];
            toPhys(as.E, &al[0], xc);
            toPhys(bs.E, &al[2], yc);
            c.kernel(xc, as.label, yc, bs.label, m, kv);
            basis(&al[0], psix);
            basis(&al[2], psiy);
            for (int a = 0; a < nEqual; a++) psixy[a] = psix[a] - psiy[a];
            for (int a = nEqual; a < 3; a++) { psixy[a] = psix[a]; psixy[a + shift] = -psiy[a]; }
            for (int a = 0; a < nLocal; a++)
                for (int ao = 0; ao < od; ao++)
                    for (int b = 0; b < nLocal; b++)
                        for (int bo = 0; bo < od; bo++) {
                            double f = kv[od * ao + bo] * det * psixy[a] * psixy[b] * q.dg[k] * as.absDet * bs.absDet;
                            if (a < 3) {
                                if (b < 3) tL[od * 3 * (argA[a] * od + ao) + argA[b] * od + bo] += f;
                                else tN[od * 3 * (argA[a] * od + ao) + argB[b - shift] * od + bo] -= f;
                            } else {
                                if (b < 3) tNP[od * 3 * (argA[b] * od + ao) + argB[a - shift] * od + bo] -= f;
                                else tLP[od * 3 * (argB[a - shift] * od + ao) + argB[b - shift] * od + bo] += f;
                            }
                        }
        }
    }
    return 1;
}

// src/integration.cpp:82-185 integrate_weakSingular
int integrate_weakSingular(const Elem& aT, const Elem& bT, const Quad& q, const Mesh& m, const Conf& c,
                           double* tL, double* tN, double* tLP, double* tNP) {
    Elem as, bs;
    int argA[3], argB[3];
    const int nEqual = join(m, aT, bT, as, bs, argA, argB);
    if (nEqual < 1 || nEqual > 3) { std::fprintf(stderr, "oracle: join nEqual=%d\n", nEqual); std::abort(); }
    const int od = m.outdim, cs = nEqual - 1, nT = ntraffo(cs);
    double xc[2], yc[2], kv[4], psix[3], psiy[3];
    for (int t = 0; t < nT; t++) {
        for (int k = 0; k < q.nPg; k++) {
            double det = q.detWeak[cs][t * q.nPg + k];
            const double* alC = &q.alphaC[cs][(t * q.nPg + k) * 4];
            const double* al = &q.alpha[cs][(t * q.nPg + k) * 4];
            toPhys(as.E, &alC[0], xc);
            toPhys(bs.E, &alC[2], yc);
            c.kernel(xc, as.label, yc, bs.label, m, kv);
            basis(&al[0], psix);
            basis(&al[2], psiy);
            for (int a = 0; a < 3; a++)
                for (int ao = 0; ao < od; ao++)
                    for (int b = 0; b < 3; b++)
                        for (int bo = 0; bo < od; bo++) {
                            double f = kv[od * ao + bo] * det * q.dg[k] * as.absDet * bs.absDet;
                            tL[od * 3 * (argA[a] * od + ao) + argA[b] * od + bo] += f * psix[a] * psix[b];
                            tLP[od * 3 * (argB[a] * od + ao) + argB[b] * od + bo] += f * psiy[a] * psiy[b];
                            tN[od * 3 * (argA[a] * od + ao) + argB[b] * od + bo] += f * psix[a] * psiy[b];
                            tNP[od * 3 * (argA[a] * od + ao) + argB[b] * od + bo] += f * psix[a] * psiy[b];
                        }
        }
    }
    return 1;
}

typedef int (*IntegrateFn)(const Elem&, const Elem&, const Quad&, const Mesh&, const Conf&, double*, double*,
                           double*, double*);
IntegrateFn integrator(int id) {
    switch (id) {
        case 0: return integrate_retriangulate;
        case 1: return integrate_baryCenter;
        case 2: return integrate_fractional;
        case 3: return integrate_weakSingular;
    }
    return nullptr;
}

// ---------------------------------------------------------------- quadrature tables
// src/MeshTypes.cpp:32-209 Duffy maps.  Each returns the map's Jacobian factor.
double traffo(int cs, int t, double* a) {
    const double xi = a[0], e1 = a[1], e2 = a[2], e3 = a[3];
    if (cs == 0) {
        if (t == 0) { a[0] = xi; a[1] = e1 * xi; a[2] = e2 * xi; a[3] = e2 * e3 * xi; }
        else { a[0] = xi * e2; a[1] = xi * e2 * e3; a[2] = xi; a[3] = xi * e1; }
        return std::pow(xi, 3) * e2;
    }
    if (cs == 1) {
        switch (t) {
            case 0: a[0] = xi; a[1] = xi * e1 * e3; a[2] = xi * (1. - e1 * e2); a[3] = xi * e1 * (1. - e2);
                return std::pow(xi, 3) * (e1 * e1);
            case 1: a[0] = xi; a[1] = xi * e1; a[2] = xi * (1. - e1 * e2 * e3); a[3] = xi * e1 * e2 * (1. - e3); break;
            case 2: a[0] = xi * (1. - e1 * e2); a[1] = xi * e1 * (1. - e2); a[2] = xi; a[3] = xi * e1 * e2 * e3; break;
            case 3: a[0] = xi * (1. - e1 * e2 * e3); a[1] = xi * e1 * e2 * (1. - e3); a[2] = xi; a[3] = xi * e1; break;
            case 4: a[0] = xi * (1. - e1 * e2 * e3); a[1] = xi * e1 * (1. - e2 * e3); a[2] = xi; a[3] = xi * e1 * e2; break;
        }
        return std::pow(xi, 3) * (e1 * e1) * e2;
    }
    switch (t) {
        case 0: a[0] = xi; a[1] = xi * (1. - e1 + e1 * e2); a[2] = xi * (1. - e1 * e2 * e3); a[3] = xi * (1. - e1); break;
        case 1: a[0] = xi * (1. - e1 * e2 * e3); a[1] = xi * (1. - e1); a[2] = xi; a[3] = xi * (1. - e1 + e1 * e2); break;
        case 2: a[0] = xi; a[1] = xi * e1 * (1. - e2 + e2 * e3); a[2] = xi * (1. - e1 * e2); a[3] = xi * e1 * (1. - e2); break;
        case 3: a[0] = xi * (1. - e1 * e2); a[1] = xi * e1 * (1. - e2); a[2] = xi; a[3] = xi * e1 * (1. - e2 + e2 * e3); break;
        case 4: a[0] = xi * (1. - e1 * e2 * e3); a[1] = xi * e1 * (1. - e2 * e3); a[2] = xi; a[3] = xi * e1 * (1. - e2); break;
        case 5: a[0] = xi; a[1] = xi * e1 * (1. - e2); a[2] = xi * (1. - e1 * e2 * e3); a[3] = xi * e1 * (1. - e2 * e3); break;
    }
    return std::pow(xi, 3) * (e1 * e1) * e2;
}

// src/MeshTypes.cpp:252-328 initializeQuadrule (+ scale/mirror src/mathhelpers.cpp:88-97)
void initQuadrule(Quad& q, double s) {
    const double SCALEDET = 0.0625;
    for (int cs = 0; cs < 3; cs++) {
        int nT = ntraffo(cs);
        q.alpha[cs].assign((size_t)nT * q.nPg * 4, 0.0);
        q.alphaC[cs].assign((size_t)nT * q.nPg * 4, 0.0);
        q.detFrac[cs].assign((size_t)nT * q.nPg, 0.0);
        q.detWeak[cs].assign((size_t)nT * q.nPg, 0.0);
        for (int t = 0; t < nT; t++) {
            for (int k = 0; k < q.nPg; k++) {
                double* al = &q.alpha[cs][((size_t)t * q.nPg + k) * 4];
                double* ac = &q.alphaC[cs][((size_t)t * q.nPg + k) * 4];
                for (int j = 0; j < 4; j++) { al[j] = q.Pg[4 * k + j]; ac[j] = q.Pg[4 * k + j]; }
                for (int j = 0; j < 4; j++) { al[j] = al[j] * 0.5 + 0.5; ac[j] = ac[j] * 0.5 + 0.5; }
                double dw = std::pow(ac[0], 3 - 2 - 2 * s);
                double df = std::pow(ac[0], 3 - 2 * s);
                ac[0] = 1.0;
                traffo(cs, t, al);
                double tdet = traffo(cs, t, ac);
                dw *= tdet * SCALEDET;
                df *= tdet * SCALEDET;
                q.detWeak[cs][(size_t)t * q.nPg + k] = dw;
                q.detFrac[cs][(size_t)t * q.nPg + k] = df;
                al[0] = al[0] - al[1]; al[2] = al[2] - al[3];
                ac[0] = ac[0] - ac[1]; ac[2] = ac[2] - ac[3];
            }
        }
    }
}

// ---------------------------------------------------------------- dual graph
// Replaces METIS_MeshToDual(ncommon=1) (src/Cassemble.cpp:389-417): neighbours = elements
// sharing >= 1 vertex, self excluded, each row sorted ascending.
void dualGraph(Mesh& m) {
    std::vector<int> vcnt(m.nV + 1, 0);
    for (int e = 0; e < m.nE; e++)
        for (int j = 0; j < 3; j++) vcnt[m.tri[3 * e + j] + 1]++;
    for (int v = 0; v < m.nV; v++) vcnt[v + 1] += vcnt[v];
    std::vector<int> vel(vcnt[m.nV]), fill(vcnt.begin(), vcnt.end() - 1);
    for (int e = 0; e < m.nE; e++)
        for (int j = 0; j < 3; j++) vel[fill[m.tri[3 * e + j]]++] = e;
    m.xadj.assign(m.nE + 1, 0);
    m.adjncy.clear();
    std::vector<int> row;
    for (int e = 0; e < m.nE; e++) {
        row.clear();
        for (int j = 0; j < 3; j++) {
            long v = m.tri[3 * e + j];
            for (int p = vcnt[v]; p < vcnt[v + 1]; p++)
                if (vel[p] != e) row.push_back(vel[p]);
        }
        std::sort(row.begin(), row.end());
        row.erase(std::unique(row.begin(), row.end()), row.end());
        m.adjncy.insert(m.adjncy.end(), row.begin(), row.end());
        m.xadj[e + 1] = (int)m.adjncy.size();
    }
}

void initElement(const Mesh& m, int Tdx, Elem& T) {
    for (int k = 0; k < 3; k++) {
        long v = m.tri[3 * Tdx + k];
        T.E[2 * k] = m.verts[2 * v];
        T.E[2 * k + 1] = m.verts[2 * v + 1];
    }
    T.absDet = absDet(T.E);
    T.label = m.triLabel[Tdx];
    T.Tdx = Tdx;
}

// ---------------------------------------------------------------- configuration lookup
KernelFn lookupKernel(const std::string& s) {
    if (s == "constant") return kernel_constant;
    if (s == "parabola") return kernel_parabola;
    if (s == "fractional") return kernel_fractional;
    if (s == "linearPrototypeMicroelastic") return kernel_linearPrototypeMicroelastic;
    if (s == "linearPrototypeMicroelasticField") return kernelField_linearPrototypeMicroelastic;
    return nullptr;
}
int lookupMethod(const std::string& s) {
    if (s == "retriangulate") return 0;
    if (s == "baryCenter") return 1;
    if (s == "fractional") return 2;
    if (s == "weakSingular") return 3;
    return -1;
}
ForceFn lookupForce(const std::string& s) {
    if (s == "linear") return f_linear;
    if (s == "constant") return f_constant;
    if (s == "linearField") return fField_linear;
    return nullptr;
}

}  // namespace

// ================================================================ C ABI (ctypes)
// src/mathhelpers.cpp:19-38 evaluateZeta: size of the intersection of the indicator rows of
// aTdx and bTdx (both sorted ascending, as std::set_intersection requires).
static double evaluateZeta(const long* indices, const long* indptr, long aTdx, long bTdx) {
    long i = indptr[aTdx], ie = indptr[aTdx + 1], j = indptr[bTdx], je = indptr[bTdx + 1], n = 0;
    while (i < ie && j < je) {
        if (indices[i] < indices[j]) i++;
        else if (indices[j] < indices[i]) j++;
        else { n++; i++; j++; }
    }
    return (double)n;
}

extern "C" {

struct OracleResult {
    long K = 0;
    std::vector<long> rowptr, colidx;
    std::vector<double> values;
    // neighbour lists: per outer element, bT indices in BFS order, flags bit0 = close layer
    std::vector<long> nbr_ptr;
    std::vector<int> nbr;
    std::vector<unsigned char> nbr_close;
    long n_pairs = 0, n_probes = 0, n_close = 0;
    double seconds = 0;
};

static thread_local char g_err[512];
// When set, every addend is scattered as its absolute value: the resulting matrix bounds the
// rounding error an entry can carry (condition-aware test tolerances).
static int g_abs = 0;
void oracle_set_abs_mode(int on) { g_abs = on; }
const char* oracle_last_error() { return g_err; }

// src/Cassemble.cpp:360-462 (par_assemble, "system") + :464-855 (par_system).
// Returns nullptr on an invalid configuration (the reference aborts).
OracleResult* oracle_par_system(int K, const long* tri, const long* triLabel, const double* verts,
                                const long* vertLabel, int nE, int nV, const double* Px, int nPx, const double* dx,
                                const double* Py, int nPy, const double* dy, double sqdelta, int is_DG,
                                const char* kernel, const char* method_remote, const char* method_close,
                                int caps, int outdim, const double* Pg, int degree, const double* dg,
                                double fractional_s, int fullCC, int nthreads, int keep_neighbours,
                                int outer_begin, int outer_end, const long* zeta_indices, const long* zeta_indptr,
                                long nZeta, int outer_stride) {
    Mesh m;
    m.nE = nE; m.nV = nV; m.outdim = outdim; m.K = K; m.is_DG = is_DG;
    m.tri = tri; m.triLabel = triLabel; m.verts = verts; m.vertLabel = vertLabel;
    m.sqdelta = sqdelta; m.delta = std::sqrt(sqdelta); m.fractional_s = fractional_s;
    Quad q;
    q.Px = Px; q.dx = dx; q.Py = Py; q.dy = dy; q.nPx = nPx; q.nPy = nPy;
    q.Pg = Pg; q.dg = dg; q.degree = degree; q.nPg = (int)std::pow(degree, 4);
    Conf c;
    c.kernel = lookupKernel(kernel);
    c.remote = lookupMethod(method_remote);
    c.close = lookupMethod(method_close);
    c.caps = caps != 0;
    c.fullCC = fullCC != 0;
    std::string ks(kernel);
    c.singular = (ks == "fractional" || ks == "linearPrototypeMicroelastic" || ks == "linearPrototypeMicroelasticField");
    if (!c.kernel || c.remote < 0 || c.close < 0) { std::snprintf(g_err, sizeof g_err, "unknown kernel/method"); return nullptr; }
    if (c.remote >= 2) { std::snprintf(g_err, sizeof g_err, "fractional/weakSingular are close-only"); return nullptr; }
    if ((c.close >= 2) && degree <= 0) { std::snprintf(g_err, sizeof g_err, "tensor Gauss degree required"); return nullptr; }
    if (c.close == 2 && is_DG) { std::snprintf(g_err, sizeof g_err, "fractional close pairs need CG"); return nullptr; }
    if ((ks == "linearPrototypeMicroelastic" || ks == "linearPrototypeMicroelasticField") && fractional_s != -0.5) {
        std::snprintf(g_err, sizeof g_err, "microelastic kernels need fractional_s=-0.5"); return nullptr;
    }
    if (is_DG ? (outdim * nE * 3 != K) : (outdim * nV != K)) { std::snprintf(g_err, sizeof g_err, "K mismatch"); return nullptr; }

    auto t0 = std::chrono::steady_clock::now();
    dualGraph(m);
    if (q.nPg > 0) initQuadrule(q, fractional_s);
    q.psix.resize(3 * nPx); q.psiy.resize(3 * nPy);
    for (int h = 0; h < nPx; h++) basis(&Px[2 * h], &q.psix[3 * h]);
    for (int h = 0; h < nPy; h++) basis(&Py[2 * h], &q.psiy[3 * h]);

    IntegrateFn integ_remote = integrator(c.remote), integ_close = integrator(c.close);
    const int nparts = nthreads > 0 ? nthreads : omp_get_max_threads();
    OracleResult* res = new OracleResult();
    res->K = K;
    std::vector<std::vector<std::pair<unsigned long, double>>> parts(nparts);
    std::vector<std::vector<int>> nbrs_t(keep_neighbours ? nE : 0);
    std::vector<std::vector<unsigned char>> close_t(keep_neighbours ? nE : 0);
    long pairs = 0, probes = 0, closec = 0;
    const int dV = 3, od = outdim, nb = dV * dV * od * od;
#pragma omp parallel num_threads(nparts) reduction(+ : pairs, probes, closec)
    {
        const int tid = omp_get_thread_num();
        std::map<unsigned long, double> Ad;
        std::vector<int> visited(nE, 0);
        std::queue<int> queue;
        Elem aT, bT;
        long aDG[3], bDG[3];
        double tL[36], tN[36], tLP[36], tNP[36];
        // contiguous block partition (replaces METIS_PartMeshDual, src/Cassemble.cpp:483-503)
        const long span0 = (outer_end > outer_begin) ? outer_begin : 0;
        const long span = (outer_end > outer_begin) ? (long)outer_end - outer_begin : (long)nE;
        const long lo = span0 + span * tid / nparts, hi = span0 + span * (tid + 1) / nparts;
        for (int aTdx = 0; aTdx < nE; aTdx++) {  // every thread scans all elements (src/Cassemble.cpp:555-559)
            // optional sample of the outer loop [outer_begin, outer_end) for timing (CPU baseline)
            if (outer_end > outer_begin && (aTdx < outer_begin || aTdx >= outer_end)) continue;
            // optional strided sample (every outer_stride-th element of the whole mesh)
            if (outer_stride > 1 && aTdx % outer_stride != 0) continue;
            const bool mine = (aTdx >= lo && aTdx < hi);
            if (!(mine && triLabel[aTdx])) continue;
            const long* aAdx;
            if (is_DG) { for (int j = 0; j < 3; j++) aDG[j] = 3L * aTdx + j; aAdx = aDG; }
            else aAdx = &tri[3 * aTdx];
            initElement(m, aTdx, aT);
            queue.push(aTdx);
            std::fill(visited.begin(), visited.end(), 0);  // O(nE) reset, src/Cassemble.cpp:597
            bool first = c.singular;
            while (!queue.empty()) {
                int sTdx = queue.front();
                queue.pop();
                int start = m.xadj[sTdx], end = m.xadj[sTdx + 1];
                for (int n = -1; start + n != end; n++) {
                    int bTdx = (n == -1) ? aTdx : m.adjncy[start + n];
                    if (!visited[bTdx]) {
                        if (triLabel[bTdx]) {
                            initElement(m, bTdx, bT);
                            const long* bAdx;
                            if (is_DG) { for (int j = 0; j < 3; j++) bDG[j] = 3L * bTdx + j; bAdx = bDG; }
                            else bAdx = &tri[3 * bTdx];
                            for (int i = 0; i < nb; i++) { tL[i] = 0; tN[i] = 0; tLP[i] = 0; tNP[i] = 0; }
                            int interact = first ? integ_close(aT, bT, q, m, c, tL, tN, tLP, tNP)
                                                 : integ_remote(aT, bT, q, m, c, tL, tN, tLP, tNP);
                            if (interact) {
                                pairs++;
                                if (first) closec++;
                                if (keep_neighbours) { nbrs_t[aTdx].push_back(bTdx); close_t[aTdx].push_back(first ? 1 : 0); }
                                queue.push(bTdx);
                                // src/Cassemble.cpp:721-770 scatter; degree of freedom (node, component) is
                                // od * node + component (interleaved, src/Cassemble.cpp:736-766)
                                // domain decomposition, src/Cassemble.cpp:648-656: weight 1 / zeta
                                const double weight =
                                    nZeta ? 1. / evaluateZeta(zeta_indices, zeta_indptr, aTdx, bTdx) : 1.;
                                auto dof = [&](long node, int comp) -> unsigned long {
                                    if (g_plant & PLANT_BLOCKED_LAYOUT) return (unsigned long)comp * (K / od) + node;
                                    return (unsigned long)(od * node + comp);
                                };
                                for (int a = 0; a < dV * od; a++) {
                                    for (int b = 0; b < dV * od; b++) {
                                        unsigned long Adx;
                                        if (is_DG || vertLabel[aAdx[a / od]] > 0) {
                                            Adx = dof(aAdx[a / od], a % od) * K + dof(aAdx[b / od], b % od);
                                            Ad[Adx] += (g_abs ? std::fabs(tL[dV * od * a + b]) : tL[dV * od * a + b]) * weight;
                                            Adx = dof(aAdx[a / od], a % od) * K + dof(bAdx[b / od], b % od);
                                            Ad[Adx] += (g_abs ? std::fabs(tN[dV * od * a + b]) : -tN[dV * od * a + b]) * weight;
                                        }
                                        if (is_DG || vertLabel[bAdx[b / od]] > 0) {
                                            Adx = dof(bAdx[b / od], b % od) * K + dof(bAdx[a / od], a % od);
                                            Ad[Adx] += (g_abs ? std::fabs(tLP[dV * od * a + b]) : tLP[dV * od * a + b]) * weight;
                                            Adx = dof(bAdx[b / od], b % od) * K + dof(aAdx[a / od], a % od);
                                            Ad[Adx] += (g_abs ? std::fabs(tNP[dV * od * a + b]) : -tNP[dV * od * a + b]) * weight;
                                        }
                                    }
                                }
                            } else {
                                probes++;
                            }
                        } else {
                            queue.push(bTdx);  // label-0 elements are traversed only (src/Cassemble.cpp:773-775)
                        }
                    }
                    visited[bTdx] = 1;
                }
                first = false;
            }
        }
        parts[tid].assign(Ad.begin(), Ad.end());
    }
    // merge + arma::sp_mat(add duplicates) (src/Cassemble.cpp:805-851): sort keys, sum duplicates
    std::vector<std::pair<unsigned long, double>> all;
    for (auto& p : parts) all.insert(all.end(), p.begin(), p.end());
    std::stable_sort(all.begin(), all.end(),
                     [](const std::pair<unsigned long, double>& a, const std::pair<unsigned long, double>& b) { return a.first < b.first; });
    res->rowptr.assign((size_t)K + 1, 0);
    size_t i = 0;
    while (i < all.size()) {
        unsigned long key = all[i].first;
        double v = 0;
        while (i < all.size() && all[i].first == key) { v += all[i].second; i++; }
        if (v != 0.0) {  // Armadillo's check_for_zeros drops exact zeros
            long r = (long)(key / (unsigned long)K), cidx = (long)(key % (unsigned long)K);
            res->colidx.push_back(cidx);
            res->values.push_back(v);
            res->rowptr[r + 1]++;
        }
    }
    for (long r = 0; r < K; r++) res->rowptr[r + 1] += res->rowptr[r];
    auto t1 = std::chrono::steady_clock::now();
    res->seconds = std::chrono::duration<double>(t1 - t0).count();
    res->n_pairs = pairs; res->n_probes = probes; res->n_close = closec;
    if (keep_neighbours) {
        res->nbr_ptr.assign((size_t)nE + 1, 0);
        for (int e = 0; e < nE; e++) {
            res->nbr_ptr[e + 1] = res->nbr_ptr[e] + (long)nbrs_t[e].size();
            res->nbr.insert(res->nbr.end(), nbrs_t[e].begin(), nbrs_t[e].end());
            res->nbr_close.insert(res->nbr_close.end(), close_t[e].begin(), close_t[e].end());
        }
    }
    return res;
}

long oracle_result_nnz(const OracleResult* r) { return (long)r->values.size(); }
long oracle_result_npairs(const OracleResult* r) { return r->n_pairs; }
long oracle_result_nprobes(const OracleResult* r) { return r->n_probes; }
long oracle_result_nclose(const OracleResult* r) { return r->n_close; }
double oracle_result_seconds(const OracleResult* r) { return r->seconds; }
void oracle_result_csr(const OracleResult* r, long* rowptr, long* colidx, double* values) {
    std::memcpy(rowptr, r->rowptr.data(), sizeof(long) * r->rowptr.size());
    std::memcpy(colidx, r->colidx.data(), sizeof(long) * r->colidx.size());
    std::memcpy(values, r->values.data(), sizeof(double) * r->values.size());
}
long oracle_result_nbr_total(const OracleResult* r) { return (long)r->nbr.size(); }
void oracle_result_neighbours(const OracleResult* r, long* ptr, int* nbr, unsigned char* isclose) {
    std::memcpy(ptr, r->nbr_ptr.data(), sizeof(long) * r->nbr_ptr.size());
    std::memcpy(nbr, r->nbr.data(), sizeof(int) * r->nbr.size());
    std::memcpy(isclose, r->nbr_close.data(), r->nbr_close.size());
}
void oracle_result_free(OracleResult* r) { delete r; }

// Dual graph (for tests): writes xadj[nE+1]; adjncy may be null to query the size.
long oracle_dual_graph(const long* tri, int nE, int nV, int* xadj, int* adjncy) {
    Mesh m; m.tri = tri; m.nE = nE; m.nV = nV;
    dualGraph(m);
    if (xadj) std::memcpy(xadj, m.xadj.data(), sizeof(int) * m.xadj.size());
    if (adjncy) std::memcpy(adjncy, m.adjncy.data(), sizeof(int) * m.adjncy.size());
    return (long)m.adjncy.size();
}

// src/integration.cpp:1775-1894 exported method_retriangulate
int oracle_method_retriangulate(const double* xc, const double* TE, double sqdelta, double* list, int caps) {
    return method_retriangulate(xc, TE, sqdelta, list, caps);
}
int oracle_toRef(const double* E, const double* x, double* p) { return toRef(E, x, p); }
void oracle_toPhys(const double* E, const double* p, double* x) { toPhys(E, p, x); }
int oracle_solve2x2(const double* A, const double* b, double* x) { return solve2x2(A, b, x); }

// Kernel point value (for closed-form tests)
int oracle_kernel(const char* kernel, const double* x, const double* y, double sqdelta, double fractional_s, double* out) {
    Mesh m; m.sqdelta = sqdelta; m.delta = std::sqrt(sqdelta); m.fractional_s = fractional_s;
    KernelFn k = lookupKernel(kernel);
    if (!k) return -1;
    k(x, 1, y, 1, m, out);
    return 0;
}

// One pair through one integration routine (tests).  method: 0 retriangulate, 1 baryCenter,
// 2 fractional, 3 weakSingular.  Blocks are (3*outdim)^2 each.
int oracle_integrate_pair(int method, const long* tri, const double* verts, int nV, int aTdx, int bTdx,
                          const double* Px, int nPx, const double* dx, const double* Py, int nPy, const double* dy,
                          double sqdelta, const char* kernel, int caps, int outdim, const double* Pg, int degree,
                          const double* dg, double fractional_s, double* tL, double* tN, double* tLP, double* tNP) {
    Mesh m; m.tri = tri; m.verts = verts; m.nV = nV; m.outdim = outdim;
    m.sqdelta = sqdelta; m.delta = std::sqrt(sqdelta); m.fractional_s = fractional_s;
    static const long lab[1] = {1};
    (void)lab;
    std::vector<long> labels(std::max(aTdx, bTdx) + 1, 1);
    m.triLabel = labels.data();
    Quad q; q.Px = Px; q.dx = dx; q.Py = Py; q.dy = dy; q.nPx = nPx; q.nPy = nPy;
    q.Pg = Pg; q.dg = dg; q.degree = degree; q.nPg = degree > 0 ? (int)std::pow(degree, 4) : 0;
    if (q.nPg) initQuadrule(q, fractional_s);
    q.psix.resize(3 * nPx); q.psiy.resize(3 * nPy);
    for (int h = 0; h < nPx; h++) basis(&Px[2 * h], &q.psix[3 * h]);
    for (int h = 0; h < nPy; h++) basis(&Py[2 * h], &q.psiy[3 * h]);
    Conf c; c.kernel = lookupKernel(kernel); c.caps = caps != 0;
    if (!c.kernel) return -1;
    IntegrateFn f = integrator(method);
    if (!f) return -1;
    Elem aT, bT;
    initElement(m, aTdx, aT);
    initElement(m, bTdx, bT);
    int nb = 9 * outdim * outdim;
    for (int i = 0; i < nb; i++) { tL[i] = 0; tN[i] = 0; tLP[i] = 0; tNP[i] = 0; }
    return f(aT, bT, q, m, c, tL, tN, tLP, tNP);
}

// Duffy tables (tests): case 0 vertex (2 maps), 1 edge (5), 2 identical (6).
// alpha/alphaC: [nT][nPg][4]; detFrac/detWeak: [nT][nPg]
int oracle_quadrule(const double* Pg, int degree, double s, int cs, double* alpha, double* alphaC, double* detFrac,
                    double* detWeak) {
    Quad q; q.Pg = Pg; q.degree = degree; q.nPg = (int)std::pow(degree, 4);
    initQuadrule(q, s);
    std::memcpy(alpha, q.alpha[cs].data(), sizeof(double) * q.alpha[cs].size());
    std::memcpy(alphaC, q.alphaC[cs].data(), sizeof(double) * q.alphaC[cs].size());
    std::memcpy(detFrac, q.detFrac[cs].data(), sizeof(double) * q.detFrac[cs].size());
    std::memcpy(detWeak, q.detWeak[cs].data(), sizeof(double) * q.detWeak[cs].size());
    return ntraffo(cs);
}

// src/Cassemble.cpp:857-928 par_forcing (+ compute_f :149-164)
int oracle_par_forcing(int K, const long* tri, const long* triLabel, const double* verts, const long* vertLabel,
                       int nE, const double* Px, int nPx, const double* dx, int is_DG, const char* force, int outdim,
                       double* fd, const long* zeta_indices, const long* zeta_indptr, long nZeta) {
    ForceFn f = lookupForce(force);
    if (!f) return -1;
    std::vector<double> psix(3 * nPx);
    for (int h = 0; h < nPx; h++) basis(&Px[2 * h], &psix[3 * h]);
    for (int i = 0; i < K; i++) fd[i] = 0;
#pragma omp parallel for
    for (int aTdx = 0; aTdx < nE; aTdx++) {
        if (triLabel[aTdx] <= 0) continue;
        long aDG[3];
        const long* aAdx;
        if (is_DG) { for (int j = 0; j < 3; j++) aDG[j] = 3L * aTdx + j; aAdx = aDG; }
        else aAdx = &tri[3 * aTdx];
        double E[6];
        for (int k = 0; k < 3; k++) { E[2 * k] = verts[2 * tri[3 * aTdx + k]]; E[2 * k + 1] = verts[2 * tri[3 * aTdx + k] + 1]; }
        double det = absDet(E);
        double termf[6] = {0, 0, 0, 0, 0, 0}, x[2], fv[2];
        for (int a = 0; a < 3 * outdim; a++)
            for (int i = 0; i < nPx; i++) {
                toPhys(E, &Px[2 * i], x);
                f(x, fv);
                termf[a] += psix[3 * i + a / outdim] * fv[a % outdim] * det * dx[i];
            }
        // domain decomposition, src/Cassemble.cpp:906-914: weight 1 / zeta(aT, aT)
        const double weight = nZeta ? 1. / evaluateZeta(zeta_indices, zeta_indptr, aTdx, aTdx) : 1.;
        for (int a = 0; a < 3 * outdim; a++)
            if (is_DG || vertLabel[aAdx[a / outdim]] > 0) {
#pragma omp atomic update
                fd[outdim * aAdx[a / outdim] + a % outdim] += termf[a] * weight;
            }
    }
    return 0;
}

// src/Cassemble.cpp:166-246 par_evaluateMass (vd must be zero on entry; J elements are visited)
int oracle_par_evaluateMass(double* vd, const double* ud, const long* tri, const long* triLabel, const double* verts,
                            const long* vertLabel, int J, int nP, const double* P, const double* dx, int outdim,
                            int is_DG) {
    std::vector<double> psi(3 * nP);
    for (int k = 0; k < nP; k++) {
        double t[3];
        basis(&P[2 * k], t);
        for (int kk = 0; kk < 3; kk++) psi[nP * kk + k] = t[kk];
    }
#pragma omp parallel for
    for (int aTdx = 0; aTdx < J; aTdx++) {
        if (triLabel[aTdx] <= 0) continue;
        long aDG[3];
        const long* aAdx;
        if (is_DG) { for (int j = 0; j < 3; j++) aDG[j] = 3L * aTdx + j; aAdx = aDG; }
        else aAdx = &tri[3 * aTdx];
        double E[6];
        for (int k = 0; k < 3; k++) { E[2 * k] = verts[2 * tri[3 * aTdx + k]]; E[2 * k + 1] = verts[2 * tri[3 * aTdx + k] + 1]; }
        double det = absDet(E);
        for (int a = 0; a < 3; a++) {
            if (!(is_DG || vertLabel[aAdx[a]] > 0)) continue;
            for (int ao = 0; ao < outdim; ao++)
                for (int b = 0; b < 3; b++) {
                    if (!(is_DG || vertLabel[aAdx[b]] > 0)) continue;
                    for (int bo = 0; bo < outdim; bo++)
                        for (int j = 0; j < nP; j++) {
#pragma omp atomic update
                            vd[outdim * aAdx[a] + ao] += psi[nP * a + j] * psi[nP * b + j] * det * dx[j] * ud[outdim * aAdx[b] + bo];
                        }
                }
        }
    }
    return 0;
}

}  // extern "C"
