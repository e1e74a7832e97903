// Device-side geometry, kernels and basis functions of the nonlocal assembly (gfx950).
//
// The library is compiled with -ffp-contract=off: the interaction predicates below decide
// which element pairs interact (neighbour lists, sparsity pattern) and must round exactly as
// the reference's FP64 CPU code, whose x86-64 build has no fused multiply-add.  Sums inside
// the quadrature loops opt back into contraction locally where the result is only compared
// to a tolerance.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define NLFEM_DEV __device__ __forceinline__

namespace nlfem {

// Element record, one 64-byte line per element: vertex coordinates (x0,y0,x1,y1,x2,y2),
// |det| of the affine map and the element label.
struct alignas(64) ElemRec {
    double E[6];
    double absDet;
    int label;
    int pad;
};

enum KernelId : int { KER_CONSTANT = 0, KER_PARABOLA = 1, KER_FRACTIONAL = 2, KER_MICRO = 3, KER_MICRO_FIELD = 4 };
enum MethodId : int { M_RETRIANGULATE = 0, M_BARYCENTER = 1, M_FRACTIONAL = 2, M_WEAKSINGULAR = 3 };

// Pair classes stored in bits 29..30 of a neighbour-list entry.
constexpr int CLS_SHIFT = 29;
constexpr uint32_t IDX_MASK = (1u << CLS_SHIFT) - 1u;
constexpr uint32_t CLS_CUT = 0u;       // general pair: retriangulate per outer point
constexpr uint32_t CLS_INTERIOR = 1u;  // inner element inside every ball B(x_k, delta) (with margin)
constexpr uint32_t CLS_CLOSE = 2u;     // first BFS layer of a singular kernel (tensor Gauss)

// src/mathhelpers.cpp:362-368
NLFEM_DEV double absolute(double v) { return v < 0 ? -v : v; }

// src/mathhelpers.cpp:131-147 signDet
NLFEM_DEV double signDet(const double* E) {
    double m00 = E[2] - E[0], m01 = E[4] - E[0];
    double m10 = E[3] - E[1], m11 = E[5] - E[1];
    double det = m00 * m11 - m01 * m10;
    return det > 0 ? 1. : (det < 0 ? -1. : 0.0);
}

// src/mathhelpers.cpp:101-109 absDet
NLFEM_DEV double absDet(const double* E) {
    double m00 = E[2] - E[0], m01 = E[4] - E[0];
    double m10 = E[3] - E[1], m11 = E[5] - E[1];
    return absolute(m00 * m11 - m01 * m10);
}

// src/mathhelpers.cpp:217-225 toPhys: ((p0*(E1-E0)) + p1*(E2-E0)) + E0
NLFEM_DEV void toPhys(const double* E, double p0, double p1, double* out) {
    out[0] = (p0 * (E[2] - E[0]) + p1 * (E[4] - E[0])) + E[0];
    out[1] = (p0 * (E[3] - E[1]) + p1 * (E[5] - E[1])) + E[1];
}

// src/mathhelpers.cpp:41-71 solve2x2 (pivoted LU) on M = [E1-E0 | E2-E0], used by toRef
// (src/mathhelpers.cpp:227-240).  Returns false when the reference would abort.
NLFEM_DEV bool solve2x2(const double* A, const double* b, double* x) {
    int dx0 = 0, dx1 = 1;
    if (absolute(A[0]) < absolute(A[2])) { dx0 = 1; dx1 = 0; }
    if (absolute(A[2 * dx0]) < 1e-15) return false;
    double l = A[2 * dx1] / A[2 * dx0];
    double u = A[2 * dx1 + 1] - l * A[2 * dx0 + 1];
    if (u == 0) return false;
    x[1] = (b[dx1] - l * b[dx0]) / u;
    x[0] = (b[dx0] - A[2 * dx0 + 1] * x[1]) / A[2 * dx0];
    return true;
}

NLFEM_DEV bool toRef(const double* E, const double* x, double* p) {
    double M[4] = {E[2] - E[0], E[4] - E[0], E[3] - E[1], E[5] - E[1]};
    double b[2] = {x[0] - E[0], x[1] - E[1]};
    return solve2x2(M, b, p);
}

// src/mathhelpers.cpp:272-279 vec_sqL2dist
NLFEM_DEV double sqdist(double x0, double x1, double y0, double y1) {
    double d0 = x0 - y0, d1 = x1 - y1;
    return d0 * d0 + d1 * d1;
}

// Inverse affine map of an element, for evaluating basis functions at physical points:
// p = Minv * (y - E0).  Rounds differently from the pivoted solve of toRef (about 1 ulp).
struct InvMap {
    double e0, e1, i00, i01, i10, i11;
};
NLFEM_DEV InvMap invmap(const double* E) {
    double m00 = E[2] - E[0], m01 = E[4] - E[0];
    double m10 = E[3] - E[1], m11 = E[5] - E[1];
    double det = m00 * m11 - m01 * m10;
    double r = 1.0 / det;
    InvMap im;
    im.e0 = E[0]; im.e1 = E[1];
    im.i00 = m11 * r; im.i01 = -m01 * r;
    im.i10 = -m10 * r; im.i11 = m00 * r;
    return im;
}

// Quotient n/d from the correctly rounded reciprocal r = 1/d and one exact residual step:
// q = n r, e = fma(-q, d, n) (exact), result = fma(e, r, q).  With r = RN(1/d) this is the
// correctly rounded quotient (Markstein's correction), i.e. it equals IEEE n/d; it replaces a
// full FP64 division (about a dozen instructions) by three.
NLFEM_DEV double div_by_recip(double n, double d, double r) {
    const double q = n * r;
    const double e = __builtin_fma(-q, d, n);
    return __builtin_fma(e, r, q);
}

// toRef of src/mathhelpers.cpp:227-240 split into its per-element part (pivot choice, l, u,
// reciprocals of the pivots) and its per-point part, which rounds exactly as solve2x2
// (src/mathhelpers.cpp:41-71): the numerators are formed without fused operations and the
// two divisions are correctly rounded.
struct PivotLU {
    double e0, e1, l, u, a0, a1, ru, ra0;  // a0 = A[2 dx0], a1 = A[2 dx0 + 1], ru = 1/u, ra0 = 1/a0
    int dx0;
    NLFEM_DEV void solve(double x0, double x1, double& p0, double& p1) const {
        const double b0 = x0 - e0, b1 = x1 - e1;
        const double bd0 = dx0 ? b1 : b0, bd1 = dx0 ? b0 : b1;
        const double n1 = bd1 - l * bd0;
        p1 = div_by_recip(n1, u, ru);
        const double n0 = bd0 - a1 * p1;
        p0 = div_by_recip(n0, a0, ra0);
    }
};
NLFEM_DEV PivotLU pivot_lu(const double* E) {
    const double A[4] = {E[2] - E[0], E[4] - E[0], E[3] - E[1], E[5] - E[1]};
    PivotLU r;
    r.dx0 = absolute(A[0]) < absolute(A[2]) ? 1 : 0;
    const int dx1 = 1 - r.dx0;
    r.e0 = E[0];
    r.e1 = E[1];
    r.a0 = A[2 * r.dx0];
    r.a1 = A[2 * r.dx0 + 1];
    r.l = A[2 * dx1] / A[2 * r.dx0];
    r.u = A[2 * dx1 + 1] - r.l * A[2 * r.dx0 + 1];
    r.ru = 1.0 / r.u;
    r.ra0 = 1.0 / r.a0;
    return r;
}

// ----------------------------------------------------------------------------------------
// method_retriangulate (src/integration.cpp:1393-1512) with placePointOnCap (:1353-1377),
// inTriangle (:1283-1298) and rightNormal (src/mathhelpers.cpp:74-78), restated as a
// stream: the polygon points R_0, R_1, ... are produced in the reference's order and every
// fan triangle (R_0, R_{j-1}, R_j), j >= 2, is handed to `emit` as soon as R_j exists, so no
// point array is kept (no dynamically indexed private memory).  Returns the number of
// polygon points; the reference's return value is max(points - 2, 0) triangles.
// With EARLY, returns as soon as 3 points exist (the interaction predicate only needs to
// know whether the fan is non-empty; points are never removed).
// ----------------------------------------------------------------------------------------
struct NoEmit {
    NLFEM_DEV void operator()(double, double, double, double, double, double) const {}
};

template <bool EARLY, class Emit>
NLFEM_DEV int retriangulate(double xc0, double xc1, const double* TE, double sqdelta, bool caps, Emit& emit) {
    const double orientation = -signDet(TE);
    double na0 = (TE[3] - TE[1]) * orientation, na1 = (TE[0] - TE[2]) * orientation;
    double nb0 = (TE[5] - TE[3]) * orientation, nb1 = (TE[2] - TE[4]) * orientation;
    double nc0 = (TE[1] - TE[5]) * orientation, nc1 = (TE[4] - TE[0]) * orientation;
    int n = 0;
    bool onEdge = false, firstOnVertex = true;
    double f0 = 0, f1 = 0, p0 = 0, p1 = 0;

    auto append = [&](double q0, double q1) {
        if (n == 0) { f0 = q0; f1 = q1; }
        else if (n >= 2) emit(f0, f1, p0, p1, q0, q1);
        p0 = q0; p1 = q1;
        n++;
    };
    // placePointOnCap(prev, ycur): point x + delta * n/|n| with n the right normal of
    // (ycur -> prev) scaled by the orientation; kept if it lies in the triangle.
    auto cap = [&](double c0, double c1) {
        double s0 = (p1 - c1) * orientation;
        double s1 = (c0 - p0) * orientation;
        double scaling = sqrt(sqdelta / (s0 * s0 + s1 * s1));
        s0 = s0 * scaling;
        s1 = s1 * scaling;
        double y0 = xc0 + s0, y1 = xc1 + s1;
        bool a = (na0 * (y0 - TE[0]) + na1 * (y1 - TE[1])) >= 0;
        bool b = (nb0 * (y0 - TE[2]) + nb1 * (y1 - TE[3])) >= 0;
        bool c = (nc0 * (y0 - TE[4]) + nc1 * (y1 - TE[5])) >= 0;
        if (a && b && c) append(y0, y1);
    };

#pragma unroll
    for (int k = 0; k < 3; k++) {
        const double P0 = TE[2 * k], P1 = TE[2 * k + 1];
        const double Q0 = TE[(2 * k + 2) % 6], Q1 = TE[(2 * k + 3) % 6];
        const double a0 = Q0 - xc0, a1 = Q1 - xc1;
        const double b0 = P0 - Q0, b1 = P1 - Q1;
        if (sqdist(P0, P1, xc0, xc1) <= sqdelta) {
            append(P0, P1);
            onEdge = false;
            if (EARLY && n >= 3) return n;
        }
        const double ab = a0 * b0 + a1 * b1;
        const double aa = a0 * a0 + a1 * a1;
        const double bb = b0 * b0 + b1 * b1;
        const double v = ab * ab - (aa - sqdelta) * bb;
        if (v >= 0) {
            // the three divisions by bb use one correctly rounded reciprocal and a Markstein
            // correction each, which rounds as the division (bb > 0: elements are not degenerate)
            const double rbb = 1.0 / bb;
            const double term1 = div_by_recip(-ab, bb, rbb);
            const double term2 = div_by_recip(sqrt(v), bb, rbb);
            double lam1, lam2;
            if (term1 > 0) {
                lam1 = term1 + term2;
                lam2 = div_by_recip(1 / lam1 * (aa - sqdelta), bb, rbb);
            } else {
                lam2 = term1 - term2;
                lam1 = div_by_recip(1 / lam2 * (aa - sqdelta), bb, rbb);
            }
            if ((0 <= lam1) && (lam1 <= 1)) {
                const double y0 = lam1 * b0 + Q0, y1 = lam1 * b1 + Q1;
                firstOnVertex = firstOnVertex && (n != 0);
                if (onEdge && caps) cap(y0, y1);
                append(y0, y1);
                onEdge = true;
                if (EARLY && n >= 3) return n;
            }
            const double dl = lam1 - lam2;
            if ((0 <= lam2) && (lam2 <= 1) && (dl * dl > 0)) {
                const double y0 = lam2 * b0 + Q0, y1 = lam2 * b1 + Q1;
                firstOnVertex = firstOnVertex && (n != 0);
                if (onEdge && caps) cap(y0, y1);
                append(y0, y1);
                onEdge = true;
                if (EARLY && n >= 3) return n;
            }
        }
    }
    if (onEdge && (!firstOnVertex && n > 1) && caps) cap(f0, f1);
    return n;
}

// src/integration.cpp:1262-1280 method_baryCenter predicate (+ baryCenter
// src/mathhelpers.cpp:184-196): ((E0 + E1) + E2) / 3 per coordinate.
NLFEM_DEV bool bary_interacts(double xc0, double xc1, const double* TE, double sqdelta) {
    double b0 = ((0.0 + TE[0]) + TE[2]) + TE[4];
    double b1 = ((0.0 + TE[1]) + TE[3]) + TE[5];
    b0 = b0 / 3.0;
    b1 = b1 / 3.0;
    return !(sqdist(xc0, xc1, b0, b1) > sqdelta);
}

// ----------------------------------------------------------------------------------------
// Kernels (src/model.cpp).  All in-scope kernels are symmetric, k(x,y) = k(y,x).
// KernelParams holds the per-assembly constants, computed once on the host.
// ----------------------------------------------------------------------------------------
struct KernelParams {
    double sqdelta;
    double c;         // kernel constant (see host code for each kernel)
    double exponent;  // fractional: 2 + 2s
};

// Scalar kernel value k(x,y) given z = x - y.
template <int KER>
NLFEM_DEV double kernel_scalar(double z0, double z1, const KernelParams& kp) {
    if constexpr (KER == KER_CONSTANT) {
        return kp.c;                                            // src/model.cpp:25-28
    } else if constexpr (KER == KER_PARABOLA) {
        return kp.c * (kp.sqdelta - (z0 * z0 + z1 * z1));       // src/model.cpp:52-60
    } else if constexpr (KER == KER_FRACTIONAL) {
        return kp.c * (1.0 / pow(sqrt(z0 * z0 + z1 * z1), kp.exponent));  // src/model.cpp:126-138
    } else if constexpr (KER == KER_MICRO) {
        return kp.c * (1.0 / sqrt(z0 * z0 + z1 * z1));          // src/model.cpp:116-124
    } else {
        return 0.0;
    }
}

// Matrix-valued microelastic kernel (src/model.cpp:140-153), row-major [k00,k01,k10,k11].
NLFEM_DEV void kernel_field(double z0, double z1, const KernelParams& kp, double* k) {
    double r = sqrt(z0 * z0 + z1 * z1);
    double den = 1.0 / (r * r * r);
    double cd = kp.c * den;
    k[0] = cd * z0 * z0;
    k[1] = cd * z0 * z1;
    k[2] = cd * z1 * z0;
    k[3] = cd * z1 * z1;
}

}  // namespace nlfem
