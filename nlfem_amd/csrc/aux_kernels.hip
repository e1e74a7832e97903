// Support kernels around the stiffness assembly: load vector (par_forcing), mass-matrix
// action (par_evaluateMass) and the batched geometry helpers behind the reference's debug
// entry points (method_retriangulate, toRef, toPhys, solve2x2).
#include <hip/hip_runtime.h>
#include <stdexcept>
#include <string>
#include <vector>

#include "nlfem_device.hpp"

namespace nlfem {

#define HIPCHECK_AUX(x)                                                                                      \
    do {                                                                                                     \
        hipError_t e_ = (x);                                                                                 \
        if (e_ != hipSuccess)                                                                                \
            throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) + " (" #x ")");        \
    } while (0)

enum ForceId : int { F_LINEAR = 0, F_CONSTANT = 1, F_LINEAR_FIELD = 2 };

// src/model.cpp:240-242 f_linear, :201-203 f_constant, :204-210 fField_linear
NLFEM_DEV void forcing(int f, const double* x, double* out) {
    if (f == F_LINEAR) out[0] = -2. * (x[1] + 1.);
    else if (f == F_CONSTANT) out[0] = 1.0;
    else {
        const double c = M_PI / 2.0;
        out[0] = -c * (1.0 + 2 * x[0]);
        out[1] = -c * x[1];
    }
}

struct AuxQuad {
    int nP;
    double P[64], w[32];
};

// par_forcing (src/Cassemble.cpp:857-928) with compute_f (:149-164): one thread per element,
// FP64 atomics into the load vector.
// weight: per-element domain-decomposition weight 1 / |Z_aT| (src/Cassemble.cpp:906-914), or null.
__global__ void k_forcing(int nE, const long* tri, const long* lab, const double* verts, const long* vlab,
                          int is_DG, int od, int f, AuxQuad q, const double* weight, double* fd) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nE || lab[e] <= 0) return;
    double E[6];
    long node[3];
    for (int j = 0; j < 3; j++) {
        long v = tri[3 * (long)e + j];
        E[2 * j] = verts[2 * v];
        E[2 * j + 1] = verts[2 * v + 1];
        node[j] = is_DG ? 3L * e + j : v;
    }
    const double det = absDet(E);
    double termf[6] = {0, 0, 0, 0, 0, 0};
    for (int a = 0; a < 3 * od; a++)
        for (int i = 0; i < q.nP; i++) {
            double x[2], fv[2];
            toPhys(E, q.P[2 * i], q.P[2 * i + 1], x);
            forcing(f, x, fv);
            const double p0 = q.P[2 * i], p1 = q.P[2 * i + 1];
            const double psi[3] = {1 - p0 - p1, p0, p1};
            termf[a] += psi[a / od] * fv[a % od] * det * q.w[i];
        }
    const double w = weight ? weight[e] : 1.0;
    for (int a = 0; a < 3 * od; a++)
        if (is_DG || vlab[node[a / od]] > 0) atomicAdd(&fd[od * node[a / od] + a % od], termf[a] * w);
}

// par_evaluateMass (src/Cassemble.cpp:166-246): v = M u over the first J elements.
__global__ void k_mass(int J, const long* tri, const long* lab, const double* verts, const long* vlab, int is_DG,
                       int od, AuxQuad q, const double* ud, double* vd) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= J || lab[e] <= 0) return;
    double E[6];
    long node[3];
    for (int j = 0; j < 3; j++) {
        long v = tri[3 * (long)e + j];
        E[2 * j] = verts[2 * v];
        E[2 * j + 1] = verts[2 * v + 1];
        node[j] = is_DG ? 3L * e + j : v;
    }
    const double det = absDet(E);
    for (int a = 0; a < 3; a++) {
        if (!(is_DG || vlab[node[a]] > 0)) continue;
        for (int ao = 0; ao < od; ao++) {
            double acc = 0;
            for (int b = 0; b < 3; b++) {
                if (!(is_DG || vlab[node[b]] > 0)) continue;
                for (int bo = 0; bo < od; bo++)
                    for (int j = 0; j < q.nP; j++) {
                        const double p0 = q.P[2 * j], p1 = q.P[2 * j + 1];
                        const double psi[3] = {1 - p0 - p1, p0, p1};
                        acc += psi[a] * psi[b] * det * q.w[j] * ud[od * node[b] + bo];
                    }
            }
            atomicAdd(&vd[od * node[a] + ao], acc);
        }
    }
}

// Batched method_retriangulate: the reference's output layout (triangles of 3 points, 6 doubles
// each, at most 9*3*2 doubles) and return value (number of triangles).
struct ListEmit {
    double* out;
    int nt;
    NLFEM_DEV void operator()(double r0, double r1, double s0, double s1, double t0, double t1) {
        if (nt < 9) {
            double* o = out + 6 * nt;
            o[0] = r0; o[1] = r1; o[2] = s0; o[3] = s1; o[4] = t0; o[5] = t1;
        }
        nt++;
    }
};
__global__ void k_retriangulate(int n, const double* x, const double* TE, double sqdelta, int caps, double* lists,
                                int* ntri) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ListEmit em{lists + 54 * (long)i, 0};
    int np = retriangulate<false>(x[2 * i], x[2 * i + 1], TE + 6 * (long)i, sqdelta, caps != 0, em);
    ntri[i] = np >= 3 ? np - 2 : 0;
}

// Batched toRef / toPhys / solve2x2 (src/mathhelpers.cpp:41-71,210-240).
__global__ void k_geom(int n, int op, const double* E, const double* p, double* out, int* ok) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (op == 0) {
        ok[i] = toRef(E + 6 * (long)i, p + 2 * (long)i, out + 2 * (long)i) ? 1 : 0;
    } else if (op == 1) {
        toPhys(E + 6 * (long)i, p[2 * i], p[2 * i + 1], out + 2 * (long)i);
        ok[i] = 1;
    } else {
        ok[i] = solve2x2(E + 4 * (long)i, p + 2 * (long)i, out + 2 * (long)i) ? 1 : 0;
    }
}

// ---------------------------------------------------------------- host wrappers
template <typename T>
static T* up(const T* h, size_t n, hipStream_t st) {
    T* d;
    HIPCHECK_AUX(hipMalloc(&d, sizeof(T) * (n ? n : 1)));
    if (n) HIPCHECK_AUX(hipMemcpyAsync(d, h, sizeof(T) * n, hipMemcpyHostToDevice, st));
    return d;
}

static AuxQuad aux_quad(const double* P, const double* w, int nP) {
    if (nP > 32) throw std::runtime_error("quadrature rule with more than 32 points");
    AuxQuad q;
    q.nP = nP;
    for (int i = 0; i < 2 * nP; i++) q.P[i] = P[i];
    for (int i = 0; i < nP; i++) q.w[i] = w[i];
    return q;
}

void run_forcing(int K, const long* tri, const long* lab, const double* verts, const long* vlab, int nE, int nV,
                 const double* Px, int nPx, const double* dx, int is_DG, int od, int f, const double* weight,
                 double* fd_host) {
    hipStream_t st;
    HIPCHECK_AUX(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    long* d_tri = up(tri, 3 * (size_t)nE, st);
    long* d_lab = up(lab, (size_t)nE, st);
    double* d_verts = up(verts, 2 * (size_t)nV, st);
    long* d_vlab = up(vlab, (size_t)nV, st);
    double* d_w = weight ? up(weight, (size_t)nE, st) : nullptr;
    double* d_fd;
    HIPCHECK_AUX(hipMalloc(&d_fd, sizeof(double) * (size_t)K));
    HIPCHECK_AUX(hipMemsetAsync(d_fd, 0, sizeof(double) * (size_t)K, st));
    k_forcing<<<(nE + 255) / 256, 256, 0, st>>>(nE, d_tri, d_lab, d_verts, d_vlab, is_DG, od, f,
                                                aux_quad(Px, dx, nPx), d_w, d_fd);
    HIPCHECK_AUX(hipGetLastError());
    HIPCHECK_AUX(hipMemcpyAsync(fd_host, d_fd, sizeof(double) * (size_t)K, hipMemcpyDeviceToHost, st));
    HIPCHECK_AUX(hipStreamSynchronize(st));
    (void)hipFree(d_tri); (void)hipFree(d_lab); (void)hipFree(d_verts); (void)hipFree(d_vlab); (void)hipFree(d_fd);
    if (d_w) (void)hipFree(d_w);
    (void)hipStreamDestroy(st);
}

void run_mass(int K, const long* tri, const long* lab, const double* verts, const long* vlab, int J, int nE, int nV,
              const double* P, int nP, const double* dx, int is_DG, int od, const double* ud, double* vd_host) {
    hipStream_t st;
    HIPCHECK_AUX(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    long* d_tri = up(tri, 3 * (size_t)nE, st);
    long* d_lab = up(lab, (size_t)nE, st);
    double* d_verts = up(verts, 2 * (size_t)nV, st);
    long* d_vlab = up(vlab, (size_t)nV, st);
    double* d_ud = up(ud, (size_t)K, st);
    double* d_vd = up(vd_host, (size_t)K, st);
    k_mass<<<(J + 255) / 256, 256, 0, st>>>(J, d_tri, d_lab, d_verts, d_vlab, is_DG, od, aux_quad(P, dx, nP), d_ud, d_vd);
    HIPCHECK_AUX(hipGetLastError());
    HIPCHECK_AUX(hipMemcpyAsync(vd_host, d_vd, sizeof(double) * (size_t)K, hipMemcpyDeviceToHost, st));
    HIPCHECK_AUX(hipStreamSynchronize(st));
    (void)hipFree(d_tri); (void)hipFree(d_lab); (void)hipFree(d_verts); (void)hipFree(d_vlab);
    (void)hipFree(d_ud); (void)hipFree(d_vd);
    (void)hipStreamDestroy(st);
}

void run_retriangulate(int n, const double* x, const double* TE, double sqdelta, int caps, double* lists, int* ntri) {
    hipStream_t st;
    HIPCHECK_AUX(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    double* d_x = up(x, 2 * (size_t)n, st);
    double* d_TE = up(TE, 6 * (size_t)n, st);
    double* d_l;
    int* d_n;
    HIPCHECK_AUX(hipMalloc(&d_l, sizeof(double) * 54 * (size_t)(n ? n : 1)));
    HIPCHECK_AUX(hipMalloc(&d_n, sizeof(int) * (size_t)(n ? n : 1)));
    HIPCHECK_AUX(hipMemsetAsync(d_l, 0, sizeof(double) * 54 * (size_t)(n ? n : 1), st));
    if (n) k_retriangulate<<<(n + 255) / 256, 256, 0, st>>>(n, d_x, d_TE, sqdelta, caps, d_l, d_n);
    HIPCHECK_AUX(hipGetLastError());
    HIPCHECK_AUX(hipMemcpyAsync(lists, d_l, sizeof(double) * 54 * (size_t)n, hipMemcpyDeviceToHost, st));
    HIPCHECK_AUX(hipMemcpyAsync(ntri, d_n, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, st));
    HIPCHECK_AUX(hipStreamSynchronize(st));
    (void)hipFree(d_x); (void)hipFree(d_TE); (void)hipFree(d_l); (void)hipFree(d_n);
    (void)hipStreamDestroy(st);
}

void run_geom(int n, int op, const double* E, const double* p, double* out, int* ok) {
    hipStream_t st;
    HIPCHECK_AUX(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    const int ne = op == 2 ? 4 : 6;
    double* d_E = up(E, (size_t)ne * n, st);
    double* d_p = up(p, 2 * (size_t)n, st);
    double* d_o;
    int* d_ok;
    HIPCHECK_AUX(hipMalloc(&d_o, sizeof(double) * 2 * (size_t)(n ? n : 1)));
    HIPCHECK_AUX(hipMalloc(&d_ok, sizeof(int) * (size_t)(n ? n : 1)));
    if (n) k_geom<<<(n + 255) / 256, 256, 0, st>>>(n, op, d_E, d_p, d_o, d_ok);
    HIPCHECK_AUX(hipGetLastError());
    HIPCHECK_AUX(hipMemcpyAsync(out, d_o, sizeof(double) * 2 * (size_t)n, hipMemcpyDeviceToHost, st));
    HIPCHECK_AUX(hipMemcpyAsync(ok, d_ok, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, st));
    HIPCHECK_AUX(hipStreamSynchronize(st));
    (void)hipFree(d_E); (void)hipFree(d_p); (void)hipFree(d_o); (void)hipFree(d_ok);
    (void)hipStreamDestroy(st);
}

}  // namespace nlfem
