// C ABI of libnlfem_amd.so (include/nlfem_amd.h) and the C++ drop-in of the reference's
// Cassemble.h entry points (include/Cassemble.h of this repository).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <atomic>
#include <cmath>
#include <memory>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/nlfem_amd.h"
#include "plan.hpp"

namespace nlfem {
void plan_setup(Plan& P, const long* tri64, const long* lab64, const double* verts, const long* vlab64);
void plan_assemble(Plan& P, const Estimates& est);
void plan_free(Plan& P);
void plan_set_surrogate(Plan& P, int nnet, const float* params, const float* norm, float tol, double smin,
                        double smax);
long long plan_surrogate_samples(Plan& P, long long maxn, unsigned stride, unsigned salt, double smin, double smax,
                                 float* feat, float* targ);
void run_forcing(int K, const long* tri, const long* lab, const double* verts, const long* vlab, int nE, int nV,
                 const double* Px, int nPx, const double* dx, int is_DG, int od, int f, const double* weight,
                 double* fd_host);
void plan_set_zeta(Plan& P, const long* indices, const long* indptr, long nZeta);
void run_mass(int K, const long* tri, const long* lab, const double* verts, const long* vlab, int J, int nE, int nV,
              const double* P, int nP, const double* dx, int is_DG, int od, const double* ud, double* vd_host);
void run_retriangulate(int n, const double* x, const double* TE, double sqdelta, int caps, double* lists, int* ntri);
void run_geom(int n, int op, const double* E, const double* p, double* out, int* ok);
namespace mlp {
void mlp_forward(const float* params, const float* norm, const float* X, long long n, float* Y);
void mlp_forward_noalloc(const float* d_params, const float* d_norm, const float* d_X, long long n, float* d_Y,
                         hipStream_t st);
void mlp_grad(const float* params, const float* norm, const float* X, const float* T, long long n, float* grad,
              float* loss);
void mlp_train(float* params, const float* norm, const float* X, const float* T, long long n, int epochs, int batch,
               float lr_max, unsigned seed, float* loss_hist);
}  // namespace mlp
}  // namespace nlfem

using namespace nlfem;

struct nlfem_plan {
    Plan P;
    Estimates est;
};

static thread_local std::string g_err;

static int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define GUARD_BEGIN try {
#define GUARD_END                                                     \
    }                                                                 \
    catch (const std::invalid_argument& e) { return fail(NLFEM_EINVAL, e.what()); } \
    catch (const std::length_error& e) { return fail(NLFEM_ECAPACITY, e.what()); }  \
    catch (const std::exception& e) { return fail(NLFEM_EDEVICE, e.what()); }

extern "C" const char* nlfem_last_error(void) { return g_err.c_str(); }

// lookup_configuration (src/Cassemble.cpp:28-147) for the in-scope kernels and methods.
static int kernel_id(const std::string& s) {
    if (s == "constant") return KER_CONSTANT;
    if (s == "parabola") return KER_PARABOLA;
    if (s == "fractional") return KER_FRACTIONAL;
    if (s == "linearPrototypeMicroelastic") return KER_MICRO;
    if (s == "linearPrototypeMicroelasticField") return KER_MICRO_FIELD;
    return -1;
}
static int method_id(const std::string& s) {
    if (s == "retriangulate") return M_RETRIANGULATE;
    if (s == "baryCenter") return M_BARYCENTER;
    if (s == "fractional") return M_FRACTIONAL;
    if (s == "weakSingular") return M_WEAKSINGULAR;
    return -1;
}
static int force_id(const std::string& s) {
    if (s == "linear") return 0;
    if (s == "constant") return 1;
    if (s == "linearField") return 2;
    return -1;
}

// Input checks of chk_Mesh / chk_Conf / chk_BasisFunction (src/checks.cpp:36-113), raised as
// std::invalid_argument instead of abort().
static void check_config(const PlanConfig& c, const std::string& kernel, const std::string& mr,
                         const std::string& mc) {
    if (c.kernel < 0) throw std::invalid_argument("unknown kernel '" + kernel + "'");
    if (c.method_remote < 0) throw std::invalid_argument("unknown integration method '" + mr + "'");
    if (c.method_close < 0) throw std::invalid_argument("unknown integration method '" + mc + "'");
    if (c.method_remote == M_FRACTIONAL || c.method_remote == M_WEAKSINGULAR)
        throw std::invalid_argument(
            "The fractional and fractional_weakSingular integration routines are meant for close elements only.");
    if ((c.method_close == M_FRACTIONAL || c.method_close == M_WEAKSINGULAR) && c.degree <= 0)
        throw std::invalid_argument("You chose an integration routing for singular kernels, but no quadrature rule for it.");
    if (c.method_close == M_FRACTIONAL && c.is_DG)
        throw std::invalid_argument("The fractional integration routine cannot be used for DG ansatz spaces.");
    if ((c.kernel == KER_MICRO || c.kernel == KER_MICRO_FIELD) && c.fractional_s != -0.5)
        throw std::invalid_argument("The fractional_s corresponding to the linearPrototypeMicroelastic kernels is s=-0.5.");
    if ((c.kernel == KER_MICRO_FIELD) != (c.outdim == 2))
        throw std::invalid_argument("outdim does not match the kernel's output dimension");
    if (c.outdim != 1 && c.outdim != 2) throw std::invalid_argument("outdim must be 1 or 2");
    if (c.is_DG ? (c.outdim * c.nE * 3 != c.K) : (c.outdim * c.nV != c.K))
        throw std::invalid_argument("Matrix dimension does not match #basis functions and output dimension.");
    if (c.q.nPx <= 0 || c.q.nPy <= 0 || c.q.nPx > MAXQ || c.q.nPy > MAXQ)
        throw std::invalid_argument("quadrature rules must have 1..32 points");
    if (!(c.sqdelta > 0)) throw std::invalid_argument("sqdelta must be positive");
    double s0 = 0, s1 = 0;  // chk_BasisFunction: integral of psi_0 = 1/6
    for (int i = 0; i < c.q.nPx; i++) s0 += c.q.dx[i] * c.q.psix[3 * i];
    for (int i = 0; i < c.q.nPy; i++) s1 += c.q.dy[i] * c.q.psiy[3 * i];
    if (std::fabs(s0 - 1.0 / 6.0) >= 1e-8) throw std::invalid_argument("Wrong integral of basis function w.r.t. weights dx.");
    if (std::fabs(s1 - 1.0 / 6.0) >= 1e-8) throw std::invalid_argument("Wrong integral of basis function w.r.t. weights dy.");
}

static nlfem_plan* make_plan(int K, const long* tri, const long* lab, const double* verts, const long* vlab, int nE,
                             int nV, const double* Px, int nPx, const double* dx, const double* Py, int nPy,
                             const double* dy, double sqdelta, int is_DG, const char* kernel, const char* mr,
                             const char* mc, int caps, int outdim, const double* Pg, int degree, const double* dg,
                             double fractional_s, int fullCC, int precision) {
    if (nE <= 0 || nV <= 0) throw std::invalid_argument("empty mesh");
    if (nE >= (1 << 29)) throw std::invalid_argument("more than 2^29 elements");
    nlfem_plan* h = new nlfem_plan();
    PlanConfig& c = h->P.cfg;
    c.K = K; c.nE = nE; c.nV = nV; c.outdim = outdim; c.is_DG = is_DG ? 1 : 0;
    std::string ks = kernel ? kernel : "", ms = mr ? mr : "", cs = mc ? mc : "";
    c.kernel = kernel_id(ks);
    c.method_remote = method_id(ms);
    c.method_close = method_id(cs);
    c.caps = caps ? 1 : 0;
    c.singular = (c.kernel == KER_FRACTIONAL || c.kernel == KER_MICRO || c.kernel == KER_MICRO_FIELD) ? 1 : 0;
    c.fullCC = fullCC ? 1 : 0;
    c.sqdelta = sqdelta;
    c.fractional_s = fractional_s;
    c.degree = degree;
    c.nPg = degree > 0 ? degree * degree * degree * degree : 0;
    c.precision = precision;
    c.q.nPx = nPx;
    c.q.nPy = nPy;
    if (nPx > MAXQ || nPy > MAXQ || nPx <= 0 || nPy <= 0) {
        delete h;
        throw std::invalid_argument("quadrature rules must have 1..32 points");
    }
    for (int i = 0; i < nPx; i++) {
        c.q.Px[2 * i] = Px[2 * i]; c.q.Px[2 * i + 1] = Px[2 * i + 1]; c.q.dx[i] = dx[i];
        c.q.psix[3 * i] = 1 - Px[2 * i] - Px[2 * i + 1]; c.q.psix[3 * i + 1] = Px[2 * i]; c.q.psix[3 * i + 2] = Px[2 * i + 1];
    }
    for (int i = 0; i < nPy; i++) {
        c.q.Py[2 * i] = Py[2 * i]; c.q.Py[2 * i + 1] = Py[2 * i + 1]; c.q.dy[i] = dy[i];
        c.q.psiy[3 * i] = 1 - Py[2 * i] - Py[2 * i + 1]; c.q.psiy[3 * i + 1] = Py[2 * i]; c.q.psiy[3 * i + 2] = Py[2 * i + 1];
    }
    if (c.nPg > 0) {
        if (!Pg || !dg) { delete h; throw std::invalid_argument("tensor Gauss points missing"); }
        c.Pg.assign(Pg, Pg + 4 * (size_t)c.nPg);
        c.dg.assign(dg, dg + (size_t)c.nPg);
    }
    try {
        check_config(c, ks, ms, cs);
    } catch (...) {
        delete h;
        throw;
    }
    // geometric estimates for the BFS / pattern capacities, over the elements that are
    // integrated: label-0 elements are only traversed (src/Cassemble.cpp:630-633,773-775), so
    // they may be degenerate (the reference tutorial's artificial elements join boundary lines
    // to one far vertex, some of them with zero area) and do not size any table
    double hmax = 0, amin = 1e300, asum = 0;
    long nlab = 0;
    for (int e = 0; e < nE; e++) {
        const long* t = tri + 3 * (size_t)e;
        double E[6];
        for (int j = 0; j < 3; j++) {
            if (t[j] < 0 || t[j] >= nV) { delete h; throw std::invalid_argument("element vertex index out of range"); }
            E[2 * j] = verts[2 * t[j]]; E[2 * j + 1] = verts[2 * t[j] + 1];
        }
        if (lab[e] == 0) continue;
        nlab++;
        for (int j = 0; j < 3; j++) {
            int k = (j + 1) % 3;
            double d = std::hypot(E[2 * k] - E[2 * j], E[2 * k + 1] - E[2 * j + 1]);
            hmax = std::max(hmax, d);
        }
        double a = 0.5 * std::fabs((E[2] - E[0]) * (E[5] - E[1]) - (E[4] - E[0]) * (E[3] - E[1]));
        if (!(a > 0)) { delete h; throw std::invalid_argument("degenerate element (zero area)"); }
        amin = std::min(amin, a);
        asum += a;
    }
    if (nlab == 0) { delete h; throw std::invalid_argument("no element with a nonzero label"); }
    h->est.hmax = hmax;
    h->est.amin = amin;
    h->est.amean = asum / nlab;
    try {
        plan_setup(h->P, tri, lab, verts, vlab);
    } catch (...) {
        plan_free(h->P);
        delete h;
        throw;
    }
    return h;
}

extern "C" int nlfem_plan_create(nlfem_plan** out, int K, const long* ptrTriangles, const long* ptrLabelTriangles,
                                 const double* ptrVerts, const long* ptrLabelVerts, int nE, int nV, const double* Px,
                                 int nPx, const double* dx, const double* Py, int nPy, const double* dy, double sqdelta,
                                 int is_DG, const char* kernel, const char* mr, const char* mc, int caps, int outdim,
                                 const double* Pg, int degree, const double* dg, double fractional_s, int fullCC,
                                 int precision) {
    GUARD_BEGIN
    if (!out) return fail(NLFEM_EINVAL, "null output handle");
    *out = make_plan(K, ptrTriangles, ptrLabelTriangles, ptrVerts, ptrLabelVerts, nE, nV, Px, nPx, dx, Py, nPy, dy,
                     sqdelta, is_DG, kernel, mr, mc, caps, outdim, Pg, degree, dg, fractional_s, fullCC, precision);
    return NLFEM_OK;
    GUARD_END
}

extern "C" int nlfem_plan_assemble(nlfem_plan* plan) {
    GUARD_BEGIN
    if (!plan) return fail(NLFEM_EINVAL, "null plan");
    plan_assemble(plan->P, plan->est);
    return NLFEM_OK;
    GUARD_END
}

extern "C" int nlfem_plan_csr_info(const nlfem_plan* plan, long long* K, long long* nnz) {
    if (!plan || !plan->P.assembled) return fail(NLFEM_EINVAL, "plan not assembled");
    if (K) *K = plan->P.cfg.K;
    if (nnz) *nnz = plan->P.nnz;
    return NLFEM_OK;
}

// Device-to-host copy of several device ranges into caller memory that may be pageable.
// A DMA into pageable memory runs at a few GB/s; instead the ranges stream in chunks through
// two pinned staging buffers (allocated once per process): while the DMA of chunk i + 1 fills
// one buffer, host threads copy chunk i out of the other.  Destinations that are pinned
// already are copied directly.
struct D2HSeg {
    void* dst;
    const void* src;
    size_t bytes;
};

static int copy_d2h_staged(hipStream_t st, const D2HSeg* seg, int nseg) {
    // the copy into fresh pageable pages is bound by their first-touch faults (35 GB/s with 8
    // threads, where the DMA into the pinned buffers runs at 51 GB/s; 16 threads were no faster)
    constexpr size_t CHUNK = 32u << 20;
    constexpr int NTHREADS = 8;
    static std::mutex mtx;
    static char* stage[2] = {nullptr, nullptr};
    std::lock_guard<std::mutex> lock(mtx);
    std::vector<D2HSeg> todo;
    for (int s = 0; s < nseg; s++) {
        if (seg[s].bytes == 0) continue;
        hipPointerAttribute_t at;
        const bool pinned = hipPointerGetAttributes(&at, seg[s].dst) == hipSuccess &&
                            (at.type == hipMemoryTypeHost || at.type == hipMemoryTypeDevice);
        (void)hipGetLastError();  // pageable memory reports an error above
        if (pinned) {
            if (hipMemcpyAsync(seg[s].dst, seg[s].src, seg[s].bytes, hipMemcpyDeviceToHost, st) != hipSuccess)
                return NLFEM_EDEVICE;
        } else {
            todo.push_back(seg[s]);
        }
    }
    if (hipStreamSynchronize(st) != hipSuccess) return NLFEM_EDEVICE;
    if (todo.empty()) return NLFEM_OK;
    if (!stage[0]) {
        for (int b = 0; b < 2; b++)
            if (hipHostMalloc((void**)&stage[b], CHUNK, hipHostMallocDefault) != hipSuccess) {
                stage[b] = nullptr;
                return NLFEM_EDEVICE;
            }
    }
    // chunk list over the segments
    struct Chunk { char* dst; const char* src; size_t bytes; };
    std::vector<Chunk> ch;
    for (const D2HSeg& s : todo)
        for (size_t o = 0; o < s.bytes; o += CHUNK)
            ch.push_back({(char*)s.dst + o, (const char*)s.src + o, std::min(CHUNK, s.bytes - o)});
    hipEvent_t ev[2];
    if (hipEventCreateWithFlags(&ev[0], hipEventDisableTiming) != hipSuccess) return NLFEM_EDEVICE;
    if (hipEventCreateWithFlags(&ev[1], hipEventDisableTiming) != hipSuccess) {
        (void)hipEventDestroy(ev[0]);
        return NLFEM_EDEVICE;
    }
    int rc = NLFEM_OK;
    auto issue = [&](size_t i) {
        const int b = (int)(i & 1);
        if (hipMemcpyAsync(stage[b], ch[i].src, ch[i].bytes, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipEventRecord(ev[b], st) != hipSuccess)
            rc = NLFEM_EDEVICE;
    };
    // NTHREADS workers for the whole copy, each with its slice of every chunk: chunk i is copied out
    // once its DMA has completed (ready >= i), and chunk i + 1's DMA reuses the buffer of chunk
    // i - 1 once every worker has copied its slice of that chunk (copied[i - 1] == NTHREADS)
    const size_t nch = ch.size();
    std::unique_ptr<std::atomic<int>[]> copied(new std::atomic<int>[nch]);
    for (size_t i = 0; i < nch; i++) copied[i].store(0);
    std::atomic<long long> ready{-1};
    std::atomic<bool> abort_copy{false};
    auto copy_slice = [&](size_t i, int t) {
        const size_t n = ch[i].bytes, part = (n + NTHREADS - 1) / NTHREADS;
        const size_t o = std::min(n, t * part), e = std::min(n, o + part);
        if (e > o) std::memcpy(ch[i].dst + o, stage[i & 1] + o, e - o);
        copied[i].fetch_add(1);
    };
    std::thread th[NTHREADS - 1];
    for (int t = 1; t < NTHREADS; t++)
        th[t - 1] = std::thread([&, t] {
            for (size_t i = 0; i < nch; i++) {
                while (ready.load() < (long long)i) {
                    if (abort_copy.load()) return;
                    std::this_thread::yield();
                }
                copy_slice(i, t);
            }
        });
    issue(0);
    for (size_t i = 0; i < nch && rc == NLFEM_OK; i++) {
        if (i + 1 < nch) {
            if (i >= 1)
                while (copied[i - 1].load() < NTHREADS) std::this_thread::yield();
            issue(i + 1);
        }
        if (rc != NLFEM_OK || hipEventSynchronize(ev[i & 1]) != hipSuccess) {
            rc = NLFEM_EDEVICE;
            break;
        }
        ready.store((long long)i);
        copy_slice(i, 0);
    }
    if (rc != NLFEM_OK) abort_copy.store(true);
    for (auto& t : th) t.join();
    (void)hipStreamSynchronize(st);
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    return rc;
}

extern "C" int nlfem_plan_csr_copy(const nlfem_plan* plan, long long* rowptr, int* col, double* val) {
    GUARD_BEGIN
    if (!plan || !plan->P.assembled) return fail(NLFEM_EINVAL, "plan not assembled");
    const Plan& P = plan->P;
    const D2HSeg seg[3] = {{rowptr, P.d_orow, sizeof(long long) * ((size_t)P.cfg.K + 1)},
                           {col, P.d_ocol, sizeof(int) * (size_t)P.nnz},
                           {val, P.d_oval, sizeof(double) * (size_t)P.nnz}};
    if (copy_d2h_staged(P.stream, seg, 3) != NLFEM_OK) return fail(NLFEM_EDEVICE, "device to host copy failed");
    return NLFEM_OK;
    GUARD_END
}

extern "C" int nlfem_plan_csr_device(const nlfem_plan* plan, long long** rowptr, int** col, double** val) {
    if (!plan || !plan->P.assembled) return fail(NLFEM_EINVAL, "plan not assembled");
    if (rowptr) *rowptr = plan->P.d_orow;
    if (col) *col = plan->P.d_ocol;
    if (val) *val = plan->P.d_oval;
    return NLFEM_OK;
}

extern "C" int nlfem_plan_neighbours_total(const nlfem_plan* plan, long long* total) {
    GUARD_BEGIN
    if (!plan || !plan->P.assembled) return fail(NLFEM_EINVAL, "plan not assembled");
    const Plan& P = plan->P;
    std::vector<int> f(P.cfg.nE), b(P.cfg.nE);
    if (hipMemcpy(f.data(), P.d_cntF, sizeof(int) * P.cfg.nE, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(b.data(), P.d_cntB, sizeof(int) * P.cfg.nE, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(NLFEM_EDEVICE, "device to host copy failed");
    long long t = 0;
    for (int e = 0; e < P.cfg.nE; e++) t += f[e] + b[e];
    *total = t;
    return NLFEM_OK;
    GUARD_END
}

extern "C" int nlfem_plan_neighbours(const nlfem_plan* plan, long long* ptr, int* nbr, unsigned char* cls) {
    GUARD_BEGIN
    if (!plan || !plan->P.assembled) return fail(NLFEM_EINVAL, "plan not assembled");
    const Plan& P = plan->P;
    const int nE = P.cfg.nE, cap = P.cap;
    std::vector<int> f(nE), b(nE);
    std::vector<uint32_t> raw((size_t)nE * cap);
    if (hipMemcpy(f.data(), P.d_cntF, sizeof(int) * nE, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(b.data(), P.d_cntB, sizeof(int) * nE, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(raw.data(), P.d_nbr, sizeof(uint32_t) * raw.size(), hipMemcpyDeviceToHost) != hipSuccess)
        return fail(NLFEM_EDEVICE, "device to host copy failed");
    long long o = 0;
    ptr[0] = 0;
    for (int e = 0; e < nE; e++) {
        const uint32_t* r = &raw[(size_t)e * cap];
        for (int p = 0; p < f[e] + b[e]; p++) {
            uint32_t v = p < f[e] ? r[p] : r[cap - 1 - (p - f[e])];
            nbr[o] = (int)(v & IDX_MASK);
            cls[o] = (v >> 31) ? 3 : (unsigned char)((v >> CLS_SHIFT) & 3u);
            o++;
        }
        ptr[e + 1] = o;
    }
    return NLFEM_OK;
    GUARD_END
}

extern "C" int nlfem_plan_stats(const nlfem_plan* plan, double* out, int n) {
    if (!plan) return fail(NLFEM_EINVAL, "null plan");
    const Stats& s = plan->P.stats;
    double v[14 + PH_COUNT];
    v[0] = (double)s.pairs; v[1] = (double)s.probes; v[2] = (double)s.close; v[3] = (double)s.interior;
    v[4] = (double)s.cut; v[5] = (double)s.bary; v[6] = (double)s.nnz_nodes; v[7] = (double)s.nnz;
    v[8] = (double)s.cap; v[9] = (double)s.max_list; v[10] = (double)s.subtri;
    for (int ph = PH_DUAL; ph < PH_COUNT; ph++) v[11 + ph - PH_DUAL] = s.ms[ph];
    int nv = 11 + PH_COUNT - PH_DUAL;
    v[nv++] = (double)s.sur_acc;
    v[nv++] = (double)s.sur_rej;
    v[nv++] = (double)s.sur_items;
    for (int i = 0; i < n && i < nv; i++) out[i] = v[i];
    return NLFEM_OK;
}

extern "C" int nlfem_plan_set_estimates(nlfem_plan* plan, double hmax, double amin, double amean) {
    if (!plan) return fail(NLFEM_EINVAL, "null plan");
    if (!(hmax > 0) || !(amin > 0) || !(amean >= amin)) return fail(NLFEM_EINVAL, "invalid size estimates");
    plan->est.hmax = hmax;
    plan->est.amin = amin;
    plan->est.amean = amean;
    plan->P.bfs_visit = 0;  // re-derive the BFS capacities from the new estimates
    return NLFEM_OK;
}

extern "C" int nlfem_plan_force_paths(nlfem_plan* plan, int flags) {
    if (!plan) return fail(NLFEM_EINVAL, "null plan");
    if (flags & ~31) return fail(NLFEM_EINVAL, "unknown path flags");
    if (flags & NLFEM_PATH_BFS_WAVE) plan->P.bfs_wave = true;
    if (flags & NLFEM_PATH_BFS_HASH) plan->P.bfs_nobm = true;
    if (flags & NLFEM_PATH_PATTERN_HASH) plan->P.pat_nobm = true;
    if (flags & NLFEM_PATH_TRANSPOSE_CSR) plan->P.tpad_off = true;
    if (flags & NLFEM_PATH_TRANSPOSE_UNSTAGED) plan->P.tunstaged = true;
    return NLFEM_OK;
}

extern "C" int nlfem_plan_set_cut_items(nlfem_plan* plan, int items) {
    if (!plan) return fail(NLFEM_EINVAL, "null plan");
    if (items != 0 && items != 64 && items != 128 && items != 256)
        return fail(NLFEM_EINVAL, "items per chunk must be 0 (one item per lane), 64, 128 or 256");
    plan->P.cut_items = items;
    return NLFEM_OK;
}

extern "C" int nlfem_plan_autotune(nlfem_plan* plan, int* chosen, double* ms) {
    GUARD_BEGIN
    if (!plan) return fail(NLFEM_EINVAL, "null plan");
    Plan& P = plan->P;
    // candidates: chunks of 64, 128, 256 items for k_cut, and 0 = k_cutw (one item per lane)
    static const int cand[4] = {64, 128, 256, 0};
    double t[4] = {0, 0, 0, 0};
    int best = 1;
    P.cut_items = cand[best];
    plan_assemble(P, plan->est);
    // scalar kernels with at most 8 outer points take the 8-lane kernel k_cut8 for any
    // nonzero value: it is timed once, as 128, against k_cutw
    const bool cut8 = P.cut8_used;
    if (P.cut_items_used == cand[best] || cut8) {
        t[best] = P.stats.ms[PH_CUT];
        for (int i : {0, 2, 3}) {
            if (cut8 && i != 3) continue;
            P.cut_items = cand[i];
            plan_assemble(P, plan->est);
            const bool ran = cand[i] ? P.cut_items_used == cand[i] : P.cutw_used;
            if (!ran) continue;  // the outer rule or the LDS does not allow it
            t[i] = P.stats.ms[PH_CUT];
            if (t[i] < t[best]) best = i;
        }
    }
    P.cut_items = cand[best];
    if (chosen) *chosen = cand[best];
    if (ms)
        for (int i = 0; i < 4; i++) ms[i] = t[i];
    return NLFEM_OK;
    GUARD_END
}

extern "C" int nlfem_plan_set_surrogate(nlfem_plan* plan, int nnet, const float* params, const float* norm,
                                        float tol, double smin, double smax) {
    if (!plan) return fail(NLFEM_EINVAL, "null plan");
    GUARD_BEGIN
    plan_set_surrogate(plan->P, nnet, params, norm, tol, smin, smax);
    return NLFEM_OK;
    GUARD_END
}

extern "C" int nlfem_plan_surrogate_samples(nlfem_plan* plan, long long maxn, unsigned stride, unsigned salt,
                                            double smin, double smax, float* feat, float* targ, long long* count) {
    if (!plan || !count) return fail(NLFEM_EINVAL, "null argument");
    GUARD_BEGIN
    *count = plan_surrogate_samples(plan->P, maxn, stride, salt, smin, smax, feat, targ);
    return NLFEM_OK;
    GUARD_END
}

extern "C" void nlfem_plan_destroy(nlfem_plan* plan) {
    if (!plan) return;
    plan_free(plan->P);
    delete plan;
}

// ARMA_SPM_BIN_FN008: header, "n_rows n_cols nnz", values, row indices, column pointers of a
// CSC matrix (read back by cython/nlfem.pyx:60-101).  The file holds A^T in CSC, which is A
// in CSR: values, column indices and row pointers of A.
static int write_arma_spmat(const std::string& path, long long K, const std::vector<long long>& rowptr,
                            const std::vector<int>& col, const std::vector<double>& val) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return fail(NLFEM_EIO, "cannot open " + path);
    std::fprintf(f, "ARMA_SPM_BIN_FN008\n%lld %lld %lld\n", K, K, (long long)val.size());
    std::fwrite(val.data(), sizeof(double), val.size(), f);
    std::vector<unsigned long long> u(col.begin(), col.end());
    std::fwrite(u.data(), sizeof(unsigned long long), u.size(), f);
    std::vector<unsigned long long> p(rowptr.begin(), rowptr.end());
    std::fwrite(p.data(), sizeof(unsigned long long), p.size(), f);
    if (std::fclose(f) != 0) return fail(NLFEM_EIO, "cannot write " + path);
    return NLFEM_OK;
}

static int write_arma_mat(const std::string& path, long long K, const std::vector<double>& v) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return fail(NLFEM_EIO, "cannot open " + path);
    std::fprintf(f, "ARMA_MAT_BIN_FN008\n%lld 1\n", K);
    std::fwrite(v.data(), sizeof(double), v.size(), f);
    if (std::fclose(f) != 0) return fail(NLFEM_EIO, "cannot write " + path);
    return NLFEM_OK;
}

extern "C" int nlfem_load_vector(int K, const long* tri, const long* lab, const double* verts, const long* vlab,
                                 int nE, int nV, const double* Px, int nPx, const double* dx, int is_DG,
                                 const char* fname, int outdim, double* fd, const long* zi, const long* zp,
                                 long nZeta) {
    GUARD_BEGIN
    int f = force_id(fname ? fname : "");
    if (f < 0) return fail(NLFEM_EINVAL, std::string("unknown forcing '") + (fname ? fname : "") + "'");
    if ((f == 2) != (outdim == 2)) return fail(NLFEM_EINVAL, "forcing output dimension does not match outdim");
    if (is_DG ? (outdim * nE * 3 != K) : (outdim * nV != K)) return fail(NLFEM_EINVAL, "K mismatch");
    std::vector<double> w;
    if (nZeta > 0) {
        // evaluateZeta(aT, aT) is the length of aT's indicator row (src/Cassemble.cpp:906-914)
        if (!zi || !zp) return fail(NLFEM_EINVAL, "nZeta > 0 needs the ZetaIndicator arrays");
        w.resize((size_t)nE);
        for (int e = 0; e < nE; e++) w[e] = 1.0 / (double)(zp[e + 1] - zp[e]);
    }
    run_forcing(K, tri, lab, verts, vlab, nE, nV, Px, nPx, dx, is_DG, outdim, f, w.empty() ? nullptr : w.data(), fd);
    return NLFEM_OK;
    GUARD_END
}

extern "C" int nlfem_plan_set_zeta(nlfem_plan* plan, const long* indices, const long* indptr, long nZeta) {
    if (!plan) return fail(NLFEM_EINVAL, "null plan");
    GUARD_BEGIN
    plan_set_zeta(plan->P, indices, indptr, nZeta);
    return NLFEM_OK;
    GUARD_END
}

extern "C" int nlfem_par_assemble(const char* compute, const char* path_spAd, const char* path_fd, int K_Omega, int K,
                                  const long* ptrTriangles, const long* ptrLabelTriangles, const double* ptrVerts,
                                  const long* ptrLabelVerts, int nE, int nE_Omega, int nV, int nV_Omega,
                                  const double* Px, int nPx, const double* dx, const double* Py, int nPy,
                                  const double* dy, double sqdelta, int is_DG, int is_NeumannBoundary,
                                  const char* str_model_kernel, const char* str_model_f,
                                  const char* str_integration_method_remote, const char* str_integration_method_close,
                                  int is_PlacePointOnCap, int dim, int outdim, const long* zi, const long* zp,
                                  long nZeta, const double* Pg, int degree, const double* dg, double maxDiameter,
                                  double fractional_s, int is_fullConnectedComponentSearch, int verbose) {
    (void)K_Omega; (void)nV_Omega; (void)is_NeumannBoundary; (void)maxDiameter;
    GUARD_BEGIN
    if (dim != 2) return fail(NLFEM_EINVAL, "only dim == 2 is supported");
    if (nZeta > 0 && (!zi || !zp)) return fail(NLFEM_EINVAL, "nZeta > 0 needs the ZetaIndicator arrays");
    std::string comp = compute ? compute : "";
    if (comp == "system") {
        long long cnt = 0;
        for (int e = 0; e < nE; e++) cnt += ptrLabelTriangles[e] > 0;
        if (cnt != nE_Omega) return fail(NLFEM_EINVAL, "Number of elements with label>0 does not coincide with nE_Omega.");
        nlfem_plan* h = make_plan(K, ptrTriangles, ptrLabelTriangles, ptrVerts, ptrLabelVerts, nE, nV, Px, nPx, dx,
                                  Py, nPy, dy, sqdelta, is_DG, str_model_kernel, str_integration_method_remote,
                                  str_integration_method_close, is_PlacePointOnCap, outdim, Pg, degree, dg,
                                  fractional_s, is_fullConnectedComponentSearch, 0);
        int rc = nZeta > 0 ? nlfem_plan_set_zeta(h, zi, zp, nZeta) : NLFEM_OK;
        if (rc == NLFEM_OK) rc = nlfem_plan_assemble(h);
        if (rc == NLFEM_OK) {
            long long nnz = h->P.nnz;
            std::vector<long long> rp((size_t)K + 1);
            std::vector<int> col((size_t)nnz);
            std::vector<double> val((size_t)nnz);
            rc = nlfem_plan_csr_copy(h, rp.data(), col.data(), val.data());
            if (rc == NLFEM_OK) {
                if (verbose) std::printf("K %d\nNNZ %lld\n", K, nnz);
                rc = write_arma_spmat(path_spAd ? path_spAd : "", K, rp, col, val);
            }
        }
        nlfem_plan_destroy(h);
        return rc;
    }
    if (comp == "forcing") {
        std::vector<double> fd((size_t)K);
        int rc = nlfem_load_vector(K, ptrTriangles, ptrLabelTriangles, ptrVerts, ptrLabelVerts, nE, nV, Px, nPx, dx,
                                   is_DG, str_model_f, outdim, fd.data(), zi, zp, nZeta);
        if (rc != NLFEM_OK) return rc;
        return write_arma_mat(path_fd ? path_fd : "", K, fd);
    }
    return NLFEM_OK;  // other strings do nothing (src/Cassemble.cpp:422-457)
    GUARD_END
}

extern "C" int nlfem_par_evaluateMass(double* vd, const double* ud, const long* Elements, const long* ElementLabels,
                                      const double* Verts, const long* VertexLabels, int K_Omega, int J, int nP,
                                      const double* P, const double* dx, int dim, int outdim, int is_DG, int nE,
                                      int nV) {
    (void)K_Omega;
    GUARD_BEGIN
    if (dim != 2) return fail(NLFEM_EINVAL, "only dim == 2 is supported");
    if (J > nE) return fail(NLFEM_EINVAL, "J exceeds the number of elements");
    const int K = is_DG ? 3 * nE * outdim : nV * outdim;
    run_mass(K, Elements, ElementLabels, Verts, VertexLabels, J, nE, nV, P, nP, dx, is_DG, outdim, ud, vd);
    return NLFEM_OK;
    GUARD_END
}

extern "C" int nlfem_retriangulate_batch(int n, const double* x, const double* TE, double sqdelta, int caps,
                                         double* lists, int* ntri) {
    GUARD_BEGIN
    run_retriangulate(n, x, TE, sqdelta, caps, lists, ntri);
    return NLFEM_OK;
    GUARD_END
}

extern "C" int nlfem_method_retriangulate(const double* xCenter, const double* TE, double sqdelta,
                                          double* reTriangleList, int isPlacePointOnCap) {
    int nt = 0;
    int rc = nlfem_retriangulate_batch(1, xCenter, TE, sqdelta, isPlacePointOnCap, reTriangleList, &nt);
    return rc == NLFEM_OK ? nt : -rc;
}

extern "C" int nlfem_toRef(const double* E, const double* phys_x, double* ref_p) {
    GUARD_BEGIN
    int ok = 0;
    run_geom(1, 0, E, phys_x, ref_p, &ok);
    return ok ? NLFEM_OK : fail(NLFEM_EINVAL, "in solve2x2. Matrix not invertible.");
    GUARD_END
}

extern "C" int nlfem_toPhys(const double* E, const double* p, int dim, double* out_x) {
    GUARD_BEGIN
    if (dim != 2) return fail(NLFEM_EINVAL, "only dim == 2 is supported");
    int ok = 0;
    run_geom(1, 1, E, p, out_x, &ok);
    return NLFEM_OK;
    GUARD_END
}

extern "C" int nlfem_solve2x2(const double* A, const double* b, double* x) {
    GUARD_BEGIN
    int ok = 0;
    run_geom(1, 2, A, b, x, &ok);
    return ok ? NLFEM_OK : fail(NLFEM_EINVAL, "in solve2x2. Matrix not invertible.");
    GUARD_END
}

// ------------------------------------------------------------------ C++ drop-in (Cassemble.h)
// Same signatures as the reference (include/Cassemble.h:62-71,125-140,202-204); errors
// print the message and abort(), as the reference does.
#include "../../include/Cassemble.h"

static void die_if(int rc) {
    if (rc != NLFEM_OK) {
        std::fprintf(stderr, "ERROR in nlfem_amd: %s\n", nlfem_last_error());
        std::abort();
    }
}

void par_assemble(std::string compute, std::string path_spAd, std::string path_fd, int K_Omega, int K,
                  const long* ptrTriangles, const long* ptrLabelTriangles, const double* ptrVerts,
                  const long* ptrLabelVerts, int nE, int nE_Omega, int nV, int nV_Omega, const double* Px, int nPx,
                  const double* dx, const double* Py, int nPy, const double* dy, double sqdelta,
                  int is_DiscontinuousGalerkin, int is_NeumannBoundary, std::string str_model_kernel,
                  std::string str_model_f, std::string str_integration_method_remote,
                  std::string str_integration_method_close, int is_PlacePointOnCap, int dim, int outdim,
                  const long* ptrZetaIndicator_indices, const long* ptrZetaIndicator_indptr, long nZeta,
                  const double* Pg, int degree, const double* dg, double maxDiameter, double fractional_s,
                  int is_fullConnectedComponentSearch, int verbose) {
    die_if(nlfem_par_assemble(compute.c_str(), path_spAd.c_str(), path_fd.c_str(), K_Omega, K, ptrTriangles,
                              ptrLabelTriangles, ptrVerts, ptrLabelVerts, nE, nE_Omega, nV, nV_Omega, Px, nPx, dx, Py,
                              nPy, dy, sqdelta, is_DiscontinuousGalerkin, is_NeumannBoundary,
                              str_model_kernel.c_str(), str_model_f.c_str(), str_integration_method_remote.c_str(),
                              str_integration_method_close.c_str(), is_PlacePointOnCap, dim, outdim,
                              ptrZetaIndicator_indices, ptrZetaIndicator_indptr, nZeta, Pg, degree, dg, maxDiameter,
                              fractional_s, is_fullConnectedComponentSearch, verbose));
}

void par_evaluateMass(double* vd, const double* ud, long* Elements, const long* ElementLabels, const double* Verts,
                      const long* VertexLabels, int K_Omega, int J, int nP, double* P, const double* dx, int dim,
                      int outdim, int is_DiscontinuousGalerkin) {
    // The reference signature carries no nE / nV: J elements are read, and the vertex count is
    // bounded by the largest index those elements use.
    long nVmax = 0;
    for (long i = 0; i < 3L * J; i++) nVmax = std::max(nVmax, Elements[i] + 1);
    die_if(nlfem_par_evaluateMass(vd, ud, Elements, ElementLabels, Verts, VertexLabels, K_Omega, J, nP, P, dx, dim,
                                  outdim, is_DiscontinuousGalerkin, J, (int)nVmax));
}

int method_retriangulate(const double* xCenter, const double* TE, double sqdelta, double* reTriangleList,
                         int isPlacePointOnCap) {
    int r = nlfem_method_retriangulate(xCenter, TE, sqdelta, reTriangleList, isPlacePointOnCap);
    if (r < 0) die_if(-r);
    return r;
}

int method_retriangulateInfty(const double*, const double*, double, double*, int) {
    // The reference ships this as a commented-out stub (src/integration.cpp:1513-1519) and
    // needs CGAL for the L-infinity ball; it is not part of this build.
    std::fprintf(stderr, "ERROR in nlfem_amd: method_retriangulateInfty is not implemented.\n");
    std::abort();
}

void toRef(const double* E, const double* phys_x, double* ref_p) { die_if(nlfem_toRef(E, phys_x, ref_p)); }
void toPhys(const double* E, const double* p, double* out_x) { die_if(nlfem_toPhys(E, p, 2, out_x)); }
void toPhys(const double* E, const double* p, int dim, double* out_x) { die_if(nlfem_toPhys(E, p, dim, out_x)); }
void solve2x2(const double* A, const double* b, double* x) { die_if(nlfem_solve2x2(A, b, x)); }

// ------------------------------------------------------------------ surrogate network
extern "C" int nlfem_mlp_info(int* nin, int* nhidden, int* nout, int* nparam, int* nnorm) {
    if (!nin || !nhidden || !nout || !nparam || !nnorm) return fail(NLFEM_EINVAL, "null pointer");
    *nin = 12; *nhidden = 64; *nout = 9; *nparam = 9737; *nnorm = 42;
    return NLFEM_OK;
}

extern "C" int nlfem_mlp_forward(const float* params, const float* norm, const float* X, long long n, float* Y) {
    if (!params || !norm || (n > 0 && (!X || !Y)) || n < 0) return fail(NLFEM_EINVAL, "invalid arguments");
    GUARD_BEGIN
    mlp::mlp_forward(params, norm, X, n, Y);
    return NLFEM_OK;
    GUARD_END
}

extern "C" int nlfem_mlp_forward_noalloc(const float* d_params, const float* d_norm, const float* d_X, long long n,
                                         float* d_Y, void* stream) {
    if (!d_params || !d_norm || (n > 0 && (!d_X || !d_Y))) return fail(NLFEM_EINVAL, "invalid arguments");
    GUARD_BEGIN
    mlp::mlp_forward_noalloc(d_params, d_norm, d_X, n, d_Y, (hipStream_t)stream);
    return NLFEM_OK;
    GUARD_END
}

extern "C" int nlfem_mlp_grad(const float* params, const float* norm, const float* X, const float* T, long long n,
                              float* grad, float* loss) {
    if (!params || !norm || !X || !T || !grad || !loss || n <= 0) return fail(NLFEM_EINVAL, "invalid arguments");
    GUARD_BEGIN
    mlp::mlp_grad(params, norm, X, T, n, grad, loss);
    return NLFEM_OK;
    GUARD_END
}

extern "C" int nlfem_mlp_train(float* params, const float* norm, const float* X, const float* T, long long n,
                               int epochs, int batch, float lr_max, unsigned seed, float* loss_hist) {
    if (!params || !norm || !X || !T || n <= 0 || !(lr_max > 0)) return fail(NLFEM_EINVAL, "invalid arguments");
    GUARD_BEGIN
    mlp::mlp_train(params, norm, X, T, n, epochs, batch, lr_max, seed, loss_hist);
    return NLFEM_OK;
    GUARD_END
}
