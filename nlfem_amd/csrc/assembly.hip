// Nonlocal stiffness assembly on one MI355X (gfx950): device kernels and the host
// orchestration of one plan.
//
// Pipeline (every phase stream-ordered on the plan's stream):
//   setup      element records (coordinates + |det| + label, 64 B per element)
//   dual       vertex->element incidence and the vertex-sharing dual graph
//   bfs        one wavefront per outer element: level-synchronous BFS over the dual graph
//              with an LDS hash set as the visited set; interaction predicates are the
//              reference's FP64 retriangulation / barycentre tests, bit-exact.  Output is a
//              padded neighbour list per outer element split into a front segment
//              (interior + close pairs) and a back segment (cut pairs), so the quadrature
//              kernels see uniform pair classes.
//   transpose  inverse neighbour relation (pairs grouped by inner element)
//   pattern    node-row sparsity pattern of the full symmetric matrix, built in integers
//              (count pass, scan, fill pass with an LDS bitonic sort per row)
//   positions  transpose-position map and per-element 3x3 position tables
//   close      tensor-Gauss Duffy quadrature of first-layer pairs (singular kernels)
//   interior   pairs whose inner element lies in every ball B(x_k, delta)
//   cut        pairs cut by the ball boundary: streamed retriangulation per outer point
//   compact    drop exact zeros (and unlabelled rows, which receive no contributions),
//              expand od x od blocks, write the scalar CSR
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "nlfem_device.hpp"
#include "plan.hpp"
#include "mlp.hpp"

namespace nlfem {

#define HIPCHECK(x)                                                                                          \
    do {                                                                                                     \
        hipError_t e_ = (x);                                                                                 \
        if (e_ != hipSuccess)                                                                                \
            throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " __FILE__ ":" + \
                                     std::to_string(__LINE__) + " (" #x ")");                                \
    } while (0)

constexpr int CNT_PAIRS = 0, CNT_PROBES = 1, CNT_CLOSE = 2, CNT_INTERIOR = 3, CNT_CUT = 4, CNT_BARY = 5,
              CNT_OVERFLOW = 6, CNT_ERROR = 7, CNT_MAXLIST = 8, CNT_MAXROW = 9, CNT_MAXT = 10, CNT_SUBTRI = 11,
              CNT_SUR_ACC = 12, CNT_SUR_REJ = 13, CNT_SUR_ITEMS = 14, CNT_SAMPLES = 15, CNT_MAXOFF = 16,
              CNT_LABEL0 = 17;
constexpr int NCNT = 24;  // counters allocated (CNT_* < NCNT)

// ======================================================================= small utilities
__global__ void k_fill_int(int* p, int v, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Exclusive scan of n counts into n+1 offsets: block-local scan (1024 items per block of
// 256 threads), recursive scan of block sums, offset add.
template <typename TIn>
__global__ __launch_bounds__(256) void k_scan_block(const TIn* in, long long* out, long long* bsum, long long n) {
    __shared__ long long s[256];
    const long long base = (long long)blockIdx.x * 1024 + threadIdx.x * 4;
    long long v[4], loc = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        v[i] = (base + i < n) ? (long long)in[base + i] : 0;
        loc += v[i];
    }
    s[threadIdx.x] = loc;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        long long t = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0;
        __syncthreads();
        s[threadIdx.x] += t;
        __syncthreads();
    }
    long long ex = s[threadIdx.x] - loc;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        if (base + i < n) out[base + i] = ex;
        ex += v[i];
    }
    if (threadIdx.x == 255) bsum[blockIdx.x] = s[255];
}

__global__ __launch_bounds__(256) void k_scan_add(long long* out, const long long* boff, long long n) {
    const long long base = (long long)blockIdx.x * 1024 + threadIdx.x * 4;
    const long long o = boff[blockIdx.x];
#pragma unroll
    for (int i = 0; i < 4; i++)
        if (base + i < n) out[base + i] += o;
}

__global__ void k_set_last(long long* out, long long n, const long long* src) { out[n] = *src; }

// scratch must hold 2*ceil(n/1024)+2 + (recursion) entries; Plan::scan_tmp is sized for that.
template <typename TIn>
static void scan_exclusive(const TIn* in, long long* out, long long n, long long* scratch, hipStream_t st) {
    long long nb = (n + 1023) / 1024;
    if (nb == 0) {
        HIPCHECK(hipMemsetAsync(out, 0, sizeof(long long), st));
        return;
    }
    long long* bsum = scratch;
    long long* boff = scratch + nb;
    k_scan_block<TIn><<<(unsigned)nb, 256, 0, st>>>(in, out, bsum, n);
    HIPCHECK(hipGetLastError());
    if (nb == 1) {
        k_set_last<<<1, 1, 0, st>>>(out, n, bsum);
        return;
    }
    scan_exclusive<long long>(bsum, boff, nb, boff + nb + 1, st);
    k_scan_add<<<(unsigned)nb, 256, 0, st>>>(out, boff, n);
    k_set_last<<<1, 1, 0, st>>>(out, n, boff + nb);
    HIPCHECK(hipGetLastError());
}

static unsigned nblk(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// LDS hash set (open addressing, linear probing, empty = -1).  Returns 1 if key was
// inserted, 0 if present, -1 if the table is full.
NLFEM_DEV int lds_insert(int* h, int H, int key) {
    uint32_t s = ((uint32_t)key * 2654435761u) & (uint32_t)(H - 1);
    for (int p = 0; p < H; p++) {
        int old = atomicCAS(&h[s], -1, key);
        if (old == -1) return 1;
        if (old == key) return 0;
        s = (s + 1) & (uint32_t)(H - 1);
    }
    return -1;
}
NLFEM_DEV int lds_slot(int* h, int H, int key) {  // slot of key, inserting it if absent; -1 if full
    uint32_t s = ((uint32_t)key * 2654435761u) & (uint32_t)(H - 1);
    for (int p = 0; p < H; p++) {
        int old = atomicCAS(&h[s], -1, key);
        if (old == -1 || old == key) return (int)s;
        s = (s + 1) & (uint32_t)(H - 1);
    }
    return -1;
}
NLFEM_DEV int lds_find(const int* h, int H, int key) {
    uint32_t s = ((uint32_t)key * 2654435761u) & (uint32_t)(H - 1);
    for (int p = 0; p < H; p++) {
        int k = h[s];
        if (k == key) return (int)s;
        if (k == -1) return -1;
        s = (s + 1) & (uint32_t)(H - 1);
    }
    return -1;
}

// Exclusive prefix sum over a 256-thread block (tmp: 4 ints of LDS); total returned in *tot.
NLFEM_DEV int block_exscan256(int v, int* tmp, int* tot) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = v;
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) tmp[w] = x;
    __syncthreads();
    int off = 0, all = 0;
    for (int i = 0; i < 4; i++) {
        if (i < w) off += tmp[i];
        all += tmp[i];
    }
    __syncthreads();
    *tot = all;
    return off + x - v;
}

// Node (degree of freedom) j of element T: the vertex for CG, 3T+j for DG.
NLFEM_DEV int node_of(const int* tri, int T, int j, int is_DG) { return is_DG ? 3 * T + j : tri[3 * T + j]; }
NLFEM_DEV bool row_labelled(const int* vlabel, int node, int is_DG) { return is_DG || vlabel[node] > 0; }

// ======================================================================= setup
__global__ void k_setup(int nE, const long* tri64, const long* lab64, const double* verts, int* tri, ElemRec* elem,
                        unsigned long long* counters, int nV) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nE) return;
    ElemRec r;
    bool bad = false;
    for (int j = 0; j < 3; j++) {
        long v = tri64[3 * (long)e + j];
        if (v < 0 || v >= nV) { bad = true; v = 0; }
        tri[3 * e + j] = (int)v;
        r.E[2 * j] = verts[2 * v];
        r.E[2 * j + 1] = verts[2 * v + 1];
    }
    r.absDet = absDet(r.E);
    r.label = (int)lab64[e];
    r.pad = 0;
    elem[e] = r;
    if (bad) atomicOr(&counters[CNT_ERROR], 1ull);
}

__global__ void k_vlabel(int nV, const long* lab64, int* vlabel) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < nV) vlabel[v] = (int)lab64[v];
}

// ======================================================================= dual graph
__global__ void k_vinc_count(int nE, const int* tri, int* vcnt) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nE) return;
    for (int j = 0; j < 3; j++) atomicAdd(&vcnt[tri[3 * e + j]], 1);
}
__global__ void k_vinc_fill(int nE, const int* tri, const long long* vptr, int* vfill, int* vel) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nE) return;
    for (int j = 0; j < 3; j++) {
        int v = tri[3 * e + j];
        int s = atomicAdd(&vfill[v], 1);
        vel[vptr[v] + s] = e;
    }
}
__global__ void k_sort_vinc(int nV, const long long* vptr, int* vel) {  // sort incidence rows (insertion sort)
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nV) return;
    long long b = vptr[v], e = vptr[v + 1];
    for (long long i = b + 1; i < e; i++) {
        int x = vel[i];
        long long j = i - 1;
        while (j >= b && vel[j] > x) { vel[j + 1] = vel[j]; j--; }
        vel[j + 1] = x;
    }
}

// Dual graph row of element e: elements sharing >= 1 vertex, self excluded, ascending
// (replaces METIS_MeshToDual(ncommon=1), src/Cassemble.cpp:389-417).  The row is the union
// of the three vertices' incidence rows, which are sorted, so a three-way merge yields it in
// order without a buffer: rows of any length (one artificial vertex shared by hundreds of
// label-0 elements, as in the reference's tutorial meshes) need no capacity.
template <bool FILL>
__global__ void k_dual(int nE, const int* tri, const long long* vptr, const int* vel, int* acnt,
                       const long long* xadj, int* adj, unsigned long long* counters) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nE) return;
    long long p[3], q[3];
    for (int j = 0; j < 3; j++) {
        const int v = tri[3 * e + j];
        p[j] = vptr[v];
        q[j] = vptr[v + 1];
    }
    long long o = FILL ? xadj[e] : 0;
    int n = 0, last = -1;
    for (;;) {
        int f = INT_MAX;
        for (int j = 0; j < 3; j++)
            if (p[j] < q[j]) f = min(f, vel[p[j]]);
        if (f == INT_MAX) break;
        for (int j = 0; j < 3; j++)
            if (p[j] < q[j] && vel[p[j]] == f) p[j]++;
        if (f == e || f == last) continue;  // self; an element listed twice at one vertex
        last = f;
        if (FILL) adj[o + n] = f;
        n++;
    }
    if (!FILL) {
        acnt[e] = n;
        atomicMax(&counters[CNT_MAXT], (unsigned long long)n);
    }
}

// Padded rows of the dual graph (W entries, -1 beyond the degree); also the graph's index
// bandwidth max |e - adj(e)| into *bw.
__global__ void k_adjpad(int nE, const long long* xadj, const int* adj, int W, int* adjp, unsigned long long* bw) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nE) return;
    const long long b = xadj[e], d = xadj[e + 1] - b;
    int m = 0;
    for (long long j = 0; j < d; j++) m = max(m, abs(adj[b + j] - e));
    if (adjp)
        for (int j = 0; j < W; j++) adjp[(size_t)e * W + j] = j < d ? adj[b + j] : -1;
    atomicMax(bw, (unsigned long long)m);
}

// ======================================================================= BFS
struct BfsArgs {
    int nE;
    const ElemRec* elem;
    const long long* xadj;
    const int* adj;
    const int* adjp;  // dual graph padded to W4*4 entries per row (-1 = none), or null
    int W4;           // int4 chunks per padded row (0: use the CSR)
    QuadArgs q;
    double sqdelta, sq_in, sq_cut, sq_far;
    int caps, remote, level0, fullCC;
    int cap, H, F;    // list capacity; visited-set words (hash slots or bitmap words); level buffers
    int whalf;        // bitmap window: element e of outer element aT is bit e - (aT - whalf)
    uint32_t* nbr;
    int* cntF;
    int* cntB;
    int2* closeList;
    long long closeCap;
    unsigned long long* counters;
    const int* ord;   // processing order of the outer elements (tile_order), or null
};

// Level-0 method codes: M_RETRIANGULATE, M_BARYCENTER, or 2 = tensor Gauss (always interacts).
// Returns the pair class or -1 when the pair does not interact.
// Outer quadrature points of aT and a bounding circle of them, shared by the wavefront.
struct OuterPts {
    double x[2 * MAXQ];
    double c0, c1;  // centre of the outer points
    double rin2;    // squared radius around c inside which a vertex is in every ball (with margin), or -1
    double rout;    // an inner element whose bounding circle stays farther than this from c meets no ball
};

NLFEM_DEV void outer_points(const BfsArgs& a, const ElemRec& A, OuterPts& op) {
    double c0 = 0, c1 = 0;
    for (int k = 0; k < a.q.nPx; k++) {
        toPhys(A.E, a.q.Px[2 * k], a.q.Px[2 * k + 1], &op.x[2 * k]);
        c0 += op.x[2 * k];
        c1 += op.x[2 * k + 1];
    }
    c0 /= a.q.nPx;
    c1 /= a.q.nPx;
    double r2 = 0;
    for (int k = 0; k < a.q.nPx; k++) r2 = fmax(r2, sqdist(op.x[2 * k], op.x[2 * k + 1], c0, c1));
    const double rx = sqrt(r2) * (1 + 1e-12) + 1e-300;
    const double delta = sqrt(a.sqdelta);
    const double rin = delta * (1 - 1e-7) - rx;
    op.c0 = c0;
    op.c1 = c1;
    op.rin2 = rin > 0 ? rin * rin : -1.0;
    op.rout = delta * (1 + 1e-6) + rx;
}

// Level-0 method codes: M_RETRIANGULATE, M_BARYCENTER, or 2 = tensor Gauss (always interacts).
// classify_quick returns the pair class, -1 when the pair does not interact, or CLS_HARD when
// only retriangulations can tell (the exact test of the hard candidates in k_bfs and k_bfs_h:
// some outer point's polygon has >= 3 points, skipping an outer point whose ball stays
// farther than delta (1 + 1e-9) from all three edges of bT, which has no polygon point).  The
// conservative tests use margins (1e-7, 1e-6, 1e-9 relative) far above rounding, so they only
// ever agree with the exact test:
//  - every vertex of bT within delta(1-1e-7) - r of the outer points' centre: interior;
//  - the bounding circle of bT farther than delta(1+1e-6) + r: every vertex is outside every
//    ball and no edge reaches one, so the reference's polygon is empty for every k;
//  - every vertex within sq_in of every outer point: interior (the polygon is bT itself);
//  - a vertex strictly inside some ball (squared distance <= sq_cut): it is a polygon point of
//    that ball and each of its two edges leaves the ball once, so the polygon has >= 3
//    points and the pair interacts.
constexpr int CLS_HARD = -2;

NLFEM_DEV int classify_quick(const BfsArgs& a, const OuterPts& op, const ElemRec& B, int method) {
    if (method == 2) return (int)CLS_CLOSE;
    if (method == M_BARYCENTER) {
        for (int k = 0; k < a.q.nPx; k++)
            if (bary_interacts(op.x[2 * k], op.x[2 * k + 1], B.E, a.sqdelta)) return 3;  // baryCenter pair
        return a.fullCC ? 3 : -1;
    }
    double dmax = 0;
    for (int j = 0; j < 3; j++) dmax = fmax(dmax, sqdist(B.E[2 * j], B.E[2 * j + 1], op.c0, op.c1));
    if (dmax <= op.rin2) return (int)CLS_INTERIOR;
    {
        const double b0 = (B.E[0] + B.E[2] + B.E[4]) / 3.0, b1 = (B.E[1] + B.E[3] + B.E[5]) / 3.0;
        double rb2 = 0;
        for (int j = 0; j < 3; j++) rb2 = fmax(rb2, sqdist(B.E[2 * j], B.E[2 * j + 1], b0, b1));
        const double rb = sqrt(rb2) * (1 + 1e-12);
        if (sqrt(sqdist(b0, b1, op.c0, op.c1)) - rb > op.rout) return a.fullCC ? (int)CLS_CUT : -1;
    }
    bool interior = true, inside = false;
    for (int k = 0; k < a.q.nPx; k++)
        for (int j = 0; j < 3; j++) {
            const double d = sqdist(B.E[2 * j], B.E[2 * j + 1], op.x[2 * k], op.x[2 * k + 1]);
            interior = interior && (d <= a.sq_in);
            inside = inside || (d <= a.sq_cut);
        }
    if (interior) return (int)CLS_INTERIOR;
    if (inside) return (int)CLS_CUT;
    return CLS_HARD;
}

// Whether the segment P -> Q stays farther than sqrt(sqf) from (x0, x1), without a division:
// with w = x - P, D = Q - P and t = w.D, the squared distance is |w|^2 (t <= 0), |x - Q|^2
// (t >= |D|^2), or (|w|^2 |D|^2 - t^2) / |D|^2, compared as (|w|^2 |D|^2 - t^2) > sqf |D|^2.
// Used as a conservative skip with sqf = delta^2 (1 + 1e-9): its rounding (relative 1e-16
// times (|w| / distance)^2, about 1e-16 near the threshold) is far below the margin.
NLFEM_DEV bool seg_far(double x0, double x1, double P0, double P1, double Q0, double Q1, double sqf) {
    const double D0 = Q0 - P0, D1 = Q1 - P1;
    const double w0 = x0 - P0, w1 = x1 - P1;
    const double t = w0 * D0 + w1 * D1, dd = D0 * D0 + D1 * D1, ww = w0 * w0 + w1 * w1;
    const double u0 = x0 - Q0, u1 = x1 - Q1;
    const double uu = u0 * u0 + u1 * u1;
    return t <= 0 ? ww > sqf : (t >= dd ? uu > sqf : ww * dd - t * t > sqf * dd);
}
// All three edges of the triangle E stay farther than sqrt(sqf) from (x0, x1).
NLFEM_DEV bool tri_far(double x0, double x1, const double* E, double sqf) {
    return seg_far(x0, x1, E[0], E[1], E[2], E[3], sqf) && seg_far(x0, x1, E[2], E[3], E[4], E[5], sqf) &&
           seg_far(x0, x1, E[4], E[5], E[0], E[1], sqf);
}

// Block index with the XCD-major order: workgroups are dispatched to the 8 XCDs round-robin
// (hardware block i runs on XCD i % 8), so the hardware blocks of one XCD take one contiguous
// range of logical blocks.  Neighbouring outer elements write into the same rows of shared
// arrays; keeping them on one XCD lets those partial lines merge in that XCD's L2 instead of
// being written back by several L2s.  (The quadrature kernels and the BFS keep the
// round-robin order, which spreads the spatially uneven work better: measured slower with
// this order.)
constexpr int NXCD = 8;
NLFEM_DEV int xcd_block(int b, int nb) {
    const int q = nb / NXCD, r = nb % NXCD, x = b % NXCD, slot = b / NXCD;
    return x * q + min(x, r) + slot;
}

// Padded dual-graph rows of up to 4 * BFS_W4MAX entries are loaded at once; longer rows use the CSR.
constexpr int BFS_W4MAX = 4;

// Position of this lane among the lanes of the wavefront for which p holds; total = their number.
NLFEM_DEV int wave_rank(bool p, int& total) {
    const unsigned long long m = __ballot(p);
    total = __popcll(m);
    return __popcll(m & ((1ull << __lane_id()) - 1ull));
}

// Visited set of one outer element: an LDS hash of element indices (BM = false), or a bitmap
// over the index window [aT - whalf, aT + whalf) when the dual graph's index bandwidth makes
// that window small (BM = true; an index outside it reports overflow and the host falls back
// to the hash).  Returns 1 when e is new, 0 when already visited, -1 on overflow.
template <bool BM>
NLFEM_DEV int visit_insert(int* vs, const BfsArgs& a, int aT, int e) {
    if constexpr (BM) {
        const int idx = e - (aT - a.whalf);
        if ((unsigned)idx >= (unsigned)(32 * a.H)) return -1;
        const unsigned bit = 1u << (idx & 31);
        const unsigned old = atomicOr((unsigned*)&vs[idx >> 5], bit);
        return (old & bit) ? 0 : 1;
    } else {
        return lds_insert(vs, a.H, e);
    }
}

// One wavefront (one block of 64 lanes) per outer element.  Equivalent to the reference's
// queue BFS (src/Cassemble.cpp:593-783): every element adjacent to an expanded element is
// tested exactly once; interacting and label-0 elements are expanded; the first layer
// {aT} u adj(aT) uses the close method.  The resulting set does not depend on visit order.
// All counters are wavefront-uniform registers, appended to with ballot ranks (no LDS atomics);
// the candidates that only a retriangulation can settle are compacted in place and tested
// densely after the cheap tests of the level, one lane per (candidate, outer point).
template <bool BM>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_bfs(BfsArgs a) {
    extern __shared__ int lds[];
    int* hash = lds;
    int* test = hash + a.H;
    int* next = test + a.F;
    unsigned* hflag = reinterpret_cast<unsigned*>(next + a.F);  // one bit per hard candidate
    __shared__ OuterPts op;
    const int lane = threadIdx.x;
    const int aT = blockIdx.x;
    if (aT >= a.nE) return;
    const ElemRec A = a.elem[aT];
    if (A.label == 0) {
        if (lane == 0) { a.cntF[aT] = 0; a.cntB[aT] = 0; }
        return;
    }
    for (int i = lane; i < a.H; i += 64) hash[i] = BM ? 0 : -1;
    if (lane == 0) outer_points(a, A, op);
    __syncthreads();
    int nTest = 0, nNext = 0, nTot = 0, nF = 0, nB = 0;
    bool ovf = false, wovf = false;  // wovf: an index fell outside the bitmap window
    unsigned long long cProbe = 0, cClose = 0, cInt = 0, cCut = 0, cBary = 0;
    // level 0 candidates: aT itself, then its dual-graph neighbours
    const long long x0 = a.xadj[aT];
    const int deg = (int)(a.xadj[aT + 1] - x0);
    if (1 + deg > a.F) {
        ovf = true;
    } else {
        if (lane == 0) { visit_insert<BM>(hash, a, aT, aT); test[0] = aT; }
        __syncthreads();
        bool lo = false;
        for (int j = lane; j < deg; j += 64) {
            const int b = a.adj[x0 + j];
            if (visit_insert<BM>(hash, a, aT, b) < 0) lo = true;
            test[1 + j] = b;
        }
        ovf = __any(lo);
        wovf = BM && ovf;
        nTest = 1 + deg;
    }
    __syncthreads();

    // record a tested candidate (cls >= 0: interacting) and queue it for expansion
    auto record = [&](int b, int cls, bool push) {
        const bool rec = cls >= 0;
        const bool isF = cls == (int)CLS_INTERIOR || cls == (int)CLS_CLOSE;
        int nr, nf, nb;
        wave_rank(rec, nr);
        const int rf = wave_rank(rec && isF, nf);
        const int rb = wave_rank(rec && !isF, nb);
        if (nTot + nr > a.cap) ovf = true;
        if (rec && !ovf) {
            const int slot = isF ? nF + rf : a.cap - 1 - (nB + rb);
            const uint32_t code = (cls == 3) ? CLS_CUT : (uint32_t)cls;
            a.nbr[(size_t)aT * a.cap + slot] = (uint32_t)b | (code << CLS_SHIFT) | (cls == 3 ? (1u << 31) : 0u);
            if (cls == (int)CLS_CLOSE) {
                const unsigned long long c = atomicAdd(&a.counters[CNT_CLOSE], 1ull);
                if ((long long)c < a.closeCap) a.closeList[c] = make_int2(aT, b);
            }
        }
        nTot += nr; nF += nf; nB += nb;
        cClose += cls == (int)CLS_CLOSE; cInt += cls == (int)CLS_INTERIOR; cCut += cls == (int)CLS_CUT;
        cBary += cls == 3;
        const bool q = push && b != aT;
        int nq;
        const int rq = wave_rank(q, nq);
        if (nNext + nq > a.F) ovf = true;
        else if (q) next[nNext + rq] = b;
        nNext += nq;
    };

    int level = 0;
    while (!ovf) {
        const int method = (level == 0) ? a.level0 : a.remote;
        // cheap tests; undecided candidates are compacted to the front of test[] (every index
        // written lies below the end of the round, whose entries have all been read)
        int nHard = 0;
        for (int r0 = 0; r0 < nTest; r0 += 64) {
            const int j = r0 + lane;
            const bool valid = j < nTest;
            const int b = valid ? test[j] : -1;
            int cls = -1;
            bool push = false, hard = false;
            if (valid) {
                const ElemRec B = a.elem[b];
                if (B.label == 0) {
                    push = true;  // label-0 elements are traversed only (src/Cassemble.cpp:773-775)
                } else {
                    cls = classify_quick(a, op, B, method);
                    hard = cls == CLS_HARD;
                    if (hard) cls = -1;
                    push = cls >= 0;
                    cProbe += !hard && cls < 0;  // tested and not interacting
                }
            }
            int nh;
            const int rh = wave_rank(hard, nh);
            if (hard) test[nHard + rh] = b;
            nHard += nh;
            record(b, cls, push);
        }
        __syncthreads();
        if (nHard > 0) {
            // the exact test (the reference predicate: some outer point's polygon has >= 3
            // points) with one lane per (hard candidate, outer point), OR-ed into the
            // candidate's bit
            const int nPx = a.q.nPx;
            for (int i = lane; i <= nHard / 32; i += 64) hflag[i] = 0u;
            __syncthreads();
            for (int it = lane; it < nHard * nPx; it += 64) {
                const int j = it / nPx, k = it - j * nPx;
                const ElemRec B = a.elem[test[j]];
                const double x0 = op.x[2 * k], x1 = op.x[2 * k + 1];
                if (!tri_far(x0, x1, B.E, a.sq_far)) {
                    NoEmit ne;
                    if (retriangulate<true>(x0, x1, B.E, a.sqdelta, a.caps != 0, ne) >= 3)
                        atomicOr(&hflag[j >> 5], 1u << (j & 31));
                }
            }
            __syncthreads();
            for (int r0 = 0; r0 < nHard; r0 += 64) {
                const int j = r0 + lane;
                const bool valid = j < nHard;
                const int b = valid ? test[j] : -1;
                int cls = -1;
                if (valid) {
                    cls = (((hflag[j >> 5] >> (j & 31)) & 1u) || a.fullCC) ? (int)CLS_CUT : -1;
                    cProbe += cls < 0;
                }
                record(b, cls, cls >= 0);
            }
        }
        ovf = __any(ovf);
        __syncthreads();
        if (nNext == 0 || ovf) break;
        // discover the next level: unvisited neighbours of the expanded elements
        nTest = 0;
        bool lovf = false;
        auto append = [&](int e, bool isnew) {  // wavefront-uniform call
            int nt;
            const int rt = wave_rank(isnew, nt);
            if (isnew && nTest + rt < a.F) test[nTest + rt] = e;
            nTest += nt;
        };
        auto visit = [&](int e) {  // wavefront-uniform call; e < 0 for no candidate
            bool isnew = false;
            if (e >= 0) {
                const int r = visit_insert<BM>(hash, a, aT, e);
                if (r < 0) lovf = true;
                isnew = r == 1;
            }
            append(e, isnew);
        };
        for (int r0 = 0; r0 < nNext; r0 += 64) {
            const int j = r0 + lane;
            const int s = j < nNext ? next[j] : -1;
            if (a.W4 > 0) {
                // padded row: all loads issued before the first insert
                const int4* row = reinterpret_cast<const int4*>(a.adjp) + (size_t)(s < 0 ? 0 : s) * a.W4;
                int4 v[BFS_W4MAX];
#pragma unroll
                for (int c = 0; c < BFS_W4MAX; c++)
                    if (c < a.W4) v[c] = s >= 0 ? row[c] : make_int4(-1, -1, -1, -1);
                if constexpr (BM) {
                    // all bitmap updates of the row are issued before any result is used
                    int ev[4 * BFS_W4MAX];
                    unsigned old[4 * BFS_W4MAX];
#pragma unroll
                    for (int c = 0; c < BFS_W4MAX; c++) {
                        ev[4 * c] = v[c].x; ev[4 * c + 1] = v[c].y; ev[4 * c + 2] = v[c].z; ev[4 * c + 3] = v[c].w;
                    }
                    const int base = aT - a.whalf;
#pragma unroll
                    for (int q = 0; q < 4 * BFS_W4MAX; q++) {
                        old[q] = 0xffffffffu;
                        if (q < 4 * a.W4 && ev[q] >= 0) {
                            const int idx = ev[q] - base;
                            if ((unsigned)idx >= (unsigned)(32 * a.H)) lovf = true;
                            else old[q] = atomicOr((unsigned*)&hash[idx >> 5], 1u << (idx & 31));
                        }
                    }
#pragma unroll
                    for (int q = 0; q < 4 * BFS_W4MAX; q++) {
                        if (q >= 4 * a.W4) break;
                        const int idx = ev[q] - base;
                        const bool isnew = ev[q] >= 0 && !((old[q] >> (idx & 31)) & 1u);
                        append(ev[q], isnew);
                    }
                } else {
#pragma unroll
                    for (int c = 0; c < BFS_W4MAX; c++) {
                        if (c >= a.W4) break;
                        visit(v[c].x);
                        visit(v[c].y);
                        visit(v[c].z);
                        visit(v[c].w);
                    }
                }
            } else {
                const long long p0 = s >= 0 ? a.xadj[s] : 0;
                const int len = s >= 0 ? (int)(a.xadj[s + 1] - p0) : 0;
                int mlen = len;
                for (int m = 32; m > 0; m >>= 1) mlen = max(mlen, __shfl_xor(mlen, m, 64));
                for (int p = 0; p < mlen; p++) visit(p < len ? a.adj[p0 + p] : -1);
            }
        }
        const bool al = __any(lovf);
        if (BM && al) wovf = true;
        ovf = al || nTest > a.F;
        nNext = 0;
        __syncthreads();
        level++;
    }
    // wave reductions of the statistics
    for (int m = 32; m > 0; m >>= 1) {
        cProbe += __shfl_xor(cProbe, m, 64);
        cClose += __shfl_xor(cClose, m, 64);
        cInt += __shfl_xor(cInt, m, 64);
        cCut += __shfl_xor(cCut, m, 64);
        cBary += __shfl_xor(cBary, m, 64);
    }
    if (lane == 0) {
        if (ovf) {
            atomicOr(&a.counters[CNT_OVERFLOW], wovf ? 2ull : 1ull);
            a.cntF[aT] = 0;
            a.cntB[aT] = 0;
        } else {
            a.cntF[aT] = nF;
            a.cntB[aT] = nB;
            atomicMax(&a.counters[CNT_MAXLIST], (unsigned long long)(nF + nB));
        }
        atomicAdd(&a.counters[CNT_PAIRS], cClose + cInt + cCut + cBary);
        atomicAdd(&a.counters[CNT_PROBES], cProbe);
        atomicAdd(&a.counters[CNT_INTERIOR], cInt);
        atomicAdd(&a.counters[CNT_CUT], cCut);
        atomicAdd(&a.counters[CNT_BARY], cBary);
    }
}

// Two outer elements per wavefront: half-wavefront h (lanes 32 h .. 32 h + 31) runs the BFS of
// outer element 2 blockIdx.x + h with the bitmap visited set, exactly as k_bfs does with a
// whole wavefront (same tests, same expansion, same list layout).  The BFS is bound by the
// latency of its dependent levels, not by arithmetic, and its register demand limits the
// resident wavefronts, so two searches per wavefront double the searches in flight.  Control
// flow stays wavefront-uniform: loops run to the larger of the two halves' bounds with per-half
// predicates, counters are half-uniform, and ranks come from ballots masked to the half.
// The level buffers hold 16-bit offsets into the bitmap window (every element stored there
// has passed the window test), which needs a window of at most 65,536 elements.
NLFEM_DEV unsigned long long half_mask() { return (__lane_id() & 32) ? 0xffffffff00000000ull : 0xffffffffull; }
NLFEM_DEV int half_rank(bool p, int& total) {
    const unsigned long long m = __ballot(p) & half_mask();
    total = __popcll(m);
    return __popcll(m & ((1ull << __lane_id()) - 1ull));
}
NLFEM_DEV bool half_any(bool p) { return (__ballot(p) & half_mask()) != 0ull; }
NLFEM_DEV int wave_max2(int v) { return max(__shfl(v, 0, 64), __shfl(v, 32, 64)); }

// Capped at 128 registers (4 wavefronts per SIMD, 14 spilled): the exact test of the hard
// candidates runs one outer point per lane, which keeps its registers low enough that the
// extra occupancy pays (jittered h = 1/128 mesh, delta = 0.1: 1.47 -> 1.33 ms; h = 1/256,
// delta = 0.05: unchanged).
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_bfs_h(BfsArgs a) {
    extern __shared__ int ldsh[];
    __shared__ OuterPts ops[2];
    const int lane = threadIdx.x;
    const int hf = lane >> 5, hl = lane & 31;
    int* hash = ldsh + hf * (a.H + a.F + a.F / 32 + 1);
    unsigned short* test = reinterpret_cast<unsigned short*>(hash + a.H);
    unsigned short* next = test + a.F;
    unsigned* hflag = reinterpret_cast<unsigned*>(hash + a.H + a.F);  // one bit per hard candidate
    OuterPts& op = ops[hf];
    const int slot = 2 * blockIdx.x + hf;  // the two outer elements: consecutive ones of the tile order
    const bool live = slot < a.nE;
    const int aT = live && a.ord ? a.ord[slot] : slot;
    const int aTs = live ? aT : 0;
    const ElemRec A = a.elem[aTs];
    const bool skip = !live || A.label == 0;
    const int base = aT - a.whalf;  // element e is bit / offset e - base of the window
    if (live && A.label == 0 && hl == 0) { a.cntF[aT] = 0; a.cntB[aT] = 0; }
    for (int i = hl; i < a.H; i += 32) hash[i] = 0;
    if (hl == 0 && !skip) outer_points(a, A, op);
    __syncthreads();
    int nTest = 0, nNext = 0, nTot = 0, nF = 0, nB = 0;
    bool ovf = false, wovf = false;
    // per-lane pair counts (each far below 2^16): probes; interior | close << 16; cut | bary << 16
    unsigned cProbe = 0, cIC = 0, cCB = 0;
    // level 0 candidates: aT itself, then its dual-graph neighbours
    const long long x0 = a.xadj[aTs];
    const int deg = skip ? 0 : (int)(a.xadj[aTs + 1] - x0);
    bool lo = false;
    if (!skip && 1 + deg <= a.F) {
        if (hl == 0) { visit_insert<true>(hash, a, aT, aT); test[0] = (unsigned short)a.whalf; }
        for (int j = hl; j < deg; j += 32) {
            const int b = a.adj[x0 + j];
            const bool inw = visit_insert<true>(hash, a, aT, b) >= 0;
            if (!inw) lo = true;
            test[1 + j] = inw ? (unsigned short)(b - base) : (unsigned short)0;
        }
    }
    lo = half_any(lo);
    ovf = (!skip && 1 + deg > a.F) || lo;
    wovf = lo;
    nTest = skip ? 0 : 1 + deg;
    __syncthreads();

    auto record = [&](int b, int cls, bool push) {
        const bool rec = cls >= 0;
        const bool isF = cls == (int)CLS_INTERIOR || cls == (int)CLS_CLOSE;
        int nr, nf, nb;
        half_rank(rec, nr);
        const int rf = half_rank(rec && isF, nf);
        const int rb = half_rank(rec && !isF, nb);
        if (nTot + nr > a.cap) ovf = true;
        if (rec && !ovf) {
            const int slot = isF ? nF + rf : a.cap - 1 - (nB + rb);
            const uint32_t code = (cls == 3) ? CLS_CUT : (uint32_t)cls;
            a.nbr[(size_t)aT * a.cap + slot] = (uint32_t)b | (code << CLS_SHIFT) | (cls == 3 ? (1u << 31) : 0u);
            if (cls == (int)CLS_CLOSE) {
                const unsigned long long c = atomicAdd(&a.counters[CNT_CLOSE], 1ull);
                if ((long long)c < a.closeCap) a.closeList[c] = make_int2(aT, b);
            }
        }
        nTot += nr; nF += nf; nB += nb;
        cIC += (cls == (int)CLS_INTERIOR ? 1u : 0u) + (cls == (int)CLS_CLOSE ? 0x10000u : 0u);
        cCB += (cls == (int)CLS_CUT ? 1u : 0u) + (cls == 3 ? 0x10000u : 0u);
        const bool q = push && b != aT;
        int nq;
        const int rq = half_rank(q, nq);
        if (nNext + nq > a.F) ovf = true;
        else if (q) next[nNext + rq] = (unsigned short)(b - base);
        nNext += nq;
    };

    bool done = skip || ovf;
    int level = 0;
    while (__any(!done)) {
        const int method = (level == 0) ? a.level0 : a.remote;
        int nHard = 0;
        const int nT = done ? 0 : nTest;
        const int mT = wave_max2(nT);
        for (int r0 = 0; r0 < mT; r0 += 32) {
            const int j = r0 + hl;
            const bool valid = j < nT;
            const int b = valid ? base + (int)test[j] : -1;
            int cls = -1;
            bool push = false, hard = false;
            if (valid) {
                const ElemRec B = a.elem[b];
                if (B.label == 0) {
                    push = true;  // label-0 elements are traversed only (src/Cassemble.cpp:773-775)
                } else {
                    cls = classify_quick(a, op, B, method);
                    hard = cls == CLS_HARD;
                    if (hard) cls = -1;
                    push = cls >= 0;
                    cProbe += !hard && cls < 0;
                }
            }
            int nh;
            const int rh = half_rank(hard, nh);
            if (hard) test[nHard + rh] = (unsigned short)(b - base);
            nHard += nh;
            record(b, cls, push);
        }
        __syncthreads();
        const int mH = wave_max2(nHard);
        if (mH > 0) {
            // the exact test (the reference predicate: some outer point's polygon has >= 3
            // points) with one lane per (hard candidate, outer point): each lane runs one
            // retriangulation and ORs a hit into its candidate's bit, instead of one lane per
            // candidate running up to nPx of them one after another
            const int nPx = a.q.nPx;
            for (int i = hl; i <= nHard / 32; i += 32) hflag[i] = 0u;
            __syncthreads();
            const int nItems = nHard * nPx;
            const int mI = wave_max2(nItems);
            for (int r0 = 0; r0 < mI; r0 += 32) {
                const int it = r0 + hl;
                if (it < nItems) {
                    const int j = it / nPx, k = it - j * nPx;
                    const ElemRec B = a.elem[base + (int)test[j]];
                    const double x0 = op.x[2 * k], x1 = op.x[2 * k + 1];
                    if (!tri_far(x0, x1, B.E, a.sq_far)) {
                        NoEmit ne;
                        if (retriangulate<true>(x0, x1, B.E, a.sqdelta, a.caps != 0, ne) >= 3)
                            atomicOr(&hflag[j >> 5], 1u << (j & 31));
                    }
                }
            }
            __syncthreads();
            for (int r0 = 0; r0 < mH; r0 += 32) {
                const int j = r0 + hl;
                const bool valid = j < nHard;
                const int b = valid ? base + (int)test[j] : -1;
                int cls = -1;
                if (valid) {
                    cls = (((hflag[j >> 5] >> (j & 31)) & 1u) || a.fullCC) ? (int)CLS_CUT : -1;
                    cProbe += cls < 0;
                }
                record(b, cls, cls >= 0);
            }
        }
        ovf = half_any(ovf);
        __syncthreads();
        if (nNext == 0 || ovf) done = true;
        if (!__any(!done)) break;
        // discover the next level: unvisited neighbours of the expanded elements
        nTest = 0;
        bool lovf = false;
        auto append = [&](int e, bool isnew) {
            int nt;
            const int rt = half_rank(isnew, nt);
            if (isnew && nTest + rt < a.F) test[nTest + rt] = (unsigned short)(e - base);
            nTest += nt;
        };
        const int nN = done ? 0 : nNext;
        const int mN = wave_max2(nN);
        for (int r0 = 0; r0 < mN; r0 += 32) {
            const int j = r0 + hl;
            const int s = j < nN ? base + (int)next[j] : -1;
            const int4* row = reinterpret_cast<const int4*>(a.adjp) + (size_t)(s < 0 ? 0 : s) * a.W4;
            int4 v[BFS_W4MAX];
#pragma unroll
            for (int c = 0; c < BFS_W4MAX; c++)
                if (c < a.W4) v[c] = s >= 0 ? row[c] : make_int4(-1, -1, -1, -1);
            int ev[4 * BFS_W4MAX];
            unsigned old[4 * BFS_W4MAX];
#pragma unroll
            for (int c = 0; c < BFS_W4MAX; c++) {
                ev[4 * c] = v[c].x; ev[4 * c + 1] = v[c].y; ev[4 * c + 2] = v[c].z; ev[4 * c + 3] = v[c].w;
            }
#pragma unroll
            for (int q = 0; q < 4 * BFS_W4MAX; q++) {
                old[q] = 0xffffffffu;
                if (q < 4 * a.W4 && ev[q] >= 0) {
                    const int idx = ev[q] - base;
                    if ((unsigned)idx >= (unsigned)(32 * a.H)) lovf = true;
                    else old[q] = atomicOr((unsigned*)&hash[idx >> 5], 1u << (idx & 31));
                }
            }
#pragma unroll
            for (int q = 0; q < 4 * BFS_W4MAX; q++) {
                if (q >= 4 * a.W4) break;
                const int idx = ev[q] - base;
                const bool isnew = ev[q] >= 0 && !((old[q] >> (idx & 31)) & 1u);
                append(ev[q], isnew);
            }
        }
        const bool al = half_any(lovf);
        if (al) wovf = true;
        if (!done) {
            ovf = al || nTest > a.F;
            if (ovf) done = true;
        }
        nNext = 0;
        __syncthreads();
        level++;
    }
    unsigned long long cInt = cIC & 0xffffu, cClose = cIC >> 16, cCut = cCB & 0xffffu, cBary = cCB >> 16;
    unsigned long long cPr = cProbe;
    for (int m = 16; m > 0; m >>= 1) {
        cPr += __shfl_xor(cPr, m, 64);
        cClose += __shfl_xor(cClose, m, 64);
        cInt += __shfl_xor(cInt, m, 64);
        cCut += __shfl_xor(cCut, m, 64);
        cBary += __shfl_xor(cBary, m, 64);
    }
    if (hl == 0 && !skip) {
        if (ovf) {
            atomicOr(&a.counters[CNT_OVERFLOW], wovf ? 2ull : 1ull);
            a.cntF[aT] = 0;
            a.cntB[aT] = 0;
        } else {
            a.cntF[aT] = nF;
            a.cntB[aT] = nB;
            atomicMax(&a.counters[CNT_MAXLIST], (unsigned long long)(nF + nB));
        }
        atomicAdd(&a.counters[CNT_PAIRS], cClose + cInt + cCut + cBary);
        atomicAdd(&a.counters[CNT_PROBES], cPr);
        atomicAdd(&a.counters[CNT_INTERIOR], cInt);
        atomicAdd(&a.counters[CNT_CUT], cCut);
        atomicAdd(&a.counters[CNT_BARY], cBary);
    }
}

NLFEM_DEV int list_index(int p, int nF, int cap) { return p < nF ? p : cap - 1 - (p - nF); }

// Largest |b - aT| over the neighbour lists (CNT_MAXOFF) and whether the mesh has label-0
// elements (CNT_LABEL0).  Without label-0 elements every element a search visits is a listed
// element or a dual-graph neighbour of one, so this extent plus the dual graph's index
// bandwidth bounds the search's index window.  Each wavefront walks LE_PER outer elements, and
// each block issues one atomic per counter (one atomic per outer element, all on one word,
// serialised in the L2: 1.5 ms on 131,072 elements).
constexpr int LE_PER = 16;
__global__ __launch_bounds__(256) void k_list_extent(int nE, int cap, const uint32_t* nbr, const int* cntF,
                                                     const int* cntB, const ElemRec* elem,
                                                     unsigned long long* counters) {
    __shared__ int wm[4], wl[4];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int a0 = (blockIdx.x * 4 + w) * LE_PER;
    int m = 0, l0 = 0;
    for (int aT = a0; aT < min(a0 + LE_PER, nE); aT++) {
        if (lane == 0 && elem[aT].label == 0) l0 = 1;
        const int nF = cntF[aT], n = nF + cntB[aT];
        for (int p = lane; p < n; p += 64) {
            const int b = (int)(nbr[(size_t)aT * cap + list_index(p, nF, cap)] & IDX_MASK);
            m = max(m, abs(b - aT));
        }
    }
    for (int k = 32; k > 0; k >>= 1) m = max(m, __shfl_xor(m, k, 64));
    if (lane == 0) { wm[w] = m; wl[w] = l0; }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int bm = max(max(wm[0], wm[1]), max(wm[2], wm[3]));
        if (bm > 0) atomicMax(&counters[CNT_MAXOFF], (unsigned long long)bm);
        if (wl[0] | wl[1] | wl[2] | wl[3]) atomicOr(&counters[CNT_LABEL0], 1ull);
    }
}

// ======================================================================= transpose
// Transposed relation.  Row b lists the outer elements aT of the pairs (aT, b): pairs from
// the front segments (interior, close) from the row's start, cut pairs from its end.  The cut
// entries also record the pair's index in the cut-pair enumeration (coff[aT] + its position
// in aT's back segment).  A block takes tg consecutive outer elements (of the tile order),
// whose lists share most of their inner elements: per-b counts are summed in an LDS hash first,
// so each block issues one global atomic per distinct b and segment (with a bounded probe; keys
// that find no slot use global atomics per entry).
constexpr int TH = 4096;
// Insert with at most 64 probes; -1 if none free.  A key recurs up to tg times per block, so
// each probe reads the slot first and takes the compare-and-swap only on an empty slot.
NLFEM_DEV int lds_slot_b(int* h, int key) {
    uint32_t s = ((uint32_t)key * 2654435761u) & (uint32_t)(TH - 1);
    for (int p = 0; p < 64; p++) {
        const int cur = h[s];
        if (cur == key) return (int)s;
        if (cur == -1) {
            const int old = atomicCAS(&h[s], -1, key);
            if (old == -1 || old == key) return (int)s;
        }
        s = (s + 1) & (uint32_t)(TH - 1);
    }
    return -1;
}
NLFEM_DEV int lds_find_b(const int* h, int key) {
    uint32_t s = ((uint32_t)key * 2654435761u) & (uint32_t)(TH - 1);
    for (int p = 0; p < 64; p++) {
        int k = h[s];
        if (k == key) return (int)s;
        if (k == -1) return -1;
        s = (s + 1) & (uint32_t)(TH - 1);
    }
    return -1;
}

// Counting pass for exact CSR rows: tcnt[b] += entries of row b, one global atomic per distinct b
// of the block (per entry for keys that find no LDS slot).
__global__ __launch_bounds__(256) void k_transpose_count(int nE, int cap, int tg, const uint32_t* nbr,
                                                         const int* cntF, const int* cntB, int* tcnt,
                                                         const int* ord) {
    __shared__ int hk[TH];
    __shared__ int hf[TH];
    const int tid = threadIdx.x;
    for (int i = tid; i < TH; i += 256) { hk[i] = -1; hf[i] = 0; }
    __syncthreads();
    const int a0 = xcd_block(blockIdx.x, gridDim.x) * tg, a1 = min(a0 + tg, nE);
    constexpr int TU = 4;
    for (int ia = a0; ia < a1; ia++) {
        const int aT = ord ? ord[ia] : ia;
        const int nF = cntF[aT], n = nF + cntB[aT];
        for (int p0 = tid; p0 < n; p0 += TU * 256) {
            uint32_t ev[TU];
#pragma unroll
            for (int u = 0; u < TU; u++) {
                const int p = p0 + u * 256;
                ev[u] = p < n ? nbr[(size_t)aT * cap + list_index(p, nF, cap)] : 0u;
            }
#pragma unroll
            for (int u = 0; u < TU; u++) {
                const int p = p0 + u * 256;
                if (p >= n) break;
                const int b = (int)(ev[u] & IDX_MASK);
                const int sl = lds_slot_b(hk, b);
                if (sl >= 0) atomicAdd(&hf[sl], 1);
                else atomicAdd(&tcnt[b], 1);
            }
        }
    }
    __syncthreads();
    for (int i = tid; i < TH; i += 256)
        if (hk[i] >= 0) atomicAdd(&tcnt[hk[i]], hf[i]);
}

// Fill pass with row runs.  Row b occupies [tptr[b], tptr[b+1]): the exact CSR rows of the
// counting pass, or padded rows of capT entries (row b at b capT; capT = 0: CSR rows), whose
// overflow k_tcheck reports; tfill / tfillB end as row b's front and cut-entry counts.  A block takes tg <= 8 consecutive outer elements k of the tile order.
//   1  LDS hash of the distinct inner elements b of their lists.  Word m[b]: bit k for a front
//      entry of outer element k that is an interior pair, bit 8 + k for a close pair, and the
//      number of back (cut) entries in bits 16..19.
//   2  Exclusive scan of the back counts over the slots (offsets in bits 20..31), then each
//      back entry goes to its slot's range of an LDS array: outer element k, bit 31 of the
//      entry, its index in k's back segment.  The slot's count is then base(next slot) - base.
//   3  Per slot: one global atomic per segment reserves the block's ranks in row b, and the
//      thread writes the block's front entries of the row, then its back entries, as runs of
//      consecutive stores that merge in the L2.  (Entries placed one outer element at a time
//      reached a row's line in passes far apart, and the lines left the L2 partially written.)
// Entries whose key found no hash slot, unexpected classes and blocks with more back entries
// than the LDS array holds take per-entry global atomics.
constexpr int TS = 4000;  // back entries staged per block (offsets and indices fit 12 bits)
static_assert(TH == 16 * 256, "the offset scan gives each thread 16 slots");
__global__ __launch_bounds__(256) void k_transpose_fill(int nE, int cap, int tg, const uint32_t* nbr, const int* cntF,
                                                        const int* cntB, const long long* tptr, int* tfill,
                                                        int* tadj, const long long* coff, int* tfillB, int* tcut,
                                                        int capT, const int* ord, int tsmax) {
    __shared__ int hk[TH];
    __shared__ unsigned hm[TH];
    __shared__ unsigned short bs[TS];  // outer element k, bit 31 of the entry, index in k's back segment
    __shared__ int sa[8], sF[8], sN[8], scan_tmp[4];
    __shared__ long long sco[8];
    const int tid = threadIdx.x;
    for (int i = tid; i < TH; i += 256) { hk[i] = -1; hm[i] = 0u; }
    const int a0 = xcd_block(blockIdx.x, gridDim.x) * tg, nk = min(tg, nE - a0);
    if (tid < nk) {
        const int aT = ord ? ord[a0 + tid] : a0 + tid;
        sa[tid] = aT;
        sF[tid] = cntF[aT];
        sN[tid] = sF[tid] + cntB[aT];
        sco[tid] = coff[aT];
    }
    __syncthreads();
    int nbt = 0;
    for (int k = 0; k < nk; k++) nbt += sN[k] - sF[k];
    const bool stage = nbt <= tsmax;  // tsmax <= TS
    auto row_range = [&](int b, long long& r0, long long& r1) {
        r0 = capT ? (long long)b * capT : tptr[b];
        r1 = capT ? r0 + capT : tptr[b + 1];
    };
    auto is_staged_back = [](uint32_t e) { return ((e >> CLS_SHIFT) & 3u) == CLS_CUT; };
    constexpr int TU = 8;  // list loads in flight per thread (4: 2.02 ms, 8: 1.96 ms at h = 1/256)
    // 1: hash, front bits and back counts
    for (int k = 0; k < nk; k++) {
        const int aT = sa[k], nF = sF[k], n = sN[k];
        for (int p0 = tid; p0 < n; p0 += TU * 256) {
            uint32_t ev[TU];
#pragma unroll
            for (int u = 0; u < TU; u++) {
                const int p = p0 + u * 256;
                ev[u] = p < n ? nbr[(size_t)aT * cap + list_index(p, nF, cap)] : 0u;
            }
#pragma unroll
            for (int u = 0; u < TU; u++) {
                const int p = p0 + u * 256;
                if (p >= n) break;
                const uint32_t e = ev[u];
                const int b = (int)(e & IDX_MASK);
                if (p < nF) {
                    const uint32_t cls = (e >> CLS_SHIFT) & 3u;
                    const int sl = lds_slot_b(hk, b);
                    if (sl >= 0 && !(e >> 31) && (cls == CLS_INTERIOR || cls == CLS_CLOSE)) {
                        atomicOr(&hm[sl], 1u << (k + (cls == CLS_CLOSE ? 8 : 0)));
                    } else {
                        long long r0, r1;
                        row_range(b, r0, r1);
                        const long long r = atomicAdd(&tfill[b], 1);
                        if (r < r1 - r0) tadj[r0 + r] = (int)((uint32_t)aT | (e & ~IDX_MASK));
                    }
                } else if (stage && is_staged_back(e)) {
                    const int sl = lds_slot_b(hk, b);
                    if (sl >= 0) atomicAdd(&hm[sl], 1u << 16);
                }
            }
        }
    }
    __syncthreads();
    // 2: offsets of the back ranges (slots 16 t .. 16 t + 15 per thread, in slot order)
    int nstaged = 0;
    if (stage) {
        unsigned w[16];
        int loc = 0;
#pragma unroll
        for (int j = 0; j < 16; j++) {
            w[j] = hm[16 * tid + j];
            loc += (int)((w[j] >> 16) & 15u);
        }
        int off = block_exscan256(loc, scan_tmp, &nstaged);
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const int c = (int)((w[j] >> 16) & 15u);
            hm[16 * tid + j] = (w[j] & 0xfffffu) | ((unsigned)off << 20);
            off += c;
        }
        __syncthreads();
    }
    // back entries: into the slot's LDS range (staged) or to their rows (per entry)
    for (int k = 0; k < nk; k++) {
        const int aT = sa[k], nF = sF[k], nB = sN[k] - nF;
        const long long co = sco[k];
        for (int j0 = tid; j0 < nB; j0 += TU * 256) {
            uint32_t ev[TU];
#pragma unroll
            for (int u = 0; u < TU; u++) {
                const int j = j0 + u * 256;
                ev[u] = j < nB ? nbr[(size_t)aT * cap + cap - 1 - j] : 0u;
            }
#pragma unroll
            for (int u = 0; u < TU; u++) {
                const int j = j0 + u * 256;
                if (j >= nB) break;
                const uint32_t e = ev[u];
                const int b = (int)(e & IDX_MASK);
                const int sl = stage && is_staged_back(e) ? lds_find_b(hk, b) : -1;
                if (sl >= 0) {
                    const unsigned old = atomicSub(&hm[sl], 1u << 16);
                    bs[(old >> 20) + ((old >> 16) & 15u) - 1u] = (unsigned short)((unsigned)k | ((e >> 31) << 3) | ((unsigned)j << 4));
                } else {
                    long long r0, r1;
                    row_range(b, r0, r1);
                    const long long r = atomicAdd(&tfillB[b], 1);
                    if (r >= r1 - r0) continue;
                    const long long q = r1 - 1 - r;
                    tcut[q] = (int)(co + j);
                    tadj[q] = (int)((uint32_t)aT | (e & ~IDX_MASK));
                }
            }
        }
    }
    __syncthreads();
    // 3: reserve and write the runs, RU slots per thread at a time (their atomics in flight
    // together)
    constexpr int RU = 4;
    for (int i0 = tid; i0 < TH; i0 += RU * 256) {
        int bb[RU], rf[RU], rb[RU];
#pragma unroll
        for (int u = 0; u < RU; u++) {
            const int i = i0 + u * 256;
            bb[u] = hk[i];
            rf[u] = rb[u] = 0;
            if (bb[u] < 0) continue;
            const unsigned w = hm[i];
            const int f = __popc(w & 0xffffu);
            const int kb = stage ? (i + 1 < TH ? (int)(hm[i + 1] >> 20) : nstaged) - (int)(w >> 20) : 0;
            if (f) rf[u] = atomicAdd(&tfill[bb[u]], f);
            if (kb) rb[u] = atomicAdd(&tfillB[bb[u]], kb);
        }
#pragma unroll
        for (int u = 0; u < RU; u++) {
            const int i = i0 + u * 256, b = bb[u];
            if (b < 0) continue;
            const unsigned w = hm[i], mf = w & 0xffffu;
            const int base = (int)(w >> 20);
            const int kb = stage ? (i + 1 < TH ? (int)(hm[i + 1] >> 20) : nstaged) - base : 0;
            long long r0, r1;
            row_range(b, r0, r1);
            const long long room = r1 - r0;
            unsigned mm = (mf | (mf >> 8)) & 0xffu;
            for (long long r = rf[u]; mm; mm &= mm - 1u, r++) {
                const int k = __ffs(mm) - 1;
                const uint32_t cls = (mf >> (8 + k)) & 1u ? CLS_CLOSE : CLS_INTERIOR;
                if (r < room) tadj[r0 + r] = (int)((uint32_t)sa[k] | (cls << CLS_SHIFT));
            }
            for (int c = 0; c < kb; c++) {
                const long long r = (long long)rb[u] + c;
                if (r >= room) break;
                const unsigned v = bs[base + c];
                const int k = (int)(v & 7u);
                const long long q = r1 - 1 - r;
                tcut[q] = (int)(sco[k] + (v >> 4));
                tadj[q] = (int)((uint32_t)sa[k] | ((v >> 3) & 1u) << 31);
            }
        }
    }
}

// Padded transposed rows: row b starts at b * capT.
__global__ void k_tptr_padded(int nE, int capT, long long* tptr) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b <= nE) tptr[b] = (long long)b * capT;
}
// A padded row whose front and back entries do not fit its capacity: overflow flag 16.
__global__ void k_tcheck(int nE, int capT, const int* tF, const int* tB, unsigned long long* counters) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nE && (long long)tF[b] + tB[b] > capT) atomicOr(&counters[CNT_OVERFLOW], 16ull);
}

// ======================================================================= pattern
struct PatArgs {
    int nNodes, is_DG, cap;
    const int* tri;
    const long long* vptr;
    const int* vel;
    const uint32_t* nbr;
    const int* cntF;
    const int* cntB;
    const long long* tptr;
    const int* tadj;
    const int* tF;   // front (interior, close) entries of each transposed row
    const int* tB;   // back (cut) entries of each transposed row
    int EH, DH, SB;  // element hash, dof hash, sort buffer sizes (powers of two)
    int* tmp;        // unsorted rows, RC entries each
    int RC;
    int whalf;       // bitmap element set: element U of row r is bit U - (T0 - whalf), T0 = first element at r
    int EL;          // bitmap mode: capacity of the compacted element list (a multiple of 256)
    int* rowcnt;
    const long long* rowptr;
    int* col;
    unsigned long long* counters;
};

// Row r of the full pattern = union over elements T holding node r, and over
// U in N(T) u N^T(T), of nodes(T) u nodes(U).  One block of 256 threads per node row: an LDS
// hash of the elements, then of their nodes; each new node is appended (unsorted) to the
// row's slot of tmp (RC entries per row) and the row length to rowcnt.
// BM: the element set is a bitmap over an index window around the row's first element
// (EH words; an element outside it reports overflow 8 and the host reruns with the hash).
// pattern_row builds row r with its tables eh (EH ints) and dh (DH ints, then EL ints in bitmap
// mode) in LDS (k_pattern) or, for rows too large for LDS, in a global-memory slice per block
// (k_pattern_g, a grid of a few blocks per CU looping over the rows).
template <bool BM>
NLFEM_DEV void pattern_row(const PatArgs& a, const int r, int* eh, int* dh) {
    __shared__ int nDof, ovf;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    for (int i = tid; i < a.EH; i += 256) eh[i] = BM ? 0 : -1;
    for (int i = tid; i < a.DH; i += 256) dh[i] = -1;
    if (tid == 0) { nDof = 0; ovf = 0; }
    __syncthreads();
    long long t0, t1;
    if (a.is_DG) { t0 = r / 3; t1 = t0 + 1; }
    else { t0 = a.vptr[r]; t1 = a.vptr[r + 1]; }
    const int base = BM ? (a.is_DG ? (int)t0 : a.vel[t0]) - a.whalf : 0;
    auto einsert = [&](int U) {
        if constexpr (BM) {
            // most inserts repeat an element (the six elements of a row share most of their
            // lists): a plain read first, which lanes on one word share, and the atomic only
            // for a bit not yet set
            const int idx = U - base;
            if ((unsigned)idx >= (unsigned)(32 * a.EH)) ovf = 8;
            else if (!(((unsigned)eh[idx >> 5] >> (idx & 31)) & 1u)) atomicOr((unsigned*)&eh[idx >> 5], 1u << (idx & 31));
        } else {
            if (lds_insert(eh, a.EH, U) < 0) ovf = 1;
        }
    };
    for (long long ti = t0; ti < t1; ti++) {
        const int T = a.is_DG ? (int)ti : a.vel[ti];
        const int nF = a.cntF[T], n = nF + a.cntB[T];
        const long long tb = a.tptr[T], te = a.tptr[T + 1];
        const int ntF = a.tF[T], nt = ntF + a.tB[T];
        if (n + nt > 0 && tid == 0) einsert(T);
        // the entries in groups of PU per thread, their loads issued before the inserts (the
        // loop is bound by the latency of these loads; on the h = 1/256 unit square with
        // delta = 0.05: one per group 2.63 ms, two 2.27, four 2.47, eight 2.62)
        const int tot = n + nt;
        constexpr int PU = 2;
        for (int p0 = tid; p0 < tot; p0 += PU * 256) {
            int Us[PU];
#pragma unroll
            for (int u = 0; u < PU; u++) {
                const int p = p0 + u * 256;
                int U = -1;
                if (p < n) U = (int)(a.nbr[(size_t)T * a.cap + list_index(p, nF, a.cap)] & IDX_MASK);
                else if (p < tot) {
                    const int e = p - n;
                    U = (int)((uint32_t)a.tadj[e < ntF ? tb + e : te - 1 - (e - ntF)] & IDX_MASK);
                }
                Us[u] = U;
            }
#pragma unroll
            for (int u = 0; u < PU; u++)
                if (Us[u] >= 0) einsert(Us[u]);
        }
    }
    __syncthreads();
    int* row = a.tmp + (size_t)r * a.RC;
    int nslots = a.EH;
    const int* elist = eh;
    if constexpr (BM) {
        // compact the set bits into a list of elements (block-wide exclusive scan of the counts)
        int* el = dh + a.DH;
        __shared__ int wtot[4];
        __shared__ int ntotal;
        int cnt = 0;
        for (int i = tid; i < a.EH; i += 256) cnt += __popc((unsigned)eh[i]);
        int v = cnt;
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(v, off, 64);
            if (lane >= off) v += y;
        }
        if (lane == 63) wtot[tid >> 6] = v;
        __syncthreads();
        int woff = 0;
        for (int w = 0; w < (tid >> 6); w++) woff += wtot[w];
        if (tid == 255) ntotal = woff + v;
        int o = woff + v - cnt;
        if (woff + v > a.EL) ovf = 1;
        else
            for (int i = tid; i < a.EH; i += 256) {
                unsigned bits = (unsigned)eh[i];
                while (bits) {
                    const int b = __ffs(bits) - 1;
                    bits &= bits - 1;
                    el[o++] = base + 32 * i + b;
                }
            }
        __syncthreads();
        nslots = (min(ntotal, a.EL) + 255) & ~255;
        for (int i = ntotal + tid; i < nslots; i += 256) el[i] = -1;
        __syncthreads();
        elist = el;
    }
    // node set.  CG rows whose nodes span at most 32 DH indices: a bitmap of that window in the
    // DH words, compacted in index order, so the row comes out sorted; otherwise the LDS hash.
    __shared__ int nlo, nhi;
    if (tid == 0) { nlo = INT_MAX; nhi = -1; }
    __syncthreads();
    if (!a.is_DG) {
        int mylo = INT_MAX, myhi = -1;
        for (int i = tid; i < nslots; i += 256) {
            const int U = elist[i];
            if (U < 0) continue;
            for (int j = 0; j < 3; j++) {
                const int d = a.tri[3 * U + j];
                mylo = min(mylo, d);
                myhi = max(myhi, d);
            }
        }
        for (int m = 32; m > 0; m >>= 1) {
            mylo = min(mylo, __shfl_xor(mylo, m, 64));
            myhi = max(myhi, __shfl_xor(myhi, m, 64));
        }
        if (lane == 0) { atomicMin(&nlo, mylo); atomicMax(&nhi, myhi); }
    }
    __syncthreads();
    const long long nspan = (long long)nhi - nlo + 1;
    if (!a.is_DG && nhi >= nlo && nspan <= 32LL * a.DH) {
        const int nw = (int)((nspan + 31) / 32);
        for (int i = tid; i < nw; i += 256) dh[i] = 0;
        __syncthreads();
        for (int i = tid; i < nslots; i += 256) {
            const int U = elist[i];
            if (U < 0) continue;
            for (int j = 0; j < 3; j++) {
                const unsigned d = (unsigned)(a.tri[3 * U + j] - nlo);
                atomicOr((unsigned*)&dh[d >> 5], 1u << (d & 31));
            }
        }
        __syncthreads();
        const int per = (nw + 255) / 256, w0 = min(tid * per, nw), w1 = min(w0 + per, nw);
        int c = 0;
        for (int i = w0; i < w1; i++) c += __popc((unsigned)dh[i]);
        __shared__ int stmp[4];
        int tot;
        int q = block_exscan256(c, stmp, &tot);
        for (int i = w0; i < w1; i++) {
            unsigned bits = (unsigned)dh[i];
            while (bits) {
                const int b = __ffs(bits) - 1;
                bits &= bits - 1;
                if (q < a.RC) row[q] = nlo + 32 * i + b;
                q++;
            }
        }
        if (tid == 0) nDof = tot;
    } else
    for (int i = tid; i < nslots; i += 256) {  // a multiple of 256: uniform trip count
        const int U = elist[i];
        for (int j = 0; j < 3; j++) {
            bool isnew = false;
            int d = -1;
            if (U >= 0) {
                d = node_of(a.tri, U, j, a.is_DG);
                const int rr = lds_insert(dh, a.DH, d);
                if (rr < 0) ovf = 1;
                isnew = rr == 1;
            }
            const unsigned long long m = __ballot(isnew);
            int base = 0;
            if (lane == 0 && m) base = atomicAdd(&nDof, __popcll(m));
            base = __shfl(base, 0, 64);
            if (isnew) {
                const int q = base + __popcll(m & ((1ull << lane) - 1ull));
                if (q < a.RC) row[q] = d;
            }
        }
    }
    __syncthreads();
    if (tid == 0) {
        const int n = nDof;
        if (ovf) {
            atomicOr(&a.counters[CNT_OVERFLOW], ovf == 8 ? 8ull : 2ull);
            a.rowcnt[r] = 0;
        } else {
            if (n > a.RC) atomicOr(&a.counters[CNT_OVERFLOW], 4ull);
            a.rowcnt[r] = n;
            atomicMax(&a.counters[CNT_MAXROW], (unsigned long long)n);
        }
    }
}

template <bool BM>
__global__ __launch_bounds__(256) void k_pattern(PatArgs a) {
    extern __shared__ int lds[];
    pattern_row<BM>(a, blockIdx.x, lds, lds + a.EH);
}

// Hashed element and node sets in global memory: block b owns the slice gtab[b (EH + DH) ..).
__global__ __launch_bounds__(256) void k_pattern_g(PatArgs a, int* gtab) {
    int* eh = gtab + (size_t)blockIdx.x * (size_t)(a.EH + a.DH);
    for (int r = blockIdx.x; r < a.nNodes; r += gridDim.x) {
        pattern_row<false>(a, r, eh, eh + a.EH);
        __syncthreads();
    }
}

// Sorted column indices of each row from its unsorted slot in tmp: LDS bitonic sort.
__global__ __launch_bounds__(256) void k_pattern_sort(PatArgs a) {
    extern __shared__ int sb[];
    const int r = blockIdx.x;
    const int tid = threadIdx.x;
    const long long o = a.rowptr[r];
    const int n = (int)(a.rowptr[r + 1] - o);
    const int* row = a.tmp + (size_t)r * a.RC;
    // rows built from a node bitmap are sorted already
    __shared__ int unsorted;
    if (tid == 0) unsorted = 0;
    __syncthreads();
    for (int i = tid; i + 1 < n; i += 256)
        if (row[i] > row[i + 1]) unsorted = 1;
    __syncthreads();
    if (!unsorted) {
        for (int i = tid; i < n; i += 256) a.col[o + i] = row[i];
        return;
    }
    int m = 1;
    while (m < n) m <<= 1;
    for (int i = tid; i < m; i += 256) sb[i] = i < n ? row[i] : 0x7fffffff;
    __syncthreads();
    for (int k = 2; k <= m; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = tid; i < m; i += 256) {
                int ixj = i ^ j;
                if (ixj > i) {
                    int x = sb[i], y = sb[ixj];
                    bool up = ((i & k) == 0);
                    if ((x > y) == up) { sb[i] = y; sb[ixj] = x; }
                }
            }
            __syncthreads();
        }
    }
    for (int i = tid; i < n; i += 256) a.col[o + i] = sb[i];
}

// Position of column c in node row r (binary search in the sorted row); -1 if absent.
NLFEM_DEV long long find_pos(const long long* rowptr, const int* col, int r, int c) {
    long long lo = rowptr[r], hi = rowptr[r + 1] - 1;
    while (lo <= hi) {
        long long mid = (lo + hi) >> 1;
        int v = col[mid];
        if (v == c) return mid;
        if (v < c) lo = mid + 1;
        else hi = mid - 1;
    }
    return -1;
}

// Position of column c in node row r, searched from a guessed rank g: one probe, then a
// galloping search towards c and a binary search inside the bracket it finds; -1 if absent.
NLFEM_DEV long long find_pos_from(const long long* rowptr, const int* col, int r, int c, long long g) {
    const long long lo = rowptr[r], n = rowptr[r + 1] - lo;
    if (n <= 0) return -1;
    g = g < 0 ? 0 : (g >= n ? n - 1 : g);
    const int v = col[lo + g];
    if (v == c) return lo + g;
    long long a, b;  // c lies strictly between positions a and b of the row, -1 <= a < b <= n
    if (v < c) {
        a = g;
        long long step = 1;
        while (a + step < n && col[lo + a + step] < c) { a += step; step <<= 1; }
        b = a + step < n ? a + step : n;
        if (b < n && col[lo + b] == c) return lo + b;
    } else {
        b = g;
        long long step = 1;
        while (b - step >= 0 && col[lo + b - step] > c) { b -= step; step <<= 1; }
        a = b - step >= 0 ? b - step : -1;
        if (a >= 0 && col[lo + a] == c) return lo + a;
    }
    long long l = a + 1, h = b - 1;
    while (l <= h) {
        const long long m = (l + h) >> 1;
        const int x = col[lo + m];
        if (x == c) return lo + m;
        if (x < c) l = m + 1;
        else h = m - 1;
    }
    return -1;
}

__global__ void k_row_of(int nNodes, const long long* rowptr, int* rowOf) {
    int r = blockIdx.x;
    for (long long p = rowptr[r] + threadIdx.x; p < rowptr[r + 1]; p += blockDim.x) rowOf[p] = r;
}
// The pattern is symmetric: entry p = (r, c) with r < c finds its transpose q = (c, r) by a
// binary search in row c and sets both tpos[p] = q and tpos[q] = p; diagonal entries are
// their own transpose, and entries below the diagonal are set by their partner.
__global__ void k_tpos(long long nnz, const long long* rowptr, const int* col, const int* rowOf, int* tpos,
                       unsigned long long* counters) {
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nnz) return;
    const int r = rowOf[p], c = col[p];
    if (r > c) return;
    if (r == c) { tpos[p] = (int)p; return; }
    // rows of neighbouring nodes share the neighbourhood's offsets, so the rank of r in row c
    // mirrors the rank of c in row r: len(c) - 1 - rank_r(c) (exact for interior nodes of a
    // regular mesh, close otherwise)
    const long long g = (rowptr[c + 1] - rowptr[c]) - 1 - (p - rowptr[r]);
    long long q = find_pos_from(rowptr, col, c, r, g);
    if (q < 0) { atomicOr(&counters[CNT_ERROR], 4ull); tpos[p] = -1; return; }
    tpos[p] = (int)q;
    tpos[q] = (int)p;
}
__global__ void k_posTT(int nE, const int* tri, int is_DG, const long long* rowptr, const int* col, int* posTT,
                        unsigned long long* counters) {
    int T = blockIdx.x * blockDim.x + threadIdx.x;
    if (T >= nE) return;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            long long q = find_pos(rowptr, col, node_of(tri, T, i, is_DG), node_of(tri, T, j, is_DG));
            posTT[9 * T + 3 * i + j] = (int)q;  // -1 when T takes part in no pair
        }
}

// ======================================================================= pair quadrature
// Cut-pair surrogate (mlp.hpp): networks, acceptance rule, exact full-triangle moments of
// the inner rule, the per-element lists of pairs handed back to the exact method, and the
// sample buffers of the training-data mode.
struct SurArgs {
    const mlp::bf16x4* wf;  // [nnet][NFRAG_FWD * 64]
    const float* bias;      // [nnet][BIAS_PAD]
    const float* norm;      // [nnet][NNORM]
    int nnet;               // networks in the ensemble (1 or 2)
    float tol;              // largest accepted spread of the ensemble (ratio units)
    double smin, smax;      // domain: sqrt|det bT| / delta
    double q0, q1[3], q2[6];  // sum_i dy_i, sum_i dy_i psi_b(P_i), sum_i dy_i psi_a psi_b(P_i) (a <= b)
    int* rlist;             // rejected pairs: rlist[roff[aT] + r], r < rcnt[aT]
    const long long* roff;
    int* rcnt;
    int list_in;            // the exact cut kernel reads its pairs from rlist / rcnt
    float* sfeat;           // training samples [smaxn][12] and targets [smaxn][ntarg]
    float* starg;
    int ntarg;              // 9 (scalar constant kernel) or 18 (microelastic field kernel)
    long long smaxn;
    unsigned stride, salt;  // keep an item when hash(aT, bT, k, salt) % stride == 0
};

struct NumArgs {
    int nE, cap, is_DG, caps, HN, od;
    double sq_in, sq_far;  // sqdelta (1 - 1e-9), sqdelta (1 + 1e-9)
    // sums of the outer and inner rules (constant kernel: every block of an interior pair
    // factorises into c |aT| |bT| times products of these):
    // p1_a = sum_k dx_k psi_a(P_k), p2_ab = sum_k dx_k psi_a psi_b (P_k) (a <= b), pdx = sum_k dx_k,
    // q0 = sum_i dy_i, q1_b = sum_i dy_i psi_b(P_i), q2_ab = sum_i dy_i psi_a psi_b (P_i)
    double p1[3], p2[6], pdx, q0, q1[3], q2[6];
    double qm[5];  // inner-rule moments: sum_i dy_i (P0, P1, P0 P0, P0 P1, P1 P1)
    const ElemRec* elem;
    const int* tri;
    const int* vlabel;
    const uint32_t* nbr;
    const int* cntF;
    const int* cntB;
    const long long* rowptr;
    const int* col;
    const int* tpos;
    const int* posTT;
    double* val;
    QuadArgs q;
    KernelParams kp;
    unsigned long long* counters;
    // interior pairs, owner computes: inverse relation and per-(element, local vertex) partial rows
    const long long* tptr;
    const int* tadj;
    const int* tF;  // front entries of each transposed row (interior and close pairs)
    const long long* poff;
    double* P;
    SurArgs sur;
    // cut pairs: termLocalPrime staged per pair (6 symmetric values at (coff[aT] + p) * 6) for
    // k_lp_gather; null: added to the matrix directly
    const long long* coff;
    double* stage;
    // domain-decomposition weights (null: 1): subdomain rows of the elements, sorted
    const long long* zptr;
    const int* zidx;
    // mirror sums of k_cutw (NC summed components per node-pair entry), merged by k_mirror
    double* mir;
};

// Weight of pair (a, b) under a domain decomposition: 1 / |Z_a n Z_b|, the number of subdomains
// the two elements share (evaluateZeta, src/mathhelpers.cpp:19-38, applied to all four blocks at
// src/Cassemble.cpp:648-656,741-766); an empty intersection gives 1/0 = inf, as there.
NLFEM_DEV double zeta_weight(const long long* zp, const int* zi, int a, int b) {
    long long i = zp[a], j = zp[b];
    const long long ie = zp[a + 1], je = zp[b + 1];
    int n = 0;
    while (i < ie && j < je) {
        const int x = zi[i], y = zi[j];
        if (x < y) i++;
        else if (y < x) j++;
        else { n++; i++; j++; }
    }
    return 1.0 / (double)n;
}

// 1 / sqrt(r2) for r2 > 0: the hardware estimate (v_rsq_f64, relative error below 2^-22)
// refined by two Newton-Raphson steps y <- y + y (1 - r2 y^2) / 2, which leave an error of a
// few units in the last place (the reference's 1 / sqrt and 1 / pow(sqrt, 3) round once or
// twice; the difference is far below the FP64 tolerance of the assembled entries).
NLFEM_DEV double rsqrt_nr(double r2) {
#if defined(__HIP_DEVICE_COMPILE__)
    double y = __builtin_amdgcn_rsq(r2);
#else
    double y = 1.0 / std::sqrt(r2);
#endif
    for (int it = 0; it < 2; it++) {
        const double e = fma(-(r2 * y), y, 1.0);
        y = fma(0.5 * y, e, y);
    }
    return y;
}

// Kernel value(s) at z = x - y: OO entries (row-major od x od).  The microelastic kernels'
// 1 / |z| and 1 / |z|^3 come from rsqrt_nr in FP64.
template <int KER, int OO, typename T>
NLFEM_DEV void kernel_eval(T z0, T z1, const KernelParams& kp, T* K) {
    constexpr bool F64 = sizeof(T) == sizeof(double);
    if constexpr (KER == KER_MICRO_FIELD) {
        T cd;
        if constexpr (F64) {
            const double ri = rsqrt_nr(z0 * z0 + z1 * z1);
            cd = kp.c * (ri * ri * ri);
        } else {
            T r = sqrt(z0 * z0 + z1 * z1);
            cd = (T)kp.c * ((T)1.0 / (r * r * r));
        }
        K[0] = cd * z0 * z0;
        K[1] = cd * z0 * z1;
        K[2] = cd * z1 * z0;
        K[3] = cd * z1 * z1;
    } else {
        if constexpr (KER == KER_CONSTANT) K[0] = (T)kp.c;
        else if constexpr (KER == KER_PARABOLA) K[0] = (T)kp.c * ((T)kp.sqdelta - (z0 * z0 + z1 * z1));
        // fractional, src/model.cpp:126-138: c / |z|^(2+2s), evaluated as c 2^(-(1+s) log2 |z|^2)
        // (FP64: a few ulps from the reference's pow of the square root)
        else if constexpr (KER == KER_FRACTIONAL && F64) K[0] = kp.c * exp2(-0.5 * kp.exponent * log2(z0 * z0 + z1 * z1));
        else if constexpr (KER == KER_FRACTIONAL)
            K[0] = (T)kp.c * exp2f((float)(-0.5 * kp.exponent) * log2f(z0 * z0 + z1 * z1));
        else if constexpr (F64) K[0] = kp.c * rsqrt_nr(z0 * z0 + z1 * z1);
        else K[0] = (T)kp.c * ((T)1.0 / sqrt(z0 * z0 + z1 * z1));
    }
}

// The basis of bT on a fan triangle rT = (R0, R1, R2).  toRef is affine, so its values at
// the inner rule's points P follow from its values at the corners (the reference's pivoted
// toRef, src/mathhelpers.cpp:227-240, at each corner):
// p(P) = p(R0) + P0 (p(R1) - p(R0)) + P1 (p(R2) - p(R0)), which is what toRef(toPhys(rT, P))
// gives up to rounding (a few units in the last place of psi in [0, 1]).
struct FanBasis {
    double a0, b0, c0, a1, b1, c1;
    NLFEM_DEV FanBasis(const PivotLU& lu, const double* rT) {
        double q0[3], q1[3];
        for (int c = 0; c < 3; c++) lu.solve(rT[2 * c], rT[2 * c + 1], q0[c], q1[c]);
        a0 = q0[0]; b0 = q0[1] - q0[0]; c0 = q0[2] - q0[0];
        a1 = q1[0]; b1 = q1[1] - q1[0]; c1 = q1[2] - q1[0];
    }
    // from the corners' reference coordinates (q0[c], q1[c]) = toRef(corner c), already solved
    NLFEM_DEV FanBasis(double q00, double q10, double q01, double q11, double q02, double q12) {
        a0 = q00; b0 = q01 - q00; c0 = q02 - q00;
        a1 = q10; b1 = q11 - q10; c1 = q12 - q10;
    }
    NLFEM_DEV void at(double P0, double P1, double& p0, double& p1) const {
        p0 = fma(P1, c0, fma(P0, b0, a0));
        p1 = fma(P1, c1, fma(P0, b1, a1));
    }
};

// Accumulators of one pair.  tN is termNonloc; for symmetric kernels termNonlocPrime equals
// termNonloc bit for bit (k(y,x) is computed from the same |z| / z z^T products), so it is not
// stored.  tLP is termLocalPrime.  tLacc accumulates termLocal over all pairs of a thread.
template <int OD>
struct PairBlocks {
    static constexpr int OO = OD * OD, NB = 9 * OO;
    double tN[NB];
    double tLP[NB];
};

// Contribution of one inner quadrature point y (with weight w = dy_i * |det rT|, basis values
// psi at y, kernel K = k(x_k, y)) to the inner sums of outer point k, and to termLocalPrime.
// Follows src/integration.cpp:424-476.
// sLP holds the six symmetric products psi_a psi_b (a <= b) times k w, summed over the inner
// points of one outer point; add_lp scales them by dx_k |aT| into termLocalPrime.
// KSYM (a symmetric kernel matrix, OD = 2): component 2 equals component 1 and is not summed
// into inL / inN (the caller reads component 1 in its place).
template <int OD, typename T, bool WL = true, bool WN = true, bool WLP = true, bool KSYM = false>
NLFEM_DEV void inner_point(const T* K, T w, const T* psi, double* inL, double* inN, double* sLP) {
#pragma clang fp contract(fast)
    constexpr int OO = OD * OD;
    T kw[OO];
#pragma unroll
    for (int o = 0; o < OO; o++) {
        kw[o] = K[o] * w;
        if (WL && !(KSYM && o == 2)) inL[o] += (double)kw[o];
    }
    if (WN) {
#pragma unroll
        for (int b = 0; b < 3; b++)
#pragma unroll
            for (int o = 0; o < OO; o++)
                if (!(KSYM && o == 2)) inN[b * OO + o] += (double)(psi[b] * kw[o]);
    }
    if (WLP) {
        const T pp[6] = {psi[0] * psi[0], psi[0] * psi[1], psi[0] * psi[2], psi[1] * psi[1], psi[1] * psi[2],
                         psi[2] * psi[2]};
#pragma unroll
        for (int q = 0; q < 6; q++)
#pragma unroll
            for (int o = 0; o < OO; o++) sLP[q * OO + o] += (double)(pp[q] * kw[o]);
    }
}

// Index of the symmetric pair (a, b) among (0,0) (0,1) (0,2) (1,1) (1,2) (2,2).
NLFEM_DEV constexpr int sym6(int a, int b) {
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    return lo == 0 ? hi : (lo == 1 ? 2 + hi : 5);
}

// termLocalPrime[a*OO*3 + (o/OD)*OD*3 + b*OD + o%OD] += psi_a psi_b k'[o] dy rTdet dx_k |aT|
// (src/integration.cpp:462-476), with the inner sum taken first.
template <int OD>
NLFEM_DEV void add_lp(const double* sLP, double wk, double* tLP) {
#pragma clang fp contract(fast)
    constexpr int OO = OD * OD;
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++)
#pragma unroll
            for (int o = 0; o < OO; o++)
                tLP[a * OO * 3 + (o / OD) * OD * 3 + b * OD + o % OD] += wk * sLP[sym6(a, b) * OO + o];
}

// Outer accumulation after outer point k (src/integration.cpp:508-537):
// termLocal += |aT| psi_a(x) psi_b(x) dx_k innerLocal, termNonloc += |aT| psi_a(x) dx_k innerNonloc.
template <int OD, bool WL = true>
NLFEM_DEV void outer_point(const double* psx, double ak, const double* inL, const double* inN, double* tL,
                           double* tN) {
#pragma clang fp contract(fast)
    constexpr int OO = OD * OD, N = 3 * OD;
#pragma unroll
    for (int a = 0; a < N; a++)
#pragma unroll
        for (int b = 0; b < N; b++) {
            if (WL) tL[N * a + b] += ak * psx[a / OD] * psx[b / OD] * inL[OD * (a % OD) + (b % OD)];
            tN[N * a + b] += ak * psx[a / OD] * inN[(a % OD) * OD + OO * (b / OD) + (b % OD)];
        }
}

// Interior pair: the retriangulation of B(x_k,delta) n bT is bT itself for every k, so
// the inner points are bT(Py_i) with basis values psiy_i (src/integration.cpp:383-538 with
// Rdx = 1 and reTriangle = bT).
template <int KER, int OD, typename T, bool WL = true, bool WLP = true>
NLFEM_DEV void pair_interior(const NumArgs& a, const ElemRec& A, const ElemRec& B, double* tL, PairBlocks<OD>& pb) {
    constexpr int OO = OD * OD;
    for (int i = 0; i < 9 * OO; i++) { pb.tN[i] = 0; pb.tLP[i] = 0; }
    const double absA = A.absDet, absB = B.absDet;
    for (int k = 0; k < a.q.nPx; k++) {
        double x[2];
        toPhys(A.E, a.q.Px[2 * k], a.q.Px[2 * k + 1], x);
        double inL[OO], inN[3 * OO];
        for (int o = 0; o < OO; o++) inL[o] = 0;
        for (int o = 0; o < 3 * OO; o++) inN[o] = 0;
        const double wk = a.q.dx[k] * absA;
        double sLP[6 * OO];
        for (int o = 0; o < 6 * OO; o++) sLP[o] = 0;
        for (int i = 0; i < a.q.nPy; i++) {
            double y[2];
            toPhys(B.E, a.q.Py[2 * i], a.q.Py[2 * i + 1], y);
            T K[OO];
            kernel_eval<KER, OO, T>((T)(x[0] - y[0]), (T)(x[1] - y[1]), a.kp, K);
            T psi[3] = {(T)a.q.psiy[3 * i], (T)a.q.psiy[3 * i + 1], (T)a.q.psiy[3 * i + 2]};
            inner_point<OD, T, WL, true, WLP>(K, (T)(a.q.dy[i] * absB), psi, inL, inN, sLP);
        }
        if (WLP) add_lp<OD>(sLP, wk, pb.tLP);
        double psx[3] = {a.q.psix[3 * k], a.q.psix[3 * k + 1], a.q.psix[3 * k + 2]};
        outer_point<OD, WL>(psx, absA * a.q.dx[k], inL, inN, tL, pb.tN);
    }
}

// General (cut) pair: streamed retriangulation per outer point, inner rule on every fan
// triangle, basis values by the inverse affine map of bT (src/integration.cpp:383-538).
template <int KER, int OD, typename T>
NLFEM_DEV void pair_cut(const NumArgs& a, const ElemRec& A, const ElemRec& B, double* tL, PairBlocks<OD>& pb,
                        unsigned& ntri) {
    constexpr int OO = OD * OD;
    for (int i = 0; i < 9 * OO; i++) { pb.tN[i] = 0; pb.tLP[i] = 0; }
    const double absA = A.absDet;
    const InvMap im = invmap(B.E);
    const PivotLU lu = pivot_lu(B.E);
    for (int k = 0; k < a.q.nPx; k++) {
        double x[2];
        toPhys(A.E, a.q.Px[2 * k], a.q.Px[2 * k + 1], x);
        double inL[OO], inN[3 * OO];
        for (int o = 0; o < OO; o++) inL[o] = 0;
        for (int o = 0; o < 3 * OO; o++) inN[o] = 0;
        const double wk = a.q.dx[k] * absA;
        double sLP[6 * OO];
        for (int o = 0; o < 6 * OO; o++) sLP[o] = 0;
        auto emit = [&](double r0, double r1, double s0, double s1, double t0, double t1) {
            const double rT[6] = {r0, r1, s0, s1, t0, t1};
            const double rTdet = absDet(rT);
            ntri++;
            const FanBasis fb(lu, rT);
            for (int i = 0; i < a.q.nPy; i++) {
                double y[2];
                toPhys(rT, a.q.Py[2 * i], a.q.Py[2 * i + 1], y);
                T K[OO];
                kernel_eval<KER, OO, T>((T)(x[0] - y[0]), (T)(x[1] - y[1]), a.kp, K);
                double p0, p1;
                if constexpr (sizeof(T) == sizeof(double)) {
                    fb.at(a.q.Py[2 * i], a.q.Py[2 * i + 1], p0, p1);
                } else {
                    const double d0 = y[0] - im.e0, d1 = y[1] - im.e1;
                    p0 = im.i00 * d0 + im.i01 * d1;
                    p1 = im.i10 * d0 + im.i11 * d1;
                }
                T psi[3] = {(T)(1 - p0 - p1), (T)p0, (T)p1};
                inner_point<OD, T>(K, (T)(a.q.dy[i] * rTdet), psi, inL, inN, sLP);
            }
        };
        retriangulate<false>(x[0], x[1], B.E, a.kp.sqdelta, a.caps != 0, emit);
        add_lp<OD>(sLP, wk, pb.tLP);
        double psx[3] = {a.q.psix[3 * k], a.q.psix[3 * k + 1], a.q.psix[3 * k + 2]};
        outer_point<OD>(psx, absA * a.q.dx[k], inL, inN, tL, pb.tN);
    }
}


// baryCenter pair (src/integration.cpp:910-1018 with method_baryCenter :1262-1280): the whole
// inner element counts when its barycentre is in B(x_k, delta); factor 2 on termLocal and
// termNonloc, the Prime blocks stay zero.
template <int KER, int OD, typename T>
NLFEM_DEV void pair_bary(const NumArgs& a, const ElemRec& A, const ElemRec& B, double* tL, PairBlocks<OD>& pb) {
    constexpr int OO = OD * OD;
    for (int i = 0; i < 9 * OO; i++) { pb.tN[i] = 0; pb.tLP[i] = 0; }
    const double absA = A.absDet, absB = B.absDet;
    double dummy[9 * OO];
    for (int k = 0; k < a.q.nPx; k++) {
        double x[2];
        toPhys(A.E, a.q.Px[2 * k], a.q.Px[2 * k + 1], x);
        double inL[OO], inN[3 * OO];
        for (int o = 0; o < OO; o++) inL[o] = 0;
        for (int o = 0; o < 3 * OO; o++) inN[o] = 0;
        if (bary_interacts(x[0], x[1], B.E, a.kp.sqdelta)) {
            for (int i = 0; i < a.q.nPy; i++) {
                double y[2];
                toPhys(B.E, a.q.Py[2 * i], a.q.Py[2 * i + 1], y);
                T K[OO];
                kernel_eval<KER, OO, T>((T)(x[0] - y[0]), (T)(x[1] - y[1]), a.kp, K);
                T psi[3] = {(T)a.q.psiy[3 * i], (T)a.q.psiy[3 * i + 1], (T)a.q.psiy[3 * i + 2]};
                inner_point<OD, T, true, true, false>(K, (T)(a.q.dy[i] * absB), psi, inL, inN, dummy);
            }
        }
        double psx[3] = {a.q.psix[3 * k], a.q.psix[3 * k + 1], a.q.psix[3 * k + 2]};
        outer_point<OD>(psx, 2 * absA * a.q.dx[k], inL, inN, tL, pb.tN);
    }
}

NLFEM_DEV double shfl_xor_d(double v, int m) { return __shfl_xor(v, m, 64); }

// Scatter of the blocks of pair (aT, b) (src/Cassemble.cpp:721-770) into the node-CSR
// values with FP64 atomics.  Positions of (row rA[i], col node) come from the LDS map of the
// three rows of aT, transposed entries through tpos, inner-element entries through posTT.
template <int OD>
NLFEM_DEV void scatter_pair(const NumArgs& a, const int* rA, const bool* labA, int b, bool prime,
                            const PairBlocks<OD>& pb, const int* hk, const int* hv, bool useHash) {
    constexpr int OO = OD * OD, N = 3 * OD;
    int rB[3];
    bool labB[3];
    long long pos[3][3];  // pos[i][j]: entry (rA[i], rB[j])
    for (int j = 0; j < 3; j++) {
        rB[j] = node_of(a.tri, b, j, a.is_DG);
        labB[j] = row_labelled(a.vlabel, rB[j], a.is_DG);
    }
    for (int j = 0; j < 3; j++) {
        int s = useHash ? lds_find(hk, a.HN, rB[j]) : -1;
        for (int i = 0; i < 3; i++) {
            long long p;
            if (useHash) p = (s >= 0) ? a.rowptr[rA[i]] + hv[3 * s + i] : -1;
            else p = find_pos(a.rowptr, a.col, rA[i], rB[j]);
            pos[i][j] = p;
            if (p < 0) atomicOr(&a.counters[CNT_ERROR], 8ull);
        }
    }
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) {
            const long long p = pos[i][j];
            if (p < 0) continue;
            if (labA[i]) {  // -termNonloc into (aAdx[i], bAdx[j])
#pragma unroll
                for (int ao = 0; ao < OD; ao++)
#pragma unroll
                    for (int bo = 0; bo < OD; bo++)
                        atomicAdd(&a.val[p * OO + ao * OD + bo], -pb.tN[N * (i * OD + ao) + j * OD + bo]);
            }
            if (prime && labB[j]) {  // -termNonlocPrime into (bAdx[j], aAdx[i])
                const long long q = a.tpos[p];
#pragma unroll
                for (int ao = 0; ao < OD; ao++)
#pragma unroll
                    for (int bo = 0; bo < OD; bo++)
                        atomicAdd(&a.val[q * OO + bo * OD + ao], -pb.tN[N * (i * OD + ao) + j * OD + bo]);
            }
        }
    }
    if (prime) {  // +termLocalPrime into (bAdx[j], bAdx[i])
        for (int j = 0; j < 3; j++) {
            if (!labB[j]) continue;
            for (int i = 0; i < 3; i++) {
                const long long p = a.posTT[9 * b + 3 * j + i];
#pragma unroll
                for (int ao = 0; ao < OD; ao++)
#pragma unroll
                    for (int bo = 0; bo < OD; bo++)
                        atomicAdd(&a.val[p * OO + bo * OD + ao], pb.tLP[N * (i * OD + ao) + j * OD + bo]);
            }
        }
    }
}


// Scatter of the blocks of a chunk of cut pairs (src/Cassemble.cpp:721-770), one thread per
// (pair, i, j) of a 256-thread block: -termNonloc into (aAdx[i], bAdx[j]) and, transposed,
// into (bAdx[j], aAdx[i]); +termLocalPrime into (bAdx[j], bAdx[i]).  Positions of row aAdx[i]
// come from the block's LDS column map (hk, hv), transposed entries through tpos, inner
// element entries through posTT.  Pairs with pbs[p] < 0 are skipped.
template <int OD, typename HV>
NLFEM_DEV void cut_scatter(const NumArgs& a, int np, const int* pbs, const double* pacc, const int* rA,
                           const bool* labA, const int* hk, const HV* hv) {
    constexpr int OO = OD * OD, NB = 9 * OO, N = 3 * OD;
    // one thread per (pair, block (i, j), component (ao, bo))
    for (int t = threadIdx.x; t < np * 9 * OO; t += 256) {
        const int p = t / (9 * OO), r = t - p * 9 * OO;
        const int ij = r / OO, o = r - ij * OO;
        const int i = ij / 3, j = ij % 3, ao = o / OD, bo = o % OD;
        const int b = pbs[p];
        if (b < 0) continue;
        const double* tN = pacc + p * 2 * NB;
        const double* tLP = tN + NB;
        const int rBj = node_of(a.tri, b, j, a.is_DG);
        const bool labBj = row_labelled(a.vlabel, rBj, a.is_DG);
        const int s = lds_find(hk, a.HN, rBj);
        if (s < 0) { atomicOr(&a.counters[CNT_ERROR], 8ull); continue; }
        const long long pos = a.rowptr[rA[i]] + hv[3 * s + i];  // entry (aAdx[i], bAdx[j])
        const double v = tN[N * (i * OD + ao) + j * OD + bo];
        if (labA[i]) atomicAdd(&a.val[pos * OO + ao * OD + bo], -v);
        if (labBj) {
            atomicAdd(&a.val[(long long)a.tpos[pos] * OO + bo * OD + ao], -v);
            // +termLocalPrime[N (i OD+ao) + j OD+bo] into (bAdx[j], bAdx[i])
            const long long q = a.posTT[9 * b + 3 * j + i];
            atomicAdd(&a.val[q * OO + bo * OD + ao], tLP[N * (i * OD + ao) + j * OD + bo]);
        }
    }
}

// termLocal of outer element aT: per-thread sums reduced over the 256-thread block, then
// added into (aAdx[i], aAdx[j]).
template <int OD>
NLFEM_DEV void cut_flush_local(const NumArgs& a, int aT, const double* tL, const bool* labA, double (*red)[9 * OD * OD]) {
    constexpr int OO = OD * OD, NB = 9 * OO, N = 3 * OD;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
#pragma unroll
    for (int i = 0; i < NB; i++) {
        double v = tL[i];
        for (int m = 32; m > 0; m >>= 1) v += shfl_xor_d(v, m);
        if (lane == 0) red[w][i] = v;
    }
    __syncthreads();
    for (int t = tid; t < NB; t += 256) {
        const double v = red[0][t] + red[1][t] + red[2][t] + red[3][t];
        const int ra = t / N, cb = t % N;
        const int i = ra / OD, ao = ra % OD, j = cb / OD, bo = cb % OD;
        if (!labA[i]) continue;
        const long long p = a.posTT[9 * aT + 3 * i + j];
        atomicAdd(&a.val[p * OO + ao * OD + bo], v);
    }
}

// One block of 256 threads per outer element; threads stride over the element's
// neighbour-list segment.  SEG 0: front segment (interior pairs; close pairs are left to the
// tensor-Gauss kernel).  SEG 1: back segment (cut pairs, or baryCenter pairs).
// ZW: every pair weighted by its domain-decomposition weight (zeta_weight): the pair's termLocal
// goes through a per-pair buffer so that it can be scaled before it joins the thread's sum.
template <int KER, int OD, int SEG, typename T, bool ZW = false>
__global__ __launch_bounds__(256) void k_pairs(NumArgs a) {
    constexpr int OO = OD * OD, NB = 9 * OO;
    extern __shared__ int lds[];
    int* hk = lds;
    int* hv = hk + a.HN;
    __shared__ double red[4][NB];
    __shared__ unsigned subtri;
    const int aT = blockIdx.x;
    const int tid = threadIdx.x;
    const ElemRec A = a.elem[aT];
    if (A.label == 0) return;
    const int nF = a.cntF[aT];
    const int n = (SEG == 0) ? nF : a.cntB[aT];
    if (n == 0) return;
    int rA[3];
    bool labA[3];
    for (int j = 0; j < 3; j++) {
        rA[j] = node_of(a.tri, aT, j, a.is_DG);
        labA[j] = row_labelled(a.vlabel, rA[j], a.is_DG);
    }
    const bool useHash = a.HN > 0;
    if (useHash) {
        for (int i = tid; i < a.HN; i += 256) hk[i] = -1;
        __syncthreads();
        for (int j = 0; j < 3; j++) {
            const long long r0 = a.rowptr[rA[j]];
            const int len = (int)(a.rowptr[rA[j] + 1] - r0);
            for (int i = tid; i < len; i += 256) {
                int s = lds_slot(hk, a.HN, a.col[r0 + i]);
                if (s >= 0) hv[3 * s + j] = i;
                else atomicOr(&a.counters[CNT_ERROR], 16ull);
            }
        }
        __syncthreads();
    }
    double tL[NB];
#pragma unroll
    for (int i = 0; i < NB; i++) tL[i] = 0;
    unsigned ntri = 0;
    if (tid == 0) subtri = 0;
    for (int p = tid; p < n; p += 256) {
        const uint32_t e = a.nbr[(size_t)aT * a.cap + (SEG == 0 ? p : a.cap - 1 - p)];
        const uint32_t cls = (e >> CLS_SHIFT) & 3u;
        const int b = (int)(e & IDX_MASK);
        if (SEG == 0 && cls == CLS_CLOSE) continue;
        const ElemRec B = a.elem[b];
        PairBlocks<OD> pb;
        bool prime = true;
        double tLz[ZW ? NB : 1];
        double* tLp = tL;
        if constexpr (ZW) {
#pragma unroll
            for (int i = 0; i < NB; i++) tLz[i] = 0;
            tLp = tLz;
        }
        if constexpr (SEG == 0) {
            pair_interior<KER, OD, T>(a, A, B, tLp, pb);
        } else {
            if (e >> 31) {
                pair_bary<KER, OD, T>(a, A, B, tLp, pb);
                prime = false;
            } else {
                pair_cut<KER, OD, T>(a, A, B, tLp, pb, ntri);
            }
        }
        if constexpr (ZW) {
            const double w = zeta_weight(a.zptr, a.zidx, aT, b);
#pragma unroll
            for (int i = 0; i < NB; i++) {
                tL[i] += w * tLz[i];
                pb.tN[i] *= w;
                pb.tLP[i] *= w;
            }
        }
        scatter_pair<OD>(a, rA, labA, b, prime, pb, hk, hv, useHash);
    }
    // block reduction of termLocal, then +termLocal into (aAdx[i], aAdx[j])
    const int lane = tid & 63, w = tid >> 6;
    if (SEG == 1) atomicAdd(&subtri, ntri);
#pragma unroll
    for (int i = 0; i < NB; i++) {
        double v = tL[i];
        for (int m = 32; m > 0; m >>= 1) v += shfl_xor_d(v, m);
        if (lane == 0) red[w][i] = v;
    }
    __syncthreads();
    constexpr int N = 3 * OD;
    if (SEG == 1 && tid == 0) atomicAdd(&a.counters[CNT_SUBTRI], (unsigned long long)subtri);
    for (int t = tid; t < NB; t += 256) {
        const double v = red[0][t] + red[1][t] + red[2][t] + red[3][t];
        const int ra = t / N, cb = t % N;  // termLocal[N*a + b]
        const int i = ra / OD, ao = ra % OD, j = cb / OD, bo = cb % OD;
        if (!labA[i]) continue;
        const long long p = a.posTT[9 * aT + 3 * i + j];
        atomicAdd(&a.val[p * OO + ao * OD + bo], v);
    }
}

// ======================================================================= interior pairs, owner computes
// One block per element T.  Every interior pair is integrated twice, once by each of its
// elements: as outer element T supplies termLocal and termNonloc (rows of T's nodes), as
// inner element T supplies termLocalPrime and termNonlocPrime (also rows of T's nodes,
// src/Cassemble.cpp:751-768).  All four land in LDS accumulators keyed by column (one slot
// per column of the union of T's three rows); the block then writes T's three partial rows
// with plain coalesced stores into P, and k_merge adds the partial rows of the elements
// around each node.  No global atomics.
constexpr unsigned long long ERR_HASH = 128ull;

// Column map of an outer element's node rows: node id -> index into the block's column
// accumulators (< H).  Rows are sorted, so their nodes lie in [first, last] of the rows.
// Rank mode, when that window spans at most 16 H nodes: a bitmap of the window (WW words)
// and per-word prefix counts (WW ints) in the same H ints of LDS; a node's index is its rank
// among the rows' nodes, found with two LDS reads and a popcount.  Otherwise an LDS hash of
// H slots (index = slot, found by probing).  Build with colmap_build by all 256 threads.
struct ColMap {
    int* h;
    int H, base, WW;  // WW > 0: rank mode
    NLFEM_DEV int find(int n) const {
        if (WW > 0) {
            const unsigned d = (unsigned)(n - base);
            if (d >= 32u * (unsigned)WW) return -1;
            const unsigned w = (unsigned)h[d >> 5], m = 1u << (d & 31);
            const int r = (w & m) ? h[WW + (d >> 5)] + __popc(w & (m - 1)) : -1;
            return r < H ? r : -1;  // more nodes than accumulators: reported by colmap_build
        }
        return lds_find(h, H, n);
    }
};
// rows: use[j] selects which of the three rows (rp[j], rlen[j] entries of col) are mapped.
NLFEM_DEV ColMap colmap_build(int* h, int H, const int* col, const long long* rp, const int* rlen, const bool* use,
                              int* tmp, unsigned long long* counters) {
    ColMap cm;
    cm.h = h;
    cm.H = H;
    int lo = INT_MAX, hi = -1;
    for (int j = 0; j < 3; j++)
        if (use[j] && rlen[j] > 0) {
            lo = min(lo, col[rp[j]]);
            hi = max(hi, col[rp[j] + rlen[j] - 1]);
        }
    const long long span = (long long)hi - lo + 1;
    cm.base = lo;
    cm.WW = (hi >= lo && span <= 16LL * H) ? (int)((span + 31) / 32) : 0;
    const int tid = threadIdx.x;
    if (cm.WW > 0) {
        for (int i = tid; i < cm.WW; i += 256) h[i] = 0;
        __syncthreads();
        for (int j = 0; j < 3; j++)
            if (use[j])
                for (int i = tid; i < rlen[j]; i += 256) {
                    const unsigned d = (unsigned)(col[rp[j] + i] - lo);
                    atomicOr(&h[d >> 5], (int)(1u << (d & 31)));
                }
        __syncthreads();
        const int per = (cm.WW + 255) / 256, w0 = min(tid * per, cm.WW), w1 = min(w0 + per, cm.WW);
        int c = 0;
        for (int i = w0; i < w1; i++) c += __popc((unsigned)h[i]);
        int tot;
        int off = block_exscan256(c, tmp, &tot);
        for (int i = w0; i < w1; i++) {
            h[cm.WW + i] = off;
            off += __popc((unsigned)h[i]);
        }
        if (tid == 0 && tot > H) atomicOr(&counters[CNT_ERROR], ERR_HASH);
    } else {
        // the hash uses the largest power of two within the H slots (H need not be one: the
        // rank mode sizes the table by the rows, not by a power of two)
        const int Hh = 1 << (31 - __clz(H));
        cm.H = Hh;
        for (int i = tid; i < Hh; i += 256) h[i] = -1;
        __syncthreads();
        for (int j = 0; j < 3; j++)
            if (use[j])
                for (int i = tid; i < rlen[j]; i += 256)
                    if (lds_slot(h, Hh, col[rp[j] + i]) < 0) atomicOr(&counters[CNT_ERROR], ERR_HASH);
    }
    __syncthreads();
    return cm;
}

template <int KER, int OD, typename T, int NPY = 0>
__global__ __launch_bounds__(256) void k_interior(NumArgs a) {
    constexpr int OO = OD * OD, NB = 9 * OO, N = 3 * OD;
    // the field kernel with the unrolled inner rule sums k01 = k10 once, so every block it adds
    // has equal components (0, 1) and (1, 0): the accumulators keep 00, 01, 11 (NCA = 3)
    constexpr bool KS3 = NPY > 0 && KER == KER_MICRO_FIELD && OD == 2;
    constexpr int NCA = KS3 ? 3 : OO;
    auto aix = [](int ao, int bo) { return KS3 ? ao + bo : ao * OD + bo; };
    extern __shared__ double ldsd[];
    double* acc = ldsd;                                  // [HN][3][NCA]
    int* hk = (int*)(acc + (size_t)a.HN * 3 * NCA);      // [HN]
    __shared__ double red[4][2 * NB];
    const int T0 = blockIdx.x;
    const int tid = threadIdx.x;
    const ElemRec A = a.elem[T0];
    if (A.label == 0) return;
    int rA[3];
    bool labA[3];
    for (int j = 0; j < 3; j++) {
        rA[j] = node_of(a.tri, T0, j, a.is_DG);
        labA[j] = row_labelled(a.vlabel, rA[j], a.is_DG);
    }
    if (!(labA[0] || labA[1] || labA[2])) return;
    __shared__ int scan_tmp[4];
    for (int i = tid; i < a.HN * 3 * NCA; i += 256) acc[i] = 0.0;
    long long rp[3];
    int rlen[3];
    for (int j = 0; j < 3; j++) {
        rp[j] = a.rowptr[rA[j]];
        rlen[j] = (int)(a.rowptr[rA[j] + 1] - rp[j]);
    }
    const ColMap cm = colmap_build(hk, a.HN, a.col, rp, rlen, labA, scan_tmp, a.counters);
    // the scalar parabola kernel: moment sums (PFACT); NPY > 0: the other general kernels with an
    // inner rule of NPY points
    constexpr bool PFACT = KER == KER_PARABOLA && OD == 1;
    constexpr bool FIXED = NPY > 0 && !(KER == KER_CONSTANT && OD == 1) && !PFACT;
    const int nPx = a.q.nPx;
    const int lane = tid & 63, w = tid >> 6;
    if constexpr (PFACT) {
        // Parabola kernel K(x, y) = c (delta^2 - |x - y|^2) (src/model.cpp:52-60).  With x = o + xt
        // and y = o + yt about T's first vertex o, a rule sum of weights f(x_k) g(y_l) K(x_k, y_l)
        // over the outer points x_k and inner points y_l (src/integration.cpp:508-537) is
        //   c [delta^2 F0 G0 - F2 G0 + 2 F1 . G1 - F0 G2],
        // with F0, F1, F2 = sum_k f(x_k) (1, xt_k, |xt_k|^2) and G likewise over the inner points.
        // The moments of T's own points are formed once per block, the partner's per pair: about
        // 300 flop per pair and role instead of 49 kernel evaluations.  In FP64 for both
        // precisions (every term is of the size of delta^2 F0 G0; an interior pair keeps
        // |x - y| below delta, so no sum cancels by more than a few digits).
        __shared__ double FT[9][4];  // T as outer element: f = |T| dx_k psi_i psi_j (x_k) (sym6), |T| dx_k psi_i (x_k) (6 + i)
        __shared__ double GT[9][4];  // T as inner element: g = |T| dy_l psi_a psi_b (y_l) (sym6), |T| dy_l psi_j (y_l) (6 + j)
        const double o0 = A.E[0], o1 = A.E[1];
        const double c = a.kp.c, d2 = a.kp.sqdelta;
        auto combine = [&](const double* F, const double* G) {
            return c * (d2 * F[0] * G[0] - F[3] * G[0] + 2.0 * (F[1] * G[1] + F[2] * G[2]) - F[0] * G[3]);
        };
        if (tid < 18) {
            const bool outer = tid < 9;
            const int f = outer ? tid : tid - 9;
            const int np = outer ? nPx : a.q.nPy;
            // basis indices of the function: (i, j) with sym6(i, j) = f, or psi_{f - 6}
            int bi = 0, bj = 0;
            if (f < 6)
                for (int i = 0; i < 3; i++)
                    for (int j = i; j < 3; j++)
                        if (sym6(i, j) == f) { bi = i; bj = j; }
            double m[4] = {0, 0, 0, 0};
            for (int k = 0; k < np; k++) {
                const double* P = outer ? a.q.Px : a.q.Py;
                const double* ps = outer ? a.q.psix : a.q.psiy;
                double x[2];
                toPhys(A.E, P[2 * k], P[2 * k + 1], x);
                const double t0 = x[0] - o0, t1 = x[1] - o1;
                const double wgt = A.absDet * (outer ? a.q.dx[k] : a.q.dy[k]) *
                                   (f < 6 ? ps[3 * k + bi] * ps[3 * k + bj] : ps[3 * k + f - 6]);
                m[0] += wgt;
                m[1] += wgt * t0;
                m[2] += wgt * t1;
                m[3] += wgt * (t0 * t0 + t1 * t1);
            }
            double* dst = outer ? FT[f] : GT[f];
            for (int e = 0; e < 4; e++) dst[e] = m[e];
        }
        __syncthreads();
        // T as outer element: termNonloc into (T_i, U_j) per pair; termLocal (six symmetric sums)
        // is linear in the partners' moments of g = |U| dy_l, summed over the pairs (SG)
        double SG[4] = {0, 0, 0, 0};
        const int nF = a.cntF[T0];
        for (int p = tid; p < nF; p += 256) {
            const uint32_t e = a.nbr[(size_t)T0 * a.cap + p];
            if (((e >> CLS_SHIFT) & 3u) != CLS_INTERIOR) continue;
            const int U = (int)(e & IDX_MASK);
            const ElemRec B = a.elem[U];
            double G[4][4];  // g = |U| dy_l psi_j (y_l) (G[j]); G[3]: g = |U| dy_l
#pragma unroll
            for (int s = 0; s < 4; s++)
#pragma unroll
                for (int t = 0; t < 4; t++) G[s][t] = 0;
            for (int l = 0; l < a.q.nPy; l++) {
                double y[2];
                toPhys(B.E, a.q.Py[2 * l], a.q.Py[2 * l + 1], y);
                const double t0 = y[0] - o0, t1 = y[1] - o1, r2 = t0 * t0 + t1 * t1;
                const double v = a.q.dy[l] * B.absDet;
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    const double g = s == 3 ? v : v * a.q.psiy[3 * l + s];
                    G[s][0] += g;
                    G[s][1] += g * t0;
                    G[s][2] += g * t1;
                    G[s][3] += g * r2;
                }
            }
#pragma unroll
            for (int t = 0; t < 4; t++) SG[t] += G[3][t];
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const int s = cm.find(node_of(a.tri, U, j, a.is_DG));
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    if (!labA[i]) continue;
                    if (s < 0) { atomicOr(&a.counters[CNT_ERROR], ERR_HASH); continue; }
                    atomicAdd(&acc[s * 3 + i], -combine(FT[6 + i], G[j]));
                }
            }
        }
        // T as inner element: pairs (U, T).  termLocalPrime is linear in U's weight moments
        // W = sum_k |U| dx_k (1, xt_k, |xt_k|^2), summed over the pairs; termNonlocPrime into
        // (T_j, U_i) per pair from F = sum_k |U| dx_k psi_i (x_k) (1, xt_k, |xt_k|^2)
        double W[4] = {0, 0, 0, 0};
        const long long tb = a.tptr[T0], te = tb + a.tF[T0];
        for (long long q = tb + tid; q < te; q += 256) {
            const uint32_t e = (uint32_t)a.tadj[q];
            if ((e >> 31) || ((e >> CLS_SHIFT) & 3u) != CLS_INTERIOR) continue;
            const int U = (int)(e & IDX_MASK);
            const ElemRec B = a.elem[U];
            double F[3][4];
#pragma unroll
            for (int s = 0; s < 3; s++)
#pragma unroll
                for (int t = 0; t < 4; t++) F[s][t] = 0;
            for (int k = 0; k < nPx; k++) {
                double x[2];
                toPhys(B.E, a.q.Px[2 * k], a.q.Px[2 * k + 1], x);
                const double t0 = x[0] - o0, t1 = x[1] - o1, r2 = t0 * t0 + t1 * t1;
                const double wk = a.q.dx[k] * B.absDet;
                W[0] += wk;
                W[1] += wk * t0;
                W[2] += wk * t1;
                W[3] += wk * r2;
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    const double f = wk * a.q.psix[3 * k + i];
                    F[i][0] += f;
                    F[i][1] += f * t0;
                    F[i][2] += f * t1;
                    F[i][3] += f * r2;
                }
            }
#pragma unroll
            for (int i = 0; i < 3; i++) {
                const int s = cm.find(node_of(a.tri, U, i, a.is_DG));
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    if (!labA[j]) continue;
                    if (s < 0) { atomicOr(&a.counters[CNT_ERROR], ERR_HASH); continue; }
                    atomicAdd(&acc[s * 3 + j], -combine(F[i], GT[6 + j]));
                }
            }
        }
#pragma unroll
        for (int t = 0; t < 4; t++)
            for (int m = 32; m > 0; m >>= 1) {
                SG[t] += shfl_xor_d(SG[t], m);
                W[t] += shfl_xor_d(W[t], m);
            }
        if (lane == 0) {
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    red[w][3 * i + j] = combine(FT[sym6(i, j)], SG);
                    red[w][NB + 3 * i + j] = combine(W, GT[sym6(i, j)]);
                }
        }
    } else if constexpr (FIXED) {
        // Inner rule of NPY points unrolled.  Per block: the outer points xo and inner points yi
        // of T, and the outer weights wxa[k][i] = |T| dx_k psi_i(x_k), wl[k][sym(i,j)] =
        // |T| dx_k psi_i(x_k) psi_j(x_k) that outer_point applies.
        __shared__ double xo[2 * MAXQ], yi[2 * NPY], wxa[3 * MAXQ], wl[6 * MAXQ];
        if (tid < nPx) {
            toPhys(A.E, a.q.Px[2 * tid], a.q.Px[2 * tid + 1], &xo[2 * tid]);
            const double ak = A.absDet * a.q.dx[tid];
            for (int i = 0; i < 3; i++) {
                wxa[3 * tid + i] = ak * a.q.psix[3 * tid + i];
                for (int j = i; j < 3; j++) wl[6 * tid + sym6(i, j)] = ak * a.q.psix[3 * tid + i] * a.q.psix[3 * tid + j];
            }
        }
        if (tid < NPY) toPhys(A.E, a.q.Py[2 * tid], a.q.Py[2 * tid + 1], &yi[2 * tid]);
        __syncthreads();
        // the field kernel's matrix is symmetric: component 2 (k10) equals component 1 (k01), so
        // it is not summed (KS), and every block entry of component 2 reads its component-1 twin
        constexpr bool KS = KER == KER_MICRO_FIELD && OD == 2;
        auto so = [](int o) { return (KS && o == 2) ? 1 : o; };
        // position in tN of entry (row T_i comp ao, col U_j comp bo), component 2 -> its twin
        auto tni = [](int i, int ao, int j, int bo) {
            return (KS && ao == 1 && bo == 0) ? N * (i * OD) + j * OD + 1 : N * (i * OD + ao) + j * OD + bo;
        };
        // T as outer element (pair_interior with termLocal, without termLocalPrime).  termLocal
        // depends on (i, j) through psi_i psi_j only: six symmetric sums per component.
        double tL6[6 * OO];
#pragma unroll
        for (int t = 0; t < 6 * OO; t++) tL6[t] = 0;
        const int nF = a.cntF[T0];
        for (int p = tid; p < nF; p += 256) {
            const uint32_t e = a.nbr[(size_t)T0 * a.cap + p];
            if (((e >> CLS_SHIFT) & 3u) != CLS_INTERIOR) continue;
            const int U = (int)(e & IDX_MASK);
            const ElemRec B = a.elem[U];
            double yu[2 * NPY];
#pragma unroll
            for (int l = 0; l < NPY; l++) toPhys(B.E, a.q.Py[2 * l], a.q.Py[2 * l + 1], &yu[2 * l]);
            double tN[NB];
#pragma unroll
            for (int t = 0; t < NB; t++) tN[t] = 0;
            for (int k = 0; k < nPx; k++) {
                const double x0 = xo[2 * k], x1 = xo[2 * k + 1];
                double inL[OO], inN[3 * OO];
#pragma unroll
                for (int o = 0; o < OO; o++) inL[o] = 0;
#pragma unroll
                for (int o = 0; o < 3 * OO; o++) inN[o] = 0;
#pragma unroll
                for (int l = 0; l < NPY; l++) {
                    T K[OO];
                    kernel_eval<KER, OO, T>((T)(x0 - yu[2 * l]), (T)(x1 - yu[2 * l + 1]), a.kp, K);
                    const T psi[3] = {(T)a.q.psiy[3 * l], (T)a.q.psiy[3 * l + 1], (T)a.q.psiy[3 * l + 2]};
                    inner_point<OD, T, true, true, false, KS>(K, (T)(a.q.dy[l] * B.absDet), psi, inL, inN, nullptr);
                }
                {
#pragma clang fp contract(fast)
#pragma unroll
                    for (int q = 0; q < 6; q++)
#pragma unroll
                        for (int o = 0; o < OO; o++)
                            if (!(KS && o == 2)) tL6[q * OO + o] += wl[6 * k + q] * inL[o];
#pragma unroll
                    for (int r = 0; r < N; r++)
#pragma unroll
                        for (int c = 0; c < N; c++)
                            if (!(KS && r % OD == 1 && c % OD == 0))
                                tN[N * r + c] += wxa[3 * k + r / OD] * inN[(r % OD) * OD + OO * (c / OD) + (c % OD)];
                }
            }
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const int s = cm.find(node_of(a.tri, U, j, a.is_DG));
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    if (!labA[i]) continue;
                    if (s < 0) { atomicOr(&a.counters[CNT_ERROR], ERR_HASH); continue; }
#pragma unroll
                    for (int ao = 0; ao < OD; ao++)
#pragma unroll
                        for (int bo = 0; bo < OD; bo++)
                            if (!(KS3 && ao == 1 && bo == 0))
                                atomicAdd(&acc[(s * 3 + i) * NCA + aix(ao, bo)], -tN[tni(i, ao, j, bo)]);
                }
            }
        }
#pragma unroll
        for (int t = 0; t < 6 * OO; t++) {
            if (KS && t % OO == 2) continue;
            double v = tL6[t];
            for (int m = 32; m > 0; m >>= 1) v += shfl_xor_d(v, m);
            tL6[t] = v;
        }
        if (lane == 0) {
#pragma unroll
            for (int r = 0; r < N; r++)
#pragma unroll
                for (int c = 0; c < N; c++) red[w][N * r + c] = tL6[sym6(r / OD, c / OD) * OO + so(OD * (r % OD) + (c % OD))];
        }
        // T as inner element: pairs (U, T).  termLocalPrime of T sums, over the pairs and the
        // outer points x_k of U, dx_k |U| psi_a psi_b(y_l) dy_l |T| k(x_k, y_l); the factor
        // psi_a psi_b(y_l) dy_l |T| depends on l only, so M[l] = sum dx_k |U| k(x_k, y_l) is
        // accumulated and the six products are applied once.
        double M[NPY * OO];
#pragma unroll
        for (int t = 0; t < NPY * OO; t++) M[t] = 0;
        const long long tb = a.tptr[T0], te = tb + a.tF[T0];
        for (long long q = tb + tid; q < te; q += 256) {
            const uint32_t e = (uint32_t)a.tadj[q];
            if ((e >> 31) || ((e >> CLS_SHIFT) & 3u) != CLS_INTERIOR) continue;
            const int U = (int)(e & IDX_MASK);
            const ElemRec B = a.elem[U];
            double tN[NB];
#pragma unroll
            for (int t = 0; t < NB; t++) tN[t] = 0;
            for (int k = 0; k < nPx; k++) {
                double x[2];
                toPhys(B.E, a.q.Px[2 * k], a.q.Px[2 * k + 1], x);
                const double wk = a.q.dx[k] * B.absDet;
                double inN[3 * OO];
#pragma unroll
                for (int o = 0; o < 3 * OO; o++) inN[o] = 0;
#pragma unroll
                for (int l = 0; l < NPY; l++) {
                    T K[OO];
                    kernel_eval<KER, OO, T>((T)(x[0] - yi[2 * l]), (T)(x[1] - yi[2 * l + 1]), a.kp, K);
                    const T psi[3] = {(T)a.q.psiy[3 * l], (T)a.q.psiy[3 * l + 1], (T)a.q.psiy[3 * l + 2]};
                    inner_point<OD, T, false, true, false, KS>(K, (T)(a.q.dy[l] * A.absDet), psi, nullptr, inN, nullptr);
                    {
#pragma clang fp contract(fast)
#pragma unroll
                        for (int o = 0; o < OO; o++)
                            if (!(KS && o == 2)) M[l * OO + o] += wk * (double)K[o];
                    }
                }
                {
#pragma clang fp contract(fast)
                    const double ak = B.absDet * a.q.dx[k];
                    double wi[3];
#pragma unroll
                    for (int i = 0; i < 3; i++) wi[i] = ak * a.q.psix[3 * k + i];
#pragma unroll
                    for (int r = 0; r < N; r++)
#pragma unroll
                        for (int c = 0; c < N; c++)
                            if (!(KS && r % OD == 1 && c % OD == 0))
                                tN[N * r + c] += wi[r / OD] * inN[(r % OD) * OD + OO * (c / OD) + (c % OD)];
                }
            }
#pragma unroll
            for (int i = 0; i < 3; i++) {  // -termNonlocPrime into (T_j, U_i)
                const int s = cm.find(node_of(a.tri, U, i, a.is_DG));
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    if (!labA[j]) continue;
                    if (s < 0) { atomicOr(&a.counters[CNT_ERROR], ERR_HASH); continue; }
#pragma unroll
                    for (int ao = 0; ao < OD; ao++)
#pragma unroll
                        for (int bo = 0; bo < OD; bo++)
                            if (!(KS3 && bo == 1 && ao == 0))
                                atomicAdd(&acc[(s * 3 + j) * NCA + aix(bo, ao)], -tN[tni(i, ao, j, bo)]);
                }
            }
        }
#pragma unroll
        for (int t = 0; t < NPY * OO; t++) {
            if (KS && t % OO == 2) continue;
            double v = M[t];
            for (int m = 32; m > 0; m >>= 1) v += shfl_xor_d(v, m);
            M[t] = v;
        }
        if (lane == 0) {
            // termLocalPrime[a * 3 OO + (o / OD) * 3 OD + b * OD + o % OD], as add_lp lays it out
#pragma unroll
            for (int ia = 0; ia < 3; ia++)
#pragma unroll
                for (int ib = 0; ib < 3; ib++)
#pragma unroll
                    for (int o = 0; o < OO; o++) {
                        double v = 0;
#pragma unroll
                        for (int l = 0; l < NPY; l++)
                            v += a.q.psiy[3 * l + ia] * a.q.psiy[3 * l + ib] * (a.q.dy[l] * A.absDet) * M[l * OO + so(o)];
                        red[w][NB + ia * OO * 3 + (o / OD) * OD * 3 + ib * OD + o % OD] = v;
                    }
        }
    } else {
    double tL[NB], tLP[NB];
#pragma unroll
    for (int i = 0; i < NB; i++) { tL[i] = 0; tLP[i] = 0; }
    // constant scalar kernel: the blocks of an interior pair are c |aT| |bT| times rule sums
    // (termNonloc_ab = p1_a q1_b, termLocal_ab = q0 p2_ab, termLocalPrime_ab = pdx q2_ab), so
    // per pair only |bT| is needed; termLocal and termLocalPrime take the sum of |bT|
    constexpr bool FACT = (KER == KER_CONSTANT && OD == 1);
    double sumB = 0, sumU = 0;
    // T as outer element: front segment, interior class
    const int nF = a.cntF[T0];
    if constexpr (FACT) {
        // Pairs in groups of IU per thread: the group's list entries, then its elements' |det| and
        // nodes, are loaded before any LDS update, so the dependent loads of a group overlap
        // (h = 1/256, delta = 0.05: 3.23 ms with 1 pair, 3.11 with 2, 3.22 with 3, 3.15 with 4,
        // 3.23 with 8, the interior phase with k_merge).
        // Role 0, T as outer element (front segment of its list): column U_j collects w q1_j,
        // row T_i takes p1_i of it below.  Role 1, T as inner element (front entries of its
        // transposed row): -termNonlocPrime into (T_j, U_i), column U_i collects w p1_i.
        constexpr int IU = 2;
        const long long tb = a.tptr[T0];
        const int nT = a.tF[T0];
#pragma unroll
        for (int role = 0; role < 2; role++) {
            const int n = role == 0 ? nF : nT;
            for (int p0 = tid; p0 < n; p0 += IU * 256) {
                int Us[IU];
#pragma unroll
                for (int u = 0; u < IU; u++) {
                    const int p = p0 + u * 256;
                    Us[u] = -1;
                    if (p < n) {
                        const uint32_t e = role == 0 ? a.nbr[(size_t)T0 * a.cap + p] : (uint32_t)a.tadj[tb + p];
                        if (!(e >> 31) && ((e >> CLS_SHIFT) & 3u) == CLS_INTERIOR) Us[u] = (int)(e & IDX_MASK);
                    }
                }
                double dU[IU];
                int nd[IU][3];
#pragma unroll
                for (int u = 0; u < IU; u++) {
                    const int U = Us[u] < 0 ? T0 : Us[u];
                    dU[u] = a.elem[U].absDet;
#pragma unroll
                    for (int j = 0; j < 3; j++) nd[u][j] = node_of(a.tri, U, j, a.is_DG);
                }
#pragma unroll
                for (int u = 0; u < IU; u++) {
                    if (Us[u] < 0) continue;
                    const double w = role == 0 ? a.kp.c * A.absDet * dU[u] : a.kp.c * dU[u] * A.absDet;
                    if (role == 0) sumB += dU[u];
                    else sumU += dU[u];
#pragma unroll
                    for (int j = 0; j < 3; j++) {
                        const int s = cm.find(nd[u][j]);
                        if (s < 0) { atomicOr(&a.counters[CNT_ERROR], ERR_HASH); continue; }
                        atomicAdd(&acc[s * 3 + role], w * (role == 0 ? a.q1[j] : a.p1[j]));
                    }
                }
            }
        }
        __syncthreads();
        // rows of T: acc[s][i] = -(p1_i Wout_s + q1_i Win_s) for assembled rows
        for (int sl = tid; sl < a.HN; sl += 256) {
            const double wo = acc[sl * 3 + 0], wi = acc[sl * 3 + 1];
            for (int i = 0; i < 3; i++) acc[sl * 3 + i] = labA[i] ? -(a.p1[i] * wo + a.q1[i] * wi) : 0.0;
        }
    }
    for (int p = tid; p < (FACT ? 0 : nF); p += 256) {
        const uint32_t e = a.nbr[(size_t)T0 * a.cap + p];
        if (((e >> CLS_SHIFT) & 3u) != CLS_INTERIOR) continue;
        const int U = (int)(e & IDX_MASK);
        const ElemRec B = a.elem[U];
        PairBlocks<OD> pb;
        pair_interior<KER, OD, T, true, false>(a, A, B, tL, pb);
        for (int j = 0; j < 3; j++) {
            const int s = cm.find(node_of(a.tri, U, j, a.is_DG));
            for (int i = 0; i < 3; i++) {
                if (!labA[i]) continue;
                if (s < 0) { atomicOr(&a.counters[CNT_ERROR], ERR_HASH); continue; }
#pragma unroll
                for (int ao = 0; ao < OD; ao++)
#pragma unroll
                    for (int bo = 0; bo < OD; bo++)
                        atomicAdd(&acc[(s * 3 + i) * OO + ao * OD + bo], -pb.tN[N * (i * OD + ao) + j * OD + bo]);
            }
        }
    }
    // T as inner element: pairs (U, T) with T in the interior front segment of U
    const long long tb = a.tptr[T0], te = FACT ? tb : tb + a.tF[T0];
    for (long long q = tb + tid; q < te; q += 256) {
        const uint32_t e = (uint32_t)a.tadj[q];
        if ((e >> 31) || ((e >> CLS_SHIFT) & 3u) != CLS_INTERIOR) continue;
        const int U = (int)(e & IDX_MASK);
        const ElemRec B = a.elem[U];
        PairBlocks<OD> pb;
        pair_interior<KER, OD, T, false, true>(a, B, A, nullptr, pb);
#pragma unroll
        for (int t = 0; t < NB; t++) tLP[t] += pb.tLP[t];
        for (int i = 0; i < 3; i++) {  // -termNonlocPrime into (T_j, U_i)
            const int s = cm.find(node_of(a.tri, U, i, a.is_DG));
            for (int j = 0; j < 3; j++) {
                if (!labA[j]) continue;
                if (s < 0) { atomicOr(&a.counters[CNT_ERROR], ERR_HASH); continue; }
#pragma unroll
                for (int ao = 0; ao < OD; ao++)
#pragma unroll
                    for (int bo = 0; bo < OD; bo++)
                        atomicAdd(&acc[(s * 3 + j) * OO + bo * OD + ao], -pb.tN[N * (i * OD + ao) + j * OD + bo]);
            }
        }
    }
    if constexpr (FACT) {
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                tL[3 * i + j] = a.kp.c * A.absDet * sumB * a.q0 * a.p2[sym6(i, j)];
                tLP[3 * i + j] = a.kp.c * sumU * A.absDet * a.pdx * a.q2[sym6(i, j)];
            }
    }
    // block reduction of termLocal and termLocalPrime
#pragma unroll
    for (int i = 0; i < NB; i++) {
        double v = tL[i], u = tLP[i];
        for (int m = 32; m > 0; m >>= 1) { v += shfl_xor_d(v, m); u += shfl_xor_d(u, m); }
        if (lane == 0) { red[w][i] = v; red[w][NB + i] = u; }
    }
    }
    __syncthreads();
    for (int t = tid; t < 2 * NB; t += 256) {
        const double v = red[0][t] + red[1][t] + red[2][t] + red[3][t];
        const int e = t % NB, ra = e / N, cb = e % N;
        const int i = ra / OD, ao = ra % OD, j = cb / OD, bo = cb % OD;
        if (t < NB) {  // +termLocal into (T_i, T_j)
            if (!labA[i]) continue;
            const int s = cm.find(rA[j]);
            if (s >= 0 && !(KS3 && ao == 1 && bo == 0)) atomicAdd(&acc[(s * 3 + i) * NCA + aix(ao, bo)], v);
        } else {       // +termLocalPrime[N a + b] into (T_b, T_a): row T_j comp bo, col T_i comp ao
            if (!labA[j]) continue;
            const int s = cm.find(rA[i]);
            if (s >= 0 && !(KS3 && bo == 1 && ao == 0)) atomicAdd(&acc[(s * 3 + j) * NCA + aix(bo, ao)], v);
        }
    }
    __syncthreads();
    // partial rows of T's nodes, in column order
    for (int j = 0; j < 3; j++) {
        if (!labA[j]) continue;
        const long long r0 = rp[j];
        const int len = rlen[j];
        double* dst = a.P + a.poff[3 * T0 + j] * OO;
        for (int i = tid; i < len; i += 256) {
            const int s = cm.find(a.col[r0 + i]);
#pragma unroll
            for (int o = 0; o < OO; o++)
                dst[(long long)i * OO + o] = (s >= 0) ? acc[(s * 3 + j) * NCA + aix(o / OD, o % OD)] : 0.0;
        }
    }
}

// Partial-row offsets: plen[3T+i] = row length of node(T,i) when T takes part in assembly
// and the node's row is assembled, else 0.
__global__ void k_plen(int nE, const ElemRec* elem, const int* tri, const int* vlabel, int is_DG,
                       const long long* rowptr, int* plen) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nE) return;
    const bool act = elem[e].label != 0;
    for (int j = 0; j < 3; j++) {
        const int r = node_of(tri, e, j, is_DG);
        plen[3 * e + j] = (act && row_labelled(vlabel, r, is_DG)) ? (int)(rowptr[r + 1] - rowptr[r]) : 0;
    }
}

// Node row r = sum of the partial rows of the elements holding node r.  The node's incidence
// row is taken in chunks of MERGE_CHUNK elements (one chunk on regular meshes; a vertex shared
// by more elements, such as an artificial vertex joining label-0 elements, takes several, the
// later ones adding to the values the first one stored).
constexpr int MERGE_CHUNK = 128;
__global__ __launch_bounds__(256) void k_merge(int nNodes, int oo, int is_DG, const ElemRec* elem, const int* tri,
                                               const int* vlabel, const long long* vptr, const int* vel,
                                               const long long* rowptr, const long long* poff, const double* P,
                                               double* val) {
    const int r = blockIdx.x;
    if (!row_labelled(vlabel, r, is_DG)) return;
    const long long r0 = rowptr[r];
    const long long len = (rowptr[r + 1] - r0) * oo;
    __shared__ long long base[MERGE_CHUNK];
    __shared__ int nb;
    const long long qs = is_DG ? 0 : vptr[r], qe = is_DG ? 1 : vptr[r + 1];
    for (long long c0 = qs; c0 < qe; c0 += MERGE_CHUNK) {
        if (threadIdx.x < 64) {
            // the chunk's elements with a nonzero label, in incidence order: one lane per
            // element (their loads overlap), positions from ballot ranks
            const int lane = threadIdx.x;
            int n = 0;
            if (is_DG) {
                const int T = r / 3;
                if (elem[T].label != 0) {
                    if (lane == 0) base[0] = poff[r] * oo;
                    n = 1;
                }
            } else {
                const long long q1 = min(qe, c0 + MERGE_CHUNK);
                for (long long qb = c0; qb < q1; qb += 64) {
                    const long long q = qb + lane;
                    bool ok = false;
                    long long b = 0;
                    if (q < q1) {
                        const int T = vel[q];
                        if (elem[T].label != 0) {
                            const int li = tri[3 * T] == r ? 0 : (tri[3 * T + 1] == r ? 1 : 2);
                            b = poff[3 * T + li] * oo;
                            ok = true;
                        }
                    }
                    const unsigned long long m = __ballot(ok);
                    const int rk = n + __popcll(m & ((1ull << lane) - 1ull));
                    if (ok) base[rk] = b;
                    n += __popcll(m);
                }
            }
            if (lane == 0) nb = n;
        }
        __syncthreads();
        const bool first = c0 == qs;
        // the partial rows' loads issued eight at a time, summed in row order (h = 1/256:
        // interior phase with this merge 3.12 -> 3.02 ms against one load per row at a time)
        const int n = nb;
        for (long long p = threadIdx.x; p < len; p += 256) {
            double v = 0.0;
            for (int t0 = 0; t0 < n; t0 += 8) {
                double u[8];
#pragma unroll
                for (int t = 0; t < 8; t++) u[t] = t0 + t < n ? P[base[t0 + t] + p] : 0.0;
#pragma unroll
                for (int t = 0; t < 8; t++)
                    if (t0 + t < n) v += u[t];
            }
            if (first) val[r0 * oo + p] = v;
            else val[r0 * oo + p] += v;
        }
        __syncthreads();
    }
}

// ======================================================================= cut pairs, two phases
// One block of 256 threads per outer element aT, cut pairs in chunks of ITEMS / nPx pairs
// (ITEMS = 64, 128 or 256 items per chunk; the plan can time all three and keep the fastest).
//   A  one thread per (pair, outer point k): streamed retriangulation; the polygon points go
//      to an LDS list (at most 9 points, as in the reference's R[18]).
//   B  one thread per fan triangle (R0, R_j, R_j+1) of any item of the chunk: the inner
//      rule on the triangle; innerLocal / innerNonloc of the item and termLocalPrime of the
//      pair accumulate in LDS (FP64 atomics).  Every thread has the same work (nPy points),
//      whatever the polygon sizes.
//   C  one thread per item: outer accumulation (src/integration.cpp:508-537) into termLocal
//      (registers) and termNonloc of the pair (LDS).
//   D  scatter of the chunk's pairs, one thread per (pair, i, j).
constexpr int CUT_ITEMS = 128;    // default items per chunk
constexpr int CUT_MAXP = 9;
constexpr int CUT_MAXPAIRS = 32;  // pairs per chunk; needs nPx >= ITEMS / CUT_MAXPAIRS
// Dynamic LDS of k_cut: polygons (shared with the pair sums), per-item sums, the outer rule's
// tables, the chunk's LU factors, and the column map (int keys, 16-bit row positions).
static size_t cut_lds_bytes(int nPx, int oo, int HN, int items = CUT_ITEMS) {
    const int cp = std::min(items / std::max(nPx, 1), CUT_MAXPAIRS);
    return 16 * (size_t)items * CUT_MAXP + 8 * ((size_t)items * 10 * oo + (size_t)nPx * (oo + 6)) +
           sizeof(PivotLU) * (size_t)cp + 10 * (size_t)HN;
}
// Whether the outer rule fits a chunk of `items` items (at least one pair, at most CUT_MAXPAIRS).
static bool cut_items_fit(int nPx, int items) { return nPx <= items && nPx * CUT_MAXPAIRS >= items; }

// Surrogate item (outer point x, inner element E): ZERO when B(x, delta) misses E, FULL when
// it contains E, PARTIAL otherwise; OUT when the item lies outside the network's domain
// (x inside E, or E's size relative to delta outside [smin, smax]).  For PARTIAL items f
// receives the 12 network inputs: E's vertices relative to x divided by delta, rotated so
// that E's centroid lies on the +x axis (6), (|u_i|^2 - 1) / s (3) and (dist(0, edge) - 1) / s
// for the edges 01, 12, 20 (3), with s = sqrt|det| in delta units.
enum SurItem : int { SUR_ZERO = 0, SUR_FULL = 1, SUR_PARTIAL = 2, SUR_OUT = 3 };

NLFEM_DEV int sur_item(double x0, double x1, const double* E, double sqdelta, double idelta, double smin,
                       double smax, float* f, double* rot = nullptr) {
    double u[6], r2[3];
    int nin = 0;
    for (int i = 0; i < 3; i++) {
        const double d0 = E[2 * i] - x0, d1 = E[2 * i + 1] - x1;
        r2[i] = d0 * d0 + d1 * d1;
        nin += r2[i] <= sqdelta;
        u[2 * i] = d0 * idelta;
        u[2 * i + 1] = d1 * idelta;
    }
    if (nin == 3) return SUR_FULL;
    double ed[3];
    double dmin = 1e300;
    for (int e = 0; e < 3; e++) {
        const int i = e, j = (e + 1) % 3;
        const double P0 = u[2 * i], P1 = u[2 * i + 1], D0 = u[2 * j] - P0, D1 = u[2 * j + 1] - P1;
        const double dd = D0 * D0 + D1 * D1;
        double t = dd > 0 ? -(P0 * D0 + P1 * D1) / dd : 0.0;
        t = t < 0 ? 0 : (t > 1 ? 1 : t);
        const double C0 = P0 + t * D0, C1 = P1 + t * D1;
        ed[e] = sqrt(C0 * C0 + C1 * C1);
        dmin = ed[e] < dmin ? ed[e] : dmin;
    }
    // is x inside E?  (signs of the three edge cross products relative to the origin)
    const double c0 = u[0] * u[3] - u[1] * u[2], c1 = u[2] * u[5] - u[3] * u[4], c2 = u[4] * u[1] - u[5] * u[0];
    const bool inside = (c0 >= 0 && c1 >= 0 && c2 >= 0) || (c0 <= 0 && c1 <= 0 && c2 <= 0);
    if (inside) return SUR_OUT;
    if (nin == 0 && dmin > 1.0) return SUR_ZERO;
    const double det = fabs((u[2] - u[0]) * (u[5] - u[1]) - (u[4] - u[0]) * (u[3] - u[1]));
    const double sz = sqrt(det);
    if (!(sz >= smin && sz <= smax)) return SUR_OUT;
    const double g0 = (u[0] + u[2] + u[4]) / 3.0, g1 = (u[1] + u[3] + u[5]) / 3.0;
    const double gn = sqrt(g0 * g0 + g1 * g1);
    const double cs = g0 / gn, sn = g1 / gn, is = 1.0 / sz;
    if (rot) { rot[0] = cs; rot[1] = sn; }
    for (int i = 0; i < 3; i++) {
        f[2 * i] = (float)(cs * u[2 * i] + sn * u[2 * i + 1]);
        f[2 * i + 1] = (float)(-sn * u[2 * i] + cs * u[2 * i + 1]);
        f[6 + i] = (float)((u[2 * i] * u[2 * i] + u[2 * i + 1] * u[2 * i + 1] - 1.0) * is);
        f[9 + i] = (float)((ed[i] - 1.0) * is);
    }
    return SUR_PARTIAL;
}

NLFEM_DEV unsigned sur_hash(unsigned aT, unsigned b, unsigned k, unsigned salt) {
    unsigned h = aT * 0x9E3779B1u ^ (b + 0x7F4A7C15u) * 0x85EBCA77u ^ (k + 1u) * 0xC2B2AE3Du ^ salt;
    h ^= h >> 16; h *= 0x7FEB352Du; h ^= h >> 15; h *= 0x846CA68Bu; h ^= h >> 16;
    return h;
}

// A symmetric 2 x 2 tensor (xx, xy, yy) into the surrogate's frame, Q T Q^T with
// Q = [[cs, sn], [-sn, cs]] (the rotation that puts bT's centroid on +x), and back (Q^T T Q).
NLFEM_DEV void field_to_frame(double xx, double xy, double yy, double cs, double sn, double* t) {
    t[0] = cs * cs * xx + 2.0 * cs * sn * xy + sn * sn * yy;
    t[1] = -cs * sn * xx + (cs * cs - sn * sn) * xy + cs * sn * yy;
    t[2] = sn * sn * xx - 2.0 * cs * sn * xy + cs * cs * yy;
}
NLFEM_DEV void field_from_frame(double xx, double xy, double yy, double cs, double sn, double* t) {
    field_to_frame(xx, xy, yy, cs, -sn, t);
}
// Scale of the field surrogate's outputs: the inner rule's sums over the whole of bT of
// psi_a psi_b times the trace of the field kernel (c / |x - y|), in units of c |det bT| / delta:
// F_q = sum_i dy_i psi_a psi_b(P_i) delta / |x - y_i| for the six index pairs q = (a <= b).
// Every component of a partial item's tensor lies in [-F_q, F_q] (|z_i z_j| <= |z|^2, and the
// integrands of the trace are positive), so the networks learn ratios in [-1, 1].
NLFEM_DEV void field_scale(double x0, double x1, const double* E, const QuadArgs& q, double delta, double* F) {
#pragma unroll
    for (int i = 0; i < 6; i++) F[i] = 0.0;
    for (int i = 0; i < q.nPy; i++) {
        double y[2];
        toPhys(E, q.Py[2 * i], q.Py[2 * i + 1], y);
        const double d0 = x0 - y[0], d1 = x1 - y[1];
        const double w = q.dy[i] * delta / sqrt(d0 * d0 + d1 * d1);
        const double* ps = &q.psiy[3 * i];
        F[0] += w * ps[0] * ps[0];
        F[1] += w * ps[0] * ps[1];
        F[2] += w * ps[0] * ps[2];
        F[3] += w * ps[1] * ps[1];
        F[4] += w * ps[1] * ps[2];
        F[5] += w * ps[2] * ps[2];
    }
}

struct PointCollector {
    double2* pts;
    int nt;
    NLFEM_DEV void operator()(double r0, double r1, double s0, double s1, double t0, double t1) {
        if (nt == 0) {
            pts[0] = make_double2(r0, r1);
            pts[1] = make_double2(s0, s1);
            pts[2] = make_double2(t0, t1);
        } else if (nt + 2 < CUT_MAXP) {
            pts[nt + 2] = make_double2(t0, t1);
        }
        nt++;
    }
};

// Outer accumulation of a chunk of cut pairs from the per-item inner sums iacc (10*OO per
// item: innerLocal, innerNonloc, the six psi_a psi_b products), as outer_point and add_lp
// do it per pair (src/integration.cpp:508-537), with one thread per (pair, entry):
// termNonloc and termLocalPrime into pacc.  termLocal only goes to the outer element's own
// rows, so tls[k*OO + o] collects innerLocal over all pairs of the block per outer point.
// Inner sums of one outer point for the constant kernel, accumulated from the fan triangles
// (R0, R_j, R_j+1) as the retriangulation emits them.  The basis of bT is affine, psi_b(y) =
// psi_b(R0) + grad psi_b . (y - R0), so on a fan triangle the inner rule's sums of 1, psi_b
// and psi_a psi_b over its pulled-back points are linear in the rule's moments (q0, qm) of
// the corner offsets d = R_j - R0.  Per fan triangle only the geometric moments
//   m0 = sum |det| q0-part,  m1_c = sum |det| (d1_c mx + d2_c my),
//   m2_cd = sum |det| (d1_c d1_d mxx + (d1_c d2_d + d2_c d1_d) mxy + d2_c d2_d myy)
// are accumulated (offsets from the fan apex, physical coordinates); finish() applies the
// basis once per item: psi at the apex from the reference's pivoted toRef, the gradients
// from the inverse Jacobian of bT.  The results agree with the point-by-point sums of
// integrate_retriangulate (src/integration.cpp:418-506) to rounding.
struct ConstantItem {
    const NumArgs& a;
    const PivotLU& lu;
    const double* E;  // vertices of bT
    double inL = 0, inN[3] = {0, 0, 0}, sLP[6] = {0, 0, 0, 0, 0, 0};
    double m0 = 0, m1[2] = {0, 0}, m2[3] = {0, 0, 0};
    double f0 = 0, f1 = 0, e0 = 0, e1 = 0;  // apex R0, offset of the previous corner
    int nt = 0;
    NLFEM_DEV ConstantItem(const NumArgs& a_, const PivotLU& lu_, const double* E_) : a(a_), lu(lu_), E(E_) {}
    NLFEM_DEV void operator()(double r0, double r1, double s0, double s1, double t0, double t1) {
#pragma clang fp contract(fast)
        if (nt == 0) {
            f0 = r0; f1 = r1;
            e0 = s0 - r0; e1 = s1 - r1;
        }
        nt++;
        const double d0 = t0 - f0, d1 = t1 - f1;
        const double D = absolute(e0 * d1 - d0 * e1);
        const double mx = a.qm[0], my = a.qm[1], mxx = a.qm[2], mxy = a.qm[3], myy = a.qm[4];
        const double u0 = e0 * mxx + d0 * mxy, u1 = e1 * mxx + d1 * mxy;
        const double v0 = e0 * mxy + d0 * myy, v1 = e1 * mxy + d1 * myy;
        m0 += D;
        m1[0] += D * (e0 * mx + d0 * my);
        m1[1] += D * (e1 * mx + d1 * my);
        m2[0] += D * (e0 * u0 + d0 * v0);
        m2[1] += D * (e1 * u0 + d1 * v0);
        m2[2] += D * (e1 * u1 + d1 * v1);
        e0 = d0; e1 = d1;
    }
    NLFEM_DEV void finish() {
#pragma clang fp contract(fast)
        if (nt == 0) return;
        double p0, p1;
        lu.solve(f0, f1, p0, p1);
        const double g[3] = {1 - p0 - p1, p0, p1};
        const InvMap im = invmap(E);
        // gradients: psi_1 = p0, psi_2 = p1, psi_0 = 1 - p0 - p1
        const double G[3][2] = {{-(im.i00 + im.i10), -(im.i01 + im.i11)}, {im.i00, im.i01}, {im.i10, im.i11}};
        const double c = a.kp.c, cm0 = c * (a.q0 * m0);
        double A[3];
        for (int b = 0; b < 3; b++) A[b] = G[b][0] * m1[0] + G[b][1] * m1[1];
        inL = cm0;
        for (int b = 0; b < 3; b++) inN[b] = g[b] * cm0 + c * A[b];
        int m = 0;
        for (int u = 0; u < 3; u++) {
            const double h0 = G[u][0] * m2[0] + G[u][1] * m2[1], h1 = G[u][0] * m2[1] + G[u][1] * m2[2];
            for (int v = u; v < 3; v++, m++)
                sLP[m] = g[u] * g[v] * cm0 + c * (g[u] * A[v] + g[v] * A[u] + h0 * G[v][0] + h1 * G[v][1]);
        }
    }
};

// wa[3 k + i] = |aT| dx_k psi_i(x_k) and wd[k] = dx_k |aT| are the block's outer weights in LDS.
// KSYM: the kernel matrix is symmetric (k01 = k10, the microelastic field kernel), so phase B
// accumulates component 1 only and component 2 is read from component 1.
template <int OD, bool KSYM = false>
NLFEM_DEV void cut_outer(const NumArgs& a, int np, const double* wa, const double* wd, const double* iacc,
                         double* pacc, double* tls) {
    constexpr int OO = OD * OD, NB = 9 * OO, N = 3 * OD;
    auto so = [](int o) { return (KSYM && o == 2) ? 1 : o; };
    const int nPx = a.q.nPx;
    const int nw = np * 2 * NB;
    for (int t = threadIdx.x; t < nw + nPx * OO; t += 256) {
        if (t < nw) {
            // one entry of one pair: termNonloc[N ra + cb] (outer_point) or termLocalPrime (add_lp)
            const int p = t / (2 * NB), e = t - p * 2 * NB;
            const double* ia = iacc + (size_t)p * nPx * 10 * OO;
            double v = 0;
            if (e < NB) {
                const int ra = e / N, cb = e - ra * N;
                const int ai = ra / OD, src = OO + so((ra % OD) * OD + (cb % OD)) + OO * (cb / OD);
#pragma unroll 4
                for (int k = 0; k < nPx; k++) {
#pragma clang fp contract(fast)
                    v += wa[3 * k + ai] * ia[k * 10 * OO + src];
                }
            } else {
                const int f = e - NB;  // a * 3 OO + (o / OD) * 3 OD + b * OD + o % OD
                const int aa = f / (3 * OO), r1 = f - aa * 3 * OO;
                const int oq = r1 / (3 * OD), r2 = r1 - oq * 3 * OD;
                const int bb = r2 / OD, o = oq * OD + (r2 - bb * OD);
                const int src = 4 * OO + sym6(aa, bb) * OO + so(o);
#pragma unroll 4
                for (int k = 0; k < nPx; k++) {
#pragma clang fp contract(fast)
                    v += wd[k] * ia[k * 10 * OO + src];
                }
            }
            pacc[t] = v;
        } else {
            // innerLocal summed over the chunk's pairs per (outer point, component)
            const int k = (t - nw) / OO, o = (t - nw) % OO;
            double v = 0;
            for (int p = 0; p < np; p++) v += iacc[((size_t)p * nPx + k) * 10 * OO + so(o)];
            tls[k * OO + o] += v;
        }
    }
}

// MODE 0 assembles; MODE 1 writes training samples of the surrogate (partial items, exact
// FP64 inner sums as targets) and assembles nothing.
// The FP64 outdim-2 instances fit 168 registers without spilling (3 wavefronts per SIMD where
// the chunk's LDS allows it: 64 items per chunk) instead of taking 192.
template <int KER, int OD, typename T, int MODE = 0, int ITEMS = CUT_ITEMS>
__global__ __launch_bounds__(256, (OD == 2 && sizeof(T) == 8) ? 3 : 1) void k_cut(NumArgs a) {
    constexpr int OO = OD * OD, NB = 9 * OO, N = 3 * OD;
    static_assert(ITEMS == 64 || ITEMS == 128 || ITEMS == 256, "items per chunk: 64, 128 or 256");
    constexpr int IW = ITEMS / 64;  // wavefronts that hold an item each
    // dynamic LDS, cut_lds_bytes() on the host: the polygons of phases A-B share their space
    // with the pair sums of phases C-D (CP * 2 NB <= 2 * ITEMS * CUT_MAXP doubles)
    const int nPx_ = a.q.nPx;
    const int CP = min(ITEMS / nPx_, CUT_MAXPAIRS);                 // pairs per chunk
    extern __shared__ double ldsc[];
    double2* pts = (double2*)ldsc;                                  // [ITEMS][CUT_MAXP]
    double* pacc = ldsc;                                            // [CP][2*NB]: termNonloc, termLocalPrime
    double* iacc = (double*)(pts + ITEMS * CUT_MAXP);              // [ITEMS][10*OO]: innerLocal, innerNonloc, sLP
    double* tls = iacc + ITEMS * 10 * OO;                           // [nPx][OO]
    double* xk = tls + nPx_ * OO;                                   // [nPx][2] outer points of aT
    double* wa = xk + 2 * nPx_;                                     // [nPx][3] |aT| dx_k psi_i(x_k)
    double* wd = wa + 3 * nPx_;                                     // [nPx] dx_k |aT|
    PivotLU* lus = (PivotLU*)(wd + nPx_);                           // [CP]
    int* hk = (int*)(lus + CP);                                     // [HN]
    unsigned short* hv = (unsigned short*)(hk + a.HN);              // [HN][3] positions in aT's rows
    __shared__ int ntri[ITEMS];
    __shared__ int tpre[ITEMS + 1];
    __shared__ int pb_[CUT_MAXPAIRS];                               // inner element of each chunk pair
    __shared__ unsigned char tmap[ITEMS * (CUT_MAXP - 2)];          // fan triangle -> item
    __shared__ unsigned subtri;
    __shared__ int wsum[4];
    const int aT = blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const ElemRec A = a.elem[aT];
    if (A.label == 0) return;
    const int nB = a.sur.list_in ? a.sur.rcnt[aT] : a.cntB[aT];
    if (nB == 0) return;
    const int nPx = a.q.nPx;
    int rA[3];
    bool labA[3];
    for (int j = 0; j < 3; j++) {
        rA[j] = node_of(a.tri, aT, j, a.is_DG);
        labA[j] = row_labelled(a.vlabel, rA[j], a.is_DG);
    }
    for (int i = tid; i < a.HN; i += 256) hk[i] = -1;
    for (int i = tid; i < nPx * OO; i += 256) tls[i] = 0.0;
    if (tid < nPx) {
        toPhys(A.E, a.q.Px[2 * tid], a.q.Px[2 * tid + 1], &xk[2 * tid]);
        const double c1 = A.absDet * a.q.dx[tid];
        for (int i = 0; i < 3; i++) wa[3 * tid + i] = c1 * a.q.psix[3 * tid + i];
        wd[tid] = a.q.dx[tid] * A.absDet;
    }
    if (tid == 0) subtri = 0;
    __syncthreads();
    for (int j = 0; j < 3; j++) {
        const long long r0 = a.rowptr[rA[j]];
        const int len = (int)(a.rowptr[rA[j] + 1] - r0);
        for (int i = tid; i < len; i += 256) {
            int s = lds_slot(hk, a.HN, a.col[r0 + i]);
            if (s >= 0) hv[3 * s + j] = (unsigned short)i;
            else atomicOr(&a.counters[CNT_ERROR], 16ull);
        }
    }
    const double absA = A.absDet;
    unsigned mytri = 0;
    for (int c0 = 0; c0 < nB; c0 += CP) {
        const int np = min(CP, nB - c0);
        const int nit = np * nPx;
        __syncthreads();
        for (int i = tid; i < ITEMS * 10 * OO; i += 256) iacc[i] = 0.0;
        if (tid < np) {
            const int b = a.sur.list_in ? a.sur.rlist[a.sur.roff[aT] + c0 + tid]
                                        : (int)(a.nbr[(size_t)aT * a.cap + a.cap - 1 - (c0 + tid)] & IDX_MASK);
            pb_[tid] = b;
            lus[tid] = pivot_lu(a.elem[b].E);
        }
        __syncthreads();
        // A: polygons; an outer point whose ball stays farther than delta (1 + 1e-9) from all
        // three edges of bT has an empty polygon and skips the retriangulation.
        if (tid < nit) {
            const int p = tid / nPx, k = tid % nPx;
            const ElemRec B = a.elem[pb_[p]];
            const double x[2] = {xk[2 * k], xk[2 * k + 1]};
            int nt = 0;
            if (!tri_far(x[0], x[1], B.E, a.sq_far)) {
                if constexpr (KER == KER_CONSTANT && OD == 1) {  // integrated here: no fan triangles for phase B
                    ConstantItem ci(a, lus[p], B.E);
                    retriangulate<false>(x[0], x[1], B.E, a.kp.sqdelta, a.caps != 0, ci);
                    ci.finish();
                    mytri += ci.nt;
                    double* ia = iacc + tid * 10;
                    ia[0] = ci.inL;
                    for (int q = 0; q < 3; q++) ia[1 + q] = ci.inN[q];
                    for (int q = 0; q < 6; q++) ia[4 + q] = ci.sLP[q];
                } else {
                    PointCollector pc{pts + tid * CUT_MAXP, 0};
                    retriangulate<false>(x[0], x[1], B.E, a.kp.sqdelta, a.caps != 0, pc);
                    if (pc.nt > CUT_MAXP - 2) atomicOr(&a.counters[CNT_ERROR], 256ull);
                    nt = min(pc.nt, CUT_MAXP - 2);
                    mytri += nt;
                }
            }
            ntri[tid] = nt;
        } else if (tid < ITEMS) {
            ntri[tid] = 0;
        }
        __syncthreads();
        {  // inclusive scan of ntri over the first IW waves (one item per thread)
            int v = (tid < ITEMS) ? ntri[tid] : 0;
            for (int off = 1; off < 64; off <<= 1) {
                const int y = __shfl_up(v, off, 64);
                if (lane >= off) v += y;
            }
            if (lane == 63 && w < IW) wsum[w] = v;
            __syncthreads();
            for (int q = 0; q < w && q < IW; q++) v += wsum[q];
            if (tid < ITEMS) {
                tpre[tid + 1] = v;
                for (int q = v - ntri[tid]; q < v; q++) tmap[q] = (unsigned char)tid;
            }
            if (tid == 0) tpre[0] = 0;
        }
        __syncthreads();
        // B: fan triangles
        const int ntot = tpre[nit];
        for (int t = tid; t < ntot; t += 256) {
            const int it = tmap[t], j = t - tpre[it];
            const int p = it / nPx, k = it % nPx;
            const double2 R0 = pts[it * CUT_MAXP], R1 = pts[it * CUT_MAXP + j + 1], R2 = pts[it * CUT_MAXP + j + 2];
            const double rT[6] = {R0.x, R0.y, R1.x, R1.y, R2.x, R2.y};
            const double rTdet = absDet(rT);
            const double x[2] = {xk[2 * k], xk[2 * k + 1]};
            const FanBasis fb(lus[p], rT);
            double inL[OO], inN[3 * OO], sLP[6 * OO];
#pragma unroll
            for (int o = 0; o < OO; o++) inL[o] = 0;
#pragma unroll
            for (int o = 0; o < 3 * OO; o++) inN[o] = 0;
#pragma unroll
            for (int o = 0; o < 6 * OO; o++) sLP[o] = 0;
            for (int i = 0; i < a.q.nPy; i++) {
                double y[2];
                toPhys(rT, a.q.Py[2 * i], a.q.Py[2 * i + 1], y);
                T K[OO];
                kernel_eval<KER, OO, T>((T)(x[0] - y[0]), (T)(x[1] - y[1]), a.kp, K);
                double p0, p1;
                fb.at(a.q.Py[2 * i], a.q.Py[2 * i + 1], p0, p1);
                T psi[3] = {(T)(1 - p0 - p1), (T)p0, (T)p1};
                inner_point<OD, T>(K, (T)(a.q.dy[i] * rTdet), psi, inL, inN, sLP);
            }
            double* ia = iacc + it * 10 * OO;  // the ~2.5 triangles of an item share these slots
            // a symmetric kernel matrix: component 2 equals component 1 and is not summed
            constexpr bool KSYM = KER == KER_MICRO_FIELD;
#pragma unroll
            for (int o = 0; o < OO; o++)
                if (!(KSYM && o == 2)) atomicAdd(&ia[o], inL[o]);
#pragma unroll
            for (int o = 0; o < 3 * OO; o++)
                if (!(KSYM && o % OO == 2)) atomicAdd(&ia[OO + o], inN[o]);
#pragma unroll
            for (int o = 0; o < 6 * OO; o++)
                if (!(KSYM && o % OO == 2)) atomicAdd(&ia[4 * OO + o], sLP[o]);
        }
        __syncthreads();
        if constexpr (MODE == 1) {  // training samples: ratios of the exact sums to the full-triangle sums
            if (tid < nit && OD == 1) {
                const int p = tid / nPx, k = tid % nPx;
                const int b = pb_[p];
                if (sur_hash((unsigned)aT, (unsigned)b, (unsigned)k, a.sur.salt) % a.sur.stride == 0) {
                    const ElemRec B = a.elem[b];
                    double x[2];
                    toPhys(A.E, a.q.Px[2 * k], a.q.Px[2 * k + 1], x);
                    float f[12];
                    if (sur_item(x[0], x[1], B.E, a.kp.sqdelta, 1.0 / sqrt(a.kp.sqdelta), a.sur.smin, a.sur.smax, f) ==
                        SUR_PARTIAL) {
                        const long long idx = (long long)atomicAdd(&a.counters[CNT_SAMPLES], 1ull);
                        if (idx < a.sur.smaxn) {
                            const double* ia = iacc + tid * 10;
                            const double cd = a.kp.c * B.absDet;
                            for (int i = 0; i < 12; i++) a.sur.sfeat[idx * 12 + i] = f[i];
                            for (int q = 0; q < 6; q++) a.sur.starg[idx * 9 + q] = (float)(ia[4 + q] / (cd * a.sur.q2[q]));
                            for (int bb = 0; bb < 3; bb++)
                                a.sur.starg[idx * 9 + 6 + bb] = (float)(ia[1 + bb] / (cd * a.sur.q1[bb]));
                        }
                    }
                }
            }
            if (tid < nit && OD == 2 && KER == KER_MICRO_FIELD) {
                // field kernel: the six symmetric 2 x 2 tensors sLP_ab = sum psi_a psi_b k(y, x) w in
                // the network's frame (bT's centroid on +x), as ratios to field_scale's F_q
                const int p = tid / nPx, k = tid % nPx;
                const int b = pb_[p];
                if (sur_hash((unsigned)aT, (unsigned)b, (unsigned)k, a.sur.salt) % a.sur.stride == 0) {
                    const ElemRec B = a.elem[b];
                    const double idelta = 1.0 / sqrt(a.kp.sqdelta);
                    float f[12];
                    double rot[2];
                    if (sur_item(xk[2 * k], xk[2 * k + 1], B.E, a.kp.sqdelta, idelta, a.sur.smin, a.sur.smax, f,
                                 rot) == SUR_PARTIAL) {
                        const long long idx = (long long)atomicAdd(&a.counters[CNT_SAMPLES], 1ull);
                        if (idx < a.sur.smaxn) {
                            const double* ia = iacc + tid * 10 * OO;
                            const double inorm = 1.0 / (a.kp.c * B.absDet * idelta);
                            double Fq[6];
                            field_scale(xk[2 * k], xk[2 * k + 1], B.E, a.q, 1.0 / idelta, Fq);
                            for (int i = 0; i < 12; i++) a.sur.sfeat[idx * 12 + i] = f[i];
                            for (int q = 0; q < 6; q++) {
                                double t[3];
                                field_to_frame(ia[4 * OO + q * OO + 0], ia[4 * OO + q * OO + 1],
                                               ia[4 * OO + q * OO + 3], rot[0], rot[1], t);
                                for (int o = 0; o < 3; o++)
                                    a.sur.starg[idx * 18 + 3 * q + o] = (float)(t[o] * inorm / Fq[q]);
                            }
                        }
                    }
                }
            }
            continue;
        }
        // C: outer accumulation (src/integration.cpp:508-537), one thread per (pair, entry) of
        // termNonloc, termLocalPrime and termLocal; termLocal is summed over the block in LDS
        cut_outer<OD, KER == KER_MICRO_FIELD>(a, np, wa, wd, iacc, pacc, tls);
        __syncthreads();
        // D: scatter (src/Cassemble.cpp:721-770)
        cut_scatter<OD>(a, np, pb_, pacc, rA, labA, hk, hv);
    }
    if constexpr (MODE == 1) return;
    atomicAdd(&subtri, mytri);
    __syncthreads();
    if (tid == 0) atomicAdd(&a.counters[CNT_SUBTRI], (unsigned long long)subtri);
    // termLocal[N ra + cb] = sum_k |aT| dx_k psi_a(x_k) psi_b(x_k) innerLocal_k, into (aAdx[i], aAdx[j])
    for (int t = tid; t < NB; t += 256) {
        const int ra = t / N, cb = t % N;
        const int i = ra / OD, ao = ra % OD, j = cb / OD, bo = cb % OD;
        if (!labA[i]) continue;
        double v = 0;
        for (int k = 0; k < nPx; k++)
            v += absA * a.q.dx[k] * a.q.psix[3 * k + i] * a.q.psix[3 * k + j] * tls[k * OO + OD * ao + bo];
        const long long p = a.posTT[9 * aT + 3 * i + j];
        atomicAdd(&a.val[p * OO + ao * OD + bo], v);
    }
}

// Cut pairs with one item (pair x outer point) per lane, for the cases k_cut8 does not take
// (outdim 2, DG, outer rules of 9..16 points): one 256-thread block per outer element, groups
// of nPx lanes (one group per cut pair, one lane per outer point: 4 pairs per wavefront with
// the 16-point rule), every wavefront walking its own pairs without block barriers.
//   - A lane retriangulates B(x_k, delta) n bT into its LDS point list (PointCollector, one
//     column of pts per lane), then integrates its fan triangles (R0, R_t+1, R_t+2) in one loop
//     to the largest triangle count of the wavefront: the inner rule (kernel, pulled-back
//     basis, the sums of inner_point) is issued once per triangle index, not once per place in
//     the retriangulation where a point can be emitted.
//   - The lane's shares of termNonloc (|aT| dx_k psi_i(x_k) innerNonloc_j) and termLocalPrime
//     (dx_k |aT| sLP) are summed per pair with LDS FP64 atomics into the group's slot (no
//     staging of per-item sums); innerLocal stays in the lane over all its pairs.
//   - The group's lanes then scatter the pair's entries as cut_scatter does
//     (src/Cassemble.cpp:721-770), and termLocal is formed once per block at the end.
// KSYM (the microelastic field kernel, k01 = k10): components 00, 01, 11 are summed.
template <int OD, int KSYM_>
struct CutwLayout {
    static constexpr int OO = OD * OD;
    static constexpr int NC = KSYM_ ? 3 : OO;   // summed components per 2 x 2 block
    static constexpr int PACC = 15 * NC;        // termNonloc 9 NC, termLocalPrime 6 NC
};
static size_t cutw_lds_bytes(int nPx, int od, bool ksym, int HN) {
    const int nc = ksym ? 3 : od * od;
    const int G = 64 / (nPx == 16 ? 16 : (nPx <= 8 ? 8 : std::max(nPx, 1)));
    return 16 * (size_t)CUT_MAXP * 256 + 8 * (size_t)4 * G * 15 * nc + 10 * (size_t)HN;
}
static bool cutw_fits(int nPx) { return nPx >= 1 && nPx <= 16; }

// Polygon points of one lane in column `tid` of pts[CUT_MAXP][256].
struct PointColumn {
    double2* col;  // &pts[tid]
    int nt;
    NLFEM_DEV void operator()(double r0, double r1, double s0, double s1, double t0, double t1) {
        if (nt == 0) {
            col[0] = make_double2(r0, r1);
            col[256] = make_double2(s0, s1);
            col[512] = make_double2(t0, t1);
        } else if (nt + 2 < CUT_MAXP) {
            col[(nt + 2) * 256] = make_double2(t0, t1);
        }
        nt++;
    }
};

// Sum of NV values over the GSZ lanes of a group (GSZ = 8 or 16, groups aligned to 8 or 16
// lanes), scattered: after log2(GSZ) xor steps, each of which keeps half of the remaining values
// (the lower half where the step's bit of k is 0), lane k holds the sums of values
// k NV / GSZ .. k NV / GSZ + NV / GSZ - 1 in out[].  share(e) gives this lane's value e.
template <int GSZ, int NV, class F>
NLFEM_DEV void group_reduce_scatter(F&& share, int k, double (&out)[NV / GSZ]) {
    static_assert((GSZ == 8 || GSZ == 16) && NV % GSZ == 0, "groups of 8 or 16 lanes");
    constexpr int H = NV / 2;
    double r[H];
    {
        const bool up = (k & (GSZ / 2)) != 0;
#pragma unroll
        for (int e = 0; e < H; e++) {
            const double lo = share(e), hi = share(H + e);
            r[e] = (up ? hi : lo) + __shfl_xor(up ? lo : hi, GSZ / 2, 64);
        }
    }
#pragma unroll
    for (int m = GSZ / 4, h = H / 2; m >= 1; m >>= 1, h >>= 1) {
        const bool up = (k & m) != 0;
#pragma unroll
        for (int e = 0; e < h; e++) {
            const double snd = up ? r[e] : r[h + e];
            r[e] = (up ? r[h + e] : r[e]) + __shfl_xor(snd, m, 64);
        }
    }
#pragma unroll
    for (int u = 0; u < NV / GSZ; u++) out[u] = r[u];
}

NLFEM_DEV void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// GSZ = 16 or 8: groups of GSZ lanes (an outer rule of 16 points, or of at most 8 with the lanes
// beyond it idle), aligned to DPP rows, whose pair sums are reduced and scattered across the
// group's lanes in registers (group_reduce_scatter); GSZ = 0: groups of nPx lanes whose pair
// sums go through LDS atomics.
template <int KER, int OD, typename T, int GSZ>
__global__ __launch_bounds__(256, (OD == 2 && sizeof(T) == 8) ? (GSZ == 16 ? 2 : 3) : 1) void k_cutw(NumArgs a) {
    constexpr int OO = OD * OD, N = 3 * OD, NB = 9 * OO;
    constexpr bool KSYM = KER == KER_MICRO_FIELD;
    using L = CutwLayout<OD, KSYM>;
    constexpr int NC = L::NC, PACC = L::PACC;
    // summed component c -> component of the OD x OD kernel block (KSYM: 00, 01, 11)
    auto oc = [](int c) { return (KSYM && c == 2) ? 3 : c; };
    // component o = ao OD + bo -> summed component
    auto co = [](int o) { return KSYM ? (o == 3 ? 2 : (o == 2 ? 1 : o)) : o; };
    extern __shared__ double ldsw[];
    double2* pts = (double2*)ldsw;                            // [CUT_MAXP][256]
    double* pacc = (double*)(pts + CUT_MAXP * 256);           // [4][G][PACC]
    const int nPx = GSZ == 16 ? 16 : a.q.nPx;
    const int GS = GSZ ? GSZ : nPx, G = 64 / GS;
    int* hk = (int*)(pacc + 4 * G * PACC);                    // [HN]
    unsigned short* hv = (unsigned short*)(hk + a.HN);        // [HN][3] positions in aT's rows
    __shared__ double red[4][NB];
    const int aT = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const ElemRec A = a.elem[aT];
    if (A.label == 0) return;
    const int nB = a.sur.list_in ? a.sur.rcnt[aT] : a.cntB[aT];
    if (nB == 0) return;
    const int gl = lane / GS, k = lane - gl * GS;
    const bool lane_on = gl < G;
    const bool kin = lane_on && k < nPx;  // the lane carries outer point k
    int rA[3];
    bool labA[3];
    for (int j = 0; j < 3; j++) {
        rA[j] = node_of(a.tri, aT, j, a.is_DG);
        labA[j] = row_labelled(a.vlabel, rA[j], a.is_DG);
    }
    for (int i = tid; i < a.HN; i += 256) hk[i] = -1;
    __syncthreads();
    for (int j = 0; j < 3; j++) {
        const long long r0 = a.rowptr[rA[j]];
        const int len = (int)(a.rowptr[rA[j] + 1] - r0);
        for (int i = tid; i < len; i += 256) {
            int s = lds_slot(hk, a.HN, a.col[r0 + i]);
            if (s >= 0) hv[3 * s + j] = (unsigned short)i;
            else atomicOr(&a.counters[CNT_ERROR], 16ull);
        }
    }
    __syncthreads();
    // this lane's outer point x_k and weights |aT| dx_k psi_i(x_k), dx_k |aT|
    double x[2] = {0, 0}, wa[3] = {0, 0, 0}, psx[3] = {0, 0, 0}, wd = 0;
    if (kin) {
        toPhys(A.E, a.q.Px[2 * k], a.q.Px[2 * k + 1], x);
        const double c1 = A.absDet * a.q.dx[k];
        for (int i = 0; i < 3; i++) {
            psx[i] = a.q.psix[3 * k + i];
            wa[i] = c1 * psx[i];
        }
        wd = a.q.dx[k] * A.absDet;
    }
    double inLs[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) inLs[c] = 0;
    unsigned mytri = 0;
    double* pa = pacc + (w * G + (lane_on ? gl : 0)) * PACC;
    double2* col = pts + tid;
    int b = 0;
    // Sum e of the current pair (layout: termNonloc (3 i + j) NC + c, then termLocalPrime
    // 9 NC + q NC + c) into the mirror sums M (src/Cassemble.cpp:721-770).  The reference adds
    // -termNonloc (i, j) to (aAdx[i], bAdx[j]) and to (bAdx[j], aAdx[i]), and termLocalPrime (i, j)
    // to (bAdx[j], bAdx[i]) for all i, j; for a symmetric kernel both are symmetric, so M takes
    // -termNonloc once at (aAdx[i], bAdx[j]) and termLocalPrime of i <= j once at (bAdx[j],
    // bAdx[i]) (halved on the diagonal), and k_mirror adds M + M^T into the rows that are
    // assembled: 15 NC atomics per pair instead of 27 OO.
    auto scatter_sum = [&](int e, double v) {
        if (e >= PACC || v == 0.0) return;
        if (e < 9 * NC) {
            const int ij = e / NC, c = e - ij * NC, i = ij / 3, j = ij - 3 * i;
            const int sl = lds_find(hk, a.HN, node_of(a.tri, b, j, a.is_DG));
            if (sl < 0) { atomicOr(&a.counters[CNT_ERROR], 8ull); return; }
            const long long pos = a.rowptr[rA[i]] + hv[3 * sl + i];  // entry (aAdx[i], bAdx[j])
            atomicAdd(&a.mir[pos * NC + c], -v);
        } else {
            const int q = (e - 9 * NC) / NC, c = (e - 9 * NC) - q * NC;
            const int i0 = q < 3 ? 0 : (q < 5 ? 1 : 2);
            const int j0 = q < 3 ? q : (q < 5 ? q - 2 : 2);
            const long long qp = a.posTT[9 * b + 3 * j0 + i0];  // entry (bAdx[j0], bAdx[i0])
            atomicAdd(&a.mir[qp * NC + c], i0 == j0 ? 0.5 * v : v);
        }
    };
    for (int c0 = 0; c0 < nB; c0 += 4 * G) {
        if (c0 + w * G >= nB) break;  // the whole wavefront has no pair left
        const int p = c0 + w * G + gl;
        const bool pon = lane_on && p < nB;
        b = !pon ? 0
                 : (a.sur.list_in ? a.sur.rlist[a.sur.roff[aT] + p]
                                  : (int)(a.nbr[(size_t)aT * a.cap + a.cap - 1 - p] & IDX_MASK));
        const ElemRec B = a.elem[b];
        int nt = 0;
        if (pon && kin && !tri_far(x[0], x[1], B.E, a.sq_far)) {
            PointColumn pc{col, 0};
            retriangulate<false>(x[0], x[1], B.E, a.kp.sqdelta, a.caps != 0, pc);
            if (pc.nt > CUT_MAXP - 2) atomicOr(&a.counters[CNT_ERROR], 256ull);
            nt = min(pc.nt, CUT_MAXP - 2);
        }
        mytri += nt;
        int mt = nt;
#pragma unroll
        for (int m = 32; m > 0; m >>= 1) mt = max(mt, __shfl_xor(mt, m, 64));
        double inL[OO], inN[3 * OO], sLP[6 * OO];
#pragma unroll
        for (int o = 0; o < OO; o++) inL[o] = 0;
#pragma unroll
        for (int o = 0; o < 3 * OO; o++) inN[o] = 0;
#pragma unroll
        for (int o = 0; o < 6 * OO; o++) sLP[o] = 0;
        if (mt > 0) {
            const PivotLU lu = pivot_lu(B.E);
            // the fan shares its apex R0, and triangle t's second corner is triangle t - 1's third:
            // one toRef per triangle (the same solves, so the same values as three per triangle)
            const double2 R0 = col[0];
            double2 R1 = col[256];
            double q00, q10, q01, q11;
            lu.solve(R0.x, R0.y, q00, q10);
            lu.solve(R1.x, R1.y, q01, q11);
            for (int t = 0; t < mt; t++) {
                if (t < nt) {
                    const double2 R2 = col[(t + 2) * 256];
                    double q02, q12;
                    lu.solve(R2.x, R2.y, q02, q12);
                    const double rT[6] = {R0.x, R0.y, R1.x, R1.y, R2.x, R2.y};
                    const double rTdet = absDet(rT);
                    const FanBasis fb(q00, q10, q01, q11, q02, q12);
                    R1 = R2;
                    q01 = q02;
                    q11 = q12;
                    for (int i = 0; i < a.q.nPy; i++) {
                        double y[2];
                        toPhys(rT, a.q.Py[2 * i], a.q.Py[2 * i + 1], y);
                        T K[OO];
                        kernel_eval<KER, OO, T>((T)(x[0] - y[0]), (T)(x[1] - y[1]), a.kp, K);
                        double p0, p1;
                        fb.at(a.q.Py[2 * i], a.q.Py[2 * i + 1], p0, p1);
                        T psi[3] = {(T)(1 - p0 - p1), (T)p0, (T)p1};
                        inner_point<OD, T, true, true, true, KSYM>(K, (T)(a.q.dy[i] * rTdet), psi, inL, inN, sLP);
                    }
                }
            }
        }
        if constexpr (GSZ != 0) {
            // the 15 NC shares of the pair (padded to a multiple of GSZ), summed over the
            // group's lanes by a reduce-scatter; lane k ends with sums k NV / GSZ .. and
            // scatters them itself
            constexpr int NV = (PACC + GSZ - 1) / GSZ * GSZ;
            auto share = [&](int e) -> double {
#pragma clang fp contract(fast)
                if (!pon || nt == 0) return 0.0;
                if (e < 9 * NC) {
                    const int ij = e / NC, c = e - ij * NC;
                    return wa[ij / 3] * inN[(ij % 3) * OO + oc(c)];
                }
                if (e < 15 * NC) {
                    const int q = (e - 9 * NC) / NC, c = (e - 9 * NC) - q * NC;
                    return wd * sLP[q * OO + oc(c)];
                }
                return 0.0;
            };
            double r[NV / GSZ];
            group_reduce_scatter<GSZ, NV>(share, k, r);
            if (pon && nt > 0) {
#pragma unroll
                for (int c = 0; c < NC; c++) inLs[c] += inL[oc(c)];
            }
            if (pon) {
#pragma unroll
                for (int u = 0; u < NV / GSZ; u++) scatter_sum(k * (NV / GSZ) + u, r[u]);
            }
        } else {
            // the group's pair sums: cleared, then every lane adds its shares
            if (pon)
                for (int e = k; e < PACC; e += GS) pa[e] = 0.0;
            wave_sync_lds();
            if (pon && nt > 0) {
#pragma clang fp contract(fast)
#pragma unroll
                for (int i = 0; i < 3; i++)
#pragma unroll
                    for (int j = 0; j < 3; j++)
#pragma unroll
                        for (int c = 0; c < NC; c++) atomicAdd(&pa[(3 * i + j) * NC + c], wa[i] * inN[j * OO + oc(c)]);
#pragma unroll
                for (int q = 0; q < 6; q++)
#pragma unroll
                    for (int c = 0; c < NC; c++) atomicAdd(&pa[9 * NC + q * NC + c], wd * sLP[q * OO + oc(c)]);
#pragma unroll
                for (int c = 0; c < NC; c++) inLs[c] += inL[oc(c)];
            }
            wave_sync_lds();
            // the group's lanes scatter the pair's sums
            if (pon)
                for (int e = k; e < PACC; e += GS) scatter_sum(e, pa[e]);
            wave_sync_lds();
        }
    }
    unsigned tri = mytri;
    for (int m = 32; m > 0; m >>= 1) tri += __shfl_xor(tri, m, 64);
    if (lane == 0 && tri) atomicAdd(&a.counters[CNT_SUBTRI], (unsigned long long)tri);
    // termLocal[N (i OD + ao) + j OD + bo] = sum_k |aT| dx_k psi_i psi_j (x_k) innerLocal_k[(ao, bo)]
    double tL[NB];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int o = 0; o < OO; o++)
                tL[N * (i * OD + o / OD) + j * OD + o % OD] = lane_on ? wa[i] * psx[j] * inLs[co(o)] : 0.0;
    cut_flush_local<OD>(a, aT, tL, labA, red);
}

// Mirror sums of k_cutw into the matrix: entry (r, c), component (x, y) gets M(r, c)[(x, y)] +
// M(c, r)[(y, x)] when row r is assembled (a labelled vertex, or DG).  One thread per entry, after
// the cut kernel; KSYM: M holds the components 00, 01 (= 10), 11.
template <int OD, bool KSYM>
__global__ __launch_bounds__(256) void k_mirror(long long nnz, const int* rowOf, const int* tpos, const int* vlabel,
                                                int is_DG, const double* mir, double* val) {
    constexpr int OO = OD * OD, NC = KSYM ? 3 : OO;
    const long long p = blockIdx.x * 256LL + threadIdx.x;
    if (p >= nnz || !row_labelled(vlabel, rowOf[p], is_DG)) return;
    const long long tp = tpos[p];
    auto ci = [](int x, int y) { return KSYM ? x + y : x * OD + y; };
#pragma unroll
    for (int o = 0; o < OO; o++) {
        const int x = o / OD, y = o % OD;
        const double v = mir[p * NC + ci(x, y)] + mir[tp * NC + ci(y, x)];
        if (v != 0.0) val[p * OO + o] += v;
    }
}

// Cut pairs, scalar kernels with at most 8 outer points: one 256-thread block per outer
// element; groups of 8 or nPx lanes, one group per cut pair and one lane per outer point, so
// a chunk of pairs needs no block barrier.  Each lane retriangulates B(x_k, delta) n bT and
// integrates the fan triangles as they are emitted (constant kernel: ConstantItem's
// geometric moments; otherwise the inner rule point by point, as the reference's
// integrate_retriangulate), forms its outer point's share of termNonloc and termLocalPrime
// (src/integration.cpp:508-537), and the group sums them in three shuffle steps.  Where the
// blocks go (src/Cassemble.cpp:721-770):
//   -termNonloc (i, j) to (aAdx[i], bAdx[j]) and (bAdx[j], aAdx[i]): summed per column node
//      bAdx[j] in LDS over all pairs of the block (a table of the partners' nodes), then added
//      to the matrix once per block while walking aT's three rows;
//   termLocalPrime (6 symmetric values) to (bAdx[j], bAdx[i]): staged per pair for
//      k_lp_gather (or added directly when no staging buffer is given);
//   termLocal: each lane keeps innerLocal of its outer point over all pairs, and the block
//      forms termLocal once at the end.
template <int KER, typename T>
__device__ __forceinline__ void cut8_body(const NumArgs& a) {
    extern __shared__ double ldsc8[];
    double* acc = ldsc8;                 // [HN][3] termNonloc (i, node) summed over pairs
    int* hk = (int*)(acc + 3 * a.HN);    // [HN] column map of aT's three rows
    __shared__ double sred[256];
    __shared__ int scan_tmp[4];
    const int aT = blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const ElemRec A = a.elem[aT];
    if (A.label == 0) return;
    const int nB = a.sur.list_in ? a.sur.rcnt[aT] : a.cntB[aT];
    if (nB == 0) return;
    const int nPx = a.q.nPx;
    // groups of GS lanes, one per cut pair, one lane per outer point.  The constant kernel's
    // items are cheap and its groups are 8 lanes wide (a power of two: the sums are
    // all-reduced with xor shuffles and the scatter is shared by the lanes); the other
    // kernels' items dominate, and groups of nPx lanes (9 per wavefront for the 7-point rule)
    // leave fewer lanes idle (tree sum into the group's first lane, which scatters).
    constexpr bool G8 = KER == KER_CONSTANT;
    const int GS = G8 ? 8 : nPx;
    const int G = 64 / GS;
    const int gl = lane / GS, k = lane - gl * GS;
    const bool lane_on = gl < G;
    const int g = w * G + gl;
    int rA[3];
    bool labA[3];
    for (int j = 0; j < 3; j++) {
        rA[j] = node_of(a.tri, aT, j, a.is_DG);
        labA[j] = row_labelled(a.vlabel, rA[j], a.is_DG);
    }
    for (int i = tid; i < a.HN; i += 256) { acc[3 * i] = 0; acc[3 * i + 1] = 0; acc[3 * i + 2] = 0; }
    long long rp[3];
    int rlen[3];
    for (int j = 0; j < 3; j++) {
        rp[j] = a.rowptr[rA[j]];
        rlen[j] = (int)(a.rowptr[rA[j] + 1] - rp[j]);
    }
    // every node of a cut partner lies in all three rows of aT
    const bool all3[3] = {true, true, true};
    const ColMap cm = colmap_build(hk, a.HN, a.col, rp, rlen, all3, scan_tmp, a.counters);
    const bool kin = lane_on && k < nPx;
    double x[2] = {0, 0};
    if (kin) toPhys(A.E, a.q.Px[2 * (kin ? k : 0)], a.q.Px[2 * (kin ? k : 0) + 1], x);
    const double ak = kin ? A.absDet * a.q.dx[k] : 0.0;   // |aT| dx_k
    const double wk = kin ? a.q.dx[k] * A.absDet : 0.0;   // dx_k |aT| (add_lp's order)
    double psx[3] = {0, 0, 0};
    if (kin) for (int i = 0; i < 3; i++) psx[i] = a.q.psix[3 * k + i];
    double sumL = 0;  // innerLocal of this lane's outer point over its pairs
    unsigned mytri = 0;
    double* const stage = a.sur.list_in ? nullptr : a.stage;
    for (int c0 = 0; c0 < nB; c0 += 4 * G) {
        const int p = c0 + g;
        if (!lane_on || p >= nB) break;  // whole groups leave together
        const int b = a.sur.list_in ? a.sur.rlist[a.sur.roff[aT] + p]
                                    : (int)(a.nbr[(size_t)aT * a.cap + a.cap - 1 - p] & IDX_MASK);
        const ElemRec B = a.elem[b];
        double inL = 0, inN[3] = {0, 0, 0}, sLP[6] = {0, 0, 0, 0, 0, 0};
        if (kin) {
            if (!tri_far(x[0], x[1], B.E, a.sq_far)) {
                const PivotLU lu = pivot_lu(B.E);
                if constexpr (KER == KER_CONSTANT) {
                    ConstantItem ci(a, lu, B.E);
                    retriangulate<false>(x[0], x[1], B.E, a.kp.sqdelta, a.caps != 0, ci);
                    ci.finish();
                    mytri += ci.nt;
                    inL = ci.inL;
                    for (int q = 0; q < 3; q++) inN[q] = ci.inN[q];
                    for (int q = 0; q < 6; q++) sLP[q] = ci.sLP[q];
                } else {
                    auto emit = [&](double r0, double r1, double s0, double s1, double t0, double t1) {
                        const double rT[6] = {r0, r1, s0, s1, t0, t1};
                        const double rTdet = absDet(rT);
                        mytri++;
                        const FanBasis fb(lu, rT);
                        for (int i = 0; i < a.q.nPy; i++) {
                            double y[2];
                            toPhys(rT, a.q.Py[2 * i], a.q.Py[2 * i + 1], y);
                            T K[1];
                            kernel_eval<KER, 1, T>((T)(x[0] - y[0]), (T)(x[1] - y[1]), a.kp, K);
                            double p0, p1;
                            fb.at(a.q.Py[2 * i], a.q.Py[2 * i + 1], p0, p1);
                            T psi[3] = {(T)(1 - p0 - p1), (T)p0, (T)p1};
                            inner_point<1, T>(K, (T)(a.q.dy[i] * rTdet), psi, &inL, inN, sLP);
                        }
                    };
                    retriangulate<false>(x[0], x[1], B.E, a.kp.sqdelta, a.caps != 0, emit);
                }
            }
        }
        sumL += inL;
        // this outer point's share of termNonloc (9) and termLocalPrime (6, symmetric), summed
        // over the group
        double tN[9], tLP[6];
        {
#pragma clang fp contract(fast)
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) tN[3 * i + j] = ak * psx[i] * inN[j];
            for (int m = 0; m < 6; m++) tLP[m] = wk * sLP[m];
        }
        if constexpr (G8) {
#pragma unroll
            for (int m = 1; m < 8; m <<= 1) {
#pragma unroll
                for (int e = 0; e < 9; e++) tN[e] += shfl_xor_d(tN[e], m);
#pragma unroll
                for (int e = 0; e < 6; e++) tLP[e] += shfl_xor_d(tLP[e], m);
            }
            // -termNonloc (i, j) goes to the entries (aAdx[i], bAdx[j]) and (bAdx[j], aAdx[i]):
            // lane j < 3 adds column j of this pair to node bAdx[j]'s accumulators, which reach
            // the matrix once per block below
            if (k < 3) {
                const int sl = cm.find(node_of(a.tri, b, k, a.is_DG));
                if (sl < 0) atomicOr(&a.counters[CNT_ERROR], 8ull);
                else
                    for (int i = 0; i < 3; i++)
                        atomicAdd(&acc[3 * sl + i], k == 0 ? tN[3 * i] : (k == 1 ? tN[3 * i + 1] : tN[3 * i + 2]));
            }
            // termLocalPrime: staged for k_lp_gather, or lane k adds entry (i, j) = k (and lane 0
            // also entry 8) to (bAdx[j], bAdx[i])
            if (stage) {
                if (k < 6) {
                    double v = tLP[0];
#pragma unroll
                    for (int m = 1; m < 6; m++) v = k == m ? tLP[m] : v;
                    stage[(a.coff[aT] + p) * 6 + k] = v;
                }
            } else {
                for (int e = k; e < 9; e += 8) {
                    const int i = e / 3, j = e % 3;
                    double v = tLP[0];
#pragma unroll
                    for (int m = 1; m < 6; m++) v = sym6(i, j) == m ? tLP[m] : v;
                    if (row_labelled(a.vlabel, node_of(a.tri, b, j, a.is_DG), a.is_DG))
                        atomicAdd(&a.val[a.posTT[9 * b + 3 * j + i]], v);
                }
            }
        } else {
            // tree sum over the group's lanes into its first lane (partners k + 1, k + 2, k + 4)
#pragma unroll
            for (int d = 1; d < 8; d <<= 1) {
                const bool take = (k & (2 * d - 1)) == 0 && k + d < GS;
                const int src = min(lane + d, 63);
#pragma unroll
                for (int e = 0; e < 9; e++) {
                    const double o = __shfl(tN[e], src, 64);
                    if (take) tN[e] += o;
                }
#pragma unroll
                for (int e = 0; e < 6; e++) {
                    const double o = __shfl(tLP[e], src, 64);
                    if (take) tLP[e] += o;
                }
            }
            if (k == 0) {
                for (int j = 0; j < 3; j++) {
                    const int sl = cm.find(node_of(a.tri, b, j, a.is_DG));
                    if (sl < 0) { atomicOr(&a.counters[CNT_ERROR], 8ull); continue; }
                    for (int i = 0; i < 3; i++) atomicAdd(&acc[3 * sl + i], tN[3 * i + j]);
                }
                if (stage) {
                    for (int m = 0; m < 6; m++) stage[(a.coff[aT] + p) * 6 + m] = tLP[m];
                } else {
                    for (int j = 0; j < 3; j++) {
                        if (!row_labelled(a.vlabel, node_of(a.tri, b, j, a.is_DG), a.is_DG)) continue;
                        for (int i = 0; i < 3; i++) atomicAdd(&a.val[a.posTT[9 * b + 3 * j + i]], tLP[sym6(i, j)]);
                    }
                }
            }
        }
    }
    // termLocal[i][j] = sum_k |aT| dx_k psi_i(x_k) psi_j(x_k) innerLocal_k: innerLocal_k summed
    // over the lanes of outer point k below
    sred[tid] = kin ? sumL : 0.0;
    unsigned tri = mytri;
    for (int m = 32; m > 0; m >>= 1) tri += __shfl_xor(tri, m, 64);
    if (lane == 0 && tri) atomicAdd(&a.counters[CNT_SUBTRI], (unsigned long long)tri);
    __syncthreads();
    // the accumulated columns reach the matrix: walk the three rows of aT
    for (int i = 0; i < 3; i++)
        for (int o = tid; o < rlen[i]; o += 256) {
            const int node = a.col[rp[i] + o];
            const int sl = cm.find(node);
            if (sl < 0) continue;
            const double v = acc[3 * sl + i];
            if (v == 0.0) continue;
            if (labA[i]) atomicAdd(&a.val[rp[i] + o], -v);
            if (row_labelled(a.vlabel, node, a.is_DG)) atomicAdd(&a.val[a.tpos[rp[i] + o]], -v);
        }
    if (tid < 9) {
        const int i = tid / 3, j = tid % 3;
        if (labA[i]) {
            double v = 0;
            for (int kk = 0; kk < nPx; kk++) {
                double S = 0;
                for (int ww = 0; ww < 4; ww++)
                    for (int q = 0; q < G; q++) S += sred[64 * ww + q * GS + kk];
                v += A.absDet * a.q.dx[kk] * a.q.psix[3 * kk + i] * a.q.psix[3 * kk + j] * S;
            }
            atomicAdd(&a.val[a.posTT[9 * aT + 3 * i + j]], v);
        }
    }
}

// The kernels around cut8_body.  Its register demand (about 160 VGPRs unconstrained) allows
// 3 wavefronts per SIMD; capped at 128 for 4 wavefronts, the constant and parabola instances
// spill 2 and 12 registers and run 10 % and 4 % faster, while the fractional kernel's would
// spill ~200.
template <int KER, typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_cut8_o4(NumArgs a) {
    cut8_body<KER, T>(a);
}
template <int KER, typename T>
__global__ __launch_bounds__(256) void k_cut8(NumArgs a) {
    cut8_body<KER, T>(a);
}

// termLocalPrime of the cut pairs staged by k_cut8, summed per inner element: one wavefront
// per element b over the cut entries at the end of its transposed row, then the nine entries
// (bAdx[j], bAdx[i]) of rows that are assembled (src/Cassemble.cpp:756-767).
__global__ __launch_bounds__(256) void k_lp_gather(int nE, int is_DG, const int* tri, const int* vlabel,
                                                   const long long* tptr, const int* tcB, const int* tcut,
                                                   const double* stage, const int* posTT, double* val) {
    const int b = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (b >= nE) return;
    const long long e1 = tptr[b + 1], e0 = e1 - tcB[b];
    if (e0 == e1) return;
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (long long q = e0 + lane; q < e1; q += 64) {
        const double2* st = reinterpret_cast<const double2*>(stage + (size_t)tcut[q] * 6);
        const double2 v0 = st[0], v1 = st[1], v2 = st[2];
        s[0] += v0.x; s[1] += v0.y; s[2] += v1.x; s[3] += v1.y; s[4] += v2.x; s[5] += v2.y;
    }
#pragma unroll
    for (int m = 32; m > 0; m >>= 1)
#pragma unroll
        for (int e = 0; e < 6; e++) s[e] += shfl_xor_d(s[e], m);
    if (lane < 9) {
        const int i = lane / 3, j = lane % 3;
        double v = s[0];
#pragma unroll
        for (int m = 1; m < 6; m++) v = sym6(i, j) == m ? s[m] : v;
        if (row_labelled(vlabel, node_of(tri, b, j, is_DG), is_DG)) atomicAdd(&val[posTT[9 * b + 3 * j + i]], v);
    }
}

// Cut pairs through the surrogate (constant kernel, scalar), one 256-thread block per outer
// element, chunks of up to SUR_MAXPAIRS pairs:
//   A  thread per (pair, outer point) item: ZERO / FULL / PARTIAL / OUT (sur_item); the
//      features of PARTIAL items are compacted into LDS
//   B  each wavefront runs 16-item tiles through the network(s) with BF16 MFMA
//   C  thread per pair: a pair with an OUT item or an item whose ensemble spread exceeds
//      tol goes to the exact method (rlist); otherwise its inner sums are rebuilt from the
//      full-triangle moments and the predicted ratios (M := row sums of S) and accumulated
//      over the outer points exactly as in the exact cut kernel
//   D  scatter of the accepted pairs (cut_scatter), termLocal at the end (cut_flush_local).
constexpr int SUR_ITEMS = 256;
constexpr int SUR_MAXPAIRS = 36;

template <int KER, int NNET>
__global__ __launch_bounds__(256) void k_cut_sur(NumArgs a) {
    constexpr int OD = 1, NB = 9;
    extern __shared__ int ldss[];
    int* hk = ldss;       // [HN]
    int* hv = hk + a.HN;  // [HN][3]
    __shared__ float feat[SUR_ITEMS][12];
    __shared__ float outp[SUR_ITEMS][9];
    __shared__ float spread[SUR_ITEMS][4];
    __shared__ unsigned char icls[SUR_ITEMS];
    __shared__ short islot[SUR_ITEMS];
    __shared__ int pb_[SUR_MAXPAIRS];
    __shared__ int prej[SUR_MAXPAIRS];
    __shared__ double pacc[SUR_MAXPAIRS * 2 * NB];
    __shared__ double red[4][NB];
    __shared__ int npart, nrej, nacc;
    const int aT = blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const ElemRec A = a.elem[aT];
    if (A.label == 0) return;
    const int nB = a.cntB[aT];
    if (nB == 0) {
        if (tid == 0) a.sur.rcnt[aT] = 0;
        return;
    }
    const int nPx = a.q.nPx;
    const int CP = min(SUR_ITEMS / nPx, SUR_MAXPAIRS);
    const double idelta = 1.0 / sqrt(a.kp.sqdelta);
    int rA[3];
    bool labA[3];
    for (int j = 0; j < 3; j++) {
        rA[j] = node_of(a.tri, aT, j, a.is_DG);
        labA[j] = row_labelled(a.vlabel, rA[j], a.is_DG);
    }
    for (int i = tid; i < a.HN; i += 256) hk[i] = -1;
    if (tid == 0) { nrej = 0; nacc = 0; }
    __syncthreads();
    for (int j = 0; j < 3; j++) {
        const long long r0 = a.rowptr[rA[j]];
        const int len = (int)(a.rowptr[rA[j] + 1] - r0);
        for (int i = tid; i < len; i += 256) {
            int s = lds_slot(hk, a.HN, a.col[r0 + i]);
            if (s >= 0) hv[3 * s + j] = i;
            else atomicOr(&a.counters[CNT_ERROR], 16ull);
        }
    }
    double tL[NB];
#pragma unroll
    for (int i = 0; i < NB; i++) tL[i] = 0;
    const double absA = A.absDet;
    unsigned myitems = 0;
    for (int c0 = 0; c0 < nB; c0 += CP) {
        const int np = min(CP, nB - c0);
        const int nit = np * nPx;
        __syncthreads();
        if (tid == 0) npart = 0;
        if (tid < np) {
            pb_[tid] = (int)(a.nbr[(size_t)aT * a.cap + a.cap - 1 - (c0 + tid)] & IDX_MASK);
            prej[tid] = 0;
        }
        __syncthreads();
        // A: classify the items, compact the features of the partial ones
        if (tid < nit) {
            const int p = tid / nPx, k = tid % nPx;
            const ElemRec B = a.elem[pb_[p]];
            double x[2];
            toPhys(A.E, a.q.Px[2 * k], a.q.Px[2 * k + 1], x);
            float f[12];
            const int c = sur_item(x[0], x[1], B.E, a.kp.sqdelta, idelta, a.sur.smin, a.sur.smax, f);
            icls[tid] = (unsigned char)c;
            if (c == SUR_PARTIAL) {
                const int sl = atomicAdd(&npart, 1);
                islot[tid] = (short)sl;
#pragma unroll
                for (int i = 0; i < 12; i++) feat[sl][i] = f[i];
            } else if (c == SUR_OUT) {
                prej[p] = 1;
            }
        }
        __syncthreads();
        // B: network tiles of 16 items per wavefront
        const int npi = npart;
        myitems += (tid == 0) ? npi : 0;
        {
            const int j = lane & 15, q = 4 * (lane >> 4);
            for (int tile = w; tile * 16 < npi; tile += 4) {
                const int sl = tile * 16 + j;
                const bool valid = sl < npi;
                float mo[4] = {0, 0, 0, 0}, sp = 0;
#pragma unroll 1
                for (int net = 0; net < NNET; net++) {
                    const float* nm = a.sur.norm + net * mlp::NNORM;
                    mlp::f32x4 xin = {0, 0, 0, 0};
                    if (valid && q < mlp::NIN) {
#pragma unroll
                        for (int i = 0; i < 4; i++) xin[i] = (feat[sl][q + i] - nm[mlp::NORM_MU + q + i]) * nm[mlp::NORM_ISD + q + i];
                    }
                    const mlp::f32x4 y = mlp::forward_tile(a.sur.wf + (size_t)net * mlp::NFRAG_FWD * 64,
                                                           a.sur.bias + net * mlp::BIAS_PAD, xin, lane);
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const int oi = q + i < mlp::NOUT ? q + i : 0;
                        const float v = y[i] * nm[mlp::NORM_OSD + oi] + nm[mlp::NORM_OMU + oi];
                        if (net == 0) {
                            mo[i] = v;
                        } else {  // two networks: mean and largest disagreement
                            if (q + i < mlp::NOUT) sp = fmaxf(sp, fabsf(mo[i] - v));
                            mo[i] = 0.5f * (mo[i] + v);
                        }
                    }
                }
                if (valid) {
#pragma unroll
                    for (int i = 0; i < 4; i++)
                        if (q + i < mlp::NOUT) outp[sl][q + i] = mo[i];
                    spread[sl][lane >> 4] = sp;
                }
            }
        }
        __syncthreads();
        // C: acceptance and outer accumulation, one thread per pair
        if (tid < np) {
            const int p = tid;
            bool rej = prej[p] != 0;
            for (int k = 0; k < nPx && !rej; k++) {
                const int it = p * nPx + k;
                if (icls[it] == SUR_PARTIAL) {
                    const int sl = islot[it];
                    const float sp = fmaxf(fmaxf(spread[sl][0], spread[sl][1]), fmaxf(spread[sl][2], spread[sl][3]));
                    if (!(sp <= a.sur.tol)) rej = true;
                }
            }
            const int b = pb_[p];
            if (rej) {
                const int r = atomicAdd(&nrej, 1);
                a.sur.rlist[a.sur.roff[aT] + r] = b;
                pb_[p] = -1;
            } else {
                atomicAdd(&nacc, 1);
                const double cd = a.kp.c * a.elem[b].absDet;
                double tN[NB], tLP[NB];
#pragma unroll
                for (int o = 0; o < NB; o++) { tN[o] = 0; tLP[o] = 0; }
                for (int k = 0; k < nPx; k++) {
                    const int it = p * nPx + k;
                    const int c = icls[it];
                    if (c == SUR_ZERO) continue;
                    double inL, inN[3], sLP[6];
                    if (c == SUR_FULL) {
                        inL = cd * a.sur.q0;
#pragma unroll
                        for (int bb = 0; bb < 3; bb++) inN[bb] = cd * a.sur.q1[bb];
#pragma unroll
                        for (int qq = 0; qq < 6; qq++) sLP[qq] = cd * a.sur.q2[qq];
                    } else {
                        const float* y = outp[islot[it]];
#pragma unroll
                        for (int qq = 0; qq < 6; qq++) {
                            // ratios of moments of a subset of bT lie in (0, 1]; the floor keeps
                            // a partially covered item from producing an exact zero entry
                            const float r = fminf(fmaxf(y[qq], 1e-7f), 1.0f);
                            sLP[qq] = cd * a.sur.q2[qq] * (double)r;
                        }
                        // M_a := sum_b S_ab keeps the partition of unity of the pair exact
                        inN[0] = sLP[0] + sLP[1] + sLP[2];
                        inN[1] = sLP[1] + sLP[3] + sLP[4];
                        inN[2] = sLP[2] + sLP[4] + sLP[5];
                        inL = inN[0] + inN[1] + inN[2];
                    }
                    double psx[3] = {a.q.psix[3 * k], a.q.psix[3 * k + 1], a.q.psix[3 * k + 2]};
                    outer_point<OD>(psx, absA * a.q.dx[k], &inL, inN, tL, tN);
                    add_lp<OD>(sLP, a.q.dx[k] * absA, tLP);
                }
                double* pa = pacc + p * 2 * NB;
#pragma unroll
                for (int o = 0; o < NB; o++) { pa[o] = tN[o]; pa[NB + o] = tLP[o]; }
            }
        }
        __syncthreads();
        cut_scatter<OD>(a, np, pb_, pacc, rA, labA, hk, hv);
    }
    __syncthreads();
    if (tid == 0) {
        a.sur.rcnt[aT] = nrej;
        atomicAdd(&a.counters[CNT_SUR_ACC], (unsigned long long)nacc);
        atomicAdd(&a.counters[CNT_SUR_REJ], (unsigned long long)nrej);
        atomicAdd(&a.counters[CNT_SUR_ITEMS], (unsigned long long)myitems);
    }
    cut_flush_local<OD>(a, aT, tL, labA, red);
}

// ======================================================================= close pairs
// ======================================================================= surrogate, field kernel
// Cut pairs of the microelastic field kernel (outdim 2) through the surrogate.  One block of 256
// threads per outer element, chunks of SURF_ITEMS items (pair x outer point):
//   A  one thread per item: its class (sur_item); a FULL item gets the exact inner sums over bT
//      (the inner rule on bT, as an interior pair does); a PARTIAL item queues its 12 features.
//   B  the two networks on tiles of 16 partial items per wavefront: network 0 returns the
//      tensors sLP_ab of the symmetric index pairs (00, 01, 02), network 1 those of (11, 12, 22),
//      each as (xx, xy, yy) in the item's frame and as ratios to field_scale's F_q.
//   C  one thread per partial item: the smallest change of the 18 outputs (least squares) that
//      makes sum_b sLP_ab R u_b = 0 for every a (u_b: bT's vertices relative to x, R the
//      rotation by 90 degrees).  Every exact quadrature sum satisfies these six equations
//      because z^T R z = 0 at every point, and with innerNonloc_a = sum_b sLP_ab and innerLocal
//      = sum_a innerNonloc_a the assembled matrix annihilates rotations and translations of the
//      mesh exactly.  The size of that change relative to the outputs is the residual indicator:
//      above tol (or for an OUT item) the pair goes to the exact kernel.  Then the tensors go
//      back to physical units and axes.
//   D  rejected pairs are listed and their items cleared; the outer accumulation (cut_outer),
//      the scatter (cut_scatter) and termLocal follow k_cut.
constexpr int SURF_ITEMS = 128;

static size_t surf_lds_bytes(int nPx, int HN) {
    const int cp = std::min(SURF_ITEMS / std::max(nPx, 1), CUT_MAXPAIRS);
    return 8 * ((size_t)SURF_ITEMS * 40 + (size_t)cp * 72 + (size_t)nPx * 10) + 10 * (size_t)HN;
}

// x := G^-1 r for a symmetric positive definite 6 x 6 matrix G (Cholesky, in place).
NLFEM_DEV void spd6_solve(double (&G)[6][6], double (&r)[6]) {
#pragma unroll
    for (int j = 0; j < 6; j++) {
        double d = G[j][j];
#pragma unroll
        for (int k = 0; k < j; k++) d -= G[j][k] * G[j][k];
        d = sqrt(fmax(d, 1e-300));
        G[j][j] = d;
#pragma unroll
        for (int i = j + 1; i < 6; i++) {
            double s = G[i][j];
#pragma unroll
            for (int k = 0; k < j; k++) s -= G[i][k] * G[j][k];
            G[i][j] = s / d;
        }
    }
#pragma unroll
    for (int i = 0; i < 6; i++) {
        double s = r[i];
#pragma unroll
        for (int k = 0; k < i; k++) s -= G[i][k] * r[k];
        r[i] = s / G[i][i];
    }
#pragma unroll
    for (int i = 5; i >= 0; i--) {
        double s = r[i];
#pragma unroll
        for (int k = i + 1; k < 6; k++) s -= G[k][i] * r[k];
        r[i] = s / G[i][i];
    }
}

// NNET = 2: one network per half; NNET = 4: two per half (networks 2h and 2h + 1 for half h),
// their mean is used and their largest disagreement on a ratio is the indicator instead.
template <int KER, int NNET>
__global__ __launch_bounds__(256) void k_cut_sur_field(NumArgs a) {
    static_assert(NNET == 2 || NNET == 4, "two halves, one or two networks each");
    constexpr int NE = NNET / 2;
    constexpr int OD = 2, OO = 4, NB = 36, N = 6;
    const int nPx = a.q.nPx;
    const int CP = min(SURF_ITEMS / nPx, CUT_MAXPAIRS);
    extern __shared__ double ldsf[];
    double* iacc = ldsf;                          // [SURF_ITEMS][10 OO]: innerLocal, innerNonloc, sLP
    double* pacc = iacc + SURF_ITEMS * 10 * OO;   // [CP][2 NB]: termNonloc, termLocalPrime
    double* tls = pacc + CP * 2 * NB;             // [nPx][OO]
    double* xk = tls + nPx * OO;                  // [nPx][2]
    double* wa = xk + 2 * nPx;                    // [nPx][3] |aT| dx_k psi_i(x_k)
    double* wd = wa + 3 * nPx;                    // [nPx] dx_k |aT|
    int* hk = (int*)(wd + nPx);                   // [HN]
    unsigned short* hv = (unsigned short*)(hk + a.HN);
    __shared__ float feat[SURF_ITEMS][12];
    __shared__ float outp[SURF_ITEMS][18];
    __shared__ double rotd[SURF_ITEMS][2];
    __shared__ float spread[SURF_ITEMS][4];
    __shared__ short pitem[SURF_ITEMS];  // partial slot -> item
    __shared__ int pb_[CUT_MAXPAIRS];
    __shared__ int prej[CUT_MAXPAIRS];
    __shared__ int npart, nrej, nacc;
    const int aT = blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const ElemRec A = a.elem[aT];
    if (A.label == 0) return;
    const int nB = a.cntB[aT];
    if (nB == 0) {
        if (tid == 0) a.sur.rcnt[aT] = 0;
        return;
    }
    const double idelta = 1.0 / sqrt(a.kp.sqdelta);
    int rA[3];
    bool labA[3];
    for (int j = 0; j < 3; j++) {
        rA[j] = node_of(a.tri, aT, j, a.is_DG);
        labA[j] = row_labelled(a.vlabel, rA[j], a.is_DG);
    }
    for (int i = tid; i < a.HN; i += 256) hk[i] = -1;
    for (int i = tid; i < nPx * OO; i += 256) tls[i] = 0.0;
    if (tid < nPx) {
        toPhys(A.E, a.q.Px[2 * tid], a.q.Px[2 * tid + 1], &xk[2 * tid]);
        const double c1 = A.absDet * a.q.dx[tid];
        for (int i = 0; i < 3; i++) wa[3 * tid + i] = c1 * a.q.psix[3 * tid + i];
        wd[tid] = a.q.dx[tid] * A.absDet;
    }
    if (tid == 0) { nrej = 0; nacc = 0; }
    __syncthreads();
    for (int j = 0; j < 3; j++) {
        const long long r0 = a.rowptr[rA[j]];
        const int len = (int)(a.rowptr[rA[j] + 1] - r0);
        for (int i = tid; i < len; i += 256) {
            int s = lds_slot(hk, a.HN, a.col[r0 + i]);
            if (s >= 0) hv[3 * s + j] = (unsigned short)i;
            else atomicOr(&a.counters[CNT_ERROR], 16ull);
        }
    }
    unsigned myitems = 0;
    for (int c0 = 0; c0 < nB; c0 += CP) {
        const int np = min(CP, nB - c0);
        const int nit = np * nPx;
        __syncthreads();
        for (int i = tid; i < SURF_ITEMS * 10 * OO; i += 256) iacc[i] = 0.0;
        if (tid == 0) npart = 0;
        if (tid < np) {
            pb_[tid] = (int)(a.nbr[(size_t)aT * a.cap + a.cap - 1 - (c0 + tid)] & IDX_MASK);
            prej[tid] = 0;
        }
        __syncthreads();
        // A: classes; exact sums of FULL items; features of PARTIAL items
        if (tid < nit) {
            const int p = tid / nPx, k = tid % nPx;
            const ElemRec B = a.elem[pb_[p]];
            const double x0 = xk[2 * k], x1 = xk[2 * k + 1];
            float f[12];
            double rot[2];
            const int c = sur_item(x0, x1, B.E, a.kp.sqdelta, idelta, a.sur.smin, a.sur.smax, f, rot);
            if (c == SUR_PARTIAL) {
                const int sl = atomicAdd(&npart, 1);
                pitem[sl] = (short)tid;
#pragma unroll
                for (int i = 0; i < 12; i++) feat[sl][i] = f[i];
                rotd[sl][0] = rot[0];
                rotd[sl][1] = rot[1];
            } else if (c == SUR_OUT) {
                prej[p] = 1;
            } else if (c == SUR_FULL) {
                double inL[OO], inN[3 * OO], sLP[6 * OO];
#pragma unroll
                for (int o = 0; o < OO; o++) inL[o] = 0;
#pragma unroll
                for (int o = 0; o < 3 * OO; o++) inN[o] = 0;
#pragma unroll
                for (int o = 0; o < 6 * OO; o++) sLP[o] = 0;
                for (int i = 0; i < a.q.nPy; i++) {
                    double y[2];
                    toPhys(B.E, a.q.Py[2 * i], a.q.Py[2 * i + 1], y);
                    double K[OO];
                    kernel_eval<KER, OO, double>(x0 - y[0], x1 - y[1], a.kp, K);
                    const double psi[3] = {a.q.psiy[3 * i], a.q.psiy[3 * i + 1], a.q.psiy[3 * i + 2]};
                    inner_point<OD, double, true, true, true, true>(K, a.q.dy[i] * B.absDet, psi, inL, inN, sLP);
                }
                double* ia = iacc + tid * 10 * OO;
#pragma unroll
                for (int o = 0; o < OO; o++) ia[o] = inL[o];
#pragma unroll
                for (int o = 0; o < 3 * OO; o++) ia[OO + o] = inN[o];
#pragma unroll
                for (int o = 0; o < 6 * OO; o++) ia[4 * OO + o] = sLP[o];
            }
        }
        __syncthreads();
        // B: networks 0 and 1 (the two halves of the 18 outputs), 16 items per wavefront
        const int npi = npart;
        myitems += (tid == 0) ? npi : 0;
        {
            const int j = lane & 15, q = 4 * (lane >> 4);
            for (int tile = w; tile * 16 < npi; tile += 4) {
                const int sl = tile * 16 + j;
                const bool valid = sl < npi;
                float sp = 0;
#pragma unroll 1
                for (int h = 0; h < 2; h++) {
                    float mo[4] = {0, 0, 0, 0};
#pragma unroll 1
                    for (int m = 0; m < NE; m++) {
                        const int net = h * NE + m;
                        const float* nm = a.sur.norm + net * mlp::NNORM;
                        mlp::f32x4 xin = {0, 0, 0, 0};
                        if (valid && q < mlp::NIN) {
#pragma unroll
                            for (int i = 0; i < 4; i++) xin[i] = (feat[sl][q + i] - nm[mlp::NORM_MU + q + i]) * nm[mlp::NORM_ISD + q + i];
                        }
                        const mlp::f32x4 y = mlp::forward_tile(a.sur.wf + (size_t)net * mlp::NFRAG_FWD * 64,
                                                               a.sur.bias + net * mlp::BIAS_PAD, xin, lane);
#pragma unroll
                        for (int i = 0; i < 4; i++) {
                            const int oi = q + i < mlp::NOUT ? q + i : 0;
                            const float v = y[i] * nm[mlp::NORM_OSD + oi] + nm[mlp::NORM_OMU + oi];
                            if (m == 0) {
                                mo[i] = v;
                            } else {
                                if (q + i < mlp::NOUT) sp = fmaxf(sp, fabsf(mo[i] - v));
                                mo[i] = 0.5f * (mo[i] + v);
                            }
                        }
                    }
                    if (valid) {
#pragma unroll
                        for (int i = 0; i < 4; i++)
                            if (q + i < mlp::NOUT) outp[sl][9 * h + q + i] = mo[i];
                    }
                }
                if (valid) spread[sl][lane >> 4] = sp;
            }
        }
        __syncthreads();
        // C: constraint projection, indicator, back to physical units and axes
        for (int sl = tid; sl < npi; sl += 256) {
            const int it = pitem[sl], p = it / nPx, k = it % nPx;
            const ElemRec B = a.elem[pb_[p]];
            const double cs = rotd[sl][0], sn = rotd[sl][1];
            double Fq[6];
            field_scale(xk[2 * k], xk[2 * k + 1], B.E, a.q, 1.0 / idelta, Fq);
            double s[18];  // in units of c |det bT| / delta
#pragma unroll
            for (int i = 0; i < 18; i++) s[i] = (double)outp[sl][i] * Fq[i / 3];
            // v_b = R u_b, u_b = bT's vertex b relative to x in delta units and frame axes (FP64,
            // so that the constraints hold for the true geometry)
            double vx[3], vy[3];
#pragma unroll
            for (int b = 0; b < 3; b++) {
                const double d0 = (B.E[2 * b] - xk[2 * k]) * idelta, d1 = (B.E[2 * b + 1] - xk[2 * k + 1]) * idelta;
                const double u0 = cs * d0 + sn * d1, u1 = -sn * d0 + cs * d1;
                vx[b] = -u1;
                vy[b] = u0;
            }
            // constraint rows (a, x) = sum_b S_ab,xx v_b,x + S_ab,xy v_b,y;
            //                 (a, y) = sum_b S_ab,xy v_b,x + S_ab,yy v_b,y
            double C[6][18];
#pragma unroll
            for (int r = 0; r < 6; r++)
#pragma unroll
                for (int c = 0; c < 18; c++) C[r][c] = 0.0;
#pragma unroll
            for (int aa = 0; aa < 3; aa++)
#pragma unroll
                for (int b = 0; b < 3; b++) {
                    const int qq = sym6(aa, b);
                    C[2 * aa][3 * qq] += vx[b];
                    C[2 * aa][3 * qq + 1] += vy[b];
                    C[2 * aa + 1][3 * qq + 1] += vx[b];
                    C[2 * aa + 1][3 * qq + 2] += vy[b];
                }
            double r[6], G[6][6];
            double tr = 0;
#pragma unroll
            for (int i = 0; i < 6; i++) {
                double v = 0;
#pragma unroll
                for (int c = 0; c < 18; c++) v += C[i][c] * s[c];
                r[i] = v;
#pragma unroll
                for (int j2 = 0; j2 < 6; j2++) {
                    double g = 0;
#pragma unroll
                    for (int c = 0; c < 18; c++) g += C[i][c] * C[j2][c];
                    G[i][j2] = g;
                }
                tr += G[i][i];
            }
#pragma unroll
            for (int i = 0; i < 6; i++) G[i][i] += 1e-12 * tr;
            spd6_solve(G, r);
            double dmax = 0, smax = 0;
#pragma unroll
            for (int c = 0; c < 18; c++) {
                double d = 0;
#pragma unroll
                for (int i = 0; i < 6; i++) d += C[i][c] * r[i];
                smax = fmax(smax, fabs(s[c]));
                s[c] -= d;
                dmax = fmax(dmax, fabs(d));
            }
            if constexpr (NE == 2) {
                const float sp = fmaxf(fmaxf(spread[sl][0], spread[sl][1]), fmaxf(spread[sl][2], spread[sl][3]));
                if (!(sp <= a.sur.tol)) prej[p] = 1;
            } else {
                if (!(dmax <= a.sur.tol * fmax(smax, 1e-30))) prej[p] = 1;
            }
            const double sc = a.kp.c * B.absDet * idelta;
            double* ia = iacc + it * 10 * OO;
            double sLP[6][OO];
#pragma unroll
            for (int qq = 0; qq < 6; qq++) {
                double t[3];
                field_from_frame(s[3 * qq], s[3 * qq + 1], s[3 * qq + 2], cs, sn, t);
                sLP[qq][0] = t[0] * sc;
                sLP[qq][1] = t[1] * sc;
                sLP[qq][2] = t[1] * sc;
                sLP[qq][3] = t[2] * sc;
            }
#pragma unroll
            for (int qq = 0; qq < 6; qq++)
#pragma unroll
                for (int o = 0; o < OO; o++) ia[4 * OO + qq * OO + o] = sLP[qq][o];
            double inL[OO] = {0, 0, 0, 0};
#pragma unroll
            for (int b = 0; b < 3; b++)
#pragma unroll
                for (int o = 0; o < OO; o++) {
                    const double v = sLP[sym6(0, b)][o] + sLP[sym6(1, b)][o] + sLP[sym6(2, b)][o];
                    ia[OO + b * OO + o] = v;
                    inL[o] += v;
                }
#pragma unroll
            for (int o = 0; o < OO; o++) ia[o] = inL[o];
        }
        __syncthreads();
        // D: rejected pairs go to the exact kernel; their items are cleared
        if (tid < np) {
            if (prej[tid]) {
                const int rr = atomicAdd(&nrej, 1);
                a.sur.rlist[a.sur.roff[aT] + rr] = pb_[tid];
            } else {
                atomicAdd(&nacc, 1);
            }
        }
        __syncthreads();
        for (int t = tid; t < nit * 10 * OO; t += 256)
            if (prej[t / (10 * OO * nPx)]) iacc[t] = 0.0;
        __syncthreads();
        if (tid < np && prej[tid]) pb_[tid] = -1;
        cut_outer<OD, true>(a, np, wa, wd, iacc, pacc, tls);
        __syncthreads();
        cut_scatter<OD>(a, np, pb_, pacc, rA, labA, hk, hv);
    }
    __syncthreads();
    if (tid == 0) {
        a.sur.rcnt[aT] = nrej;
        atomicAdd(&a.counters[CNT_SUR_ACC], (unsigned long long)nacc);
        atomicAdd(&a.counters[CNT_SUR_REJ], (unsigned long long)nrej);
        atomicAdd(&a.counters[CNT_SUR_ITEMS], (unsigned long long)myitems);
    }
    // termLocal[N ra + cb] = sum_k |aT| dx_k psi_a(x_k) psi_b(x_k) innerLocal_k (accepted pairs)
    for (int t = tid; t < NB; t += 256) {
        const int ra = t / N, cb = t % N;
        const int i = ra / OD, ao = ra % OD, j = cb / OD, bo = cb % OD;
        if (!labA[i]) continue;
        double v = 0;  // (cut_outer stored component 1 in slot 2 as well)
        for (int k = 0; k < nPx; k++)
            v += A.absDet * a.q.dx[k] * a.q.psix[3 * k + i] * a.q.psix[3 * k + j] * tls[k * OO + OD * ao + bo];
        const long long p = a.posTT[9 * aT + 3 * i + j];
        atomicAdd(&a.val[p * OO + ao * OD + bo], v);
    }
}

struct CloseArgs {
    int is_DG, nPg, od;
    const ElemRec* elem;
    const int* tri;
    const int* vlabel;
    const int2* list;
    const double* tg;  // Duffy tables, 10 doubles per (case, traffo, point)
    const double* dg;
    const long long* rowptr;
    const int* col;
    double* val;
    KernelParams kp;
    unsigned long long* counters;
    const long long* zptr;  // domain-decomposition weights (null: 1)
    const int* zidx;
};

// Tensor-Gauss quadrature of a first-layer pair with a singular kernel: integrate_fractional
// (src/integration.cpp:187-288) or integrate_weakSingular (:82-185) with join (:1913-1958).
// One wavefront per pair; lanes stride over (traffo, point); wave reduction; lanes scatter.
template <int KER, int OD, bool WEAK>
__global__ __launch_bounds__(64) void k_close(CloseArgs a) {
    constexpr int OO = OD * OD, N = 3 * OD, NB = 9 * OO;
    constexpr int NACC = WEAK ? 3 * NB : 25 * OO;
    __shared__ double blk[4][NB];
    __shared__ double accs[NACC];
    const int lane = threadIdx.x;
    const int2 pr = a.list[blockIdx.x];
    const int aT = pr.x, bT = pr.y;
    // join: shared vertices first, in the order found (src/integration.cpp:1913-1958)
    int aV[3], bV[3], aS[3], bS[3], argA[3], argB[3], AinB[3] = {0, 0, 0}, BinA[3] = {0, 0, 0};
    for (int j = 0; j < 3; j++) { aV[j] = a.tri[3 * aT + j]; bV[j] = a.tri[3 * bT + j]; }
    int nEqual = 0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            if (aV[i] == bV[j]) {
                AinB[i]++; BinA[j]++;
                aS[nEqual] = aV[i]; argA[nEqual] = i;
                bS[nEqual] = bV[j]; argB[nEqual] = j;
                nEqual++;
            }
    if (nEqual < 1 || nEqual > 3) {
        if (lane == 0) atomicOr(&a.counters[CNT_ERROR], 32ull);
        return;
    }
    {
        int ia = 0, ib = 0;
        for (int i = 0; i < 3; i++) {
            if (!AinB[i]) { aS[nEqual + ia] = aV[i]; argA[nEqual + ia] = i; ia++; }
            if (!BinA[i]) { bS[nEqual + ib] = bV[i]; argB[nEqual + ib] = i; ib++; }
        }
    }
    const ElemRec EA = a.elem[aT], EB = a.elem[bT];
    double as[6], bs[6];
    for (int j = 0; j < 3; j++) {
        int ja = 0, jb = 0;
        for (int q = 0; q < 3; q++) { if (aV[q] == aS[j]) ja = q; if (bV[q] == bS[j]) jb = q; }
        as[2 * j] = EA.E[2 * ja]; as[2 * j + 1] = EA.E[2 * ja + 1];
        bs[2 * j] = EB.E[2 * jb]; bs[2 * j + 1] = EB.E[2 * jb + 1];
    }
    const double absA = absDet(as), absB = absDet(bs);
    const int cs = nEqual - 1;
    const int nT = cs == 0 ? 2 : (cs == 1 ? 5 : 6);
    const long long base = (cs == 0 ? 0 : (cs == 1 ? 2 : 7)) * (long long)a.nPg * 10;
    const int nLocal = 6 - nEqual, shift = 3 - nEqual;
    double acc[NACC];
#pragma unroll
    for (int i = 0; i < NACC; i++) acc[i] = 0;
    // integrate_fractional: the products psixy_i psixy_j are symmetric in (i, j) and vanish
    // beyond the pair's 6 - nEqual local basis functions, so only the lower triangle of the
    // first NL rows is summed (NL = 3, 4, 5 for identical, edge-sharing, vertex-sharing pairs)
    // and mirrored after the wave reduction
    auto frac_points = [&](auto nlc) {
        constexpr int NL = decltype(nlc)::value;
        for (int it = lane; it < nT * a.nPg; it += 64) {
            const int k = it % a.nPg;
            const double* t = a.tg + base + (long long)it * 10;
            double xc[2], yc[2];
            toPhys(as, t[4], t[5], xc);
            toPhys(bs, t[6], t[7], yc);
            double K[OO];
            kernel_eval<KER, OO, double>(xc[0] - yc[0], xc[1] - yc[1], a.kp, K);
            const double wdg = t[8] * a.dg[k];
            const double px[3] = {1 - t[4] - t[5], t[4], t[5]};
            const double py[3] = {1 - t[6] - t[7], t[6], t[7]};
            double pxy[5];
#pragma unroll
            for (int i = 0; i < NL; i++) {
                const int ie = i < nEqual ? i : (i < 3 ? i : i - shift);
                pxy[i] = i < nEqual ? px[i] - py[i] : (i < 3 ? px[i] : -py[ie]);
            }
#pragma unroll
            for (int o = 0; o < OO; o++) {
                const double f = K[o] * wdg;
#pragma unroll
                for (int i = 0; i < NL; i++) {
                    const double fi = f * pxy[i];
#pragma unroll
                    for (int j = 0; j <= i; j++) acc[(i * 5 + j) * OO + o] += fi * pxy[j];
                }
            }
        }
    };
    if constexpr (!WEAK) {
        if (nLocal == 3) frac_points(std::integral_constant<int, 3>{});
        else if (nLocal == 4) frac_points(std::integral_constant<int, 4>{});
        else frac_points(std::integral_constant<int, 5>{});
    }
    for (int it = lane; WEAK && it < nT * a.nPg; it += 64) {
        const int k = it % a.nPg;
        const double* t = a.tg + base + (long long)it * 10;
        const double det = WEAK ? t[9] : t[8];
        double xc[2], yc[2];
        toPhys(as, t[4], t[5], xc);
        toPhys(bs, t[6], t[7], yc);
        double K[OO];
        kernel_eval<KER, OO, double>(xc[0] - yc[0], xc[1] - yc[1], a.kp, K);
        const double wdg = det * a.dg[k];
        if constexpr (WEAK) {
            double px[3] = {1 - t[0] - t[1], t[0], t[1]};
            double py[3] = {1 - t[2] - t[3], t[2], t[3]};
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++)
#pragma unroll
                    for (int o = 0; o < OO; o++) {
                        double f = K[o] * wdg;
                        acc[(i * 3 + j) * OO + o] += f * px[i] * px[j];
                        acc[NB + (i * 3 + j) * OO + o] += f * py[i] * py[j];
                        acc[2 * NB + (i * 3 + j) * OO + o] += f * px[i] * py[j];
                    }
        }
    }
    // wave reduction into LDS (fractional: the lower triangle, mirrored)
#pragma unroll
    for (int i = 0; i < NACC; i++) {
        if (!WEAK && (i / OO) % 5 > (i / OO) / 5) continue;  // j > i
        double v = acc[i];
        for (int m = 32; m > 0; m >>= 1) v += shfl_xor_d(v, m);
        if (lane == 0) {
            accs[i] = v;
            if (!WEAK) accs[((i / OO) % 5 * 5 + (i / OO) / 5) * OO + i % OO] = v;
        }
    }
    for (int i = lane; i < 4 * NB; i += 64) (&blk[0][0])[i] = 0;
    __syncthreads();
    // every block carries |aT| |bT|, and with a domain decomposition the pair's weight
    const double ab = absA * absB * (a.zptr ? zeta_weight(a.zptr, a.zidx, aT, bT) : 1.0);
    if (lane == 0) {
        double* tL = blk[0];
        double* tN = blk[1];
        double* tLP = blk[2];
        double* tNP = blk[3];
        if constexpr (WEAK) {
            for (int i = 0; i < 3; i++)
                for (int ao = 0; ao < OD; ao++)
                    for (int j = 0; j < 3; j++)
                        for (int bo = 0; bo < OD; bo++) {
                            const int o = ao * OD + bo;
                            tL[N * (argA[i] * OD + ao) + argA[j] * OD + bo] += accs[(i * 3 + j) * OO + o] * ab;
                            tLP[N * (argB[i] * OD + ao) + argB[j] * OD + bo] += accs[NB + (i * 3 + j) * OO + o] * ab;
                            tN[N * (argA[i] * OD + ao) + argB[j] * OD + bo] += accs[2 * NB + (i * 3 + j) * OO + o] * ab;
                            tNP[N * (argA[i] * OD + ao) + argB[j] * OD + bo] += accs[2 * NB + (i * 3 + j) * OO + o] * ab;
                        }
        } else {
            for (int i = 0; i < nLocal; i++)
                for (int ao = 0; ao < OD; ao++)
                    for (int j = 0; j < nLocal; j++)
                        for (int bo = 0; bo < OD; bo++) {
                            const double f = accs[(i * 5 + j) * OO + ao * OD + bo] * ab;
                            if (i < 3) {
                                if (j < 3) tL[N * (argA[i] * OD + ao) + argA[j] * OD + bo] += f;
                                else tN[N * (argA[i] * OD + ao) + argB[j - shift] * OD + bo] -= f;
                            } else {
                                if (j < 3) tNP[N * (argA[j] * OD + ao) + argB[i - shift] * OD + bo] -= f;
                                else tLP[N * (argB[i - shift] * OD + ao) + argB[j - shift] * OD + bo] += f;
                            }
                        }
        }
    }
    __syncthreads();
    // scatter (src/Cassemble.cpp:721-770); lanes split the 4*NB entries
    int nA[3], nB[3];
    for (int j = 0; j < 3; j++) { nA[j] = node_of(a.tri, aT, j, a.is_DG); nB[j] = node_of(a.tri, bT, j, a.is_DG); }
    for (int t = lane; t < 4 * NB; t += 64) {
        const int X = t / NB, e = t % NB, ra = e / N, cb = e % N;
        const double v = blk[X][e];
        if (v == 0.0) continue;
        int rn, rc, cn, cc;
        double sgn;
        if (X == 0) { rn = nA[ra / OD]; rc = ra % OD; cn = nA[cb / OD]; cc = cb % OD; sgn = 1; }
        else if (X == 1) { rn = nA[ra / OD]; rc = ra % OD; cn = nB[cb / OD]; cc = cb % OD; sgn = -1; }
        else if (X == 2) { rn = nB[cb / OD]; rc = cb % OD; cn = nB[ra / OD]; cc = ra % OD; sgn = 1; }
        else { rn = nB[cb / OD]; rc = cb % OD; cn = nA[ra / OD]; cc = ra % OD; sgn = -1; }
        if (!row_labelled(a.vlabel, rn, a.is_DG)) continue;
        const long long p = find_pos(a.rowptr, a.col, rn, cn);
        if (p < 0) { atomicOr(&a.counters[CNT_ERROR], 64ull); continue; }
        atomicAdd(&a.val[p * OO + rc * OD + cc], sgn * v);
    }
}

// ======================================================================= compaction
// One wavefront per node row: scalar rows od*r + o keep the nonzero entries, in column order.
template <bool FILL>
__global__ __launch_bounds__(64) void k_compact(int nNodes, int od, const long long* rowptr, const int* col,
                                                const double* val, int* ocnt, const long long* orow, int* ocol,
                                                double* oval) {
    const int r = blockIdx.x;
    const int lane = threadIdx.x;
    const long long r0 = rowptr[r], len = rowptr[r + 1] - r0;
    const int oo = od * od;
    for (int o = 0; o < od; o++) {
        const long long total = len * od;
        long long written = 0;
        long long outBase = FILL ? orow[(long long)r * od + o] : 0;
        for (long long t0 = 0; t0 < total; t0 += 64) {
            const long long t = t0 + lane;
            bool nz = false;
            long long p = 0;
            int oc = 0;
            if (t < total) {
                p = r0 + t / od;
                oc = (int)(t % od);
                nz = val[p * oo + o * od + oc] != 0.0;
            }
            const unsigned long long m = __ballot(nz);
            if (FILL && nz) {
                const long long idx = outBase + written + __popcll(m & ((1ull << lane) - 1ull));
                ocol[idx] = col[p] * od + oc;
                oval[idx] = val[p * oo + o * od + oc];
            }
            written += __popcll(m);
        }
        if (!FILL && lane == 0) ocnt[(long long)r * od + o] = (int)written;
    }
}

// ======================================================================= host side
static long long next_pow2(long long v) {
    long long p = 1;
    while (p < v) p <<= 1;
    return p;
}

static void ev_begin(Plan& P, int ph) {
    if (P.timing) HIPCHECK(hipEventRecord(P.ev[2 * ph], P.stream));
}
static void ev_end(Plan& P, int ph) {
    if (P.timing) HIPCHECK(hipEventRecord(P.ev[2 * ph + 1], P.stream));
}

template <typename Tp>
static void dalloc(Tp*& p, size_t n) {
    if (p) { (void)hipFree(p); p = nullptr; }
    if (n == 0) n = 1;
    HIPCHECK(hipMalloc((void**)&p, n * sizeof(Tp)));
}

// Plan-owned buffer of at least n elements: kept when it is large enough (its contents are
// not preserved otherwise).
template <typename Tp>
static void dreserve(Plan& P, Tp*& p, size_t n) {
    if (n == 0) n = 1;
    size_t& cap = P.dcap[(const void*)&p];
    if (p && cap >= n) return;
    if (p) { (void)hipFree(p); p = nullptr; }
    HIPCHECK(hipMalloc((void**)&p, n * sizeof(Tp)));
    cap = n;
}

static void read_counters(Plan& P, unsigned long long* h) {
    HIPCHECK(hipMemcpyAsync(h, P.d_counters, NCNT * sizeof(unsigned long long), hipMemcpyDeviceToHost, P.stream));
    HIPCHECK(hipStreamSynchronize(P.stream));
}

// Duffy tables (src/MeshTypes.cpp:252-328, maps :32-209, scale/mirror
// src/mathhelpers.cpp:88-97): per (case, traffo, point) 10 doubles
// [alpha(4), alphaCanceled(4), detFractional, detWeak].
static void duffy_map(int cs, int t, double* a, double* jac) {
    const double xi = a[0], e1 = a[1], e2 = a[2], e3 = a[3];
    double j;
    if (cs == 0) {
        if (t == 0) { a[0] = xi; a[1] = e1 * xi; a[2] = e2 * xi; a[3] = e2 * e3 * xi; }
        else { a[0] = xi * e2; a[1] = xi * e2 * e3; a[2] = xi; a[3] = xi * e1; }
        j = std::pow(xi, 3) * e2;
    } else if (cs == 1) {
        switch (t) {
            case 0: a[0] = xi; a[1] = xi * e1 * e3; a[2] = xi * (1. - e1 * e2); a[3] = xi * e1 * (1. - e2); break;
            case 1: a[0] = xi; a[1] = xi * e1; a[2] = xi * (1. - e1 * e2 * e3); a[3] = xi * e1 * e2 * (1. - e3); break;
            case 2: a[0] = xi * (1. - e1 * e2); a[1] = xi * e1 * (1. - e2); a[2] = xi; a[3] = xi * e1 * e2 * e3; break;
            case 3: a[0] = xi * (1. - e1 * e2 * e3); a[1] = xi * e1 * e2 * (1. - e3); a[2] = xi; a[3] = xi * e1; break;
            default: a[0] = xi * (1. - e1 * e2 * e3); a[1] = xi * e1 * (1. - e2 * e3); a[2] = xi; a[3] = xi * e1 * e2; break;
        }
        j = (t == 0) ? std::pow(xi, 3) * (e1 * e1) : std::pow(xi, 3) * (e1 * e1) * e2;
    } else {
        switch (t) {
            case 0: a[0] = xi; a[1] = xi * (1. - e1 + e1 * e2); a[2] = xi * (1. - e1 * e2 * e3); a[3] = xi * (1. - e1); break;
            case 1: a[0] = xi * (1. - e1 * e2 * e3); a[1] = xi * (1. - e1); a[2] = xi; a[3] = xi * (1. - e1 + e1 * e2); break;
            case 2: a[0] = xi; a[1] = xi * e1 * (1. - e2 + e2 * e3); a[2] = xi * (1. - e1 * e2); a[3] = xi * e1 * (1. - e2); break;
            case 3: a[0] = xi * (1. - e1 * e2); a[1] = xi * e1 * (1. - e2); a[2] = xi; a[3] = xi * e1 * (1. - e2 + e2 * e3); break;
            case 4: a[0] = xi * (1. - e1 * e2 * e3); a[1] = xi * e1 * (1. - e2 * e3); a[2] = xi; a[3] = xi * e1 * (1. - e2); break;
            default: a[0] = xi; a[1] = xi * e1 * (1. - e2); a[2] = xi * (1. - e1 * e2 * e3); a[3] = xi * e1 * (1. - e2 * e3); break;
        }
        j = std::pow(xi, 3) * (e1 * e1) * e2;
    }
    *jac = j;
}

std::vector<double> duffy_tables(const std::vector<double>& Pg, int nPg, double s) {
    std::vector<double> out((size_t)13 * nPg * 10);
    size_t o = 0;
    for (int cs = 0; cs < 3; cs++) {
        const int nT = cs == 0 ? 2 : (cs == 1 ? 5 : 6);
        for (int t = 0; t < nT; t++)
            for (int k = 0; k < nPg; k++) {
                double al[4], ac[4], jac;
                for (int j = 0; j < 4; j++) { al[j] = Pg[4 * k + j] * 0.5 + 0.5; ac[j] = al[j]; }
                double dw = std::pow(ac[0], 3 - 2 - 2 * s);
                double df = std::pow(ac[0], 3 - 2 * s);
                ac[0] = 1.0;
                duffy_map(cs, t, al, &jac);
                duffy_map(cs, t, ac, &jac);
                dw *= jac * 0.0625;
                df *= jac * 0.0625;
                al[0] = al[0] - al[1]; al[2] = al[2] - al[3];
                ac[0] = ac[0] - ac[1]; ac[2] = ac[2] - ac[3];
                double* r = &out[o];
                for (int j = 0; j < 4; j++) { r[j] = al[j]; r[4 + j] = ac[j]; }
                r[8] = df;
                r[9] = dw;
                o += 10;
            }
    }
    return out;
}

KernelParams kernel_params(const PlanConfig& c) {
    KernelParams kp;
    kp.sqdelta = c.sqdelta;
    const double delta = std::sqrt(c.sqdelta);
    kp.exponent = 0;
    switch (c.kernel) {
        case KER_CONSTANT: kp.c = 4 / (M_PI * (c.sqdelta * c.sqdelta)); break;               // src/model.cpp:25-28
        case KER_PARABOLA: kp.c = 12. / (M_PI * std::pow(c.sqdelta, 3)); break;              // :52-60
        case KER_FRACTIONAL:                                                                 // :126-138
            kp.exponent = 2 + 2. * c.fractional_s;
            kp.c = (2. - 2. * c.fractional_s) / (M_PI * std::pow(delta, 2. - 2. * c.fractional_s));
            break;
        case KER_MICRO: kp.c = 3.0 / (M_PI * std::pow(delta, 3)); break;                     // :116-124
        default: kp.c = 3.0 / std::pow(delta, 3); break;                                     // :140-153
    }
    return kp;
}

// Upload the mesh and build the element records.
// A processing order of the outer elements that groups them by square tiles of the bounding box
// (about 16 x 16 element sizes each, row-major over the tiles), each tile's elements in index
// order.  A kernel whose block b takes the elements ord[tg b .. tg b + tg) with the blocks of one
// XCD on one contiguous range then has its concurrent blocks working in a compact patch of the
// mesh instead of a strip of whole mesh rows, so the rows they write share the XCD's L2.
static std::vector<int> tile_order(int nE, int nV, const long* tri, const double* v) {
    std::vector<int> ord(nE);
    double x0 = 1e300, x1 = -1e300, y0 = 1e300, y1 = -1e300;
    for (int i = 0; i < nV; i++) {
        x0 = std::min(x0, v[2 * i]); x1 = std::max(x1, v[2 * i]);
        y0 = std::min(y0, v[2 * i + 1]); y1 = std::max(y1, v[2 * i + 1]);
    }
    const double area = (x1 - x0) * (y1 - y0);
    double cell = 16.0 * std::sqrt(2.0 * area / std::max(nE, 1));
    if (!(cell > 0) || !std::isfinite(cell)) {
        for (int i = 0; i < nE; i++) ord[i] = i;
        return ord;
    }
    long long ntx = 0, nty = 0;
    for (;;) {
        ntx = (long long)((x1 - x0) / cell) + 1;
        nty = (long long)((y1 - y0) / cell) + 1;
        if (ntx * nty <= (1LL << 22)) break;
        cell *= 2.0;
    }
    std::vector<int> key(nE);
    std::vector<int> start((size_t)(ntx * nty) + 1, 0);
    for (int e = 0; e < nE; e++) {
        double cx = 0, cy = 0;
        for (int j = 0; j < 3; j++) {
            const long vi = tri[3 * (size_t)e + j];
            cx += v[2 * vi];
            cy += v[2 * vi + 1];
        }
        const long long tx = std::min<long long>(ntx - 1, std::max<long long>(0, (long long)((cx / 3 - x0) / cell)));
        const long long ty = std::min<long long>(nty - 1, std::max<long long>(0, (long long)((cy / 3 - y0) / cell)));
        key[e] = (int)(ty * ntx + tx);
        start[key[e] + 1]++;
    }
    for (size_t t = 1; t < start.size(); t++) start[t] += start[t - 1];
    for (int e = 0; e < nE; e++) ord[start[key[e]]++] = e;
    return ord;
}

void plan_setup(Plan& P, const long* tri64, const long* lab64, const double* verts, const long* vlab64) {
    const PlanConfig& c = P.cfg;
    P.oo = c.outdim * c.outdim;
    P.nNodes = c.is_DG ? 3 * c.nE : c.nV;
    HIPCHECK(hipStreamCreateWithFlags(&P.stream, hipStreamNonBlocking));
    for (int i = 0; i < 2 * PH_COUNT; i++) HIPCHECK(hipEventCreate(&P.ev[i]));
    long* d_tri64;
    long* d_lab64;
    long* d_vlab64;
    double* d_verts;
    HIPCHECK(hipMalloc(&d_tri64, sizeof(long) * 3 * (size_t)c.nE));
    HIPCHECK(hipMalloc(&d_lab64, sizeof(long) * (size_t)c.nE));
    HIPCHECK(hipMalloc(&d_vlab64, sizeof(long) * (size_t)c.nV));
    HIPCHECK(hipMalloc(&d_verts, sizeof(double) * 2 * (size_t)c.nV));
    HIPCHECK(hipMemcpyAsync(d_tri64, tri64, sizeof(long) * 3 * (size_t)c.nE, hipMemcpyHostToDevice, P.stream));
    HIPCHECK(hipMemcpyAsync(d_lab64, lab64, sizeof(long) * (size_t)c.nE, hipMemcpyHostToDevice, P.stream));
    HIPCHECK(hipMemcpyAsync(d_vlab64, vlab64, sizeof(long) * (size_t)c.nV, hipMemcpyHostToDevice, P.stream));
    HIPCHECK(hipMemcpyAsync(d_verts, verts, sizeof(double) * 2 * (size_t)c.nV, hipMemcpyHostToDevice, P.stream));
    dalloc(P.d_tri, 3 * (size_t)c.nE);
    dalloc(P.d_elem, (size_t)c.nE);
    dalloc(P.d_vlabel, (size_t)c.nV);
    dalloc(P.d_counters, NCNT);
    HIPCHECK(hipMemsetAsync(P.d_counters, 0, NCNT * sizeof(unsigned long long), P.stream));
    k_setup<<<nblk(c.nE, 256), 256, 0, P.stream>>>(c.nE, d_tri64, d_lab64, d_verts, P.d_tri, P.d_elem,
                                                   P.d_counters, c.nV);
    k_vlabel<<<nblk(c.nV, 256), 256, 0, P.stream>>>(c.nV, d_vlab64, P.d_vlabel);
    HIPCHECK(hipGetLastError());
    const std::vector<int> ord = tile_order(c.nE, c.nV, tri64, verts);
    dalloc(P.d_ord, (size_t)c.nE);
    HIPCHECK(hipMemcpyAsync(P.d_ord, ord.data(), sizeof(int) * (size_t)c.nE, hipMemcpyHostToDevice, P.stream));
    if (c.nPg > 0) {
        std::vector<double> tg = duffy_tables(c.Pg, c.nPg, c.fractional_s);
        dalloc(P.d_tg, tg.size());
        dalloc(P.d_dg, (size_t)c.nPg);
        HIPCHECK(hipMemcpyAsync(P.d_tg, tg.data(), sizeof(double) * tg.size(), hipMemcpyHostToDevice, P.stream));
        HIPCHECK(hipMemcpyAsync(P.d_dg, c.dg.data(), sizeof(double) * c.nPg, hipMemcpyHostToDevice, P.stream));
    }
    const size_t scan_n = 4u << 20;
    HIPCHECK(hipMalloc(&P.d_scan_tmp, scan_n * sizeof(long long)));
    P.scan_tmp_bytes = scan_n * sizeof(long long);
    HIPCHECK(hipStreamSynchronize(P.stream));
    unsigned long long h[NCNT];
    read_counters(P, h);
    (void)hipFree(d_tri64);
    (void)hipFree(d_lab64);
    (void)hipFree(d_vlab64);
    (void)hipFree(d_verts);
    if (h[CNT_ERROR]) throw std::runtime_error("element references a vertex index out of range");
}

static long long* scan_scratch(Plan& P) { return (long long*)P.d_scan_tmp; }

namespace mlp {
void upload(const float* params, const float* norm, uint16_t* d_wf, float* d_bias, float* d_norm, hipStream_t st);
}

static NumArgs base_numargs(const Plan& P) {
    const PlanConfig& c = P.cfg;
    NumArgs na{};
    na.nE = c.nE; na.cap = P.cap; na.is_DG = c.is_DG; na.caps = c.caps; na.od = c.outdim;
    na.elem = P.d_elem; na.tri = P.d_tri; na.vlabel = P.d_vlabel; na.nbr = P.d_nbr; na.cntF = P.d_cntF;
    na.cntB = P.d_cntB; na.rowptr = P.d_rowptr; na.col = P.d_col; na.tpos = P.d_tpos; na.posTT = P.d_posTT;
    na.val = P.d_val; na.q = c.q; na.kp = kernel_params(c); na.counters = P.d_counters;
    na.sq_in = c.sqdelta * (1.0 - 1e-9);
    na.sq_far = c.sqdelta * (1.0 + 1e-9);
    {
        const QuadArgs& q = c.q;
        na.pdx = 0; na.q0 = 0;
        for (int i = 0; i < 3; i++) { na.p1[i] = 0; na.q1[i] = 0; }
        for (int i = 0; i < 6; i++) { na.p2[i] = 0; na.q2[i] = 0; }
        for (int k = 0; k < q.nPx; k++) {
            const double* ps = &q.psix[3 * k];
            const double pp[6] = {ps[0] * ps[0], ps[0] * ps[1], ps[0] * ps[2], ps[1] * ps[1], ps[1] * ps[2], ps[2] * ps[2]};
            na.pdx += q.dx[k];
            for (int b = 0; b < 3; b++) na.p1[b] += q.dx[k] * ps[b];
            for (int m = 0; m < 6; m++) na.p2[m] += q.dx[k] * pp[m];
        }
        for (int m = 0; m < 5; m++) na.qm[m] = 0;
        for (int i = 0; i < q.nPy; i++) {
            const double* ps = &q.psiy[3 * i];
            const double pp[6] = {ps[0] * ps[0], ps[0] * ps[1], ps[0] * ps[2], ps[1] * ps[1], ps[1] * ps[2], ps[2] * ps[2]};
            na.q0 += q.dy[i];
            for (int b = 0; b < 3; b++) na.q1[b] += q.dy[i] * ps[b];
            for (int m = 0; m < 6; m++) na.q2[m] += q.dy[i] * pp[m];
            const double P0 = q.Py[2 * i], P1 = q.Py[2 * i + 1];
            na.qm[0] += q.dy[i] * P0;
            na.qm[1] += q.dy[i] * P1;
            na.qm[2] += q.dy[i] * P0 * P0;
            na.qm[3] += q.dy[i] * P0 * P1;
            na.qm[4] += q.dy[i] * P1 * P1;
        }
    }
    na.sur = SurArgs{};
    na.sur.stride = 1;
    na.zptr = P.d_zptr;
    na.zidx = P.d_zidx;
    return na;
}

// The surrogate covers cut pairs of the scalar constant kernel with the retriangulate method
// and the 7-point outer rule it was trained with.
static bool surrogate_applies(const Plan& P) {
    const PlanConfig& c = P.cfg;
    if (P.sur.nnet == 0 || c.method_remote != M_RETRIANGULATE || P.stats.bary > 0 || P.d_zptr) return false;
    const long long HN = next_pow2((long long)(1.3 * std::max(P.maxRow, 1)));
    if (c.kernel == KER_CONSTANT && c.outdim == 1)
        return !c.singular && c.q.nPx <= SUR_ITEMS / 4 && HN * 16 <= 160 * 1024 && HN * 28 <= 160 * 1024;
    // field kernel: the two networks are the two halves of the 18 outputs
    return c.kernel == KER_MICRO_FIELD && c.outdim == 2 && (P.sur.nnet == 2 || P.sur.nnet == 4) && !c.is_DG &&
           c.q.nPx <= SURF_ITEMS && c.q.nPx <= CUT_ITEMS && c.q.nPx * CUT_MAXPAIRS >= CUT_ITEMS &&
           surf_lds_bytes(c.q.nPx, (int)HN) <= 160 * 1024 && cut_lds_bytes(c.q.nPx, 4, (int)HN, CUT_ITEMS) <= 160 * 1024 &&
           P.maxRow <= 65535;
}

static void set_surrogate_args(const Plan& P, SurArgs& s) {
    const QuadArgs& q = P.cfg.q;
    s.wf = (const mlp::bf16x4*)P.sur.d_wf;
    s.bias = P.sur.d_bias;
    s.norm = P.sur.d_norm;
    s.nnet = P.sur.nnet;
    s.tol = P.sur.tol;
    s.smin = P.sur.smin;
    s.smax = P.sur.smax;
    // full-triangle sums of the inner rule in the order of inner_point
    s.q0 = 0;
    for (int i = 0; i < 3; i++) s.q1[i] = 0;
    for (int i = 0; i < 6; i++) s.q2[i] = 0;
    for (int i = 0; i < q.nPy; i++) {
        const double* ps = &q.psiy[3 * i];
        const double pp[6] = {ps[0] * ps[0], ps[0] * ps[1], ps[0] * ps[2], ps[1] * ps[1], ps[1] * ps[2], ps[2] * ps[2]};
        s.q0 += q.dy[i];
        for (int b = 0; b < 3; b++) s.q1[b] += ps[b] * q.dy[i];
        for (int k = 0; k < 6; k++) s.q2[k] += pp[k] * q.dy[i];
    }
    s.rlist = P.d_rlist;
    s.roff = P.d_roff;
    s.rcnt = P.d_rcnt;
    s.list_in = 0;
    s.stride = 1;
}

void plan_assemble(Plan& P, const Estimates& est) {
    const PlanConfig& c = P.cfg;
    const hipStream_t st = P.stream;
    unsigned long long h[NCNT];
    HIPCHECK(hipMemsetAsync(P.d_counters, 0, NCNT * sizeof(unsigned long long), st));

    // ---------------------------------------------------------------- dual graph
    ev_begin(P, PH_DUAL);
    int*& d_cnt = P.d_wcnt;
    dreserve(P, d_cnt, (size_t)std::max(c.nE, c.nV) + 1);
    HIPCHECK(hipMemsetAsync(d_cnt, 0, sizeof(int) * (size_t)c.nV, st));
    dreserve(P, P.d_vptr, (size_t)c.nV + 1);
    dreserve(P, P.d_vel, 3 * (size_t)c.nE);
    k_vinc_count<<<nblk(c.nE, 256), 256, 0, st>>>(c.nE, P.d_tri, d_cnt);
    scan_exclusive<int>(d_cnt, P.d_vptr, c.nV, scan_scratch(P), st);
    HIPCHECK(hipMemsetAsync(d_cnt, 0, sizeof(int) * (size_t)c.nV, st));
    k_vinc_fill<<<nblk(c.nE, 256), 256, 0, st>>>(c.nE, P.d_tri, P.d_vptr, d_cnt, P.d_vel);
    k_sort_vinc<<<nblk(c.nV, 256), 256, 0, st>>>(c.nV, P.d_vptr, P.d_vel);
    dreserve(P, P.d_xadj, (size_t)c.nE + 1);
    k_dual<false><<<nblk(c.nE, 128), 128, 0, st>>>(c.nE, P.d_tri, P.d_vptr, P.d_vel, d_cnt, nullptr, nullptr,
                                                   P.d_counters);
    scan_exclusive<int>(d_cnt, P.d_xadj, c.nE, scan_scratch(P), st);
    long long nadj = 0;
    HIPCHECK(hipMemcpyAsync(&nadj, P.d_xadj + c.nE, sizeof(long long), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    dreserve(P, P.d_adj, (size_t)nadj);
    k_dual<true><<<nblk(c.nE, 128), 128, 0, st>>>(c.nE, P.d_tri, P.d_vptr, P.d_vel, nullptr, P.d_xadj, P.d_adj,
                                                  P.d_counters);
    HIPCHECK(hipGetLastError());
    read_counters(P, h);
    const int maxdeg = (int)h[CNT_MAXT];
    int W4 = (maxdeg + 3) / 4;
    if (W4 > BFS_W4MAX) W4 = 0;  // irregular meshes: BFS walks the CSR rows
    if (W4 > 0) dreserve(P, P.d_adjp, (size_t)c.nE * 4 * W4);
    HIPCHECK(hipMemsetAsync(P.d_counters + CNT_MAXROW, 0, sizeof(unsigned long long), st));
    k_adjpad<<<nblk(c.nE, 256), 256, 0, st>>>(c.nE, P.d_xadj, P.d_adj, 4 * W4, W4 > 0 ? P.d_adjp : nullptr,
                                               P.d_counters + CNT_MAXROW);
    HIPCHECK(hipGetLastError());
    read_counters(P, h);
    const long long bandwidth = (long long)h[CNT_MAXROW];
    ev_end(P, PH_DUAL);

    // ---------------------------------------------------------------- BFS
    const double delta = std::sqrt(c.sqdelta);
    double area_est = M_PI * std::pow(delta + 2.0 * est.hmax, 2) / est.aref();
    long long visit = c.fullCC ? (long long)c.nE : (long long)(1.3 * area_est) + 128;
    if (visit > c.nE) visit = c.nE;
    dreserve(P, P.d_cntF, (size_t)c.nE);
    dreserve(P, P.d_cntB, (size_t)c.nE);
    const long long closeCap = (long long)c.nE * 32 + 1;
    dreserve(P, P.d_closeList, (size_t)closeCap);
    BfsArgs ba;
    ba.nE = c.nE; ba.elem = P.d_elem; ba.xadj = P.d_xadj; ba.adj = P.d_adj; ba.q = c.q;
    ba.adjp = P.d_adjp; ba.W4 = W4;
    ba.sqdelta = c.sqdelta;
    ba.sq_in = c.sqdelta * (1.0 - 1e-9);
    ba.sq_cut = c.sqdelta * (1.0 - 1e-9);
    ba.sq_far = c.sqdelta * (1.0 + 1e-9);
    ba.caps = c.caps;
    ba.remote = c.method_remote;
    ba.level0 = c.singular ? (c.method_close >= 2 ? 2 : c.method_close) : c.method_remote;
    ba.fullCC = c.fullCC;
    ba.cntF = nullptr;
    ba.ord = P.d_ord;
    ba.closeCap = closeCap;
    // LDS sizing: the visited hash holds ~1.05x the elements of the delta+2h disc (load <= ~0.65
    // after rounding up to a power of two), the level buffers ~8 sqrt(disc) (a ring of the
    // disc).  Small tables keep ~13 single-wave blocks per CU resident; overflow of any table is
    // detected and the BFS is rerun with everything doubled.
    long long hvis = c.fullCC ? (long long)c.nE : (long long)(1.05 * area_est) + 64;
    long long fvis = (long long)(8.0 * std::sqrt(area_est)) + 64;
    if (P.bfs_visit > 0) { visit = P.bfs_visit; hvis = P.bfs_hvis; fvis = P.bfs_fvis; }  // sizes that worked before
    // Bitmap visited set: after L levels the BFS has moved at most L * bandwidth indices from
    // aT, and it needs at most about (delta + 2 hmax) / hmin levels (hmin: smallest element
    // height), so a window of +-bandwidth * (levels + 3) covers the search.  An index outside
    // the window is reported and the BFS reruns with the hash.
    const double hmin = 2.0 * est.amin / est.hmax;
    const long long levels = (long long)std::ceil((delta + 2.0 * est.hmax) / hmin) + 3;
    // After a first BFS the plan knows the largest index offset any search visited
    // (CNT_MAXOFF): later assemblies of the same mesh visit the same elements, so that extent
    // is an exact window.  The level bound above takes the smallest element height of the
    // mesh, far below the typical one on jittered or mixed meshes.  The measured window is
    // taken where the estimated one keeps two searches from sharing a wavefront (16-bit
    // offsets) or lets the LDS hold fewer than the 12 wavefronts per CU the registers allow;
    // otherwise the estimate stays (a narrower window measured 2 % slower on the regular C4
    // mesh, C2's jittered mesh 23 % faster).
    long long whalf = bandwidth * levels + 1;
    if (P.bfs_maxoff > 0) {
        const long long hw = (2 * whalf + 31) / 32;
        const long long fw = std::max<long long>(next_pow2(std::max<long long>(fvis, 256)), 256);
        if (32 * hw > 65536 || 12 * 2 * (long long)sizeof(int) * (hw + fw) > 160 * 1024)
            whalf = std::min(whalf, P.bfs_maxoff + 1);
    }
    const long long bm_words = (2 * whalf + 31) / 32;
    bool use_bm = !c.fullCC && W4 > 0 && bm_words <= 8192 && !P.bfs_nobm;
    for (int attempt = 0;; attempt++) {
        long long cap = std::max<long long>(visit, 16);
        long long H = next_pow2(std::min<long long>(std::max<long long>(hvis, 64), 2 * cap));
        long long F = std::min<long long>(next_pow2(std::max<long long>(fvis, 256)), next_pow2(cap));
        if (use_bm) H = bm_words;
        // two outer elements per wavefront (k_bfs_h) with the bitmap when both searches' tables
        // stay small
        // (16-bit level buffers: a window of at most 65,536 elements)
        // (k_bfs_h: per search the visited bitmap, two level buffers of F 16-bit offsets and a bit
        // per hard candidate)
        const size_t hwords = (size_t)(H + F + F / 32 + 1);
        bool half = use_bm && !P.bfs_wave && 32 * H <= 65536 && 2 * sizeof(int) * hwords <= 64 * 1024;
        size_t ldsb = half ? 2 * sizeof(int) * hwords : sizeof(int) * (size_t)(H + 2 * F + F / 32 + 1);
        if (ldsb > 160 * 1024) {
            // too large for LDS: fall back to the largest table that fits and rely on overflow detection
            H = use_bm ? H : 16384;
            F = std::min<long long>(F, 8192);
            ldsb = sizeof(int) * (size_t)(H + 2 * F + F / 32 + 1);
        }
        ba.whalf = (int)whalf;
        P.cap = (int)cap;
        dreserve(P, P.d_nbr, (size_t)c.nE * (size_t)cap);
        ba.cap = (int)cap; ba.H = (int)H; ba.F = (int)F;
        ba.nbr = P.d_nbr; ba.cntF = P.d_cntF; ba.cntB = P.d_cntB; ba.closeList = P.d_closeList;
        ba.counters = P.d_counters;
        HIPCHECK(hipMemsetAsync(P.d_counters, 0, NCNT * sizeof(unsigned long long), st));
        ev_begin(P, PH_BFS);
        if (half) {
            if (ldsb > 64 * 1024) (void)hipFuncSetAttribute((const void*)k_bfs_h, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ldsb);
            k_bfs_h<<<(c.nE + 1) / 2, 64, ldsb, st>>>(ba);
        } else if (use_bm) {
            if (ldsb > 64 * 1024) (void)hipFuncSetAttribute((const void*)k_bfs<true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ldsb);
            k_bfs<true><<<c.nE, 64, ldsb, st>>>(ba);
        } else {
            if (ldsb > 64 * 1024) (void)hipFuncSetAttribute((const void*)k_bfs<false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ldsb);
            k_bfs<false><<<c.nE, 64, ldsb, st>>>(ba);
        }
        HIPCHECK(hipGetLastError());
        ev_end(P, PH_BFS);
        read_counters(P, h);
        if (!h[CNT_OVERFLOW]) {
            P.bfs_visit = visit; P.bfs_hvis = hvis; P.bfs_fvis = fvis;
            break;
        }
        if (use_bm && (h[CNT_OVERFLOW] & 2ull)) {  // left the bitmap window: hash from now on
            use_bm = false;
            P.bfs_nobm = true;
            continue;
        }
        if (attempt >= 3 || (H >= 16384 && ldsb > 64 * 1024 && visit >= c.nE))
            throw std::length_error("BFS capacity overflow: interaction horizon too large for this build");
        visit = std::min<long long>(2 * visit, c.nE);
        hvis = std::min<long long>(2 * hvis, 2 * (long long)c.nE);
        fvis *= 2;
    }
    if (P.bfs_maxoff == 0 && !c.fullCC) {
        // the index extent of the searches, measured once per plan from the lists (-1: the mesh
        // has label-0 elements, which searches traverse without listing them)
        k_list_extent<<<nblk(c.nE, 4 * LE_PER), 256, 0, st>>>(c.nE, P.cap, P.d_nbr, P.d_cntF, P.d_cntB, P.d_elem,
                                                              P.d_counters);
        HIPCHECK(hipGetLastError());
        unsigned long long hx[NCNT];
        read_counters(P, hx);
        P.bfs_maxoff = hx[CNT_LABEL0] ? -1 : (long long)hx[CNT_MAXOFF] + bandwidth;
    }
    P.stats.pairs = h[CNT_PAIRS];
    P.stats.probes = h[CNT_PROBES];
    P.stats.close = h[CNT_CLOSE];
    P.stats.interior = h[CNT_INTERIOR];
    P.stats.cut = h[CNT_CUT];
    P.stats.bary = h[CNT_BARY];
    P.stats.max_list = (int)h[CNT_MAXLIST];
    P.stats.cap = P.cap;
    const long long nclose = (long long)h[CNT_CLOSE];
    if (nclose > closeCap) throw std::length_error("close pair list overflow");

    // ---------------------------------------------------------------- transpose
    ev_begin(P, PH_TRANSPOSE);
    // outer elements per transpose block: with a local element numbering (the BFS stayed in its
    // index window) consecutive elements share most of their lists, and eight lists of up to
    // 1400 entries keep the LDS hash under half full
    const int tg = (use_bm && P.stats.max_list <= 1400) ? 8 : 1;  // k_transpose_fill: at most 8
    const int tblocks = (c.nE + tg - 1) / tg;
    const int tsmax = P.tunstaged ? -1 : TS;
    dreserve(P, P.d_tptr, (size_t)c.nE + 1);
    dreserve(P, P.d_coff, (size_t)c.nE + 1);
    dreserve(P, P.d_tcB, (size_t)c.nE);
    dreserve(P, P.d_tF, (size_t)c.nE);
    scan_exclusive<int>(P.d_cntB, P.d_coff, c.nE, scan_scratch(P), st);
    // Padded rows of the neighbour lists' capacity first (the relation is nearly symmetric, so
    // a transposed row rarely outgrows it): one pass, no counting.  A row that overflows
    // makes this and later assemblies of the plan count first and fill exact CSR rows.
    bool tdone = false;
    if (!P.tpad_off) {
        dreserve(P, P.d_tadj, (size_t)c.nE * P.cap);
        dreserve(P, P.d_tcut, (size_t)c.nE * P.cap);
        k_tptr_padded<<<nblk(c.nE + 1, 256), 256, 0, st>>>(c.nE, P.cap, P.d_tptr);
        HIPCHECK(hipMemsetAsync(P.d_tF, 0, sizeof(int) * (size_t)c.nE, st));
        HIPCHECK(hipMemsetAsync(P.d_tcB, 0, sizeof(int) * (size_t)c.nE, st));
        k_transpose_fill<<<tblocks, 256, 0, st>>>(c.nE, P.cap, tg, P.d_nbr, P.d_cntF, P.d_cntB, P.d_tptr, P.d_tF,
                                                  P.d_tadj, P.d_coff, P.d_tcB, P.d_tcut, P.cap, P.d_ord, tsmax);
        k_tcheck<<<nblk(c.nE, 256), 256, 0, st>>>(c.nE, P.cap, P.d_tF, P.d_tcB, P.d_counters);
        HIPCHECK(hipGetLastError());
        read_counters(P, h);
        if (h[CNT_OVERFLOW] & 16ull) {
            P.tpad_off = true;
            unsigned long long zero = 0;
            HIPCHECK(hipMemcpyAsync(P.d_counters + CNT_OVERFLOW, &zero, sizeof(zero), hipMemcpyHostToDevice, st));
        } else {
            tdone = true;
        }
    }
    if (!tdone) {
        HIPCHECK(hipMemsetAsync(d_cnt, 0, sizeof(int) * (size_t)c.nE, st));
        k_transpose_count<<<tblocks, 256, 0, st>>>(c.nE, P.cap, tg, P.d_nbr, P.d_cntF, P.d_cntB, d_cnt, P.d_ord);
        scan_exclusive<int>(d_cnt, P.d_tptr, c.nE, scan_scratch(P), st);
        long long ntadj = 0;
        HIPCHECK(hipMemcpyAsync(&ntadj, P.d_tptr + c.nE, sizeof(long long), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        dreserve(P, P.d_tadj, (size_t)ntadj);
        dreserve(P, P.d_tcut, (size_t)ntadj);
        HIPCHECK(hipMemsetAsync(P.d_tF, 0, sizeof(int) * (size_t)c.nE, st));
        HIPCHECK(hipMemsetAsync(P.d_tcB, 0, sizeof(int) * (size_t)c.nE, st));
        k_transpose_fill<<<tblocks, 256, 0, st>>>(c.nE, P.cap, tg, P.d_nbr, P.d_cntF, P.d_cntB, P.d_tptr, P.d_tF,
                                                  P.d_tadj, P.d_coff, P.d_tcB, P.d_tcut, 0, P.d_ord, tsmax);
    }
    HIPCHECK(hipGetLastError());
    ev_end(P, PH_TRANSPOSE);

    // ---------------------------------------------------------------- pattern
    ev_begin(P, PH_PATTERN);
    const double pat_est = M_PI * std::pow(delta + 3.0 * est.hmax, 2) / est.aref();
    long long ecap = c.fullCC ? c.nE : std::min<long long>((long long)(1.4 * pat_est) + 256, c.nE);
    PatArgs pa;
    pa.nNodes = P.nNodes; pa.is_DG = c.is_DG; pa.cap = P.cap; pa.tri = P.d_tri; pa.vptr = P.d_vptr;
    pa.vel = P.d_vel; pa.nbr = P.d_nbr; pa.cntF = P.d_cntF; pa.cntB = P.d_cntB; pa.tptr = P.d_tptr;
    pa.tadj = P.d_tadj; pa.tF = P.d_tF; pa.tB = P.d_tcB; pa.counters = P.d_counters;
    int*& d_rowcnt = P.d_wrowcnt;
    dreserve(P, d_rowcnt, (size_t)P.nNodes + 1);
    dreserve(P, P.d_rowptr, (size_t)P.nNodes + 1);
    int*& d_ptmp = P.d_wptmp;
    long long rc = next_pow2((long long)(0.6 * (c.is_DG ? 4 * ecap : ecap)) + 64);
    // element set as a bitmap: the row's elements lie within levels + 2 dual-graph steps of
    // the row's first element, each step moving at most `bandwidth` indices
    // (with a measured list extent m: the row's elements are the list entries of the elements T
    // that share the row's vertex, |T - first| <= bandwidth, so they lie within m + bandwidth)
    const long long pwhalf = P.bfs_maxoff > 0 ? std::min(bandwidth * (levels + 2), P.bfs_maxoff) + 1
                                              : bandwidth * (levels + 2) + 1;
    const long long pwords = next_pow2((2 * pwhalf + 31) / 32 + 8);  // 32*pwords is a multiple of 256
    bool pat_bm = !c.fullCC && !P.pat_nobm && pwords <= 4096;
    pa.whalf = (int)pwhalf;
    for (int attempt = 0;; attempt++) {
        pa.EH = pat_bm ? (int)pwords : (int)next_pow2(2 * ecap);
        pa.DH = (int)next_pow2(c.is_DG ? 4 * ecap : ecap);
        pa.EL = pat_bm ? (int)((ecap + 255) & ~255LL) : 0;
        if (rc > pa.DH) rc = pa.DH;
        pa.SB = 0;
        pa.RC = (int)rc;
        dreserve(P, d_ptmp, (size_t)P.nNodes * (size_t)rc);
        pa.tmp = d_ptmp;
        pa.rowcnt = d_rowcnt;
        pa.rowptr = nullptr;
        pa.col = nullptr;
        size_t ldsb = sizeof(int) * (size_t)(pa.EH + pa.DH + pa.EL);
        bool glob = false;  // rows too large for LDS tables: hashed sets in global memory
        if (ldsb > 160 * 1024) {
            if (pat_bm) { pat_bm = false; P.pat_nobm = true; continue; }
            glob = true;
        }
        HIPCHECK(hipMemsetAsync(P.d_counters, 0, NCNT * sizeof(unsigned long long), st));
        if (glob) {
            const int G = (int)std::min<long long>(P.nNodes, 1024);
            dreserve(P, P.d_gtab, (size_t)G * (size_t)(pa.EH + pa.DH));
            k_pattern_g<<<G, 256, 0, st>>>(pa, P.d_gtab);
        } else if (pat_bm) {
            if (ldsb > 64 * 1024) (void)hipFuncSetAttribute((const void*)k_pattern<true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ldsb);
            k_pattern<true><<<P.nNodes, 256, ldsb, st>>>(pa);
        } else {
            if (ldsb > 64 * 1024) (void)hipFuncSetAttribute((const void*)k_pattern<false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ldsb);
            k_pattern<false><<<P.nNodes, 256, ldsb, st>>>(pa);
        }
        HIPCHECK(hipGetLastError());
        read_counters(P, h);
        if (!h[CNT_OVERFLOW]) break;
        if (attempt >= 5) throw std::length_error("pattern capacity overflow");
        if (pat_bm && (h[CNT_OVERFLOW] & 8ull)) {  // left the bitmap window: hash from now on
            pat_bm = false;
            P.pat_nobm = true;
            continue;
        }
        if (h[CNT_OVERFLOW] & 2ull) ecap *= 2;   // a hash table filled up
        else rc *= 2;                            // a row outgrew its slot
    }
    P.maxRow = (int)h[CNT_MAXROW];
    scan_exclusive<int>(d_rowcnt, P.d_rowptr, P.nNodes, scan_scratch(P), st);
    HIPCHECK(hipMemcpyAsync(&P.nnz_nodes, P.d_rowptr + P.nNodes, sizeof(long long), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    if (P.nnz_nodes >= (1ll << 31)) throw std::length_error("pattern exceeds 2^31 node entries");
    dreserve(P, P.d_col, (size_t)P.nnz_nodes);
    pa.SB = (int)next_pow2(std::max(P.maxRow, 1));
    pa.rowptr = P.d_rowptr;
    pa.col = P.d_col;
    {
        size_t ldsb = sizeof(int) * (size_t)pa.SB;
        if (ldsb > 160 * 1024) throw std::length_error("pattern row too long for the LDS sort");
        if (ldsb > 64 * 1024) (void)hipFuncSetAttribute((const void*)k_pattern_sort, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ldsb);
        k_pattern_sort<<<P.nNodes, 256, ldsb, st>>>(pa);
        HIPCHECK(hipGetLastError());
    }
    ev_end(P, PH_PATTERN);

    // ---------------------------------------------------------------- positions
    ev_begin(P, PH_POSITIONS);
    int*& d_rowOf = P.d_wrowOf;
    dreserve(P, d_rowOf, (size_t)P.nnz_nodes);
    dreserve(P, P.d_tpos, (size_t)P.nnz_nodes);
    dreserve(P, P.d_posTT, 9 * (size_t)c.nE);
    k_row_of<<<P.nNodes, 256, 0, st>>>(P.nNodes, P.d_rowptr, d_rowOf);
    k_tpos<<<nblk(P.nnz_nodes, 256), 256, 0, st>>>(P.nnz_nodes, P.d_rowptr, P.d_col, d_rowOf, P.d_tpos, P.d_counters);
    k_posTT<<<nblk(c.nE, 256), 256, 0, st>>>(c.nE, P.d_tri, c.is_DG, P.d_rowptr, P.d_col, P.d_posTT, P.d_counters);
    dreserve(P, P.d_val, (size_t)P.nnz_nodes * P.oo);
    HIPCHECK(hipMemsetAsync(P.d_val, 0, sizeof(double) * (size_t)P.nnz_nodes * P.oo, st));
    HIPCHECK(hipGetLastError());
    ev_end(P, PH_POSITIONS);

    // ---------------------------------------------------------------- quadrature
    NumArgs na = base_numargs(P);
    // LDS map of the three rows of the outer element: 16 bytes per slot
    long long HN = next_pow2((long long)(2.5 * std::max(P.maxRow, 1)));
    na.HN = (HN * 16 <= 64 * 1024) ? (int)HN : 0;
    const size_t nlds = na.HN > 0 ? (size_t)na.HN * 16 : 0;

#define LAUNCH_PAIRS(KER, OD, SEG)                                                                          \
    do {                                                                                                    \
        if (c.precision == 1) k_pairs<KER, OD, SEG, float><<<c.nE, 256, nlds, st>>>(na);                    \
        else k_pairs<KER, OD, SEG, double><<<c.nE, 256, nlds, st>>>(na);                                    \
    } while (0)
#define DISPATCH_PAIRS(SEG)                                                        \
    do {                                                                           \
        switch (c.kernel) {                                                        \
            case KER_CONSTANT: LAUNCH_PAIRS(KER_CONSTANT, 1, SEG); break;          \
            case KER_PARABOLA: LAUNCH_PAIRS(KER_PARABOLA, 1, SEG); break;          \
            case KER_FRACTIONAL: LAUNCH_PAIRS(KER_FRACTIONAL, 1, SEG); break;      \
            case KER_MICRO: LAUNCH_PAIRS(KER_MICRO, 1, SEG); break;                \
            default: LAUNCH_PAIRS(KER_MICRO_FIELD, 2, SEG); break;                 \
        }                                                                          \
    } while (0)
    // With domain-decomposition weights every pair is integrated and scattered by the per-pair
    // kernel, its four blocks scaled by the pair's weight (FP64 atomics into val).
#define LAUNCH_PAIRS_Z(KER, OD, SEG)                                                                        \
    do {                                                                                                    \
        if (c.precision == 1) k_pairs<KER, OD, SEG, float, true><<<c.nE, 256, nlds, st>>>(na);              \
        else k_pairs<KER, OD, SEG, double, true><<<c.nE, 256, nlds, st>>>(na);                              \
    } while (0)
#define DISPATCH_PAIRS_Z(SEG)                                                      \
    do {                                                                           \
        switch (c.kernel) {                                                        \
            case KER_CONSTANT: LAUNCH_PAIRS_Z(KER_CONSTANT, 1, SEG); break;        \
            case KER_PARABOLA: LAUNCH_PAIRS_Z(KER_PARABOLA, 1, SEG); break;        \
            case KER_FRACTIONAL: LAUNCH_PAIRS_Z(KER_FRACTIONAL, 1, SEG); break;    \
            case KER_MICRO: LAUNCH_PAIRS_Z(KER_MICRO, 1, SEG); break;              \
            default: LAUNCH_PAIRS_Z(KER_MICRO_FIELD, 2, SEG); break;               \
        }                                                                          \
    } while (0)

    // interior pairs: owner computes into partial rows, then merge (plain stores into val).
    // Horizons whose rows outgrow the LDS column table (about 36 elements per horizon for scalar
    // kernels, 16 for outdim 2) take the per-pair kernel with global atomics instead.
    ev_begin(P, PH_INTERIOR);
    const bool int_fits = (size_t)((((long long)(1.3 * std::max(P.maxRow, 1)) + 63) / 64 * 64)) * (24 * P.oo + 4) <=
                          160 * 1024;
    if (P.d_zptr) {
        DISPATCH_PAIRS_Z(0);
        HIPCHECK(hipGetLastError());
    } else if (!int_fits) {
        DISPATCH_PAIRS(0);
        HIPCHECK(hipGetLastError());
    } else {
        int*& d_plen = P.d_wplen;
        dreserve(P, d_plen, 3 * (size_t)c.nE);
        dreserve(P, P.d_poff, 3 * (size_t)c.nE + 1);
        k_plen<<<nblk(c.nE, 256), 256, 0, st>>>(c.nE, P.d_elem, P.d_tri, P.d_vlabel, c.is_DG, P.d_rowptr, d_plen);
        scan_exclusive<int>(d_plen, P.d_poff, 3LL * c.nE, scan_scratch(P), st);
        long long ptotal = 0;
        HIPCHECK(hipMemcpyAsync(&ptotal, P.d_poff + 3LL * c.nE, sizeof(long long), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        if (P.ptotal < ptotal * P.oo) {
            dreserve(P, P.d_P, (size_t)ptotal * P.oo);
            P.ptotal = ptotal * P.oo;
        }
        na.tptr = P.d_tptr; na.tadj = P.d_tadj; na.tF = P.d_tF; na.poff = P.d_poff; na.P = P.d_P;
        // accumulator slots: 1.3x the longest row, in multiples of 64 (the column map's rank mode
        // needs one slot per distinct node of the element's three rows; an overflow is reported
        // and the table doubled)
        long long hn = ((long long)(1.3 * std::max(P.maxRow, 1)) + 63) / 64 * 64;
        for (int attempt = 0;; attempt++) {
            NumArgs ni = na;
            ni.HN = (int)hn;
            // the field kernel with the 7-point inner rule accumulates 3 components (k01 = k10)
            const int nca = (c.kernel == KER_MICRO_FIELD && c.q.nPy == 7) ? 3 : P.oo;
            const size_t lb = (size_t)hn * (24 * nca + 4);
            if (lb > 160 * 1024) throw std::length_error("interior LDS table exceeds 160 KiB");
#define LAUNCH_INT_N(KER, OD, NPY)                                                                                  \
    do {                                                                                                            \
        if (c.precision == 1) {                                                                                     \
            if (lb > 64 * 1024) (void)hipFuncSetAttribute((const void*)k_interior<KER, OD, float, NPY>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lb); \
            k_interior<KER, OD, float, NPY><<<c.nE, 256, lb, st>>>(ni);                                             \
        } else {                                                                                                    \
            if (lb > 64 * 1024) (void)hipFuncSetAttribute((const void*)k_interior<KER, OD, double, NPY>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lb); \
            k_interior<KER, OD, double, NPY><<<c.nE, 256, lb, st>>>(ni);                                            \
        }                                                                                                           \
    } while (0)
// the general kernels with the 7-point inner rule: the rule unrolled
#define LAUNCH_INT(KER, OD)                                                                                         \
    do {                                                                                                            \
        if (KER != KER_CONSTANT && c.q.nPy == 7) LAUNCH_INT_N(KER, OD, 7);                                          \
        else LAUNCH_INT_N(KER, OD, 0);                                                                              \
    } while (0)
            switch (c.kernel) {
                case KER_CONSTANT: LAUNCH_INT(KER_CONSTANT, 1); break;
                case KER_PARABOLA: LAUNCH_INT(KER_PARABOLA, 1); break;
                case KER_FRACTIONAL: LAUNCH_INT(KER_FRACTIONAL, 1); break;
                case KER_MICRO: LAUNCH_INT(KER_MICRO, 1); break;
                default: LAUNCH_INT(KER_MICRO_FIELD, 2); break;
            }
#undef LAUNCH_INT
#undef LAUNCH_INT_N
            HIPCHECK(hipGetLastError());
            read_counters(P, h);
            if (!(h[CNT_ERROR] & ERR_HASH)) break;
            if (attempt >= 2) throw std::length_error("interior LDS column table overflow");
            // clear only the overflow bit: an inconsistency reported earlier (a transposed
            // position missing) must still reach the final check
            const unsigned long long keep = h[CNT_ERROR] & ~ERR_HASH;
            HIPCHECK(hipMemcpyAsync(P.d_counters + CNT_ERROR, &keep, sizeof(keep), hipMemcpyHostToDevice, st));
            HIPCHECK(hipStreamSynchronize(st));
            hn *= 2;
        }
        k_merge<<<P.nNodes, 256, 0, st>>>(P.nNodes, P.oo, c.is_DG, P.d_elem, P.d_tri, P.d_vlabel, P.d_vptr, P.d_vel,
                                          P.d_rowptr, P.d_poff, P.d_P, P.d_val);
        HIPCHECK(hipGetLastError());
    }
    ev_end(P, PH_INTERIOR);

    if (nclose > 0) {
        ev_begin(P, PH_CLOSE);
        CloseArgs ca;
        ca.is_DG = c.is_DG; ca.nPg = c.nPg; ca.od = c.outdim; ca.elem = P.d_elem; ca.tri = P.d_tri;
        ca.vlabel = P.d_vlabel; ca.list = P.d_closeList; ca.tg = P.d_tg; ca.dg = P.d_dg; ca.rowptr = P.d_rowptr;
        ca.col = P.d_col; ca.val = P.d_val; ca.kp = na.kp; ca.counters = P.d_counters;
        ca.zptr = P.d_zptr; ca.zidx = P.d_zidx;
        const bool weak = c.method_close == M_WEAKSINGULAR;
#define LAUNCH_CLOSE(KER, OD)                                                                     \
    do {                                                                                          \
        if (weak) k_close<KER, OD, true><<<(unsigned)nclose, 64, 0, st>>>(ca);                     \
        else k_close<KER, OD, false><<<(unsigned)nclose, 64, 0, st>>>(ca);                         \
    } while (0)
        switch (c.kernel) {
            case KER_CONSTANT: LAUNCH_CLOSE(KER_CONSTANT, 1); break;
            case KER_PARABOLA: LAUNCH_CLOSE(KER_PARABOLA, 1); break;
            case KER_FRACTIONAL: LAUNCH_CLOSE(KER_FRACTIONAL, 1); break;
            case KER_MICRO: LAUNCH_CLOSE(KER_MICRO, 1); break;
            default: LAUNCH_CLOSE(KER_MICRO_FIELD, 2); break;
        }
#undef LAUNCH_CLOSE
        HIPCHECK(hipGetLastError());
        ev_end(P, PH_CLOSE);
    }

    ev_begin(P, PH_CUT);
    P.cut_items_used = 0;
    P.cutw_used = false;
    P.cut8_used = false;
    bool gather = false;  // termLocalPrime of the cut pairs staged for k_lp_gather
    P.stats.sur_acc = P.stats.sur_rej = P.stats.sur_items = 0;
    const bool use_sur = surrogate_applies(P);
    // LDS of the cut kernels: the column map of aT's rows (k_cut8: 28 B per slot; k_cut: its
    // chunk buffers and 10 B per slot); rows beyond it (very large horizons) take the per-pair
    // kernel, whose scatter searches the rows in global memory
    const int cut_HN = (int)next_pow2((long long)(1.3 * std::max(P.maxRow, 1)));
    // k_cutw (cut_items 0) where its LDS fits; the constant kernel keeps k_cut8, whose items
    // are integrated from polygon moments
    const bool cutw_ok = P.cut_items == 0 && cutw_fits(c.q.nPx) && P.maxRow <= 65535 &&
                         cutw_lds_bytes(c.q.nPx, c.outdim, c.kernel == KER_MICRO_FIELD, cut_HN) <= 160 * 1024;
    const bool use_cut8 = c.outdim == 1 && c.q.nPx <= 8 && !c.is_DG && (c.kernel == KER_CONSTANT || !cutw_ok);
    const bool cut_fits = use_cut8 ? (size_t)cut_HN * 28 <= 160 * 1024
                                   : cutw_ok || (cut_lds_bytes(c.q.nPx, P.oo, cut_HN, CUT_ITEMS) <= 160 * 1024 &&
                                                 P.maxRow <= 65535);
    if (P.d_zptr) {
        DISPATCH_PAIRS_Z(1);
    } else if (P.stats.bary > 0 || c.q.nPx > CUT_ITEMS || c.q.nPx * CUT_MAXPAIRS < CUT_ITEMS || !cut_fits) {
        DISPATCH_PAIRS(1);
    } else {
        NumArgs nc = na;
        if (use_sur) {
            // surrogate first; the pairs it hands back go through the exact kernel below
            dreserve(P, P.d_roff, (size_t)c.nE + 1);
            dreserve(P, P.d_rcnt, (size_t)c.nE);
            scan_exclusive<int>(P.d_cntB, P.d_roff, c.nE, scan_scratch(P), st);
            long long ncut = 0;
            HIPCHECK(hipMemcpyAsync(&ncut, P.d_roff + c.nE, sizeof(long long), hipMemcpyDeviceToHost, st));
            HIPCHECK(hipStreamSynchronize(st));
            if (P.rlist_cap < ncut || !P.d_rlist) {
                dreserve(P, P.d_rlist, (size_t)std::max<long long>(ncut, 1));
                P.rlist_cap = std::max<long long>(ncut, 1);
            }
            NumArgs ns = na;
            ns.HN = (int)next_pow2((long long)(1.3 * std::max(P.maxRow, 1)));
            set_surrogate_args(P, ns.sur);
            const size_t ls = (size_t)ns.HN * 16;
            if (c.kernel == KER_MICRO_FIELD) {
                const size_t lf = surf_lds_bytes(c.q.nPx, ns.HN);
                if (lf > 160 * 1024 || P.maxRow > 65535) throw std::length_error("surrogate LDS exceeds 160 KiB");  // unreachable: see surrogate_applies
                if (P.sur.nnet == 4) {
                    if (lf > 64 * 1024) (void)hipFuncSetAttribute((const void*)k_cut_sur_field<KER_MICRO_FIELD, 4>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lf);
                    k_cut_sur_field<KER_MICRO_FIELD, 4><<<c.nE, 256, lf, st>>>(ns);
                } else {
                    if (lf > 64 * 1024) (void)hipFuncSetAttribute((const void*)k_cut_sur_field<KER_MICRO_FIELD, 2>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lf);
                    k_cut_sur_field<KER_MICRO_FIELD, 2><<<c.nE, 256, lf, st>>>(ns);
                }
            } else if (P.sur.nnet == 2) {
                if (ls > 64 * 1024) (void)hipFuncSetAttribute((const void*)k_cut_sur<KER_CONSTANT, 2>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ls);
                k_cut_sur<KER_CONSTANT, 2><<<c.nE, 256, ls, st>>>(ns);
            } else {
                if (ls > 64 * 1024) (void)hipFuncSetAttribute((const void*)k_cut_sur<KER_CONSTANT, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ls);
                k_cut_sur<KER_CONSTANT, 1><<<c.nE, 256, ls, st>>>(ns);
            }
            HIPCHECK(hipGetLastError());
            nc.sur = ns.sur;
            nc.sur.list_in = 1;
        }
        nc.HN = cut_HN;
        // items per chunk of k_cut: the plan's choice when the outer rule and the LDS allow it
        int items = CUT_ITEMS;
        if (cut_items_fit(c.q.nPx, P.cut_items) && cut_lds_bytes(c.q.nPx, P.oo, nc.HN, P.cut_items) <= 160 * 1024)
            items = P.cut_items;
        const size_t lb = cut_lds_bytes(c.q.nPx, P.oo, nc.HN, items);
        if (lb > 160 * 1024 || P.maxRow > 65535) throw std::length_error("cut-pair LDS exceeds 160 KiB");
#define LAUNCH_CUT_I(KER, OD, T, IT)                                                                                  \
    do {                                                                                                             \
        if (lb > 64 * 1024) (void)hipFuncSetAttribute((const void*)k_cut<KER, OD, T, 0, IT>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lb); \
        k_cut<KER, OD, T, 0, IT><<<c.nE, 256, lb, st>>>(nc);                                                         \
    } while (0)
#define LAUNCH_CUT_T(KER, OD, T)                                      \
    do {                                                             \
        if (items == 64) LAUNCH_CUT_I(KER, OD, T, 64);               \
        else if (items == 256) LAUNCH_CUT_I(KER, OD, T, 256);        \
        else LAUNCH_CUT_I(KER, OD, T, 128);                          \
    } while (0)
#define LAUNCH_CUT(KER, OD)                                           \
    do {                                                             \
        P.cut_items_used = items;                                    \
        if (c.precision == 1) LAUNCH_CUT_T(KER, OD, float);          \
        else LAUNCH_CUT_T(KER, OD, double);                          \
    } while (0)
#define LAUNCH_CUT8_K(KFN, KER)                                                                           \
    do {                                                                                                  \
        const size_t l8 = (size_t)nc.HN * 28;                                                             \
        if (c.precision == 1) {                                                                           \
            if (l8 > 64 * 1024) (void)hipFuncSetAttribute((const void*)KFN<KER, float>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)l8); \
            KFN<KER, float><<<c.nE, 256, l8, st>>>(nc);                                                   \
        } else {                                                                                          \
            if (l8 > 64 * 1024) (void)hipFuncSetAttribute((const void*)KFN<KER, double>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)l8); \
            KFN<KER, double><<<c.nE, 256, l8, st>>>(nc);                                                  \
        }                                                                                                 \
    } while (0)
#define LAUNCH_CUT8(KER) LAUNCH_CUT8_K(k_cut8, KER)
        if (use_cut8) {
            P.cut8_used = true;
            long long ncut = 0;
            HIPCHECK(hipMemcpyAsync(&ncut, P.d_coff + c.nE, sizeof(long long), hipMemcpyDeviceToHost, st));
            HIPCHECK(hipStreamSynchronize(st));
            if (!use_sur) {
                // termLocalPrime staged per cut pair, summed per inner element by k_lp_gather
                if (ncut <= INT_MAX) {  // the transposed cut entries hold int32 pair indices
                    if (P.stage_cap < ncut || !P.d_stage) {
                        dreserve(P, P.d_stage, (size_t)std::max<long long>(ncut, 1) * 6);
                        P.stage_cap = std::max<long long>(ncut, 1);
                    }
                    nc.coff = P.d_coff;
                    nc.stage = P.d_stage;
                }
            }
            switch (c.kernel) {
                case KER_CONSTANT: LAUNCH_CUT8_K(k_cut8_o4, KER_CONSTANT); break;
                case KER_PARABOLA: LAUNCH_CUT8_K(k_cut8_o4, KER_PARABOLA); break;
                case KER_FRACTIONAL: LAUNCH_CUT8(KER_FRACTIONAL); break;
                default: LAUNCH_CUT8(KER_MICRO); break;
            }
            gather = nc.stage != nullptr;
        } else if (cutw_ok) {
            // one item per lane, no chunk barriers (k_cutw)
            const size_t lw = cutw_lds_bytes(c.q.nPx, c.outdim, c.kernel == KER_MICRO_FIELD, nc.HN);
#define LAUNCH_CUTW_T(KER, OD, T)                                                                                   \
    do {                                                                                                           \
        if (c.q.nPx == 16) {                                                                                       \
            if (lw > 64 * 1024) (void)hipFuncSetAttribute((const void*)k_cutw<KER, OD, T, 16>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lw); \
            k_cutw<KER, OD, T, 16><<<c.nE, 256, lw, st>>>(nc);                                                     \
        } else if (c.q.nPx <= 8) {                                                                                 \
            if (lw > 64 * 1024) (void)hipFuncSetAttribute((const void*)k_cutw<KER, OD, T, 8>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lw); \
            k_cutw<KER, OD, T, 8><<<c.nE, 256, lw, st>>>(nc);                                                      \
        } else {                                                                                                   \
            if (lw > 64 * 1024) (void)hipFuncSetAttribute((const void*)k_cutw<KER, OD, T, 0>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lw); \
            k_cutw<KER, OD, T, 0><<<c.nE, 256, lw, st>>>(nc);                                                      \
        }                                                                                                          \
    } while (0)
#define LAUNCH_CUTW(KER, OD)                                       \
    do {                                                          \
        if (c.precision == 1) LAUNCH_CUTW_T(KER, OD, float);      \
        else LAUNCH_CUTW_T(KER, OD, double);                      \
    } while (0)
            P.cutw_used = true;
            const int nmc = c.kernel == KER_MICRO_FIELD ? 3 : P.oo;
            dreserve(P, P.d_mir, (size_t)P.nnz_nodes * nmc);
            HIPCHECK(hipMemsetAsync(P.d_mir, 0, sizeof(double) * (size_t)P.nnz_nodes * nmc, st));
            nc.mir = P.d_mir;
            switch (c.kernel) {
                case KER_CONSTANT: LAUNCH_CUTW(KER_CONSTANT, 1); break;
                case KER_PARABOLA: LAUNCH_CUTW(KER_PARABOLA, 1); break;
                case KER_FRACTIONAL: LAUNCH_CUTW(KER_FRACTIONAL, 1); break;
                case KER_MICRO: LAUNCH_CUTW(KER_MICRO, 1); break;
                default: LAUNCH_CUTW(KER_MICRO_FIELD, 2); break;
            }
            if (c.kernel == KER_MICRO_FIELD)
                k_mirror<2, true><<<nblk(P.nnz_nodes, 256), 256, 0, st>>>(P.nnz_nodes, P.d_wrowOf, P.d_tpos, P.d_vlabel,
                                                                          c.is_DG, P.d_mir, P.d_val);
            else
                k_mirror<1, false><<<nblk(P.nnz_nodes, 256), 256, 0, st>>>(P.nnz_nodes, P.d_wrowOf, P.d_tpos,
                                                                           P.d_vlabel, c.is_DG, P.d_mir, P.d_val);
#undef LAUNCH_CUTW
#undef LAUNCH_CUTW_T
        } else {
            switch (c.kernel) {
                case KER_CONSTANT: LAUNCH_CUT(KER_CONSTANT, 1); break;
                case KER_PARABOLA: LAUNCH_CUT(KER_PARABOLA, 1); break;
                case KER_FRACTIONAL: LAUNCH_CUT(KER_FRACTIONAL, 1); break;
                case KER_MICRO: LAUNCH_CUT(KER_MICRO, 1); break;
                default: LAUNCH_CUT(KER_MICRO_FIELD, 2); break;
            }
        }
#undef LAUNCH_CUT
#undef LAUNCH_CUT_T
#undef LAUNCH_CUT_I
#undef LAUNCH_CUT8
#undef LAUNCH_CUT8_K
    }
    HIPCHECK(hipGetLastError());
    ev_end(P, PH_CUT);
    ev_begin(P, PH_GATHER);
    if (gather)
        k_lp_gather<<<(c.nE + 3) / 4, 256, 0, st>>>(c.nE, c.is_DG, P.d_tri, P.d_vlabel, P.d_tptr, P.d_tcB, P.d_tcut,
                                                     P.d_stage, P.d_posTT, P.d_val);
    HIPCHECK(hipGetLastError());
    ev_end(P, PH_GATHER);
#undef DISPATCH_PAIRS
#undef LAUNCH_PAIRS
#undef DISPATCH_PAIRS_Z
#undef LAUNCH_PAIRS_Z

    // ---------------------------------------------------------------- compaction
    ev_begin(P, PH_COMPACT);
    const long long K = c.K;
    int*& d_ocnt = P.d_wocnt;
    dreserve(P, d_ocnt, (size_t)K + 1);
    k_compact<false><<<P.nNodes, 64, 0, st>>>(P.nNodes, c.outdim, P.d_rowptr, P.d_col, P.d_val, d_ocnt, nullptr,
                                              nullptr, nullptr);
    dreserve(P, P.d_orow, (size_t)K + 1);
    scan_exclusive<int>(d_ocnt, P.d_orow, K, scan_scratch(P), st);
    HIPCHECK(hipMemcpyAsync(&P.nnz, P.d_orow + K, sizeof(long long), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    dreserve(P, P.d_ocol, (size_t)P.nnz);
    dreserve(P, P.d_oval, (size_t)P.nnz);
    k_compact<true><<<P.nNodes, 64, 0, st>>>(P.nNodes, c.outdim, P.d_rowptr, P.d_col, P.d_val, nullptr, P.d_orow,
                                             P.d_ocol, P.d_oval);
    HIPCHECK(hipGetLastError());
    ev_end(P, PH_COMPACT);
    read_counters(P, h);
    if (h[CNT_ERROR]) throw std::runtime_error("assembly consistency error (code " + std::to_string(h[CNT_ERROR]) + ")");
    P.stats.nnz_nodes = P.nnz_nodes;
    P.stats.nnz = P.nnz;
    P.stats.subtri = h[CNT_SUBTRI];
    P.stats.sur_acc = h[CNT_SUR_ACC];
    P.stats.sur_rej = h[CNT_SUR_REJ];
    P.stats.sur_items = h[CNT_SUR_ITEMS];
    if (P.timing) {
        for (int ph = PH_DUAL; ph < PH_COUNT; ph++) {
            float ms = 0;
            if (ph == PH_CLOSE && nclose == 0) { P.stats.ms[ph] = 0; continue; }
            HIPCHECK(hipEventElapsedTime(&ms, P.ev[2 * ph], P.ev[2 * ph + 1]));
            P.stats.ms[ph] = ms;
        }
    }
    P.assembled = true;
}

void plan_set_surrogate(Plan& P, int nnet, const float* params, const float* norm, float tol, double smin,
                        double smax) {
    if (nnet < 0 || nnet > 4 || (nnet > 2 && P.cfg.kernel != KER_MICRO_FIELD) || (nnet == 3))
        throw std::invalid_argument("the surrogate has one or two networks (field kernel: two or four)");
    if (nnet > 0 && (!params || !norm)) throw std::invalid_argument("surrogate parameters missing");
    if (nnet > 0 && !(tol >= 0)) throw std::invalid_argument("surrogate tolerance must be >= 0");
    if (nnet > 0 && !(smin >= 0 && smax > smin)) throw std::invalid_argument("invalid surrogate domain");
    if (!P.sur.d_wf) {
        HIPCHECK(hipMalloc((void**)&P.sur.d_wf, 4 * mlp::NFRAG_FWD * 256 * sizeof(uint16_t)));
        HIPCHECK(hipMalloc((void**)&P.sur.d_bias, 4 * mlp::BIAS_PAD * sizeof(float)));
        HIPCHECK(hipMalloc((void**)&P.sur.d_norm, 4 * mlp::NNORM * sizeof(float)));
    }
    for (int n = 0; n < nnet; n++)
        mlp::upload(params + (size_t)n * mlp::NPARAM, norm + (size_t)n * mlp::NNORM,
                    P.sur.d_wf + (size_t)n * mlp::NFRAG_FWD * 256, P.sur.d_bias + n * mlp::BIAS_PAD,
                    P.sur.d_norm + n * mlp::NNORM, P.stream);
    HIPCHECK(hipStreamSynchronize(P.stream));
    P.sur.nnet = nnet;
    P.sur.tol = tol;
    P.sur.smin = smin;
    P.sur.smax = smax;
}

// Training samples of the surrogate from the exact FP64 cut-pair sums of the assembled plan:
// partial items whose hash(aT, bT, k, salt) is a multiple of stride.  Returns the number of
// qualifying items (those beyond maxn are not stored).
long long plan_surrogate_samples(Plan& P, long long maxn, unsigned stride, unsigned salt, double smin, double smax,
                                 float* feat, float* targ) {
    const PlanConfig& c = P.cfg;
    if (!P.assembled) throw std::invalid_argument("assemble the plan before drawing surrogate samples");
    const bool field = c.kernel == KER_MICRO_FIELD && c.outdim == 2 && !c.is_DG;
    if (!((c.kernel == KER_CONSTANT && c.outdim == 1) || field) || c.method_remote != M_RETRIANGULATE ||
        P.stats.bary > 0)
        throw std::invalid_argument(
            "surrogate samples need the constant or the microelastic field kernel (CG) with the retriangulate method");
    const int NT = field ? 18 : 9;
    if (c.q.nPx > CUT_ITEMS || c.q.nPx * CUT_MAXPAIRS < CUT_ITEMS)
        throw std::invalid_argument("outer rule size not supported by the cut kernel");
    if (stride == 0 || maxn < 0 || (maxn > 0 && (!feat || !targ)) || !(smax > smin))
        throw std::invalid_argument("invalid sample arguments");
    const hipStream_t st = P.stream;
    float *dF = nullptr, *dT = nullptr;
    HIPCHECK(hipMalloc((void**)&dF, (size_t)std::max<long long>(maxn, 1) * 12 * sizeof(float)));
    HIPCHECK(hipMalloc((void**)&dT, (size_t)std::max<long long>(maxn, 1) * NT * sizeof(float)));
    NumArgs nc = base_numargs(P);
    nc.HN = (int)next_pow2((long long)(1.3 * std::max(P.maxRow, 1)));
    nc.sur.sfeat = dF;
    nc.sur.starg = dT;
    nc.sur.smaxn = maxn;
    nc.sur.stride = stride;
    nc.sur.salt = salt;
    nc.sur.smin = smin;
    nc.sur.smax = smax;
    set_surrogate_args(P, nc.sur);
    nc.sur.sfeat = dF; nc.sur.starg = dT; nc.sur.smaxn = maxn; nc.sur.stride = stride; nc.sur.salt = salt;
    nc.sur.smin = smin; nc.sur.smax = smax;
    nc.sur.ntarg = NT;
    const size_t lb = cut_lds_bytes(c.q.nPx, P.oo, nc.HN);
    if (lb > 160 * 1024 || P.maxRow > 65535) throw std::length_error("cut-pair LDS exceeds 160 KiB");
    unsigned long long h[NCNT];
    try {
        HIPCHECK(hipMemsetAsync(P.d_counters + CNT_SAMPLES, 0, sizeof(unsigned long long), st));
        if (field) {
            if (lb > 64 * 1024) (void)hipFuncSetAttribute((const void*)k_cut<KER_MICRO_FIELD, 2, double, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lb);
            k_cut<KER_MICRO_FIELD, 2, double, 1><<<c.nE, 256, lb, st>>>(nc);
        } else {
            if (lb > 64 * 1024) (void)hipFuncSetAttribute((const void*)k_cut<KER_CONSTANT, 1, double, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lb);
            k_cut<KER_CONSTANT, 1, double, 1><<<c.nE, 256, lb, st>>>(nc);
        }
        HIPCHECK(hipGetLastError());
        read_counters(P, h);
        const long long n = (long long)h[CNT_SAMPLES];
        const long long m = std::min(n, maxn);
        if (m > 0) {
            HIPCHECK(hipMemcpy(feat, dF, (size_t)m * 12 * sizeof(float), hipMemcpyDeviceToHost));
            HIPCHECK(hipMemcpy(targ, dT, (size_t)m * NT * sizeof(float), hipMemcpyDeviceToHost));
        }
        (void)hipFree(dF);
        (void)hipFree(dT);
        return n;
    } catch (...) {
        (void)hipFree(dF);
        (void)hipFree(dT);
        throw;
    }
}

// Domain-decomposition indicator (mesh["ZetaIndicator"], cython/nlfem.pyx:205-214): CSR rows of
// the subdomains of each element; rows are sorted here (the reference's set_intersection
// expects sorted rows).  nZeta == 0 removes it.
void plan_set_zeta(Plan& P, const long* indices, const long* indptr, long nZeta) {
    const int nE = P.cfg.nE;
    if (P.d_zptr) { (void)hipFree(P.d_zptr); P.d_zptr = nullptr; }
    if (P.d_zidx) { (void)hipFree(P.d_zidx); P.d_zidx = nullptr; }
    P.nZeta = 0;
    if (nZeta <= 0) return;
    if (!indices || !indptr) throw std::invalid_argument("ZetaIndicator arrays missing");
    if (indptr[0] != 0) throw std::invalid_argument("ZetaIndicator indptr must start at 0");
    std::vector<long long> zp((size_t)nE + 1);
    for (int e = 0; e <= nE; e++) {
        zp[e] = indptr[e];
        if (e > 0 && zp[e] < zp[e - 1]) throw std::invalid_argument("ZetaIndicator indptr must be non-decreasing");
    }
    if (zp[nE] >= (1ll << 31)) throw std::invalid_argument("ZetaIndicator has too many entries");
    std::vector<int> zi((size_t)std::max<long long>(zp[nE], 1));
    for (long long i = 0; i < zp[nE]; i++) {
        if (indices[i] < 0 || indices[i] >= nZeta) throw std::invalid_argument("ZetaIndicator column out of range");
        zi[i] = (int)indices[i];
    }
    for (int e = 0; e < nE; e++) std::sort(zi.begin() + zp[e], zi.begin() + zp[e + 1]);
    dalloc(P.d_zptr, (size_t)nE + 1);
    dalloc(P.d_zidx, zi.size());
    HIPCHECK(hipMemcpyAsync(P.d_zptr, zp.data(), sizeof(long long) * zp.size(), hipMemcpyHostToDevice, P.stream));
    HIPCHECK(hipMemcpyAsync(P.d_zidx, zi.data(), sizeof(int) * zi.size(), hipMemcpyHostToDevice, P.stream));
    HIPCHECK(hipStreamSynchronize(P.stream));
    P.nZeta = nZeta;
}

void plan_free(Plan& P) {
    void* sp[] = {P.sur.d_wf, P.sur.d_bias, P.sur.d_norm, P.d_rlist, P.d_roff, P.d_rcnt};
    for (void* p : sp)
        if (p) (void)hipFree(p);
    void* ptrs[] = {P.d_tri, P.d_ord, P.d_mir, P.d_vlabel, P.d_elem, P.d_vptr, P.d_vel, P.d_xadj, P.d_adj, P.d_nbr, P.d_cntF,
                    P.d_cntB, P.d_closeList, P.d_counters, P.d_tptr, P.d_tadj, P.d_rowptr, P.d_col, P.d_tpos,
                    P.d_posTT, P.d_val, P.d_tg, P.d_dg, P.d_orow, P.d_ocol, P.d_oval, P.d_scan_tmp,
                    P.d_poff, P.d_P, P.d_adjp, P.d_coff, P.d_tcB, P.d_tcut, P.d_stage, P.d_tF, P.d_zptr,
                    P.d_zidx, P.d_gtab};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    void* wp[] = {P.d_wcnt, P.d_wrowcnt, P.d_wptmp, P.d_wrowOf, P.d_wplen, P.d_wocnt};
    for (void* p : wp)
        if (p) (void)hipFree(p);
    for (int i = 0; i < 2 * PH_COUNT; i++)
        if (P.ev[i]) (void)hipEventDestroy(P.ev[i]);
    if (P.stream) (void)hipStreamDestroy(P.stream);
}

}  // namespace nlfem
