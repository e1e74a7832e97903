// Host-side plan of one stiffness assembly on one GPU: configuration, device buffers,
// phase timings.  The plan owns everything; all entry points take it by argument, so the
// library is reentrant per plan (the reference keeps global function pointers).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <unordered_map>
#include <vector>

#include "nlfem_device.hpp"

namespace nlfem {

constexpr int MAXQ = 32;  // max points of the outer / inner triangle rules

struct QuadArgs {
    int nPx, nPy;
    double Px[2 * MAXQ], dx[MAXQ], Py[2 * MAXQ], dy[MAXQ];
    double psix[3 * MAXQ], psiy[3 * MAXQ];
};

struct PlanConfig {
    int K = 0, nE = 0, nV = 0, outdim = 1, is_DG = 0;
    int kernel = KER_CONSTANT, method_remote = M_RETRIANGULATE, method_close = M_RETRIANGULATE;
    int caps = 1, singular = 0, fullCC = 0;
    double sqdelta = 0, fractional_s = -1;
    int degree = 0, nPg = 0;
    int precision = 0;  // 0: FP64 throughout; 1: mixed (FP32 quadrature sums, FP64 accumulation)
    QuadArgs q;
    std::vector<double> Pg, dg;
};

// Phase timings in milliseconds (hipEvent pairs on the plan's stream).
enum Phase : int {
    PH_SETUP = 0, PH_DUAL, PH_BFS, PH_TRANSPOSE, PH_PATTERN, PH_POSITIONS, PH_CLOSE, PH_INTERIOR, PH_CUT,
    PH_COMPACT, PH_GATHER, PH_COUNT
};

struct Stats {
    unsigned long long pairs = 0, probes = 0, close = 0, interior = 0, cut = 0, bary = 0, subtri = 0;
    unsigned long long sur_acc = 0, sur_rej = 0, sur_items = 0;  // cut pairs by the surrogate / handed back; network items
    long long nnz_nodes = 0, nnz = 0;
    int cap = 0, max_list = 0;
    float ms[PH_COUNT] = {0};
};

// Geometric size estimates for the LDS capacities (grown and retried on overflow).
struct Estimates {
    double hmax = 0, amin = 1e300, amean = 1.0;
    // element area for counting elements in a disc: the mean, unless the mesh is nearly
    // uniform below it (capacities are grown and the kernels rerun on overflow)
    double aref() const { return amin > 0.6 * amean ? amin : 0.6 * amean; }
};

// Surrogate networks for cut pairs (constant kernel): packed BF16 fragments, biases and
// normalisation per network, acceptance threshold and domain.
struct SurrogateModel {
    int nnet = 0;
    float tol = 0;
    double smin = 0, smax = 0;
    uint16_t* d_wf = nullptr;
    float* d_bias = nullptr;
    float* d_norm = nullptr;
};

struct Plan {
    PlanConfig cfg;
    SurrogateModel sur;
    int* d_rlist = nullptr;       // cut pairs handed back to the exact method
    long long* d_roff = nullptr;  // [nE+1]
    int* d_rcnt = nullptr;        // [nE]
    long long rlist_cap = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[2 * PH_COUNT] = {};
    bool timing = true;
    int oo = 1;
    int nNodes = 0;  // CG: nV, DG: 3 nE

    // mesh
    int* d_tri = nullptr;        // [nE*3] vertex ids
    int* d_ord = nullptr;        // [nE] outer elements grouped by spatial tiles (tile_order)
    double* d_mir = nullptr;     // [nnz_nodes][NC] mirror sums of the per-lane cut kernel
    int* d_vlabel = nullptr;     // [nV]
    ElemRec* d_elem = nullptr;   // [nE]
    // vertex -> element incidence
    long long* d_vptr = nullptr; // [nV+1]
    int* d_vel = nullptr;
    // dual graph
    long long* d_xadj = nullptr; // [nE+1]
    int* d_adj = nullptr;
    int* d_adjp = nullptr;       // dual graph rows padded to a multiple of 4 (BFS)
    // neighbour lists (padded: cap slots per outer element; front = interior/close, back = cut/bary)
    int cap = 0;
    uint32_t* d_nbr = nullptr;
    int* d_cntF = nullptr;
    int* d_cntB = nullptr;
    int2* d_closeList = nullptr;
    unsigned long long* d_counters = nullptr;  // [16]: stats + close count + flags
    // transpose of the neighbour relation
    long long* d_tptr = nullptr;
    int* d_tadj = nullptr;
    long long* d_coff = nullptr;  // [nE+1] exclusive scan of the cut-pair counts
    int* d_tcB = nullptr;         // [nE] cut entries at the end of each transposed row
    int* d_tF = nullptr;          // [nE] interior and close entries at the start of each row
    bool tpad_off = false;        // a padded transposed row overflowed once: count first
    bool tunstaged = false;       // cut entries of the transposed rows placed one at a time
    int* d_tcut = nullptr;        // [ntadj] cut entries: index of the pair in the cut enumeration
    double* d_stage = nullptr;    // [ncut*6] termLocalPrime per cut pair
    long long stage_cap = 0;
    // node CSR (pattern of the full symmetric matrix in node space, od x od blocks per entry)
    long long* d_rowptr = nullptr;  // [nNodes+1]
    int* d_col = nullptr;           // [nnz_nodes]
    int* d_tpos = nullptr;          // [nnz_nodes] position of the transposed entry
    int* d_posTT = nullptr;         // [nE*9] positions of (node(T,i), node(T,j))
    double* d_val = nullptr;        // [nnz_nodes*oo]
    long long* d_poff = nullptr;    // [3 nE + 1] offsets of the partial rows of (element, local node)
    double* d_P = nullptr;          // partial rows of interior pairs (owner computes)
    long long ptotal = 0;           // allocated doubles in d_P
    long long nnz_nodes = 0;
    int maxRow = 0;
    // Duffy tables [case][traffo][k]: alpha (4), alphaCanceled (4), det (1)
    double* d_tg = nullptr;
    double* d_dg = nullptr;
    // final scalar CSR
    long long* d_orow = nullptr;  // [K+1]
    int* d_ocol = nullptr;
    double* d_oval = nullptr;
    long long nnz = 0;
    // domain-decomposition indicator (nZeta > 0): CSR rows of subdomains per element, sorted
    long long* d_zptr = nullptr;  // [nE+1], null: every pair has weight 1
    int* d_zidx = nullptr;
    long nZeta = 0;
    // scratch
    void* d_scan_tmp = nullptr;
    size_t scan_tmp_bytes = 0;
    int* d_wcnt = nullptr;     // per-element / per-vertex counters
    int* d_wrowcnt = nullptr;  // pattern row lengths
    int* d_wptmp = nullptr;    // unsorted pattern rows
    int* d_wrowOf = nullptr;   // row of each node-CSR entry
    int* d_wplen = nullptr;    // partial-row lengths
    int* d_wocnt = nullptr;    // output row lengths
    int* d_gtab = nullptr;     // pattern hash tables in global memory (rows too large for LDS)
    // allocated capacity (elements) of each device buffer above, keyed by the member's address:
    // repeated assemblies reuse buffers that are large enough
    std::unordered_map<const void*, size_t> dcap;

    long long bfs_visit = 0, bfs_hvis = 0, bfs_fvis = 0;  // BFS capacities that fit (reused)
    long long bfs_maxoff = 0;  // bound of |visited element - outer element| from the first BFS (0: not yet measured, -1: none)
    int cut_items = 0;        // cut kernel of the k_cut cases: 0 = k_cutw (one item per lane), or k_cut with 64, 128, 256 items per chunk
    int cut_items_used = 0;   // items per chunk k_cut ran with in the last assembly (0: k_cut not used)
    bool cutw_used = false;   // the last assembly's cut pairs went through k_cutw (cut_items = 0)
    bool cut8_used = false;   // the last assembly's cut pairs went through k_cut8
    bool bfs_nobm = false;  // the bitmap visited set overflowed its window once: use the hash
    bool bfs_wave = false;  // one outer element per wavefront in the bitmap BFS (k_bfs instead of k_bfs_h)
    bool pat_nobm = false;  // the same for the pattern's element set

    Stats stats;
    bool assembled = false;
};

}  // namespace nlfem
