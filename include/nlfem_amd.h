/* nlfem_amd.h -- C ABI of the MI355X nonlocal stiffness assembly (libnlfem_amd.so).
 *
 * Plain pointers and sizes only.  Every function returns 0 on success and a nonzero status
 * on failure; nlfem_last_error() then describes the failure (thread-local).  Host pointers
 * are borrowed for the duration of the call, read-only unless documented as output.
 * The library allocates its device memory itself (hipMalloc) and runs on the current HIP
 * device; every plan owns one non-blocking stream with default priority.
 *
 * Reference interfaces (cagau/nlfem) replaced by each entry point are cited per function.
 */
#ifndef NLFEM_AMD_H
#define NLFEM_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes */
#define NLFEM_OK 0
#define NLFEM_EINVAL 1   /* invalid configuration; the reference would abort() */
#define NLFEM_EDEVICE 2  /* HIP runtime / device failure */
#define NLFEM_ECAPACITY 3 /* a capacity estimate could not be met */
#define NLFEM_EIO 4      /* result file could not be written */

const char* nlfem_last_error(void);

/* Drop-in for par_assemble (include/Cassemble.h:125-140; Cython binding
 * include/Cassemble.pxd:12-25).  compute = "system" writes the stiffness matrix A^T in
 * Armadillo ARMA_SPM_BIN_FN008 format to path_spAd (the reference's output contract,
 * src/Cassemble.cpp:850-851); "forcing" writes the load vector as ARMA_MAT_BIN_FN008 to
 * path_fd (src/Cassemble.cpp:927).  std::string arguments become NUL-terminated strings.
 * K_Omega, nE_Omega (checked), nV_Omega, is_NeumannBoundary, str_model_f (system),
 * maxDiameter and verbose keep the reference's (lack of) effect on values.  Only dim == 2 is
 * supported (NLFEM_EINVAL otherwise).  nZeta > 0 applies the domain-decomposition weights
 * 1 / |Z_aT n Z_bT| of the CSR indicator (ptrZetaIndicator_indices, _indptr; nE rows, nZeta
 * columns) to every pair, and 1 / |Z_aT| to every load contribution (src/mathhelpers.cpp:19-38,
 * src/Cassemble.cpp:648-656,906-914). */
int nlfem_par_assemble(const char* compute, const char* path_spAd, const char* path_fd, int K_Omega, int K,
                       const long* ptrTriangles, const long* ptrLabelTriangles, const double* ptrVerts,
                       const long* ptrLabelVerts, int nE, int nE_Omega, int nV, int nV_Omega, const double* Px,
                       int nPx, const double* dx, const double* Py, int nPy, const double* dy, double sqdelta,
                       int is_DiscontinuousGalerkin, int is_NeumannBoundary, const char* str_model_kernel,
                       const char* str_model_f, const char* str_integration_method_remote,
                       const char* str_integration_method_close, int is_PlacePointOnCap, int dim, int outdim,
                       const long* ptrZetaIndicator_indices, const long* ptrZetaIndicator_indptr, long nZeta,
                       const double* Pg, int degree, const double* dg, double maxDiameter, double fractional_s,
                       int is_fullConnectedComponentSearch, int verbose);

/* Drop-in for par_evaluateMass (include/Cassemble.h:202-204): vd += M ud over the first J
 * elements (the reference's Python caller passes J = nE_Omega, cython/nlfem.pyx:674). */
int nlfem_par_evaluateMass(double* vd, const double* ud, const long* Elements, const long* ElementLabels,
                           const double* Verts, const long* VertexLabels, int K_Omega, int J, int nP, const double* P,
                           const double* dx, int dim, int outdim, int is_DiscontinuousGalerkin, int nE, int nV);

/* Debug helpers, as exported to Cython (include/Cassemble.pxd:30-38; include/Cassemble.h:62-71).
 * They run the device implementation used by the assembly kernels. */
int nlfem_method_retriangulate(const double* xCenter, const double* TE, double sqdelta, double* reTriangleList,
                               int isPlacePointOnCap); /* returns #triangles (>= 0) or -status */
int nlfem_toRef(const double* E, const double* phys_x, double* ref_p);
int nlfem_toPhys(const double* E, const double* p, int dim, double* out_x);
int nlfem_solve2x2(const double* A, const double* b, double* x);

/* Batched variant of method_retriangulate: n centres x[2n], triangles TE[6n]; lists[54n]
 * (9 triangles x 3 points x 2 coordinates each), ntri[n].  No counterpart in the reference;
 * used to check the device predicate against CPU results in bulk. */
int nlfem_retriangulate_batch(int n, const double* x, const double* TE, double sqdelta, int isPlacePointOnCap,
                              double* lists, int* ntri);

/* ------------------------------------------------------------------------------------
 * Device-resident plan API (no counterpart in the reference, whose result crosses back
 * through a file).  create uploads the mesh and configuration; assemble runs the whole
 * pipeline on the GPU (dual graph, BFS neighbour lists, integer sparsity pattern, pair
 * quadrature, compaction); the CSR of A stays in HBM until copied out.  precision: 0 = FP64,
 * 1 = FP32 quadrature sums with FP64 accumulation and FP64 geometry.
 * ------------------------------------------------------------------------------------ */
typedef struct nlfem_plan nlfem_plan;

int nlfem_plan_create(nlfem_plan** out, int K, const long* ptrTriangles, const long* ptrLabelTriangles,
                      const double* ptrVerts, const long* ptrLabelVerts, int nE, int nV, const double* Px, int nPx,
                      const double* dx, const double* Py, int nPy, const double* dy, double sqdelta,
                      int is_DiscontinuousGalerkin, const char* str_model_kernel,
                      const char* str_integration_method_remote, const char* str_integration_method_close,
                      int is_PlacePointOnCap, int outdim, const double* Pg, int degree, const double* dg,
                      double fractional_s, int is_fullConnectedComponentSearch, int precision);
int nlfem_plan_assemble(nlfem_plan* plan);
/* K (matrix order) and nnz of the assembled CSR of A (rows sorted, exact zeros dropped). */
int nlfem_plan_csr_info(const nlfem_plan* plan, long long* K, long long* nnz);
/* Copies rowptr[K+1], col[nnz], val[nnz] to host memory.  Pageable destinations are filled
 * through two pinned staging buffers (32 MB each, allocated on first use and kept for the
 * process), with the DMA of one chunk overlapping the host copy of the previous one; pinned
 * destinations are copied directly.  Calls are serialised by a lock. */
int nlfem_plan_csr_copy(const nlfem_plan* plan, long long* rowptr, int* col, double* val);
/* Device pointers of the CSR (valid until the next assemble or destroy). */
int nlfem_plan_csr_device(const nlfem_plan* plan, long long** rowptr, int** col, double** val);
/* Neighbour lists: total number of (aT, bT) pairs, then ptr[nE+1], nbr[total], cls[total]
 * (0 cut, 1 interior, 2 close, 3 baryCenter).  Order within a list is unspecified. */
int nlfem_plan_neighbours_total(const nlfem_plan* plan, long long* total);
int nlfem_plan_neighbours(const nlfem_plan* plan, long long* ptr, int* nbr, unsigned char* cls);
/* Statistics: out[0..] = pairs, probes, close, interior, cut, bary, nnz_nodes, nnz, cap,
 * max_list, subtri (fan triangles integrated by the cut kernel), then per-phase milliseconds
 * (dual, bfs, transpose, pattern, positions, close, interior, cut, compact, gather: the
 * per-inner-element sum of the cut pairs' staged termLocalPrime, which runs between cut and
 * compact).  n = number of doubles available in out; returns NLFEM_OK. */
int nlfem_plan_stats(const nlfem_plan* plan, double* out, int n);
/* Cut pairs through the surrogate network (no counterpart in the reference; retriangulate).
 * Constant kernel: nnet = 1 or 2 networks (params[nnet][9737], norm[nnet][42], see the
 * surrogate section below), the ensemble; a pair goes to the exact method when one of its
 * items lies outside [smin, smax] (sqrt|det bT| / delta) or the two networks disagree by more
 * than tol on an output ratio.  Microelastic field kernel (outdim 2, CG): nnet = 2, network 0
 * and 1 give the first and the last nine of the 18 item outputs; the outputs are corrected to
 * satisfy the rotation constraints exactly, and a pair goes to the exact method when an item
 * lies outside [smin, smax] or its correction exceeds tol relative to its outputs.
 * nnet = 0 switches the surrogate off.  Stats gain (after the phase times) the pairs taken by
 * the surrogate, the pairs handed back, and the network items evaluated. */
/* Override the mesh-size estimates (longest edge, smallest and mean element area) from which
 * the BFS and pattern table capacities are derived; for meshes made of several disconnected
 * parts of different resolution, pass the values of the part with the most elements per
 * horizon.  Capacities still grow on overflow. */
int nlfem_plan_set_estimates(nlfem_plan* plan, double hmax, double amin, double amean);
/* Force the general data structures that the plan otherwise uses only when the element
 * numbering is not local or a fixed-size layout overflows (no counterpart in the reference;
 * for checking those paths and for meshes whose numbering is known to be poor).  Bits:
 * NLFEM_PATH_BFS_HASH: hashed BFS visited sets; NLFEM_PATH_PATTERN_HASH: hashed pattern
 * element sets; NLFEM_PATH_TRANSPOSE_CSR: transposed relation counted into exact CSR rows;
 * NLFEM_PATH_BFS_WAVE: one outer element per wavefront in the bitmap BFS (by default two
 * outer elements share a wavefront, 32 lanes each); NLFEM_PATH_TRANSPOSE_UNSTAGED: cut
 * entries of the transposed rows placed one at a time with global atomics (otherwise taken by
 * blocks whose cut entries exceed the LDS staging array). */
#define NLFEM_PATH_BFS_HASH 1
#define NLFEM_PATH_PATTERN_HASH 2
#define NLFEM_PATH_TRANSPOSE_CSR 4
#define NLFEM_PATH_BFS_WAVE 8
#define NLFEM_PATH_TRANSPOSE_UNSTAGED 16
int nlfem_plan_force_paths(nlfem_plan* plan, int flags);
/* Cut kernel for every kernel but the constant one and for outdim 2 and DG: 0 (default) the
 * per-lane kernel (one item per lane, each wavefront on its own pairs, no chunk barriers;
 * outer rules of up to 16 points); otherwise the chunked kernel with 64, 128 or 256 items
 * (pair x outer point) per chunk, i.e. 4, 8 or 16 pairs per 256-thread block with the 16-point
 * rule (larger chunks keep more threads busy per chunk and need more LDS per block), except
 * that scalar CG kernels with at most 8 outer points take the 8-lane kernel for any nonzero
 * value.  No counterpart in the reference.  A value that the outer rule or the LDS does not
 * allow falls back to 128 at launch. */
int nlfem_plan_set_cut_items(nlfem_plan* plan, int items);
/* Autotune of the cut kernel: assembles once with 128 items per chunk and, when the chunked
 * kernel ran, once with each of 64, 256 and 0 (the per-lane kernel) that fits (when the 8-lane
 * kernel ran instead, once with 0 only); keeps the value with the shortest cut phase.  *chosen (may be NULL) gets it, ms[4] (may be NULL) the
 * cut-phase times of 64, 128, 256 and 0 (0 where not tried).  The plan holds a complete
 * assembly afterwards (values do not depend on the choice beyond summation order).  No
 * counterpart in the reference. */
int nlfem_plan_autotune(nlfem_plan* plan, int* chosen, double* ms);
int nlfem_plan_set_surrogate(nlfem_plan* plan, int nnet, const float* params, const float* norm, float tol,
                             double smin, double smax);
/* Training samples from the exact FP64 cut-pair sums of an assembled plan: feat[maxn][12],
 * targ[maxn][9] (constant kernel: ratios to the full-triangle sums) or targ[maxn][18]
 * (microelastic field kernel: the six symmetric tensors sum psi_a psi_b k w of the item in its
 * frame, in units of c |det bT| / delta); items kept when a hash of (aT, bT, k, salt) is a
 * multiple of stride.  *count = qualifying items (may exceed maxn). */
int nlfem_plan_surrogate_samples(nlfem_plan* plan, long long maxn, unsigned stride, unsigned salt, double smin,
                                 double smax, float* feat, float* targ, long long* count);
void nlfem_plan_destroy(nlfem_plan* plan);

/* Load vector on the device (par_forcing, src/Cassemble.cpp:857-928): fd[K].  nZeta > 0 weighs
 * element aT's contribution by 1 / |Z_aT| (src/Cassemble.cpp:906-914); nZeta = 0 ignores the
 * indicator arrays (may be NULL). */
int nlfem_load_vector(int K, const long* ptrTriangles, const long* ptrLabelTriangles, const double* ptrVerts,
                      const long* ptrLabelVerts, int nE, int nV, const double* Px, int nPx, const double* dx,
                      int is_DiscontinuousGalerkin, const char* str_model_f, int outdim, double* fd,
                      const long* ptrZetaIndicator_indices, const long* ptrZetaIndicator_indptr, long nZeta);
/* Domain-decomposition weights of a plan (the nZeta > 0 case of par_assemble,
 * src/mathhelpers.cpp:19-38 applied at src/Cassemble.cpp:648-656,741-766): indicator rows
 * indices[indptr[e] .. indptr[e+1]) of the nE elements, columns < nZeta; every pair's four
 * blocks are scaled by 1 / |Z_aT n Z_bT| (an empty intersection gives inf, as in the reference).
 * nZeta = 0 removes them.  Weighted plans integrate every pair with the per-pair kernels
 * (slower than the default path) and do not use the surrogate. */
int nlfem_plan_set_zeta(nlfem_plan* plan, const long* indices, const long* indptr, long nZeta);

/* ------------------------------------------------------------------------------------
 * Cut-pair surrogate network (no counterpart in the reference): a 12 -> 64 -> 64 -> 64 -> 9
 * perceptron with SiLU activations evaluated with BF16 MFMA and FP32 accumulation.
 * params[9737] = W1[64][12] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[9][64] b4[9];
 * norm[42] = input mean[12], input 1/std[12], output mean[9], output std[9].
 * X[n][12] inputs, T[n][9] targets, Y[n][9] outputs, all host memory.
 * ------------------------------------------------------------------------------------ */
int nlfem_mlp_info(int* nin, int* nhidden, int* nout, int* nparam, int* nnorm);
int nlfem_mlp_forward(const float* params, const float* norm, const float* X, long long n, float* Y);
/* Allocation-free variant of nlfem_mlp_forward.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  All pointers are DEVICE memory owned by the caller:
 * d_params[9737], d_norm[42], d_X[n][12], d_Y[n][9]; stream is a hipStream_t (NULL: the
 * default stream); the call only enqueues work (no synchronisation).  It calls no hipMalloc,
 * hipFree or other allocating routine: the packed BF16 network lives in a fixed device array
 * of the library, 39,744 bytes per device, held until the process ends.  n must lie in
 * [0, 2^31 - 1] (NLFEM_EINVAL otherwise).  Every call on a device repacks that device's one
 * copy, so calls that overlap on one device (different streams or threads) race: serialise
 * them.  Results equal nlfem_mlp_forward's bit for bit (same kernels). */
int nlfem_mlp_forward_noalloc(const float* d_params, const float* d_norm, const float* d_X, long long n, float* d_Y,
                              void* stream);
/* Gradient of the mean squared error in normalised output units over all n samples. */
int nlfem_mlp_grad(const float* params, const float* norm, const float* X, const float* T, long long n,
                   float* grad, float* loss);
/* Adam (warm-up + cosine decay), minibatches of `batch` (multiple of 64) per epoch;
 * params is updated in place; loss_hist[epochs] (may be NULL) gets the mean minibatch loss. */
int nlfem_mlp_train(float* params, const float* norm, const float* X, const float* T, long long n, int epochs,
                    int batch, float lr_max, unsigned seed, float* loss_hist);

#ifdef __cplusplus
}
#endif

#endif /* NLFEM_AMD_H */
