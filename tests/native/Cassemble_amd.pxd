# The reference's Cassemble.pxd (include/Cassemble.pxd:11-40) pointed at the C ABI of
# libnlfem_amd.so, as a maintainer would bind it; argument list of par_assemble unchanged
# (include/Cassemble.h:125-140), strings as const char*, a status instead of abort().
cdef extern from "nlfem_amd.h":
    int nlfem_par_assemble(const char* compute, const char* path_spAd, const char* path_fd,
                           int K_Omega, int K, const long* ptrTriangles, const long* ptrLabelTriangles,
                           const double* ptrVerts, const long* ptrLabelVerts, int nE, int nE_Omega,
                           int nV, int nV_Omega, const double* Px, int nPx, const double* dx,
                           const double* Py, int nPy, const double* dy, double sqdelta,
                           int is_DiscontinuousGalerkin, int is_NeumannBoundary,
                           const char* str_model_kernel, const char* str_model_f,
                           const char* str_integration_method_remote,
                           const char* str_integration_method_close, int is_PlacePointOnCap,
                           int dim, int outdim, const long* ptrZetaIndicator_indices,
                           const long* ptrZetaIndicator_indptr, long nZeta, const double* Pg,
                           int degree, const double* dg, double maxDiameter, double fractional_s,
                           int is_fullConnectedComponentSearch, int verbose)
    const char* nlfem_last_error()
