// Cassemble.h -- C++ drop-in for the reference's boundary header (cagau/nlfem
// include/Cassemble.h:62-71,125-140,202-204), implemented by libnlfem_amd.so on the GPU.
// Same names, argument order and meaning; invalid configurations print a message and
// abort(), as in the reference.  The C ABI underneath is include/nlfem_amd.h.
#ifndef NLFEM_AMD_CASSEMBLE_H
#define NLFEM_AMD_CASSEMBLE_H
#include <string>

// include/Cassemble.h:62-64 (exported TE overload, src/integration.cpp:1775-1894)
int method_retriangulate(const double* xCenter, const double* TE, double sqdelta, double* reTriangleList,
                         int isPlacePointOnCap);
// include/Cassemble.h:65-67; a stub in the reference too (src/integration.cpp:1513-1519): aborts.
int method_retriangulateInfty(const double* xCenter, const double* TE, double sqdelta, double* reTriangleList,
                              int isPlacePointOnCap);
// include/Cassemble.h:68-71
void toRef(const double* E, const double* phys_x, double* ref_p);
void toPhys(const double* E, const double* p, double* out_x);
void toPhys(const double* E, const double* p, int dim, double* out_x);
void solve2x2(const double* A, const double* b, double* x);

// include/Cassemble.h:125-140
void par_assemble(std::string compute, std::string path_spAd, std::string path_fd, int K_Omega, int K,
                  const long* ptrTriangles, const long* ptrLabelTriangles, const double* ptrVerts,
                  const long* ptrLabelVerts, int nE, int nE_Omega, int nV, int nV_Omega, const double* Px, int nPx,
                  const double* dx, const double* Py, int nPy, const double* dy, double sqdelta,
                  int is_DiscontinuousGalerkin, int is_NeumannBoundary, std::string str_model_kernel,
                  std::string str_model_f, std::string str_integration_method_remote,
                  std::string str_integration_method_close, int is_PlacePointOnCap, int dim, int outdim,
                  const long* ptrZetaIndicator_indices, const long* ptrZetaIndicator_indptr, long nZeta = 0,
                  const double* Pg = nullptr, int degree = 0, const double* dg = nullptr, double maxDiameter = 0.0,
                  double fractional_s = -1.0, int is_fullConnectedComponentSearch = 0, int verbose = 0);

// include/Cassemble.h:202-204
void par_evaluateMass(double* vd, const double* ud, long* Elements, const long* ElementLabels, const double* Verts,
                      const long* VertexLabels, int K_Omega, int J, int nP, double* P, const double* dx, int dim,
                      int outdim, int is_DiscontinuousGalerkin);

#endif
