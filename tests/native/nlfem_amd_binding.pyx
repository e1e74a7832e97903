# cython: language_level=3
# The call site of the reference's stiffnessMatrix (cython/nlfem.pyx:242-266) over the C ABI:
# the same arguments, the std::string paths passed as char pointers, and a nonzero status
# raised as RuntimeError.  Compiled by tests/native/build.py to check that the binding shown
# for maintainers builds against include/nlfem_amd.h and runs.
cimport Cassemble_amd as C


def par_assemble_system(str path, long K_Omega, long K, long[::1] elements, long[::1] elementLabels,
                        double[::1] vertices, long[::1] vertexLabels, int nE, int nE_Omega, int nV,
                        int nV_Omega, double[::1] Px, double[::1] dx, double[::1] Py, double[::1] dy,
                        double sqdelta, int is_DG, str kernel, str method_remote, str method_close,
                        int caps, int dim, int outdim, double fractional_s=-1.0):
    cdef long zero = 0
    cdef bytes p = path.encode(), k = kernel.encode(), mr = method_remote.encode()
    cdef bytes mc = method_close.encode()
    rc = C.nlfem_par_assemble(b"system", p, b"", K_Omega, K, &elements[0], &elementLabels[0],
                              &vertices[0], &vertexLabels[0], nE, nE_Omega, nV, nV_Omega,
                              &Px[0], Px.shape[0] // 2, &dx[0], &Py[0], Py.shape[0] // 2, &dy[0],
                              sqdelta, is_DG, 0, k, b"", mr, mc, caps, dim, outdim,
                              &zero, &zero, 0, NULL, 0, NULL, 0.0, fractional_s, 0, 0)
    if rc:
        raise RuntimeError(C.nlfem_last_error().decode())
    return rc
