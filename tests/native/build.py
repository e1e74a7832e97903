"""Builds the native callers of the boundary used by tests/test_native_boundary.py, in-tree
under tests/native/_build (run by __graft_entry__.build(); a CPU-only step):

- nlfem_amd_binding.so: the Cython binding of the C ABI that INTEGRATION.md shows (cython
  3, g++), linked against nlfem_amd/libnlfem_amd.so;
- dropin_main: a C++ program calling the std::string drop-in par_assemble of include/Cassemble.h,
  and libdropin.so, the same caller as a shared library (dropin_run) for an in-process call.

Both find the library through an rpath relative to their own location, so the built tree can
be moved.
"""
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "_build")
LIBDIR = os.path.join(ROOT, "nlfem_amd")
RPATH = "-Wl,-rpath,$ORIGIN/../../../nlfem_amd"


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{' '.join(cmd)} failed:\n{r.stderr[-4000:]}")


def build():
    os.makedirs(OUT, exist_ok=True)
    inc = os.path.join(ROOT, "include")
    cpp = os.path.join(OUT, "nlfem_amd_binding.cpp")
    _run([sys.executable, "-m", "cython", "--cplus", "-3", "-I", HERE, os.path.join(HERE, "nlfem_amd_binding.pyx"),
          "-o", cpp])
    ext = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    _run(["g++", "-O2", "-shared", "-fPIC", "-std=c++17", "-I", inc, "-I", sysconfig.get_paths()["include"], cpp,
          "-o", os.path.join(OUT, "nlfem_amd_binding" + ext), "-L", LIBDIR, "-lnlfem_amd", RPATH])
    _run(["g++", "-O2", "-std=c++17", "-I", inc, os.path.join(HERE, "dropin_main.cpp"), "-o",
          os.path.join(OUT, "dropin_main"), "-L", LIBDIR, "-lnlfem_amd", RPATH])
    _run(["g++", "-O2", "-shared", "-fPIC", "-std=c++17", "-DDROPIN_LIB", "-I", inc, os.path.join(HERE, "dropin_main.cpp"),
          "-o", os.path.join(OUT, "libdropin.so"), "-L", LIBDIR, "-lnlfem_amd", RPATH])
    return OUT


if __name__ == "__main__":
    print(build())
