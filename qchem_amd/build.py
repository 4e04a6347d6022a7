"""In-tree build of the native libraries (gfx950).

    libqcjk.so         HIP kernels + the C ABI of include/qcjk.h      (hipcc --offload-arch=gfx950)
    libqchem_host.so   C++20 host side: the Orbital_HF_IBS / WholeSystemFock_IBS adapter and the RHF
                       driver (Hamiltonian, DIIS, SCF loop) that calls the engine through the C ABI

Run ``python -m qchem_amd.build`` (or ``__graft_entry__.build()``).  Always recompiles.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
INCLUDE = os.path.join(ROOT, "include")
OBJ = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC",
    # packed FP32 instructions are disabled for every HIP translation unit
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wno-unused-result", "-Wno-unused-value",
    "-I" + INCLUDE, "-I" + CSRC,
]
# longest compiles first (they bound the parallel build)
HIP_SOURCES = ["qcjk_k_heavy2r_c.hip", "qcjk_k_heavy2r_d.hip", "qcjk_k_heavy2r_b.hip", "qcjk_k_heavy2r_a.hip",
               "qcjk_k_lightc_b.hip", "qcjk_k_light_c.hip", "qcjk_k_light_b.hip", "qcjk_k_lightc.hip", "qcjk_k_light_a.hip",
               "qcjk_k_heavy2_c.hip", "qcjk_k_heavy2_d.hip", "qcjk_k_heavy2_b.hip", "qcjk_k_heavy2_a.hip",
               "qcjk_k_heavy.hip", "qcjk_eri.hip", "qcjk_setup.hip", "qcjk_api.hip"]
ML_SOURCE = "qcml.hip"
HOST_SOURCES = ["linalg.cpp", "scf.cpp", "capi.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), r.stdout))
    return r.stdout


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    for src in HIP_SOURCES:
        obj = os.path.join(OBJ, src.replace(".hip", ".o"))
        jobs.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
    hip_objs = [j[-1] for j in jobs]
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        for out in ex.map(_run, jobs):
            if verbose and out.strip():
                print(out)
    lib = os.path.join(HERE, "libqcjk.so")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", lib + ".tmp"] + hip_objs)
    os.replace(lib + ".tmp", lib)

    # learned-guess network: its own library (include/qcml.h)
    mobj = os.path.join(OBJ, "qcml.o")
    _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, ML_SOURCE), "-o", mobj])
    mlib = os.path.join(HERE, "libqcml.so")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", mlib + ".tmp", mobj])
    os.replace(mlib + ".tmp", mlib)

    cxx = shutil.which("g++") or "c++"
    host_objs = []
    for src in HOST_SOURCES:
        obj = os.path.join(OBJ, src.replace(".cpp", ".o"))
        # x86-64-v3 (AVX2 + FMA): every EPYC host of an MI355X has it; the dense host algebra of
        # the SCF loop (products, tridiagonal eigensolver) runs ~3x faster with it
        _run([cxx, "-std=c++20", "-O3", "-march=x86-64-v3", "-fPIC", "-Wall", "-I" + INCLUDE, "-I" + HOST, "-c",
              os.path.join(HOST, src), "-o", obj])
        host_objs.append(obj)
    hlib = os.path.join(HERE, "libqchem_host.so")
    _run([cxx, "-shared", "-fPIC", "-o", hlib + ".tmp"] + host_objs +
         ["-L" + HERE, "-lqcjk", "-Wl,-rpath,$ORIGIN"])
    os.replace(hlib + ".tmp", hlib)
    return lib, hlib, mlib


if __name__ == "__main__":
    for p in build(verbose="-v" in sys.argv):
        print(p)
