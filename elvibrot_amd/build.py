"""Build of the HIP library ``libsg4hpsi.so`` (gfx950) in-tree.

Called by ``__graft_entry__.build()``; compiles every time it is called.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsg4hpsi.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-munsafe-fp-atomics",
    # packed FP32 VALU instructions are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc():
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.exists(c) or c == "hipcc":
            return c


SOURCES = ("sg4_setup.cpp", "sg4_hpsi.hip", "sg4_k_split_f64.hip", "sg4_k_split_f32.hip", "sg4_k_mfma.hip", "sg4_cheby.hip",
           "sg4_k_grid_f64.hip", "sg4_k_grid_f32.hip", "sg4_block.hip")


def build(verbose=False):
    """Compile the translation units in parallel into _build/, then link."""
    obj_dir = os.path.join(HERE, "_build", "obj")
    os.makedirs(obj_dir, exist_ok=True)
    procs = []
    for f in SOURCES:
        obj = os.path.join(obj_dir, f + ".o")
        cmd = [hipcc()] + HIPCC_FLAGS + ["-c", "-o", obj, os.path.join(CSRC, f)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((f, obj, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                               stderr=subprocess.STDOUT, text=True)))
    objs, failed = [], []
    for f, obj, pr in procs:
        out, _ = pr.communicate()
        if pr.returncode != 0:
            failed.append(f)
            sys.stderr.write(out)
        objs.append(obj)
    if failed:
        raise RuntimeError("hipcc failed on " + ", ".join(failed))
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-shared", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed linking libsg4hpsi.so")
    os.replace(tmp, LIB)
    return LIB


FORTRAN = os.path.join(HERE, "fortran")
ROCM_LIB = "/opt/rocm/lib"
FBUILD = os.path.join(HERE, "_build")


def build_fortran(verbose=False):
    """iso_c_binding module + test driver with amdflang; the driver finds
    libsg4hpsi.so through an $ORIGIN-relative rpath."""
    os.makedirs(FBUILD, exist_ok=True)
    fc = "/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else "amdflang"
    mod_o = os.path.join(FBUILD, "mod_sg4_gpu.o")
    cmds = [[fc, "-O2", "-c", os.path.join(FORTRAN, "mod_sg4_gpu.f90"), "-J", FBUILD, "-o", mod_o]]
    exes = []
    for drv in ("test_sg4_fortran", "test_sg4_reflayout"):
        drv_o = os.path.join(FBUILD, drv + ".o")
        exe = os.path.join(FBUILD, drv)
        cmds += [[fc, "-O2", "-c", os.path.join(FORTRAN, drv + ".f90"), "-J", FBUILD, "-o", drv_o],
                 # the reference-layout driver allocates device blocks itself (hipMalloc)
                 [fc, drv_o, mod_o, "-L" + HERE, "-lsg4hpsi", "-L" + ROCM_LIB, "-lamdhip64",
                  "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath," + ROCM_LIB, "-o", exe]]
        exes.append(exe)
    for c in cmds:
        if verbose:
            print(" ".join(c), flush=True)
        r = subprocess.run(c, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("amdflang failed")
    return exes


if __name__ == "__main__":
    print(build(verbose=True))
