"""Build the gfx950 shared library in-tree with hipcc.

The library is ``efdcplus8.5_amd/efdc_amd/libefdc_hip.so``.  All HIP code
is compiled for gfx950 only, with packed-FP32 instructions disabled and
FP contraction off (the FP64 kernels then round exactly like the CPU
restatement they are compared with).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(HERE, "libefdc_hip.so")
SOURCES = ["efdc_ctx.hip", "efdc_puv.hip", "efdc_tran.hip", "efdc_cgp.hip",
           "efdc_seg.hip", "efdc_adj.hip", "efdc_uvw.hip", "efdc_bal.hip",
           "efdc_io.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def build(verbose: bool = False, jobs: int = 4) -> str:
    """Compile every source to an object, then link the shared library."""
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                            stderr=subprocess.STDOUT)))
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        # hipcc's host pass does not know the packed-fp32 feature: harmless
        lines = [ln for ln in text.splitlines()
                 if "packed-fp32-ops" not in ln and ln.strip()]
        if p.returncode != 0:
            failed.append(src)
            print(text, file=sys.stderr)
        elif verbose and lines:
            print("\n".join(lines))
    if failed:
        raise RuntimeError(f"hipcc failed for {failed}")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    return LIB


FORTRAN = os.path.join(PKG, "fortran")
FBIN = os.path.join(FORTRAN, "_bin")
FC = os.environ.get("EFDC_FC", "/opt/rocm/llvm/bin/flang")


def build_fortran(global_mod_dir: str | None = None) -> str:
    """Compile the ISO_C_BINDING module efdc_hip and link the Fortran
    check program tests/fortran/congrad_check.f90 against the library.

    With ``global_mod_dir`` (a directory holding the host's global.mod),
    the drop-in CALPUV2C/CONGRAD/CALCONC are compiled too, as the host's
    build would compile them.  Returns the path of the check program.
    """
    os.makedirs(FBIN, exist_ok=True)
    mod = os.path.join(FBIN, "mod")
    os.makedirs(mod, exist_ok=True)
    iface = os.path.join(FORTRAN, "efdc_hip.f90")
    subprocess.check_call([FC, "-O2", "-fPIC", "-c", "-module-dir", mod, iface,
                           "-o", os.path.join(mod, "efdc_hip.o")])
    if global_mod_dir:
        subprocess.check_call([FC, "-O2", "-fPIC", "-c", "-fdefault-real-8", "-cpp",
                               "-module-dir", mod, "-I", global_mod_dir,
                               os.path.join(FORTRAN, "efdc_hip_dropin.f90"),
                               "-o", os.path.join(mod, "efdc_hip_dropin.o")])
        gobj = os.path.join(global_mod_dir, "var_global_mod.o")
        if os.path.exists(gobj):
            build_dropin_harness(global_mod_dir, mod, gobj)
    exe = os.path.join(FBIN, "congrad_check")
    subprocess.check_call([FC, "-O2", "-I", mod,
                           os.path.join(ROOT, "tests", "fortran", "congrad_check.f90"),
                           os.path.join(mod, "efdc_hip.o"),
                           "-L" + HERE, "-lefdc_hip",
                           "-Wl,-rpath,$ORIGIN/../../efdc_amd", "-o", exe])
    return exe


HARNESS = os.path.join(FBIN, "libdropin_harness.so")


def build_dropin_harness(global_mod_dir: str, mod: str, gobj: str) -> str:
    """The test harness that runs the drop-ins through the host's MODULE
    GLOBAL (tests/fortran/gen_dropin_harness.py) as a shared library."""
    import importlib.util
    gen_path = os.path.join(ROOT, "tests", "fortran", "gen_dropin_harness.py")
    spec = importlib.util.spec_from_file_location("gen_dropin_harness", gen_path)
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    src = os.path.join(FBIN, "dropin_harness.f90")
    open(src, "w").write(gen.gen())
    hobj = os.path.join(FBIN, "dropin_harness.o")
    subprocess.check_call([FC, "-O2", "-fPIC", "-c", "-fdefault-real-8", "-cpp",
                           "-module-dir", mod, "-I", global_mod_dir, "-I", mod, src,
                           "-o", hobj])
    tmp = HARNESS + ".tmp"
    subprocess.check_call([FC, "-shared", "-fPIC", gobj, os.path.join(mod, "efdc_hip.o"),
                           os.path.join(mod, "efdc_hip_dropin.o"), hobj,
                           "-L" + HERE, "-lefdc_hip",
                           "-Wl,-rpath,$ORIGIN/../../efdc_amd", "-o", tmp])
    os.replace(tmp, HARNESS)
    return HARNESS


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_fortran())
