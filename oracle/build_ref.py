"""Build oracle/_ref/libefdc_ref.so from the reference's own Fortran units.

Compiles, unmodified and where they lie under /root/reference/EFDC,
var_global_mod.f90, caltran.f90, calfqc.f90, caldiff.f90, setopenbc.f90
and setopenbc2.f90 with flang (-fdefault-real-8: the ReleaseDP64 REAL
kind), plus the generated harness oracle/_ref/ref_driver.f90.  All
outputs go to oracle/_ref/.  Used only by tests as the checker of the C
restatement.  CONGRAD, CALPUV2C and CALCONC are not built: they need the
Intel IFPORT library (through EFDCOUT, DSTIME, HEAT_MODULE), which this
image lacks.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_ref")
SRC = "/root/reference/EFDC"
UNITS = ["var_global_mod.f90", "caldiff.f90", "calfqc.f90", "caltran.f90",
         "setopenbc.f90", "setopenbc2.f90"]
FLANG = os.environ.get("FLANG", "/opt/rocm/llvm/bin/flang")
FLAGS = ["-O2", "-fPIC", "-cpp", "-fdefault-real-8", "-ffp-contract=off"]


def build() -> str:
    if not os.path.isdir(SRC):
        raise SystemExit("reference sources not present; nothing to build")
    os.makedirs(OUT, exist_ok=True)
    sys.path.insert(0, HERE)
    import gen_ref_driver
    drv = os.path.join(OUT, "ref_driver.f90")
    open(drv, "w").write(gen_ref_driver.gen())
    objs = []
    for u in UNITS + [drv]:
        src = u if os.path.isabs(u) else os.path.join(SRC, u)
        obj = os.path.join(OUT, os.path.basename(u).replace(".f90", ".o"))
        subprocess.check_call([FLANG, *FLAGS, "-module-dir", OUT, "-c", src,
                               "-o", obj], cwd=OUT)
        objs.append(obj)
    lib = os.path.join(OUT, "libefdc_ref.so")
    subprocess.check_call([FLANG, "-shared", "-fPIC", *objs, "-o", lib], cwd=OUT)
    return lib


if __name__ == "__main__":
    print(build())
