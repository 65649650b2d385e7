"""The Fortran side of the boundary: the ISO_C_BINDING module
efdcplus8.5_amd/fortran/efdc_hip.f90, the drop-in CALPUV2C/CONGRAD/CALCONC
(efdc_hip_dropin.f90) and the check program tests/fortran/congrad_check.f90.

CPU tests: the Fortran parameters equal the header's enum values; every
bind(C) derived type has the size and member offsets of its C struct (both
compilers print them); the drop-in routines compile against the
reference's GLOBAL module; the check program links and runs (without a GPU
it must stop with the no-device status).  GPU test: the check program
solves a manufactured CONGRAD system in both CG modes through the Fortran
interface.
"""
import os
import re
import subprocess

import pytest

from conftest import ROOT

FORTRAN = os.path.join(ROOT, "efdcplus8.5_amd", "fortran")
IFACE = os.path.join(FORTRAN, "efdc_hip.f90")
DROPIN = os.path.join(FORTRAN, "efdc_hip_dropin.f90")
CHECK = os.path.join(FORTRAN, "_bin", "congrad_check")
FC = os.environ.get("EFDC_FC", "/opt/rocm/llvm/bin/flang")
GLOBAL_MOD = os.path.join(ROOT, "oracle", "_ref", "global.mod")

STRUCTS = {
    # C struct -> (Fortran type, members in order)
    "efdc_grid_t": ["LA", "LCM", "KC", "KCM", "LWC", "LEC", "LSC", "LNC",
                    "LNWC", "LSEC"],
    "efdc_params_t": ["G", "RITB1", "SNLT", "HDRY", "TCON_TBEGIN", "ITERM",
                      "RSQM", "cg_rel_tol", "cg_mode", "store",
                      "mixed_inner_rtol", "mixed_max_outer", "cg_kernel",
                      "mixed_precond_degree", "tran_shape", "ISDRY", "NDRYSTP"],
    "efdc_pbc_side_t": ["n", "L", "ispb", "pser", "pcb", "psb"],
    "efdc_open_bc_t": ["side", "mtide", "tcp"],
    "efdc_cbc_side_t": ["n", "L", "cb", "csert", "ntscr"],
    "efdc_conc_bc_t": ["side", "wtci"],
    "efdc_sources_t": ["n", "LQS", "qss", "qsercell", "cqs", "csert"],
    "efdc_constituent_t": ["mo", "isadac", "isfct", "rkdye", "is_dye", "wset",
                           "probdep", "wese"],
    "efdc_step_t": ["dt", "n", "niter"],
    "efdc_cg_stats_t": ["iters", "outer", "rsq", "rsqm_used", "ccmnm"],
    "efdc_balance_t": ["volume", "mass", "nconst"],
}


def _fortran_parameters():
    text = open(IFACE).read()
    text = re.sub(r"!.*", "", text)
    text = text.replace("&\n", " ")
    out = {}
    for m in re.finditer(r"parameter\s*::([^\n]*)", text):
        for k, v in re.findall(r"(EFDC_\w+)\s*=\s*(-?\d+)", m.group(1)):
            out[k] = int(v)
    return out


def test_fortran_parameters_match_header():
    from efdc_amd import abi
    f = _fortran_parameters()
    arrays = {k: v for k, v in f.items() if k[5:] in abi.ARRAY_IDS}
    assert {k[5:]: v for k, v in arrays.items()} == abi.ARRAY_IDS
    for k, v in abi.STATUS.items():
        assert f[k] == v, k
    for name in ("efdc_cg_mode", "efdc_store", "efdc_cg_kernel"):
        for k, v in abi._parse_enum(name).items():
            assert f[k] == v, k


def _need_fc():
    if not os.path.exists(FC):
        pytest.skip("no Fortran compiler")


def test_bindc_layouts_match_c_structs(tmp_path):
    _need_fc()
    fprog = ["program layout", "  use iso_c_binding", "  use efdc_hip",
             "  implicit none"]
    for s in STRUCTS:
        fprog.append(f"  type({s}), target :: v_{s}")
    for s, members in STRUCTS.items():
        fprog.append(f"  print '(A,1X,A,1X,I0)', '{s}', 'SIZE', "
                     f"c_sizeof(v_{s})")
        for m in members:
            ref = f"v_{s}%{m}" + ("(1)" if m in ("side", "mass") else "")
            fprog.append(
                f"  print '(A,1X,A,1X,I0)', '{s}', '{m}', "
                f"transfer(c_loc({ref}), 0_c_intptr_t) - "
                f"transfer(c_loc(v_{s}), 0_c_intptr_t)")
    fprog.append("end program")
    (tmp_path / "layout.f90").write_text("\n".join(fprog) + "\n")
    mod = tmp_path / "mod"
    mod.mkdir()
    subprocess.check_call([FC, "-c", "-module-dir", str(mod), IFACE, "-o",
                           str(mod / "efdc_hip.o")])
    subprocess.check_call([FC, "-I", str(mod), str(tmp_path / "layout.f90"),
                           str(mod / "efdc_hip.o"),
                           # the layout program calls no library function
                           "-Wl,--unresolved-symbols=ignore-all", "-o",
                           str(tmp_path / "flayout")])
    fout = subprocess.run([str(tmp_path / "flayout")], capture_output=True,
                          text=True, check=True).stdout.split("\n")

    cprog = ["#include <stdio.h>", "#include <stddef.h>",
             '#include "efdc_hip.h"', "int main(void) {"]
    for s, members in STRUCTS.items():
        cprog.append(f'  printf("%s SIZE %zu\\n", "{s}", sizeof({s}));')
        for m in members:
            cprog.append(f'  printf("%s %s %zu\\n", "{s}", "{m}", '
                         f"offsetof({s}, {m}));")
    cprog.append("  return 0; }")
    (tmp_path / "layout.c").write_text("\n".join(cprog) + "\n")
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"),
                           str(tmp_path / "layout.c"), "-o",
                           str(tmp_path / "clayout")])
    cout = subprocess.run([str(tmp_path / "clayout")], capture_output=True,
                          text=True, check=True).stdout.split("\n")
    norm = lambda lines: sorted(" ".join(ln.split()) for ln in lines if ln.strip())
    assert norm(fout) == norm(cout)


def test_dropin_compiles_against_reference_global(tmp_path):
    _need_fc()
    if not os.path.exists(GLOBAL_MOD):
        pytest.skip("reference GLOBAL module not built (oracle/build_ref.py)")
    mod = tmp_path / "mod"
    mod.mkdir()
    subprocess.check_call([FC, "-c", "-module-dir", str(mod), IFACE, "-o",
                           str(mod / "efdc_hip.o")])
    r = subprocess.run([FC, "-c", "-fdefault-real-8", "-cpp", "-module-dir",
                        str(mod), "-I", os.path.dirname(GLOBAL_MOD), DROPIN,
                        "-o", str(tmp_path / "dropin.o")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    nm = subprocess.run(["nm", str(tmp_path / "dropin.o")], capture_output=True,
                        text=True, check=True).stdout
    for sym in ("calpuv2c_", "congrad_", "calconc_"):
        assert re.search(r"\bT " + sym + r"$", nm, re.M), sym


def test_check_program_links():
    if not os.path.exists(CHECK):
        pytest.skip("Fortran check program not built (build_fortran)")
    r = subprocess.run([CHECK], capture_output=True, text=True, timeout=120)
    if r.returncode != 0:
        # a machine without a GPU reaches the library and gets its status
        assert "efdc_ctx_create failed  7 no HIP device visible" in r.stdout, r.stdout
    else:
        assert "FORTRAN_CONGRAD_OK" in r.stdout


@pytest.mark.gpu
def test_fortran_congrad_on_gpu():
    assert os.path.exists(CHECK), "build the Fortran check program first"
    r = subprocess.run([CHECK], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FORTRAN_CONGRAD_OK" in r.stdout
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith("FORTRAN_CONGRAD mode=")]
    assert len(lines) == 2
