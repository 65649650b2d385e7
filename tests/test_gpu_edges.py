"""The HIP path against the oracle on the boundary cases the reference's
units have (the CPU tests hold the oracle to the reference's own compiled
CALTRAN and SETOPENBC on the same kinds of input):

* full steps of a basin open on all four sides: clamped rows (ISPB 0) and
  radiation rows (ISPB 1, 2, EFDC/setopenbc.f90:55-61,98-103,142-147,
  186-191) on W/E/S/N, a corner cell in two pressure sides and in two
  concentration sides, two tidal constituents, recovery ramps NTSCR 0..3,
  anti-diffusion with FCT, Manning drag and a dye source; the first
  step's assembly bit for bit, then P, depths and transports within the
  FP64 CG's round-off and the constituents within 1e-9 relative; the
  FP32-refined solve within the 1e-6 m bar;
* CALCONC alone on random inputs that reach every branch of CALTRAN's
  open-boundary code on all four sides (inflow and outflow, the salinity
  bottom clamp, ramps), sources with QSERCELL, and anti-diffusion with and
  without FCT: FP64 bit for bit over several steps.
"""
import numpy as np
import pytest

from conftest import eta_err, wet_index
from edge_cases import edge_basin, random_case

pytestmark = pytest.mark.gpu

ASSEMBLY = ["FUHDYE", "FVHDXE", "FP1", "FP", "CS", "CW", "CE", "CN", "CC",
            "CCS", "CCW", "CCE", "CCN", "CCC", "CCCI", "FPTMP"]


@pytest.mark.parametrize("cg_kernel", [0, 1])
@pytest.mark.parametrize("seed", [21, 22])
def test_four_open_sides_match_oracle(oracle_mod, seed, cg_kernel):
    from efdc_amd.lib import Context
    case = edge_basin(seed)
    ref = case.copy()
    o = oracle_mod.Oracle(ref, rsqm=1e-24, iterm=2000)
    ctx = Context(case.grid, RSQM=1e-24, ITERM=2000, cg_kernel=cg_kernel).upload_case(case)
    i = wet_index(case)
    kc = case.grid.KC
    for n in range(1, 121):
        so = o.step(n)
        sg = ctx.step(case.dt, n)
        assert so["status"] == 0 and sg["status"] == 0, n
        if n == 1:
            for name in ASSEMBLY:
                np.testing.assert_array_equal(ctx.get(name)[i], ref.h[name][i], err_msg=name)
        if n % 20 == 0 or n < 3:
            assert eta_err(ctx.get("P")[i], ref.h["P"][i]) <= 1e-9, n
            for name in ("HP", "UHDYE", "VHDXE", "HU", "HV"):
                a, b = ctx.get(name)[i], ref.h[name][i]
                assert np.max(np.abs(a - b)) <= 1e-9 * max(1.0, np.max(np.abs(b))), (n, name)
            for j, c in enumerate(ref.consts):
                con, con1 = ctx.get_constituent(j)
                scale = max(1e-30, np.max(np.abs(c["con"][:kc])))
                assert np.max(np.abs(con[:kc] - c["con"][:kc])) / scale <= 1e-9, (n, j)
                assert np.max(np.abs(con1[:kc] - c["con1"][:kc])) / scale <= 1e-9, (n, j)
    ctx.close()


def test_four_open_sides_mixed_solve_within_bar(oracle_mod):
    from efdc_amd.lib import CG_MIXED, Context
    case = edge_basin(23)
    ref = case.copy()
    o = oracle_mod.Oracle(ref, rsqm=1e-24, iterm=2000)
    ctx = Context(case.grid, RSQM=0.0, cg_rel_tol=1e-12, ITERM=2000,
                  cg_mode=CG_MIXED).upload_case(case)
    i = wet_index(case)
    for n in range(1, 81):
        o.step(n)
        ctx.step(case.dt, n)
    d = eta_err(ctx.get("P")[i], ref.h["P"][i])
    print("mixed, four open sides, 80 steps: max |d eta| =", d)
    assert d <= 1e-6
    ctx.close()


@pytest.mark.parametrize("isadac,isfct", [(0, 0), (1, 0), (1, 1)])
@pytest.mark.parametrize("seed", [1, 2])
def test_calconc_all_boundary_branches_bitexact(oracle_mod, seed, isadac, isfct):
    from efdc_amd.lib import Context
    case = random_case(seed, isadac=isadac, isfct=isfct)
    ref = case.copy()
    o = oracle_mod.Oracle(ref)
    ctx = Context(case.grid).upload_case(case)
    kc = case.grid.KC
    for niter in (1, 2, 3, 7):
        o.calconc(niter=niter)
        ctx.calconc(case.dt, niter)
        for j, c in enumerate(ref.consts):
            con, con1 = ctx.get_constituent(j)
            np.testing.assert_array_equal(con[:kc], c["con"][:kc],
                                          err_msg=f"CON mo={c['mo']} niter={niter}")
            np.testing.assert_array_equal(con1[:kc], c["con1"][:kc],
                                          err_msg=f"CON1 mo={c['mo']} niter={niter}")
    ctx.close()
