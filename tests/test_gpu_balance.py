"""The two-level balances on the GPU (efdc_bal2t_begin / efdc_bal2t_end,
EFDC/bal2t.f90 BAL2T1, BAL2T3A, BAL2T2, BAL2T5) against the oracle's:
the same tidal basin with a point source and a withdrawal, 40 steps in
FP64 storage, with the depth-uniform closure and with the internal mode.

The period totals are FP64 reductions in another order than the oracle's
serial sums, so they agree to a few ulps of the totals; the outflow
totals (sums of O(sources + boundary cells) terms per step) agree to a
few ulps of the terms.  The errors END - (BEG - OUT) close as the
oracle's do (round-off for the volume, and for the mass with a uniform
concentration).  With FP32 storage the totals are sums of FP32 values
and agree within the FP32 bar.
"""
import numpy as np
import pytest

from test_oracle_balance import MASS_TOL, VOL_TOL, _basin

pytestmark = pytest.mark.gpu

KEYS = ("volbeg", "volend", "volout", "massbeg", "massend", "massout")


def _run(oracle_mod, internal=False, uniform=False, store=0, nsteps=40):
    from efdc_amd.lib import Context
    case = _basin(internal, uniform)
    ref = case.copy()
    o = oracle_mod.Oracle(ref, rsqm=1e-26, iterm=3000)
    ctx = Context(case.grid, RSQM=1e-26, ITERM=3000, store=store).upload_case(case)
    o.bal2t_begin()
    ctx.bal2t_begin()
    for n in range(1, nsteps + 1):
        so = o.step(n)
        sg = ctx.step(case.dt, n, internal=internal)
        assert so["status"] == 0 and sg["status"] == 0, n
    bo = o.bal2t_end()
    bg = ctx.bal2t_end()
    ctx.close()
    assert bg["nsteps"] == nsteps
    return bo, bg


def _mass_scale(b, j):
    """The size of constituent j's budget (a dye that starts at zero has
    no mass at the start, only what the source brings in)."""
    return max(abs(b["massbeg"][j]), abs(b["massend"][j]), abs(b["massout"][j]))


@pytest.mark.parametrize("internal", [False, True])
def test_period_matches_oracle_fp64(oracle_mod, internal):
    bo, bg = _run(oracle_mod, internal=internal)
    for k in ("volbeg", "volend"):
        np.testing.assert_allclose(bg[k], bo[k], rtol=1e-13, atol=0, err_msg=k)
    for k in ("massbeg", "massend"):
        for j in range(len(bo[k])):
            assert abs(bg[k][j] - bo[k][j]) <= 1e-13 * _mass_scale(bo, j), (k, j)
    # outflow totals: compared on the scale of the totals they change
    np.testing.assert_allclose(bg["volout"], bo["volout"], rtol=0,
                               atol=1e-13 * abs(bo["volbeg"]))
    for j in range(len(bo["massout"])):
        d = abs(bg["massout"][j] - bo["massout"][j])
        assert d <= 1e-12 * _mass_scale(bo, j), (j, bg["massout"][j], bo["massout"][j])
    # the volume budget closes on the GPU as in the oracle
    assert abs(bg["volerr"]) <= VOL_TOL * abs(bg["volout"]) + 1e-13 * abs(bg["volbeg"])
    assert abs(bo["volout"]) > 1.0


def test_uniform_concentration_mass_closes(oracle_mod):
    bo, bg = _run(oracle_mod, uniform=True)
    for j in range(len(bg["massend"])):
        assert abs(bg["massout"][j]) > 1.0
        assert abs(bg["masserr"][j]) <= MASS_TOL * abs(bg["massend"][j]) + \
            1e-13 * abs(bg["massbeg"][j]), j
        d = abs(bg["massout"][j] - bo["massout"][j])
        assert d <= 1e-12 * _mass_scale(bo, j), (j, bg["massout"][j], bo["massout"][j])


def test_period_fp32_storage(oracle_mod):
    bo, bg = _run(oracle_mod, store=1)
    for k in ("volbeg", "volend"):
        np.testing.assert_allclose(bg[k], bo[k], rtol=1e-5, err_msg=k)
    for k in ("massbeg", "massend"):
        for j in range(len(bo[k])):
            assert abs(bg[k][j] - bo[k][j]) <= 1e-5 * _mass_scale(bo, j), (k, j)
    np.testing.assert_allclose(bg["volout"], bo["volout"], rtol=0, atol=1e-5 * abs(bo["volbeg"]))


def test_end_without_begin_refused():
    from efdc_amd import grids
    from efdc_amd.lib import Context, EfdcError
    case = grids.tidal_basin(8, 6, 2, seed=3)
    ctx = Context(case.grid).upload_case(case)
    with pytest.raises(EfdcError):
        ctx.bal2t_end()
    ctx.bal2t_begin()
    ctx.bal2t_end()
    with pytest.raises(EfdcError):
        ctx.bal2t_end()
    ctx.close()
