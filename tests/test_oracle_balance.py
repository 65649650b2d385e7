"""Pins of the oracle's two-level balances (ISBAL >= 1, EFDC/bal2t.f90):
BAL2T1's masked totals at the start of a period, the source terms of
BAL2T3A and the open-boundary fluxes of BAL2T2 that every CALCONC adds
to the outflow totals, and BAL2T5's totals and errors END - (BEG - OUT).

* volume: the depth update is the discrete continuity of the same face
  flows the balance adds up (UHDY2/VHDX2 across the faces of the
  open-boundary cells, QSUME of the sources), so the volume error of a
  period is round-off, for a tidal basin with a point source and a
  withdrawal, with the depth-uniform closure and with the internal mode;
* mass: with a uniform concentration everywhere (initial field, sources,
  boundary inflow) every flux carries that concentration, so the mass
  error is round-off too.  (With a varying field the reference's BAL2T2
  takes the upwind concentration after the transport, :193-214, while the
  transport moved the one before it, so its mass error is O(DT) per step;
  the checks use the uniform field for that reason.)

Planted mistakes (ORC_PLANT 30-32 in efdc_oracle.c) each fail one of
these.
"""
import os

import numpy as np
import pytest

VOL_TOL = 1e-12
MASS_TOL = 1e-11


def _basin(internal=False, uniform=False):
    """The tidal basin (west open boundary, M2 tide, dye source), a
    withdrawal source of -5e-4 m^3/s added in a corner cell, NTSCR = 0."""
    from efdc_amd import grids
    case = grids.tidal_basin(20, 10, 3, seed=13, dt=10.0, ntscr=0, amp=0.3, period=3000.0)
    g, k, h = case.grid, case.k, case.h
    src = case.sources
    lw = int(g.LIJ[8, 15])
    kc, nc = g.KC, len(case.consts)
    q = -5e-4 * k["DZC"][:kc, lw - 1]
    src["LQS"] = np.append(src["LQS"], np.int32(lw)).astype(np.int32)
    src["qss"] = np.vstack([src["qss"], q])
    src["qsercell"] = np.vstack([src["qsercell"], np.zeros(kc)])
    src["cqs"] = np.concatenate([src["cqs"], np.zeros((1, kc, nc))])
    src["csert"] = np.concatenate([src["csert"], np.zeros((1, kc, nc))])
    h["QSUME"][lw - 1] += q.sum()
    k["QSUM"][:kc, lw - 1] += q
    if internal:
        grids.add_internal_mode(case, av=1e-2)
    if uniform:
        for c in case.consts:
            c["con"][:] = 10.0
            c["con1"][:] = 10.0
        src["cqs"][:] = 10.0
        case.conc_bc["side"][1]["cb"][:] = 10.0
    return case


def period_errors(oracle_mod, lib_path=None, internal=False, uniform=False, nsteps=40):
    case = _basin(internal, uniform)
    o = oracle_mod.Oracle(case, rsqm=1e-26, iterm=3000, lib_path=lib_path)
    o.bal2t_begin()
    for n in range(1, nsteps + 1):
        st = o.step(n)
        assert st["status"] == 0
    b = o.bal2t_end()
    vol = abs(b["volerr"]) / max(abs(b["volout"]), 1e-300)
    mass = [abs(e) / max(abs(m), 1e-300) for e, m in zip(b["masserr"], b["massend"])]
    return b, vol, mass


def volume_error(oracle_mod, lib_path=None, internal=False):
    return period_errors(oracle_mod, lib_path, internal)[1]


def mass_error(oracle_mod, lib_path=None):
    return max(period_errors(oracle_mod, lib_path, uniform=True)[2])


@pytest.mark.parametrize("internal", [False, True])
def test_volume_balance_closes(oracle_mod, internal):
    b, vol, _ = period_errors(oracle_mod, internal=internal)
    # the period moves real water: outflow and sources both matter
    assert abs(b["volout"]) > 1.0
    assert vol <= VOL_TOL


def test_uniform_concentration_mass_balance_closes(oracle_mod):
    b, _, mass = period_errors(oracle_mod, uniform=True)
    assert all(abs(m) > 1.0 for m in b["massout"])
    assert max(mass) <= MASS_TOL


CHECKS = {
    "volume": lambda m, p: volume_error(m, p) > VOL_TOL,
    "mass": lambda m, p: mass_error(m, p) > MASS_TOL,
}

PLANTS = {
    30: "volume",   # west face flux with the wrong sign
    31: "mass",     # withdrawal left out of the source mass
    32: "volume",   # open-boundary cells counted in the volume
}


@pytest.fixture(scope="module")
def plant_dir(tmp_path_factory):
    return tmp_path_factory.mktemp("bal_plants")


@pytest.mark.parametrize("plant", sorted(PLANTS))
def test_planted_mistakes_fail(oracle_mod, plant_dir, plant):
    path = oracle_mod.build(plant=plant, out=os.path.join(str(plant_dir), f"p{plant}.so"))
    assert CHECKS[PLANTS[plant]](oracle_mod, path), PLANTS[plant]
