"""Pins of the oracle's CALPUV2C (EFDC/calpuv2c.f90:164-571,896-914) by
what the scheme fixes, chosen so that a wrong factor, sign or neighbour in
any of its formulas fails one of them:

* implicit consistency: the P that CONGRAD returns (before :898 replaces
  it) equals g*(HP+BELV) with the HP that the continuity update (:475)
  makes from the face transports (:426) recomputed from that P.  The
  coefficients (:329-341), FP1 (:289), FP (:320) and the transport update
  are one discretisation of continuity plus momentum only if their
  factors agree; checked to round-off on a tidal basin with Manning drag
  and a point source;
* seiche period: the scheme is the trapezoidal rule applied to the linear
  shallow-water system, so a free standing wave of a closed channel of
  uniform depth H and N cells advances by theta = 2 atan(omega dt/2) per
  step, omega = sqrt(g H) (2/dx) sin(pi/(2N)); along x this pins HU
  (:905), along y HV (:906) and the magnitude of the coupling constant;
  on a sloping bed omega comes from the eigenproblem of the staggered
  operator with face depths the mean of the two cells, which pins the
  neighbour HU is averaged with;
* energy: without drag the trapezoidal rule keeps
  E = sum g DXYP eta^2/2 + sum UHDYE^2/(2 HRUO HU) up to the drift of the
  depth-dependent HU; with Manning drag (quadratic bottom stress through
  TBX, :217) E falls at every step, which pins the sign of the stress.

test_planted_mistakes_fail builds the oracle with each ORC_PLANT mistake
(efdc_oracle.c) and shows the check named for it fails.
"""
import math
import os

import numpy as np
import pytest

from conftest import G

DX = 100.0


# ------------------------------------------------------------------ cases

def _tidal(lib_path, oracle_mod):
    from efdc_amd import grids
    case = grids.tidal_basin(24, 12, 3, seed=5)
    return case, oracle_mod.Oracle(case, rsqm=1e-28, iterm=3000, lib_path=lib_path)


def _channel(n, depth, along):
    """One row (along x) or one column (along y) of n wet cells."""
    from efdc_amd import grids
    wet = np.zeros((3, n + 2), bool) if along == "x" else np.zeros((n + 2, 3), bool)
    if along == "x":
        wet[1, 1:-1] = True
    else:
        wet[1:-1, 1] = True
    g = grids.build_grid(wet, 1)
    h, k, w = grids._empty_arrays(g)
    grids.fill_geometry(g, h, k, w, DX, np.ones(1))
    L = np.arange(2, g.LA + 1) - 1
    pos = (g.IL[L] - 2) if along == "x" else (g.JL[L] - 2)
    H = depth(pos) if callable(depth) else np.full(len(L), float(depth))
    return g, h, k, w, L, pos, H


def _at_rest_plus(g, h, k, w, L, eta, H, dt, manning=None):
    from efdc_amd import grids
    h["BELV"][L] = -H
    h["HP"][L] = H + eta
    h["H1P"][:] = h["HP"]
    h["H2P"][:] = h["HP"]
    h["P"][L] = G * eta
    h["P1"][L] = h["P"][L]
    grids.update_depth_arrays(g, h)
    grids.layer_depths(g, h, k)
    k["H1PK"][:] = k["HPK"]
    nm = None if manning is None else np.full(g.LCM, manning)
    return grids.Case(name="channel", grid=g, h=h, k=k, w=w, consts=[], dt=dt,
                      manning_n=nm)


# ----------------------------------------------------------------- checks

def consistency(oracle_mod, lib_path=None, nsteps=6):
    """max |P from CONGRAD - g(HP+BELV)| over nsteps, relative to the
    largest change of P in a step."""
    case, o = _tidal(lib_path, oracle_mod)
    L = np.arange(2, case.grid.LA + 1) - 1
    worst = dmax = 0.0
    for n in range(1, nsteps + 1):
        o.closure_pre()
        p_old = case.h["P"][L].copy()
        side = case.copy()
        so = oracle_mod.Oracle(side, rsqm=1e-28, iterm=3000, lib_path=lib_path)
        so.assemble(n=n)
        so.congrad()
        o.calpuv2c(n=n)
        o.closure_post()
        p_fin = G * (case.h["HP"][L] + case.h["BELV"][L])
        worst = max(worst, float(np.max(np.abs(side.h["P"][L] - p_fin))))
        dmax = max(dmax, float(np.max(np.abs(p_fin - p_old))))
    return worst / dmax if np.isfinite(worst) else math.inf


def seiche_error(oracle_mod, lib_path=None, along="x", n=32, H=10.0, slope=0.0,
                 amp=1e-3, dt=5.0, periods=2.0):
    """|theta_measured / theta_expected - 1| for the fundamental standing
    wave of a closed channel; theta from the wave's amplitude a_n by the
    three-term recurrence a_{n+1} + a_{n-1} = 2 cos(theta) a_n (least
    squares over two periods)."""
    depth = (lambda p: H + slope * (p - 0.5 * (n - 1))) if slope else H
    g, h, k, w, L, pos, Hc = _channel(n, depth, along)
    if slope:
        D = np.zeros((n - 1, n))
        D[np.arange(n - 1), np.arange(n - 1)] = -1.0
        D[np.arange(n - 1), np.arange(1, n)] = 1.0
        Hf = 0.5 * (Hc[:-1] + Hc[1:])
        ev, V = np.linalg.eigh((G / DX ** 2) * D.T @ np.diag(Hf) @ D)
        om, phi = math.sqrt(ev[1]), V[:, 1]
    else:
        om = math.sqrt(G * H) * (2.0 / DX) * math.sin(math.pi / (2 * n))
        phi = np.cos(math.pi * (pos + 0.5) / n)
    phi = phi / np.max(np.abs(phi))
    case = _at_rest_plus(g, h, k, w, L, amp * phi, Hc, dt)
    o = oracle_mod.Oracle(case, rsqm=1e-30, iterm=5000, lib_path=lib_path)
    nsteps = int(periods * 2.0 * math.pi / om / dt)
    a = [float(phi @ case.h["P"][L]) / G / float(phi @ phi)]
    for s in range(1, nsteps + 1):
        o.calpuv2c(n=s)
        a.append(float(phi @ case.h["P"][L]) / G / float(phi @ phi))
    a = np.array(a)
    x, y = a[1:-1], a[2:] + a[:-2]
    c = float(x @ y) / float(2.0 * (x @ x))
    if not np.isfinite(c) or abs(c) > 1.0:
        return math.inf            # no oscillation at all
    return abs(math.acos(c) / (2.0 * math.atan(om * dt / 2.0)) - 1.0)


def _energy(case, L):
    h = case.h
    eta = h["P"][L] / G
    pe = 0.5 * G * float(np.sum(h["DXYP"][L] * eta ** 2))
    f = h["SUB"][L] > 0
    ke = float(np.sum(h["UHDYE"][L][f] ** 2 / (2.0 * h["HRUO"][L][f] * h["HU"][L][f])))
    return pe + ke


def energy_series(oracle_mod, lib_path=None, manning=None, n=32, H=3.0, amp=0.1,
                  dt=5.0, periods=3):
    g, h, k, w, L, pos, Hc = _channel(n, H, "x")
    case = _at_rest_plus(g, h, k, w, L, amp * np.cos(math.pi * (pos + 0.5) / n), Hc,
                         dt, manning=manning)
    o = oracle_mod.Oracle(case, rsqm=1e-30, iterm=5000, lib_path=lib_path)
    om = math.sqrt(G * H) * (2.0 / DX) * math.sin(math.pi / (2 * n))
    E = [_energy(case, L)]
    for s in range(1, int(periods * 2.0 * math.pi / om / dt) + 1):
        if manning is not None:
            o.closure_pre()
        o.calpuv2c(n=s)
        E.append(_energy(case, L))
    return np.array(E)


def energy_without_drag_drift(oracle_mod, lib_path=None):
    E = energy_series(oracle_mod, lib_path)
    return float(np.max(np.abs(E / E[0] - 1.0)))


def energy_with_drag_largest_rise(oracle_mod, lib_path=None):
    """Largest relative rise of E in a step with Manning n = 0.04 (< 0
    when E falls at every step)."""
    E = energy_series(oracle_mod, lib_path, manning=0.04)
    return float(np.max(np.diff(E) / E[:-1]))


# ------------------------------------------------------- the checker's pins

CONSISTENCY_TOL = 1e-12
SEICHE_TOL = 1e-8           # uniform depth: exact dispersion relation
SLOPE_TOL = 1e-6            # sloping bed: amplitude 1e-3 m on 6-14 m
DRIFT_TOL = 1e-3            # 0.1 m on 3 m: the drift of HU with the depth


def test_implicit_consistency(oracle_mod):
    assert consistency(oracle_mod) <= CONSISTENCY_TOL


@pytest.mark.parametrize("along", ["x", "y"])
def test_seiche_period_uniform_depth(oracle_mod, along):
    assert seiche_error(oracle_mod, along=along) <= SEICHE_TOL


def test_seiche_period_sloping_bed(oracle_mod):
    assert seiche_error(oracle_mod, slope=0.25) <= SLOPE_TOL


def test_energy_kept_without_drag_and_falls_with_it(oracle_mod):
    assert energy_without_drag_drift(oracle_mod) <= DRIFT_TOL
    E = energy_series(oracle_mod, manning=0.04)
    assert np.all(np.diff(E) < 0.0)
    assert E[-1] < 0.5 * E[0]


# ------------------------------------------------------- planted mistakes

CHECKS = {
    "consistency": lambda m, p: consistency(m, p) > CONSISTENCY_TOL,
    "seiche_x": lambda m, p: seiche_error(m, p, along="x") > SEICHE_TOL,
    "seiche_y": lambda m, p: seiche_error(m, p, along="y") > SEICHE_TOL,
    "seiche_slope": lambda m, p: seiche_error(m, p, slope=0.25) > SLOPE_TOL,
    "drag_energy": lambda m, p: energy_with_drag_largest_rise(m, p) >= 0.0,
}

PLANTS = {
    1: "consistency",     # coefficient constant -G DELT/2
    2: "consistency",     # FP1 with G
    3: "consistency",     # FP with G
    4: "consistency",     # face transports with DELT
    5: "seiche_slope",    # HU from the east neighbour
    6: "seiche_y",        # HV without 1/2
    7: "drag_energy",     # bottom stress added
    8: "consistency",     # QSUME once in FP
    9: "seiche_x",        # continuity with DELT
}


@pytest.fixture(scope="module")
def plant_dir(tmp_path_factory):
    return tmp_path_factory.mktemp("plants")


@pytest.mark.parametrize("plant", sorted(PLANTS))
def test_planted_mistakes_fail(oracle_mod, plant_dir, plant):
    path = oracle_mod.build(plant=plant, out=os.path.join(str(plant_dir), f"p{plant}.so"))
    assert CHECKS[PLANTS[plant]](oracle_mod, path), PLANTS[plant]
