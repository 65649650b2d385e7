"""Pins of the C oracle by what mathematics fixes, for the units whose
reference sources cannot be built here (CONGRAD, CALPUV2C and CALCONC's
vertical diffusion need Intel's IFPORT through EFDCOUT/DSTIME/
HEAT_MODULE):

* CONGRAD: the solution equals scipy's sparse direct solve; the iterates
  equal scipy's preconditioned CG after the same number of iterations;
* CALPUV2C assembly: exact symmetry CS(L)=CN(LS), CW(L)=CE(LW); row sums
  CC+CS+CW+CE+CN = DXYP/dt; strict diagonal dominance (hence SPD);
  clamped open-boundary rows reproduce the tide exactly;
* CALPUV2C as a whole: a lake at rest stays at rest; a closed basin keeps
  its volume to round-off; a point source adds exactly Q*dt of volume;
* CALTRAN: a uniform concentration stays uniform under a discretely
  divergence-free flow; total mass is conserved in a closed basin; at
  Courant number 1 the upwind scheme shifts the field by one cell;
* vertical diffusion: equals scipy's banded solve of the standard
  finite-volume implicit diffusion matrix; conserves column mass;
* sources: the dye mass equals load x time;
* dye decay (EFDC/calconc.f90:581-616): with no flow the dye follows the
  closed forms C*(1/(1+k dt))^n, C*exp(-k dt)^n and C + n dt/86400 (age);
* failure paths: ITERM exceeded and negative depth return their status.
"""
import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp
import scipy.sparse.linalg as spl

from conftest import G, wet_index


def lake(n=40, seed=3, **kw):
    from efdc_amd import grids
    return grids.external_case(grids.lake_mask(n, n, seed), seed=seed, **kw)


def system_matrix(case, scaled=True):
    g = case.grid
    i = wet_index(case)
    names = (("CCC", None), ("CCS", "LSC"), ("CCN", "LNC"), ("CCW", "LWC"),
             ("CCE", "LEC")) if scaled else (("CC", None), ("CS", "LSC"),
                                              ("CN", "LNC"), ("CW", "LWC"),
                                              ("CE", "LEC"))
    rows, cols, vals = [], [], []
    for name, mp in names:
        cc = i if mp is None else g.maps[mp][i] - 1
        keep = (cc >= 1) & (cc <= g.LA - 1)
        rows.append(i[keep] - 1)
        cols.append(cc[keep] - 1)
        vals.append(case.h[name][i][keep])
    n = g.LA - 1
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows),
                          np.concatenate(cols))), shape=(n, n))


# ------------------------------------------------------------------ CONGRAD

def test_congrad_solution_equals_sparse_direct_solve(oracle_mod):
    case = lake(48, seed=3)
    o = oracle_mod.Oracle(case, rsqm=1e-28, iterm=3000)
    o.assemble()
    A = system_matrix(case)
    i = wet_index(case)
    x = spl.spsolve(A.tocsc(), case.h["FPTMP"][i])
    st = o.congrad()
    assert st["status"] == 0
    assert np.max(np.abs(case.h["P"][i] - x)) <= 1e-12 * np.max(np.abs(x))


@pytest.mark.parametrize("k", [1, 2, 5, 12])
def test_congrad_iterates_equal_scipy_jacobi_pcg(oracle_mod, k):
    """After k iterations from the same start, CONGRAD's P equals the
    k-th iterate of scipy's CG preconditioned with diag(A)^-1."""
    case = lake(32, seed=7)
    o = oracle_mod.Oracle(case, rsqm=0.0, iterm=k)
    o.assemble()
    A = system_matrix(case)
    i = wet_index(case)
    b = case.h["FPTMP"][i].copy()
    x0 = case.h["P"][i].copy()
    Minv = spl.LinearOperator(A.shape, matvec=lambda v: case.h["CCCI"][i] * v)
    xs, _ = spl.cg(A, b, x0=x0, rtol=0.0, atol=0.0, maxiter=k, M=Minv)
    st = o.congrad()
    assert st["iters"] == k + 1 and st["status"] != 0  # ITERM reached
    d = np.max(np.abs(case.h["P"][i] - xs)) / np.max(np.abs(xs - x0))
    assert d <= 1e-10, d


def test_congrad_stops_on_rsqm_like_the_reference(oracle_mod):
    """The stop test is sum r^2 <= RSQM on the scaled system."""
    case = lake(32, seed=9)
    o = oracle_mod.Oracle(case, rsqm=1e-12, iterm=500)
    o.assemble()
    st = o.congrad()
    A = system_matrix(case)
    i = wet_index(case)
    r = case.h["FPTMP"][i] - A @ case.h["P"][i]
    assert st["rsq"] <= 1e-12
    assert abs(float(r @ r) - st["rsq"]) <= 1e-6 * 1e-12 + 1e-3 * st["rsq"]


# -------------------------------------------------------------- assembly

def test_assembly_symmetric_rowsum_and_dominant(oracle_mod):
    case = lake(40, seed=4)
    o = oracle_mod.Oracle(case)
    o.assemble()
    g = case.grid
    i = wet_index(case)
    h = case.h
    ls = g.maps["LSC"][i] - 1
    lw = g.maps["LWC"][i] - 1
    ws = ls != g.LC - 1
    ww = lw != g.LC - 1
    np.testing.assert_array_equal(h["CS"][i][ws], h["CN"][ls[ws]])
    np.testing.assert_array_equal(h["CW"][i][ww], h["CE"][lw[ww]])
    rs = h["CC"][i] + h["CS"][i] + h["CW"][i] + h["CE"][i] + h["CN"][i]
    np.testing.assert_allclose(rs, h["DXYP"][i] / case.dt, rtol=1e-13)
    off = np.abs(h["CS"][i]) + np.abs(h["CW"][i]) + np.abs(h["CE"][i]) + np.abs(h["CN"][i])
    assert np.all(h["CC"][i] > off)
    A = system_matrix(case)
    assert abs(A - A.T).max() == 0.0
    # scaling by the minimum diagonal
    assert np.min(h["CCC"][i]) == 1.0


def test_clamped_open_boundary_reproduces_tide(oracle_mod):
    from efdc_amd import grids
    case = grids.tidal_basin(24, 12, 3, seed=5)
    o = oracle_mod.Oracle(case, rsqm=1e-26, iterm=1000)
    n = 37
    o.calpuv2c(n=n)
    pi2 = 2.0 * 3.1415926535898   # EFDC/aaefdc.f90:650-651
    tn = case.dt * n
    expect = G * 0.5 * np.cos(pi2 * np.fmod(tn, 44712.0) / 44712.0)
    lb = case.open_bc["side"][0]["L"] - 1
    np.testing.assert_allclose(case.h["P"][lb], expect, rtol=1e-12)
    # the boundary rows are decoupled from the interior
    lec = case.grid.maps["LEC"][lb] - 1
    assert np.all(case.h["CW"][lec] == 0.0)


# ------------------------------------------------------------ CALPUV2C

def test_lake_at_rest_stays_at_rest(oracle_mod):
    case = lake(32, seed=2, eta_sigma=0.0)
    i = wet_index(case)
    hp0 = case.h["HP"].copy()
    o = oracle_mod.Oracle(case, rsqm=1e-20)
    st = o.calpuv2c(n=1)
    assert st["iters"] == 0 and st["status"] == 0   # RPCG == 0 (congrad.f90:88)
    np.testing.assert_array_equal(case.h["HP"][i], hp0[i])
    assert np.all(case.h["UHDYE"][i] == 0.0) and np.all(case.h["VHDXE"][i] == 0.0)
    assert np.max(np.abs(case.h["P"][i])) <= 1e-12


def test_closed_basin_volume_conserved(oracle_mod):
    case = lake(40, seed=6)
    i = wet_index(case)
    v0 = float(np.sum(case.h["DXYP"][i] * case.h["HP"][i]))
    o = oracle_mod.Oracle(case, rsqm=1e-9)       # holds for any CG accuracy
    for n in range(1, 6):
        o.calpuv2c(n=n)
    v1 = float(np.sum(case.h["DXYP"][i] * case.h["HP"][i]))
    assert abs(v1 - v0) <= 1e-13 * v0
    assert np.max(np.abs(case.h["UHDYE"][i])) > 0  # the lake did move


def test_point_source_adds_its_volume(oracle_mod):
    from efdc_amd import grids
    case = grids.tidal_basin(20, 10, 2, seed=8)
    case.open_bc = None
    case.conc_bc = None
    i = wet_index(case)
    v0 = float(np.sum(case.h["DXYP"][i] * case.h["HP"][i]))
    o = oracle_mod.Oracle(case, rsqm=1e-20)
    for n in range(1, 11):
        o.calpuv2c(n=n)
    v1 = float(np.sum(case.h["DXYP"][i] * case.h["HP"][i]))
    assert abs((v1 - v0) - 10 * case.dt * 1e-3) <= 1e-14 * v0


# ------------------------------------------------------------- transport

def no_source_transport(seed=2, nx=24, ny=16, kc=5):
    from efdc_amd import grids
    case = grids.transport_case(nx, ny, kc, seed=seed, sed_ws=0.0)
    case.sources = None
    case.k["QSUM"][:] = 0.0
    return case


def test_uniform_concentration_is_preserved(oracle_mod):
    case = no_source_transport()
    kc = case.grid.KC
    i = wet_index(case)
    for c in case.consts:
        c["con"][:kc, i] = 7.5
        c["con1"][:] = c["con"]
    o = oracle_mod.Oracle(case)
    for n in range(1, 4):
        o.calconc(niter=n)
    for c in case.consts:
        if c["is_dye"]:
            continue
        assert np.max(np.abs(c["con"][:kc, i] - 7.5)) <= 1e-12 * 7.5


def test_closed_basin_mass_conserved(oracle_mod):
    case = no_source_transport(seed=4)
    o = oracle_mod.Oracle(case)
    _, m0 = o.balance()
    for n in range(1, 6):
        o.calconc(niter=n)
    _, m1 = o.balance()
    np.testing.assert_allclose(m1, m0, rtol=1e-12, atol=1e-9)


def test_courant_one_shifts_by_one_cell(oracle_mod):
    from efdc_amd import grids
    nx, kc, dx, dt, h = 30, 1, 100.0, 5.0, 2.0
    wet = np.zeros((3, nx + 2), bool)
    wet[1, 1:-1] = True
    g = grids.build_grid(wet, kc)
    hh, k, w = grids._empty_arrays(g)
    grids.fill_geometry(g, hh, k, w, dx, np.ones(kc))
    i = wet_index(type("C", (), {"grid": g})())
    hh["HP"][i] = h
    hh["HPI"][i] = 1 / h
    k["HPK"][0, i] = h
    k["H1PK"][0, i] = h
    k["HPKI"][0, i] = 1 / h
    U = h / (dt * hh["DXYIP"][i][0])          # Courant number 1
    k["UHDY2"][0, i] = U * hh["SUB"][i]
    rng = np.random.default_rng(0)
    con = np.zeros((g.KCM, g.LCM))
    con[0, i] = rng.uniform(1, 2, len(i))
    case = grids.Case(name="c1", grid=g, h=hh, k=k, w=w, dt=dt,
                      consts=[dict(mo=4, isadac=0, isfct=0, rkdye=0.0, is_dye=0,
                                   con=con, con1=con.copy())])
    old = con[0, i].copy()
    oracle_mod.Oracle(case).calconc(niter=1)
    new = case.consts[0]["con"][0, i]
    np.testing.assert_allclose(new[1:-1], old[:-2], rtol=1e-14)
    assert new[0] == 0.0


def test_vertical_diffusion_equals_finite_volume_banded_solve(oracle_mod):
    """Implicit diffusion (I - dt D) C' = C with the standard flux form
    D: layer thickness H*DZC_K, interface spacing H*DZG_K, interface
    diffusivity Kz_K = AB_K*H (EFDC carries AB scaled by 1/H)."""
    from efdc_amd import grids
    kc = 7
    case = grids.tidal_basin(6, 4, kc, seed=3)
    rng = np.random.default_rng(1)
    dzck = rng.uniform(0.5, 2.0, kc)
    grids.fill_geometry(case.grid, case.h, case.k, case.w, 100.0, dzck)
    i = wet_index(case)
    case.k["AB"][:kc - 1, i] = np.exp(rng.uniform(np.log(1e-5), np.log(1e-1), (kc - 1, len(i))))
    case.h["HPI"][i] = 1.0 / case.h["HP"][i]
    for c in case.consts:
        c["con"][:kc, i] = rng.uniform(0, 10, (kc, len(i)))
    before = [c["con"].copy() for c in case.consts]
    o = oracle_mod.Oracle(case)
    o.vertical_diffusion()
    dt = case.dt
    dzc = case.k["DZC"][:kc, i[0]]
    dzg = 0.5 * (dzc[:-1] + dzc[1:])
    for c, b in zip(case.consts, before):
        for col in (0, 5, len(i) - 1):
            L = i[col]
            H = case.h["HP"][L]
            kz = case.k["AB"][:kc - 1, L] * H
            ab = np.zeros((3, kc))
            for K in range(kc):
                up = dt * kz[K] / (H * dzc[K] * H * dzg[K]) if K < kc - 1 else 0.0
                lo = dt * kz[K - 1] / (H * dzc[K] * H * dzg[K - 1]) if K > 0 else 0.0
                ab[1, K] = 1 + up + lo
                if K < kc - 1:
                    ab[0, K + 1] = -up
                if K > 0:
                    ab[2, K - 1] = -lo
            x = sla.solve_banded((1, 1), ab, b[:kc, L])
            np.testing.assert_allclose(c["con"][:kc, L], x, rtol=1e-12)
            # column mass
            assert abs(np.sum(dzc * c["con"][:kc, L]) - np.sum(dzc * b[:kc, L])) <= \
                1e-13 * np.sum(dzc * b[:kc, L])


def test_dye_mass_equals_load_times_time(oracle_mod):
    from efdc_amd import grids
    case = grids.tidal_basin(20, 10, 3, seed=12)
    case.open_bc = None
    case.conc_bc = None
    o = oracle_mod.Oracle(case, rsqm=1e-22, iterm=500)
    for n in range(1, 21):
        o.step(n)
    _, m = o.balance()
    assert abs(m[1] - 20 * case.dt * 1.0) <= 1e-9 * 100


@pytest.mark.parametrize("rk", [2e-4, -5e-5, 1000.0])
def test_dye_decay_closed_forms(oracle_mod, rk):
    """No flow, no sources, no vertical exchange: CALCONC changes the dye
    only through the decay/age step."""
    from efdc_amd import grids
    case = grids.transport_case(24, 12, 3, seed=5)
    case.sources = None
    for name in ("UHDY2", "VHDX2", "QSUM", "AB", "U2", "V2"):
        case.k[name][:] = 0.0
    case.w["W2"][:] = 0.0
    case.k["H1PK"][:] = case.k["HPK"]
    dye = case.consts[2]
    assert dye["is_dye"] == 1
    rng = np.random.default_rng(4)
    kc, L = case.grid.KC, wet_index(case)
    dye["con"][:kc, L] = rng.uniform(0.5, 2.0, (kc, len(L)))
    dye["con1"][:] = dye["con"]
    dye["rkdye"] = rk
    c0 = dye["con"][:kc, L].copy()
    o = oracle_mod.Oracle(case)
    n = 10
    for it in range(1, n + 1):
        o.calconc(niter=it)
    got = case.consts[2]["con"][:kc, L]
    if rk == 1000.0:
        want = c0 + n * (case.dt / 86400.0)
    elif rk < 0:
        want = c0 * np.exp(-rk * case.dt) ** n
    else:
        want = c0 * (1.0 / (1.0 + case.dt * rk)) ** n
    np.testing.assert_allclose(got, want, rtol=1e-13)


# ---------------------------------------------------------- failure paths

# ------------------------------------------------------------- settling

def _settling_case(kc=6, ws=2e-3, probdep=0.0, wese=0.0, seed=8):
    """A transport case whose sediment class settles; random non-negative
    and a few negative concentrations (the sweep clips them to zero)."""
    from efdc_amd import grids
    case = grids.transport_case(12, 8, kc, seed=seed, sed_ws=ws, sed_probdep=probdep,
                                sed_wese=wese, dt=600.0)
    rng = np.random.default_rng(seed)
    L = wet_index(case)
    sed = case.consts[3]
    sed["con"][:kc, L] = rng.uniform(-0.02, 0.5, (kc, len(L)))
    # layer thicknesses vary by column and layer
    case.k["HPKI"][:kc, L] = 1.0 / rng.uniform(0.5, 3.0, (kc, len(L)))
    return case, L


def test_settling_conserves_mass_with_closed_bed(oracle_mod):
    """No deposition, no resuspension: the implicit upwind fluxes only move
    mass downwards, so sum_K SED*H_K equals sum_K max(SED,0)*H_K."""
    case, L = _settling_case()
    kc = case.grid.KC
    H = 1.0 / case.k["HPKI"][:kc, L]
    before = np.sum(np.maximum(case.consts[3]["con"][:kc, L], 0.0) * H, axis=0)
    o = oracle_mod.Oracle(case)
    o.settle()
    after = case.consts[3]["con"][:kc, L]
    assert np.all(after >= 0.0)
    np.testing.assert_allclose(np.sum(after * H, axis=0), before, rtol=1e-13, atol=0)


def test_settling_equals_bidiagonal_solve(oracle_mod):
    """The sweep is the solution of the implicit upwind system
    (H_K/dt) (c_K - c0_K) = w c_{K+1} - w_K c_K per column (w_K = ws above
    the bed, probdep*ws at the bed, plus the resuspension flux into the
    bottom layer), solved here as a dense linear system."""
    ws, pd, we = 2e-3, 0.6, 1e-5
    case, L = _settling_case(ws=ws, probdep=pd, wese=we)
    kc, dt = case.grid.KC, case.dt
    H = 1.0 / case.k["HPKI"][:kc, L]
    c0 = np.maximum(case.consts[3]["con"][:kc, L], 0.0)
    o = oracle_mod.Oracle(case)
    o.settle()
    got = case.consts[3]["con"][:kc, L]
    for j in range(0, len(L), 7):
        A = np.zeros((kc, kc))
        b = c0[:, j] * H[:, j] / dt
        for K in range(kc):
            A[K, K] = H[K, j] / dt + (ws if K > 0 else pd * ws)
            if K + 1 < kc:
                A[K, K + 1] = -ws
        b[0] += we
        want = np.linalg.solve(A, b)
        np.testing.assert_allclose(got[:, j], want, rtol=1e-12, atol=1e-15)


def test_settling_without_deposition_collects_mass_at_the_bed(oracle_mod):
    case, L = _settling_case(ws=5e-3)
    kc = case.grid.KC
    H = 1.0 / case.k["HPKI"][:kc, L]
    total = np.sum(np.maximum(case.consts[3]["con"][:kc, L], 0.0) * H, axis=0)
    o = oracle_mod.Oracle(case)
    for _ in range(400):
        o.settle()
    c = case.consts[3]["con"][:kc, L]
    assert np.max(c[1:]) <= 1e-6 * np.max(c[0])
    np.testing.assert_allclose(c[0] * H[0], total, rtol=1e-9)


def test_cg_maximum_iterations_status(oracle_mod):
    case = lake(24, seed=1)
    o = oracle_mod.Oracle(case, rsqm=1e-30, iterm=2)
    st = o.calpuv2c(n=1)
    assert st["status"] == 3        # EFDC_ERR_CG_MAXITER


def test_negative_depth_status(oracle_mod):
    case = lake(24, seed=1)
    i = wet_index(case)
    case.h["QSUME"][i[len(i) // 2]] = -1e5   # a sink far beyond the cell volume
    o = oracle_mod.Oracle(case, rsqm=1e-20, iterm=500)
    st = o.calpuv2c(n=1)
    assert st["status"] == 4        # EFDC_ERR_NEGDEPTH
