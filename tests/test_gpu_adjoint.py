"""Discrete adjoint of the hot-path step (efdc_adj_*) on a small tidal basin.

The misfit J is the squared difference of layer-mean dye at observed cells
(next to three dye sources, every 4 steps) from synthetic data.  Checks:

* loads: J is quadratic in the loads (the dye is linear in them and does
  not act on the flow), so a central difference of any size is exact up to
  rounding; the adjoint's directional derivative must match it to 1e-8,
  with the forward model on the GPU and with the FP64 CPU oracle (the CPU
  tangent of J along the direction);
* Manning n: J depends on n through the bottom drag, the external solve,
  the transports and the advection; central differences with a relative
  step of 1e-3 (O(h^2) truncation) must match to the 1e-4 bar, on the GPU
  and with the CPU oracle;
* the checkpoint interval changes nothing (same bits), nor does keeping
  each step's recording instead of running the step again;
* torch.autograd.gradcheck of the autograd.Function wrapper;
* the FP32-storage forward gives the FP64-storage gradient to 1e-3;
* with settling on the observed constituent (the implicit sweep of
  EFDC/calsed.f90:232-256 with deposition and resuspension): the load
  gradient against central differences of the GPU model and of the CPU
  oracle, and the Manning gradient (through HPKI of the sweep) against
  central differences;
* with anti-diffusion, with and without FCT (EFDC/caltran.f90:473-786), on
  a dye block whose edges the limiter has to clip (Courant numbers ~0.3; J
  changes 20-fold with the limiter): the same checks, the Manning one with a
  step small enough for the strongly nonlinear J.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NSRC = 3


def _setup(nsteps=24, seed=7, kc=3, nx=24, ny=12, dt=5.0):
    from efdc_amd import grids
    base = grids.tidal_basin(nx, ny, kc, seed=seed, dt=dt)
    g = base.grid
    rng = np.random.default_rng(seed)
    west = set(int(x) for x in base.open_bc["side"][0]["L"])
    inner = [L for L in range(2, g.LA + 1) if L not in west and g.IL[L - 1] > 3]
    src = np.sort(rng.choice(inner, size=NSRC, replace=False)).astype(np.int32)
    obs = []
    for L in src:
        obs += [L, g.maps["LEC"][L - 1], g.maps["LNC"][L - 1], g.maps["LWC"][L - 1]]
    obs = np.array(sorted(set(int(x) for x in obs if x != g.LC)), np.int32)
    case = grids.with_dye_sources(base, src)
    true = rng.uniform(0.0, 5.0, (nsteps, NSRC))
    loads0 = rng.uniform(0.0, 5.0, (nsteps, NSRC))
    return case, obs, true, loads0, rng


def _problem(case, obs, true, rng, nsteps=24, **kw):
    from efdc_amd import adjoint
    prob = adjoint.DyeMisfit(case, nsteps=nsteps, obs_L=obs, obs_every=4, **kw)
    y = prob.observations(true)
    prob.data = y * (1.0 + 0.01 * rng.standard_normal(y.shape))
    return prob


def _cpu_misfit(case, obs, loads, nman, data, nsteps, obs_every=4, dye=1):
    from oracle import oracle as orc
    c = case.copy()
    c.manning_n = np.array(nman, np.float64)
    o = orc.Oracle(c, rsqm=0.0, iterm=4000, rel_tol=1e-13)
    kc = c.grid.KC
    q = np.asarray(case.sources["qss"]).reshape(NSRC, kc).sum(axis=1)
    cqs = o.src_cqs.reshape(NSRC, kc, -1)
    dz = c.k["DZC"][:kc, obs - 1]
    J = 0.0
    for n in range(1, nsteps + 1):
        cqs[:, :, dye] = (loads[n - 1] / q)[:, None]
        o.step(n)
        if n % obs_every == 0:
            y = np.sum(dz * c.consts[dye]["con"][:kc, obs - 1], axis=0)
            J += float(np.sum((y - data[n // obs_every - 1]) ** 2))
    return J


@pytest.fixture(scope="module")
def prob_fp64():
    case, obs, true, loads0, rng = _setup()
    prob = _problem(case, obs, true, rng, ckpt=5)
    J, gl, gn, st = prob.gradient(loads0)
    yield dict(prob=prob, case=case, obs=obs, loads0=loads0, J=J, gl=gl, gn=gn, st=st,
               rng=rng)
    prob.close()


def test_load_gradient_vs_central_difference(prob_fp64):
    p = prob_fp64
    prob, l0 = p["prob"], p["loads0"]
    assert p["J"] > 0 and np.any(p["gl"] != 0)
    for trial in range(2):
        d = p["rng"].standard_normal(l0.shape)
        fd = (prob.misfit(l0 + d) - prob.misfit(l0 - d)) / 2.0
        ad = float(np.sum(p["gl"] * d))
        print(f"loads: adjoint {ad:.12e} fd {fd:.12e} rel {abs(ad - fd) / abs(fd):.2e}")
        assert abs(ad - fd) <= 1e-8 * abs(fd)
    # single entries: the last load of a source before an observation
    for s in range(NSRC):
        e = np.zeros_like(l0)
        e[19, s] = 1.0
        fd = (prob.misfit(l0 + e) - prob.misfit(l0 - e)) / 2.0
        assert abs(p["gl"][19, s] - fd) <= 1e-8 * max(abs(fd), 1e-30) + 1e-300


def test_manning_gradient_vs_central_difference(prob_fp64):
    p = prob_fp64
    prob, case, l0 = p["prob"], p["case"], p["loads0"]
    n0 = case.manning_n.copy()
    g = case.grid
    L = np.arange(2, g.LA + 1) - 1
    d = np.zeros(g.LCM)
    d[L] = p["rng"].uniform(-1.0, 1.0, len(L)) * n0[L]
    h = 1e-3
    fd = (prob.misfit(l0, n0 + h * d) - prob.misfit(l0, n0 - h * d)) / (2 * h)
    ad = float(np.sum(p["gn"] * d))
    print(f"manning: adjoint {ad:.10e} fd {fd:.10e} rel {abs(ad - fd) / abs(fd):.2e}")
    assert abs(fd) > 0
    assert abs(ad - fd) <= 1e-4 * abs(fd)


def test_gradients_vs_cpu_oracle_tangent(prob_fp64):
    """The FP64 oracle run of the same forward model: its central
    difference along a load direction is the exact tangent of J (J is
    quadratic in the loads); along a Manning direction it is the tangent up
    to O(h^2)."""
    p = prob_fp64
    prob, case, l0, obs = p["prob"], p["case"], p["loads0"], p["obs"]
    n0 = case.manning_n.copy()
    data = prob.data
    d = p["rng"].standard_normal(l0.shape)
    fd = (_cpu_misfit(case, obs, l0 + d, n0, data, 24) -
          _cpu_misfit(case, obs, l0 - d, n0, data, 24)) / 2.0
    ad = float(np.sum(p["gl"] * d))
    print(f"cpu loads: adjoint {ad:.12e} oracle {fd:.12e}")
    assert abs(ad - fd) <= 1e-7 * abs(fd)
    g = case.grid
    L = np.arange(2, g.LA + 1) - 1
    dn = np.zeros(g.LCM)
    dn[L] = p["rng"].uniform(-1.0, 1.0, len(L)) * n0[L]
    h = 1e-3
    fdn = (_cpu_misfit(case, obs, l0, n0 + h * dn, data, 24) -
           _cpu_misfit(case, obs, l0, n0 - h * dn, data, 24)) / (2 * h)
    adn = float(np.sum(p["gn"] * dn))
    print(f"cpu manning: adjoint {adn:.10e} oracle {fdn:.10e}")
    assert abs(adn - fdn) <= 1e-4 * abs(fdn)


def test_checkpoint_interval_changes_nothing(prob_fp64):
    p = prob_fp64
    prob = p["prob"]
    prob.ckpt = 24
    J, gl, gn, _ = prob.gradient(p["loads0"])
    prob.ckpt = 5
    assert J == p["J"]
    assert np.array_equal(gl, p["gl"]) and np.array_equal(gn, p["gn"])


def test_kept_recordings_equal_rerun_steps(prob_fp64):
    """Reversing each step from the state after it and its saved recording
    (efdc_adj_record_save/load) gives the bits of reversing it right after
    running it again from the state before it."""
    p = prob_fp64
    J, gl, gn, _ = p["prob"].gradient(p["loads0"], rerun_steps=True)
    assert J == p["J"]
    assert np.array_equal(gl, p["gl"]) and np.array_equal(gn, p["gn"])


def test_fp32_storage_forward_gradient():
    from efdc_amd import lib
    case, obs, true, loads0, rng = _setup()
    p64 = _problem(case, obs, true, np.random.default_rng(3), ckpt=6)
    data = p64.data
    _, gl64, gn64, _ = p64.gradient(loads0)
    p64.close()
    from efdc_amd import adjoint
    p32 = adjoint.DyeMisfit(case, nsteps=24, obs_L=obs, obs_every=4, ckpt=6,
                            store=lib.STORE_FP32)
    p32.data = data
    _, gl32, gn32, _ = p32.gradient(loads0)
    p32.close()
    el = np.max(np.abs(gl32 - gl64)) / np.max(np.abs(gl64))
    en = np.max(np.abs(gn32 - gn64)) / np.max(np.abs(gn64))
    print("fp32 storage: load grad rel", el, "manning grad rel", en)
    assert el <= 1e-3 and en <= 1e-2


def test_torch_gradcheck_of_autograd_function():
    import torch
    from efdc_amd import adjoint
    case, obs, true, loads0, rng = _setup(nsteps=8)
    prob = _problem(case, obs, true[:8], rng, nsteps=8, ckpt=3)
    cidx = adjoint.coarse_index(case.grid, 2, 2)
    f = adjoint.autograd_function()
    loads = torch.tensor(loads0[:8], dtype=torch.float64, requires_grad=True)
    nc = torch.full((2, 2), 0.025, dtype=torch.float64, requires_grad=True)
    J = f.apply(loads, nc, prob, cidx)
    J.backward()
    assert loads.grad is not None and nc.grad is not None
    scale = float(J.detach()) + 1e-30
    ok = torch.autograd.gradcheck(lambda a, b: f.apply(a, b, prob, cidx), (loads, nc),
                                  eps=1e-6, atol=1e-6 * scale, rtol=1e-4,
                                  nondet_tol=0.0, fast_mode=False)
    assert ok
    prob.close()


def test_seed_is_ordered_after_pending_work():
    """efdc_adj_seed_con called while the forward step's transport kernels
    are still queued on the context's (non-blocking) stream: the seed list
    must reach the device after everything queued before it, so exactly
    the listed entries of lambda change."""
    from efdc_amd import grids, lib
    base = grids.tidal_basin(160, 160, 3, seed=3)
    g = base.grid
    case = grids.with_dye_sources(base, [int(g.LIJ[80, 80])])
    ctx = lib.Context(g).upload_case(case)
    ctx.adj_enable([1])
    ctx.adj_reset()
    L = np.array([int(g.LIJ[40, 40]), int(g.LIJ[100, 120]), int(g.LIJ[40, 40])], np.int32)
    K = np.array([1, 3, 1], np.int32)
    val = np.array([1.5, -2.0, 0.25])
    for n in (1, 2):
        ctx.step(case.dt, n)              # returns with CALCONC still queued
        ctx.adj_seed_con(1, L, K, val)
    lam = ctx.adj_get(lib.ADJ_CON + 1)
    want = np.zeros_like(lam)
    for l_, k_, v_ in zip(L, K, val):
        want[k_ - 1, l_ - 1] += 2 * v_
    assert np.array_equal(lam, want)
    ctx.close()


def _settling_case():
    case, obs, true, loads0, rng = _setup(seed=11)
    d = case.consts[1]
    d["wset"], d["probdep"], d["wese"] = 1e-3, 0.5, 1e-5
    return case, obs, true, loads0, rng


@pytest.fixture(scope="module")
def prob_settle():
    case, obs, true, loads0, rng = _settling_case()
    prob = _problem(case, obs, true, rng, ckpt=6)
    J, gl, gn, st = prob.gradient(loads0)
    yield dict(prob=prob, case=case, obs=obs, loads0=loads0, J=J, gl=gl, gn=gn, rng=rng)
    prob.close()


def test_settling_changes_the_misfit(prob_settle):
    """The settling acts: the same loads without it give another J."""
    p = prob_settle
    case = p["case"].copy()
    case.consts[1]["wset"] = 0.0
    from efdc_amd import adjoint
    q = adjoint.DyeMisfit(case, nsteps=24, obs_L=p["obs"], obs_every=4, ckpt=6)
    q.data = p["prob"].data
    J0 = q.misfit(p["loads0"])
    q.close()
    print("J with settling", p["J"], "without", J0)
    assert abs(J0 - p["J"]) > 1e-3 * abs(p["J"])


def test_settling_load_gradient_vs_central_differences(prob_settle):
    """The settled dye is affine in the loads (no clip is active where the
    dye is positive), so J is quadratic and a central difference is exact
    up to rounding, on the GPU model and on the FP64 CPU oracle."""
    p = prob_settle
    prob, l0 = p["prob"], p["loads0"]
    d = p["rng"].standard_normal(l0.shape)
    fd = (prob.misfit(l0 + d) - prob.misfit(l0 - d)) / 2.0
    ad = float(np.sum(p["gl"] * d))
    print(f"settling loads: adjoint {ad:.12e} fd {fd:.12e}")
    assert abs(ad - fd) <= 1e-8 * abs(fd)
    n0 = p["case"].manning_n.copy()
    fdc = (_cpu_misfit(p["case"], p["obs"], l0 + d, n0, prob.data, 24) -
           _cpu_misfit(p["case"], p["obs"], l0 - d, n0, prob.data, 24)) / 2.0
    print(f"settling loads: oracle {fdc:.12e}")
    assert abs(ad - fdc) <= 1e-7 * abs(fdc)


def test_settling_manning_gradient_vs_central_difference(prob_settle):
    p = prob_settle
    prob, case, l0 = p["prob"], p["case"], p["loads0"]
    n0 = case.manning_n.copy()
    g = case.grid
    L = np.arange(2, g.LA + 1) - 1
    d = np.zeros(g.LCM)
    d[L] = p["rng"].uniform(-1.0, 1.0, len(L)) * n0[L]
    h = 1e-3
    fd = (prob.misfit(l0, n0 + h * d) - prob.misfit(l0, n0 - h * d)) / (2 * h)
    ad = float(np.sum(p["gn"] * d))
    print(f"settling manning: adjoint {ad:.10e} fd {fd:.10e}")
    assert abs(fd) > 0
    assert abs(ad - fd) <= 1e-4 * abs(fd)


def _adac_case(isfct):
    """Anti-diffusion on the observed dye, which starts as a block of 4
    units in the middle of the basin, so that the antidiffusive fluxes
    overshoot at its edges and the FCT limiter acts."""
    case, obs, true, loads0, rng = _setup(seed=5, dt=60.0)
    g = case.grid
    d = case.consts[1]
    d["isadac"], d["isfct"] = 1, isfct
    L = np.arange(2, g.LA + 1)
    I, J = g.IL[L - 1], g.JL[L - 1]
    inb = (I >= 8) & (I <= 15) & (J >= 4) & (J <= 8)
    blk = L[inb]
    d["con"][:g.KC, blk - 1] = 4.0
    d["con1"][:g.KC, blk - 1] = 4.0
    # the cells around the block are observed too
    ring = L[(I >= 7) & (I <= 16) & (J >= 3) & (J <= 9) & ~inb]
    obs = np.array(sorted(set(obs.tolist()) | set(int(x) for x in ring)), np.int32)
    return case, obs, true, loads0, rng


@pytest.fixture(scope="module", params=[0, 1], ids=["adac", "adac_fct"])
def prob_adac(request):
    case, obs, true, loads0, rng = _adac_case(request.param)
    prob = _problem(case, obs, true, rng, ckpt=6)
    J, gl, gn, st = prob.gradient(loads0)
    yield dict(prob=prob, case=case, obs=obs, loads0=loads0, J=J, gl=gl, gn=gn, rng=rng,
               isfct=request.param)
    prob.close()


def test_adac_load_gradient_vs_central_differences(prob_adac):
    """Anti-diffusion (and FCT) make the dye nonlinear in the loads: the
    central difference of the GPU model with steps of 1e-3 of the loads
    (O(h^2) truncation, no MAX/MIN branch changes within the step) and the
    same difference of the FP64 CPU oracle must match the adjoint's
    directional derivative."""
    p = prob_adac
    prob, l0 = p["prob"], p["loads0"]
    assert np.any(p["gl"] != 0)
    h = 1e-3
    for trial in range(2):
        d = p["rng"].standard_normal(l0.shape)
        fd = (prob.misfit(l0 + h * d) - prob.misfit(l0 - h * d)) / (2 * h)
        ad = float(np.sum(p["gl"] * d))
        print(f"adac isfct={p['isfct']} loads: adjoint {ad:.12e} fd {fd:.12e} "
              f"rel {abs(ad - fd) / abs(fd):.2e}")
        assert abs(ad - fd) <= 1e-5 * abs(fd)
    n0 = p["case"].manning_n.copy()
    fdc = (_cpu_misfit(p["case"], p["obs"], l0 + h * d, n0, prob.data, 24) -
           _cpu_misfit(p["case"], p["obs"], l0 - h * d, n0, prob.data, 24)) / (2 * h)
    print(f"adac isfct={p['isfct']} loads: oracle {fdc:.12e}")
    assert abs(ad - fdc) <= 1e-5 * abs(fdc)


def test_fct_limiter_acts(prob_adac):
    """The limiter changes the misfit (so the FCT terms of the reverse are
    exercised by the gradient checks)."""
    p = prob_adac
    if not p["isfct"]:
        pytest.skip("the FCT case compares itself with the plain one")
    from efdc_amd import adjoint
    case, obs, true, loads0, rng = _adac_case(0)
    q = adjoint.DyeMisfit(case, nsteps=24, obs_L=obs, obs_every=4, ckpt=6)
    q.data = p["prob"].data
    J0 = q.misfit(p["loads0"])
    q.close()
    print("J with FCT", p["J"], "without", J0)
    assert abs(J0 - p["J"]) > 1e-3 * abs(p["J"])


def test_adac_manning_gradient_vs_central_difference(prob_adac):
    p = prob_adac
    prob, case, l0 = p["prob"], p["case"], p["loads0"]
    n0 = case.manning_n.copy()
    g = case.grid
    L = np.arange(2, g.LA + 1) - 1
    d = np.zeros(g.LCM)
    d[L] = p["rng"].uniform(-1.0, 1.0, len(L)) * n0[L]
    # the central difference converges as h^2 to the adjoint (1e-5 of it
    # at h = 1e-3, 1e-7 at 1e-4, 1e-8 at 1e-5, with and without FCT): the
    # anti-diffusion makes J strongly nonlinear in n, so the step is small
    h = 1e-5
    fd = (prob.misfit(l0, n0 + h * d) - prob.misfit(l0, n0 - h * d)) / (2 * h)
    ad = float(np.sum(p["gn"] * d))
    print(f"adac isfct={p['isfct']} manning: adjoint {ad:.10e} fd {fd:.10e} "
          f"rel {abs(ad - fd) / abs(fd):.2e}")
    assert abs(fd) > 0
    assert abs(ad - fd) <= 1e-6 * abs(fd)
