"""The two launch strategies of CONGRAD against the oracle.

* multi-kernel: two launches per iteration (EFDC_CG_KERNEL_MULTI);
* persistent: one cooperative launch per solve with the rows held in
  registers (EFDC_CG_KERNEL_PERSISTENT).

Both run the FP64 Jacobi-PCG of the reference (|d eta| <= 1e-9 m against
the oracle at RSQM=1e-24, iteration count within one) and the FP32-refined
solve (|d eta| <= 1e-6 m).  The persistent solve is also run on a
column-major renumbering of the cells (neighbour maps that are not the
row-major L+-1 pattern, so the general four-map operator is used), on a
grid so small that one workgroup holds every row, and on a grid whose
north/south neighbours lie beyond the workgroup's LDS window (the
automatic choice then falls back to the multi-kernel solve, the forced
persistent choice is refused).  The persistent mixed solve runs in both its
forms (Chronopoulos-Gear and pipelined).
"""
import dataclasses

import numpy as np
import pytest

from conftest import eta_err, wet_index

pytestmark = pytest.mark.gpu

TIGHT = 1e-24
SYSTEM = ["CCC", "CCS", "CCW", "CCE", "CCN", "CCCI", "FPTMP", "P"]


def _lake_system(oracle_mod, n, seed):
    from efdc_amd import grids
    wet = grids.lake_mask(n, n, seed)
    case = grids.external_case(wet, seed=seed)
    ref = case.copy()
    o = oracle_mod.Oracle(ref, rsqm=TIGHT, iterm=4000)
    o.assemble()
    sys0 = {k: ref.h[k].copy() for k in SYSTEM}
    o.congrad()
    return case, sys0, ref.h["P"].copy()


def _solve(grid, system, pipeline=None, **kw):
    from efdc_amd.lib import Context
    kw.setdefault("RSQM", TIGHT)
    kw.setdefault("ITERM", 4000)
    ctx = Context(grid, **kw)
    if pipeline is not None:
        ctx.set_cg_pipeline(pipeline)
        assert ctx.cg_pipeline() == pipeline
    for k, v in system.items():
        ctx.set(k, v)
    out = ctx.congrad()
    p = ctx.get("P")
    ctx.close()
    return out, p


@pytest.mark.parametrize("kernel", ["multi", "persistent"])
@pytest.mark.parametrize("n,seed", [(24, 3), (64, 5), (200, 7)])
def test_fp64_strategies_match_oracle(oracle_mod, kernel, n, seed):
    from efdc_amd import lib
    ck = {"multi": lib.CG_KERNEL_MULTI, "persistent": lib.CG_KERNEL_PERSISTENT}[kernel]
    case, system, p_ref = _lake_system(oracle_mod, n, seed)
    out, p = _solve(case.grid, system, cg_kernel=ck)
    i = wet_index(case)
    d = eta_err(p[i], p_ref[i])
    print(f"{kernel} n={n} wet={case.grid.nwet} iters {out['iters']} deta {d:.3e}")
    assert d <= 1e-9


@pytest.mark.parametrize("kernel", ["multi", "persistent"])
def test_mixed_strategies_within_1e6_m(oracle_mod, kernel):
    from efdc_amd import lib
    ck = {"multi": lib.CG_KERNEL_MULTI, "persistent": lib.CG_KERNEL_PERSISTENT}[kernel]
    case, system, p_ref = _lake_system(oracle_mod, 160, 9)
    out, p = _solve(case.grid, system, cg_kernel=ck, cg_mode=lib.CG_MIXED,
                    RSQM=1e-20)
    i = wet_index(case)
    d = eta_err(p[i], p_ref[i])
    print(f"mixed {kernel}: outer {out['outer']} inner {out['iters']} deta {d:.3e}")
    assert out["outer"] >= 1
    assert d <= 1e-6


def test_persistent_equals_multi_iterations(oracle_mod):
    """Same FP64 recurrences: the two launch strategies take the same
    number of iterations (up to one, from the order of the dot sums)."""
    from efdc_amd import lib
    case, system, _ = _lake_system(oracle_mod, 128, 12)
    om, pm = _solve(case.grid, system, cg_kernel=lib.CG_KERNEL_MULTI)
    op, pp = _solve(case.grid, system, cg_kernel=lib.CG_KERNEL_PERSISTENT)
    i = wet_index(case)
    assert abs(om["iters"] - op["iters"]) <= 1
    assert eta_err(pm[i], pp[i]) <= 1e-9


def _column_major(case, system):
    """The same cells numbered column by column (I outer, J inner)."""
    from efdc_amd.grids import Grid
    g = case.grid
    old = np.arange(2, g.LA + 1)
    key = np.lexsort((g.JL[old - 1], g.IL[old - 1]))   # I major, J minor
    new_of_old = np.zeros(g.LCM + 1, dtype=np.int64)
    new_of_old[old[key]] = np.arange(2, g.LA + 1)
    LC = g.LC
    new_of_old[1] = 1
    new_of_old[LC] = LC
    maps = {}
    for name, m in g.maps.items():
        nm = np.zeros_like(m)
        nm[0] = m[0]
        nm[g.LA] = m[g.LA]
        nm[LC] = m[LC]
        src = old[key]                       # old L in new order
        nm[np.arange(2, g.LA + 1) - 1] = new_of_old[m[src - 1]]
        maps[name] = nm.astype(np.int32)
    g2 = dataclasses.replace(g, maps=maps)
    sys2 = {}
    for k, v in system.items():
        w = v.copy()
        w[np.arange(2, g.LA + 1) - 1] = v[old[key] - 1]
        sys2[k] = w
    return g2, sys2, old[key]


def test_persistent_general_maps(oracle_mod):
    from efdc_amd import lib
    case, system, p_ref = _lake_system(oracle_mod, 96, 4)
    g2, sys2, src = _column_major(case, system)
    assert not np.all((g2.maps["LWC"][1:g2.LA] == np.arange(1, g2.LA)) |
                      (g2.maps["LWC"][1:g2.LA] == g2.LC))
    out, p = _solve(g2, sys2, cg_kernel=lib.CG_KERNEL_PERSISTENT)
    newi = np.arange(2, g2.LA + 1) - 1
    d = eta_err(p[newi], p_ref[src - 1])
    print("column-major persistent: iters", out["iters"], "deta", d)
    assert d <= 1e-9


@pytest.mark.parametrize("n,seed,deg", [(12, 2, 3), (60, 4, 1), (160, 9, 3), (300, 11, 2)])
def test_mixed_persistent_forms(oracle_mod, n, seed, deg):
    """The Chronopoulos-Gear form (1) and its pipelined form (2) of the
    persistent mixed solve: both within 1e-6 m of the FP64 oracle, with
    inner iteration counts that differ by rounding only (the iterates are
    the same in exact arithmetic); one workgroup, several, and different
    preconditioner degrees."""
    from efdc_amd import lib
    case, system, p_ref = _lake_system(oracle_mod, n, seed)
    i = wet_index(case)
    res = {}
    for form in (1, 2):
        out, p = _solve(case.grid, system, pipeline=form, cg_kernel=lib.CG_KERNEL_PERSISTENT,
                        cg_mode=lib.CG_MIXED, RSQM=1e-20, mixed_precond_degree=deg)
        d = eta_err(p[i], p_ref[i])
        print(f"n={n} wet={case.grid.nwet} form {form}: outer {out['outer']} "
              f"inner {out['iters']} deta {d:.3e}")
        assert d <= 1e-6
        res[form] = out["iters"]
    assert abs(res[1] - res[2]) <= max(2, res[1] // 10)


def test_persistent_single_workgroup(oracle_mod):
    from efdc_amd import lib
    case, system, p_ref = _lake_system(oracle_mod, 12, 2)
    assert case.grid.nwet < 512
    for mode in (lib.CG_FP64, lib.CG_MIXED):
        out, p = _solve(case.grid, system, cg_kernel=lib.CG_KERNEL_PERSISTENT,
                        cg_mode=mode, RSQM=TIGHT if mode == lib.CG_FP64 else 1e-22)
        i = wet_index(case)
        d = eta_err(p[i], p_ref[i])
        print("single workgroup mode", mode, "iters", out["iters"], "deta", d)
        assert d <= (1e-9 if mode == lib.CG_FP64 else 1e-6)


def test_wide_rows_fall_back_or_refuse(oracle_mod):
    """Two wet rows of 4200 cells: a north neighbour lies 4200 cells away,
    beyond the persistent solve's window."""
    from efdc_amd import grids, lib
    wet = np.zeros((4, 4202), dtype=bool)
    wet[1:3, 1:4201] = True
    case = grids.external_case(wet, seed=1)
    ref = case.copy()
    o = oracle_mod.Oracle(ref, rsqm=TIGHT, iterm=4000)
    o.assemble()
    system = {k: ref.h[k].copy() for k in SYSTEM}
    o.congrad()
    out, p = _solve(case.grid, system, cg_kernel=lib.CG_KERNEL_AUTO)
    i = wet_index(case)
    assert eta_err(p[i], ref.h["P"][i]) <= 1e-9
    with pytest.raises(lib.EfdcError) as e:
        _solve(case.grid, system, cg_kernel=lib.CG_KERNEL_PERSISTENT)
    assert e.value.status == 1


def _solve_launch(grid, system, mode, spin_limit=0, **kw):
    from efdc_amd.lib import Context
    kw.setdefault("RSQM", TIGHT)
    kw.setdefault("ITERM", 4000)
    ctx = Context(grid, **kw)
    ctx.set_cg_launch(mode, spin_limit)
    for k, v in system.items():
        ctx.set(k, v)
    out = ctx.congrad()
    p = ctx.get("P")
    launch = ctx.cg_launch()
    ctx.close()
    return out, p, launch


@pytest.mark.parametrize("mixed", [False, True])
def test_plain_and_cooperative_launch_same_bits(oracle_mod, mixed):
    """The persistent solve launched plainly (the default) and cooperatively
    runs the same kernel on the same grid: the same P, bit for bit."""
    from efdc_amd import lib
    case, system, p_ref = _lake_system(oracle_mod, 200, 7)
    kw = dict(cg_kernel=lib.CG_KERNEL_PERSISTENT)
    if mixed:
        kw.update(cg_mode=lib.CG_MIXED, RSQM=1e-20)
    o0, p0, l0 = _solve_launch(case.grid, system, 0, **kw)
    o1, p1, l1 = _solve_launch(case.grid, system, 1, **kw)
    print("plain", o0["iters"], l0, "cooperative", o1["iters"], l1)
    assert l0[0] == 0 and l1 == (1, 0)
    assert o0["iters"] == o1["iters"]
    i = wet_index(case)
    np.testing.assert_array_equal(p0[i], p1[i])
    assert eta_err(p0[i], p_ref[i]) <= (1e-6 if mixed else 1e-9)


@pytest.mark.parametrize("mixed", [False, True])
def test_plain_launch_repeated_after_its_waits_give_up(oracle_mod, mixed):
    """A spin limit of one polling round makes the plain launch's waits give
    up as soon as a workgroup is not there yet (what a grid that is not
    co-resident would see); the solve is then repeated cooperatively from
    the P the first launch reached and meets the same bar against the
    oracle."""
    from efdc_amd import lib
    case, system, p_ref = _lake_system(oracle_mod, 200, 7)
    kw = dict(cg_kernel=lib.CG_KERNEL_PERSISTENT)
    if mixed:
        kw.update(cg_mode=lib.CG_MIXED, RSQM=1e-20)
    out, p, launch = _solve_launch(case.grid, system, 0, spin_limit=1, **kw)
    i = wet_index(case)
    d = eta_err(p[i], p_ref[i])
    print(f"mixed={mixed}: launch {launch} iters {out['iters']} deta {d:.3e}")
    assert launch[1] >= 1
    assert d <= (1e-6 if mixed else 1e-9)


@pytest.mark.parametrize("mixed", [False, True])
def test_step_repeats_the_solve_and_the_passes_after_it(oracle_mod, mixed):
    """The same inside CALPUV2C steps, where the abort is read after the
    step: the solve and the passes after it are run again, and P, HP and the
    transports meet the oracle's bars of the parity tests (1e-9; 1e-6 for
    the FP32-refined solve)."""
    from efdc_amd import grids, lib
    # enough cells for about a hundred workgroups, whose arrivals at an
    # exchange are spread more than one polling round apart
    case = grids.tidal_basin(320, 160, 1, seed=11)
    ref = case.copy()
    o = oracle_mod.Oracle(ref, rsqm=TIGHT, iterm=2000)
    kw = dict(cg_mode=lib.CG_MIXED, RSQM=1e-20) if mixed else dict(RSQM=TIGHT)
    ctx = lib.Context(case.grid, ITERM=2000, cg_kernel=lib.CG_KERNEL_PERSISTENT,
                      **kw).upload_case(case)
    ctx.set_cg_launch(0, 1)
    bar = 1e-6 if mixed else 1e-9
    i = wet_index(case)
    for n in (1, 2, 3):
        o.step(n)
        ctx.step(case.dt, n)
        d = eta_err(ctx.get("P")[i], ref.h["P"][i])
        print(f"mixed={mixed} step {n}: launch {ctx.cg_launch()} deta {d:.3e}")
        assert d <= bar
        for name in ("HP", "UHDYE", "VHDXE", "HU", "HV"):
            a, b = ctx.get(name)[i], ref.h[name][i]
            assert np.max(np.abs(a - b)) <= bar * max(1.0, np.max(np.abs(b))), name
        # keep the two runs on the same inputs for the next step
        for name in ["P", "HP", "UHDYE", "VHDXE", "HU", "HV", "HUI", "HVI",
                     "HPI", "UHE", "VHE", "UHDY1E", "VHDX1E", "H1P"]:
            ctx.set(name, ref.h[name])
    assert ctx.cg_launch()[1] >= 1
    ctx.close()


def _solve64(grid, system, pipelined, **kw):
    from efdc_amd.lib import Context
    from efdc_amd import lib
    kw.setdefault("RSQM", TIGHT)
    kw.setdefault("ITERM", 4000)
    ctx = Context(grid, cg_kernel=lib.CG_KERNEL_PERSISTENT, **kw)
    ctx.set_cg64_pipeline(pipelined)
    assert ctx.cg64_pipeline() == pipelined
    for k, v in system.items():
        ctx.set(k, v)
    out = ctx.congrad()
    p = ctx.get("P")
    ctx.close()
    return out, p


@pytest.mark.parametrize("n,seed", [(12, 2), (24, 3), (64, 5), (200, 7)])
def test_fp64_pipelined_form_matches_oracle(oracle_mod, n, seed):
    """The FP64 persistent solve in the pipelined order (one exchange per
    iteration) against the oracle, |d eta| <= 1e-9 m, and its iteration
    count within one of the reference's recurrences (same iterates in exact
    arithmetic, same stop test)."""
    case, system, p_ref = _lake_system(oracle_mod, n, seed)
    o_ref, p_rec = _solve64(case.grid, system, False)
    o_pl, p_pl = _solve64(case.grid, system, True)
    i = wet_index(case)
    d = eta_err(p_pl[i], p_ref[i])
    print(f"n={n} wet={case.grid.nwet}: iters {o_ref['iters']} (recurrences) "
          f"{o_pl['iters']} (pipelined), deta {d:.3e}, "
          f"vs recurrences {eta_err(p_pl[i], p_rec[i]):.3e}")
    assert abs(o_pl["iters"] - o_ref["iters"]) <= 1
    assert d <= 1e-9


def test_fp64_pipelined_form_general_maps(oracle_mod):
    """The same on the column-major renumbering (the general four-map
    operator, maps that are not the row-major pattern)."""
    case, system, p_ref = _lake_system(oracle_mod, 96, 4)
    g2, sys2, src = _column_major(case, system)
    out, p = _solve64(g2, sys2, True)
    newi = np.arange(2, g2.LA + 1) - 1
    d = eta_err(p[newi], p_ref[src - 1])
    print("column-major, pipelined FP64: iters", out["iters"], "deta", d)
    assert d <= 1e-9


def test_fp64_pipelined_form_in_steps(oracle_mod):
    """CALPUV2C steps of a tidal basin (a clamped open boundary: rows that
    are not symmetric) with the pipelined FP64 solve: P, HP and the face
    transports against the oracle at the parity tests' bars."""
    from efdc_amd import grids, lib
    case = grids.tidal_basin(160, 80, 1, seed=12)
    ref = case.copy()
    o = oracle_mod.Oracle(ref, rsqm=TIGHT, iterm=2000)
    ctx = lib.Context(case.grid, RSQM=TIGHT, ITERM=2000,
                      cg_kernel=lib.CG_KERNEL_PERSISTENT).upload_case(case)
    ctx.set_cg64_pipeline(True)
    i = wet_index(case)
    for n in (1, 2, 3):
        o.step(n)
        st = ctx.step(case.dt, n)
        d = eta_err(ctx.get("P")[i], ref.h["P"][i])
        print(f"step {n}: deta {d:.3e}")
        assert d <= 1e-9
        for name in ("HP", "UHDYE", "VHDXE", "HU", "HV"):
            a, b = ctx.get(name)[i], ref.h[name][i]
            assert np.max(np.abs(a - b)) <= 1e-9 * max(1.0, np.max(np.abs(b))), name
    ctx.close()
