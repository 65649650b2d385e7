"""Ensembles of independent basins in one context (efdc_set_segments).

Each basin of a concatenated case must give the same bits as the same basin
run alone with the FP64 multi-kernel CG: every pointwise kernel sees the same
inputs, the CCMNM scaling minimum is taken per basin, and the per-basin CG
tiles start at the basin's first cell, so the dot-product partials and their
reduction order are those of the basin alone.  Balances are per basin (FP64
sums in another order: equal to 1e-13 relative).
"""
import numpy as np
import pytest

from conftest import eta_err

pytestmark = pytest.mark.gpu

RSQM = 1e-20
H_FIELDS = ["P", "P1", "HP", "H1P", "HU", "HV", "UHDYE", "VHDXE", "UHDY1E", "VHDX1E",
            "CCC", "CCS", "CCW", "FPTMP"]


def _ctx(case, **kw):
    from efdc_amd.lib import CG_FP64, CG_KERNEL_MULTI, Context
    kw.setdefault("RSQM", RSQM)
    kw.setdefault("ITERM", 2000)
    kw.setdefault("cg_kernel", CG_KERNEL_MULTI)
    kw.setdefault("cg_mode", CG_FP64)
    return Context(case.grid, **kw).upload_case(case)


def _tidal_ensemble():
    from efdc_amd import grids
    basins = [grids.tidal_basin(40, 20, 3, seed=21, amp=0.5),
              grids.tidal_basin(24, 36, 3, seed=22, amp=0.3, depth=(3.0, 12.0)),
              grids.tidal_basin(64, 12, 3, seed=23, amp=0.8, dye_load=2.0),
              grids.tidal_basin(30, 30, 3, seed=24, amp=0.1, manning=0.04),
              grids.ensemble_member(0, kc=3)]
    ens, first = grids.concat_basins(basins)
    return basins, ens, first


def test_tidal_ensemble_bitwise_equals_separate_basins(oracle_mod):
    """Five basins (four rectangles, one irregular lake with a west
    boundary segment), 30 full steps: bitwise equal to the basins run
    alone, and within 1e-8 m / 1e-9 relative of the FP64 oracle run on
    each basin."""
    from efdc_amd import grids
    basins, ens, first = _tidal_ensemble()
    refs = [b.copy() for b in basins]
    oracles = [oracle_mod.Oracle(r, rsqm=RSQM, iterm=2000) for r in refs]
    singles = [_ctx(b) for b in basins]
    ctx = _ctx(ens)
    ctx.set_segments(first)
    nsteps = 30
    for n in range(1, nsteps + 1):
        for o in oracles:
            o.step(n)
        its = [s.step(b.dt, n)["iters"] for s, b in zip(singles, basins)]
        st = ctx.step(ens.dt, n)
        it_seg, status = ctx.segment_stats()
        assert list(it_seg) == its, (n, list(it_seg), its)
        assert st["iters"] == max(its)
        assert not status.any()
    kc = ens.grid.KC
    for b, (s, basin) in enumerate(zip(singles, basins)):
        sl = grids.basin_slice(ens, first, b)
        nw = basin.grid.nwet
        for name in H_FIELDS:
            got = ctx.get(name)[sl]
            want = s.get(name)[1:1 + nw]
            assert np.array_equal(got, want), (b, name, np.max(np.abs(got - want)))
        for j in range(len(basin.consts)):
            con_e, con1_e = ctx.get_constituent(j)
            con_s, con1_s = s.get_constituent(j)
            assert np.array_equal(con_e[:kc, sl], con_s[:kc, 1:1 + nw]), (b, j)
            assert np.array_equal(con1_e[:kc, sl], con1_s[:kc, 1:1 + nw]), (b, j)
    p = ctx.get("P")
    for b, (r, basin) in enumerate(zip(refs, basins)):
        sl = grids.basin_slice(ens, first, b)
        nw = basin.grid.nwet
        d = eta_err(p[sl], r.h["P"][1:1 + nw])
        assert d <= 1e-8, (b, d)
        for j, c in enumerate(r.consts):
            con, _ = ctx.get_constituent(j)
            want = c["con"][:kc, 1:1 + nw]
            rel = np.max(np.abs(con[:kc, sl] - want)) / max(1e-30, np.max(np.abs(want)))
            assert rel <= 1e-9, (b, j, rel)
    vol, mass = ctx.balance_segments()
    for b, s in enumerate(singles):
        v, m = s.balance()
        np.testing.assert_allclose(vol[b], v, rtol=1e-13)
        np.testing.assert_allclose(mass[b], m, rtol=1e-13)
    print("basins", len(basins), "cells", ens.grid.nwet, "last iters", list(it_seg))
    for s in singles:
        s.close()
    ctx.close()


def test_lake_ensemble_external_solve_bitwise():
    """Irregular lake masks (ragged rows, islands) through CALPUV2C."""
    from efdc_amd import grids
    basins = [grids.external_case(grids.lake_mask(n, n, seed), seed=seed)
              for n, seed in ((48, 31), (80, 32), (33, 33), (64, 34), (20, 35))]
    ens, first = grids.concat_basins(basins)
    singles = [_ctx(b) for b in basins]
    ctx = _ctx(ens)
    ctx.set_segments(first)
    for n in (1, 2, 3):
        its = [s.calpuv2c(b.dt, n)["iters"] for s, b in zip(singles, basins)]
        ctx.calpuv2c(ens.dt, n)
        assert list(ctx.segment_stats()[0]) == its
    for b, (s, basin) in enumerate(zip(singles, basins)):
        sl = grids.basin_slice(ens, first, b)
        nw = basin.grid.nwet
        for name in ("P", "HP", "UHDYE", "VHDXE"):
            assert np.array_equal(ctx.get(name)[sl], s.get(name)[1:1 + nw]), (b, name)
    for s in singles:
        s.close()
    ctx.close()


def test_segment_iteration_cap_is_per_basin():
    from efdc_amd.lib import EfdcError
    _, ens, first = _tidal_ensemble()
    ctx = _ctx(ens, RSQM=1e-30, ITERM=3)
    ctx.set_segments(first)
    with pytest.raises(EfdcError) as e:
        ctx.calpuv2c(ens.dt, 1)
    assert e.value.status == 3     # EFDC_ERR_CG_MAXITER
    it, status = ctx.segment_stats()
    assert list(it) == [3] * len(first)
    assert list(status) == [3] * len(first)
    ctx.close()


def test_set_segments_rejects_bad_tables():
    from efdc_amd import grids
    from efdc_amd.lib import EfdcError
    _, ens, first = _tidal_ensemble()
    ctx = _ctx(ens)
    for bad in ([3, 100], [2, int(first[2]), int(first[1])], [2, ens.grid.LA + 1]):
        with pytest.raises(EfdcError) as e:
            ctx.set_segments(bad)
        assert e.value.status == 1
    # a split inside a basin: its neighbour maps cross the split
    with pytest.raises(EfdcError) as e:
        ctx.set_segments([2, int(first[1]) + 40])
    assert e.value.status == 1 and "crosses" in str(e.value)
    ctx.close()


def test_single_segment_is_the_plain_model():
    """set_segments([2]) returns to the one-model path, bit for bit."""
    _, ens, first = _tidal_ensemble()
    a = _ctx(ens)
    b = _ctx(ens)
    b.set_segments(first)
    b.set_segments([2])
    for n in (1, 2):
        a.step(ens.dt, n)
        b.step(ens.dt, n)
    assert eta_err(a.get("P"), b.get("P")) == 0.0
    a.close()
    b.close()


def test_captured_cg_iterations_equal_launch_by_launch():
    """The FP64 multi-kernel solves replaying captured graphs of 8 and 1
    CG iterations give the bits of the launch-by-launch loop, for one model
    and for basins; the timed trial of the two keeps one of them and
    reports both."""
    _, ens, first = _tidal_ensemble()
    for segs in (False, True):
        a, b, t = _ctx(ens), _ctx(ens), _ctx(ens)
        if segs:
            for x in (a, b, t):
                x.set_segments(first)
        b.set_cg_replay(1)
        t.set_cg_replay(2)
        assert (a.cg_replay(), b.cg_replay(), t.cg_replay()) == (0, 1, 2)
        for n in range(1, 6):
            ia = a.step(ens.dt, n)["iters"]
            ib = b.step(ens.dt, n)["iters"]
            it = t.step(ens.dt, n)["iters"]
            assert ia == ib == it, (segs, n)
        assert t.cg_replay() in (0, 1)
        rep = t.tune_report()
        assert "cg_fp64 launches" in rep["shapes"] and "cg_fp64 graphs" in rep["shapes"]
        for name in ("P", "HP", "UHDYE", "VHDXE"):
            assert np.array_equal(a.get(name), b.get(name)), (segs, name)
            assert np.array_equal(a.get(name), t.get(name)), (segs, name)
        for x in (a, b, t):
            x.close()


def test_segment_tiles_and_their_trials():
    """Tiles of 256, 1024 and 2048 cells sum the basins' dot products in
    another order: within round-off of the 512-cell tile (which is the
    basins-alone order); cells = 0 times the four sizes on the next solve,
    keeps the fastest and reports the trials."""
    _, ens, first = _tidal_ensemble()
    ref = _ctx(ens)
    ref.set_segments(first)
    assert ref.segment_tile() == 512
    for n in range(1, 9):
        ref.step(ens.dt, n)
    p_ref = ref.get("P")
    L = np.arange(1, ens.grid.LA)
    for cells in (256, 1024, 2048, 0):
        ctx = _ctx(ens)
        ctx.set_segments(first)
        ctx.set_segment_tile(cells)
        for n in range(1, 9):
            ctx.step(ens.dt, n)
        if cells:
            assert ctx.segment_tile() == cells
        else:
            assert ctx.segment_tile() in (256, 512, 1024, 2048)
            rep = ctx.tune_report()
            trials = [k for k in rep["shapes"] if k.startswith("seg_cg_tile_")]
            assert len(trials) == 4
            chosen = [k for k in rep["chosen"] if k.startswith("seg_cg_tile_")]
            assert chosen == [f"seg_cg_tile_{ctx.segment_tile()}"]
        assert eta_err(ctx.get("P")[L], p_ref[L]) <= 1e-9, cells
        ctx.close()
    with pytest.raises(Exception):
        ref.set_segment_tile(300)
    ref.close()


def test_256_basins_each_equal_to_the_basin_alone():
    """256 basins of 8 shapes in one context, 3 full steps: a sample of
    them, run alone, give the same bits (the 512-cell tile)."""
    from efdc_amd import grids
    basins = [grids.tidal_basin(12 + 4 * (i % 4), 10 + 3 * (i % 2), 2, seed=300 + i,
                                amp=0.05 + 0.003 * (i % 50), dye_load=0.5 + (i % 7))
              for i in range(256)]
    ens, first = grids.concat_basins(basins)
    ctx = _ctx(ens)
    ctx.set_segments(first)
    for n in range(1, 4):
        ctx.step(ens.dt, n)
    p = ctx.get("P")
    hp = ctx.get("HP")
    for b in (0, 1, 77, 128, 200, 254, 255):
        s = _ctx(basins[b])
        for n in range(1, 4):
            s.step(basins[b].dt, n)
        sl = grids.basin_slice(ens, first, b)
        nw = basins[b].grid.nwet
        assert np.array_equal(p[sl], s.get("P")[1:1 + nw]), b
        assert np.array_equal(hp[sl], s.get("HP")[1:1 + nw]), b
        s.close()
    ctx.close()


def test_fp32_refined_solve_per_basin():
    """The FP32-refined solve of every basin (FP32 inner CG on the scaled
    operator, FP64 residuals and corrections, per-basin stop tests and
    outer rounds) on irregular lakes, whose scaled operator is symmetric:
    the mixed kernels run, and over 10 external steps each basin's η stays
    within 1e-6 m of the same basin solved alone with the FP64 CG.  (With
    open pressure boundaries the operator is not symmetric and the basins
    are solved in FP64, as one model is.)"""
    from efdc_amd import grids
    from efdc_amd.lib import CG_MIXED
    basins = [grids.external_case(grids.lake_mask(n, n, seed), seed=seed)
              for n, seed in ((48, 31), (80, 32), (33, 33), (64, 34), (20, 35))]
    ens, first = grids.concat_basins(basins)
    ctx = _ctx(ens, cg_mode=CG_MIXED, cg_rel_tol=1e-10, RSQM=0.0)
    ctx.set_segments(first)
    singles = [_ctx(b) for b in basins]
    for n in range(1, 11):
        if n == 2:
            ctx.timers(True)
        st = ctx.calpuv2c(ens.dt, n)
        if n == 2:
            names = set(ctx.read_timers())
            ctx.timers(False)
            assert {"mx_residual64", "mx_spmv_pupdate32", "mx_axpy_dots32", "mx_apply"} <= names, names
            assert "cg_spmv_pupdate" not in names
        assert st["status"] == 0, n
        for s, b in zip(singles, basins):
            s.calpuv2c(b.dt, n)
    it, status = ctx.segment_stats()
    assert list(status) == [0] * len(first) and min(it) > 0
    p = ctx.get("P")
    for b, (s, basin) in enumerate(zip(singles, basins)):
        sl = grids.basin_slice(ens, first, b)
        nw = basin.grid.nwet
        assert eta_err(p[sl], s.get("P")[1:1 + nw]) <= 1e-6, b
        s.close()
    ctx.close()
