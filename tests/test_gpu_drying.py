"""Drying on the GPU against the oracle, ISDRY 1..97 and 99 (drying and
rewetting, EFDC/calpuv2c.f90:604-716, which the GPU evaluates for all cells
at once where the reference walks them in order): the re-solve loop
(EFDC/calpuv2c.f90:302,574-602), open-boundary drying (:497-530), the
face reopening at the step start (:246-253) and the transport mask
LMASKDRY (:916-939) with the transport leaving the masked columns out.
The cell flags and the face switches must agree exactly and the number of
re-solves NCORDRY step by step; P within the FP64 CG's round-off, the
concentrations within 1e-9 relative (they inherit the CG's round-off
through the flows), and every column shape of the transport kernels must
give the same bits."""
import numpy as np
import pytest

from conftest import eta_err, wet_index

pytestmark = pytest.mark.gpu

HDRY = 0.2


def _stratified(case):
    """Layered salinity with settling, dye 1 everywhere with decay."""
    kc = case.grid.KC
    L = wet_index(case)
    sal, dye = case.consts
    for K in range(kc):
        sal["con"][K, L] = 10.0 + 5.0 * K
    sal["con1"][:] = sal["con"]
    sal["wset"], sal["probdep"], sal["wese"] = 1e-4, 0.5, 1e-6
    dye["con"][:kc, L] = 1.0
    dye["con1"][:] = dye["con"]
    dye["rkdye"] = 1e-4
    return case


@pytest.mark.parametrize("cg_kernel", [0, 1])
def test_open_boundary_drying_only_isdry98(oracle_mod, cg_kernel):
    """ISDRY = 98: only the open-boundary cells dry (EFDC/calpuv2c.f90:
    497-530); the cell pass and its re-solve are skipped, so NCORDRY stays
    0.  Flags, eta, HP and the face masks against the oracle every step of
    a basin whose open boundary dries."""
    from efdc_amd import grids
    from efdc_amd.lib import Context
    # a basin 3 m deep everywhere but a 0.6 m sill on the open-boundary
    # column: the tide (0.8 m) dries the sill, no other cell comes near HDRY
    case = grids.drying_basin(kc=3, deep=3.0, high=-3.0, shoal=0.0)
    g, h = case.grid, case.h
    lb = case.open_bc["side"][0]["L"] - 1
    h["BELV"][lb] = -0.6
    for key in ("HP", "H1P", "H2P"):
        h[key][lb] = 0.6
    h["P"][lb] = grids.G * (h["HP"][lb] + h["BELV"][lb])
    h["P1"][lb] = h["P"][lb]
    grids.update_depth_arrays(g, h)
    grids.layer_depths(g, h, case.k)
    case.k["H1PK"][:] = case.k["HPK"]
    if cg_kernel == 1:
        case = _stratified(case)
    ref = case.copy()
    o = oracle_mod.Oracle(ref, rsqm=1e-24, iterm=2000, isdry=98, hdry=HDRY)
    ctx = Context(case.grid, RSQM=1e-24, ITERM=2000, ISDRY=98, HDRY=HDRY,
                  cg_kernel=cg_kernel).upload_case(case)
    L = wet_index(case)
    obc_dried = False
    for n in range(1, 301):
        so = o.step(n)
        sg = ctx.step(case.dt, n)
        assert so["status"] == 0 and sg["status"] == 0, n
        m, d, k = ctx.get_dry_state()
        assert k == o.o.ncordry == 0, (n, k, o.o.ncordry)
        assert np.array_equal(m[L], o.lmaskdry[L]), n
        assert np.array_equal(d[L], o.iscdry[L]), n
        obc_dried |= bool(np.any(d[lb] == 1))
        if n % 50 == 0 or n < 4:
            assert eta_err(ctx.get("P")[L], ref.h["P"][L]) <= 1e-9, n
            for key in ("SUB", "SVB"):
                assert np.array_equal(ctx.get(key)[L], ref.h[key][L]), (n, key)
            assert np.max(np.abs(ctx.get("HP")[L] - ref.h["HP"][L])) <= 1e-9, n
    assert obc_dried
    ctx.close()


@pytest.mark.parametrize("isdry", [1, 99])
@pytest.mark.parametrize("cg_kernel", [0, 1])
@pytest.mark.parametrize("deep", [2.0, 0.6])
def test_drying_matches_oracle(oracle_mod, deep, cg_kernel, isdry):
    from efdc_amd import grids
    from efdc_amd.lib import Context
    case = grids.drying_basin(kc=3, deep=deep)
    if cg_kernel == 1:
        case = _stratified(case)
    ref = case.copy()
    o = oracle_mod.Oracle(ref, rsqm=1e-24, iterm=2000, isdry=isdry, hdry=HDRY)
    ctx = Context(case.grid, RSQM=1e-24, ITERM=2000, ISDRY=isdry, HDRY=HDRY,
                  cg_kernel=cg_kernel).upload_case(case)
    L = wet_index(case)
    lb = case.open_bc["side"][0]["L"] - 1
    kc = case.grid.KC
    seen = set()
    obc_dried = False
    for n in range(1, 601):
        so = o.step(n)
        sg = ctx.step(case.dt, n)
        assert so["status"] == 0 and sg["status"] == 0, n
        m, d, k = ctx.get_dry_state()
        assert k == o.o.ncordry, (n, k, o.o.ncordry)
        assert np.array_equal(m[L], o.lmaskdry[L]), n
        assert np.array_equal(d[L], o.iscdry[L]), n
        seen.add(k)
        obc_dried |= bool(np.any(d[lb] == 1))
        if n % 50 == 0 or n < 4:
            assert eta_err(ctx.get("P")[L], ref.h["P"][L]) <= 1e-9, n
            for key in ("SUB", "SVB"):
                assert np.array_equal(ctx.get(key)[L], ref.h[key][L]), (n, key)
            hp = ctx.get("HP")[L]
            assert np.max(np.abs(hp - ref.h["HP"][L])) <= 1e-9, n
            for j, c in enumerate(ref.consts):
                con, _ = ctx.get_constituent(j)
                scale = max(1e-30, np.max(np.abs(c["con"][:kc])))
                assert np.max(np.abs(con[:kc] - c["con"][:kc])) / scale <= 1e-9, (n, j)
    if deep == 2.0:
        assert max(seen) >= (2 if isdry == 1 else 1)
    else:
        assert obc_dried
    for s, c in enumerate(ref.consts):
        con, _ = ctx.get_constituent(s)
        scale = max(1e-30, np.max(np.abs(c["con"][:kc])))
        assert np.max(np.abs(con[:kc] - c["con"][:kc])) / scale <= 1e-9, s
    ctx.close()


def _sill_basin(kc=3, sill=0.6, deep=3.0):
    """A 3 m deep basin behind a 0.6 m sill at its open (west) boundary:
    with the 0.8 m tide of drying_basin the boundary cells dry at low
    water and come back, the interior never gets shallower than ~0.85 m."""
    from efdc_amd import grids
    case = grids.drying_basin(kc=kc, deep=deep, high=-deep, shoal=0.0)
    g, h, k = case.grid, case.h, case.k
    L = wet_index(case)
    belv = np.where(g.IL[L] == 2, -sill, -deep)
    for key in ("HP", "H1P", "H2P"):
        h[key][L] = -belv
    h["BELV"][L] = belv
    h["P"][L] = 0.0
    h["P1"][L] = 0.0
    grids.update_depth_arrays(g, h)
    grids.layer_depths(g, h, k)
    k["H1PK"][:] = k["HPK"]
    return case


@pytest.mark.parametrize("cg_kernel", [0, 1])
def test_open_boundary_drying_98_matches_oracle(oracle_mod, cg_kernel):
    """ISDRY = 98: only open-boundary cells dry (EFDC/calpuv2c.f90:497-530),
    the cell pass is skipped, so no step solves twice."""
    from efdc_amd.lib import Context
    case = _sill_basin()
    ref = case.copy()
    o = oracle_mod.Oracle(ref, rsqm=1e-24, iterm=2000, isdry=98, hdry=HDRY)
    ctx = Context(case.grid, RSQM=1e-24, ITERM=2000, ISDRY=98, HDRY=HDRY,
                  cg_kernel=cg_kernel).upload_case(case)
    L = wet_index(case)
    lb = case.open_bc["side"][0]["L"] - 1
    kc = case.grid.KC
    dried = returned = 0
    prev = False
    for n in range(1, 601):
        so = o.step(n)
        sg = ctx.step(case.dt, n)
        assert so["status"] == 0 and sg["status"] == 0, n
        m, d, k = ctx.get_dry_state()
        assert k == 0 and o.o.ncordry == 0, n
        assert np.array_equal(m[L], o.lmaskdry[L]), n
        assert np.array_equal(d[L], o.iscdry[L]), n
        now = bool(np.any(d[lb] == 1))
        dried += now
        returned += prev and not now
        prev = now
        if n % 50 == 0 or n < 4:
            assert eta_err(ctx.get("P")[L], ref.h["P"][L]) <= 1e-9, n
            for key in ("SUB", "SVB"):
                assert np.array_equal(ctx.get(key)[L], ref.h[key][L]), (n, key)
            for j, c in enumerate(ref.consts):
                con, _ = ctx.get_constituent(j)
                scale = max(1e-30, np.max(np.abs(c["con"][:kc])))
                assert np.max(np.abs(con[:kc] - c["con"][:kc])) / scale <= 1e-9, (n, j)
    assert dried > 100 and returned >= 1
    ctx.close()


def test_drying_transport_shapes_bitwise():
    from efdc_amd import grids
    from efdc_amd.lib import Context
    case = _stratified(grids.drying_basin(kc=3, deep=0.6))
    shapes = [1, 2, 17, 34, -1, -2]
    ctxs = [Context(case.grid, RSQM=1e-24, ITERM=2000, ISDRY=1, HDRY=HDRY,
                    tran_shape=q).upload_case(case) for q in shapes]
    for n in range(1, 301):
        for c in ctxs:
            c.step(case.dt, n)
    kc = case.grid.KC
    masked = ctxs[0].get_dry_state()[0][wet_index(case)] == 0
    assert masked.any()
    for j in range(len(case.consts)):
        a, a1 = ctxs[0].get_constituent(j)
        for c in ctxs[1:]:
            b, b1 = c.get_constituent(j)
            assert np.array_equal(a[:kc], b[:kc]) and np.array_equal(a1[:kc], b1[:kc]), j
    for c in ctxs:
        c.close()


def test_drying_is_refused_where_not_built():
    from efdc_amd import grids
    from efdc_amd.lib import Context, EfdcError
    case = grids.drying_basin()
    ctx = Context(case.grid, ISDRY=1).upload_case(case)
    with pytest.raises(EfdcError):
        ctx.adj_enable([0])
    ctx.close()


def test_rewetting_closed_basin_matches_oracle(oracle_mod):
    """ISDRY = 99 on the sloshing basin whose dry bank is wetted again."""
    import test_oracle_drying as t
    from efdc_amd.lib import Context
    case = t._closed()
    ref = case.copy()
    o = oracle_mod.Oracle(ref, rsqm=1e-26, iterm=4000, isdry=99, hdry=HDRY)
    ctx = Context(case.grid, RSQM=1e-26, ITERM=4000, ISDRY=99, HDRY=HDRY).upload_case(case)
    L = wet_index(case)
    for n in range(1, 401):
        o.step(n)
        ctx.step(case.dt, n)
        m, d, k = ctx.get_dry_state()
        assert k == o.o.ncordry and np.array_equal(m[L], o.lmaskdry[L]), n
        assert np.array_equal(d[L], o.iscdry[L]), n
    for key in ("SUB", "SVB"):
        assert np.array_equal(ctx.get(key)[L], ref.h[key][L]), key
    assert np.max(np.abs(ctx.get("HP")[L] - ref.h["HP"][L])) <= 1e-9
    ctx.close()


@pytest.mark.parametrize("isdry", [1, 99])
def test_draining_matches_oracle(oracle_mod, isdry):
    """NDRYSTP < 0: the isolated shallow pit is drained on the same step,
    with the same rate, volume, counters and depths as in the oracle."""
    import test_oracle_drying as t
    from efdc_amd.lib import Context
    case, (shallow, deep) = t._puddles()
    ref = case.copy()
    o = oracle_mod.Oracle(ref, rsqm=1e-26, iterm=4000, isdry=isdry, hdry=HDRY, ndrystp=-5)
    ctx = Context(case.grid, RSQM=1e-26, ITERM=4000, ISDRY=isdry, HDRY=HDRY,
                  NDRYSTP=-5).upload_case(case)
    L = wet_index(case)
    kc = case.grid.KC
    for n in range(1, 21):
        o.step(n)
        ctx.step(case.dt, n)
        q, v, k = ctx.get_dry_waste()
        assert np.array_equal(k[L], o.natdry[L]), n
        assert np.max(np.abs(q[L] - o.qdwaste[L])) <= 1e-9 * max(1.0, o.qdwaste.max()), n
        assert np.max(np.abs(v[L] - o.vdwaste[L])) <= 1e-9 * max(1.0, o.vdwaste.max()), n
        assert np.max(np.abs(ctx.get("HP")[L] - ref.h["HP"][L])) <= 1e-9, n
    assert v[shallow] > 0 and v[deep] == 0
    h1pk = ctx.get("H1PK")
    assert np.array_equal(h1pk[:kc, shallow], ref.k["H1PK"][:kc, shallow])
    assert np.max(np.abs(h1pk[:kc, L] - ref.k["H1PK"][:kc, L])) <= 1e-9
    ctx.close()


def test_drain_counters_continue_after_restart(oracle_mod):
    """A context created mid-run (a restart after step 3) and given the
    oracle's VDWASTE and NATDRY continues like the oracle: the shallow pit,
    isolated since step 1, drains on step 6 as in the uninterrupted run
    (with its counter back at 0 it would drain on step 9)."""
    import test_oracle_drying as t
    from efdc_amd.lib import Context
    case, (shallow, deep) = t._puddles()
    ref = case.copy()
    o = oracle_mod.Oracle(ref, rsqm=1e-26, iterm=4000, isdry=1, hdry=HDRY, ndrystp=-5)
    for n in (1, 2, 3):
        o.step(n)
    assert o.natdry[shallow] == 3 and o.vdwaste[shallow] == 0.0
    ctx = Context(case.grid, RSQM=1e-26, ITERM=4000, ISDRY=1, HDRY=HDRY,
                  NDRYSTP=-5).upload_case(ref.copy())
    ctx.set_dry_counters(o.vdwaste.copy(), o.natdry.copy())
    L = wet_index(case)
    for n in range(4, 11):
        o.step(n)
        ctx.step(case.dt, n)
        q, v, k = ctx.get_dry_waste()
        assert np.array_equal(k[L], o.natdry[L]), n
        assert np.max(np.abs(v[L] - o.vdwaste[L])) <= 1e-9 * max(1.0, o.vdwaste.max()), n
        assert np.max(np.abs(ctx.get("HP")[L] - ref.h["HP"][L])) <= 1e-9, n
        if n == 6:
            assert v[shallow] > 0
    ctx.close()


@pytest.mark.parametrize("isdry", [1, 99])
def test_antidiffusion_with_drying_matches_oracle(oracle_mod, isdry):
    """ISADAC = 1 with FCT on the drying basin: the FCT limiter sees the
    closed faces through SUB3D = SUB*SUB3DO, masked columns are left out."""
    from efdc_amd import grids
    from efdc_amd.lib import Context
    case = _stratified(grids.drying_basin(kc=3, deep=0.6, isadac=1, isfct=1))
    ref = case.copy()
    o = oracle_mod.Oracle(ref, rsqm=1e-24, iterm=2000, isdry=isdry, hdry=HDRY)
    ctx = Context(case.grid, RSQM=1e-24, ITERM=2000, ISDRY=isdry, HDRY=HDRY).upload_case(case)
    L = wet_index(case)
    kc = case.grid.KC
    for n in range(1, 301):
        o.step(n)
        ctx.step(case.dt, n)
        m, d, k = ctx.get_dry_state()
        assert k == o.o.ncordry and np.array_equal(m[L], o.lmaskdry[L]), n
        if n % 50 == 0:
            for j, c in enumerate(ref.consts):
                con, _ = ctx.get_constituent(j)
                scale = max(1e-30, np.max(np.abs(c["con"][:kc])))
                assert np.max(np.abs(con[:kc] - c["con"][:kc])) / scale <= 1e-9, (n, j)
    assert np.any(m[L] == 0)
    ctx.close()
