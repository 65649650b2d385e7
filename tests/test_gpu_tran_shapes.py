"""Launch shapes of the CALCONC column kernels (efdc_params_t.tran_shape).

k_tran_colm (1, 2 or 4 constituents per thread, the transport fields and
the tridiagonal factors computed once per layer for all of them; with the
compiler's register budget, at most 168 VGPRs, or that and the transport
fields loaded two layers ahead) and
k_tran_col (one constituent per thread at a time, 1, 2 or 4 waves) must
give the same bits for every shape: same values, same FP64 operation order.
The cases cover 2, 3 (a partly filled wave) and 4 constituents, FP64 and
FP32 storage, the open boundary with its concentration BCs, the dye
source, and the anti-diffusion pipeline (where the column kernel stops
before the vertical solve).  FP64 results are also held to the oracle.
"""
import numpy as np
import pytest

from conftest import eta_err

pytestmark = pytest.mark.gpu

SHAPES4 = [1, 2, 4, 17, 18, 20, 33, 34, -1, -2, -4]


def _ctx(case, shape, **kw):
    from efdc_amd.lib import Context
    kw.setdefault("RSQM", 1e-24)
    kw.setdefault("ITERM", 2000)
    return Context(case.grid, tran_shape=shape, **kw).upload_case(case)


def _drop_last_constituent(case):
    c = case.copy()
    c.consts = c.consts[:-1]
    s = dict(c.sources)
    s["cqs"] = np.ascontiguousarray(np.asarray(s["cqs"])[..., :-1])
    s["csert"] = np.ascontiguousarray(np.asarray(s["csert"])[..., :-1])
    c.sources = s
    return c


def _same_bits(ctxs, nconst, kc):
    base = ctxs[0]
    for other in ctxs[1:]:
        for j in range(nconst):
            a, a1 = base.get_constituent(j)
            b, b1 = other.get_constituent(j)
            assert np.array_equal(a[:kc], b[:kc]) and np.array_equal(a1[:kc], b1[:kc]), j


@pytest.mark.parametrize("store", ["fp64", "fp32"])
@pytest.mark.parametrize("nconst", [4, 3])
def test_calconc_shapes_bitwise(oracle_mod, store, nconst):
    from efdc_amd import grids, lib
    case = grids.transport_case(48, 24, 6, seed=41)
    if nconst == 3:
        case = _drop_last_constituent(case)
    st = lib.STORE_FP64 if store == "fp64" else lib.STORE_FP32
    shapes = [q for q in SHAPES4 if not (nconst == 3 and q == -4)]
    ctxs = [_ctx(case, q, store=st) for q in shapes]
    ref = case.copy()
    o = oracle_mod.Oracle(ref)
    kc = case.grid.KC
    for n in range(1, 4):
        o.calconc(niter=n)
        for c in ctxs:
            c.calconc(case.dt, n)
    _same_bits(ctxs, nconst, kc)
    if store == "fp64":
        for s, c in enumerate(ref.consts):
            con, _ = ctxs[0].get_constituent(s)
            err = np.max(np.abs(con[:kc] - c["con"][:kc]))
            assert err <= 1e-12 * max(1.0, np.max(np.abs(c["con"][:kc]))), (s, err)
    for c in ctxs:
        c.close()


@pytest.mark.parametrize("store", ["fp64", "fp32"])
def test_tidal_basin_steps_shapes_bitwise(oracle_mod, store):
    """Full steps with the open boundary (outflow clipping and the inflow
    recovery ramp), the tide and the dye source."""
    from efdc_amd import grids, lib
    case = grids.tidal_basin(40, 20, 5, seed=43)
    st = lib.STORE_FP64 if store == "fp64" else lib.STORE_FP32
    shapes = [1, 2, 17, 18, 33, 34, -1, -2]
    ctxs = [_ctx(case, q, store=st) for q in shapes]
    ref = case.copy()
    o = oracle_mod.Oracle(ref, rsqm=1e-24, iterm=2000)
    for n in range(1, 21):
        o.step(n)
        for c in ctxs:
            c.step(case.dt, n)
    kc = case.grid.KC
    _same_bits(ctxs, len(case.consts), kc)
    for c in ctxs[1:]:
        assert eta_err(ctxs[0].get("P"), c.get("P")) == 0.0
    if store == "fp64":
        for s, c in enumerate(ref.consts):
            con, _ = ctxs[0].get_constituent(s)
            scale = max(1e-30, np.max(np.abs(c["con"][:kc])))
            assert np.max(np.abs(con[:kc] - c["con"][:kc])) / scale <= 1e-9, s
    for c in ctxs:
        c.close()


def test_antidiffusion_shapes_bitwise():
    from efdc_amd import grids
    case = grids.transport_case(40, 24, 6, seed=45, isadac=1, isfct=1)
    ctxs = [_ctx(case, q) for q in SHAPES4]
    for n in range(1, 4):
        for c in ctxs:
            c.calconc(case.dt, n)
    _same_bits(ctxs, len(case.consts), case.grid.KC)
    for c in ctxs:
        c.close()


def test_shape_that_does_not_fit_is_refused():
    """Five constituents one per wave would need five waves per workgroup."""
    from efdc_amd import grids
    from efdc_amd.lib import EfdcError
    case = grids.transport_case(24, 12, 4, seed=47)
    c = case.copy()
    c.consts = c.consts + [dict(c.consts[0])]
    s = dict(c.sources)
    for key in ("cqs", "csert"):
        v = np.asarray(s[key])
        s[key] = np.ascontiguousarray(np.concatenate([v, v[..., :1]], axis=-1))
    c.sources = s
    ctx = _ctx(c, 1)
    with pytest.raises(EfdcError) as e:
        ctx.calconc(c.dt, 1)
    assert e.value.status == 1 and "tran_shape" in str(e.value)
    ctx.close()
