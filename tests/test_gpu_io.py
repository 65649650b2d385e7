"""The on-disk formats written from the device (efdc_restart_text /
efdc_restart_write, RESTOUT, EFDC/restartmod.f90:17-297; efdc_ee_open /
efdc_ee_append, HEADEROUT / WSOUT / VELOUT, EFDC/efdcout.f90:73-117,
371-497) against the oracle's writers, byte for byte:

* a state with every field the restart reads set to values that span the
  exponent range, with signed zeros, subnormals, NaN, infinities and
  exact ties planted (the values the device hands to the host formatter),
  in FP64 and FP32 storage, on two calls in a row (NLO shifted by N on
  each, as RESTOUT leaves it in memory) and through the file;
* the state after internal-mode steps with salinity, a dye and a west
  concentration boundary: the cell and source records byte for byte,
  the boundary records' NLO identical and CLO to the digits written;
* EE_WS.OUT and EE_VEL.OUT over two snapshots, with and without KSZ.
"""
import os

import numpy as np
import pytest

from conftest import wet_index

pytestmark = pytest.mark.gpu

SPECIALS = [0.0, -0.0, 5e-324, 1e-300, -2.5e-310, 1e300, np.nan, np.inf, -np.inf, 12345678.5,
            658.609375, -123456785.0, 0.999999995, 9.999999995e279, 1e-280]
H_READ = ["HP", "H1P", "BELV", "UHDYE", "UHDY1E", "VHDXE", "VHDX1E", "QSUME"]
K_READ = ["QSUM", "U", "U1", "V", "V1"]


def _base():
    from efdc_amd import grids
    return grids.add_internal_mode(grids.tidal_basin(14, 9, 4, seed=21, dt=10.0, ntscr=4), av=1e-3)


def _spread(rng, shape):
    """Values of both signs over thirty decades, with the specials planted."""
    v = 10.0 ** rng.uniform(-15, 15, shape) * rng.choice([-1.0, 1.0], shape)
    flat = v.reshape(-1)
    flat[:len(SPECIALS)] = SPECIALS
    return v


def _synthetic(rng):
    case = _base()
    g = case.grid
    for name in H_READ:
        case.h[name][:] = _spread(rng, g.LCM)
    for name in K_READ:
        case.k[name][:] = _spread(rng, case.k[name].shape)
    for c in case.consts:
        c["con"][:] = _spread(rng, c["con"].shape)
        c["con1"][:] = _spread(rng, c["con1"].shape)
    return case


def _host(case, rng):
    lcm, kc = case.grid.LCM, case.grid.KC
    return dict(hwq=_spread(rng, lcm), h2wq=_spread(rng, lcm),
                qq=_spread(rng, (kc + 1, lcm)), qq1=_spread(rng, (kc + 1, lcm)),
                qql=_spread(rng, (kc + 1, lcm)), qql1=_spread(rng, (kc + 1, lcm)),
                dml=_spread(rng, (kc + 1, lcm)), temb=_spread(rng, lcm))


def _slots(case):
    mo = [c["mo"] for c in case.consts]
    return (mo.index(1), -1, mo.index(3))


def _pull(ctx, ref):
    """The device values of every array the restart reads, into the
    oracle's case (FP32 storage rounds them on upload)."""
    kc = ref.grid.KC
    for name in H_READ:
        ref.h[name][:] = ctx.get(name)
    for name in K_READ:
        ref.k[name][:kc] = ctx.get(name)[:kc]
    for j, c in enumerate(ref.consts):
        con, con1 = ctx.get_constituent(j)
        c["con"][:kc] = con[:kc]
        c["con1"][:kc] = con1[:kc]


@pytest.mark.parametrize("store", [0, 1])
def test_restart_bytes_match_oracle(oracle_mod, store, tmp_path):
    from efdc_amd.lib import Context
    rng = np.random.default_rng(31 + store)
    case = _synthetic(rng)
    host = _host(case, rng)
    ref = case.copy()
    ctx = Context(case.grid, store=store).upload_case(case)
    _pull(ctx, ref)
    o = oracle_mod.Oracle(ref)
    kw = dict(isco=[1, 1, 0, 1, 0, 0, 0, 0, 0], istran=[0, 1, 0, 1, 0, 0, 0, 0, 0],
              slots=_slots(case), **host)
    for call, n in enumerate((40, 90)):
        tg = ctx.restart_text(n, 0.5 + call, **kw)
        to = o.restart_text(n, 0.5 + call, **kw)
        assert len(tg) == len(to)
        assert tg == to, call
    assert b"NaN" in tg and b"-Inf" in tg and b"0.10000000+301" in tg and b"-323" in tg
    path = str(tmp_path / "RESTART.OUT")
    ctx.restart_write(path, 7, 2.0, keep_nlo=True, **kw)
    assert open(path, "rb").read() == o.restart_text(7, 2.0, keep_nlo=True, **kw)
    ctx.close()


def test_restart_sediment_records_longer_than_a_stage(oracle_mod):
    """Cohesive sediment with 8 classes on 100 bed layers and KC = 64: a
    cell's records (SEDB, SED, SEDB1, SED1 per class, HBED, HBED1, VDRBED,
    VDRBED1) come to ~60 KB, more than a workgroup stages, so the writer
    takes the one-thread-per-value kernel; the text is the same."""
    from efdc_amd import grids
    from efdc_amd.lib import Context
    rng = np.random.default_rng(12)
    case = grids.add_internal_mode(grids.tidal_basin(5, 4, 64, seed=2), av=1e-3)
    case.conc_bc = None
    case.sources = None
    for _ in range(8):
        case.consts.append(grids._constituent(case.grid, 6, _spread(rng, case.k["U"].shape)))
    for name in ("U", "U1", "V", "V1", "QSUM"):
        case.k[name][:] = _spread(rng, case.k[name].shape)
    host = _host(case, rng)
    lcm, kb, ns = case.grid.LCM, 100, 8
    host.update(sedb=_spread(rng, (ns, kb, lcm)), sedb1=_spread(rng, (ns, kb, lcm)),
                hbed=_spread(rng, (kb, lcm)), hbed1=_spread(rng, (kb, lcm)),
                vdrbed=_spread(rng, (kb, lcm)), vdrbed1=_spread(rng, (kb, lcm)))
    ref = case.copy()
    ctx = Context(case.grid).upload_case(case)
    o = oracle_mod.Oracle(ref)
    s0, _, s2 = _slots(case)
    kw = dict(isco=[0, 1, 0, 1, 0, 0, 1, 0, 0], istran=[0, 1, 0, 1, 0, 0, 1, 0, 0],
              slots=(s0, -1, s2), sed_slots=list(range(2, 10)), kb=kb, kbm=kb, **host)
    tg = ctx.restart_text(3, 0.0, **kw)
    assert tg == o.restart_text(3, 0.0, **kw)
    assert max(len(x) for x in tg.split(b"\n")) == 50 * 17
    with pytest.raises(Exception, match="twice"):
        ctx.restart_text(3, 0.0, **dict(kw, sed_slots=[s2] + list(range(3, 10))))
    ctx.close()


def test_restart_after_internal_mode_steps(oracle_mod):
    from efdc_amd.lib import Context
    case = _base()
    ref = case.copy()
    o = oracle_mod.Oracle(ref, rsqm=1e-24, iterm=3000)
    ctx = Context(case.grid, RSQM=1e-24, ITERM=3000).upload_case(case)
    for n in range(1, 13):
        assert o.step(n, uvw={})["status"] == 0
        assert ctx.step(case.dt, n, internal=True, uvw={})["status"] == 0
    _pull(ctx, ref)
    rng = np.random.default_rng(5)
    host = _host(case, rng)
    kw = dict(isco=[0, 1, 0, 1, 0, 0, 0, 0, 0], istran=[0, 1, 0, 1, 0, 0, 0, 0, 0],
              slots=_slots(case), **host)
    lg = ctx.restart_text(12, 0.25, **kw).split(b"\n")
    lo = o.restart_text(12, 0.25, **kw).split(b"\n")
    assert len(lg) == len(lo)
    kc = case.grid.KC
    nbc = 0
    for a, b in zip(lg, lo):
        if a == b:
            continue
        # a boundary record: CLO is the transport's result, equal to the
        # oracle's to round-off
        assert len(a) == len(b) == 17 * kc, (a, b)
        x = np.array([float(a[j:j + 17]) for j in range(0, len(a), 17)])
        y = np.array([float(b[j:j + 17]) for j in range(0, len(b), 17)])
        np.testing.assert_allclose(x, y, rtol=2e-8, atol=1e-300)
        nbc += 1
    assert nbc <= 2 * len(case.conc_bc["side"][1]["L"])
    ctx.close()


@pytest.mark.parametrize("with_ksz", [False, True])
def test_ee_files_match_oracle(oracle_mod, tmp_path, with_ksz):
    from efdc_amd.lib import Context
    rng = np.random.default_rng(9)
    case = _synthetic(rng)
    g = case.grid
    case.w["W"][:] = _spread(rng, case.w["W"].shape)
    ref = case.copy()
    ctx = Context(g).upload_case(case)
    o = oracle_mod.Oracle(ref)
    rs = [rng.uniform(0, 1, g.LCM) for _ in range(4)]
    ksz = None
    if with_ksz:
        ksz = np.ones(g.LCM, np.int32)
        ksz[wet_index(case)[::3]] = rng.integers(1, g.KC + 1, len(wet_index(case)[::3]))
    dg, do = tmp_path / "gpu", tmp_path / "orc"
    os.makedirs(dg)
    os.makedirs(do)
    ctx.ee_open(str(dg), 14, 9, rs, igridv=1 if with_ksz else 0, ksz=ksz)
    o.ee_open(str(do), 14, 9, rs, igridv=1 if with_ksz else 0, ksz=ksz)
    for n in (5, 10):
        ctx.ee_append(str(dg), n, n * 10.0 / 86400.0, 10.0, 2)
        o.ee_append(str(do), n, n * 10.0 / 86400.0, 10.0, 2)
    for f in ("EE_WS.OUT", "EE_VEL.OUT"):
        assert open(dg / f, "rb").read() == open(do / f, "rb").read(), f
    ctx.close()


def test_restart_refusals(oracle_mod):
    from efdc_amd.lib import Context, EfdcError
    case = _base()
    ctx = Context(case.grid).upload_case(case)
    host = _host(case, np.random.default_rng(1))
    with pytest.raises(EfdcError, match="toxics"):
        ctx.restart_text(1, 0.0, isco=[0, 0, 0, 0, 0, 1, 0, 0, 0],
                         istran=[0, 0, 0, 0, 0, 1, 0, 0, 0], **host)
    with pytest.raises(EfdcError, match="HWQ"):
        ctx.restart_text(1, 0.0)
    with pytest.raises(EfdcError, match="TEMB"):
        h = dict(host)
        del h["temb"]
        ctx.restart_text(1, 0.0, isco=[0, 0, 1, 0, 0, 0, 0, 0, 0],
                         istran=[0, 0, 1, 0, 0, 0, 0, 0, 0], slots=(0, 0, 0), **h)
    with pytest.raises(EfdcError):
        ctx.ee_append("/nonexistent", 1, 0.0, 1.0, 0)
    ctx.close()
