"""Drying (ISDRY 1..97) on the GPU against the oracle: the re-solve loop
(EFDC/calpuv2c.f90:302,574-602), open-boundary drying (:497-530), the
face reopening at the step start (:246-253) and the transport mask
LMASKDRY (:916-939).  The cell flags and the face switches must agree
exactly and the number of re-solves NCORDRY step by step; P within the
FP64 CG's round-off."""
import numpy as np
import pytest

from conftest import eta_err, wet_index

pytestmark = pytest.mark.gpu

HDRY = 0.2


@pytest.mark.parametrize("cg_kernel", [0, 1])
@pytest.mark.parametrize("deep", [2.0, 0.6])
def test_drying_matches_oracle(oracle_mod, deep, cg_kernel):
    from efdc_amd import grids
    from efdc_amd.lib import Context
    case = grids.drying_basin(deep=deep)
    ref = case.copy()
    o = oracle_mod.Oracle(ref, rsqm=1e-24, iterm=2000, isdry=1, hdry=HDRY)
    ctx = Context(case.grid, RSQM=1e-24, ITERM=2000, ISDRY=1, HDRY=HDRY,
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
    if deep == 2.0:
        assert max(seen) >= 2
    else:
        assert obc_dried
    for s, c in enumerate(ref.consts):
        con, _ = ctx.get_constituent(s)
        scale = max(1e-30, np.max(np.abs(c["con"][:kc])))
        assert np.max(np.abs(con[:kc] - c["con"][:kc])) / scale <= 1e-8, s
    ctx.close()


def test_drying_is_refused_where_not_built():
    from efdc_amd import grids
    from efdc_amd.lib import Context, EfdcError
    case = grids.drying_basin()
    ctx = Context(case.grid, ISDRY=1).upload_case(case)
    with pytest.raises(EfdcError):
        ctx.adj_enable([0])
    ctx.close()
