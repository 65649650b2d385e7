"""The transport autotuner: every launch shape it times gives the same bits,
and the choice persists for the following steps."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("store", ["fp64", "fp32"])
def test_autotuned_transport_is_bitwise_the_default(store):
    from efdc_amd import grids, lib
    st = lib.STORE_FP64 if store == "fp64" else lib.STORE_FP32
    case = grids.tidal_basin(64, 32, 5, seed=8)
    a = lib.Context(case.grid, store=st).upload_case(case)
    b = lib.Context(case.grid, store=st).upload_case(case)
    a.autotune(True)
    for n in range(1, 6):
        a.step(case.dt, n)
        b.step(case.dt, n)
    rep = a.tune_report()
    print(rep)
    assert len(rep["shapes"]) >= 2 and len(rep["chosen"]) == 1
    assert rep["chosen"][0] in rep["shapes"]
    for j in range(len(case.consts)):
        ca, c1a = a.get_constituent(j)
        cb, c1b = b.get_constituent(j)
        assert np.array_equal(ca, cb) and np.array_equal(c1a, c1b)
    assert np.array_equal(a.get("P"), b.get("P"))
    a.close()
    b.close()


def test_autotuned_mixed_solve_degree():
    """With the mixed CG on the persistent solve, the tuner times the local
    preconditioner degrees from the same P, then the pipelined form at the
    fastest degree, keeps one, and the step's result stays within the
    mixed-precision bar of the FP64 solve."""
    from efdc_amd import grids, lib
    from conftest import eta_err, wet_index
    wet = grids.lake_mask(160, 160, 5)
    case = grids.external_case(wet, seed=5)
    m = lib.Context(case.grid, cg_mode=lib.CG_MIXED, RSQM=1e-20, ITERM=2000).upload_case(case)
    d = lib.Context(case.grid, cg_mode=lib.CG_FP64, RSQM=1e-24, ITERM=2000).upload_case(case)
    m.autotune(True)
    for n in (1, 2, 3):
        m.calpuv2c(case.dt, n)
        d.calpuv2c(case.dt, n)
    rep = m.tune_report()
    print(rep)
    degs = [k for k in rep["shapes"] if k.startswith("pcg_mixed degree=")]
    forms = [k for k in rep["shapes"] if k.startswith("pcg_mixed pipelined")]
    assert len(degs) == 5 and len(forms) == 1
    assert len([k for k in rep["chosen"] if k.startswith("pcg_mixed")]) == 1
    assert m.cg_pipeline() == (2 if forms[0] in rep["chosen"] else 1)
    i = wet_index(case)
    assert eta_err(m.get("P")[i], d.get("P")[i]) <= 1e-6
    m.close()
    d.close()
