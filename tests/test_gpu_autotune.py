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
    assert len(rep["shapes"]) >= 2 and rep["chosen"] in rep["shapes"]
    for j in range(len(case.consts)):
        ca, c1a = a.get_constituent(j)
        cb, c1b = b.get_constituent(j)
        assert np.array_equal(ca, cb) and np.array_equal(c1a, c1b)
    assert np.array_equal(a.get("P"), b.get("P"))
    a.close()
    b.close()
