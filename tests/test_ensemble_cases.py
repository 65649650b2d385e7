"""Host side of the batched ensemble: basin generator and concatenation."""
import numpy as np

from efdc_amd import grids


def test_ensemble_member_shapes_and_boundaries():
    for b in (0, 7, 100):
        m = grids.ensemble_member(b, kc=3)
        g = m.grid
        ny, nx = g.JC - 2, g.IC - 2
        assert 64 <= nx <= 512 and 64 <= ny <= 256
        L = np.arange(2, g.LA + 1) - 1
        d = -m.h["BELV"][L]
        assert d.min() >= 2.0 - 1e-12 and d.max() <= 30.0 + 1e-12
        assert 0.015 <= m.manning_n[1] <= 0.04
        amp = m.open_bc["side"][0]["pcb"][0] / grids.G
        assert 0.1 <= amp <= 1.0
        lb = m.open_bc["side"][0]["L"]
        assert len(lb) > 0
        # every west-boundary cell has land to the west and water to the east
        assert np.all(g.maps["LWC"][lb - 1] == g.LC)
        assert np.all(g.maps["LEC"][lb - 1] != g.LC)
        assert 2 <= m.sources["LQS"][0] <= g.LA


def test_concat_basins_keeps_each_basin():
    members = [grids.tidal_basin(20, 10, 2, seed=1),
               grids.ensemble_member(3, kc=2),
               grids.tidal_basin(12, 30, 2, seed=2, amp=0.2)]
    ens, first = grids.concat_basins(members)
    assert first[0] == 2 and ens.grid.nwet == sum(m.grid.nwet for m in members)
    for b, m in enumerate(members):
        sl = grids.basin_slice(ens, first, b)
        n = m.grid.nwet
        assert sl.stop - sl.start == n
        for key in ("HP", "BELV", "DXYP", "SUB", "QSUME"):
            assert np.array_equal(ens.h[key][sl], m.h[key][1:1 + n])
        assert np.array_equal(ens.k["DZC"][:, sl], m.k["DZC"][:, 1:1 + n])
        for j in range(len(m.consts)):
            assert np.array_equal(ens.consts[j]["con"][:, sl], m.consts[j]["con"][:, 1:1 + n])
        # maps: basin-local neighbours shifted, land stays land
        off = int(first[b]) - 2
        for name in ("LWC", "LEC", "LSC", "LNC"):
            old = m.grid.maps[name][1:1 + n]
            new = ens.grid.maps[name][sl]
            assert np.array_equal(new, np.where(old == m.grid.LC, ens.grid.LC, old + off))
    # boundary and source tables point at the same cells
    nb = [len(m.open_bc["side"][0]["L"]) for m in members]
    lb = ens.open_bc["side"][0]["L"]
    assert len(lb) == sum(nb)
    assert np.array_equal(lb[:nb[0]], members[0].open_bc["side"][0]["L"])
    assert np.array_equal(lb[nb[0]:nb[0] + nb[1]],
                          members[1].open_bc["side"][0]["L"] + int(first[1]) - 2)
    assert np.array_equal(ens.sources["LQS"],
                          [m.sources["LQS"][0] + int(f) - 2 for m, f in zip(members, first)])


def test_concat_rejects_mismatched_layers():
    import pytest
    with pytest.raises(ValueError):
        grids.concat_basins([grids.tidal_basin(10, 10, 2, seed=1),
                             grids.tidal_basin(10, 10, 3, seed=2)])
