"""efdc_congrad_noalloc, the allocation-free CONGRAD, against efdc_congrad.

The allocation-free entry point runs the FP64 multi-kernel Jacobi-PCG
(EFDC/congrad.f90:54-187) on arrays the caller holds in device memory, with
its scratch in a static device array.  It must give the bits of
efdc_congrad on a context solving with EFDC_CG_KERNEL_MULTI, on two calls
in a row (the second on a smaller lake, so the scratch holds the first
call's larger vectors when it starts), and stay within 1e-9 m of the
oracle's solution.  The test holds the device arrays through the HIP
runtime the library links (hipMalloc, hipMemcpy, a stream of its own for
the second call); the entry point itself allocates nothing.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import eta_err, wet_index

pytestmark = pytest.mark.gpu

TIGHT = 1e-24
COEF = ["CCC", "CCS", "CCW", "CCE", "CCN", "CCCI", "FPTMP"]
MAPS = ["LWC", "LEC", "LSC", "LNC"]
H2D, D2H = 1, 2


def _hip():
    """The HIP runtime the library itself links (already loaded with it)."""
    from efdc_amd import lib
    lib.load()
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    h.hipStreamDestroy.argtypes = [C.c_void_p]
    for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipStreamCreate", "hipStreamDestroy",
              "hipDeviceSynchronize"):
        getattr(h, f).restype = C.c_int
    return h


class _Dev:
    """Device copies of host arrays, freed together."""

    def __init__(self, hip):
        self.hip, self.ptrs = hip, []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), a.nbytes) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def get(self, p, n, dtype):
        out = np.zeros(n, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _lake_system(oracle_mod, n, seed):
    from efdc_amd import grids
    wet = grids.lake_mask(n, n, seed)
    case = grids.external_case(wet, seed=seed)
    ref = case.copy()
    o = oracle_mod.Oracle(ref, rsqm=TIGHT, iterm=4000)
    o.assemble()
    system = {k: ref.h[k].copy() for k in COEF + ["P"]}
    o.congrad()
    return case, system, ref.h["P"].copy()


def _context_solve(case, system):
    from efdc_amd import lib
    ctx = lib.Context(case.grid, RSQM=TIGHT, ITERM=4000, cg_kernel=lib.CG_KERNEL_MULTI)
    for k, v in system.items():
        ctx.set(k, v)
    out = ctx.congrad()
    p = ctx.get("P")
    ctx.close()
    return out, p


def _noalloc_solve(case, system, stream=None, maps=None):
    from efdc_amd import lib
    L = lib.load()
    g = case.grid
    hip = _hip()
    dev = _Dev(hip)
    maps = maps if maps is not None else {m: g.maps[m] for m in MAPS}
    dm = [dev.put(np.asarray(maps[m][:g.LA + 1], np.int32)) for m in MAPS]
    dc = [dev.put(np.asarray(system[k][:g.LA + 1], np.float64)) for k in COEF]
    p = dev.put(np.asarray(system["P"][:g.LA + 1], np.float64))
    st = lib.CgStats()
    rc = L.efdc_congrad_noalloc(g.LA, *dm, *dc, p, TIGHT, 4000, stream, C.byref(st))
    why = (L.efdc_last_error(None) or b"").decode()
    assert hip.hipDeviceSynchronize() == 0
    out = np.zeros(g.LCM)
    out[:g.LA + 1] = dev.get(p, g.LA + 1, np.float64)
    dev.free()
    return rc, why, st, out


def test_noalloc_matches_congrad_on_two_calls(oracle_mod):
    hip = _hip()
    for n, seed, own_stream in ((96, 11, False), (40, 12, True)):
        case, system, p_ref = _lake_system(oracle_mod, n, seed)
        out, p_ctx = _context_solve(case, system)
        s = C.c_void_p()
        if own_stream:
            assert hip.hipStreamCreate(C.byref(s)) == 0
        rc, why, st, p_na = _noalloc_solve(case, system, stream=s if own_stream else None)
        if own_stream:
            hip.hipStreamDestroy(s)
        i = wet_index(case)
        d = eta_err(p_na[i], p_ref[i])
        print(f"n={n} wet={case.grid.nwet} iters ctx {out['iters']} noalloc {st.iters} "
              f"deta vs oracle {d:.3e}")
        assert rc == 0, why
        assert st.iters == out["iters"]
        assert np.array_equal(p_na[i], p_ctx[i])
        assert d <= 1e-9


def test_noalloc_refuses_a_map_outside_the_grid(oracle_mod):
    from efdc_amd.abi import STATUS
    case, system, _ = _lake_system(oracle_mod, 24, 3)
    g = case.grid
    maps = {m: np.array(g.maps[m], np.int32) for m in MAPS}
    maps["LNC"][g.LA // 2] = g.LA + 7   # a wet row pointing past LC
    rc, why, _, _ = _noalloc_solve(case, system, maps=maps)
    assert rc == STATUS["EFDC_ERR_ARG"], why
    assert "outside" in why
    # the next call with good maps is not affected by the refused one
    out, p_ctx = _context_solve(case, system)
    rc, why, st, p_na = _noalloc_solve(case, system)
    i = wet_index(case)
    assert rc == 0, why
    assert st.iters == out["iters"] and np.array_equal(p_na[i], p_ctx[i])
