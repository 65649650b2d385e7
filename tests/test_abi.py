"""The C-ABI library loads and exports every entry point declared in
include/efdc_hip.h (no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "efdc_hip.h")


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*\**\s*(efdc_\w+)\s*\(",
                                 text, flags=re.M)))


def test_header_declares_entry_points():
    fns = declared_functions()
    for need in ("efdc_ctx_create", "efdc_congrad", "efdc_calpuv2c",
                 "efdc_calconc", "efdc_balance", "efdc_set_array"):
        assert need in fns


def test_library_exports_every_declared_symbol():
    from efdc_amd.lib import LIB_PATH
    assert os.path.exists(LIB_PATH), "build the library first"
    lib = ctypes.CDLL(LIB_PATH)
    missing = [f for f in declared_functions() if not hasattr(lib, f)]
    assert not missing, missing


def test_abi_version_and_defaults():
    from efdc_amd import lib
    L = lib.load()
    assert L.efdc_abi_version() == 1
    p = lib.default_params()
    assert p.G == 9.81 and p.ITERM == 200 and p.RSQM == 1e-9
    assert p.cg_mode == lib.CG_FP64 and p.store == lib.STORE_FP64


def test_invalid_params_are_refused_before_any_device_call():
    """efdc_ctx_create checks the parameters before it looks for a device,
    so a bad value is EFDC_ERR_ARG on any machine."""
    from efdc_amd import grids, lib
    from efdc_amd.abi import STATUS
    case = grids.tidal_basin(8, 4, 2)
    for bad in (dict(tran_shape=3), dict(tran_shape=-3), dict(tran_shape=8), dict(tran_shape=36),
                dict(tran_shape=-17), dict(tran_shape=48),
                dict(mixed_precond_degree=17), dict(cg_kernel=7), dict(store=5),
                dict(ISDRY=100), dict(ISDRY=-1)):
        with pytest.raises(lib.EfdcError) as e:
            lib.Context(case.grid, **bad)
        assert e.value.status == STATUS["EFDC_ERR_ARG"], bad
    assert lib.default_params().tran_shape == 0


def test_enum_values_match_layout_classes():
    from efdc_amd.abi import ARRAY_IDS, shape_class
    from efdc_amd import grids
    for k in grids.H_KEYS:
        assert shape_class(k) == "h", k
    for k in grids.K_KEYS:
        assert shape_class(k) == "k", k
    for k in grids.W_KEYS:
        assert shape_class(k) == "w", k
    assert len(set(ARRAY_IDS.values())) == len(ARRAY_IDS)


def test_context_without_device_fails_loudly():
    """No CPU fallback: on a machine without a GPU, creating a context
    reports EFDC_ERR_NODEVICE instead of computing anything."""
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    from efdc_amd import grids, lib
    from efdc_amd.abi import STATUS
    case = grids.tidal_basin(8, 4, 2)
    with pytest.raises(lib.EfdcError) as e:
        lib.Context(case.grid)
    assert e.value.status == STATUS["EFDC_ERR_NODEVICE"]


def test_fortran_module_binds_every_declared_symbol():
    """MODULE EFDC_HIP (ISO_C_BINDING) has an interface for every entry
    point of the header, so a Fortran host can call all of them."""
    f90 = os.path.join(ROOT, "efdcplus8.5_amd", "fortran", "efdc_hip.f90")
    bound = set(re.findall(r'name="(efdc_\w+)"', open(f90).read()))
    missing = [f for f in declared_functions() if f not in bound]
    assert not missing, missing


def test_noalloc_congrad_refuses_sizes_beyond_its_scratch():
    """efdc_congrad_noalloc checks LA against its static scratch before any
    device call: too large (or too small) is EFDC_ERR_ARG with a reason."""
    import ctypes as C
    from efdc_amd import lib
    from efdc_amd.abi import STATUS
    L = lib.load()
    st = lib.CgStats()
    dummy = [1] * 12  # never dereferenced: the size check comes first
    for la in (L.efdc_congrad_noalloc and 1048576, 5_000_000, 1):
        rc = L.efdc_congrad_noalloc(la, *dummy, 1e-9, 100, None, C.byref(st))
        assert rc == STATUS["EFDC_ERR_ARG"]
        assert b"LA+1" in L.efdc_last_error(None)
