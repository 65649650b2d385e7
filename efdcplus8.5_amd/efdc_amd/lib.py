"""ctypes binding of libefdc_hip.so -- the Python side of the C-ABI.

Mirrors the reference's hot-path calls: ``Context.calpuv2c`` (CALPUV2C,
EFDC/calpuv2c.f90), ``Context.congrad`` (CONGRAD, EFDC/congrad.f90) and
``Context.calconc`` (CALCONC/CALTRAN, EFDC/calconc.f90, EFDC/caltran.f90).
There is no CPU fallback: without the built library or without a GPU the
calls raise.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import sys
import weakref
from typing import Optional

import numpy as np

from . import formats as _fmt
from .abi import ARRAY_IDS, STATUS_NAMES, shape_class

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libefdc_hip.so")

PD = C.POINTER(C.c_double)
PI = C.POINTER(C.c_int32)

CG_FP64, CG_MIXED = 0, 1
CG_KERNEL_AUTO, CG_KERNEL_MULTI, CG_KERNEL_PERSISTENT = 0, 1, 2
STORE_FP64, STORE_FP32 = 0, 1


class EfdcError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class Grid(C.Structure):
    _fields_ = [("LA", C.c_int32), ("LCM", C.c_int32), ("KC", C.c_int32),
                ("KCM", C.c_int32), ("LWC", PI), ("LEC", PI), ("LSC", PI),
                ("LNC", PI), ("LNWC", PI), ("LSEC", PI)]


class Params(C.Structure):
    _fields_ = [("G", C.c_double), ("RITB1", C.c_double), ("SNLT", C.c_double),
                ("HDRY", C.c_double), ("TCON_TBEGIN", C.c_double),
                ("ITERM", C.c_int32), ("RSQM", C.c_double),
                ("cg_rel_tol", C.c_double), ("cg_mode", C.c_int32),
                ("store", C.c_int32), ("mixed_inner_rtol", C.c_double),
                ("mixed_max_outer", C.c_int32), ("cg_kernel", C.c_int32),
                ("mixed_precond_degree", C.c_int32), ("tran_shape", C.c_int32),
                ("ISDRY", C.c_int32), ("NDRYSTP", C.c_int32), ("implicit_drag", C.c_int32)]


class PbcSide(C.Structure):
    _fields_ = [("n", C.c_int32), ("L", PI), ("ispb", PI), ("pser", PD),
                ("pcb", PD), ("psb", PD)]


class OpenBC(C.Structure):
    _fields_ = [("side", PbcSide * 4), ("mtide", C.c_int32), ("tcp", PD)]


class CbcSide(C.Structure):
    _fields_ = [("n", C.c_int32), ("L", PI), ("cb", PD), ("csert", PD),
                ("ntscr", PI)]


class ConcBC(C.Structure):
    _fields_ = [("side", CbcSide * 4), ("wtci", PD)]


class Sources(C.Structure):
    _fields_ = [("n", C.c_int32), ("LQS", PI), ("qss", PD), ("qsercell", PD),
                ("cqs", PD), ("csert", PD)]


class ConstDesc(C.Structure):
    _fields_ = [("mo", C.c_int32), ("isadac", C.c_int32), ("isfct", C.c_int32),
                ("rkdye", C.c_double), ("is_dye", C.c_int32), ("wset", C.c_double),
                ("probdep", C.c_double), ("wese", C.c_double)]


class Step(C.Structure):
    _fields_ = [("dt", C.c_double), ("n", C.c_int64), ("niter", C.c_int64)]


class CgStats(C.Structure):
    _fields_ = [("iters", C.c_int32), ("outer", C.c_int32), ("rsq", C.c_double),
                ("rsqm_used", C.c_double), ("ccmnm", C.c_double)]


class TimerEntry(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("count", C.c_int64),
                ("total_ms", C.c_double)]


class Timers(C.Structure):
    _fields_ = [("n", C.c_int32), ("k", TimerEntry * 64)]


class Balance(C.Structure):
    _fields_ = [("volume", C.c_double), ("mass", C.c_double * 16),
                ("nconst", C.c_int32)]


class TuneEntry(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("ms", C.c_float), ("chosen", C.c_int32)]


class TuneReport(C.Structure):
    _fields_ = [("n", C.c_int32), ("v", TuneEntry * 8)]


class UvwOpts(C.Structure):
    _fields_ = [("AVCON1", C.c_double), ("shear", C.c_int32), ("wind", C.c_int32),
                ("uv_after", C.c_int32)]


class Bal2t(C.Structure):
    _fields_ = [("nconst", C.c_int32), ("nsteps", C.c_int64), ("volbeg", C.c_double),
                ("volend", C.c_double), ("volout", C.c_double), ("volerr", C.c_double),
                ("massbeg", C.c_double * 16), ("massend", C.c_double * 16),
                ("massout", C.c_double * 16), ("masserr", C.c_double * 16)]


class AdjOpts(C.Structure):
    _fields_ = [("nadj", C.c_int32), ("slot", C.c_int32 * 16), ("cg_rel_tol", C.c_double),
                ("cg_itmax", C.c_int32)]


class AdjStats(C.Structure):
    _fields_ = [("cg_iters", C.c_int32), ("cg_rsq", C.c_double)]


ADJ_MANNING, ADJ_HP, ADJ_UHDYE, ADJ_VHDXE, ADJ_P, ADJ_HU, ADJ_HV = range(7)
ADJ_CON = 16

_lib = None
# contexts not closed yet: closed by close_all() at interpreter exit, while
# the HIP runtime is still whole (a context destroyed from __del__ during
# interpreter teardown would free device memory and destroy its stream
# after the runtime's, or a profiler's, own exit handlers may have run)
_live = weakref.WeakSet()


def close_all():
    """Destroy every context that is still open (registered with atexit
    when the library is loaded)."""
    for ctx in list(_live):
        ctx.close()


def load(path: Optional[str] = None):
    """Load the library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    # EFDC_HIP_LIB selects another build of the same library (for example
    # an instrumented one); the default is the in-tree build
    path = path or os.environ.get("EFDC_HIP_LIB") or LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with __graft_entry__.build() "
            "(there is no CPU fallback)")
    lib = C.CDLL(path)
    P = C.c_void_p
    sig = {
        "efdc_abi_version": ([], C.c_int),
        "efdc_params_default": ([C.POINTER(Params)], None),
        "efdc_ctx_create": ([C.POINTER(Grid), C.POINTER(Params), C.POINTER(P)], C.c_int),
        "efdc_ctx_destroy": ([P], C.c_int),
        "efdc_last_error": ([P], C.c_char_p),
        "efdc_set_params": ([P, C.POINTER(Params)], C.c_int),
        "efdc_set_array": ([P, C.c_int32, PD], C.c_int),
        "efdc_get_array": ([P, C.c_int32, PD], C.c_int),
        "efdc_set_open_bc": ([P, C.POINTER(OpenBC)], C.c_int),
        "efdc_set_conc_bc": ([P, C.POINTER(ConcBC)], C.c_int),
        "efdc_set_sources": ([P, C.POINTER(Sources)], C.c_int),
        "efdc_add_constituent": ([P, C.POINTER(ConstDesc), PI], C.c_int),
        "efdc_set_constituent": ([P, C.c_int32, PD, PD], C.c_int),
        "efdc_get_constituent": ([P, C.c_int32, PD, PD], C.c_int),
        "efdc_get_constituent_at": ([P, C.c_int32, C.c_int32, PI, PI, PD], C.c_int),
        "efdc_get_dry_state": ([P, PI, PI, PI], C.c_int),
        "efdc_get_dry_waste": ([P, PD, PD, PI], C.c_int),
        "efdc_congrad": ([P, C.POINTER(CgStats)], C.c_int),
        "efdc_calpuv2c": ([P, C.POINTER(Step), C.POINTER(CgStats)], C.c_int),
        "efdc_calconc": ([P, C.POINTER(Step)], C.c_int),
        "efdc_caluvw": ([P, C.POINTER(Step), C.POINTER(UvwOpts)], C.c_int),
        "efdc_bal2t_begin": ([P], C.c_int),
        "efdc_set_segment_tile": ([P, C.c_int32], C.c_int),
        "efdc_set_cg_replay": ([P, C.c_int32], C.c_int),
        "efdc_get_cg_replay": ([P, C.POINTER(C.c_int32)], C.c_int),
        "efdc_set_cg_pipeline": ([P, C.c_int32], C.c_int),
        "efdc_get_cg_pipeline": ([P, C.POINTER(C.c_int32)], C.c_int),
        "efdc_set_cg_launch": ([P, C.c_int32, C.c_uint32], C.c_int),
        "efdc_set_dry_counters": ([P, PD, PI], C.c_int),
        "efdc_set_cg64_pipeline": ([P, C.c_int32], C.c_int),
        "efdc_get_cg64_pipeline": ([P, C.POINTER(C.c_int32)], C.c_int),
        "efdc_get_cg_launch": ([P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)], C.c_int),
        "efdc_get_segment_tile": ([P, C.POINTER(C.c_int32)], C.c_int),
        "efdc_restart_text": ([P, C.POINTER(_fmt.Restart), C.c_char_p, C.c_int64,
                               C.POINTER(C.c_int64)], C.c_int),
        "efdc_restart_write": ([P, C.POINTER(_fmt.Restart), C.c_char_p], C.c_int),
        "efdc_ee_open": ([P, C.c_char_p, C.POINTER(_fmt.EeGrid)], C.c_int),
        "efdc_ee_append": ([P, C.c_char_p, C.c_int32, C.c_double, C.c_double, C.c_int32],
                           C.c_int),
        "efdc_bal2t_end": ([P, C.POINTER(Bal2t)], C.c_int),
        "efdc_balance": ([P, C.POINTER(Balance)], C.c_int),
        "efdc_synchronize": ([P], C.c_int),
        "efdc_device_ptr": ([P, C.c_int32, C.POINTER(P), C.POINTER(C.c_int64)], C.c_int),
        "efdc_closure_set_manning": ([P, PD], C.c_int),
        "efdc_closure_pre": ([P], C.c_int),
        "efdc_closure_post": ([P], C.c_int),
        "efdc_timers_enable": ([P, C.c_int32], C.c_int),
        "efdc_timers_read": ([P, C.POINTER(Timers)], C.c_int),
        "efdc_set_source_cqs": ([P, PD], C.c_int),
        "efdc_stream_copy_gbs": ([P, C.c_size_t, C.c_int32, PD], C.c_int),
        "efdc_autotune": ([P, C.c_int32], C.c_int),
        "efdc_tune_report": ([P, C.POINTER(TuneReport)], C.c_int),
        "efdc_state_bytes": ([P, C.POINTER(C.c_size_t)], C.c_int),
        "efdc_state_save": ([P, P], C.c_int),
        "efdc_adj_record_bytes": ([P, C.POINTER(C.c_size_t)], C.c_int),
        "efdc_adj_record_save": ([P, P], C.c_int),
        "efdc_adj_record_load": ([P, P], C.c_int),
        "efdc_state_load": ([P, P], C.c_int),
        "efdc_state_alloc": ([P, C.POINTER(P)], C.c_int),
        "efdc_state_free": ([P, P], C.c_int),
        "efdc_adj_enable": ([P, C.POINTER(AdjOpts)], C.c_int),
        "efdc_adj_disable": ([P], C.c_int),
        "efdc_adj_reset": ([P], C.c_int),
        "efdc_adj_seed_con": ([P, C.c_int32, C.c_int32, PI, PI, PD], C.c_int),
        "efdc_adj_step": ([P, C.POINTER(Step), C.POINTER(AdjStats)], C.c_int),
        "efdc_adj_source_grad": ([P, PD], C.c_int),
        "efdc_adj_get": ([P, C.c_int32, PD], C.c_int),
        "efdc_set_segments": ([P, C.c_int32, PI], C.c_int),
        "efdc_segment_stats": ([P, C.c_int32, PI, PI], C.c_int),
        "efdc_balance_segments": ([P, C.c_int32, PD, PD], C.c_int),
        # device pointers (maps, 7 coefficient arrays, P), RSQM, ITERM, stream, stats
        "efdc_congrad_noalloc": ([C.c_int32] + [P] * 12 + [C.c_double, C.c_int32, P,
                                                          C.POINTER(CgStats)], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    _lib = lib
    atexit.register(close_all)
    return lib


def default_params(**kw) -> Params:
    p = Params()
    load().efdc_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _pd(a: np.ndarray):
    return a.ctypes.data_as(PD)


def _pi(a: np.ndarray):
    return a.ctypes.data_as(PI)


class Context:
    """One device context holding a case's arrays in HBM."""

    def __init__(self, grid, params: Optional[Params] = None, **kw):
        self.lib = load()
        self.grid = grid
        self._keep = []
        g = Grid()
        g.LA, g.LCM, g.KC, g.KCM = grid.LA, grid.LCM, grid.KC, grid.KCM
        for m in ("LWC", "LEC", "LSC", "LNC", "LNWC", "LSEC"):
            a = np.ascontiguousarray(grid.maps[m], np.int32)
            self._keep.append(a)
            setattr(g, m, _pi(a))
        self.params = params if params is not None else default_params(**kw)
        h = C.c_void_p()
        st = self.lib.efdc_ctx_create(C.byref(g), C.byref(self.params), C.byref(h))
        if st != 0:
            why = (self.lib.efdc_last_error(None) or b"").decode()
            raise EfdcError(st, why)
        self.h = h
        self.nconst = 0
        self.has_manning = False
        _live.add(self)

    # ----------------------------------------------------------- errors
    def _check(self, st: int, ctx="self"):
        if st != 0:
            msg = ""
            if ctx == "self" and getattr(self, "h", None):
                msg = (self.lib.efdc_last_error(self.h) or b"").decode()
            raise EfdcError(st, msg)

    def close(self):
        if getattr(self, "h", None):
            self.lib.efdc_ctx_destroy(self.h)
            self.h = None
        _live.discard(self)

    def __del__(self):
        # during interpreter teardown the atexit hook has already closed
        # every context; no HIP call is made from a finaliser then
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ----------------------------------------------------------- arrays
    def set(self, name: str, arr: np.ndarray):
        a = np.ascontiguousarray(arr, np.float64)
        self._check(self.lib.efdc_set_array(self.h, ARRAY_IDS[name], _pd(a)))

    def get(self, name: str) -> np.ndarray:
        g = self.grid
        cls = shape_class(name)
        shape = {"h": (g.LCM,), "k": (g.KCM, g.LCM), "w": (g.KCM + 1, g.LCM)}[cls]
        out = np.zeros(shape)
        self._check(self.lib.efdc_get_array(self.h, ARRAY_IDS[name], _pd(out)))
        return out

    def set_params(self, **kw):
        for k, v in kw.items():
            setattr(self.params, k, v)
        self._check(self.lib.efdc_set_params(self.h, C.byref(self.params)))

    # ----------------------------------------------------------- tables
    def set_open_bc(self, obc: dict):
        o = OpenBC()
        mt = int(obc["mtide"])
        o.mtide = mt
        tcp = np.ascontiguousarray(obc["tcp"], np.float64)
        self._keep.append(tcp)
        o.tcp = _pd(tcp)
        for s, side in enumerate(obc["side"]):
            if side is None:
                continue
            arrs = [np.ascontiguousarray(side["L"], np.int32),
                    np.ascontiguousarray(side["ispb"], np.int32),
                    np.ascontiguousarray(side["pser"], np.float64),
                    np.ascontiguousarray(side["pcb"], np.float64).reshape(-1),
                    np.ascontiguousarray(side["psb"], np.float64).reshape(-1)]
            self._keep.append(arrs)
            o.side[s].n = len(arrs[0])
            o.side[s].L, o.side[s].ispb = _pi(arrs[0]), _pi(arrs[1])
            o.side[s].pser, o.side[s].pcb, o.side[s].psb = (
                _pd(arrs[2]), _pd(arrs[3]), _pd(arrs[4]))
        self._check(self.lib.efdc_set_open_bc(self.h, C.byref(o)))

    def set_conc_bc(self, cbc: dict):
        o = ConcBC()
        w = np.ascontiguousarray(cbc["wtci"], np.float64).reshape(-1)
        self._keep.append(w)
        o.wtci = _pd(w)
        for s, side in enumerate(cbc["side"]):
            if side is None:
                continue
            arrs = [np.ascontiguousarray(side["L"], np.int32),
                    np.ascontiguousarray(side["cb"], np.float64).reshape(-1),
                    np.ascontiguousarray(side["csert"], np.float64).reshape(-1),
                    np.ascontiguousarray(side["ntscr"], np.int32)]
            self._keep.append(arrs)
            o.side[s].n = len(arrs[0])
            o.side[s].L, o.side[s].cb = _pi(arrs[0]), _pd(arrs[1])
            o.side[s].csert, o.side[s].ntscr = _pd(arrs[2]), _pi(arrs[3])
        self._check(self.lib.efdc_set_conc_bc(self.h, C.byref(o)))

    def set_sources(self, src: dict):
        o = Sources()
        arrs = [np.ascontiguousarray(src["LQS"], np.int32),
                np.ascontiguousarray(src["qss"], np.float64).reshape(-1),
                np.ascontiguousarray(src["qsercell"], np.float64).reshape(-1),
                np.ascontiguousarray(src["cqs"], np.float64).reshape(-1),
                np.ascontiguousarray(src["csert"], np.float64).reshape(-1)]
        self._keep.append(arrs)
        o.n = len(arrs[0])
        o.LQS, o.qss, o.qsercell = _pi(arrs[0]), _pd(arrs[1]), _pd(arrs[2])
        o.cqs, o.csert = _pd(arrs[3]), _pd(arrs[4])
        self._check(self.lib.efdc_set_sources(self.h, C.byref(o)))

    def add_constituent(self, mo, isadac=0, isfct=0, rkdye=0.0, is_dye=0, wset=0.0,
                        probdep=0.0, wese=0.0) -> int:
        d = ConstDesc(mo=mo, isadac=isadac, isfct=isfct, rkdye=rkdye, is_dye=is_dye,
                      wset=wset, probdep=probdep, wese=wese)
        slot = C.c_int32(-1)
        self._check(self.lib.efdc_add_constituent(self.h, C.byref(d), C.byref(slot)))
        self.nconst += 1
        return slot.value

    def set_constituent(self, slot, con, con1=None):
        a = np.ascontiguousarray(con, np.float64)
        b = None if con1 is None else np.ascontiguousarray(con1, np.float64)
        self._check(self.lib.efdc_set_constituent(
            self.h, slot, _pd(a), None if b is None else _pd(b)))

    def get_constituent(self, slot):
        g = self.grid
        con = np.zeros((g.KCM, g.LCM))
        con1 = np.zeros((g.KCM, g.LCM))
        self._check(self.lib.efdc_get_constituent(self.h, slot, _pd(con), _pd(con1)))
        return con, con1

    def get_constituent_at(self, slot, L, K) -> np.ndarray:
        """CON of constituent `slot` at cells (L, K) (1-based), FP64."""
        L = np.ascontiguousarray(L, np.int32)
        K = np.ascontiguousarray(K, np.int32)
        out = np.zeros(len(L))
        self._check(self.lib.efdc_get_constituent_at(self.h, slot, len(L), _pi(L), _pi(K),
                                                     _pd(out)))
        return out

    def get_dry_state(self, arrays=True):
        """(lmaskdry, iscdry, ncordry) after the last calpuv2c: per-cell
        arrays in L order (index L-1), 1 = wet in lmaskdry, 1 = dried this
        step in iscdry (None, None without `arrays`); ncordry = re-solves of
        that step."""
        n = self.grid.LCM
        m = np.zeros(n, np.int32) if arrays else None
        d = np.zeros(n, np.int32) if arrays else None
        k = np.zeros(1, np.int32)
        self._check(self.lib.efdc_get_dry_state(
            self.h, _pi(m) if arrays else None, _pi(d) if arrays else None, _pi(k)))
        return m, d, int(k[0])

    def get_dry_waste(self):
        """(qdwaste, vdwaste, natdry) per cell in L order (index L-1): the
        drained rate of the last step (m^3/s), the drained volume so far
        (m^3) and the isolated-step counters (NDRYSTP < 0)."""
        n = self.grid.LCM
        q, v, k = np.zeros(n), np.zeros(n), np.zeros(n, np.int32)
        self._check(self.lib.efdc_get_dry_waste(self.h, _pd(q), _pd(v), _pi(k)))
        return q, v, k

    def set_dry_counters(self, vdwaste=None, natdry=None):
        """Drained volume (m^3) and isolated-step counters per cell in L
        order (index L-1), e.g. from a restart; either may be None."""
        v = None if vdwaste is None else np.ascontiguousarray(vdwaste, np.float64)
        k = None if natdry is None else np.ascontiguousarray(natdry, np.int32)
        for a in (v, k):
            if a is not None and a.shape != (self.grid.LCM,):
                raise ValueError("set_dry_counters: arrays of LCM entries")
        self._check(self.lib.efdc_set_dry_counters(
            self.h, None if v is None else _pd(v), None if k is None else _pi(k)))

    # ------------------------------------------------------------ calls
    def congrad(self) -> dict:
        s = CgStats()
        self._check(self.lib.efdc_congrad(self.h, C.byref(s)))
        return dict(iters=s.iters, outer=s.outer, rsq=s.rsq, rsqm=s.rsqm_used)

    def calpuv2c(self, dt: float, n: int, niter: Optional[int] = None,
                 check: bool = True) -> dict:
        st = Step(dt=dt, n=n, niter=n if niter is None else niter)
        s = CgStats()
        code = self.lib.efdc_calpuv2c(self.h, C.byref(st), C.byref(s))
        if check:
            self._check(code)
        return dict(iters=s.iters, outer=s.outer, rsq=s.rsq, rsqm=s.rsqm_used,
                    ccmnm=s.ccmnm, status=code)

    def calconc(self, dt: float, n: int, niter: Optional[int] = None):
        st = Step(dt=dt, n=n, niter=n if niter is None else niter)
        self._check(self.lib.efdc_calconc(self.h, C.byref(st)))

    def bal2t_begin(self):
        """BAL2T1: open a balance period (EFDC/bal2t.f90:8-176)."""
        self._check(self.lib.efdc_bal2t_begin(self.h))

    def bal2t_end(self) -> dict:
        """BAL2T5: close it and return the totals, outflows and errors."""
        b = Bal2t()
        self._check(self.lib.efdc_bal2t_end(self.h, C.byref(b)))
        nc = b.nconst
        return dict(nsteps=b.nsteps, volbeg=b.volbeg, volend=b.volend, volout=b.volout,
                    volerr=b.volerr, massbeg=np.array(b.massbeg[:nc]),
                    massend=np.array(b.massend[:nc]), massout=np.array(b.massout[:nc]),
                    masserr=np.array(b.masserr[:nc]))

    def restart_text(self, n: int, timeday: float, **kw) -> bytes:
        """RESTOUT's RESTART.OUT text (EFDC/restartmod.f90:17-272) from the
        device state; keywords as efdc_amd.formats.restart_opts."""
        r, keep = _fmt.restart_opts(n, timeday, **kw)
        nb = C.c_int64(0)
        self._check(self.lib.efdc_restart_text(self.h, C.byref(r), None, 0, C.byref(nb)))
        buf = C.create_string_buffer(max(nb.value, 1))
        self._check(self.lib.efdc_restart_text(self.h, C.byref(r), buf, nb.value, C.byref(nb)))
        del keep
        return buf.raw[:nb.value]

    def restart_write(self, path: str, n: int, timeday: float, **kw):
        r, keep = _fmt.restart_opts(n, timeday, **kw)
        self._check(self.lib.efdc_restart_write(self.h, C.byref(r), path.encode()))
        del keep

    def ee_open(self, dirname: str, ic: int, jc: int, rssbc, igridv: int = 0, ksz=None):
        """HEADEROUT for EE_WS.OUT and EE_VEL.OUT (EFDC/efdcout.f90:73-113)."""
        g, keep = _fmt.ee_grid(ic, jc, rssbc, igridv, ksz)
        self._check(self.lib.efdc_ee_open(self.h, dirname.encode(), C.byref(g)))
        del keep

    def ee_append(self, dirname: str, n: int, eetime: float, delt: float, lminstep: int = 0):
        """WSOUT and VELOUT (EFDC/efdcout.f90:371-497)."""
        self._check(self.lib.efdc_ee_append(self.h, dirname.encode(), n, eetime, delt, lminstep))

    def balance(self):
        b = Balance()
        self._check(self.lib.efdc_balance(self.h, C.byref(b)))
        return b.volume, np.array(b.mass[:b.nconst])

    # ------------------------------------------------ snapshots, adjoint
    def state_bytes(self) -> int:
        n = C.c_size_t(0)
        self._check(self.lib.efdc_state_bytes(self.h, C.byref(n)))
        return n.value

    def state_save(self, dev_ptr: int):
        self._check(self.lib.efdc_state_save(self.h, C.c_void_p(dev_ptr)))

    def state_load(self, dev_ptr: int):
        self._check(self.lib.efdc_state_load(self.h, C.c_void_p(dev_ptr)))

    def autotune(self, on: bool = True):
        self._check(self.lib.efdc_autotune(self.h, 1 if on else 0))

    def tune_report(self) -> dict:
        r = TuneReport()
        self._check(self.lib.efdc_tune_report(self.h, C.byref(r)))
        shapes = {r.v[j].name.decode(): r.v[j].ms for j in range(r.n)}
        chosen = [r.v[j].name.decode() for j in range(r.n) if r.v[j].chosen]
        return dict(shapes=shapes, chosen=chosen)

    def stream_copy_gbs(self, mib: int = 1024, reps: int = 10) -> float:
        """Achievable HBM GB/s of a plain streaming copy of mib MiB."""
        g = C.c_double(0.0)
        self._check(self.lib.efdc_stream_copy_gbs(self.h, mib << 20, reps, C.byref(g)))
        return g.value

    def adj_record_bytes(self) -> int:
        n = C.c_size_t(0)
        self._check(self.lib.efdc_adj_record_bytes(self.h, C.byref(n)))
        return n.value

    def adj_record_save(self, dev_ptr: int):
        self._check(self.lib.efdc_adj_record_save(self.h, C.c_void_p(dev_ptr)))

    def adj_record_load(self, dev_ptr: int):
        self._check(self.lib.efdc_adj_record_load(self.h, C.c_void_p(dev_ptr)))

    def state_alloc(self) -> int:
        """A device buffer for one snapshot (free it with state_free); with
        the adjoint enabled it also has room for a step's recording at
        offset state_bytes()."""
        b = C.c_void_p()
        self._check(self.lib.efdc_state_alloc(self.h, C.byref(b)))
        return b.value

    def state_free(self, dev_ptr: int):
        self._check(self.lib.efdc_state_free(self.h, C.c_void_p(dev_ptr)))

    def set_source_cqs(self, cqs):
        a = np.ascontiguousarray(cqs, np.float64).reshape(-1)
        self._check(self.lib.efdc_set_source_cqs(self.h, _pd(a)))

    def adj_enable(self, slots, cg_rel_tol=1e-12, cg_itmax=0):
        o = AdjOpts(nadj=len(slots), cg_rel_tol=cg_rel_tol, cg_itmax=cg_itmax)
        for j, s_ in enumerate(slots):
            o.slot[j] = s_
        self._check(self.lib.efdc_adj_enable(self.h, C.byref(o)))

    def adj_reset(self):
        self._check(self.lib.efdc_adj_reset(self.h))

    def adj_seed_con(self, slot, L, K, val):
        L = np.ascontiguousarray(L, np.int32)
        K = np.ascontiguousarray(K, np.int32)
        v = np.ascontiguousarray(val, np.float64)
        self._check(self.lib.efdc_adj_seed_con(self.h, slot, len(L), _pi(L), _pi(K), _pd(v)))

    def adj_step(self, dt: float, n: int, niter: Optional[int] = None) -> dict:
        st = Step(dt=dt, n=n, niter=n if niter is None else niter)
        s = AdjStats()
        self._check(self.lib.efdc_adj_step(self.h, C.byref(st), C.byref(s)))
        return dict(cg_iters=s.cg_iters, cg_rsq=s.cg_rsq)

    def adj_source_grad(self, nsrc: int):
        g = self.grid
        out = np.zeros(nsrc * g.KC * max(self.nconst, 1))
        self._check(self.lib.efdc_adj_source_grad(self.h, _pd(out)))
        return out.reshape(nsrc, g.KC, max(self.nconst, 1))

    def adj_get(self, field: int):
        g = self.grid
        shape = (g.KCM, g.LCM) if field >= ADJ_CON else (g.LCM,)
        out = np.zeros(shape)
        self._check(self.lib.efdc_adj_get(self.h, field, _pd(out)))
        return out

    # ------------------------------------------- independent basins
    def set_segments(self, first_L):
        """Declare the basins of a concatenated case by their first L."""
        a = np.ascontiguousarray(first_L, np.int32)
        self._check(self.lib.efdc_set_segments(self.h, len(a), _pi(a)))
        self.nseg = len(a)

    def set_segment_tile(self, cells: int):
        """CG tile of the basins: 256, 512 (bitwise, the default), 1024,
        2048, or 0 to time the four on the next solve."""
        self._check(self.lib.efdc_set_segment_tile(self.h, cells))

    def set_cg_replay(self, mode: int):
        """0 launch by launch, 1 captured graphs, 2 time both on the next
        solve (FP64 multi-kernel CG)."""
        self._check(self.lib.efdc_set_cg_replay(self.h, mode))

    def cg_replay(self) -> int:
        v = C.c_int32(0)
        self._check(self.lib.efdc_get_cg_replay(self.h, C.byref(v)))
        return v.value

    def set_cg_pipeline(self, mode: int):
        """Form of the mixed persistent inner solve: 0 tuned, 1 the
        Chronopoulos-Gear form, 2 its pipelined form."""
        self._check(self.lib.efdc_set_cg_pipeline(self.h, mode))

    def cg_pipeline(self) -> int:
        v = C.c_int32(0)
        self._check(self.lib.efdc_get_cg_pipeline(self.h, C.byref(v)))
        return v.value

    def set_cg_launch(self, mode: int, spin_limit: int = 0):
        """Launch of the persistent solves: 0 plain (repeated cooperatively
        if it was not co-resident), 1 cooperative; spin_limit > 0 shortens the
        plain launches' waits (tests of the repetition)."""
        self._check(self.lib.efdc_set_cg_launch(self.h, mode, spin_limit))

    def set_cg64_pipeline(self, on: bool):
        """FP64 persistent solve in the pipelined form (one exchange per
        iteration) instead of the reference's recurrences."""
        self._check(self.lib.efdc_set_cg64_pipeline(self.h, 1 if on else 0))

    def cg64_pipeline(self) -> bool:
        v = C.c_int32(0)
        self._check(self.lib.efdc_get_cg64_pipeline(self.h, C.byref(v)))
        return bool(v.value)

    def cg_launch(self) -> tuple:
        """(mode, solves repeated cooperatively so far)."""
        m, r = C.c_int32(0), C.c_int32(0)
        self._check(self.lib.efdc_get_cg_launch(self.h, C.byref(m), C.byref(r)))
        return m.value, r.value

    def segment_tile(self) -> int:
        v = C.c_int32(0)
        self._check(self.lib.efdc_get_segment_tile(self.h, C.byref(v)))
        return v.value

    def segment_stats(self):
        """Per-basin CG iterations and status of the last solve."""
        n = self.nseg
        it = np.zeros(n, np.int32)
        stt = np.zeros(n, np.int32)
        self._check(self.lib.efdc_segment_stats(self.h, n, _pi(it), _pi(stt)))
        return it, stt

    def balance_segments(self):
        """Per-basin volume (nseg,) and constituent mass (nseg, nconst)."""
        n = self.nseg
        vol = np.zeros(n)
        mass = np.zeros((n, max(self.nconst, 1)))
        self._check(self.lib.efdc_balance_segments(self.h, n, _pd(vol), _pd(mass)))
        return vol, mass[:, :self.nconst]

    def synchronize(self):
        self._check(self.lib.efdc_synchronize(self.h))

    def timers(self, on: bool):
        self._check(self.lib.efdc_timers_enable(self.h, 1 if on else 0))

    def read_timers(self) -> dict:
        """{kernel name: (launches, total ms)} since the last read."""
        t = Timers()
        self._check(self.lib.efdc_timers_read(self.h, C.byref(t)))
        return {t.k[j].name.decode(): (t.k[j].count, t.k[j].total_ms)
                for j in range(t.n) if t.k[j].count > 0}

    def closure_pre(self):
        self._check(self.lib.efdc_closure_pre(self.h))

    def closure_post(self):
        self._check(self.lib.efdc_closure_post(self.h))

    def set_manning(self, n):
        a = np.ascontiguousarray(n, np.float64)
        self._check(self.lib.efdc_closure_set_manning(self.h, _pd(a)))
        self.has_manning = True

    # ------------------------------------------------------- whole case
    def upload_case(self, case):
        """Copy every array and table of a grids.Case to the device."""
        for dct in (case.h, case.k, case.w):
            for name, arr in dct.items():
                self.set(name, arr)
        if case.open_bc is not None:
            self.set_open_bc(case.open_bc)
        for c in case.consts:
            slot = self.add_constituent(c["mo"], c["isadac"], c["isfct"],
                                        c["rkdye"], c["is_dye"], c.get("wset", 0.0),
                                        c.get("probdep", 0.0), c.get("wese", 0.0))
            self.set_constituent(slot, c["con"], c["con1"])
        if case.conc_bc is not None and case.consts:
            self.set_conc_bc(case.conc_bc)
        if case.sources is not None:
            self.set_sources(case.sources)
        if case.manning_n is not None:
            self.set_manning(case.manning_n)
        return self

    def caluvw(self, dt: float, n: int = 1, avcon1: float = 0.0, shear: int = 1,
               wind: int = 0, uv_after: int = 1):
        """CALUVW (EFDC/caluvw.f90) with the HDMT2T steps around it
        (EFDC/hdmt2t.f90:513-527,638-655)."""
        st = Step(dt, n, n)
        op = UvwOpts(avcon1, shear, wind, uv_after)
        self._check(self.lib.efdc_caluvw(self.h, C.byref(st), C.byref(op)))

    def step(self, dt: float, n: int, internal: bool = False, uvw: Optional[dict] = None) -> dict:
        """One hot-path step in the order of EFDC/hdmt2t.f90:494-562: the
        external solve, then the internal mode (CALUVW) when `internal`,
        else the depth-uniform stand-in for it, then the transport."""
        if self.has_manning:
            self.closure_pre()
        st = self.calpuv2c(dt, n)
        if internal:
            self.caluvw(dt, n, **(uvw or {}))
        else:
            self.closure_post()
        if self.nconst:
            self.calconc(dt, n)
        return st
