"""ctypes front end of the C oracle (oracle/efdc_oracle.c).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, as the checker.  It works in place on the
numpy arrays of an efdc_amd.grids.Case.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "efdcplus8.5_amd"))
from efdc_amd.abi import ARRAY_IDS  # noqa: E402
from efdc_amd import formats as _fmt  # noqa: E402

LIB = os.path.join(HERE, "liboracle.so")
NARR = 256
MAXC = 16

PD = C.POINTER(C.c_double)
PI = C.POINTER(C.c_int32)
PL = C.POINTER(C.c_int64)


class PbcSide(C.Structure):
    _fields_ = [("n", C.c_int32), ("L", PI), ("ispb", PI), ("pser", PD),
                ("pcb", PD), ("psb", PD)]


class CbcSide(C.Structure):
    _fields_ = [("n", C.c_int32), ("L", PI), ("cb", PD), ("csert", PD),
                ("ntscr", PI), ("clo", PD), ("nlo", PL)]


class Const(C.Structure):
    _fields_ = [("mo", C.c_int32), ("isadac", C.c_int32), ("isfct", C.c_int32),
                ("is_dye", C.c_int32), ("rkdye", C.c_double), ("con", PD),
                ("con1", PD), ("wset", C.c_double), ("probdep", C.c_double),
                ("wese", C.c_double)]


class OrcCtx(C.Structure):
    _fields_ = [("LA", C.c_int32), ("LCM", C.c_int32), ("KC", C.c_int32),
                ("KCM", C.c_int32),
                ("LWC", PI), ("LEC", PI), ("LSC", PI), ("LNC", PI),
                ("LNWC", PI), ("LSEC", PI),
                ("a", PD * NARR),
                ("G", C.c_double), ("RITB1", C.c_double), ("SNLT", C.c_double),
                ("HDRY", C.c_double), ("TCON_TBEGIN", C.c_double),
                ("RSQM", C.c_double), ("CG_REL_TOL", C.c_double),
                ("ITERM", C.c_int32),
                ("nthreads", C.c_int32),
                ("pbc", PbcSide * 4), ("mtide", C.c_int32), ("tcp", PD),
                ("lopenbcdry", PI),
                ("cbc", CbcSide * 4), ("wtci", PD),
                ("nqsij", C.c_int32), ("lqs", PI), ("qss", PD),
                ("qsercell", PD), ("cqs", PD), ("csert_q", PD),
                ("nconst", C.c_int32), ("c", Const * MAXC),
                ("ISDRY", C.c_int32), ("iscdry", PI), ("lmaskdry", PI),
                ("sub1", PD), ("svb1", PD), ("ncordry", C.c_int32),
                ("NDRYSTP", C.c_int32), ("natdry", PI), ("qdwaste", PD), ("vdwaste", PD),
                ("isbal", C.c_int32), ("bal_volout", C.c_double),
                ("bal_massout", C.c_double * MAXC),
                ("IMPLICIT_DRAG", C.c_int32)]


class UvwOpts(C.Structure):
    _fields_ = [("AVCON1", C.c_double), ("shear", C.c_int32), ("wind", C.c_int32),
                ("uv_after", C.c_int32)]


class CgOut(C.Structure):
    _fields_ = [("iters", C.c_int32), ("rsq", C.c_double),
                ("ccmnm", C.c_double), ("status", C.c_int32)]


def build(force: bool = False, plant: int = 0, out: str = None) -> str:
    """Compile the oracle into ``out`` (default oracle/liboracle.so, the
    checker).  ``plant`` > 0 builds a variant with that planted mistake
    (the ORC_PLANT list in efdc_oracle.c), for the tests that show a check
    fails on it; such a variant never replaces the checker."""
    srcs = [os.path.join(HERE, f) for f in ("efdc_oracle.c", "efdc_oracle_io.c")]
    deps = srcs + [os.path.join(HERE, "efdc_oracle.h"),
                   os.path.join(ROOT, "include", "efdc_hip.h")]
    dst = out or LIB
    if plant and dst == LIB:
        raise ValueError("a planted oracle needs its own output path")
    if force or plant or not os.path.exists(dst) or \
            os.path.getmtime(dst) < max(os.path.getmtime(f) for f in deps):
        subprocess.check_call(
            ["gcc", "-O2", "-fopenmp", "-fPIC", "-shared",
             "-ffp-contract=off", "-Wall", f"-DORC_PLANT={int(plant)}", *srcs,
             "-o", dst, "-lm"])
    return dst


_libs = {}


def lib(path: str = None):
    """The oracle library at ``path`` (default: the checker, built on
    first use)."""
    path = path or LIB
    if path not in _libs:
        if path == LIB and not os.path.exists(LIB):
            build()
        _lib = C.CDLL(path)
        for fn, args in dict(
                orc_congrad=[C.POINTER(OrcCtx), C.POINTER(CgOut)],
                orc_calpuv2c=[C.POINTER(OrcCtx), C.c_double, C.c_int64,
                              C.POINTER(CgOut)],
                orc_assemble=[C.POINTER(OrcCtx), C.c_double, C.c_int64, PD],
                orc_calconc=[C.POINTER(OrcCtx), C.c_double, C.c_int64],
                orc_vertical_diffusion=[C.POINTER(OrcCtx), C.c_double],
                orc_settle=[C.POINTER(OrcCtx), C.c_double],
                orc_balance=[C.POINTER(OrcCtx), PD, PD],
                orc_closure_pre=[C.POINTER(OrcCtx), PD],
                orc_closure_post=[C.POINTER(OrcCtx)],
                orc_caluvw=[C.POINTER(OrcCtx), C.c_double, C.POINTER(UvwOpts)],
                orc_bal2t_totals=[C.POINTER(OrcCtx), PD, PD],
                orc_bal2t_begin=[C.POINTER(OrcCtx), PD, PD],
                orc_setopenbc_apply=[C.POINTER(OrcCtx), C.c_double, C.c_int64],
                orc_caltran_only=[C.POINTER(OrcCtx), C.c_int, C.c_double,
                                  C.c_int64],
                orc_restart_text=[C.POINTER(OrcCtx), C.POINTER(_fmt.Restart),
                                  C.POINTER(C.c_void_p), C.POINTER(C.c_int64)],
                orc_free=[C.c_void_p],
                orc_ee_open=[C.POINTER(OrcCtx), C.c_char_p, C.POINTER(_fmt.EeGrid)],
                orc_ee_append=[C.POINTER(OrcCtx), C.c_char_p, C.POINTER(_fmt.EeGrid),
                               C.c_int32, C.c_double, C.c_double, C.c_int32],
                orc_fmt_e17_8=[C.c_double, C.c_char_p]).items():
            f = getattr(_lib, fn)
            f.argtypes = args
            f.restype = None if fn in ("orc_settle", "orc_fmt_e17_8", "orc_free") else C.c_int
        _libs[path] = _lib
    return _libs[path]


def _pd(a):
    return a.ctypes.data_as(PD)


def _pi(a):
    return a.ctypes.data_as(PI)


class Oracle:
    """Binds a Case to the C oracle; arrays are shared, not copied."""

    def __init__(self, case, *, rsqm=1e-9, iterm=200, nthreads=1, g=9.81,
                 ritb1=1.0, snlt=0.0, hdry=0.2, tcon_tbegin=0.0, rel_tol=0.0, isdry=0,
                 ndrystp=0, implicit_drag=0, lib_path=None):
        self.case = case
        self.L = lib(lib_path)
        g_ = case.grid
        o = OrcCtx()
        self._keep = []
        o.LA, o.LCM, o.KC, o.KCM = g_.LA, g_.LCM, g_.KC, g_.KCM
        for m in ("LWC", "LEC", "LSC", "LNC", "LNWC", "LSEC"):
            arr = np.ascontiguousarray(g_.maps[m], dtype=np.int32)
            self._keep.append(arr)
            setattr(o, m, _pi(arr))
        for dct in (case.h, case.k, case.w):
            for name, arr in dct.items():
                assert arr.dtype == np.float64 and arr.flags.c_contiguous
                o.a[ARRAY_IDS[name]] = _pd(arr)
        o.G, o.RITB1, o.SNLT, o.HDRY = g, ritb1, snlt, hdry
        o.TCON_TBEGIN, o.RSQM, o.ITERM, o.nthreads = tcon_tbegin, rsqm, iterm, nthreads
        o.CG_REL_TOL = rel_tol
        o.ISDRY = isdry
        self.lopenbcdry = np.zeros(g_.LCM, np.int32)
        o.lopenbcdry = _pi(self.lopenbcdry)
        self.iscdry = np.zeros(g_.LCM, np.int32)
        self.lmaskdry = np.ones(g_.LCM, np.int32)
        self.sub1 = np.zeros(g_.LCM)
        self.svb1 = np.zeros(g_.LCM)
        o.iscdry, o.lmaskdry = _pi(self.iscdry), _pi(self.lmaskdry)
        o.sub1, o.svb1 = _pd(self.sub1), _pd(self.svb1)
        o.NDRYSTP = ndrystp
        o.IMPLICIT_DRAG = implicit_drag
        self.natdry = np.zeros(g_.LCM, np.int32)
        self.qdwaste = np.zeros(g_.LCM)
        self.vdwaste = np.zeros(g_.LCM)
        o.natdry, o.qdwaste, o.vdwaste = _pi(self.natdry), _pd(self.qdwaste), _pd(self.vdwaste)
        obc = case.open_bc
        if obc is not None:
            mt = int(obc["mtide"])
            o.mtide = mt
            tcp = np.ascontiguousarray(obc["tcp"], dtype=np.float64)
            self._keep.append(tcp)
            o.tcp = _pd(tcp)
            for s, side in enumerate(obc["side"]):
                if side is None:
                    continue
                arrs = dict(L=np.ascontiguousarray(side["L"], np.int32),
                            ispb=np.ascontiguousarray(side["ispb"], np.int32),
                            pser=np.ascontiguousarray(side["pser"], np.float64),
                            pcb=np.ascontiguousarray(side["pcb"], np.float64).reshape(-1),
                            psb=np.ascontiguousarray(side["psb"], np.float64).reshape(-1))
                self._keep.append(arrs)
                o.pbc[s].n = len(arrs["L"])
                o.pbc[s].L, o.pbc[s].ispb = _pi(arrs["L"]), _pi(arrs["ispb"])
                o.pbc[s].pser, o.pbc[s].pcb, o.pbc[s].psb = (
                    _pd(arrs["pser"]), _pd(arrs["pcb"]), _pd(arrs["psb"]))
        nc = len(case.consts)
        cbc = case.conc_bc
        self.cbc_mem = []
        if cbc is not None:
            wtci = np.ascontiguousarray(cbc["wtci"], np.float64).reshape(-1)
            self._keep.append(wtci)
            o.wtci = _pd(wtci)
            for s, side in enumerate(cbc["side"]):
                if side is None:
                    continue
                n = len(side["L"])
                arrs = dict(L=np.ascontiguousarray(side["L"], np.int32),
                            cb=np.ascontiguousarray(side["cb"], np.float64).reshape(-1),
                            csert=np.ascontiguousarray(side["csert"], np.float64).reshape(-1),
                            ntscr=np.ascontiguousarray(side["ntscr"], np.int32),
                            clo=np.zeros(n * g_.KC * max(nc, 1)),
                            nlo=np.zeros(n * g_.KC * max(nc, 1), np.int64))
                self._keep.append(arrs)
                self.cbc_mem.append(arrs)
                b = o.cbc[s]
                b.n = n
                b.L, b.cb, b.csert = _pi(arrs["L"]), _pd(arrs["cb"]), _pd(arrs["csert"])
                b.ntscr = _pi(arrs["ntscr"])
                b.clo = _pd(arrs["clo"])
                b.nlo = arrs["nlo"].ctypes.data_as(PL)
        src = case.sources
        if src is not None:
            arrs = dict(LQS=np.ascontiguousarray(src["LQS"], np.int32),
                        qss=np.ascontiguousarray(src["qss"], np.float64).reshape(-1),
                        qsercell=np.ascontiguousarray(src["qsercell"], np.float64).reshape(-1),
                        cqs=np.ascontiguousarray(src["cqs"], np.float64).reshape(-1),
                        csert=np.ascontiguousarray(src["csert"], np.float64).reshape(-1))
            self._keep.append(arrs)
            self.src_cqs = arrs["cqs"]   # shared: edit in place to change CQS
            o.nqsij = len(arrs["LQS"])
            o.lqs = _pi(arrs["LQS"])
            o.qss, o.qsercell = _pd(arrs["qss"]), _pd(arrs["qsercell"])
            o.cqs, o.csert_q = _pd(arrs["cqs"]), _pd(arrs["csert"])
        o.nconst = nc
        for i, c in enumerate(case.consts):
            o.c[i].mo, o.c[i].isadac, o.c[i].isfct = c["mo"], c["isadac"], c["isfct"]
            o.c[i].is_dye, o.c[i].rkdye = c["is_dye"], c["rkdye"]
            o.c[i].con, o.c[i].con1 = _pd(c["con"]), _pd(c["con1"])
            o.c[i].wset = c.get("wset", 0.0)
            o.c[i].probdep = c.get("probdep", 0.0)
            o.c[i].wese = c.get("wese", 0.0)
        self.o = o

    # ------------------------------------------------------------ calls
    def congrad(self):
        out = CgOut()
        self.L.orc_congrad(C.byref(self.o), C.byref(out))
        return dict(iters=out.iters, rsq=out.rsq, status=out.status)

    def assemble(self, n=1, dt=None):
        cm = C.c_double(0.0)
        self.L.orc_assemble(C.byref(self.o), dt or self.case.dt, n, C.byref(cm))
        return cm.value

    def calpuv2c(self, n=1, dt=None):
        out = CgOut()
        self.L.orc_calpuv2c(C.byref(self.o), dt or self.case.dt, n, C.byref(out))
        return dict(iters=out.iters, rsq=out.rsq, ccmnm=out.ccmnm,
                    status=out.status)

    def calconc(self, niter=1, dt=None):
        return self.L.orc_calconc(C.byref(self.o), dt or self.case.dt, niter)

    def vertical_diffusion(self, dt=None):
        return self.L.orc_vertical_diffusion(C.byref(self.o), dt or self.case.dt)

    def settle(self, dt=None):
        """Only the settling sweep of CALCONC (EFDC/calsed.f90:232-256,308-312)."""
        return self.L.orc_settle(C.byref(self.o), dt or self.case.dt)

    def balance(self):
        v = C.c_double(0.0)
        m = np.zeros(MAXC)
        self.L.orc_balance(C.byref(self.o), C.byref(v), _pd(m))
        return v.value, m[:len(self.case.consts)].copy()

    def bal2t_begin(self):
        """BAL2T1: the masked volume and mass totals at the start of a
        balance period; from now on every CALCONC adds its source
        (BAL2T3A) and open-boundary (BAL2T2) terms to the outflow totals."""
        v = C.c_double(0.0)
        m = np.zeros(MAXC)
        self.L.orc_bal2t_begin(C.byref(self.o), C.byref(v), _pd(m))
        self.o.isbal = 1
        self.bal_beg = (v.value, m[:len(self.case.consts)].copy())

    def bal2t_end(self):
        """BAL2T5: totals at the end of the period, the outflow totals and
        the errors END - (BEG - OUT) for the volume and every constituent."""
        v = C.c_double(0.0)
        m = np.zeros(MAXC)
        self.L.orc_bal2t_totals(C.byref(self.o), C.byref(v), _pd(m))
        nc = len(self.case.consts)
        vb, mb = self.bal_beg
        vo = self.o.bal_volout
        mo = np.array(self.o.bal_massout[:nc])
        self.o.isbal = 0
        return dict(volbeg=vb, volend=v.value, volout=vo, volerr=v.value - (vb - vo),
                    massbeg=mb, massend=m[:nc].copy(), massout=mo,
                    masserr=m[:nc] - (mb - mo))

    def restart_text(self, n, timeday, **kw) -> bytes:
        """RESTOUT's RESTART.OUT text (EFDC/restartmod.f90:17-272)."""
        r, keep = _fmt.restart_opts(n, timeday, **kw)
        out = C.c_void_p()
        nb = C.c_int64(0)
        st = self.L.orc_restart_text(C.byref(self.o), C.byref(r), C.byref(out), C.byref(nb))
        assert st == 0, st
        text = C.string_at(out, nb.value)
        self.L.orc_free(out)
        del keep
        return text

    def ee_open(self, dirname, ic, jc, rssbc, igridv=0, ksz=None):
        g, keep = _fmt.ee_grid(ic, jc, rssbc, igridv, ksz)
        self._ee = (g, keep)
        assert self.L.orc_ee_open(C.byref(self.o), dirname.encode(), C.byref(g)) == 0

    def ee_append(self, dirname, n, eetime, delt, lminstep=0):
        g, _ = self._ee
        assert self.L.orc_ee_append(C.byref(self.o), dirname.encode(), C.byref(g), n, eetime,
                                    delt, lminstep) == 0

    def closure_pre(self):
        nman = np.ascontiguousarray(self.case.manning_n, np.float64)
        self.L.orc_closure_pre(C.byref(self.o), _pd(nman))

    def setopenbc_apply(self, n=1, dt=None):
        self.L.orc_setopenbc_apply(C.byref(self.o), dt or self.case.dt, n)

    def caltran_only(self, c, niter=1, dt=None):
        self.L.orc_caltran_only(C.byref(self.o), c, dt or self.case.dt, niter)

    def closure_post(self):
        self.L.orc_closure_post(C.byref(self.o))

    def caluvw(self, dt=None, avcon1=0.0, shear=1, wind=0, uv_after=1):
        """CALUVW with the HDMT2T steps around it (EFDC/hdmt2t.f90:513-527,
        638-655); returns the status (4 = negative depth)."""
        op = UvwOpts(avcon1, shear, wind, uv_after)
        return self.L.orc_caluvw(C.byref(self.o), dt or self.case.dt, C.byref(op))

    def step(self, n, uvw=None):
        """One 2TL hot-path step of the harness sequence (HDMT2T order,
        EFDC/hdmt2t.f90:494-562): closure -> CALPUV2C -> the internal mode
        (CALUVW, when the case carries its arrays; else the depth-uniform
        closure) -> CALCONC."""
        if self.case.manning_n is not None:
            self.closure_pre()
        st = self.calpuv2c(n=n)
        if self.case.meta.get("internal"):
            st["uvw"] = self.caluvw(**(uvw or {}))
        else:
            self.closure_post()
        if self.case.consts:
            self.calconc(niter=n)
        return st
