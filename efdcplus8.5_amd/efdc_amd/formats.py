"""The on-disk formats of EFDC+ that the library writes: the argument
structures of efdc_restart_text / efdc_restart_write (RESTOUT,
EFDC/restartmod.f90:17-297) and efdc_ee_open / efdc_ee_append (HEADEROUT,
WSOUT, VELOUT, EFDC/efdcout.f90:73-117,371-497), and readers of the EE
streams in the order GetEFDC reads them (GETEE_WS / GETEE_VEL,
GetEFDC/src/geteeoutmod.f90:14-148).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

PD = C.POINTER(C.c_double)
PI = C.POINTER(C.c_int32)


class Restart(C.Structure):
    _fields_ = [("n", C.c_int64), ("timeday", C.c_double), ("isco", C.c_int32 * 9),
                ("istran", C.c_int32 * 9), ("isice", C.c_int32), ("slot", C.c_int32 * 3),
                ("nsed", C.c_int32), ("sed_slot", PI), ("kb", C.c_int32), ("kbm", C.c_int32),
                ("keep_nlo", C.c_int32),
                ("hwq", PD), ("h2wq", PD), ("qq", PD), ("qq1", PD), ("qql", PD), ("qql1", PD),
                ("dml", PD), ("ice_thick", PD), ("ice_temp", PD), ("temb", PD), ("sedb", PD),
                ("sedb1", PD), ("hbed", PD), ("hbed1", PD), ("vdrbed", PD), ("vdrbed1", PD)]


class EeGrid(C.Structure):
    _fields_ = [("ic", C.c_int32), ("jc", C.c_int32), ("igridv", C.c_int32), ("ksz", PI),
                ("rssbce", PD), ("rssbcw", PD), ("rssbcs", PD), ("rssbcn", PD)]


HOST_ARRAYS = ("hwq", "h2wq", "qq", "qq1", "qql", "qql1", "dml", "ice_thick", "ice_temp",
               "temb", "sedb", "sedb1", "hbed", "hbed1", "vdrbed", "vdrbed1")


def restart_opts(n: int, timeday: float, *, isco: Sequence[int] = (0,) * 9,
                 istran: Sequence[int] = (0,) * 9, isice: int = 0,
                 slots: Sequence[int] = (-1, -1, -1), sed_slots: Sequence[int] = (),
                 kb: int = 0, kbm: int = 0, keep_nlo: bool = False, **host):
    """The efdc_restart_t of a RESTOUT call; host arrays (HOST_ARRAYS) as
    float64 numpy arrays in the var_global layout: (LCM), (KC+1, LCM) for
    QQ..DML, (KBM, LCM) for the bed arrays, (NSCM, KBM, LCM) for SEDB.
    Returns (struct, arrays to keep alive)."""
    r = Restart()
    r.n, r.timeday, r.isice = n, timeday, isice
    r.isco[:] = list(isco)
    r.istran[:] = list(istran)
    r.slot[:] = list(slots)
    r.kb, r.kbm, r.keep_nlo = kb, kbm, int(bool(keep_nlo))
    keep = []
    ss = np.ascontiguousarray(np.asarray(sed_slots, dtype=np.int32))
    keep.append(ss)
    r.nsed = len(ss)
    r.sed_slot = ss.ctypes.data_as(PI) if len(ss) else None
    for name, arr in host.items():
        if name not in HOST_ARRAYS:
            raise KeyError(name)
        a = np.ascontiguousarray(arr, dtype=np.float64)
        keep.append(a)
        setattr(r, name, a.ctypes.data_as(PD))
    return r, keep


def ee_grid(ic: int, jc: int, rssbc: Sequence[np.ndarray], igridv: int = 0,
            ksz: Optional[np.ndarray] = None):
    """The efdc_ee_grid_t of HEADEROUT: RSSBCE, RSSBCW, RSSBCS, RSSBCN
    (LCM) and KSZ (LCM, None: 1 everywhere)."""
    g = EeGrid()
    g.ic, g.jc, g.igridv = ic, jc, igridv
    keep = []
    for name, arr in zip(("rssbce", "rssbcw", "rssbcs", "rssbcn"), rssbc):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        keep.append(a)
        setattr(g, name, a.ctypes.data_as(PD))
    if ksz is not None:
        k = np.ascontiguousarray(ksz, dtype=np.int32)
        keep.append(k)
        g.ksz = k.ctypes.data_as(PI)
    return g, keep


def read_ee_ws(path: str) -> dict:
    """EE_WS.OUT as GETEE_WS reads a VER >= 8400 file: the header, then per
    snapshot N, EETIME, DELT, LMINSTEP and HP(L = 2..LA)."""
    raw = open(path, "rb").read()
    ver, hsize, bsize, ic, jc, nact = np.frombuffer(raw, "<i4", 6)
    recs = []
    at = int(hsize)
    while at < len(raw):
        n = int(np.frombuffer(raw, "<i4", 1, at)[0])
        t = float(np.frombuffer(raw, "<f8", 1, at + 4)[0])
        delt = float(np.frombuffer(raw, "<f4", 1, at + 12)[0])
        lmin = int(np.frombuffer(raw, "<i4", 1, at + 16)[0])
        hp = np.frombuffer(raw, "<f4", int(nact), at + 20).copy()
        recs.append(dict(n=n, time=t, delt=delt, lminstep=lmin, hp=hp))
        at += int(bsize)
    return dict(ver=int(ver), hsize=int(hsize), bsize=int(bsize), ic=int(ic), jc=int(jc),
                nactive=int(nact), records=recs)


def read_ee_vel(path: str) -> dict:
    """EE_VEL.OUT: the header (with RSSBC*) and per snapshot N, EETIME,
    DELT and U, V, W over the active layers (CELL3D values each)."""
    raw = open(path, "rb").read()
    ver, hsize, bsize, igridv, cell3d, ic, jc, kc, nact = np.frombuffer(raw, "<i4", 9)
    lcm = (int(hsize) // 4 - 9) // 4
    rs = np.frombuffer(raw, "<f4", 4 * lcm, 36).reshape(4, lcm).copy()
    recs = []
    at = int(hsize)
    while at < len(raw):
        n = int(np.frombuffer(raw, "<i4", 1, at)[0])
        t = float(np.frombuffer(raw, "<f8", 1, at + 4)[0])
        delt = float(np.frombuffer(raw, "<f4", 1, at + 12)[0])
        uvw = np.frombuffer(raw, "<f4", 3 * int(cell3d), at + 16).reshape(3, int(cell3d)).copy()
        recs.append(dict(n=n, time=t, delt=delt, u=uvw[0], v=uvw[1], w=uvw[2]))
        at += int(bsize)
    return dict(ver=int(ver), hsize=int(hsize), bsize=int(bsize), igridv=int(igridv),
                cell3d=int(cell3d), ic=int(ic), jc=int(jc), kc=int(kc), nactive=int(nact),
                lcm=lcm, rssbc=rs, records=recs)
