"""Synthetic EFDC+ cases in the var_global_mod array layout.

Every array is a numpy FP64 (or int32 for maps) array in the layout the
Fortran host hands over through the C-ABI:

* horizontal arrays have length LCM, element L (1-based) at index L-1;
* layer arrays are numpy (KCM, LCM) C-order == Fortran (LCM,KCM), so
  element (L,K) is ``a[K-1, L-1]``;
* W2 / DZIG are (KCM+1, LCM) == Fortran (LCM,0:KCM), element (L,K) is
  ``a[K, L-1]``.

Cell numbering follows CELLMAP (EFDC/cellmap.f90:12-26): wet cells are
numbered L=2..LA row by row (J outer, I inner); L=1 and LC=LA+1 are ghost
cells and land neighbours map to LC (EFDC/cellmap.f90:83-148).  As in
SCANINP, LCM = LC+1 and KCM = KC+1 (EFDC/scaninpmod.f90:71,80).

Geometry follows EFDC/aaefdc.f90:2172-2185 (uniform dx=dy here) and the
layer constants EFDC/aaefdc.f90:1212-1232 (DZIC, DZIG, CDZKK, CDZKMK).

The closure that stands in for the internal mode (which produces the
transport inputs UHDY2/VHDX2/W2/HPK upstream of CALCONC and lies outside
the accelerated path) is the depth-uniform one: UHDY2 = 0.5*(UHDYE+UHDY1E)
*DZC, W2 = 0, HPK = HP*DZC, and Manning bottom drag
TBX = g n^2 u|u| / HU^(1/3) on the external velocity.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Dict, List, Optional

import numpy as np

G = 9.81
H_KEYS = (
    "DXYP DXYIP DXYU DXYV DXIU DYIU DXIV DYIV HRUO HRVO SUB SVB BELV "
    "P P1 HP H1P H2P HU HV HUI HVI HPI H1U H1V H1UI H1VI "
    "UHDYE VHDXE UHDY1E VHDX1E UHE VHE "
    "TSX TSY TBX TBY FCAXE FCAYE FPGXE FPGYE FXE FYE QSUME RCX RCY "
    "FUHDYE FVHDXE FP FP1 CC CS CW CE CN CCC CCS CCW CCE CCN CCCI FPTMP SUBO SVBO"
).split()
K_KEYS = ("HPK H1PK HPKI UHDY2 VHDX2 CDZKK CDZKMK AB QSUM U2 V2 SUB3D SVB3D "
          "DZC").split()
W_KEYS = ("W2", "DZIG")
# internal mode (CALUVW): bound by add_internal_mode
IH_KEYS = "STBX STBY VU UV V1U U1V DYU DXV".split()
IK_KEYS = ("UHDYF VHDXF UHDY VHDX UHDYF1 VHDXF1 UHDY1 VHDX1 U V U1 V1 UHDYF2 VHDXF2 "
           "UHDYEK VHDXEK UHDY1EK VHDX1EK DU DV AVUI AVVI").split()
IW_KEYS = ("W", "W1")


@dataclasses.dataclass
class Grid:
    """Cell numbering and neighbour maps of a masked rectangular grid."""

    IC: int
    JC: int
    LA: int
    KC: int
    IL: np.ndarray
    JL: np.ndarray
    LIJ: np.ndarray
    maps: Dict[str, np.ndarray]

    @property
    def LC(self) -> int:
        return self.LA + 1

    @property
    def LCM(self) -> int:
        return self.LA + 2

    @property
    def KCM(self) -> int:
        return self.KC + 1

    @property
    def nwet(self) -> int:
        return self.LA - 1


def build_grid(wet: np.ndarray, kc: int) -> Grid:
    """Number wet cells like CELLMAP and build LWC/LEC/LSC/LNC/LNWC/LSEC.

    ``wet`` is a (JC, IC) boolean array (J rows, I columns, 0-based); its
    outer ring must be land.  Follows EFDC/cellmap.f90:12-26 (ordering) and
    EFDC/cellmap.f90:83-148 (maps; LC for land, ghosts map to LC / 1).
    """
    wet = np.asarray(wet, dtype=bool)
    JC, IC = wet.shape
    if wet[0].any() or wet[-1].any() or wet[:, 0].any() or wet[:, -1].any():
        raise ValueError("the outer ring of the mask must be land")
    jj, ii = np.nonzero(wet)  # row-major: J outer, I inner
    nwet = len(jj)
    LA = nwet + 1
    LC = LA + 1
    LCM = LC + 1
    LIJ = np.zeros((JC, IC), dtype=np.int32)
    LIJ[jj, ii] = np.arange(2, LA + 1, dtype=np.int32)
    IL = np.zeros(LCM, dtype=np.int32)
    JL = np.zeros(LCM, dtype=np.int32)
    IL[1:LA] = ii + 1  # 1-based I, J
    JL[1:LA] = jj + 1

    def nb(dj, di):
        v = LIJ[jj + dj, ii + di]
        return np.where(v > 0, v, LC).astype(np.int32)

    maps = {}
    for name, (dj, di) in dict(LNC=(1, 0), LSC=(-1, 0), LEC=(0, 1), LWC=(0, -1),
                               LNWC=(1, -1), LSEC=(-1, 1), LNEC=(1, 1),
                               LSWC=(-1, -1)).items():
        m = np.zeros(LCM, dtype=np.int32)
        m[0] = LC       # L=1 -> LC
        m[1:LA] = nb(dj, di)
        m[LA] = 1       # L=LC -> 1
        m[LC] = LC      # padding element of LCM
        maps[name] = m
    return Grid(IC=IC, JC=JC, LA=LA, KC=kc, IL=IL, JL=JL, LIJ=LIJ, maps=maps)


@dataclasses.dataclass
class Case:
    """One EFDC+ case: grid, var_global arrays, boundary tables, scalars."""

    name: str
    grid: Grid
    h: Dict[str, np.ndarray]
    k: Dict[str, np.ndarray]
    w: Dict[str, np.ndarray]
    consts: List[dict]                 # transported constituents
    open_bc: Optional[dict] = None     # pressure BCs
    conc_bc: Optional[dict] = None     # concentration BCs
    sources: Optional[dict] = None     # flow sources
    dt: float = 5.0
    manning_n: Optional[np.ndarray] = None  # per-cell Manning n (closure)
    meta: dict = dataclasses.field(default_factory=dict)

    def copy(self) -> "Case":
        def cp(d):
            return {a: b.copy() for a, b in d.items()}
        c = dataclasses.replace(self, h=cp(self.h), k=cp(self.k), w=cp(self.w))
        c.consts = [dict(x, con=x["con"].copy(), con1=x["con1"].copy())
                    for x in self.consts]
        c.meta = dict(self.meta)
        return c


# --------------------------------------------------------------- helpers

def _smooth_noise(rng, shape, sigma):
    from scipy.ndimage import gaussian_filter
    z = rng.uniform(0.0, 1.0, size=shape)
    return gaussian_filter(z, sigma=sigma, mode="reflect")


def _rescale(z, lo, hi):
    zmin, zmax = float(z.min()), float(z.max())
    if zmax - zmin < 1e-300:
        return np.full_like(z, 0.5 * (lo + hi))
    return lo + (hi - lo) * (z - zmin) / (zmax - zmin)


def _empty_arrays(g: Grid):
    LCM, KCM = g.LCM, g.KCM
    h = {key: np.zeros(LCM) for key in H_KEYS}
    k = {key: np.zeros((KCM, LCM)) for key in K_KEYS}
    w = {key: np.zeros((KCM + 1, LCM)) for key in W_KEYS}
    return h, k, w


def _face_masks(g: Grid):
    """SUB/SVB: 1 on faces between two wet cells (EFDC SUBO/SVBO)."""
    LA, LC = g.LA, g.LC
    sub = np.zeros(g.LCM)
    svb = np.zeros(g.LCM)
    L = np.arange(2, LA + 1)
    sub[L - 1] = (g.maps["LWC"][L - 1] != LC).astype(float)
    svb[L - 1] = (g.maps["LSC"][L - 1] != LC).astype(float)
    return sub, svb


def fill_geometry(g: Grid, h, k, w, dx: float, dzck: np.ndarray):
    """Metrics (EFDC/aaefdc.f90:2172-2185) and layer constants
    (EFDC/aaefdc.f90:1212-1232) for a uniform dx=dy grid with sigma layers
    DZC(L,K)=DZCK(K)/sum(DZCK)."""
    LA, KC = g.LA, g.KC
    L = np.arange(2, LA + 1) - 1
    sub, svb = _face_masks(g)
    h["SUB"][:] = sub
    h["SVB"][:] = svb
    h["SUBO"][:] = sub
    h["SVBO"][:] = svb
    for key, val in dict(DXYP=dx * dx, DXYIP=1.0 / (dx * dx), DXYU=dx * dx,
                         DXYV=dx * dx, DXIU=1.0 / dx, DYIU=1.0 / dx,
                         DXIV=1.0 / dx, DYIV=1.0 / dx).items():
        h[key][L] = val
    h["HRUO"][L] = sub[L] * dx * (1.0 / dx)
    h["HRVO"][L] = svb[L] * dx * (1.0 / dx)
    h["RCX"][:] = 1.0
    h["RCY"][:] = 1.0
    h["RCX"][0] = h["RCY"][0] = 0.0
    h["RCX"][g.LC - 1] = h["RCY"][g.LC - 1] = 0.0
    dzc = np.asarray(dzck, float) / float(np.sum(dzck))
    for K in range(1, KC + 1):
        k["DZC"][K - 1, L] = dzc[K - 1]
    dzic = 1.0 / dzc
    dzig = np.zeros(KC + 1)
    for K in range(1, KC):
        dzig[K] = 1.0 / (0.5 * (dzc[K - 1] + dzc[K]))
    for K in range(0, KC + 1):
        w["DZIG"][K, L] = dzig[K]
    for K in range(2, KC + 1):
        k["CDZKMK"][K - 1, L] = dzig[K - 1] * dzic[K - 1]
    for K in range(1, KC):
        k["CDZKK"][K - 1, L] = dzic[K - 1] * dzig[K]
    # 3D face masks (sigma grid: every layer open where the 2D face is)
    for K in range(1, KC + 1):
        k["SUB3D"][K - 1, :] = sub
        k["SVB3D"][K - 1, :] = svb


def update_depth_arrays(g: Grid, h):
    """HU/HV and inverses from HP (EFDC/calpuv2c.f90:901-913)."""
    LA = g.LA
    L = np.arange(2, LA + 1) - 1
    LW = g.maps["LWC"][L] - 1
    LS = g.maps["LSC"][L] - 1
    dxyp, hp = h["DXYP"], h["HP"]
    h["HU"][L] = (dxyp[L] * hp[L] + dxyp[LW] * hp[LW]) * (0.5 / h["DXYU"][L])
    h["HV"][L] = (dxyp[L] * hp[L] + dxyp[LS] * hp[LS]) * (0.5 / h["DXYV"][L])
    h["HPI"][L] = 1.0 / hp[L]
    h["HUI"][L] = 1.0 / h["HU"][L]
    h["HVI"][L] = 1.0 / h["HV"][L]


def layer_depths(g: Grid, h, k):
    """HPK = HP*DZC, HPKI = 1/HPK on wet cells (EFDC/caluvw.f90:1163-1172
    form for a sigma grid)."""
    L = np.arange(2, g.LA + 1) - 1
    for K in range(g.KC):
        k["HPK"][K, L] = h["HP"][L] * k["DZC"][K, L]
        k["HPKI"][K, L] = 1.0 / k["HPK"][K, L]


def largest_component(mask: np.ndarray) -> np.ndarray:
    from scipy.ndimage import label
    lab, n = label(mask)  # 4-connectivity
    if n == 0:
        return mask
    counts = np.bincount(lab.ravel())
    counts[0] = 0
    return lab == int(np.argmax(counts))


def _pad_ring(a: np.ndarray, value=False):
    return np.pad(a, 1, mode="constant", constant_values=value)


# ------------------------------------------------------------ constructors

def lake_mask(nx: int, ny: int, seed: int, wet_fraction=0.67, sigma=16.0):
    """Irregular lake: thresholded smoothed noise, largest 4-connected
    component, padded with a land ring.  Returns (JC, IC) bool."""
    rng = np.random.default_rng(seed)
    z = _smooth_noise(rng, (ny, nx), sigma)
    thr = np.quantile(z, 1.0 - wet_fraction)
    m = z > thr
    m[0, :] = m[-1, :] = m[:, 0] = m[:, -1] = False
    m = largest_component(m)
    return _pad_ring(m)


def external_case(wet: np.ndarray, *, seed: int, dx=100.0, dt=10.0, kc=1,
                  depth=(2.0, 40.0), depth_sigma=4.0, eta_sigma=0.05,
                  eta_smooth=2.0, name="external") -> Case:
    """Config-1 style external-mode case: irregular lake, random depth, a
    random free-surface perturbation, zero transports and forcings.  One
    CALPUV2C call assembles the SPD system exactly as
    EFDC/calpuv2c.f90:164-398 does and solves it."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    g = build_grid(wet, kc)
    h, k, w = _empty_arrays(g)
    fill_geometry(g, h, k, w, dx, np.ones(kc))
    JC, IC = wet.shape
    dep = _rescale(_smooth_noise(rng, (JC, IC), depth_sigma), *depth)
    eta = gaussian_filter(rng.normal(0.0, 1.0, size=(JC, IC)), eta_smooth)
    eta *= eta_sigma / max(float(eta[wet].std()), 1e-30)
    L = np.arange(2, g.LA + 1) - 1
    jj, ii = g.JL[L] - 1, g.IL[L] - 1
    h["BELV"][L] = -dep[jj, ii]
    h["HP"][L] = dep[jj, ii] + eta[jj, ii]
    h["H1P"][:] = h["HP"]
    h["P"][L] = G * (h["HP"][L] + h["BELV"][L])
    update_depth_arrays(g, h)
    layer_depths(g, h, k)
    k["H1PK"][:] = k["HPK"]
    return Case(name=name, grid=g, h=h, k=k, w=w, consts=[], dt=dt,
                meta=dict(seed=seed, dx=dx))


def config1(seed=1001, n=1024, wet_fraction=0.67) -> Case:
    """External-mode solve at scale: n x n lattice, irregular lake mask
    (~0.67 n^2 wet cells), depths 2-40 m, dt=10 s, eta' sigma=0.05 m."""
    wet = lake_mask(n, n, seed, wet_fraction=wet_fraction)
    return external_case(wet, seed=seed, dt=10.0, name=f"cfg1-{n}")


def _constituent(g: Grid, mo: int, con: np.ndarray, **kw) -> dict:
    d = dict(mo=mo, isadac=0, isfct=0, rkdye=0.0, is_dye=0, wset=0.0, probdep=0.0,
             wese=0.0, con=con, con1=con.copy())
    d.update(kw)
    return d


def tidal_basin(nx=128, ny=64, kc=5, *, seed=1000, dx=100.0, dt=5.0,
                amp=0.5, period=44712.0, manning=0.025, ab=1e-4,
                dye_load=1.0, depth=(5.0, 20.0), depth_sigma=8.0,
                ntscr=12, isadac=0, isfct=0, name=None, wet=None,
                dzck=None) -> Case:
    """Config-0 style case: rectangular basin nx*ny, west column is a
    clamped open boundary (ISPBW=0) with an M2-like tide, other sides
    closed.  Salinity 0->30 psu W->E, a dye point source (QSS=1e-3 m^3/s
    split over layers by DZC, CQS = dye_load/1e-3) at the centre cell.

    With ``wet`` (a padded (ny+2, nx+2) mask such as lake_mask returns) the
    basin takes that shape; the open boundary is then a west segment: in
    the middle third of the wet rows, each row's westernmost wet cell whose
    east neighbour is wet, and the dye source sits at the wet cell nearest
    the centre."""
    rng = np.random.default_rng(seed)
    if wet is None:
        wet = _pad_ring(np.ones((ny, nx), dtype=bool))
    else:
        wet = np.asarray(wet, dtype=bool)
        ny, nx = wet.shape[0] - 2, wet.shape[1] - 2
    g = build_grid(wet, kc)
    h, k, w = _empty_arrays(g)
    fill_geometry(g, h, k, w, dx, np.ones(kc) if dzck is None else np.asarray(dzck, float))
    dep_i = _rescale(_smooth_noise(rng, (ny, nx), depth_sigma), *depth)
    dep = _pad_ring(dep_i, 0.0)
    L = np.arange(2, g.LA + 1) - 1
    jj, ii = g.JL[L] - 1, g.IL[L] - 1
    h["BELV"][L] = -dep[jj, ii]
    h["HP"][L] = dep[jj, ii]
    h["H1P"][:] = h["HP"]
    h["H2P"][:] = h["HP"]
    update_depth_arrays(g, h)
    layer_depths(g, h, k)
    k["H1PK"][:] = k["HPK"]
    k["AB"][:kc - 1, L] = ab if kc > 1 else 0.0
    if wet[1:-1, 1:-1].all():
        # west open boundary: first wet column (I index 1 in 0-based padded grid)
        west = np.nonzero(ii == 1)[0]
    else:
        rows = np.unique(jj)
        lo, hi = rows[len(rows) // 3], rows[(2 * len(rows)) // 3]
        west = []
        for j in range(lo, hi + 1):
            cols = np.nonzero(wet[j])[0]
            if len(cols) and wet[j, cols[0] + 1]:
                west.append(int(g.LIJ[j, cols[0]]) - 2)   # position in L
        west = np.array(west, dtype=np.int64)
    lpbw = (L[west] + 1).astype(np.int32)
    nb = len(lpbw)
    open_bc = dict(side=[dict(L=lpbw, ispb=np.zeros(nb, np.int32),
                              pser=np.zeros(nb), pcb=np.full(nb, G * amp),
                              psb=np.zeros(nb))] + [None] * 3,
                   mtide=1, tcp=np.array([period]))
    # constituents: salinity gradient, dye zero
    sal = np.zeros((g.KCM, g.LCM))
    for K in range(kc):
        sal[K, L] = 30.0 * (ii - 1) / max(nx - 1, 1)
    dye = np.zeros((g.KCM, g.LCM))
    consts = [_constituent(g, 1, sal, isadac=isadac, isfct=isfct),
              _constituent(g, 3, dye, is_dye=1, isadac=isadac, isfct=isfct)]
    nc = len(consts)
    # concentration BC on the same west cells: inflow value = west salinity
    cb = np.zeros((nb, 2, nc))
    wtci = np.zeros((kc, 2))
    for K in range(kc):
        f = (K) / max(kc - 1, 1)
        wtci[K] = (1.0 - f, f)
    conc_bc = dict(side=[None, dict(L=lpbw, cb=cb, csert=np.zeros((nb, kc, nc)),
                                    ntscr=np.full(nb, ntscr, np.int32)),
                         None, None], wtci=wtci)
    # dye point source at the centre cell (I=nx/2, J=ny/2, 1-based interior)
    lsrc = int(g.LIJ[ny // 2, nx // 2])
    if lsrc == 0:
        d2 = (jj - ny // 2) ** 2 + (ii - nx // 2) ** 2
        lsrc = int(L[int(np.argmin(d2))]) + 1
    q = 1e-3
    qss = np.zeros((1, kc))
    qss[0] = q * k["DZC"][:kc, lsrc - 1]
    cqs = np.zeros((1, kc, nc))
    cqs[0, :, 1] = dye_load / q
    cqs[0, :, 0] = sal[:kc, lsrc - 1]
    sources = dict(LQS=np.array([lsrc], np.int32), qss=qss,
                   qsercell=np.zeros((1, kc)), cqs=cqs,
                   csert=np.zeros((1, kc, nc)))
    h["QSUME"][lsrc - 1] = q
    for K in range(kc):
        k["QSUM"][K, lsrc - 1] = qss[0, K]
    nman = np.full(g.LCM, manning)
    return Case(name=name or f"basin-{nx}x{ny}x{kc}", grid=g, h=h, k=k, w=w,
                consts=consts, open_bc=open_bc, conc_bc=conc_bc,
                sources=sources, dt=dt, manning_n=nman,
                meta=dict(seed=seed, dx=dx))


def drying_basin(nx=40, ny=12, kc=2, *, seed=1005, amp=0.8, period=3000.0,
                 deep=2.0, high=0.3, hdry=0.2, dt=4.0, shoal=1.0, name=None,
                 **kw) -> Case:
    """Tidal basin for the drying cases (ISDRY > 0): the bottom rises
    linearly from `deep` m below the datum at the open (west) boundary to
    `high` m above it at the east end, so the east cells start dry
    (HP = 0.5*hdry, their faces close on the first step), the ebb of the
    tide (amplitude `amp`) dries a band west of them and the flood wets it
    again.  A Gaussian shoal of height `shoal` m (radius 2 cells) at 45% of
    the length, mid-width, dries at low water with wet cells on all four
    sides.  Otherwise as tidal_basin."""
    case = tidal_basin(nx, ny, kc, seed=seed, amp=amp, period=period, dt=dt,
                       name=name or f"drying-{nx}x{ny}x{kc}", **kw)
    g, h, k = case.grid, case.h, case.k
    L = np.arange(2, g.LA + 1) - 1
    x = (g.IL[L] - 2) / max(nx - 1, 1)
    belv = -deep + (deep + high) * x
    r2 = (g.IL[L] - 2 - 0.45 * (nx - 1)) ** 2 + (g.JL[L] - 2 - 0.5 * (ny - 1)) ** 2
    belv = belv + shoal * np.exp(-r2 / 8.0)
    hp = np.maximum(-belv, 0.5 * hdry)
    h["BELV"][L] = belv
    for key in ("HP", "H1P", "H2P"):
        h[key][L] = hp
    h["P"][L] = G * (hp + belv)
    h["P1"][L] = h["P"][L]
    update_depth_arrays(g, h)
    layer_depths(g, h, k)
    k["H1PK"][:] = k["HPK"]
    return case


def ensemble_member(b: int, kc: int = 8, seed0: int = 1003) -> Case:
    """Basin b of the batched ensemble: (nx, ny) log-uniform in
    64x64 .. 512x256, an irregular lake mask as at scale, depths 2-30 m,
    Manning n in U(0.015, 0.04), tidal amplitude in U(0.1, 1.0) m on a west
    boundary segment, KC layers, salinity and a dye source; seed seed0+b."""
    seed = seed0 + b
    rng = np.random.default_rng(seed)
    nx = int(round(math.exp(rng.uniform(math.log(64), math.log(512)))))
    ny = int(round(math.exp(rng.uniform(math.log(64), math.log(256)))))
    manning = float(rng.uniform(0.015, 0.04))
    amp = float(rng.uniform(0.1, 1.0))
    wet = lake_mask(nx, ny, seed)
    return tidal_basin(nx, ny, kc, seed=seed, amp=amp, manning=manning,
                       depth=(2.0, 30.0), wet=wet, name=f"member{b}-{nx}x{ny}")


def config3(nbasins: int = 256, kc: int = 8):
    """The batched ensemble: ``nbasins`` members concatenated; returns
    (case, first_L, members)."""
    members = [ensemble_member(b, kc) for b in range(nbasins)]
    case, first = concat_basins(members, name=f"ensemble{nbasins}x{kc}")
    return case, first, members


def with_dye_sources(case: Case, cells, q: float = 1e-3, dye_slot: int = 1) -> Case:
    """Replace the source table of a basin with one inflow per cell in
    ``cells`` (L values): QSS = q split over the layers by DZC, dye CQS = 0
    (the loads are set per step), the other constituents enter at the local
    concentration.  QSUME/QSUM follow."""
    c = case.copy()
    g = c.grid
    kc = g.KC
    nc = len(c.consts)
    cells = np.asarray(cells, np.int32)
    ns = len(cells)
    qss = np.zeros((ns, kc))
    cqs = np.zeros((ns, kc, nc))
    for s_, L in enumerate(cells):
        qss[s_] = q * c.k["DZC"][:kc, L - 1]
        for j, x in enumerate(c.consts):
            if j != dye_slot:
                cqs[s_, :, j] = x["con"][:kc, L - 1]
    c.sources = dict(LQS=cells, qss=qss, qsercell=np.zeros((ns, kc)), cqs=cqs,
                     csert=np.zeros((ns, kc, nc)))
    c.h["QSUME"][:] = 0.0
    c.k["QSUM"][:] = 0.0
    for s_, L in enumerate(cells):
        c.h["QSUME"][L - 1] += q
        c.k["QSUM"][:kc, L - 1] += qss[s_]
    return c


def adjoint_case(n=512, kc=8, nsrc=16, nobs=50, seed=1004, reach=3):
    """Adjoint workload: an n x n basin built like the config-0 basin (west
    tidal boundary, salinity, Manning n = 0.025) with ``nsrc`` dye sources
    at random interior wet cells and ``nobs`` observed wet cells drawn at
    random among the cells within ``reach`` cells (in I and J) of a source,
    where the dye of a 200-step window arrives.  Returns (case, obs_L)."""
    rng = np.random.default_rng(seed)
    base = tidal_basin(n, n, kc, seed=seed, name=f"adjoint-{n}x{n}x{kc}")
    g = base.grid
    L = np.arange(2, g.LA + 1)
    ii, jj = g.IL[L - 1], g.JL[L - 1]
    inner = L[(ii > 1 + reach) & (ii < n - reach) & (jj > reach) & (jj <= n - reach)]
    src = np.sort(rng.choice(inner, size=nsrc, replace=False))
    near = set()
    for s_ in src:
        i0, j0 = g.IL[s_ - 1], g.JL[s_ - 1]
        for dj in range(-reach, reach + 1):
            for di in range(-reach, reach + 1):
                Lq = int(g.LIJ[j0 - 1 + dj, i0 - 1 + di])   # IL, JL are 1-based
                if Lq > 0:
                    near.add(Lq)
    obs = rng.choice(np.array(sorted(near)), size=min(nobs, len(near)), replace=False)
    return with_dye_sources(base, src), np.sort(obs).astype(np.int32)


def config0(**kw) -> Case:
    return tidal_basin(128, 64, 5, seed=kw.pop("seed", 1000), name="cfg0", **kw)


def transport_case(nx=2048, ny=1024, kc=20, *, seed=1002, dx=100.0, dt=5.0,
                   depth=10.0, umax=1.0, isadac=0, isfct=0,
                   name=None, with_w=True, sed_ws=1e-4, sed_probdep=1.0,
                   sed_wese=0.0) -> Case:
    """Config-2 style transport case: all cells wet, uniform depth, a
    divergence-free layer flow from a streamfunction plus a vertically
    alternating potential flow whose column sum vanishes (non-zero W2 from
    continuity, EFDC/caluvw.f90:668-680), log-uniform AB, 4 constituents:
    salinity, temperature, dye (one source) and a cohesive sediment class
    of 0.1 kg/m^3 settling at sed_ws m/s (deposition probability
    sed_probdep, resuspension flux sed_wese; sed_ws=0: no settling)."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    wet = _pad_ring(np.ones((ny, nx), dtype=bool))
    g = build_grid(wet, kc)
    h, k, w = _empty_arrays(g)
    fill_geometry(g, h, k, w, dx, np.ones(kc))
    L = np.arange(2, g.LA + 1) - 1
    jj, ii = g.JL[L] - 1, g.IL[L] - 1
    h["BELV"][L] = -depth
    h["HP"][L] = depth
    h["H1P"][:] = h["HP"]
    update_depth_arrays(g, h)
    layer_depths(g, h, k)
    k["H1PK"][:] = k["HPK"]
    # streamfunction on cell corners (ny+1, nx+1), zero on the boundary
    psi = gaussian_filter(rng.normal(size=(ny + 1, nx + 1)), 24.0)
    taper = np.outer(np.sin(np.linspace(0, np.pi, ny + 1)),
                     np.sin(np.linspace(0, np.pi, nx + 1)))
    psi *= taper
    # U face of cell (j,i): psi(j+1,i) - psi(j,i) (corner (i-1/2, j+-1/2))
    uf = psi[1:, :-1] - psi[:-1, :-1]      # (ny, nx) west faces
    vf = -(psi[:-1, 1:] - psi[:-1, :-1])   # (ny, nx) south faces
    uf[:, 0] = 0.0
    vf[0, :] = 0.0
    phi = gaussian_filter(rng.normal(size=(ny, nx)), 16.0)
    gu = np.zeros((ny, nx))
    gv = np.zeros((ny, nx))
    gu[:, 1:] = phi[:, 1:] - phi[:, :-1]
    gv[1:, :] = phi[1:, :] - phi[:-1, :]
    dzc = 1.0 / kc
    hu = depth
    smax = max(np.abs(uf).max(), np.abs(vf).max(), 1e-30)
    scale = umax * dx * hu * dzc / smax    # max layer velocity ~ umax
    gmax = max(np.abs(gu).max(), np.abs(gv).max(), 1e-30)
    gscale = 0.5 * umax * dx * hu * dzc / gmax if with_w else 0.0
    ipad = ii - 1
    jpad = jj - 1
    for K in range(kc):
        b = math.cos(math.pi * (K + 0.5) / kc)
        k["UHDY2"][K, L] = scale * uf[jpad, ipad] + gscale * b * gu[jpad, ipad]
        k["VHDX2"][K, L] = scale * vf[jpad, ipad] + gscale * b * gv[jpad, ipad]
    # W2 from layer continuity, W2(L,0)=0, W2(L,KC)=0
    LE = g.maps["LEC"][L] - 1
    LN = g.maps["LNC"][L] - 1
    for K in range(1, kc):
        uk, vk = k["UHDY2"][K - 1], k["VHDX2"][K - 1]
        div = (uk[LE] - uk[L] + vk[LN] - vk[L]) * h["DXYIP"][L]
        w["W2"][K, L] = w["W2"][K - 1, L] - div
    w["W2"][kc, L] = 0.0
    lo, hi = math.log(1e-5), math.log(1e-2)
    k["AB"][:kc - 1, L] = np.exp(rng.uniform(lo, hi, size=(kc - 1, len(L))))
    k["U2"][:kc, L] = k["UHDY2"][:kc, L] / (dx * hu * dzc)
    k["V2"][:kc, L] = k["VHDX2"][:kc, L] / (dx * hu * dzc)

    def field(lo_, hi_, sig):
        z = _rescale(_smooth_noise(rng, (ny, nx), sig), lo_, hi_)
        a = np.zeros((g.KCM, g.LCM))
        for K in range(kc):
            a[K, L] = z[jpad, ipad] * (1.0 + 0.01 * K)
        return a

    sal = field(0.0, 35.0, 32.0)
    tem = field(4.0, 25.0, 32.0)
    dye = np.zeros((g.KCM, g.LCM))
    sed = np.zeros((g.KCM, g.LCM))
    sed[:kc, L] = 0.1
    consts = [_constituent(g, 1, sal, isadac=isadac, isfct=isfct),
              _constituent(g, 2, tem, isadac=isadac, isfct=isfct),
              _constituent(g, 3, dye, is_dye=1, isadac=isadac, isfct=isfct),
              _constituent(g, 4, sed, isadac=isadac, isfct=isfct, wset=sed_ws,
                           probdep=sed_probdep, wese=sed_wese)]
    nc = len(consts)
    lsrc = int(g.LIJ[ny // 2 + 1, nx // 2 + 1])
    qss = np.zeros((1, kc))
    cqs = np.zeros((1, kc, nc))
    qsercell = np.zeros((1, kc))
    cqs[0, :, 2] = 1000.0
    qss[0] = 1e-3 / kc
    for c in (0, 1, 3):
        cqs[0, :, c] = consts[c]["con"][:kc, lsrc - 1]
    sources = dict(LQS=np.array([lsrc], np.int32), qss=qss, qsercell=qsercell,
                   cqs=cqs, csert=np.zeros((1, kc, nc)))
    for K in range(kc):
        k["QSUM"][K, lsrc - 1] = qss[0, K]
    h["QSUME"][lsrc - 1] = qss.sum()
    # sources add volume: keep HPK consistent only where no source acts;
    # the source is tiny (1e-3 m^3/s) and the tests account for it.
    return Case(name=name or f"transport-{nx}x{ny}x{kc}", grid=g, h=h, k=k,
                w=w, consts=consts, sources=sources, dt=dt,
                meta=dict(seed=seed, dx=dx))


def config2(**kw) -> Case:
    return transport_case(2048, 1024, 20, name="cfg2", **kw)


# ------------------------------------------------------------ ensembles

def concat_basins(cases: List[Case], name: Optional[str] = None):
    """Concatenate independent basins into one case.

    The basins' masks are stacked north of one another (each keeps its land
    ring, so no neighbour map joins two basins) and numbered like CELLMAP;
    basin b's cells then form the contiguous range first_L[b] ..
    first_L[b+1]-1 of the result, in the basin's own order.  Every array,
    boundary table, source and constituent is carried over at the shifted
    L.  The basins must share KC, dt, the tidal constituents, the
    concentration-BC weights and the constituent list.  Returns the case and
    first_L (the argument of efdc_set_segments).
    """
    c0 = cases[0]
    kc = c0.grid.KC
    for c in cases[1:]:
        if c.grid.KC != kc or c.dt != c0.dt:
            raise ValueError("basins must share KC and dt")
        if [(x["mo"], x["isadac"], x["isfct"], x["rkdye"], x["is_dye"]) for x in c.consts] != \
                [(x["mo"], x["isadac"], x["isfct"], x["rkdye"], x["is_dye"]) for x in c0.consts]:
            raise ValueError("basins must transport the same constituents")
    icmax = max(c.grid.IC for c in cases)
    masks = []
    for c in cases:
        m = c.grid.LIJ > 0
        masks.append(np.pad(m, ((0, 0), (0, icmax - m.shape[1]))))
    g = build_grid(np.concatenate(masks, axis=0), kc)
    offs = np.cumsum([0] + [c.grid.nwet for c in cases])
    first_L = (2 + offs[:-1]).astype(np.int32)
    LC = g.LC
    # the stacked numbering is the concatenation of the basins' numberings
    for b, c in enumerate(cases):
        n, o = c.grid.nwet, int(offs[b])
        for mname, m in c.grid.maps.items():
            old = m[1:1 + n]
            want = np.where(old == c.grid.LC, LC, old + o)
            if not np.array_equal(g.maps[mname][1 + o:1 + o + n], want):
                raise AssertionError(f"map {mname} of basin {b} does not concatenate")

    def cat(arrs, sel):
        a0 = sel(c0)
        out = np.zeros(a0.shape[:-1] + (g.LCM,), a0.dtype)
        out[..., 0] = a0[..., 0]
        out[..., LC - 1] = a0[..., c0.grid.LC - 1]
        out[..., g.LCM - 1] = a0[..., c0.grid.LCM - 1]
        for b, c in enumerate(cases):
            n, o = c.grid.nwet, int(offs[b])
            out[..., 1 + o:1 + o + n] = sel(c)[..., 1:1 + n]
        return out

    h = {k: cat(cases, lambda c, k=k: c.h[k]) for k in c0.h}
    kk = {k: cat(cases, lambda c, k=k: c.k[k]) for k in c0.k}
    w = {k: cat(cases, lambda c, k=k: c.w[k]) for k in c0.w}
    consts = []
    for j, x in enumerate(c0.consts):
        d = dict(x)
        d["con"] = cat(cases, lambda c, j=j: c.consts[j]["con"])
        d["con1"] = cat(cases, lambda c, j=j: c.consts[j]["con1"])
        consts.append(d)

    def shifted(b, L):
        return (np.asarray(L, np.int64) + int(offs[b])).astype(np.int32)

    open_bc = None
    if any(c.open_bc is not None for c in cases):
        if any(c.open_bc is None for c in cases):
            raise ValueError("either every basin or none has open boundaries")
        mt = c0.open_bc["mtide"]
        for c in cases:
            if c.open_bc["mtide"] != mt or not np.array_equal(c.open_bc["tcp"], c0.open_bc["tcp"]):
                raise ValueError("basins must share the tidal constituents")
        sides = []
        for s in range(4):
            parts = [(b, c.open_bc["side"][s]) for b, c in enumerate(cases)
                     if c.open_bc["side"][s] is not None]
            if not parts:
                sides.append(None)
                continue
            sides.append(dict(
                L=np.concatenate([shifted(b, d["L"]) for b, d in parts]),
                **{f: np.concatenate([np.asarray(d[f]) for _, d in parts], axis=0)
                   for f in ("ispb", "pser", "pcb", "psb")}))
        open_bc = dict(side=sides, mtide=mt, tcp=np.array(c0.open_bc["tcp"]))
    conc_bc = None
    if any(c.conc_bc is not None for c in cases):
        if any(c.conc_bc is None for c in cases):
            raise ValueError("either every basin or none has concentration BCs")
        for c in cases:
            if not np.array_equal(c.conc_bc["wtci"], c0.conc_bc["wtci"]):
                raise ValueError("basins must share the concentration-BC weights")
        sides = []
        for s in range(4):
            parts = [(b, c.conc_bc["side"][s]) for b, c in enumerate(cases)
                     if c.conc_bc["side"][s] is not None]
            if not parts:
                sides.append(None)
                continue
            sides.append(dict(
                L=np.concatenate([shifted(b, d["L"]) for b, d in parts]),
                **{f: np.concatenate([np.asarray(d[f]) for _, d in parts], axis=0)
                   for f in ("cb", "csert", "ntscr")}))
        conc_bc = dict(side=sides, wtci=np.array(c0.conc_bc["wtci"]))
    sources = None
    parts = [(b, c.sources) for b, c in enumerate(cases) if c.sources is not None]
    if parts:
        sources = dict(LQS=np.concatenate([shifted(b, d["LQS"]) for b, d in parts]),
                       **{f: np.concatenate([np.asarray(d[f]) for _, d in parts], axis=0)
                          for f in ("qss", "qsercell", "cqs", "csert")})
    manning = None
    if any(c.manning_n is not None for c in cases):
        if any(c.manning_n is None for c in cases):
            raise ValueError("either every basin or none has a Manning n")
        manning = cat(cases, lambda c: c.manning_n)
    case = Case(name=name or f"ensemble-{len(cases)}", grid=g, h=h, k=kk, w=w,
                consts=consts, open_bc=open_bc, conc_bc=conc_bc, sources=sources,
                dt=c0.dt, manning_n=manning,
                meta=dict(first_L=first_L, basins=[c.name for c in cases]))
    return case, first_L


def basin_slice(case: Case, first_L, b: int) -> slice:
    """0-based index range of basin b's cells in a concatenated array."""
    lo = int(first_L[b]) - 1
    hi = int(first_L[b + 1]) - 1 if b + 1 < len(first_L) else case.grid.LA
    return slice(lo, hi)


def add_internal_mode(case: Case, *, av=1e-3, drag=None) -> Case:
    """Bind the arrays of the internal mode (CALUVW, EFDC/caluvw.f90) to a
    case, in place: the layer flows of the depth-uniform state that the
    external flow UHDYE/VHDXE gives (UHDYF = UHDYE, UHDY = UHDYE*DZC,
    U = UHDYE*HUI*DYIU per layer, W = 0) at both levels, the inverse face
    viscosities AVUI = AVVI = 1/av (EFDC/ainit.f90:211; `av` a scalar or a
    (KC-1,) vector per interface), the face widths DYU = DXV = dx, and the
    bottom-drag coefficients STBX = STBY: `drag` if given, else
    G*n^2/HU^(1/3) from the case's Manning n at the initial depth (the
    mapping of a Manning coefficient onto EFDC's quadratic drag)."""
    g = case.grid
    LCM, KCM, KC = g.LCM, g.KCM, g.KC
    L = np.arange(2, g.LA + 1) - 1
    h, k, w = case.h, case.k, case.w
    for key in IH_KEYS:
        h[key] = np.zeros(LCM)
    for key in IK_KEYS:
        k[key] = np.zeros((KCM, LCM))
    for key in IW_KEYS:
        w[key] = np.zeros((KCM + 1, LCM))
    dx = float(case.meta.get("dx", 1.0 / h["DYIU"][L[0]]))
    h["DYU"][L] = dx
    h["DXV"][L] = dx
    if drag is not None:
        h["STBX"][L] = drag
        h["STBY"][L] = drag
    elif case.manning_n is not None:
        n2 = case.manning_n[L] ** 2
        h["STBX"][L] = G * n2 / np.cbrt(h["HU"][L])
        h["STBY"][L] = G * n2 / np.cbrt(h["HV"][L])
    avk = np.broadcast_to(np.asarray(av, float), (max(KC - 1, 1),))
    for K in range(KC - 1):
        k["AVUI"][K, L] = 1.0 / avk[K]
        k["AVVI"][K, L] = 1.0 / avk[K]
    for K in range(KC):
        for lvl in ("", "1"):
            k["UHDYF" + lvl][K, L] = h["UHDYE"][L]
            k["VHDXF" + lvl][K, L] = h["VHDXE"][L]
            k["UHDY" + lvl][K, L] = h["UHDYE"][L] * k["DZC"][K, L]
            k["VHDX" + lvl][K, L] = h["VHDXE"][L] * k["DZC"][K, L]
            k["U" + lvl][K, L] = h["UHDYE"][L] * h["HUI"][L] * h["DYIU"][L]
            k["V" + lvl][K, L] = h["VHDXE"][L] * h["HVI"][L] * h["DXIV"][L]
    case.meta["internal"] = True
    return case
