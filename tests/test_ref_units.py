"""The C restatement (oracle/efdc_oracle.c) against the reference's own
units compiled from the EFDC+ sources (oracle/_ref): CALTRAN with CALFQC
(upwind advection, sources, open-boundary concentrations with recovery
ramps, anti-diffusion, FCT) and SETOPENBC + SETOPENBC2, on randomized
inputs.  Expected: bit-for-bit equality.  Skipped where the reference
build is absent (it is built only where /root/reference exists)."""
import numpy as np
import pytest

from edge_cases import random_case
from oracle import ref as R

pytestmark = pytest.mark.skipif(not R.available(), reason="oracle/_ref not built")


def load_reference(case, clo_init=None, nlo_init=None):
    g = case.grid
    kc, nc = g.KC, len(case.consts)
    nstv = max(c["mo"] for c in case.consts)
    cs = case.conc_bc["side"]
    ps = case.open_bc["side"]
    nb = [len(s["L"]) for s in cs]             # S, W, E, N
    npb = [len(s["L"]) for s in ps]            # W, E, S, N
    nq = len(case.sources["LQS"])
    ncser = sum(nb) + nq
    nbcsop = sum(npb)
    R.lib().ref_init(g.LA, kc, nstv, nb[0], nb[1], nb[2], nb[3],
                     npb[2], npb[0], npb[1], npb[3], nq, 2, ncser, nbcsop)
    for m in ("LWC", "LEC", "LSC", "LNC", "LNWC", "LSEC"):
        R.put(m, g.maps[m])
    for dct in (case.h, case.k, case.w):
        for name in ("DXYIP", "DXIU", "DYIV", "HPI", "DXYP", "H1P", "UHDY2",
                     "VHDX2", "H1PK", "HPKI", "HPK", "QSUM", "U2", "V2", "SUB3D",
                     "SVB3D", "DZC", "W2", "DZIG"):
            if name in dct:
                R.put(name, dct[name])
    csert = np.zeros((g.KCM, ncser + 1, nstv))
    sid = 0
    for s, tag in enumerate("SWEN"):
        side = cs[s]
        m = len(side["L"])
        mm = max(m, 1)
        cb = np.zeros((mm, 2, nstv))
        clo = np.zeros((mm, g.KCM, nstv))
        nlo = np.zeros((mm, g.KCM, nstv), np.int64)
        ncs = np.zeros((mm, nstv), np.int32)
        for ci, c in enumerate(case.consts):
            M = c["mo"] - 1
            cb[:m, :, M] = side["cb"][:, :, ci]
            if clo_init is not None:
                clo[:m, :kc, M] = clo_init[s][:, :, ci]
                nlo[:m, :kc, M] = nlo_init[s][:, :, ci]
            for LL in range(m):
                ncs[LL, M] = sid + LL + 1
                csert[:kc, sid + LL + 1, M] = side["csert"][LL, :, ci]
        sid += m
        R.put(f"LCB{tag}", side["L"] if m else np.zeros(1, np.int32))
        R.put(f"CB{tag}", R.fortran(cb))
        R.put(f"CLO{tag}", R.fortran(clo))
        R.put(f"NLO{tag}", R.fortran(nlo))
        R.put(f"NTSCR{tag}", side["ntscr"] if m else np.zeros(1, np.int32))
        R.put(f"NCSER{tag}", R.fortran(ncs))
    src = case.sources
    ncq = np.zeros((nq, nstv), np.int32)
    cqs = np.zeros((g.KCM, nq, nstv))
    for ci, c in enumerate(case.consts):
        M = c["mo"] - 1
        for s in range(nq):
            ncq[s, M] = sum(nb) + s + 1
            csert[:kc, sum(nb) + s + 1, M] = src["csert"][s, :, ci]
            cqs[:kc, s, M] = src["cqs"][s, :, ci]
    qss = np.zeros((g.KCM, nq))
    qss[:kc] = src["qss"].T
    qsc = np.zeros((g.KCM, nq + 1))
    qsc[:kc, 1:] = src["qsercell"].T
    R.put("CSERT", R.fortran(csert))
    wt = np.zeros((g.KCM, 2))
    wt[:kc] = case.conc_bc["wtci"]
    R.put("WTCI", R.fortran(wt))
    R.put("LQS", src["LQS"])
    R.put("NCSERQ", R.fortran(ncq))
    R.put("QSS", R.fortran(qss))
    R.put("CQS", R.fortran(cqs))
    R.put("QSERCELL", R.fortran(qsc))
    # LOBCS / LOBCS2 in the order S, W, E, N (EFDC/setbcs.f90:884-938)
    lob, lob2 = [], []
    nbmap = {0: "LEC", 1: "LWC", 2: "LNC", 3: "LSC"}
    for s in (2, 0, 1, 3):
        for LL, l in enumerate(ps[s]["L"]):
            lob.append(l)
            if ps[s]["ispb"][LL] <= 1:
                lob2.append(g.maps[nbmap[s]][l - 1])
    a = np.zeros(g.LCM, np.int32)
    a[:len(lob)] = lob
    R.put("LOBCS", a)
    a = np.zeros(g.LCM, np.int32)
    a[:len(lob2)] = lob2
    R.put("LOBCS2", a)
    lbcs = np.zeros(max(nq + nbcsop, 1), np.int32)
    lbcs[:nq] = src["LQS"]
    lbcs[nq:nq + len(lob)] = lob
    R.put("LBCS", lbcs)
    R.lib().ref_scalars(9.81, 0.2, 1.0, 0.0, len(lob2), nq + nbcsop)
    return nb


@pytest.mark.parametrize("isadac,isfct", [(0, 0), (1, 0), (1, 1)])
@pytest.mark.parametrize("seed", [1, 2])
def test_caltran_bitexact_vs_reference(oracle_mod, seed, isadac, isfct):
    case = random_case(seed, isadac=isadac, isfct=isfct)
    ref_case = case.copy()
    o = oracle_mod.Oracle(case)
    nb = load_reference(ref_case)
    g = case.grid
    kc = g.KC
    for c in ref_case.consts:
        R.lib().ref_flags(R.mvar_of(c["mo"]), c["isadac"], c["isfct"])
    for niter in (1, 2, 3, 7):
        for ci, c in enumerate(case.consts):
            o.caltran_only(ci, niter=niter)
            rc = ref_case.consts[ci]
            con = np.ascontiguousarray(rc["con"])
            con1 = np.ascontiguousarray(rc["con1"])
            R.lib().ref_caltran(R.mvar_of(rc["mo"]), rc["mo"], con.ctypes.data,
                                con1.ctypes.data, case.dt, niter)
            rc["con"][:] = con
            rc["con1"][:] = con1
            np.testing.assert_array_equal(c["con"][:kc], rc["con"][:kc],
                                          err_msg=f"CON mo={c['mo']} niter={niter}")
            np.testing.assert_array_equal(c["con1"][:kc], rc["con1"][:kc],
                                          err_msg=f"CON1 mo={c['mo']} niter={niter}")
    # open-boundary recovery memory CLO*/NLO*
    nc = len(case.consts)
    nstv = max(c["mo"] for c in case.consts)
    for s, tag in enumerate("SWEN"):
        m = nb[s]
        if m == 0:
            continue
        mm = max(m, 1)
        clo = R.get(f"CLO{tag}", mm * g.KCM * nstv).reshape((mm, g.KCM, nstv), order="F")
        nlo = R.get(f"NLO{tag}", mm * g.KCM * nstv, np.int64).reshape((mm, g.KCM, nstv), order="F")
        mem = o.cbc_mem[[i for i, sd in enumerate(case.conc_bc["side"]) if sd is not None].index(s)]
        for ci, c in enumerate(case.consts):
            M = c["mo"] - 1
            mine_c = mem["clo"].reshape(m, kc, nc)[:, :, ci]
            mine_n = mem["nlo"].reshape(m, kc, nc)[:, :, ci]
            np.testing.assert_array_equal(mine_c, clo[:m, :kc, M])
            np.testing.assert_array_equal(mine_n, nlo[:m, :kc, M])


@pytest.mark.parametrize("isadac,isfct", [(0, 0), (1, 1)])
@pytest.mark.parametrize("seed", [6, 7])
def test_caltran_with_dry_cells_bitexact_vs_reference(oracle_mod, seed, isadac, isfct):
    """CALTRAN over the wet lists (ISDRY > 0): about a fifth of the cells
    are left out by LMASKDRY (the reference's LKWET/LWET loops and the
    open-boundary CYCLEs, EFDC/caltran.f90:95-300,324-441,515), and, as
    after a drying step, the faces of a masked cell carry no flow."""
    case = random_case(seed, isadac=isadac, isfct=isfct)
    g = case.grid
    kc = g.KC
    rng = np.random.default_rng(50 + seed)
    mask = np.ones(g.LCM, np.int32)
    L = np.arange(2, g.LA + 1)
    dry = L[rng.uniform(size=len(L)) < 0.2]
    mask[dry - 1] = 0
    for l in dry:
        i, e, n_ = l - 1, g.maps["LEC"][l - 1] - 1, g.maps["LNC"][l - 1] - 1
        for key, faces in (("UHDY2", (i, e)), ("VHDX2", (i, n_)), ("U2", (i, e)),
                           ("V2", (i, n_))):
            case.k[key][:, list(faces)] = 0.0
        case.w["W2"][:, i] = 0.0
        case.h["SUB"][[i, e]] = 0.0
        case.h["SVB"][[i, n_]] = 0.0
    ref_case = case.copy()
    # the library and the oracle keep SUB3DO in SUB3D and scale it by SUB
    # when drying is on; the reference's CALPUV2C has already set
    # SUB3D = SUB*SUB3DO (EFDC/calpuv2c.f90:982-990)
    ref_case.k["SUB3D"][:] = ref_case.k["SUB3D"] * ref_case.h["SUB"][None, :]
    ref_case.k["SVB3D"][:] = ref_case.k["SVB3D"] * ref_case.h["SVB"][None, :]
    o = oracle_mod.Oracle(case, isdry=1)
    o.lmaskdry[:] = mask
    load_reference(ref_case)
    R.lib().ref_wetlist(mask.ctypes.data)
    for c in ref_case.consts:
        R.lib().ref_flags(R.mvar_of(c["mo"]), c["isadac"], c["isfct"])
    moved = False
    for niter in (1, 2, 3, 7):
        for ci, c in enumerate(case.consts):
            o.caltran_only(ci, niter=niter)
            rc = ref_case.consts[ci]
            con = np.ascontiguousarray(rc["con"])
            con1 = np.ascontiguousarray(rc["con1"])
            R.lib().ref_caltran(R.mvar_of(rc["mo"]), rc["mo"], con.ctypes.data,
                                con1.ctypes.data, case.dt, niter)
            rc["con"][:] = con
            rc["con1"][:] = con1
            np.testing.assert_array_equal(c["con"][:kc], rc["con"][:kc],
                                          err_msg=f"CON mo={c['mo']} niter={niter}")
            np.testing.assert_array_equal(c["con1"][:kc], rc["con1"][:kc],
                                          err_msg=f"CON1 mo={c['mo']} niter={niter}")
            # a masked column keeps its old level
            np.testing.assert_array_equal(c["con"][:kc, dry - 1], c["con1"][:kc, dry - 1])
            wet = L[mask[L - 1] == 1] - 1
            moved |= bool(np.any(c["con"][:kc, wet] != c["con1"][:kc, wet]))
    assert moved


@pytest.mark.parametrize("seed", [3, 4, 5])
def test_setopenbc_bitexact_vs_reference(oracle_mod, seed):
    case = random_case(seed, nconst=1)
    rng = np.random.default_rng(100 + seed)
    g = case.grid
    L = np.arange(2, g.LA + 1) - 1
    for name in ("CC", "CS", "CW", "CE", "CN", "FP", "FUHDYE", "FVHDXE"):
        case.h[name][L] = rng.normal(0, 100, len(L))
    for name in ("HU", "HV"):
        case.h[name][L] = rng.uniform(1, 20, len(L))
    case.h["BELV"][L] = rng.uniform(-5, 0.5, len(L))
    case.h["RCX"][L] = rng.uniform(0.5, 1, len(L))
    case.h["RCY"][L] = rng.uniform(0.5, 1, len(L))
    case.h["HRUO"][L] = rng.uniform(0.5, 1.5, len(L)) * case.h["SUB"][L]
    case.h["HRVO"][L] = rng.uniform(0.5, 1.5, len(L)) * case.h["SVB"][L]
    ref_case = case.copy()
    o = oracle_mod.Oracle(case, hdry=0.3, tcon_tbegin=1234.5)
    load_reference(ref_case)
    for name in ("CC", "CS", "CW", "CE", "CN", "FP", "FUHDYE", "FVHDXE", "HU",
                 "HV", "BELV", "RCX", "RCY", "HRUO", "HRVO", "DXIU", "DYIU",
                 "DXIV", "DYIV", "DXYP"):
        R.put(name, ref_case.h[name])
    ps = case.open_bc["side"]
    pser = np.zeros(sum(len(s["L"]) for s in ps) + 1)
    sid = 0
    for s, tag in enumerate("WESN"):
        side = ps[s]
        m = len(side["L"])
        R.put(f"LPB{tag}", side["L"] if m else np.zeros(1, np.int32))
        R.put(f"ISPB{tag}", side["ispb"] if m else np.zeros(1, np.int32))
        R.put(f"NPSER{tag}", np.arange(sid + 1, sid + m + 1, dtype=np.int32)
              if m else np.zeros(1, np.int32))
        pser[sid + 1:sid + m + 1] = side["pser"]
        pc = np.zeros((max(m, 1), 2))
        pc[:m] = side["pcb"]
        sb = np.zeros((max(m, 1), 2))
        sb[:m] = side["psb"]
        R.put(f"PCB{tag}", R.fortran(pc))
        R.put(f"PSB{tag}", R.fortran(sb))
        sid += m
    R.put("PSERT", pser)
    R.put("TCP", case.open_bc["tcp"])
    R.lib().ref_scalars(9.81, 0.3, 1.0, 1234.5, 0, 0)
    # some cells start inactivated (LOPENBCDRY)
    dry = np.zeros(g.LCM, np.int32)
    allb = np.concatenate([s["L"] for s in ps])
    dry[allb[rng.uniform(size=len(allb)) < 0.2] - 1] = 1
    o.lopenbcdry[:] = dry
    R.put("LOPENBCDRY", dry)
    for n in (1, 17, 4001):
        o.setopenbc_apply(n=n)
        R.lib().ref_setopenbc(case.dt, n)
        for name in ("CC", "CS", "CW", "CE", "CN", "FP"):
            got = R.get(name, g.LCM)
            np.testing.assert_array_equal(case.h[name], got, err_msg=f"{name} n={n}")
        np.testing.assert_array_equal(o.lopenbcdry, R.get("LOPENBCDRY", g.LCM, np.int32))
