"""Cases shared by the reference-unit tests and the GPU tests: random
inputs that reach every branch of the units (open boundaries on all four
sides with ISPB 0/1/2, corner cells in two sides, recovery ramps, sources
with QSERCELL, anti-diffusion and FCT)."""
import numpy as np

G = 9.81


def boundary_cells(g, mapname, rng, nmax):
    L = np.arange(2, g.LA + 1)
    cand = L[g.maps[mapname][L - 1] == g.LC]
    rng.shuffle(cand)
    return np.sort(cand[:nmax]).astype(np.int32)


def random_case(seed, kc=4, nx=16, ny=12, isadac=1, isfct=1, nconst=4):
    from efdc_amd import grids
    rng = np.random.default_rng(seed)
    wet = grids.lake_mask(nx, ny, seed, wet_fraction=0.75, sigma=2.0)
    g = grids.build_grid(wet, kc)
    h, k, w = grids._empty_arrays(g)
    grids.fill_geometry(g, h, k, w, 100.0, rng.uniform(0.5, 1.5, kc))
    L = np.arange(2, g.LA + 1) - 1
    n = len(L)
    for K in range(kc):
        k["UHDY2"][K, L] = rng.normal(0, 40, n) * h["SUB"][L]
        k["VHDX2"][K, L] = rng.normal(0, 40, n) * h["SVB"][L]
        k["H1PK"][K, L] = rng.uniform(1, 3, n)
        k["HPK"][K, L] = rng.uniform(1, 3, n)
        k["HPKI"][K, L] = 1.0 / k["HPK"][K, L]
        k["U2"][K, L] = rng.normal(0, 0.3, n)
        k["V2"][K, L] = rng.normal(0, 0.3, n)
        k["QSUM"][K, L] = np.where(rng.uniform(size=n) < 0.1, rng.normal(0, 1e-3, n), 0.0)
    for K in range(1, kc):
        w["W2"][K, L] = rng.normal(0, 2e-3, n)
    h["HPI"][L] = rng.uniform(0.1, 0.5, n)
    consts = []
    mos = [1, 2, 3, 4][:nconst]
    for mo in mos:
        con = np.zeros((g.KCM, g.LCM))
        con[:kc, L] = rng.uniform(-1, 30, (kc, n))
        consts.append(dict(mo=mo, isadac=isadac, isfct=isfct, rkdye=0.0,
                           is_dye=int(mo == 3), con=con, con1=con.copy()))
    nc = len(consts)
    sides_c = [boundary_cells(g, m, rng, 5) for m in ("LSC", "LWC", "LEC", "LNC")]
    # make sure one corner cell sits in two concentration-BC sides
    corner = [l for l in range(2, g.LA + 1)
              if g.maps["LSC"][l - 1] == g.LC and g.maps["LWC"][l - 1] == g.LC]
    if corner:
        c0 = corner[0]
        sides_c[0] = np.unique(np.append(sides_c[0], c0)).astype(np.int32)
        sides_c[1] = np.unique(np.append(sides_c[1], c0)).astype(np.int32)
    wtci = rng.uniform(0, 1, (kc, 2))
    cside = []
    for cells in sides_c:
        m = len(cells)
        cside.append(dict(L=cells, cb=rng.uniform(0, 30, (m, 2, nc)),
                          csert=rng.uniform(0, 5, (m, kc, nc)),
                          ntscr=rng.integers(0, 4, m).astype(np.int32)))
    conc_bc = dict(side=cside, wtci=wtci)
    # pressure BC lists (W, E, S, N) on the same cells: they define LOBCS
    order = {0: 1, 1: 2, 2: 0, 3: 3}  # pbc side -> conc side
    pside = []
    for s in range(4):
        cells = sides_c[order[s]]
        m = len(cells)
        pside.append(dict(L=cells, ispb=rng.integers(0, 3, m).astype(np.int32),
                          pser=rng.normal(0, 5, m), pcb=rng.normal(0, 3, (m, 2)),
                          psb=rng.normal(0, 3, (m, 2))))
    open_bc = dict(side=pside, mtide=2, tcp=np.array([44712.0, 43200.0]))
    nq = 3
    lqs = rng.choice(L + 1, nq, replace=False).astype(np.int32)
    sources = dict(LQS=lqs, qss=rng.normal(0, 1e-2, (nq, kc)),
                   qsercell=rng.normal(0, 1e-2, (nq, kc)),
                   cqs=rng.uniform(0, 50, (nq, kc, nc)),
                   csert=rng.uniform(0, 50, (nq, kc, nc)))
    return grids.Case(name="rand", grid=g, h=h, k=k, w=w, consts=consts,
                      open_bc=open_bc, conc_bc=conc_bc, sources=sources, dt=5.0)


def edge_basin(seed=21, nx=20, ny=12, kc=3):
    """A tidal basin (grids.tidal_basin, Manning drag, dye source,
    anti-diffusion with FCT) open on all four sides: the clamped west
    column, the whole east column with ISPB drawn from 0/1/2, the west half
    of the south row with radiation rows (ISPB 1/2) and the east two
    thirds of the north row with 0/1/2; the SW and NE corner cells belong
    to two pressure sides and to two concentration sides.  Two tidal
    constituents with random amplitudes, random concentration-BC values
    and recovery ramps NTSCR in 0..3."""
    from efdc_amd import grids
    rng = np.random.default_rng(seed)
    case = grids.tidal_basin(nx, ny, kc, seed=seed, isadac=1, isfct=1)
    g = case.grid
    L = np.arange(2, g.LA + 1)
    ii, jj = g.IL[L - 1], g.JL[L - 1]
    west = L[ii == 2]; east = L[ii == nx + 1]; south = L[jj == 2]; north = L[jj == ny + 1]
    south = south[:nx // 2]            # the west half, with the SW corner
    north = north[nx // 3:]            # the east two thirds, with the NE corner
    mt = 2
    tcp = np.array([44712.0, 43200.0])
    def side(cells, ispb):
        m = len(cells)
        return dict(L=cells.astype(np.int32), ispb=ispb.astype(np.int32),
                    pser=rng.normal(0, 0.05, m) * G,
                    pcb=rng.uniform(0.1, 0.3, (m, mt)) * G, psb=rng.normal(0, 0.1, (m, mt)) * G)
    pside = [side(west, np.zeros(len(west))),
             side(east, rng.integers(0, 3, len(east))),
             side(south, rng.integers(1, 3, len(south))),
             side(north, rng.integers(0, 3, len(north)))]
    case.open_bc = dict(side=pside, mtide=mt, tcp=tcp)
    nc = len(case.consts)
    cs = []
    for cells in (south, west, east, north):     # S, W, E, N
        m = len(cells)
        cs.append(dict(L=cells.astype(np.int32), cb=rng.uniform(0, 30, (m, 2, nc)),
                       csert=rng.uniform(0, 5, (m, kc, nc)),
                       ntscr=rng.integers(0, 4, m).astype(np.int32)))
    case.conc_bc = dict(side=cs, wtci=rng.uniform(0, 1, (kc, 2)))
    return case
