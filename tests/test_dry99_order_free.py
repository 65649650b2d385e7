"""The ISDRY = 99 pass (EFDC/calpuv2c.f90:604-716) walks the cells in L
order and updates the shared face switches in place; the GPU kernel
k_dry99 evaluates every cell at once from the pass-start switches (a
neighbour with a smaller L that had a case has left 0 on the shared face,
the later writer of a face wins).  This CPU test holds that order-free form
to the sequential one on random small grids with random depths, depth
changes, switches, face flows and ISCDRY: same switches, same ISCDRY, same
"solve again" flag.  (Both are written out here in Python; the GPU kernel
and the C oracle are compared with each other in tests/test_gpu_drying.py.)
"""
import numpy as np

HDRY = 0.2
HDRY2 = 2 * HDRY
HDRY90 = 0.9 * HDRY
def grid(nx, ny):
    # row-major cells L=2..LA, ghost LC; index 0 = dummy (L=1)
    LA = nx*ny + 1; LC = LA + 1
    idx = lambda j, i: 1 + j*nx + i   # device index
    lw = np.full(LC, LC-1); le = np.full(LC, LC-1); ls = np.full(LC, LC-1); ln = np.full(LC, LC-1)
    for j in range(ny):
        for i in range(nx):
            k = idx(j, i)
            if i > 0: lw[k] = idx(j, i-1)
            if i < nx-1: le[k] = idx(j, i+1)
            if j > 0: ls[k] = idx(j-1, i)
            if j < ny-1: ln[k] = idx(j+1, i)
    return LA, LC, lw, le, ls, ln
def seq(nw, lw, le, ls, ln, HP, H1P, SUB, SVB, SUBO, SVBO, U, V, isc):
    SUB = SUB.copy(); SVB = SVB.copy(); isc = isc.copy(); ic = 0
    for i in range(1, nw+1):
        w, e, s, n = lw[i], le[i], ls[i], ln[i]
        if not HP[i] <= HDRY: continue
        SW, SE, SS, SN = SUB[i], SUB[e], SVB[i], SVB[n]
        d = HP[i]-H1P[i]
        if d > 0:
            SUB[i]=SUB[e]=SVB[i]=SVB[n]=0.
            if d > (HDRY-H1P[i])*0.01:
                if SUBO[i]>0.5 and U[i]>0 and HP[w]>HDRY2: SUB[i]=SUBO[i]
                if SUBO[e]>0.5 and U[e]<0 and HP[e]>HDRY2: SUB[e]=SUBO[e]
                if SVBO[i]>0.5 and V[i]>0 and HP[s]>HDRY2: SVB[i]=SVBO[i]
                if SVBO[n]>0.5 and V[n]<0 and HP[n]>HDRY2: SVB[n]=SVBO[n]
                isc[i] = 1 if SUB[i]+SUB[e]+SVB[i]+SVB[n] < 0.5 else 0
                if abs(SUB[i]-SW)>0.5 or abs(SUB[e]-SE)>0.5 or abs(SVB[i]-SS)>0.5 or abs(SVB[n]-SN)>0.5: ic=1
            elif isc[i]==0: isc[i]=1; ic=1
        elif HP[i] < HDRY90 or H1P[i] < HDRY:
            SUB[i]=SUB[e]=SVB[i]=SVB[n]=0.
            if isc[i]==0: isc[i]=1; ic=1
    return SUB, SVB, isc, ic
def case(j, nw, HP, H1P):
    if j < 1 or j > nw: return 0
    if not HP[j] <= HDRY: return 0
    d = HP[j]-H1P[j]
    if d > 0: return 1 if d > (HDRY-H1P[j])*0.01 else 2
    return 3 if (HP[j] < HDRY90 or H1P[j] < HDRY) else 0
def par(nw, lw, le, ls, ln, HP, H1P, SUB, SVB, SUBO, SVBO, U, V, isc):
    sn = SUB.copy(); vn = SVB.copy(); isc = isc.copy(); ic = 0
    for i in range(1, nw+1):
        w, e, s, n = lw[i], le[i], ls[i], ln[i]
        ci, cw, cs = case(i,nw,HP,H1P), case(w,nw,HP,H1P), case(s,nw,HP,H1P)
        ww=we=wsv=wn=0.
        if ci == 1:
            if SUBO[i]>0.5 and U[i]>0 and HP[w]>HDRY2: ww=SUBO[i]
            if SUBO[e]>0.5 and U[e]<0 and HP[e]>HDRY2: we=SUBO[e]
            if SVBO[i]>0.5 and V[i]>0 and HP[s]>HDRY2: wsv=SVBO[i]
            if SVBO[n]>0.5 and V[n]<0 and HP[n]>HDRY2: wn=SVBO[n]
        ow=os_=0.
        if cw==1 and SUBO[i]>0.5 and U[i]<0 and HP[i]>HDRY2: ow=SUBO[i]
        if cs==1 and SVBO[i]>0.5 and V[i]<0 and HP[i]>HDRY2: os_=SVBO[i]
        fw, fs = SUB[i], SVB[i]
        if ci and cw: fw = ww if i > w else ow
        elif ci: fw = ww
        elif cw: fw = ow
        if ci and cs: fs = wsv if i > s else os_
        elif ci: fs = wsv
        elif cs: fs = os_
        sn[i]=fw; vn[i]=fs
        if ci == 1:
            ce, cn = case(e,nw,HP,H1P), case(n,nw,HP,H1P)
            s0w = 0. if (cw and w < i) else SUB[i]; s0e = 0. if (ce and e < i) else SUB[e]
            s0s = 0. if (cs and s < i) else SVB[i]; s0n = 0. if (cn and n < i) else SVB[n]
            isc[i] = 1 if ww+we+wsv+wn < 0.5 else 0
            if abs(ww-s0w)>0.5 or abs(we-s0e)>0.5 or abs(wsv-s0s)>0.5 or abs(wn-s0n)>0.5: ic=1
        elif ci and isc[i]==0: isc[i]=1; ic=1
    return sn, vn, isc, ic
def test_order_free_pass_equals_the_sequential_pass():
    rng = np.random.default_rng(0)
    bad = 0
    for t in range(600):
        nx, ny = rng.integers(1, 7), rng.integers(1, 7)
        LA, LC, lw, le, ls, ln = grid(nx, ny); nw = LA-1
        HP = rng.choice([0.05, 0.15, 0.19, 0.21, 0.3, 0.5, 1.0], LC+1); HP[0]=0; HP[LC-1]=0
        H1P = HP + rng.choice([-0.1, -0.001, 0.0, 0.001, 0.1], LC+1)
        SUBO = np.zeros(LC+1); SVBO = np.zeros(LC+1)
        for i in range(1, nw+1):
            if lw[i] != LC-1: SUBO[i] = 1.
            if ls[i] != LC-1: SVBO[i] = 1.
        SUB = SUBO * rng.integers(0, 2, LC+1); SVB = SVBO * rng.integers(0, 2, LC+1)
        U = rng.normal(size=LC+1) * SUB; V = rng.normal(size=LC+1) * SVB
        isc = rng.integers(0, 2, LC+1)
        a = seq(nw, lw, le, ls, ln, HP, H1P, SUB, SVB, SUBO, SVBO, U, V, isc)
        b = par(nw, lw, le, ls, ln, HP, H1P, SUB, SVB, SUBO, SVBO, U, V, isc)
        r = slice(1, nw+1)
        if not (np.array_equal(a[0][r], b[0][r]) and np.array_equal(a[1][r], b[1][r]) and np.array_equal(a[2][r], b[2][r]) and a[3]==b[3]):
            bad += 1
        assert bad == 0, (t, nx, ny)
