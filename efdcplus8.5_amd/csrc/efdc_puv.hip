// efdc_puv.hip -- CALPUV2C (2TL, with the drying passes of ISDRY > 0) and CONGRAD on gfx950.
//
// One thread per wet cell, 256-thread workgroups, XCD-aware tile order.
// Every kernel here is a streaming pass over L-ordered (LCM) FP64 arrays;
// the roofline is HBM (or MALL/L2 when the working set is resident).
//
// CONGRAD (EFDC/congrad.f90:54-187) runs as two kernels per iteration:
//   k_cg_a: p = CCCI*r + beta*p (recomputed for the 4 neighbours) fused
//           with Ap = A p and the partial sums of p.Ap;
//   k_cg_b: x += alpha p, r -= alpha Ap, partial sums of r.z and r.r.
// The alpha/beta scalars never leave the device: the last workgroup of a
// kernel (atomic ticket) finishes the FP64 reduction in a fixed order and
// writes the scalars for the next kernel.  The next launches read a device
// `done` flag and return immediately once the stop test (sum r^2 <= RSQM)
// has been met, so the host only polls every few iterations.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <utility>

#include "efdc_internal.h"

namespace {

struct HPtr {
  double *a[EFDC_NUM_H];
};

HPtr hptr(efdc_ctx *c) {
  HPtr h;
  for (int j = 0; j < EFDC_NUM_H; ++j) h.a[j] = c->dH[j];
  return h;
}

#define HA(id) h.a[EFDC_##id]

// Cells per thread of the CG kernels: each workgroup owns a contiguous run
// of CG_CPT*256 cells, which halves the number of workgroups (and of
// reduction arrivals) per launch.
int cg_blocks(int nwet) { return (nwet + CG_CPT * EFDC_BLOCK - 1) / (CG_CPT * EFDC_BLOCK); }
// The mixed solve's kernels move 24-32 bytes per cell: 4 cells per thread
// keep more loads in flight and quarter the blocks (and reduction arrivals).
#define MX_CPT 4
// Their tiles start at i = 0 mod 256 (whole 256-byte segments per wave);
// the dummy cell i = 0 is skipped.
int mx_blocks(int nwet) { return (nwet + MX_CPT * EFDC_BLOCK) / (MX_CPT * EFDC_BLOCK); }

// Pointwise kernels: tile t covers i = 256t .. 256t+255, so every wave's
// accesses cover whole 256-byte segments; the dummy cell i = 0 (L=1) is
// reported as nwet+1, i.e. not wet.  Launch puv_blocks(nwet) workgroups.
__device__ __forceinline__ int cell_index(int nwet, int &tile) {
  tile = xcd_tile(blockIdx.x, gridDim.x);
  const int i = tile * EFDC_BLOCK + (int)threadIdx.x;
  return i == 0 ? nwet + 1 : i;  // wet iff <= nwet
}
int puv_blocks(int nwet) { return efdc_grid_blocks(nwet + 1); }

// Buffers of the rotated time levels: pair[q] = {X, X1} after the swap,
// tri = {HP, H1P, H2P} after the rotation.
struct Rotated {
  double *pair[6][2];
  double *tri[3];
  int lc0;
};

// Part t (0..13) of giving each rotated name back the ghost entries (L=1
// and L=LC) it had before the rotation -- the reference's shifts only copy
// the wet cells: pair q = t/2 (t < 12) or the triple (t = 12, 13), at ghost
// L=1 (t even) or L=LC (t odd).
__device__ __forceinline__ void rot_ghost(const Rotated &r, int t) {
  const int g = (t & 1) ? r.lc0 : 0, q = t >> 1;
  if (q < 6) {
    const double a = r.pair[q][0][g], b = r.pair[q][1][g];
    r.pair[q][0][g] = b;
    r.pair[q][1][g] = a;
  } else {
    const double hp = r.tri[1][g], h1p = r.tri[2][g], h2p = r.tri[0][g];
    r.tri[0][g] = hp;
    r.tri[1][g] = h1p;
    r.tri[2][g] = h2p;
  }
}

// Start of a CALPUV2C step in its own launch (steps with drying or adjoint
// recording; the others do this work in the explicit and assembly passes):
// reset the flags and restore the ghost entries, one thread per part so
// that the 14 load/store round trips run side by side.
__global__ void k_step_begin(CgState *st, Rotated r) {
  const int t = threadIdx.x;
  if (t == 0) {
    st->symok = 0;
    st->negdep = 0;
  }
  if (t < 14) rot_ghost(r, t);
}

// ---------------------------------------------------------------------
// Explicit stage and time-level shifts: EFDC/calpuv2c.f90:214-220,256-290
// ---------------------------------------------------------------------
// fold = 1: this pass also does the step start of k_step_begin for the
// names it does not read at a ghost (the flags, the HU, HV, HUI, HVI pairs
// and the HP triple); the assembly then does the UHDYE and VHDXE pairs, so
// here their ghost entries still sit where the rotation left them.
template <typename T>
__global__ __launch_bounds__(EFDC_BLOCK) void k_puv_explicit(
    HPtr h, const int *__restrict__ lw, const int *__restrict__ ls,
    const int *__restrict__ le, const int *__restrict__ ln, const T *hpk,
    T *h1pk, int64_t pitch, int KC, int nwet, double dt, double G,
    double RITB1, double SNLT, int dry, int drag, Rotated rot, int fold, CgState *st) {
  if (fold && blockIdx.x == 0) {
    const int t = threadIdx.x;
    if (t == 0) {
      st->symok = 0;
      st->negdep = 0;
    }
    if (t >= 4 && t < 14) rot_ghost(rot, t);
  }
  int tile;
  const int i = cell_index(nwet, tile);
  if (i > nwet) return;
  const double DELT = dt, DELTD2 = 0.5 * dt, DELTI = 1. / DELT;
  const int w = lw[i], s = ls[i], e = le[i], n = ln[i];
  // The shifts UHDY1E=UHDYE, VHDX1E=VHDXE, H1U=HU, H1V=HV, H1UI=HUI,
  // H1VI=HVI, H2P=H1P, H1P=HP (:256-265) are buffer rotations done by the
  // host before this launch, so the step-start values are read from the
  // "1" arrays here and in every kernel up to the one that rewrites the
  // current array.
  const double *P = HA(P);
  HA(FUHDYE)[i] =
      HA(UHDY1E)[i] - DELTD2 * HA(SUB)[i] * HA(HRUO)[i] * HA(H1U)[i] * (P[i] - P[w]) +
      HA(SUB)[i] * DELT * HA(DXIU)[i] *
          (HA(DXYU)[i] * (HA(TSX)[i] - RITB1 * HA(TBX)[i]) + HA(FCAXE)[i] +
           HA(FPGXE)[i] - SNLT * HA(FXE)[i]);
  HA(FVHDXE)[i] =
      HA(VHDX1E)[i] - DELTD2 * HA(SVB)[i] * HA(HRVO)[i] * HA(H1V)[i] * (P[i] - P[s]) +
      HA(SVB)[i] * DELT * HA(DYIV)[i] *
          (HA(DXYV)[i] * (HA(TSY)[i] - RITB1 * HA(TBY)[i]) - HA(FCAYE)[i] +
           HA(FPGYE)[i] - SNLT * HA(FYE)[i]);
  if (drag) {  // implicit bottom / vegetation drag, :236-241
    HA(FUHDYE)[i] = HA(FUHDYE)[i] * HA(RCX)[i];
    HA(FVHDXE)[i] = HA(FVHDXE)[i] * HA(RCY)[i];
  }
  if (dry) {  // faces reopened after the explicit terms, :246-253
    HA(SUB)[i] = HA(SUBO)[i];
    HA(SVB)[i] = HA(SVBO)[i];
  }
  HA(P1)[i] = P[i];
  for (int k = 0; k < KC; ++k) h1pk[i + k * pitch] = hpk[i + k * pitch];
  const double C1 = 0.5 * G;
  // FP1 reads UHDYE(LE), VHDXE(LN) (:284-289): the step-start values, in the
  // "1" arrays at wet cells; a land neighbour's entry is UHDYE's own ghost
  // (the shift copies wet cells only), which the current array holds once
  // the ghosts are restored -- and the "1" array while they are not (fold)
  const bool ge = !fold && (e < 1 || e > nwet), gn = !fold && (n < 1 || n > nwet);
  const double ue = ge ? HA(UHDYE)[e] : HA(UHDY1E)[e];
  const double vn = gn ? HA(VHDXE)[n] : HA(VHDX1E)[n];
  HA(FP1)[i] = DELTI * HA(DXYP)[i] * P[i] -
               C1 * (ue - HA(UHDY1E)[i] + vn - HA(VHDX1E)[i]);
}

// ---------------------------------------------------------------------
// Assembly: EFDC/calpuv2c.f90:315-342, with the min-diagonal reduction of
// :360-376 over the cells that are not pressure open boundaries (those
// get their final CC in k_puv_bc_side; k_puv_bc_min completes the minimum).
// ---------------------------------------------------------------------
// fold = 1: the UHDYE and VHDXE pairs' ghost entries are restored here (the
// explicit pass did the other names; this pass reads neither pair).
__global__ __launch_bounds__(EFDC_BLOCK) void k_puv_assemble(
    HPtr h, const int *__restrict__ le, const int *__restrict__ ln,
    const unsigned char *__restrict__ pbcflag, int nwet, double dt, double G,
    double *part, CgState *st, unsigned *st_tk, Rotated rot, int fold) {
  if (fold && blockIdx.x == 0 && threadIdx.x < 4) rot_ghost(rot, threadIdx.x);
  int tile;
  const int i = cell_index(nwet, tile);
  double ccv = 1.E+18;
  if (i <= nwet) {
    const double DELTD2 = 0.5 * dt, DELTI = 1. / dt;
    const int e = le[i], n = ln[i];
    const double *SUB = HA(SUB), *SVB = HA(SVB);
    const double C1 = 0.5 * G;
    HA(FP)[i] = HA(FP1)[i] - C1 * (SUB[e] * HA(FUHDYE)[e] - SUB[i] * HA(FUHDYE)[i] +
                                   SVB[n] * HA(FVHDXE)[n] - SVB[i] * HA(FVHDXE)[i] -
                                   2.0 * HA(QSUME)[i]);
    const double C2 = -0.5 * DELTD2 * G;
    // HU(LE), HV(LN) (:329-342): the step-start values in the "1" arrays at
    // wet cells, the current arrays' own ghost entries at a land neighbour
    const bool ge = e < 1 || e > nwet, gn = n < 1 || n > nwet;
    const double hue = ge ? HA(HU)[e] : HA(H1U)[e];
    const double hvn = gn ? HA(HV)[n] : HA(H1V)[n];
    const double cs = C2 * SVB[i] * HA(HRVO)[i] * HA(RCY)[i] * HA(H1V)[i];
    const double cw = C2 * SUB[i] * HA(HRUO)[i] * HA(RCX)[i] * HA(H1U)[i];
    const double ce = C2 * SUB[e] * HA(HRUO)[e] * HA(RCX)[e] * hue;
    const double cn = C2 * SVB[n] * HA(HRVO)[n] * HA(RCY)[n] * hvn;
    HA(CS)[i] = cs;
    HA(CW)[i] = cw;
    HA(CE)[i] = ce;
    HA(CN)[i] = cn;
    const double cc = DELTI * HA(DXYP)[i] - cs - cw - ce - cn;
    HA(CC)[i] = cc;
    if (!pbcflag[i]) ccv = cc;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) st->symok = 0;  // reset flags
  double m;
  if (grid_min(ccv, part, st_tk + TS_PUV_MIN * TICKET_SITE, gridDim.x, tile, m) &&
      threadIdx.x == 0)
    st->ccmnm = m;
}

// ---------------------------------------------------------------------
// SETOPENBC (EFDC/setopenbc.f90:17-206), the open-boundary part of the
// min-diagonal reduction, and SETOPENBC2 (EFDC/setopenbc2.f90:17-45).
// The sides are applied in the reference order W, E, S, N, one launch per
// side (a side whose cells share a neighbour runs on one thread, in the
// reference's cell order, because of the FP(neighbour) += updates); then
// one launch over all boundary cells takes the minimum of their CC and
// zeroes the couplings into the clamped cells.
// ---------------------------------------------------------------------
struct PbcArgs {
  PbcDev side[4];
  const double *tide;  // cos(M) at [M], sin(M) at [mtide+M]
  int mtide;
  int *lopenbcdry;
};

__global__ __launch_bounds__(EFDC_BLOCK) void k_puv_bc_side(
    HPtr h, PbcArgs b, PbcDev s, int side, const int *__restrict__ lw, const int *__restrict__ le,
    const int *__restrict__ ls, const int *__restrict__ ln, double dt,
    double G, double HDRY) {
  const double D_LTD2 = 0.5 * dt, DELTI = 1. / dt, DE_TI = DELTI;
  const double HDRY2 = 0.2 * HDRY;
  double *CC = HA(CC), *CS = HA(CS), *CW = HA(CW), *CE = HA(CE), *CN = HA(CN),
         *FP = HA(FP), *DXYP = HA(DXYP), *BELV = HA(BELV), *HRUO = HA(HRUO),
         *HRVO = HA(HRVO), *RCX = HA(RCX), *RCY = HA(RCY), *HUT = HA(H1U),
         *HVT = HA(H1V), *DXIU = HA(DXIU), *DYIU = HA(DYIU), *DXIV = HA(DXIV),
         *DYIV = HA(DYIV), *FUHDYE = HA(FUHDYE), *FVHDXE = HA(FVHDXE);
  {
    int LL0 = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
    if (s.serial) {
      if (LL0 != 0) return;
      stride = 1;
    }
    for (int LL = LL0; LL < s.n; LL += stride) {
      const int i = s.L[LL];
      CC[i] = DELTI * DXYP[i];
      CS[i] = 0.;
      CW[i] = 0.;
      CE[i] = 0.;
      CN[i] = 0.;
      if (b.lopenbcdry[i]) continue;
      double fp = s.pser[LL];
      for (int M = 0; M < b.mtide; ++M)
        fp = fp + s.pcb[LL * b.mtide + M] * b.tide[M] +
             s.psb[LL * b.mtide + M] * b.tide[b.mtide + M];
      FP[i] = fp;
      const bool rad = (s.ispb[LL] == 1 || s.ispb[LL] == 2);
      if (side == 0) {
        const int e = le[i];
        if (rad) {
          const double CET = 0.5 * D_LTD2 * G * HRUO[e] * RCX[e] * HUT[e];
          const double TMP = D_LTD2 * sqrt(G * HUT[e]) * DXIU[e];
          CC[i] = CET * (1. + TMP) / TMP;
          CE[i] = -CET;
          FP[i] = CET * (2. * FP[i] - sqrt(G * HUT[e]) * FUHDYE[e] * DYIU[e] / HUT[e]) / TMP;
        } else {
          double CET;
          const double FP1G = FP[i] / G - HDRY2;
          if (FP1G < BELV[i] || FP1G < BELV[e]) {
            FP[i] = (BELV[i] + HDRY2) * G;
            CET = 0.;
            b.lopenbcdry[i] = 1;
          } else {
            CET = 0.5 * D_LTD2 * G * HRUO[e] * RCX[e] * HUT[e];
          }
          FP[e] = FP[e] + CET * FP[i];
          FP[i] = DE_TI * DXYP[i] * FP[i];
        }
      } else if (side == 1) {
        const int w = lw[i];
        if (rad) {
          const double CWT = 0.5 * D_LTD2 * G * HRUO[i] * RCX[i] * HUT[i];
          const double TMP = D_LTD2 * sqrt(G * HUT[i]) * DXIU[i];
          CC[i] = CWT * (1. + TMP) / TMP;
          CW[i] = -CWT;
          FP[i] = CWT * (2. * FP[i] + sqrt(G * HUT[i]) * FUHDYE[i] * DYIU[i] / HUT[i]) / TMP;
        } else {
          double CWT;
          const double FP1G = FP[i] / G - HDRY2;
          if (FP1G < BELV[i] || FP1G < BELV[w]) {
            FP[i] = (BELV[i] + HDRY2) * G;
            CWT = 0.;
            b.lopenbcdry[i] = 1;
          } else {
            CWT = 0.5 * D_LTD2 * G * HRUO[i] * RCX[i] * HUT[i];
          }
          FP[w] = FP[w] + CWT * FP[i];
          FP[i] = DE_TI * DXYP[i] * FP[i];
        }
      } else if (side == 2) {
        const int n = ln[i];
        if (rad) {
          const double CNT = 0.5 * D_LTD2 * G * HRVO[n] * RCY[n] * HVT[n];
          const double TMP = D_LTD2 * sqrt(G * HVT[n]) * DYIV[n];
          CC[i] = CNT * (1. + TMP) / TMP;
          CN[i] = -CNT;
          FP[i] = CNT * (2. * FP[i] - sqrt(G * HVT[n]) * FVHDXE[n] * DXIV[n] / HVT[n]) / TMP;
        } else {
          double CNT;
          const double FP1G = FP[i] / G - HDRY2;
          if (FP1G < BELV[i] || FP1G < BELV[n]) {
            FP[i] = (BELV[i] + HDRY2) * G;
            CNT = 0.;
            b.lopenbcdry[i] = 1;
          } else {
            CNT = 0.5 * D_LTD2 * G * HRVO[n] * RCY[n] * HVT[n];
          }
          FP[n] = FP[n] + CNT * FP[i];
          FP[i] = DE_TI * DXYP[i] * FP[i];
        }
      } else {
        const int s2 = ls[i];
        if (rad) {
          const double CST = 0.5 * D_LTD2 * G * HRVO[i] * RCY[i] * HVT[i];
          const double TMP = D_LTD2 * sqrt(G * HVT[i]) * DYIV[i];
          CC[i] = CST * (1. + TMP) / TMP;
          CS[i] = -CST;
          FP[i] = CST * (2. * FP[i] + sqrt(G * HVT[i]) * FVHDXE[i] * DXIV[i] / HVT[i]) / TMP;
        } else {
          double CST;
          const double FP1G = FP[i] / G - HDRY2;
          if (FP1G < BELV[i] || FP1G < BELV[s2]) {
            FP[i] = (BELV[i] + HDRY2) * G;
            CST = 0.;
            b.lopenbcdry[i] = 1;
          } else {
            CST = 0.5 * D_LTD2 * G * HRVO[i] * RCY[i] * HVT[i];
          }
          FP[s2] = FP[s2] + CST * FP[i];
          FP[i] = DE_TI * DXYP[i] * FP[i];
        }
      }
    }
  }
}

__global__ __launch_bounds__(EFDC_BLOCK) void k_puv_bc_min(
    HPtr h, PbcArgs b, int ntot, const int *__restrict__ lw, const int *__restrict__ le,
    const int *__restrict__ ls, const int *__restrict__ ln, double *part, CgState *st,
    unsigned *st_tk) {
  double *CC = HA(CC), *CS = HA(CS), *CW = HA(CW), *CE = HA(CE), *CN = HA(CN);
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  double m = 1.E+18;
  if (t < ntot) {
    // side of entry t (the sides' cells follow one another)
    const int n0 = b.side[0].n, n01 = n0 + b.side[1].n, n012 = n01 + b.side[2].n;
    int side, LL;
    const int *L, *ispb;
    if (t < n0) {
      side = 0; LL = t; L = b.side[0].L; ispb = b.side[0].ispb;
    } else if (t < n01) {
      side = 1; LL = t - n0; L = b.side[1].L; ispb = b.side[1].ispb;
    } else if (t < n012) {
      side = 2; LL = t - n01; L = b.side[2].L; ispb = b.side[2].ispb;
    } else {
      side = 3; LL = t - n012; L = b.side[3].L; ispb = b.side[3].ispb;
    }
    const int i = L[LL];
    // minimum of CC over the open-boundary cells (:365-375)
    m = CC[i];
    // SETOPENBC2: no coupling from the interior into a clamped cell
    if (ispb[LL] == 0) {
      if (side == 0) CW[le[i]] = 0.;
      else if (side == 1) CE[lw[i]] = 0.;
      else if (side == 2) CS[ln[i]] = 0.;
      else CN[ls[i]] = 0.;
    }
  }
  double mm;
  if (grid_min(m, part, st_tk + TS_PUV_MIN * TICKET_SITE, gridDim.x, blockIdx.x, mm) &&
      threadIdx.x == 0)
    st->ccmnm = fmin(st->ccmnm, mm);
}

// ---------------------------------------------------------------------
// Scaling by the minimum diagonal: EFDC/calpuv2c.f90:384-398, plus the
// check that the system is symmetric and E/W neighbours are contiguous,
// which selects the compact CG operator (CCN(L)=CCS(LN), CCE(L)=CCW(L+1)).
// ---------------------------------------------------------------------
// With dis != nullptr (one model, FP32-refined solve) the pass also forms
// the Jacobi scaling of the mixed solve: dis = 1/sqrt(CCC) and the FP32
// entries A'(i,S) = CCS d(i) d(S), A'(i,W) = CCW d(i) d(i-1), the
// neighbours' CCC recomputed as CC*CCMNMI (the same bits as theirs).
// With rctr != nullptr the first workgroup also clears the persistent
// solve's barrier counters (nctr words) and abort flag, which the solve
// launched next would otherwise clear in a launch of its own.
__global__ __launch_bounds__(EFDC_BLOCK) void k_puv_scale(
    HPtr h, const int *__restrict__ lw, const int *__restrict__ le,
    const int *__restrict__ ln, const int *__restrict__ ls, int nwet, int lc0, CgState *st,
    const double *segmin, const int *segid, double *dis, float *as32, float *aw32,
    unsigned *rctr, int *rabort, int nctr) {
  if (rctr && blockIdx.x == 0) {
    for (int q = threadIdx.x; q < nctr; q += blockDim.x) rctr[q] = 0u;
    if (threadIdx.x == 0) *rabort = 0;
  }
  int tile;
  const int i = cell_index(nwet, tile);
  if (i > nwet) return;
  // one CCMNM per model; per basin when the cells form several basins
  const double CCMNMI = 1. / (segmin ? segmin[segid[i]] : st->ccmnm);
  const double cs = HA(CS)[i], cw = HA(CW)[i], ce = HA(CE)[i], cn = HA(CN)[i];
  HA(CCS)[i] = cs * CCMNMI;
  const double ccw = cw * CCMNMI;
  HA(CCW)[i] = ccw;
  HA(CCE)[i] = ce * CCMNMI;
  HA(CCN)[i] = cn * CCMNMI;
  const double ccc = HA(CC)[i] * CCMNMI;
  HA(CCC)[i] = ccc;
  HA(FPTMP)[i] = HA(FP)[i] * CCMNMI;
  HA(CCCI)[i] = 1. / ccc;
  const int n = ln[i], e = le[i], w = lw[i];
  // the neighbours' values are loaded whatever the tests below find (every
  // index is a valid cell, lc0 included), so the loads are in flight together
  const double cs_n = HA(CS)[n], cw_e = HA(CW)[e];
  bool bad = false;
  if (n != lc0 && cn != cs_n) bad = true;
  if (e != lc0 && ce != cw_e) bad = true;
  if (w == lc0 && ccw != 0.0) bad = true;
  if (bad) atomicOr(&st->symok, 1);
  if (dis) {
    const int js = ls[i];
    const double cc_s = HA(CC)[js], cc_w = HA(CC)[i - 1];
    const double di = 1.0 / sqrt(ccc);
    const double ds = (js != lc0) ? 1.0 / sqrt(cc_s * CCMNMI) : 0.0;
    const double dw = (i - 1 > 0) ? 1.0 / sqrt(cc_w * CCMNMI) : 0.0;
    const double cs2 = (js != lc0) ? cs * CCMNMI : 0.0;
    dis[i] = di;
    as32[i] = (float)(cs2 * di * ds);
    aw32[i] = (float)(ccw * di * dw);
  }
}

// Same check on an already scaled system (standalone CONGRAD calls).
__global__ __launch_bounds__(EFDC_BLOCK) void k_cg_check(
    const double *ccs, const double *ccw, const double *cce, const double *ccn,
    const int *__restrict__ lw, const int *__restrict__ le,
    const int *__restrict__ ln, int nwet, int lc0, CgState *st) {
  int tile;
  const int i = cell_index(nwet, tile);
  if (blockIdx.x == 0 && threadIdx.x == 0) (void)0;
  if (i > nwet) return;
  const int n = ln[i], e = le[i], w = lw[i];
  bool bad = false;
  if (n != lc0 && ccn[i] != ccs[n]) bad = true;
  if (e != lc0 && cce[i] != ccw[e]) bad = true;
  if (w == lc0 && ccw[i] != 0.0) bad = true;
  if (bad) atomicOr(&st->symok, 1);
}

__global__ void k_clear_negdep(CgState *st) { st->negdep = 0; }

__global__ void k_reset_flags(CgState *st) {
  st->symok = 0;
  st->negdep = 0;
}


// ---------------------------------------------------------------------
// CONGRAD, FP64 (EFDC/congrad.f90)
// ---------------------------------------------------------------------
struct CgArgs {
  const double *ccc, *ccs, *ccw, *ccn, *cce, *ccci, *fptmp;
  double *x, *r, *ap;
  const int *lw, *le, *ls, *ln;
  int nwet, lc0;
  int rowmajor;
  CgState *st;
  double *part;
  unsigned *tk;
};

// Off-diagonal coefficients of row i in the reference's sense, with
// land neighbours contributing zero (the reference multiplies by the
// zero ghost entry of PCG).
struct Row {
  double cs, cn, cw, ce;
  int js, jn, jw, je;
};

__device__ __forceinline__ Row row_of(const CgArgs &a, int i, bool fast) {
  Row q;
  q.js = a.ls[i];
  q.jn = a.ln[i];
  q.cs = a.ccs[i];
  q.cw = a.ccw[i];
  // the coefficients are loaded whatever the neighbour (every index is a
  // valid cell, lc0 included) and selected after, so that the loads are in
  // flight together
  if (fast) {
    q.jw = i - 1;
    q.je = i + 1;
    const double cn = a.ccs[q.jn], ce = a.ccw[q.je];
    q.cn = (q.jn != a.lc0) ? cn : 0.0;
    q.ce = (q.je != a.lc0) ? ce : 0.0;
  } else {
    q.jw = a.lw[i];
    q.je = a.le[i];
    const double cn = a.ccn[i], ce = a.cce[i];
    q.cn = (q.jn != a.lc0) ? cn : 0.0;
    q.ce = (q.je != a.lc0) ? ce : 0.0;
    if (q.jw == a.lc0) q.cw = 0.0;
  }
  if (q.js == a.lc0) q.cs = 0.0;
  return q;
}

// Per-cell work of the three CG kernels over the cells i0 .. i1-1 of one
// tile (cell i0 + u*EFDC_BLOCK + threadIdx.x for u < CG_CPT).  Shared by
// the single-model kernels (tile t: i0 = 1 + t*CG_CPT*EFDC_BLOCK) and the
// per-basin kernels (tiles counted from each basin's first cell), so one
// basin of an ensemble and the same basin alone give the same bits.
template <int CPT = CG_CPT>
__device__ __forceinline__ void cg_init_cells(const CgArgs &a, int i0, int i1, bool fast,
                                              double (&v)[2]) {
  // every lane computes a row (past i1, the tile's last cell again: i0 <
  // i1 for every tile launched) and only the own rows are kept, so the loads
  // of all the thread's rows are in flight together
#pragma unroll
  for (int u = 0; u < CPT; ++u) {
    const int i = i0 + u * EFDC_BLOCK + (int)threadIdx.x;
    const bool own = i < i1;
    const int ic = own ? i : i1 - 1;
    const Row q = row_of(a, ic, fast);
    const double *P = a.x;
    // RCG = FPTMP - CCC*P - CCN*PN - CCS*PS - CCW*PW - CCE*PE (:73)
    const double fpt = a.fptmp[ic];
    const double r = fpt - a.ccc[ic] * P[ic] - q.cn * P[q.jn] - q.cs * P[q.js] -
                     q.cw * P[q.jw] - q.ce * P[q.je];
    const double z = r * a.ccci[ic];  // PCG = RCG*CCCI (:80)
    if (own) a.r[i] = r;
    v[0] += own ? r * z : 0.0;
    v[1] += own ? fpt * fpt : 0.0;
  }
}

__device__ __forceinline__ void cg_init_state(CgState *st, const double (&s)[2], double rsqm,
                                              double rel_tol) {
  st->rpcg = s[0];
  st->bnorm2 = s[1];
  st->rsqm = rel_tol > 0.0 ? rel_tol * rel_tol * s[1] : rsqm;
  st->beta = 0.0;
  st->alpha = 0.0;
  st->iter = 0;
  st->rsq = 0.0;
  st->status = 0;
  st->done = (s[0] == 0.0) ? 1 : 0;  // :88
}

__global__ __launch_bounds__(EFDC_BLOCK) void k_cg_init(CgArgs a, double rsqm,
                                                        double rel_tol) {
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const bool fast = a.rowmajor && (a.st->symok == 0);
  double v[2] = {0.0, 0.0};
  cg_init_cells(a, 1 + tile * CG_CPT * EFDC_BLOCK, a.nwet + 1, fast, v);
  double s[2];
  if (grid_sum<2>(v, a.part, a.tk + TS_CG_INIT * TICKET_SITE, gridDim.x, tile, s) &&
      threadIdx.x == 0)
    cg_init_state(a.st, s, rsqm, rel_tol);
}

__device__ __forceinline__ double pnew_at(const CgArgs &a, const double *pold,
                                          double beta, int j) {
  // PCG = TMPCG + BETA*PCG with TMPCG = CCCI*RCG (:134,158); ghosts hold 0
  // (loaded whatever j, a valid cell, and selected)
  const double v = a.ccci[j] * a.r[j] + beta * pold[j];
  return (j > 0 && j != a.lc0) ? v : 0.0;
}

template <int CPT = CG_CPT>
__device__ __forceinline__ void cg_a_cells(const CgArgs &a, int i0, int i1, bool fast,
                                           double beta, const double *pold, double *pnew,
                                           double (&v)[1]) {
#pragma unroll
  for (int u = 0; u < CPT; ++u) {
    const int i = i0 + u * EFDC_BLOCK + (int)threadIdx.x;
    const bool own = i < i1;
    const int ic = own ? i : i1 - 1;  // as in cg_init_cells
    const Row q = row_of(a, ic, fast);
    const double pc = a.ccci[ic] * a.r[ic] + beta * pold[ic];
    const double ps = pnew_at(a, pold, beta, q.js);
    const double pn = pnew_at(a, pold, beta, q.jn);
    const double pw = pnew_at(a, pold, beta, q.jw);
    const double pe = pnew_at(a, pold, beta, q.je);
    // APCG = CCC*PCG + CCS*PS + CCN*PN + CCW*PW + CCE*PE (:113)
    const double apc = a.ccc[ic] * pc + q.cs * ps + q.cn * pn + q.cw * pw + q.ce * pe;
    if (own) {
      pnew[i] = pc;
      a.ap[i] = apc;
    }
    v[0] += own ? apc * pc : 0.0;
  }
}

__global__ __launch_bounds__(EFDC_BLOCK) void k_cg_a(CgArgs a,
                                                     const double *pold,
                                                     double *pnew) {
  if (a.st->done) return;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const bool fast = a.rowmajor && (a.st->symok == 0);
  double v[1] = {0.0};
  cg_a_cells(a, 1 + tile * CG_CPT * EFDC_BLOCK, a.nwet + 1, fast, a.st->beta, pold, pnew, v);
  double s[1];
  if (grid_sum<1>(v, a.part, a.tk + TS_CG_A * TICKET_SITE, gridDim.x, tile, s) &&
      threadIdx.x == 0) {
    a.st->papcg = s[0];
    a.st->alpha = a.st->rpcg / s[0];  // :124
  }
}

template <int CPT = CG_CPT>
__device__ __forceinline__ void cg_b_cells(const CgArgs &a, int i0, int i1, double alpha,
                                           const double *pnew, double (&v)[2]) {
#pragma unroll
  for (int u = 0; u < CPT; ++u) {
    const int i = i0 + u * EFDC_BLOCK + (int)threadIdx.x;
    const bool own = i < i1;
    const int ic = own ? i : i1 - 1;  // as in cg_init_cells
    const double x = a.x[ic] + alpha * pnew[ic];  // :132
    const double r = a.r[ic] - alpha * a.ap[ic];  // :133
    const double z = a.ccci[ic] * r;              // :134
    if (own) {
      a.x[i] = x;
      a.r[i] = r;
    }
    v[0] += own ? r * z : 0.0;
    v[1] += own ? r * r : 0.0;
  }
}

// beta, rsq, the stop test and the iteration cap after the r.z, r.r sums;
// returns true when this call ended the solve.
__device__ __forceinline__ bool cg_b_state(CgState *st, const double (&s)[2], int iterm) {
  const double beta = s[0] / st->rpcg;  // :147-148
  st->beta = beta;
  st->rpcg = s[0];
  st->rsq = s[1];
  st->iter += 1;
  if (s[1] <= st->rsqm) { st->done = 1; return true; }             // :165
  if (st->iter >= iterm) { st->done = 1; st->status = 1; return true; }
  return false;
}

__global__ __launch_bounds__(EFDC_BLOCK) void k_cg_b(CgArgs a,
                                                     const double *pnew,
                                                     int iterm) {
  if (a.st->done) return;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  double v[2] = {0.0, 0.0};
  cg_b_cells(a, 1 + tile * CG_CPT * EFDC_BLOCK, a.nwet + 1, a.st->alpha, pnew, v);
  double s[2];
  if (grid_sum<2>(v, a.part, a.tk + TS_CG_B * TICKET_SITE, gridDim.x, tile, s) &&
      threadIdx.x == 0)
    (void)cg_b_state(a.st, s, iterm);
}

// ---------------------------------------------------------------------
// Independent basins: the same three kernels over per-basin tiles.  Tile b
// of the launch covers cells tiles[b].x .. tiles[b].y-1 of basin
// tiles[b].z; its partials go to part[j*ntiles + b]; the last tile of a
// basin to arrive sums the basin's partials (tiles segtile0[s] ..
// segtile0[s+1]-1, in order) exactly as the single-model reduction sums
// its blocks, and updates the basin's state.  Tiles of a basin whose solve
// has ended return at once.
// ---------------------------------------------------------------------
struct SegArgs {
  const int4 *tiles;
  const int *segtile0;
  CgState *segst;
  unsigned *segtk;
  unsigned *segdone;
  double *part;
  int ntiles;
};

// Per-basin arrival; true in the block that arrives last for its basin.
__device__ __forceinline__ bool seg_last(const SegArgs &g, int seg, int ntile_seg) {
  __shared__ bool am_last;
  if (threadIdx.x == 0) {
    __builtin_amdgcn_s_waitcnt(0);
    unsigned *t = g.segtk + (size_t)seg * TICKET_PAD;
    am_last = atomic_inc(t) == (unsigned)ntile_seg - 1u;
    if (am_last) reset_ticket(t);
  }
  __syncthreads();
  return am_last;
}

template <int NV>
__device__ __forceinline__ bool seg_sum(const SegArgs &g, int b, int seg, double (&v)[NV],
                                        double (&out)[NV]) {
  block_sum<NV>(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < NV; ++j) publish(&g.part[(size_t)j * g.ntiles + b], v[j]);
  }
  const int t0 = g.segtile0[seg], nt = g.segtile0[seg + 1] - t0;
  if (!seg_last(g, seg, nt)) return false;
  double acc[NV];
  sum_parts<NV>(g.part + t0, g.ntiles, nt, acc);
  __shared__ double res[NV];
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < NV; ++j) res[j] = acc[j];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < NV; ++j) out[j] = res[j];
  return true;
}

template <int CPT>
__global__ __launch_bounds__(EFDC_BLOCK) void k_scg_init(CgArgs a, SegArgs g, double rsqm,
                                                         double rel_tol) {
  const int b = blockIdx.x;
  const int4 t = g.tiles[b];
  const bool fast = a.rowmajor && (a.st->symok == 0);
  double v[2] = {0.0, 0.0};
  cg_init_cells<CPT>(a, t.x, t.y, fast, v);
  double s[2];
  if (seg_sum<2>(g, b, t.z, v, s) && threadIdx.x == 0) {
    CgState *st = g.segst + t.z;
    cg_init_state(st, s, rsqm, rel_tol);
    if (st->done) atomic_inc(g.segdone);
  }
}

template <int CPT>
__global__ __launch_bounds__(EFDC_BLOCK) void k_scg_a(CgArgs a, SegArgs g, const double *pold,
                                                      double *pnew) {
  const int b = blockIdx.x;
  const int4 t = g.tiles[b];
  CgState *st = g.segst + t.z;
  if (st->done) return;
  const bool fast = a.rowmajor && (a.st->symok == 0);
  double v[1] = {0.0};
  cg_a_cells<CPT>(a, t.x, t.y, fast, st->beta, pold, pnew, v);
  double s[1];
  if (seg_sum<1>(g, b, t.z, v, s) && threadIdx.x == 0) {
    st->papcg = s[0];
    st->alpha = st->rpcg / s[0];  // :124
  }
}

template <int CPT>
__global__ __launch_bounds__(EFDC_BLOCK) void k_scg_b(CgArgs a, SegArgs g, const double *pnew,
                                                      int iterm) {
  const int b = blockIdx.x;
  const int4 t = g.tiles[b];
  CgState *st = g.segst + t.z;
  if (st->done) return;
  double v[2] = {0.0, 0.0};
  cg_b_cells<CPT>(a, t.x, t.y, st->alpha, pnew, v);
  double s[2];
  if (seg_sum<2>(g, b, t.z, v, s) && threadIdx.x == 0) {
    if (cg_b_state(st, s, iterm)) atomic_inc(g.segdone);
  }
}

// Per-basin minimum of CC over all its cells after SETOPENBC: the
// reference's minimum over the interior cells (:360-364) and over the
// open-boundary cells (:365-375) taken together.  The minimum is exact in
// any order, so it is taken in two launches: one workgroup per CG tile of a
// basin (the whole chip busy), then one workgroup per basin over its tiles.
__device__ __forceinline__ double block_min(double m) {
  m = wave_min(m);
  __shared__ double sh[EFDC_BLOCK / 64];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  double mm = sh[0];
  for (int q = 1; q < (int)(blockDim.x >> 6); ++q) mm = fmin(mm, sh[q]);
  return mm;
}

__global__ __launch_bounds__(EFDC_BLOCK) void k_seg_min_tiles(const double *cc,
                                                              const int4 *tiles,
                                                              double *tmin) {
  const int4 t = tiles[blockIdx.x];
  double m = 1.E+18;
  for (int i = t.x + (int)threadIdx.x; i < t.y; i += blockDim.x) m = fmin(m, cc[i]);
  m = block_min(m);
  if (threadIdx.x == 0) tmin[blockIdx.x] = m;
}

__global__ __launch_bounds__(EFDC_BLOCK) void k_seg_min(const double *tmin, const int *segtile0,
                                                        double *segmin) {
  const int sgi = blockIdx.x;
  const int t0 = segtile0[sgi], t1 = segtile0[sgi + 1];
  double m = 1.E+18;
  for (int t = t0 + (int)threadIdx.x; t < t1; t += blockDim.x) m = fmin(m, tmin[t]);
  m = block_min(m);
  if (threadIdx.x == 0) segmin[sgi] = m;
}

__global__ void k_seg_reset(CgState *segst, int nseg, unsigned *segdone) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nseg) segst[s].done = 0;
  if (s == 0) *segdone = 0u;
}

// ---------------------------------------------------------------------
// Mixed precision: FP32 inner CG on the Jacobi-scaled operator
// A' = D^-1/2 A D^-1/2 (unit diagonal), FP64 dot products, FP64
// residual-based refinement of P.  Uses the compact symmetric operator.
// ---------------------------------------------------------------------
struct MxArgs {
  CgArgs g;
  double *dis;       // 1/sqrt(CCC)
  float *as, *aw;    // scaled off-diagonals (south, west) in FP32
  float *r32, *y32, *ap32;
};

// Per-cell work of the mixed solve, shared by the single-model kernels and
// the per-basin ones.
__device__ __forceinline__ void mx_prep_cell(const MxArgs &m, int i, double (&v)[1]) {
  const CgArgs &a = m.g;
  const double di = 1.0 / sqrt(a.ccc[i]);
  m.dis[i] = di;
  const int js = a.ls[i];
  const double ds = (js != a.lc0) ? 1.0 / sqrt(a.ccc[js]) : 0.0;
  const double dw = (i - 1 > 0) ? 1.0 / sqrt(a.ccc[i - 1]) : 0.0;
  const double cs = (js != a.lc0) ? a.ccs[i] : 0.0;
  m.as[i] = (float)(cs * di * ds);
  m.aw[i] = (float)(a.ccw[i] * di * dw);
  v[0] += a.fptmp[i] * a.fptmp[i];
}

// r64 = FPTMP - A P (FP64, compact operator), r' = D^-1/2 r64 in FP32, y = 0
__device__ __forceinline__ void mx_resid_cell(const MxArgs &m, int i, double (&v)[2]) {
  const CgArgs &a = m.g;
  const Row q = row_of(a, i, true);
  const double *P = a.x;
  const double r = a.fptmp[i] - a.ccc[i] * P[i] - q.cn * P[q.jn] -
                   q.cs * P[q.js] - q.cw * P[q.jw] - q.ce * P[q.je];
  const float rs = (float)(m.dis[i] * r);
  m.r32[i] = rs;
  m.y32[i] = 0.0f;
  v[0] += r * r;
  v[1] += (double)rs * (double)rs;
}

// (loaded whatever j, a valid cell lc0 included, and selected, so that the
// loads of a thread's cells are in flight together)
// (p = r + beta p in FP32, one fused multiply-add, for a cell's own p and
// for its neighbours' alike)
__device__ __forceinline__ float p32_at(const MxArgs &m, const float *pold,
                                        float beta, int j) {
  const float pj = fmaf(beta, pold[j], m.r32[j]);
  return (j > 0 && j != m.g.lc0) ? pj : 0.0f;
}

// own = false: a lane past the cells (i clamped to a wet cell by the
// caller) computes the row and keeps nothing
__device__ __forceinline__ void mx_a_cell(const MxArgs &m, int i, bool own, float beta,
                                          const float *pold, float *pnew, double (&v)[1]) {
  const CgArgs &a = m.g;
  const int js = a.ls[i], jn = a.ln[i];
  const float pc = p32_at(m, pold, beta, i);
  const float ps = p32_at(m, pold, beta, js);
  const float pn = p32_at(m, pold, beta, jn);
  const float pw = p32_at(m, pold, beta, i - 1);
  const float pe = p32_at(m, pold, beta, i + 1);
  const float as_n = m.as[jn], aw_e = m.aw[i + 1];
  const float an = (jn != a.lc0) ? as_n : 0.0f;
  const float ae = (i + 1 != a.lc0) ? aw_e : 0.0f;
  const float apc = pc + m.as[i] * ps + an * pn + m.aw[i] * pw + ae * pe;
  if (own) {
    pnew[i] = pc;
    m.ap32[i] = apc;
  }
  v[0] += own ? (double)apc * (double)pc : 0.0;
}

__device__ __forceinline__ void mx_b_cell(const MxArgs &m, int i, bool own, float alpha,
                                          const float *pnew, double (&v)[1]) {
  // the vector updates in FP32 (fused multiply-add); the sums stay FP64
  const float y = fmaf(alpha, pnew[i], m.y32[i]);
  const float r = fmaf(-alpha, m.ap32[i], m.r32[i]);
  if (own) {
    m.y32[i] = y;
    m.r32[i] = r;
  }
  v[0] += own ? (double)r * (double)r : 0.0;
}

__global__ __launch_bounds__(EFDC_BLOCK) void k_mx_prep(MxArgs m, double rsqm,
                                                        double rel_tol) {
  const CgArgs &a = m.g;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  double v[1] = {0.0};
  #pragma unroll
  for (int u = 0; u < MX_CPT; ++u) {
    const int i = (tile * MX_CPT + u) * EFDC_BLOCK + (int)threadIdx.x;
    if (i >= 1 && i <= a.nwet) mx_prep_cell(m, i, v);
  }
  double s[1];
  if (grid_sum_x<1>(v, a.part, a.tk + TS_MX_PREP * TICKET_SITE, gridDim.x, tile, s) &&
      threadIdx.x == 0) {
    a.st->bnorm2 = s[0];
    a.st->rsqm = rel_tol > 0.0 ? rel_tol * rel_tol * s[0] : rsqm;
    a.st->outer = 0;
    a.st->iter = 0;
    a.st->status = 0;
  }
}

// starts one inner solve
__global__ __launch_bounds__(EFDC_BLOCK) void k_mx_resid(MxArgs m) {
  const CgArgs &a = m.g;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  double v[2] = {0.0, 0.0};
  #pragma unroll
  for (int u = 0; u < MX_CPT; ++u) {
    const int i = (tile * MX_CPT + u) * EFDC_BLOCK + (int)threadIdx.x;
    if (i >= 1 && i <= a.nwet) mx_resid_cell(m, i, v);
  }
  double s[2];
  if (grid_sum_x<2>(v, a.part, a.tk + TS_MX_RESID * TICKET_SITE, gridDim.x, tile, s) &&
      threadIdx.x == 0) {
    CgState *st = a.st;
    st->rsq64 = s[0];
    st->rsq = s[0];
    st->bnorm2_in = s[1];
    st->rz_in = s[1];
    st->beta = 0.0;
    st->inner_iter = 0;
    st->done = (s[0] <= st->rsqm || s[1] == 0.0) ? 1 : 0;
  }
}

__global__ __launch_bounds__(EFDC_BLOCK) void k_mx_a(MxArgs m, const float *pold,
                                                     float *pnew) {
  const CgArgs &a = m.g;
  if (a.st->done) return;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const float beta = (float)a.st->beta;
  double v[1] = {0.0};
  #pragma unroll
  for (int u = 0; u < MX_CPT; ++u) {
    const int i = (tile * MX_CPT + u) * EFDC_BLOCK + (int)threadIdx.x;
    const bool own = i >= 1 && i <= a.nwet;
    mx_a_cell(m, own ? i : (i < 1 ? 1 : a.nwet), own, beta, pold, pnew, v);
  }
  double s[1];
  if (grid_sum_x<1>(v, a.part, a.tk + TS_MX_A * TICKET_SITE, gridDim.x, tile, s) &&
      threadIdx.x == 0)
    a.st->alpha = a.st->rz_in / s[0];
}

__global__ __launch_bounds__(EFDC_BLOCK) void k_mx_b(MxArgs m, const float *pnew,
                                                     double rtol2, int cap) {
  const CgArgs &a = m.g;
  if (a.st->done) return;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const float alpha = (float)a.st->alpha;
  double v[1] = {0.0};
  #pragma unroll
  for (int u = 0; u < MX_CPT; ++u) {
    const int i = (tile * MX_CPT + u) * EFDC_BLOCK + (int)threadIdx.x;
    const bool own = i >= 1 && i <= a.nwet;
    mx_b_cell(m, own ? i : (i < 1 ? 1 : a.nwet), own, alpha, pnew, v);
  }
  double s[1];
  if (grid_sum_x<1>(v, a.part, a.tk + TS_MX_B * TICKET_SITE, gridDim.x, tile, s) &&
      threadIdx.x == 0) {
    CgState *st = a.st;
    st->beta = s[0] / st->rz_in;
    st->rz_in = s[0];
    st->inner_iter += 1;
    st->iter += 1;
    if (s[0] <= rtol2 * st->bnorm2_in || st->inner_iter >= cap) st->done = 1;
  }
}

__global__ __launch_bounds__(EFDC_BLOCK) void k_mx_apply(MxArgs m) {
  const CgArgs &a = m.g;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
#pragma unroll
  for (int u = 0; u < MX_CPT; ++u) {
    const int i = (tile * MX_CPT + u) * EFDC_BLOCK + (int)threadIdx.x;
    if (i >= 1 && i <= a.nwet) a.x[i] = a.x[i] + m.dis[i] * (double)m.y32[i];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) a.st->outer += 1;
}

// ---------------------------------------------------------------------
// The mixed solve of independent basins: the per-cell work above over the
// per-basin tiles of 1024 cells, every reduction, state, stop test and
// outer correction per basin.  A basin whose FP64 residual meets its stop
// test is finished (fin) and skips the later rounds; the host runs outer
// rounds until every basin is finished or the round cap is reached, and
// inner iterations until every running basin's inner solve has stopped
// (counters: segfin of finished basins, segdone of stopped inner solves).
// ---------------------------------------------------------------------
struct SMxArgs {
  MxArgs m;
  SegArgs g;
  unsigned *segfin;
};

template <int CPT, typename F>
__device__ __forceinline__ void tile_cells(const int4 &t, F f) {
#pragma unroll
  for (int u = 0; u < CPT; ++u) {
    const int i = t.x + u * EFDC_BLOCK + (int)threadIdx.x;
    if (i < t.y) f(i);
  }
}

template <int CPT>
__global__ __launch_bounds__(EFDC_BLOCK) void k_smx_prep(SMxArgs x, double rsqm,
                                                         double rel_tol) {
  const int b = blockIdx.x;
  const int4 t = x.g.tiles[b];
  double v[1] = {0.0};
  tile_cells<CPT>(t, [&](int i) { mx_prep_cell(x.m, i, v); });
  double s[1];
  if (seg_sum<1>(x.g, b, t.z, v, s) && threadIdx.x == 0) {
    CgState *st = x.g.segst + t.z;
    st->bnorm2 = s[0];
    st->rsqm = rel_tol > 0.0 ? rel_tol * rel_tol * s[0] : rsqm;
    st->outer = 0;
    st->iter = 0;
    st->status = 0;
    st->fin = 0;
    st->done = 0;
  }
}

template <int CPT>
__global__ __launch_bounds__(EFDC_BLOCK) void k_smx_resid(SMxArgs x, int max_outer) {
  const int b = blockIdx.x;
  const int4 t = x.g.tiles[b];
  CgState *st = x.g.segst + t.z;
  if (st->fin) return;
  double v[2] = {0.0, 0.0};
  tile_cells<CPT>(t, [&](int i) { mx_resid_cell(x.m, i, v); });
  double s[2];
  if (seg_sum<2>(x.g, b, t.z, v, s) && threadIdx.x == 0) {
    st->rsq64 = s[0];
    st->rsq = s[0];
    st->bnorm2_in = s[1];
    st->rz_in = s[1];
    st->beta = 0.0;
    st->inner_iter = 0;
    const bool conv = s[0] <= st->rsqm;
    const bool cap = !conv && st->outer >= max_outer;
    if (conv || cap) {
      st->fin = 1;
      if (cap) st->status = 1;
      atomic_inc(x.segfin);
    }
    st->done = (conv || cap || s[1] == 0.0) ? 1 : 0;
    if (st->done) atomic_inc(x.g.segdone);
  }
}

template <int CPT>
__global__ __launch_bounds__(EFDC_BLOCK) void k_smx_a(SMxArgs x, const float *pold, float *pnew) {
  const int b = blockIdx.x;
  const int4 t = x.g.tiles[b];
  CgState *st = x.g.segst + t.z;
  if (st->done) return;
  const float beta = (float)st->beta;
  double v[1] = {0.0};
  tile_cells<CPT>(t, [&](int i) { mx_a_cell(x.m, i, true, beta, pold, pnew, v); });
  double s[1];
  if (seg_sum<1>(x.g, b, t.z, v, s) && threadIdx.x == 0) st->alpha = st->rz_in / s[0];
}

template <int CPT>
__global__ __launch_bounds__(EFDC_BLOCK) void k_smx_b(SMxArgs x, const float *pnew, double rtol2,
                                                      int cap) {
  const int b = blockIdx.x;
  const int4 t = x.g.tiles[b];
  CgState *st = x.g.segst + t.z;
  if (st->done) return;
  const float alpha = (float)st->alpha;
  double v[1] = {0.0};
  tile_cells<CPT>(t, [&](int i) { mx_b_cell(x.m, i, true, alpha, pnew, v); });
  double s[1];
  if (seg_sum<1>(x.g, b, t.z, v, s) && threadIdx.x == 0) {
    st->beta = s[0] / st->rz_in;
    st->rz_in = s[0];
    st->inner_iter += 1;
    st->iter += 1;
    if (s[0] <= rtol2 * st->bnorm2_in || st->inner_iter >= cap) {
      st->done = 1;
      atomic_inc(x.g.segdone);
    }
  }
}

// P += D^-1/2 y for the basins that ran an inner solve in this round
template <int CPT>
__global__ __launch_bounds__(EFDC_BLOCK) void k_smx_apply(SMxArgs x) {
  const int b = blockIdx.x;
  const int4 t = x.g.tiles[b];
  CgState *st = x.g.segst + t.z;
  if (st->fin) return;
  tile_cells<CPT>(t, [&](int i) {
    x.m.g.x[i] = x.m.g.x[i] + x.m.dis[i] * (double)x.m.y32[i];
  });
  if (t.w == 0 && threadIdx.x == 0) st->outer += 1;
}

__global__ void k_fill_u32x2(unsigned *a, unsigned *b) {
  if (threadIdx.x == 0) {
    *a = 0u;
    if (b) *b = 0u;
  }
}

// ---------------------------------------------------------------------
// Post-solve: EFDC/calpuv2c.f90:423-434 (face transports), :472-477
// (depths), :490-492 (open boundaries), :896-914 (P, HU, HV, inverses),
// EFDC/negdep.f90:34-45 (negative depth flag).
// ---------------------------------------------------------------------
__global__ __launch_bounds__(EFDC_BLOCK) void k_puv_post1(
    HPtr h, const int *__restrict__ lw, const int *__restrict__ ls, int nwet,
    double dt) {
  int tile;
  const int i = cell_index(nwet, tile);
  if (i > nwet) return;
  const double DELTD2 = 0.5 * dt;
  const double *P = HA(P);
  const int w = lw[i], s = ls[i];
  const double u = HA(SUB)[i] * (HA(FUHDYE)[i] - DELTD2 * HA(HRUO)[i] * HA(RCX)[i] *
                                                    HA(H1U)[i] * (P[i] - P[w]));
  const double v = HA(SVB)[i] * (HA(FVHDXE)[i] - DELTD2 * HA(HRVO)[i] * HA(RCY)[i] *
                                                    HA(H1V)[i] * (P[i] - P[s]));
  HA(UHDYE)[i] = u;
  HA(VHDXE)[i] = v;
  HA(UHE)[i] = u * HA(DYIU)[i];
  HA(VHE)[i] = v * HA(DXIV)[i];
}

__global__ __launch_bounds__(EFDC_BLOCK) void k_puv_post2(
    HPtr h, const int *__restrict__ le, const int *__restrict__ ln,
    const unsigned char *__restrict__ pbcflag, int nwet, double dt, double G) {
  int tile;
  const int i = cell_index(nwet, tile);
  if (i > nwet) return;
  const double DELTD2 = 0.5 * dt, GI = 1. / G;
  const int e = le[i], n = ln[i];
  const double *U = HA(UHDYE), *U1 = HA(UHDY1E), *V = HA(VHDXE), *V1 = HA(VHDX1E);
  double hp = HA(H1P)[i] + DELTD2 * HA(DXYIP)[i] *
                               (2. * HA(QSUME)[i] -
                                (U[e] + U1[e] - U[i] - U1[i] + V[n] + V1[n] - V[i] - V1[i]));
  if (pbcflag[i]) hp = GI * HA(P)[i] - HA(BELV)[i];
  HA(HP)[i] = hp;
  HA(P)[i] = G * (hp + HA(BELV)[i]);
}

__global__ __launch_bounds__(EFDC_BLOCK) void k_puv_post3(
    HPtr h, const int *__restrict__ lw, const int *__restrict__ ls, int nwet,
    CgState *st) {
  int tile;
  const int i = cell_index(nwet, tile);
  if (i > nwet) return;
  const int w = lw[i], s = ls[i];
  const double *HP = HA(HP), *DXYP = HA(DXYP);
  const double hu = (DXYP[i] * HP[i] + DXYP[w] * HP[w]) * (0.5 / HA(DXYU)[i]);
  const double hv = (DXYP[i] * HP[i] + DXYP[s] * HP[s]) * (0.5 / HA(DXYV)[i]);
  HA(HU)[i] = hu;
  HA(HV)[i] = hv;
  HA(HPI)[i] = 1. / HP[i];
  HA(HUI)[i] = 1. / hu;
  HA(HVI)[i] = 1. / hv;
  if (HP[i] <= 0.) st->negdep = 1;
}

// ---------------------------------------------------------------------
// Drying, ISDRY > 0: open boundaries (all), cells (1..97), cells with
// rewetting (99), draining of isolated cells (NDRYSTP < 0), LMASKDRY.
// ---------------------------------------------------------------------
// Step start, EFDC/calpuv2c.f90:169-177: ISCDRY = 0, SUB1/SVB1 = SUB/SVB.
__global__ __launch_bounds__(EFDC_BLOCK) void k_dry_begin(HPtr h, int *iscdry, double *sub1,
                                                          double *svb1, double *qdwaste,
                                                          int nwet) {
  int tile;
  const int i = cell_index(nwet, tile);
  if (i > nwet) return;
  if (qdwaste) qdwaste[i] = 0.;  // NDRYSTP < 0: the rate is per step (:181-188)
  iscdry[i] = 0;
  sub1[i] = HA(SUB)[i];
  svb1[i] = HA(SVB)[i];
}

// LOPENBCDRY = .FALSE. on the open-boundary cells, :153-158.
__global__ void k_dry_obc_reset(PbcDev s, int *lopenbcdry) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < s.n) lopenbcdry[s.L[t]] = 0;
}

// Open-boundary cells that dry, :497-530: the depth is reset and the cell
// and its faces are closed for the rest of the step.  Each thread writes
// its own cell and zeros faces it shares with its east / north
// neighbours (the same value from every writer).
__global__ void k_dry_obc(HPtr h, PbcDev s, const int *le, const int *ln, int *lopenbcdry,
                          int *iscdry, double dt, double G, double HDRY) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= s.n) return;
  const int i = s.L[t];
  double *HP = HA(HP);
  const double hp = HP[i], h1p = HA(H1P)[i];
  const bool dry = lopenbcdry[i] != 0;
  if (!((dry || (h1p < 0.9 * HDRY && hp < h1p) || hp < 0.) && iscdry[i] == 0)) return;
  double hn;
  if (hp <= 0. && dry) hn = 0.2 * HDRY;
  else hn = fmin(fmax(h1p, 0.1 * HDRY), 0.9 * HDRY);
  HP[i] = hn;
  iscdry[i] = 1;
  const int e = le[i], n = ln[i];
  HA(SUB)[i] = 0.;
  HA(SVB)[i] = 0.;
  HA(SUB)[e] = 0.;
  HA(SVB)[n] = 0.;
  HA(UHDYE)[i] = 0.;
  HA(UHDYE)[e] = 0.;
  HA(VHDXE)[i] = 0.;
  HA(VHDXE)[n] = 0.;
  lopenbcdry[i] = 1;
  const double DELTI = 1. / dt;
  HA(CC)[i] = DELTI * HA(DXYP)[i];
  const double pv = (hn + HA(BELV)[i]) * G;
  HA(P)[i] = pv;
  HA(FP1)[i] = DELTI * HA(DXYP)[i] * pv;
}

// st->drychg = 1 if any thread of the workgroup passes `mine`: one LDS
// flag per workgroup, then one plain store by thread 0 unless the flag
// already reads 1 (atomics or one store per wave on the one address
// serialise a pass in which many cells change).  Every thread of the
// workgroup must call it (two barriers).
__device__ __forceinline__ void raise_flag_once(bool mine, CgState *st) {
  __shared__ int any;
  if (threadIdx.x == 0) any = 0;
  __syncthreads();
  if (mine) any = 1;
  __syncthreads();
  if (threadIdx.x == 0 && any &&
      __hip_atomic_load(&st->drychg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
    __hip_atomic_store(&st->drychg, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Cells that dry, :574-602: HP <= HDRY closes the four faces of the cell;
// a cell that becomes dry raises st->drychg (ICORDRY = 1: solve again)
// through raise_flag_once.  Faces already closed are not written again.
__global__ __launch_bounds__(EFDC_BLOCK) void k_dry_cells(HPtr h, const int *le, const int *ln,
                                                          int *iscdry, int nwet, double HDRY,
                                                          CgState *st) {
  int tile;
  const int i = cell_index(nwet, tile);
  bool newly = false;
  if (i <= nwet && HA(HP)[i] <= HDRY) {
    if (iscdry[i] == 0) {
      iscdry[i] = 1;
      newly = true;
    }
    const int e = le[i], n = ln[i];
    double *SUB = HA(SUB), *SVB = HA(SVB);
    if (SUB[i] != 0.) SUB[i] = 0.;
    if (SVB[i] != 0.) SVB[i] = 0.;
    if (SUB[e] != 0.) SUB[e] = 0.;
    if (SVB[n] != 0.) SVB[n] = 0.;
  }
  raise_flag_once(newly, st);
}

__global__ void k_dry_flag_reset(CgState *st) { st->drychg = 0; }

// Draining of isolated wet cells, NDRYSTP < 0 (:838-893, 1010-1018): a wet
// cell (HP >= HDRY) whose four faces are closed counts its isolated steps
// in NATDRY (an open face resets it); past -NDRYSTP steps, if a face that
// is open on the full grid leads to a lower bed, or HP < 2 HDRY, HP and
// H1P go to 0.9 HDRY with the layer depths, and the volume taken out is
// booked in QDWASTE (per second) and VDWASTE.  P keeps the value of the
// last solve pass, as in the reference.  Each thread changes only its own
// cell; the neighbours' beds and the face switches are only read.
template <typename T>
__global__ __launch_bounds__(EFDC_BLOCK) void k_dry_waste(
    HPtr h, const int *lw, const int *le, const int *ls, const int *ln, int *natdry,
    double *qdwaste, double *vdwaste, T *hpk, T *hpki, T *h1pk, const T *dzc, int64_t pitch,
    int KC, int nwet, double HDRY, int ntmp, double dt) {
  int tile;
  const int i = cell_index(nwet, tile);
  if (i > nwet) return;
  double *HP = HA(HP);
  const double hp = HP[i];
  if (!(hp >= HDRY)) return;
  const int w = lw[i], e = le[i], s = ls[i], n = ln[i];
  const double *SUB = HA(SUB), *SVB = HA(SVB);
  if (SUB[i] + SUB[e] + SVB[i] + SVB[n] > 0.5) {
    natdry[i] = 0;
    return;
  }
  const int cnt = natdry[i] + 1;
  natdry[i] = cnt;
  if (cnt <= ntmp) return;
  const double *SUBO = HA(SUBO), *SVBO = HA(SVBO), *BELV = HA(BELV);
  const double b = BELV[i];
  double RVAL = 0.;
  if (SUBO[e] > 0.5 && BELV[e] < b) RVAL = RVAL + 1.;
  if (SUBO[i] > 0.5 && BELV[w] < b) RVAL = RVAL + 1.;
  if (SVBO[n] > 0.5 && BELV[n] < b) RVAL = RVAL + 1.;
  if (SVBO[i] > 0.5 && BELV[s] < b) RVAL = RVAL + 1.;
  if (!(RVAL > 0. || hp < HDRY * 2.)) return;
  const double hn = 0.90 * HDRY;
  HP[i] = hn;
  HA(H1P)[i] = hn;
  for (int k = 0; k < KC; ++k) {
    const int64_t j = i + (int64_t)k * pitch;
    const double dz = (double)dzc[j];
    const double hk = hn * dz;
    hpk[j] = (T)hk;
    hpki[j] = (T)(1. / hk);
    h1pk[j] = (T)(hn * dz);
  }
  natdry[i] = 0;
  const double dxyp = HA(DXYP)[i];
  qdwaste[i] = (1. / dt) * dxyp * (hp - hn);
  vdwaste[i] = vdwaste[i] + dxyp * (hp - hn);
}

// Drying and rewetting, ISDRY = 99 (:604-716).  The reference walks the
// cells in L order and updates the face switches in place; a cell's four
// faces are shared with its neighbours, and a dry cell (HP <= HDRY) rewrites
// all four.  Case of a dry cell (dry99_case): 1 rising fast enough to stay
// wet (faces closed, then those with inflow from a neighbour deeper than
// 2 HDRY reopened), 2 rising too slowly (faces closed), 3 falling and low
// (faces closed), 0 no change.  Whatever a neighbour wrote on a face that
// it shares with a dry cell is 0 (a face reopens only towards a cell deeper
// than 2 HDRY), so the values a cell finds on its faces when its turn comes
// are 0 where a neighbour with a smaller L had a case, else the values of
// the pass start; and the final value of a face is the write of the cell
// with the larger L among the two that wrote it.  Each thread owns the west
// and south faces of its cell and writes them to sub_n / svb_n (the pass
// start values stay readable until k_dry99_commit).
__device__ __forceinline__ int dry99_case(const double *HP, const double *H1P, int j, int nwet,
                                          double HDRY) {
  if (j < 1 || j > nwet) return 0;  // L = 2..LA only
  const double hp = HP[j], h1p = H1P[j];
  if (!(hp <= HDRY)) return 0;
  const double dh = hp - h1p;
  if (dh > 0.0) return dh > (HDRY - h1p) * 0.01 ? 1 : 2;
  return (hp < 0.9 * HDRY || h1p < HDRY) ? 3 : 0;
}

__global__ __launch_bounds__(EFDC_BLOCK) void k_dry99(HPtr h, const int *lw, const int *le,
                                                      const int *ls, const int *ln,
                                                      int *iscdry, double *sub_n,
                                                      double *svb_n, int nwet, double HDRY,
                                                      CgState *st) {
  int tile;
  const int i = cell_index(nwet, tile);
  bool chg = false;
  if (i <= nwet) {
    const double *HP = HA(HP), *H1P = HA(H1P), *SUB = HA(SUB), *SVB = HA(SVB),
                 *SUBO = HA(SUBO), *SVBO = HA(SVBO), *UHDYE = HA(UHDYE), *VHDXE = HA(VHDXE);
    const double HDRY2 = 2. * HDRY;
    const int w = lw[i], e = le[i], s = ls[i], n = ln[i];
    const int ci = dry99_case(HP, H1P, i, nwet, HDRY);
    const int cw = dry99_case(HP, H1P, w, nwet, HDRY);
    const int cs = dry99_case(HP, H1P, s, nwet, HDRY);
    // this cell's writes on its W, E, S, N faces (case 1 may reopen)
    double ww = 0., we = 0., wsv = 0., wn = 0.;
    if (ci == 1) {
      if (SUBO[i] > 0.5 && UHDYE[i] > 0.0 && HP[w] > HDRY2) ww = SUBO[i];
      if (SUBO[e] > 0.5 && UHDYE[e] < 0.0 && HP[e] > HDRY2) we = SUBO[e];
      if (SVBO[i] > 0.5 && VHDXE[i] > 0.0 && HP[s] > HDRY2) wsv = SVBO[i];
      if (SVBO[n] > 0.5 && VHDXE[n] < 0.0 && HP[n] > HDRY2) wn = SVBO[n];
    }
    // the west / south neighbours' writes on the faces they share with i
    double ow = 0., os = 0.;
    if (cw == 1 && SUBO[i] > 0.5 && UHDYE[i] < 0.0 && HP[i] > HDRY2) ow = SUBO[i];
    if (cs == 1 && SVBO[i] > 0.5 && VHDXE[i] < 0.0 && HP[i] > HDRY2) os = SVBO[i];
    // final W and S faces: the later writer (larger L) wins
    double fw = SUB[i], fs = SVB[i];
    if (ci != 0 && cw != 0) fw = (i > w) ? ww : ow;
    else if (ci != 0) fw = ww;
    else if (cw != 0) fw = ow;
    if (ci != 0 && cs != 0) fs = (i > s) ? wsv : os;
    else if (ci != 0) fs = wsv;
    else if (cs != 0) fs = os;
    sub_n[i] = fw;
    svb_n[i] = fs;
    if (ci == 1) {
      const int ce = dry99_case(HP, H1P, e, nwet, HDRY);
      const int cn = dry99_case(HP, H1P, n, nwet, HDRY);
      const double sw0 = (cw != 0 && w < i) ? 0. : SUB[i];
      const double se0 = (ce != 0 && e < i) ? 0. : SUB[e];
      const double ss0 = (cs != 0 && s < i) ? 0. : SVB[i];
      const double sn0 = (cn != 0 && n < i) ? 0. : SVB[n];
      iscdry[i] = (ww + we + wsv + wn) < 0.5 ? 1 : 0;
      chg = fabs(ww - sw0) > 0.5 || fabs(we - se0) > 0.5 || fabs(wsv - ss0) > 0.5 ||
            fabs(wn - sn0) > 0.5;
    } else if (ci != 0 && iscdry[i] == 0) {
      iscdry[i] = 1;
      chg = true;
    }
  }
  raise_flag_once(chg, st);
}

__global__ __launch_bounds__(EFDC_BLOCK) void k_dry99_commit(HPtr h, const double *sub_n,
                                                             const double *svb_n, int nwet) {
  int tile;
  const int i = cell_index(nwet, tile);
  if (i > nwet) return;
  HA(SUB)[i] = sub_n[i];
  HA(SVB)[i] = svb_n[i];
}

// Transport mask, :916-939: a cell is left out of the transport when
// HP < HDRY and each of its four faces was closed at the step start and
// is closed now.
__global__ __launch_bounds__(EFDC_BLOCK) void k_dry_mask(HPtr h, const int *le, const int *ln,
                                                         const double *sub1, const double *svb1,
                                                         int *lmask, int nwet, double HDRY) {
  int tile;
  const int i = cell_index(nwet, tile);
  if (i > nwet) return;
  int m = 1;
  if (HA(HP)[i] < HDRY) {
    const int e = le[i], n = ln[i];
    const double *SUB = HA(SUB), *SVB = HA(SVB);
    const int iue = sub1[i] < 0.5 && SUB[i] < 0.5;
    const int iuw = sub1[e] < 0.5 && SUB[e] < 0.5;
    const int ivs = svb1[i] < 0.5 && SVB[i] < 0.5;
    const int ivn = svb1[n] < 0.5 && SVB[n] < 0.5;
    if (iue + iuw + ivs + ivn == 4) m = 0;
  }
  lmask[i] = m;
}

CgArgs cg_args(efdc_ctx *c) {
  CgArgs a;
  std::memset(&a, 0, sizeof a);  // the bytes are a graph key
  a.ccc = c->dH[EFDC_CCC];
  a.ccs = c->dH[EFDC_CCS];
  a.ccw = c->dH[EFDC_CCW];
  a.ccn = c->dH[EFDC_CCN];
  a.cce = c->dH[EFDC_CCE];
  a.ccci = c->dH[EFDC_CCCI];
  a.fptmp = c->dH[EFDC_FPTMP];
  a.x = c->dH[EFDC_P];
  a.r = c->r;
  a.ap = c->ap;
  a.lw = c->dLW;
  a.le = c->dLE;
  a.ls = c->dLS;
  a.ln = c->dLN;
  a.nwet = c->n_wet();
  a.lc0 = c->LC - 1;
  a.rowmajor = c->rowmajor ? 1 : 0;
  a.st = c->st;
  a.part = c->part;
  a.tk = c->tickets;
  return a;
}

int read_state(efdc_ctx *c) {
  HIP_TRY(c, hipMemcpyAsync(c->st_host, c->st, sizeof(CgState),
                            hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return EFDC_OK;
}

// ---------------------------------------------------------------------
// Captured CG iterations: the FP64 multi-kernel solves replay hipGraphs of
// 8 and of 1 iterations (two launches each) instead of launching every
// kernel, the same number of iterations between polls of the solve state
// as the launch-by-launch loop (the last solve's count first, then 4 at a
// time).  The 8-iteration graph starts on the even direction buffer and
// ends on it; the 1-iteration graph is captured for the buffer parity it
// is replayed at (two graphs).  Iterations after the stop test are
// launches that return at once (the kernels read the state's done flag),
// so the results are those of the launch-by-launch loop bit for bit.  A
// graph is captured again when its launch arguments (the key) change.
// With the library's timers on, the kernels are launched one by one so
// that each is timed.
// ---------------------------------------------------------------------
constexpr int CG_GRAPH_ITERS = 8;

// replay `n` iterations starting at iteration `it0` with graph slots g8
// (8 iterations, even start) and g1[2] (one iteration from parity 0 / 1)
inline int cg_graph_replay(efdc_ctx *c, hipGraphExec_t g8, hipGraphExec_t g1e,
                           hipGraphExec_t g1o, int it0, int n) {
  int it = it0, left = n;
  while (left > 0) {
    if ((it & 1) == 0 && left >= CG_GRAPH_ITERS) {
      HIP_TRY(c, hipGraphLaunch(g8, c->stream));
      it += CG_GRAPH_ITERS;
      left -= CG_GRAPH_ITERS;
    } else {
      HIP_TRY(c, hipGraphLaunch((it & 1) ? g1o : g1e, c->stream));
      ++it;
      --left;
    }
  }
  return EFDC_OK;
}

template <typename F>
int cg_graph_get(efdc_ctx *c, int slot, const std::vector<unsigned char> &key, F enqueue) {
  if (c->cg_gexec[slot] && c->cg_gkey[slot] == key) return EFDC_OK;
  if (c->cg_gexec[slot]) (void)hipGraphExecDestroy(c->cg_gexec[slot]);
  c->cg_gexec[slot] = nullptr;
  c->cg_gkey[slot].clear();
  HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  enqueue();
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (e == hipSuccess) e = hipGraphInstantiate(&c->cg_gexec[slot], g, nullptr, nullptr, 0);
  if (g) (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    c->cg_gexec[slot] = nullptr;
    c->err = std::string("CG graph capture: ") + hipGetErrorString(e);
    return EFDC_ERR_HIP;
  }
  c->cg_gkey[slot] = key;
  return EFDC_OK;
}

template <typename... A>
void key_of(std::vector<unsigned char> &k, const A &...a) {
  k.clear();
  (k.insert(k.end(), (const unsigned char *)&a, (const unsigned char *)&a + sizeof(a)), ...);
}

// A timed trial of the two replay modes of a solve (`solve(graph)` runs
// it whole), each from the P the solve starts from; the faster becomes
// the context's mode and P is restored.
template <typename F>
int cg_graph_trial(efdc_ctx *c, F solve) {
  int s;
  const size_t bytes = sizeof(double) * c->LCM;
  if (!c->tune_psave && (s = dev_alloc(c, (void **)&c->tune_psave, bytes))) return s;
  HIP_TRY(c, hipMemcpyAsync(c->tune_psave, c->dH[EFDC_P], bytes, hipMemcpyDeviceToDevice,
                            c->stream));
  hipEvent_t e0, e1;
  HIP_TRY(c, hipEventCreate(&e0));
  HIP_TRY(c, hipEventCreate(&e1));
  float ms[2] = {0.f, 0.f};
  s = EFDC_OK;
  for (int g = 0; g < 2 && !s; ++g) {
    hipError_t e = hipMemcpyAsync(c->dH[EFDC_P], c->tune_psave, bytes, hipMemcpyDeviceToDevice,
                                  c->stream);
    if (e == hipSuccess) e = hipEventRecord(e0, c->stream);
    if (e != hipSuccess) {
      c->err = hipGetErrorString(e);
      s = EFDC_ERR_HIP;
      break;
    }
    if ((s = solve(g == 1))) break;
    e = hipEventRecord(e1, c->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms[g], e0, e1);
    if (e != hipSuccess) {
      c->err = hipGetErrorString(e);
      s = EFDC_ERR_HIP;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (s) return s;
  c->cg_graph = ms[1] < ms[0];
  c->cg_graph_tune = false;
  const char *names[2] = {"cg_fp64 launches", "cg_fp64 graphs"};
  for (int g = 0; g < 2; ++g)
    if (c->tune_n < 8) {
      std::snprintf(c->tune_name[c->tune_n], 32, "%s", names[g]);
      c->tune_ms[c->tune_n] = ms[g];
      c->tune_chosen[c->tune_n] = (g == 1) == c->cg_graph ? 1 : 0;
      ++c->tune_n;
    }
  HIP_TRY(c, hipMemcpyAsync(c->dH[EFDC_P], c->tune_psave, bytes, hipMemcpyDeviceToDevice,
                            c->stream));
  return EFDC_OK;
}

int cg_fp64_solve(efdc_ctx *c, bool graph) {
  const int nb = cg_blocks(c->n_wet());
  CgArgs a = cg_args(c);
  const int iterm = c->prm.ITERM;
  TIMED(c, TK_CG_INIT, k_cg_init<<<nb, EFDC_BLOCK, 0, c->stream>>>(a, c->prm.RSQM, c->prm.cg_rel_tol));
  int it = 0;
  if (graph) {
    std::vector<unsigned char> key;
    double *p0 = c->pb[0], *p1 = c->pb[1];
    int s;
    for (int q = 0; q < 3; ++q) {
      const int nit = q == 0 ? CG_GRAPH_ITERS : 1, par = q == 2 ? 1 : 0;
      key_of(key, a, p0, p1, iterm, nb, q);
      if ((s = cg_graph_get(c, q, key, [&] {
             for (int j = par; j < par + nit; ++j) {
               k_cg_a<<<nb, EFDC_BLOCK, 0, c->stream>>>(a, c->pb[j & 1], c->pb[(j + 1) & 1]);
               k_cg_b<<<nb, EFDC_BLOCK, 0, c->stream>>>(a, c->pb[(j + 1) & 1], iterm);
             }
           })))
        return s;
    }
    int chunk = c->last_iters > 0 ? c->last_iters : 8;
    for (;;) {
      const int n = std::min(chunk, iterm - it);
      if ((s = cg_graph_replay(c, c->cg_gexec[0], c->cg_gexec[1], c->cg_gexec[2], it, n)))
        return s;
      it += n;
      if ((s = read_state(c))) return s;
      if (c->st_host->done || it >= iterm) break;
      chunk = 4;
    }
  } else {
    int chunk = c->last_iters > 0 ? c->last_iters : 8;
    for (;;) {
      for (int j = 0; j < chunk && it < iterm; ++j, ++it) {
        TIMED(c, TK_CG_A, k_cg_a<<<nb, EFDC_BLOCK, 0, c->stream>>>(a, c->pb[it & 1], c->pb[(it + 1) & 1]));
        TIMED(c, TK_CG_B, k_cg_b<<<nb, EFDC_BLOCK, 0, c->stream>>>(a, c->pb[(it + 1) & 1], iterm));
      }
      int s = read_state(c);
      if (s) return s;
      if (c->st_host->done || it >= iterm) break;
      chunk = 4;
    }
  }
  c->last_iters = c->st_host->iter;
  if (c->timing) c->cg_iter_count += c->st_host->iter;
  return EFDC_OK;
}

int cg_fp64(efdc_ctx *c) {
  int s;
  if (c->cg_graph_tune && !c->timing &&
      (s = cg_graph_trial(c, [&](bool g) { return cg_fp64_solve(c, g); })))
    return s;
  return cg_fp64_solve(c, c->cg_graph && !c->timing);
}

int cg_mixed(efdc_ctx *c) {
  const int nb = mx_blocks(c->n_wet());
  MxArgs m;
  m.g = cg_args(c);
  m.dis = c->dis;
  m.as = c->as32;
  m.aw = c->aw32;
  m.r32 = c->r32;
  m.y32 = c->y32;
  m.ap32 = c->ap32;
  const double rtol = c->prm.mixed_inner_rtol > 0 ? c->prm.mixed_inner_rtol : 1e-5;
  const int cap = c->prm.ITERM;
  TIMED(c, TK_MX_PREP, k_mx_prep<<<nb, EFDC_BLOCK, 0, c->stream>>>(m, c->prm.RSQM, c->prm.cg_rel_tol));
  int outer = 0, total = 0;
  for (;;) {
    TIMED(c, TK_MX_RESID, k_mx_resid<<<nb, EFDC_BLOCK, 0, c->stream>>>(m));
    int s = read_state(c);
    if (s) return s;
    if (c->st_host->rsq64 <= c->st_host->rsqm) break;
    if (outer >= c->prm.mixed_max_outer) {
      // report through the status word like the FP64 path
      c->st_host->status = 1;
      break;
    }
    int it = 0;
    int chunk = c->last_iters > 0 ? c->last_iters : 8;
    for (;;) {
      for (int j = 0; j < chunk && it < cap; ++j, ++it) {
        TIMED(c, TK_MX_A, k_mx_a<<<nb, EFDC_BLOCK, 0, c->stream>>>(m, c->p32[it & 1], c->p32[(it + 1) & 1]));
        TIMED(c, TK_MX_B, k_mx_b<<<nb, EFDC_BLOCK, 0, c->stream>>>(m, c->p32[(it + 1) & 1], rtol * rtol, cap));
      }
      s = read_state(c);
      if (s) return s;
      if (c->st_host->done || it >= cap) break;
      chunk = 4;
    }
    c->last_iters = c->st_host->inner_iter;
    total += c->st_host->inner_iter;
    TIMED(c, TK_MX_APPLY, k_mx_apply<<<nb, EFDC_BLOCK, 0, c->stream>>>(m));
    ++outer;
  }
  c->st_host->iter = total;
  if (c->timing) c->cg_iter_count += total;
  c->st_host->outer = outer;
  return EFDC_OK;
}

// Statistics and status of a finished solve from the host copy of the state.
int cg_finish(efdc_ctx *c, bool mixed, efdc_cg_stats_t *stats) {
  const CgState &h = *c->st_host;
  if (stats) {
    stats->iters = h.iter;
    stats->outer = mixed ? h.outer : 1;
    stats->rsq = mixed ? h.rsq64 : h.rsq;
    stats->rsqm_used = h.rsqm;
    stats->ccmnm = h.ccmnm;
  }
  if (h.status != 0) {
    c->err = "CONGRAD: maximum iterations exceeded in external solution";
    return EFDC_ERR_CG_MAXITER;
  }
  return EFDC_OK;
}

// defer: a persistent solve leaves the read-back of its state to the caller
// (puv_run reads it with the rest of the step); *ran_mixed reports the mode.
int cg_dispatch(efdc_ctx *c, efdc_cg_stats_t *stats, bool defer = false,
                bool *ran_mixed = nullptr) {
  int s;
  c->pcg_deferred = false;
  if (c->nseg > 1) return seg_cg_run(c, stats, ran_mixed);
  bool mixed = (c->prm.cg_mode == EFDC_CG_MIXED);
  if (mixed) {
    // the mixed operator is the compact symmetric one; the symmetry flag
    // is read from the device once and kept (it follows from the grid,
    // the face masks and the boundary tables)
    if (c->symok_cache < 0) {
      if ((s = read_state(c))) return s;
      c->symok_cache = c->st_host->symok;
    }
    if (!c->rowmajor || c->symok_cache != 0) mixed = false;
  }
  if (ran_mixed) *ran_mixed = mixed;
  if (c->prm.cg_kernel == EFDC_CG_KERNEL_PERSISTENT && !pcg_applicable(c, mixed)) {
    c->err = "persistent CG requested but the rows do not fit one co-resident grid";
    return EFDC_ERR_ARG;
  }
  if (pcg_applicable(c, mixed) && mixed && c->tune_cg_pending &&
      c->prm.mixed_precond_degree == 0) {
    c->tune_cg_pending = false;
    if ((s = pcg_tune_degree(c))) return s;
  }
  if (pcg_applicable(c, mixed))
    s = pcg_solve(c, mixed, defer);
  else
    s = mixed ? cg_mixed(c) : cg_fp64(c);
  if (s) return s;
  if (c->pcg_deferred) return EFDC_OK;
  return cg_finish(c, mixed, stats);
}

}  // namespace

void cg_graphs_free(efdc_ctx *c) {
  for (int q = 0; q < 6; ++q) {
    if (c->cg_gexec[q]) (void)hipGraphExecDestroy(c->cg_gexec[q]);
    c->cg_gexec[q] = nullptr;
    c->cg_gkey[q].clear();
  }
}


int cg_fp64_vec(efdc_ctx *c, double *x, const double *rhs, CgState *st, CgState *st_host,
                double rel_tol, int itmax, int *iters, double *rsq) {
  const int nb = cg_blocks(c->n_wet());
  CgArgs a = cg_args(c);
  a.x = x;
  a.fptmp = rhs;
  a.st = st;
  // the operator's symmetry flag (row_of's compact form) comes from the
  // forward solve of the same coefficients
  HIP_TRY(c, hipMemcpyAsync(st, c->st, sizeof(CgState), hipMemcpyDeviceToDevice, c->stream));
  TIMED(c, TK_ADJ_CG, k_cg_init<<<nb, EFDC_BLOCK, 0, c->stream>>>(a, 0.0, rel_tol));
  int it = 0, chunk = 8;
  for (;;) {
    for (int j = 0; j < chunk && it < itmax; ++j, ++it) {
      TIMED(c, TK_ADJ_CG, k_cg_a<<<nb, EFDC_BLOCK, 0, c->stream>>>(a, c->pb[it & 1],
                                                                  c->pb[(it + 1) & 1]));
      TIMED(c, TK_ADJ_CG, k_cg_b<<<nb, EFDC_BLOCK, 0, c->stream>>>(a, c->pb[(it + 1) & 1],
                                                                  itmax));
    }
    HIP_TRY(c, hipMemcpyAsync(st_host, st, sizeof(CgState), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (st_host->done || it >= itmax) break;
    chunk = 8;
  }
  *iters = st_host->iter;
  *rsq = st_host->rsq;
  if (!st_host->done || st_host->status) {
    c->err = "adjoint CG: no convergence within the iteration cap";
    return EFDC_ERR_CG_MAXITER;
  }
  return EFDC_OK;
}

int cg_run(efdc_ctx *c, bool check_sym, efdc_cg_stats_t *stats) {
  const int nb = puv_blocks(c->n_wet());
  if (check_sym) {
    k_reset_flags<<<1, 1, 0, c->stream>>>(c->st);
    k_cg_check<<<nb, EFDC_BLOCK, 0, c->stream>>>(
        c->dH[EFDC_CCS], c->dH[EFDC_CCW], c->dH[EFDC_CCE], c->dH[EFDC_CCN],
        c->dLW, c->dLE, c->dLN, c->n_wet(), c->LC - 1, c->st);
  }
  return cg_dispatch(c, stats);
}

int puv_run(efdc_ctx *c, const efdc_step_t *step, efdc_cg_stats_t *stats) {
  const int nb = puv_blocks(c->n_wet());
  const int nwet = c->n_wet(), lc0 = c->LC - 1;
  const double dt = step->dt, G = c->prm.G;
  // time-level shifts of EFDC/calpuv2c.f90:256-265 as buffer rotations
  static const int pairs[6][2] = {{EFDC_UHDYE, EFDC_UHDY1E}, {EFDC_VHDXE, EFDC_VHDX1E},
                                  {EFDC_HU, EFDC_H1U},       {EFDC_HV, EFDC_H1V},
                                  {EFDC_HUI, EFDC_H1UI},     {EFDC_HVI, EFDC_H1VI}};
  Rotated rot;
  for (int q = 0; q < 6; ++q) {
    std::swap(c->dH[pairs[q][0]], c->dH[pairs[q][1]]);
    rot.pair[q][0] = c->dH[pairs[q][0]];
    rot.pair[q][1] = c->dH[pairs[q][1]];
  }
  {
    double *hp = c->dH[EFDC_HP], *h1p = c->dH[EFDC_H1P], *h2p = c->dH[EFDC_H2P];
    c->dH[EFDC_H2P] = h1p;
    c->dH[EFDC_H1P] = hp;
    c->dH[EFDC_HP] = h2p;
  }
  rot.tri[0] = c->dH[EFDC_HP];
  rot.tri[1] = c->dH[EFDC_H1P];
  rot.tri[2] = c->dH[EFDC_H2P];
  rot.lc0 = lc0;
  HPtr h = hptr(c);
  const bool dry = c->prm.ISDRY > 0;
  if (dry && c->nseg > 1) {
    c->err = "drying (ISDRY > 0) with several basins is not built";
    return EFDC_ERR_ARG;
  }
  if (c->prm.ISDRY == 99) {
    if (!c->mutual) {
      c->err = "ISDRY = 99 needs mutual neighbour maps (LEC(LWC(L)) = L, LNC(LSC(L)) = L)";
      return EFDC_ERR_ARG;
    }
    if (!c->sub_n) {
      const size_t n = (size_t)c->pitch;
      if (hipMalloc((void **)&c->sub_n, n * sizeof(double)) != hipSuccess ||
          hipMalloc((void **)&c->svb_n, n * sizeof(double)) != hipSuccess) {
        c->err = "hipMalloc failed (ISDRY = 99 face buffers)";
        return EFDC_ERR_NOMEM;
      }
    }
  }
  uvw_step_begin(c);
  // without drying or adjoint recording the explicit and assembly passes
  // do the step start themselves
  const int fold = (!dry && !adj_recording(c)) ? 1 : 0;
  if (!fold) k_step_begin<<<1, 64, 0, c->stream>>>(c->st, rot);
  if (dry) {
    k_dry_begin<<<nb, EFDC_BLOCK, 0, c->stream>>>(h, c->iscdry, c->sub1, c->svb1,
                                                  c->prm.NDRYSTP < 0 ? c->qdwaste : nullptr,
                                                  nwet);
    for (int q = 0; q < 4; ++q)
      if (c->pbc[q].n > 0)
        k_dry_obc_reset<<<(c->pbc[q].n + 255) / 256, 256, 0, c->stream>>>(c->pbc[q],
                                                                           c->lopenbcdry);
  }
  if (c->store == EFDC_STORE_FP64)
    TIMED(c, TK_PUV_EXPLICIT, k_puv_explicit<double><<<nb, EFDC_BLOCK, 0, c->stream>>>(
        h, c->dLW, c->dLS, c->dLE, c->dLN, (const double *)c->dK[EFDC_HPK - EFDC_HPK],
        (double *)c->dK[EFDC_H1PK - EFDC_HPK], c->pitch, c->KC, nwet, dt, G,
        c->prm.RITB1, c->prm.SNLT, dry ? 1 : 0, c->prm.implicit_drag ? 1 : 0, rot, fold,
        c->st));
  else
    TIMED(c, TK_PUV_EXPLICIT, k_puv_explicit<float><<<nb, EFDC_BLOCK, 0, c->stream>>>(
        h, c->dLW, c->dLS, c->dLE, c->dLN, (const float *)c->dK[EFDC_HPK - EFDC_HPK],
        (float *)c->dK[EFDC_H1PK - EFDC_HPK], c->pitch, c->KC, nwet, dt, G,
        c->prm.RITB1, c->prm.SNLT, dry ? 1 : 0, c->prm.implicit_drag ? 1 : 0, rot, fold,
        c->st));
  c->ncordry = 0;
  // a step that may solve more than once (drying) or records the solution
  // (adjoint) launches the persistent solve cooperatively, so that it never
  // has to be solved again after the passes that follow it
  c->pcg_force_coop = dry || adj_recording(c);
  bool ran_mixed = false;
  int cgs = EFDC_OK;
  for (;;) {
    // label 1000 of EFDC/calpuv2c.f90:302: the drying loop re-enters here
    // (fold is 0 on the re-solves of the drying loop: only steps without
    // drying fold, and they assemble once)
    TIMED(c, TK_PUV_ASSEMBLE, k_puv_assemble<<<nb, EFDC_BLOCK, 0, c->stream>>>(h, c->dLE, c->dLN, c->pbcflag,
                                                     nwet, dt, G, c->part, c->st,
                                                     c->tickets, rot, fold));
    if (c->npbc_total > 0) {
      // tidal phase on the host in FP64 like SETOPENBC (:22-32)
      const double PI2 = 2. * 3.1415926535898;  // EFDC/aaefdc.f90:650-651
      const double TN = dt * (double)step->n + c->prm.TCON_TBEGIN;
      for (int M = 0; M < c->mtide; ++M) {
        double TM = std::fmod(TN, c->tcp_host[M]);
        TM = PI2 * TM / c->tcp_host[M];
        c->tide_host[M] = std::cos(TM);
        c->tide_host[c->mtide + M] = std::sin(TM);
      }
      if (c->mtide > 0)
        HIP_TRY(c, hipMemcpyAsync(c->tide_dev, c->tide_host,
                                  sizeof(double) * 2 * c->mtide,
                                  hipMemcpyHostToDevice, c->stream));
      PbcArgs b;
      for (int s = 0; s < 4; ++s) b.side[s] = c->pbc[s];
      b.tide = c->tide_dev;
      b.mtide = c->mtide;
      b.lopenbcdry = c->lopenbcdry;
      int ntot = 0;
      for (int q = 0; q < 4; ++q) {
        if (c->pbc[q].n == 0) continue;
        ntot += c->pbc[q].n;
        const int nbs = c->pbc[q].serial ? 1 : (c->pbc[q].n + EFDC_BLOCK - 1) / EFDC_BLOCK;
        TIMED(c, TK_PUV_BC, k_puv_bc_side<<<nbs, EFDC_BLOCK, 0, c->stream>>>(
            h, b, c->pbc[q], q, c->dLW, c->dLE, c->dLS, c->dLN, dt, G, c->prm.HDRY));
      }
      TIMED(c, TK_PUV_BC, k_puv_bc_min<<<(ntot + EFDC_BLOCK - 1) / EFDC_BLOCK, EFDC_BLOCK, 0,
                                         c->stream>>>(h, b, ntot, c->dLW, c->dLE, c->dLS,
                                                      c->dLN, c->part, c->st, c->tickets));
    }
    const bool segs = c->nseg > 1;
    if (segs) {
      int s = seg_min_run(c);
      if (s) return s;
    }
    const bool coef = !segs && c->dis && c->prm.cg_mode == EFDC_CG_MIXED;
    // the persistent solve's counters, once they exist, are cleared here
    const bool rst = !segs && c->pcg_ctr != nullptr && c->pcg_abort != nullptr;
    TIMED(c, TK_PUV_SCALE, k_puv_scale<<<nb, EFDC_BLOCK, 0, c->stream>>>(
        h, c->dLW, c->dLE, c->dLN, c->dLS, nwet, lc0, c->st, segs ? c->d_segmin : nullptr,
        segs ? c->d_segid : nullptr, coef ? c->dis : nullptr, c->as32, c->aw32,
        rst ? c->pcg_ctr : nullptr, rst ? c->pcg_abort : nullptr, PCG_NCTR));
    if (rst) c->pcg_reset_done = true;
    c->mx_coef_ok = coef;
    int s = cg_dispatch(c, stats, /*defer=*/true, &ran_mixed);
    c->mx_coef_ok = false;
    if (s != EFDC_OK && s != EFDC_ERR_CG_MAXITER) return s;
    cgs = s;
    if (adj_recording(c) && (s = adj_record_ps(c))) return s;
    TIMED(c, TK_PUV_POST1, k_puv_post1<<<nb, EFDC_BLOCK, 0, c->stream>>>(h, c->dLW, c->dLS, nwet, dt));
    TIMED(c, TK_PUV_POST2, k_puv_post2<<<nb, EFDC_BLOCK, 0, c->stream>>>(h, c->dLE, c->dLN, c->pbcflag,
                                                  nwet, dt, G));
    if (!dry) break;
    // open-boundary drying (:497-530), then the cell drying and its flag
    // (:574-602); P already is G (HP + BELV) (:559-569, k_puv_post2)
    for (int q = 0; q < 4; ++q)
      if (c->pbc[q].n > 0)
        k_dry_obc<<<(c->pbc[q].n + 255) / 256, 256, 0, c->stream>>>(
            h, c->pbc[q], c->dLE, c->dLN, c->lopenbcdry, c->iscdry, dt, G, c->prm.HDRY);
    const bool cells = c->prm.ISDRY < 98 || c->prm.ISDRY == 99;
    if (cells) {
      k_dry_flag_reset<<<1, 1, 0, c->stream>>>(c->st);
      if (c->prm.ISDRY < 98) {
        k_dry_cells<<<nb, EFDC_BLOCK, 0, c->stream>>>(h, c->dLE, c->dLN, c->iscdry, nwet,
                                                      c->prm.HDRY, c->st);
      } else {
        k_dry99<<<nb, EFDC_BLOCK, 0, c->stream>>>(h, c->dLW, c->dLE, c->dLS, c->dLN,
                                                  c->iscdry, c->sub_n, c->svb_n, nwet,
                                                  c->prm.HDRY, c->st);
        k_dry99_commit<<<nb, EFDC_BLOCK, 0, c->stream>>>(h, c->sub_n, c->svb_n, nwet);
      }
    }
    HIP_TRY(c, hipGetLastError());
    if ((s = read_state(c))) return s;
    if (c->pcg_deferred) {
      c->pcg_deferred = false;
      if (c->st_host->pabort) {
        c->err = "persistent CG: an exchange between workgroups timed out";
        return EFDC_ERR_HIP;
      }
      cgs = cg_finish(c, ran_mixed, stats);
    }
    if (!cells || !c->st_host->drychg || cgs != EFDC_OK) break;
    if (++c->ncordry >= 500) {
      c->err = "NCORDRY > 500: the drying loop did not settle (EFDC/calpuv2c.f90:596-601)";
      return EFDC_ERR_STATE;
    }
  }  // drying loop
  if (dry && c->prm.NDRYSTP < 0) {
    if (c->store == EFDC_STORE_FP64)
      k_dry_waste<double><<<nb, EFDC_BLOCK, 0, c->stream>>>(
          h, c->dLW, c->dLE, c->dLS, c->dLN, c->natdry, c->qdwaste, c->vdwaste,
          (double *)c->dK[0], (double *)c->dK[EFDC_HPKI - EFDC_HPK],
          (double *)c->dK[EFDC_H1PK - EFDC_HPK], (const double *)c->dK[EFDC_DZC - EFDC_HPK],
          c->pitch, c->KC, nwet, c->prm.HDRY, -c->prm.NDRYSTP, dt);
    else
      k_dry_waste<float><<<nb, EFDC_BLOCK, 0, c->stream>>>(
          h, c->dLW, c->dLE, c->dLS, c->dLN, c->natdry, c->qdwaste, c->vdwaste,
          (float *)c->dK[0], (float *)c->dK[EFDC_HPKI - EFDC_HPK],
          (float *)c->dK[EFDC_H1PK - EFDC_HPK], (const float *)c->dK[EFDC_DZC - EFDC_HPK],
          c->pitch, c->KC, nwet, c->prm.HDRY, -c->prm.NDRYSTP, dt);
  }
  TIMED(c, TK_PUV_POST3, k_puv_post3<<<nb, EFDC_BLOCK, 0, c->stream>>>(h, c->dLW, c->dLS, nwet, c->st));
  if (dry)
    k_dry_mask<<<nb, EFDC_BLOCK, 0, c->stream>>>(h, c->dLE, c->dLN, c->sub1, c->svb1,
                                                 c->lmaskdry, nwet, c->prm.HDRY);
  int s;
  if ((s = read_state(c))) return s;
  if (c->pcg_deferred) {
    c->pcg_deferred = false;
    if (c->st_host->pabort && c->pcg_last_plain && !c->pcg_force_coop) {
      // the plain launch was not co-resident: solve again, cooperatively,
      // from the P the step reached (any P is a starting iterate of the same
      // system), then redo the passes that followed the solve
      ++c->pcg_retries;
      c->pcg_force_coop = true;
      s = cg_dispatch(c, stats, /*defer=*/false, &ran_mixed);
      c->pcg_force_coop = false;
      if (s != EFDC_OK && s != EFDC_ERR_CG_MAXITER) return s;
      cgs = s;
      k_clear_negdep<<<1, 1, 0, c->stream>>>(c->st);
      TIMED(c, TK_PUV_POST1, k_puv_post1<<<nb, EFDC_BLOCK, 0, c->stream>>>(h, c->dLW, c->dLS, nwet, dt));
      TIMED(c, TK_PUV_POST2, k_puv_post2<<<nb, EFDC_BLOCK, 0, c->stream>>>(h, c->dLE, c->dLN, c->pbcflag,
                                                    nwet, dt, G));
      TIMED(c, TK_PUV_POST3, k_puv_post3<<<nb, EFDC_BLOCK, 0, c->stream>>>(h, c->dLW, c->dLS, nwet, c->st));
      HIP_TRY(c, hipGetLastError());
      if ((s = read_state(c))) return s;
    } else if (c->st_host->pabort) {
      c->err = "persistent CG: an exchange between workgroups timed out";
      return EFDC_ERR_HIP;
    } else {
      cgs = cg_finish(c, ran_mixed, stats);
    }
  }
  if (c->st_host->symok != c->symok_cache && c->symok_cache >= 0) {
    // cannot happen for a fixed grid and boundary tables (the cache is
    // reset by their setters); refuse rather than keep a solve made with
    // the wrong operator
    c->symok_cache = c->st_host->symok;
    if (ran_mixed) {
      c->err = "the symmetry of the system changed without a setter call";
      return EFDC_ERR_STATE;
    }
  }
  if (stats) stats->ccmnm = c->st_host->ccmnm;
  if (cgs != EFDC_OK) return cgs;
  if (c->st_host->negdep) {
    c->err = "NEGDEP: negative depth after the external solution";
    return EFDC_ERR_NEGDEPTH;
  }
  return EFDC_OK;
}

// ---------------------------------------------------------------------
// Independent basins: per-basin minimum and the FP64 CG over per-basin
// tiles (the multi-kernel schedule: two launches per iteration for all
// basins together, the host polling every few iterations).
// ---------------------------------------------------------------------
int seg_min_run(efdc_ctx *c) {
  // the tile minima go to the CG partials, which the solve rewrites later
  TIMED(c, TK_PUV_SCALE, k_seg_min_tiles<<<c->ntiles, EFDC_BLOCK, 0, c->stream>>>(
      c->dH[EFDC_CC], c->d_tiles, c->d_segpart));
  TIMED(c, TK_PUV_SCALE, k_seg_min<<<c->nseg, EFDC_BLOCK, 0, c->stream>>>(
      c->d_segpart, c->d_segtile0, c->d_segmin));
  HIP_TRY(c, hipGetLastError());
  return EFDC_OK;
}

namespace {

SegArgs seg_args(efdc_ctx *c) {
  SegArgs g;
  std::memset(&g, 0, sizeof g);  // the bytes are a graph key
  g.tiles = c->d_tiles;
  g.segtile0 = c->d_segtile0;
  g.segst = c->d_segst;
  g.segtk = c->d_segtk;
  g.segdone = c->d_segdone;
  g.part = c->d_segpart;
  g.ntiles = c->ntiles;
  return g;
}

template <int CPT>
void seg_iter(efdc_ctx *c, const CgArgs &a, const SegArgs &g, int j, int iterm, bool timed) {
  const int nt = g.ntiles;
  if (timed) {
    TIMED(c, TK_CG_A, k_scg_a<CPT><<<nt, EFDC_BLOCK, 0, c->stream>>>(a, g, c->pb[j & 1],
                                                                     c->pb[(j + 1) & 1]));
    TIMED(c, TK_CG_B, k_scg_b<CPT><<<nt, EFDC_BLOCK, 0, c->stream>>>(a, g, c->pb[(j + 1) & 1],
                                                                     iterm));
  } else {
    k_scg_a<CPT><<<nt, EFDC_BLOCK, 0, c->stream>>>(a, g, c->pb[j & 1], c->pb[(j + 1) & 1]);
    k_scg_b<CPT><<<nt, EFDC_BLOCK, 0, c->stream>>>(a, g, c->pb[(j + 1) & 1], iterm);
  }
}

template <int CPT>
void seg_init(efdc_ctx *c, const CgArgs &a, const SegArgs &g) {
  k_seg_reset<<<(c->nseg + 255) / 256, 256, 0, c->stream>>>(c->d_segst, c->nseg, c->d_segdone);
  TIMED(c, TK_CG_INIT, k_scg_init<CPT><<<g.ntiles, EFDC_BLOCK, 0, c->stream>>>(
      a, g, c->prm.RSQM, c->prm.cg_rel_tol));
}

// One solve of every basin with tile set q: graph replays (or, with the
// timers on or during the tile trials, launch by launch) until every
// basin has stopped or ITERM iterations ran.
template <int CPT>
int seg_solve(efdc_ctx *c, bool graph) {
  const CgArgs a = cg_args(c);
  const SegArgs g = seg_args(c);
  const int iterm = c->prm.ITERM;
  seg_init<CPT>(c, a, g);
  int it = 0;
  unsigned done = 0;
  if (graph) {
    std::vector<unsigned char> key;
    double *p0 = c->pb[0], *p1 = c->pb[1];
    const int cpt = CPT;
    int s;
    for (int q = 0; q < 3; ++q) {
      const int nit = q == 0 ? CG_GRAPH_ITERS : 1, par = q == 2 ? 1 : 0;
      key_of(key, a, g, p0, p1, iterm, cpt, q);
      if ((s = cg_graph_get(c, 3 + q, key, [&] {
             for (int j = par; j < par + nit; ++j) seg_iter<CPT>(c, a, g, j, iterm, false);
           })))
        return s;
    }
    int chunk = c->last_iters > 0 ? c->last_iters : 8;
    for (;;) {
      const int n = std::min(chunk, iterm - it);
      if ((s = cg_graph_replay(c, c->cg_gexec[3], c->cg_gexec[4], c->cg_gexec[5], it, n)))
        return s;
      it += n;
      HIP_TRY(c, hipMemcpyAsync(&done, c->d_segdone, sizeof(unsigned), hipMemcpyDeviceToHost,
                                c->stream));
      HIP_TRY(c, hipStreamSynchronize(c->stream));
      if ((int)done >= c->nseg || it >= iterm) break;
      chunk = 4;
    }
  } else {
    int chunk = c->last_iters > 0 ? c->last_iters : 8;
    for (;;) {
      for (int j = 0; j < chunk && it < iterm; ++j, ++it) seg_iter<CPT>(c, a, g, it, iterm, true);
      HIP_TRY(c, hipMemcpyAsync(&done, c->d_segdone, sizeof(unsigned), hipMemcpyDeviceToHost,
                                c->stream));
      HIP_TRY(c, hipStreamSynchronize(c->stream));
      if ((int)done >= c->nseg || it >= iterm) break;
      chunk = 4;
    }
  }
  HIP_TRY(c, hipGetLastError());
  return EFDC_OK;
}

int seg_solve_set(efdc_ctx *c, int q, bool graph) {
  switch (q) {
    case 0: return seg_solve<1>(c, graph);
    case 1: return seg_solve<2>(c, graph);
    case 2: return seg_solve<4>(c, graph);
    default: return seg_solve<8>(c, graph);
  }
}

// The tile trials: each set solves the same system from the same P, timed
// with events on the context stream; the fastest becomes the current set.
int seg_tune_tiles(efdc_ctx *c) {
  int s;
  if (!c->tune_psave && (s = dev_alloc(c, (void **)&c->tune_psave, sizeof(double) * c->LCM)))
    return s;
  const size_t bytes = sizeof(double) * c->LCM;
  HIP_TRY(c, hipMemcpyAsync(c->tune_psave, c->dH[EFDC_P], bytes, hipMemcpyDeviceToDevice,
                            c->stream));
  hipEvent_t e0, e1;
  HIP_TRY(c, hipEventCreate(&e0));
  HIP_TRY(c, hipEventCreate(&e1));
  float best = 1e30f;
  int bq = c->seg_tset;
  const bool timing = c->timing;
  c->timing = false;
  for (int q = 0; q < 4; ++q) {
    HIP_TRY(c, hipMemcpyAsync(c->dH[EFDC_P], c->tune_psave, bytes, hipMemcpyDeviceToDevice,
                              c->stream));
    seg_select_tiles(c, q);
    HIP_TRY(c, hipEventRecord(e0, c->stream));
    if ((s = seg_solve_set(c, q, false))) break;
    HIP_TRY(c, hipEventRecord(e1, c->stream));
    HIP_TRY(c, hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_TRY(c, hipEventElapsedTime(&ms, e0, e1));
    if (c->tune_n < 8) {
      std::snprintf(c->tune_name[c->tune_n], 32, "seg_cg_tile_%d", EFDC_BLOCK << q);
      c->tune_ms[c->tune_n] = ms;
      c->tune_chosen[c->tune_n] = 0;
      ++c->tune_n;
    }
    if (ms < best) {
      best = ms;
      bq = q;
    }
  }
  c->timing = timing;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (s) return s;
  for (int j = 0; j < c->tune_n; ++j)
    if (std::strncmp(c->tune_name[j], "seg_cg_tile_", 12) == 0)
      c->tune_chosen[j] = std::atoi(c->tune_name[j] + 12) == (EFDC_BLOCK << bq) ? 1 : 0;
  seg_select_tiles(c, bq);
  HIP_TRY(c, hipMemcpyAsync(c->dH[EFDC_P], c->tune_psave, bytes, hipMemcpyDeviceToDevice,
                            c->stream));
  c->seg_tune_pending = false;
  return EFDC_OK;
}

}  // namespace

namespace {

// The FP32-refined solve of every basin (tiles of 1024 cells, the
// single-model mixed kernels' per-cell work).
int seg_mixed_run(efdc_ctx *c, efdc_cg_stats_t *stats) {
  if (!c->dis) {
    c->err = "the FP32-refined solve needs a context created with EFDC_CG_MIXED";
    return EFDC_ERR_STATE;
  }
  constexpr int Q = 2, CPT = 4;  // 1024-cell tiles
  SMxArgs x;
  std::memset(&x, 0, sizeof x);
  x.m.g = cg_args(c);
  x.m.dis = c->dis;
  x.m.as = c->as32;
  x.m.aw = c->aw32;
  x.m.r32 = c->r32;
  x.m.y32 = c->y32;
  x.m.ap32 = c->ap32;
  x.g = seg_args(c);
  x.g.tiles = c->d_tiles_s[Q];
  x.g.segtile0 = c->d_segtile0_s[Q];
  x.g.ntiles = c->ntiles_s[Q];
  x.segfin = c->d_segfin;
  const int nt = c->ntiles_s[Q], cap = c->prm.ITERM;
  const double rtol = c->prm.mixed_inner_rtol > 0 ? c->prm.mixed_inner_rtol : 1e-5;
  k_fill_u32x2<<<1, 64, 0, c->stream>>>(c->d_segfin, c->d_segdone);
  TIMED(c, TK_MX_PREP, k_smx_prep<CPT><<<nt, EFDC_BLOCK, 0, c->stream>>>(x, c->prm.RSQM,
                                                                        c->prm.cg_rel_tol));
  unsigned fin = 0, done = 0;
  for (;;) {
    const unsigned fin_before = fin;
    k_fill_u32x2<<<1, 64, 0, c->stream>>>(c->d_segdone, nullptr);
    TIMED(c, TK_MX_RESID, k_smx_resid<CPT><<<nt, EFDC_BLOCK, 0, c->stream>>>(
        x, c->prm.mixed_max_outer));
    HIP_TRY(c, hipMemcpyAsync(&fin, c->d_segfin, sizeof(unsigned), hipMemcpyDeviceToHost,
                              c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if ((int)fin >= c->nseg) break;
    const unsigned target = (unsigned)c->nseg - fin_before;
    int it = 0, chunk = c->last_iters > 0 ? c->last_iters : 8;
    for (;;) {
      for (int j = 0; j < chunk && it < cap; ++j, ++it) {
        TIMED(c, TK_MX_A, k_smx_a<CPT><<<nt, EFDC_BLOCK, 0, c->stream>>>(
            x, c->p32[it & 1], c->p32[(it + 1) & 1]));
        TIMED(c, TK_MX_B, k_smx_b<CPT><<<nt, EFDC_BLOCK, 0, c->stream>>>(
            x, c->p32[(it + 1) & 1], rtol * rtol, cap));
      }
      HIP_TRY(c, hipMemcpyAsync(&done, c->d_segdone, sizeof(unsigned), hipMemcpyDeviceToHost,
                                c->stream));
      HIP_TRY(c, hipStreamSynchronize(c->stream));
      if (done >= target || it >= cap) break;
      chunk = 4;
    }
    c->last_iters = it;
    TIMED(c, TK_MX_APPLY, k_smx_apply<CPT><<<nt, EFDC_BLOCK, 0, c->stream>>>(x));
  }
  HIP_TRY(c, hipGetLastError());
  HIP_TRY(c, hipMemcpyAsync(c->h_segst, c->d_segst, sizeof(CgState) * c->nseg,
                            hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  int imax = 0, omax = 0, bad = 0;
  double rsq = 0.0;
  for (int q = 0; q < c->nseg; ++q) {
    imax = std::max(imax, c->h_segst[q].iter);
    omax = std::max(omax, c->h_segst[q].outer);
    rsq = std::max(rsq, c->h_segst[q].rsq64);
    if (c->h_segst[q].status != 0) bad = 1;
  }
  if (c->timing) c->cg_iter_count += imax;
  if (stats) {
    stats->iters = imax;
    stats->outer = omax;
    stats->rsq = rsq;
    stats->rsqm_used = c->h_segst[0].rsqm;
    stats->ccmnm = 0.0;
  }
  if (bad) {
    c->err = "CONGRAD: the FP32-refined solve of a basin did not converge within the outer "
             "rounds";
    return EFDC_ERR_CG_MAXITER;
  }
  return EFDC_OK;
}

}  // namespace

int seg_cg_run(efdc_ctx *c, efdc_cg_stats_t *stats, bool *ran_mixed) {
  int s;
  if (ran_mixed) *ran_mixed = false;
  if (c->prm.cg_mode == EFDC_CG_MIXED) {
    // the compact symmetric operator, as for one model (else FP64)
    if (c->symok_cache < 0) {
      if ((s = read_state(c))) return s;
      c->symok_cache = c->st_host->symok;
    }
    if (c->rowmajor && c->symok_cache == 0) {
      if (ran_mixed) *ran_mixed = true;
      return seg_mixed_run(c, stats);
    }
  }
  if (c->seg_tune_pending && (s = seg_tune_tiles(c))) return s;
  if (c->cg_graph_tune && !c->timing &&
      (s = cg_graph_trial(c, [&](bool g) { return seg_solve_set(c, c->seg_tset, g); })))
    return s;
  if ((s = seg_solve_set(c, c->seg_tset, c->cg_graph && !c->timing))) return s;
  HIP_TRY(c, hipMemcpyAsync(c->h_segst, c->d_segst, sizeof(CgState) * c->nseg,
                            hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  int imax = 0, bad = 0;
  double rsq = 0.0;
  for (int q = 0; q < c->nseg; ++q) {
    imax = std::max(imax, c->h_segst[q].iter);
    rsq = std::max(rsq, c->h_segst[q].rsq);
    if (c->h_segst[q].status != 0) bad = 1;
  }
  c->last_iters = imax;
  if (c->timing) c->cg_iter_count += imax;
  if (stats) {
    stats->iters = imax;
    stats->outer = 1;
    stats->rsq = rsq;
    stats->rsqm_used = c->h_segst[0].rsqm;
    stats->ccmnm = 0.0;
  }
  if (bad) {
    c->err = "CONGRAD: maximum iterations exceeded in external solution (a basin)";
    return EFDC_ERR_CG_MAXITER;
  }
  return EFDC_OK;
}

// ---------------------------------------------------------------------
// CONGRAD without device allocation (efdc_congrad_noalloc): the FP64
// multi-kernel schedule above (k_cg_init, k_cg_a, k_cg_b) on the caller's
// device arrays, with every scratch vector, partial sum, arrival counter
// and the solve state in one static device array per GPU.
// Layout (bytes): CgState at 0, arrival counters at NA_OFF_TK, partials at
// NA_OFF_PART, then r, Ap, p (two buffers) as FP64 vectors of
// EFDC_NOALLOC_MAX_LCM entries from NA_OFF_VEC, then the 0-based
// LWC/LEC/LSC/LNC as int32 vectors of the same length.
// ---------------------------------------------------------------------
namespace {
constexpr size_t na_up(size_t b) { return (b + 255) & ~size_t(255); }
constexpr size_t NA_MAXC = EFDC_NOALLOC_MAX_LCM;
constexpr int NA_MAXB = (int)((NA_MAXC + CG_CPT * EFDC_BLOCK - 1) / (CG_CPT * EFDC_BLOCK));
constexpr size_t NA_OFF_TK = na_up(sizeof(CgState));
constexpr size_t NA_TK_BYTES = sizeof(unsigned) * TICKET_SITES * TICKET_SITE;
constexpr size_t NA_OFF_PART = NA_OFF_TK + na_up(NA_TK_BYTES);
constexpr size_t NA_OFF_VEC = 65536;
constexpr size_t NA_OFF_MAP = NA_OFF_VEC + 4 * NA_MAXC * sizeof(double);
constexpr size_t NA_BYTES = NA_OFF_MAP + 4 * NA_MAXC * sizeof(int);
static_assert(NA_OFF_PART + na_up(sizeof(double) * (2 * NA_MAXB + 16)) <= NA_OFF_VEC,
              "no-alloc header overflows its 64 KiB");
static_assert(NA_BYTES == EFDC_NOALLOC_SCRATCH_BYTES, "scratch size differs from the header");
}  // namespace

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

namespace {
// Device address of `scratch` on each GPU, looked up once per device (each
// GPU holds its own copy of the array at its own address).
std::mutex na_mu;
void *na_ptr[64] = {};

// Zero the solve state, the arrival counters and the first p buffer; turn
// the 1-based maps into 0-based indices; any wet row whose neighbour lies
// outside 2..LC raises st->status = 2 and is pointed at the ghost LC so
// that nothing reads out of bounds.
__global__ __launch_bounds__(EFDC_BLOCK) void k_na_prepare(
    unsigned char *base, const int *__restrict__ lwc, const int *__restrict__ lec,
    const int *__restrict__ lsc, const int *__restrict__ lnc, int la) {
  const int i = blockIdx.x * EFDC_BLOCK + (int)threadIdx.x;  // 0 .. la (L-1)
  if (blockIdx.x == 0) {
    unsigned *w = reinterpret_cast<unsigned *>(base);
    for (int j = threadIdx.x; j < (int)(NA_OFF_PART / sizeof(unsigned)); j += EFDC_BLOCK)
      w[j] = 0u;
  }
  if (i > la) return;
  int *m = reinterpret_cast<int *>(base + NA_OFF_MAP);
  double *p0 = reinterpret_cast<double *>(base + NA_OFF_VEC) + 2 * NA_MAXC;
  p0[i] = 0.0;
  const int *src[4] = {lwc, lec, lsc, lnc};
  const bool wet = (i >= 1 && i < la);
  bool bad = false;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    int v = src[q][i] - 1;
    if (v < 1 || v > la) {
      if (wet) bad = true;
      v = la;
    }
    m[q * NA_MAXC + i] = v;
  }
  if (bad) {
    CgState *st = reinterpret_cast<CgState *>(base);
    __hip_atomic_store(&st->status, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}
}  // namespace

int efdc_congrad_noalloc(int32_t LA, const int32_t *LWC, const int32_t *LEC,
                         const int32_t *LSC, const int32_t *LNC, const double *CCC,
                         const double *CCS, const double *CCW, const double *CCE,
                         const double *CCN, const double *CCCI, const double *FPTMP,
                         double *P, double RSQM, int32_t ITERM, void *stream,
                         efdc_cg_stats_t *stats) {
  auto fail = [](int st, const std::string &why) {
    efdc_thread_error(why.c_str());
    return st;
  };
  if (LA < 2 || (size_t)LA + 1 > NA_MAXC)
    return fail(EFDC_ERR_ARG, "efdc_congrad_noalloc: LA+1 must lie in 3.." +
                                  std::to_string(NA_MAXC));
  if (!LWC || !LEC || !LSC || !LNC || !CCC || !CCS || !CCW || !CCE || !CCN || !CCCI ||
      !FPTMP || !P)
    return fail(EFDC_ERR_ARG, "efdc_congrad_noalloc: null array");
  if (ITERM < 1 || !(RSQM >= 0.0))
    return fail(EFDC_ERR_ARG, "efdc_congrad_noalloc: ITERM < 1 or RSQM not >= 0");
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(EFDC_ERR_NODEVICE, hipGetErrorString(e));
  if (dev < 0 || dev >= 64) return fail(EFDC_ERR_ARG, "efdc_congrad_noalloc: device id >= 64");
  unsigned char *base;
  {
    std::lock_guard<std::mutex> g(na_mu);
    if (!na_ptr[dev]) {
      void *p = nullptr;
      e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p)
        return fail(EFDC_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                      hipGetErrorString(e));
      na_ptr[dev] = p;
    }
    base = static_cast<unsigned char *>(na_ptr[dev]);
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  CgState *st = reinterpret_cast<CgState *>(base);
  double *vec = reinterpret_cast<double *>(base + NA_OFF_VEC);
  int *map = reinterpret_cast<int *>(base + NA_OFF_MAP);
  CgArgs a;
  a.ccc = CCC;
  a.ccs = CCS;
  a.ccw = CCW;
  a.ccn = CCN;
  a.cce = CCE;
  a.ccci = CCCI;
  a.fptmp = FPTMP;
  a.x = P;
  a.r = vec;
  a.ap = vec + NA_MAXC;
  double *pb[2] = {vec + 2 * NA_MAXC, vec + 3 * NA_MAXC};
  a.lw = map;
  a.le = map + NA_MAXC;
  a.ls = map + 2 * NA_MAXC;
  a.ln = map + 3 * NA_MAXC;
  a.nwet = LA - 1;
  a.lc0 = LA;
  a.rowmajor = 0;  // the general four-map operator: no layout is assumed
  a.st = st;
  a.part = reinterpret_cast<double *>(base + NA_OFF_PART);
  a.tk = reinterpret_cast<unsigned *>(base + NA_OFF_TK);
  CgState h;
  auto sync_state = [&]() -> hipError_t {
    hipError_t q = hipMemcpyAsync(&h, st, sizeof(CgState), hipMemcpyDeviceToHost, s);
    if (q != hipSuccess) return q;
    return hipStreamSynchronize(s);
  };
  k_na_prepare<<<(LA + EFDC_BLOCK) / EFDC_BLOCK, EFDC_BLOCK, 0, s>>>(base, LWC, LEC, LSC, LNC,
                                                                    LA);
  if ((e = hipGetLastError()) != hipSuccess || (e = sync_state()) != hipSuccess)
    return fail(EFDC_ERR_HIP, hipGetErrorString(e));
  if (h.status == 2)
    return fail(EFDC_ERR_ARG, "efdc_congrad_noalloc: a wet row's neighbour map lies outside 2..LA+1");
  const int nb = cg_blocks(LA - 1);
  k_cg_init<<<nb, EFDC_BLOCK, 0, s>>>(a, RSQM, 0.0);
  int it = 0, chunk = 16;
  for (;;) {
    for (int j = 0; j < chunk && it < ITERM; ++j, ++it) {
      k_cg_a<<<nb, EFDC_BLOCK, 0, s>>>(a, pb[it & 1], pb[(it + 1) & 1]);
      k_cg_b<<<nb, EFDC_BLOCK, 0, s>>>(a, pb[(it + 1) & 1], ITERM);
    }
    if ((e = hipGetLastError()) != hipSuccess || (e = sync_state()) != hipSuccess)
      return fail(EFDC_ERR_HIP, hipGetErrorString(e));
    if (h.done || it >= ITERM) break;
    chunk = 8;
  }
  if (stats) {
    stats->iters = h.iter;
    stats->outer = 1;
    stats->rsq = h.rsq;
    stats->rsqm_used = h.rsqm;
    stats->ccmnm = 0.0;
  }
  if (h.status != 0)
    return fail(EFDC_ERR_CG_MAXITER,
                "CONGRAD: maximum iterations exceeded in external solution");
  return EFDC_OK;
}
