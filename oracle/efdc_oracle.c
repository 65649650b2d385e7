/*
 * efdc_oracle.c -- CPU restatement of the EFDC+ 8.5 external-mode solve and
 * constituent transport, FP64, plain C + OpenMP.
 *
 * THIS IS TEST INFRASTRUCTURE.  It is the checker of the HIP library in
 * tests/, the check of __graft_entry__.smoke() and the CPU baseline leg of
 * bench.py.  The product never links or calls it.
 *
 * Each function follows the reference subroutine it names; statement
 * order and the left-to-right evaluation order of the reference's
 * expressions are kept so that the restatement can be compared bit for bit
 * with the reference units that build here (oracle/_ref) and with the FP64
 * HIP kernels that do not reduce.  The OpenMP structure follows the
 * reference: CALPUV2C/CONGRAD loops over at most 8 static L-domains
 * (EFDC/calpuv2c.f90:61-62) with serial dot products
 * (EFDC/congrad.f90:118-125,139-149), CALTRAN one constituent per thread
 * (EFDC/calconc.f90:159-229), vertical diffusion column-parallel
 * (EFDC/calconc.f90:259).
 */
#include "efdc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define A(id) (o->a[id])
#define HX(L) ((size_t)(L) - 1)
#define KX(L, K) ((size_t)(L) - 1 + (size_t)((K) - 1) * (size_t)o->LCM)
#define WX(L, K) ((size_t)(L) - 1 + (size_t)(K) * (size_t)o->LCM)

static const double PI_EFDC = 3.1415926535898; /* EFDC/aaefdc.f90:650-651 */

/* Planted mistakes.  Built with -DORC_PLANT=n (tests/test_oracle_plants.py
   compiles such variants into a temporary directory; the checker that the
   tests use is always built with ORC_PLANT = 0), the oracle makes mistake n,
   so that a test can show that one of its checks fails on it:
     1  coefficient constant -G*DELT/2 instead of -G*DELT/4
        (EFDC/calpuv2c.f90:329)
     2  FP1 with G instead of G/2 (:289)
     3  FP with G instead of G/2 (:320)
     4  face transports with DELT instead of DELT/2 (:426)
     5  HU averaged with the east neighbour instead of the west (:905)
     6  HV without the factor 1/2 (:906)
     7  bottom stress added instead of subtracted (:217)
     8  QSUME counted once in FP instead of twice (:320)
     9  continuity with DELT instead of DELT/2 (:475)
   and in CALUVW (EFDC/caluvw.f90):
    20  layer flows not corrected to the external solution (:592-600)
    21  Sherman-Morrison correction for the bottom drag left out (:491-520)
    22  W without the external-flow split UHDYEK (:676-679)
    23  W without the factor 1/2 of the two-level flow average (:676)
    24  depth from the layer flows with the east sum added (:953)
   and in the two-level balances (EFDC/bal2t.f90):
    30  west open-boundary volume flux with the wrong sign (:224)
    31  source mass without the withdrawal term MIN(QSS,0)*C (:374)
    32  volume totals without the SPB mask of the open-boundary cells (:86) */
#ifndef ORC_PLANT
#define ORC_PLANT 0
#endif
#define PLANT(n, mistake, right) (ORC_PLANT == (n) ? (mistake) : (right))

static inline double dmax(double a, double b) { return a > b ? a : b; }
static inline double dmin(double a, double b) { return a < b ? a : b; }

/* number of CALPUV threads: NOPTIMAL = MIN(NTHREADS,8,MAX(LA/100,1))
   (EFDC/calpuv2c.f90:61) */
static int noptimal(const orc_ctx *o) {
  int nt = o->nthreads > 0 ? o->nthreads : 1;
  int m = o->LA / 100;
  if (m < 1) m = 1;
  if (nt > 8) nt = 8;
  return nt < m ? nt : m;
}

/* ===================================================================== */
/* CONGRAD: EFDC/congrad.f90:54-187                                        */
/* ===================================================================== */
int orc_congrad(orc_ctx *o, orc_cg_out *out) {
  const int LA = o->LA, LCM = o->LCM;
  double *P = A(EFDC_P), *CCC = A(EFDC_CCC), *CCS = A(EFDC_CCS),
         *CCW = A(EFDC_CCW), *CCE = A(EFDC_CCE), *CCN = A(EFDC_CCN),
         *CCCI = A(EFDC_CCCI), *FPTMP = A(EFDC_FPTMP);
  const int32_t *LNC = o->LNC, *LSC = o->LSC, *LEC = o->LEC, *LWC = o->LWC;
  double *PCG = calloc(LCM, sizeof(double)), *PN = calloc(LCM, sizeof(double)),
         *PE = calloc(LCM, sizeof(double)), *PW = calloc(LCM, sizeof(double)),
         *PS = calloc(LCM, sizeof(double)), *RCG = calloc(LCM, sizeof(double)),
         *TMPCG = calloc(LCM, sizeof(double)),
         *APCG = calloc(LCM, sizeof(double));
  const int nopt = noptimal(o);
  int iter = 0;
  double rsq = 0.0;
  double rsqm = o->RSQM;
  out->status = EFDC_OK;
  if (o->CG_REL_TOL > 0.0) { /* relative stop (the GPU contexts' cg_rel_tol) */
    double b2 = 0.0;
    for (int L = 2; L <= LA; ++L) b2 = b2 + FPTMP[HX(L)] * FPTMP[HX(L)];
    rsqm = o->CG_REL_TOL * o->CG_REL_TOL * b2;
  }

  /* initial residual, :54-77 */
#pragma omp parallel for num_threads(nopt) schedule(static)
  for (int L = 2; L <= LA; ++L) {
    PN[HX(L)] = P[HX(LNC[HX(L)])];
    PE[HX(L)] = P[HX(LEC[HX(L)])];
    PW[HX(L)] = P[HX(LWC[HX(L)])];
    PS[HX(L)] = P[HX(LSC[HX(L)])];
  }
#pragma omp parallel for num_threads(nopt) schedule(static)
  for (int L = 2; L <= LA; ++L) {
    const size_t i = HX(L);
    RCG[i] = FPTMP[i] - CCC[i] * P[i] - CCN[i] * PN[i] - CCS[i] * PS[i] -
             CCW[i] * PW[i] - CCE[i] * PE[i];
  }
  /* :79-86 */
  for (int L = 2; L <= LA; ++L) PCG[HX(L)] = RCG[HX(L)] * CCCI[HX(L)];
  double rpcg = 0.0;
  for (int L = 2; L <= LA; ++L) rpcg = rpcg + RCG[HX(L)] * PCG[HX(L)];
  if (rpcg == 0.0) { /* :88 */
    out->iters = 0;
    out->rsq = 0.0;
    goto done;
  }
  for (iter = 1; iter <= o->ITERM; ++iter) { /* :91-169 */
#pragma omp parallel for num_threads(nopt) schedule(static)
    for (int L = 2; L <= LA; ++L) {
      const size_t i = HX(L);
      PN[i] = PCG[HX(LNC[i])];
      PE[i] = PCG[HX(LEC[i])];
      PW[i] = PCG[HX(LWC[i])];
      PS[i] = PCG[HX(LSC[i])];
    }
#pragma omp parallel for num_threads(nopt) schedule(static)
    for (int L = 2; L <= LA; ++L) {
      const size_t i = HX(L);
      APCG[i] = CCC[i] * PCG[i] + CCS[i] * PS[i] + CCN[i] * PN[i] +
                CCW[i] * PW[i] + CCE[i] * PE[i];
    }
    double papcg = 0.0; /* :119-124, serial */
    for (int L = 2; L <= LA; ++L) papcg = papcg + APCG[HX(L)] * PCG[HX(L)];
    const double alpha = rpcg / papcg;
#pragma omp parallel for num_threads(nopt) schedule(static)
    for (int L = 2; L <= LA; ++L) { /* :127-137 */
      const size_t i = HX(L);
      P[i] = P[i] + alpha * PCG[i];
      RCG[i] = RCG[i] - alpha * APCG[i];
      TMPCG[i] = CCCI[i] * RCG[i];
    }
    double rpcgn = 0.0; /* :140-149, serial */
    rsq = 0.0;
    for (int L = 2; L <= LA; ++L) {
      rpcgn = rpcgn + RCG[HX(L)] * TMPCG[HX(L)];
      rsq = rsq + RCG[HX(L)] * RCG[HX(L)];
    }
    const double beta = rpcgn / rpcg;
    rpcg = rpcgn;
    if (rsq > rsqm) { /* :151-162 */
#pragma omp parallel for num_threads(nopt) schedule(static)
      for (int L = 2; L <= LA; ++L)
        PCG[HX(L)] = TMPCG[HX(L)] + beta * PCG[HX(L)];
    }
    if (rsq <= rsqm) break; /* :165-167 */
  }
  out->iters = iter; /* IDRYTBP = ITER, :187 */
  out->rsq = rsq;
  if (rsq > rsqm) out->status = EFDC_ERR_CG_MAXITER; /* :171-184 */
done:
  free(PCG); free(PN); free(PE); free(PW); free(PS); free(RCG); free(TMPCG);
  free(APCG);
  return out->status;
}

/* ===================================================================== */
/* SETOPENBC: EFDC/setopenbc.f90:1-209 (HUT=HU, HVT=HV)                    */
/* ===================================================================== */
static void orc_setopenbc(orc_ctx *o, double dt, int64_t n, double D_LTD2,
                          double DE_TI) {
  const double G = o->G, HDRY2 = 0.2 * o->HDRY, DELTI = DE_TI;
  double *CC = A(EFDC_CC), *CS = A(EFDC_CS), *CW = A(EFDC_CW),
         *CE = A(EFDC_CE), *CN = A(EFDC_CN), *FP = A(EFDC_FP),
         *DXYP = A(EFDC_DXYP), *BELV = A(EFDC_BELV), *HRUO = A(EFDC_HRUO),
         *HRVO = A(EFDC_HRVO), *RCX = A(EFDC_RCX), *RCY = A(EFDC_RCY),
         *HUT = A(EFDC_HU), *HVT = A(EFDC_HV), *DXIU = A(EFDC_DXIU),
         *DYIU = A(EFDC_DYIU), *DXIV = A(EFDC_DXIV), *DYIV = A(EFDC_DYIV),
         *FUHDYE = A(EFDC_FUHDYE), *FVHDXE = A(EFDC_FVHDXE);
  double cccos[64], sssin[64];
  const double PI2 = 2. * PI_EFDC;
  const double TN = dt * (double)n + o->TCON_TBEGIN; /* :22-26 */
  for (int M = 0; M < o->mtide && M < 64; ++M) {     /* :27-32 */
    double TM = fmod(TN, o->tcp[M]);
    TM = PI2 * TM / o->tcp[M];
    cccos[M] = cos(TM);
    sssin[M] = sin(TM);
  }
  for (int side = 0; side < 4; ++side) {
    const orc_pbc_side *s = &o->pbc[side];
    for (int LL = 0; LL < s->n; ++LL) {
      const int L = s->L[LL];
      const size_t i = HX(L);
      CC[i] = DELTI * DXYP[i];
      CS[i] = 0.; CW[i] = 0.; CE[i] = 0.; CN[i] = 0.;
      if (o->lopenbcdry[i]) continue;
      FP[i] = s->pser[LL];
      for (int M = 0; M < o->mtide; ++M)
        FP[i] = FP[i] + s->pcb[LL * o->mtide + M] * cccos[M] +
                s->psb[LL * o->mtide + M] * sssin[M];
      const int rad = (s->ispb[LL] == 1 || s->ispb[LL] == 2);
      if (side == 0) { /* WEST :55-74 */
        const size_t e = HX(o->LEC[i]);
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
            o->lopenbcdry[i] = 1;
          } else {
            CET = 0.5 * D_LTD2 * G * HRUO[e] * RCX[e] * HUT[e];
          }
          FP[e] = FP[e] + CET * FP[i];
          FP[i] = DE_TI * DXYP[i] * FP[i];
        }
      } else if (side == 1) { /* EAST :98-117 */
        const size_t w = HX(o->LWC[i]);
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
            o->lopenbcdry[i] = 1;
          } else {
            CWT = 0.5 * D_LTD2 * G * HRUO[i] * RCX[i] * HUT[i];
          }
          FP[w] = FP[w] + CWT * FP[i];
          FP[i] = DE_TI * DXYP[i] * FP[i];
        }
      } else if (side == 2) { /* SOUTH :142-161 */
        const size_t nn = HX(o->LNC[i]);
        if (rad) {
          const double CNT = 0.5 * D_LTD2 * G * HRVO[nn] * RCY[nn] * HVT[nn];
          const double TMP = D_LTD2 * sqrt(G * HVT[nn]) * DYIV[nn];
          CC[i] = CNT * (1. + TMP) / TMP;
          CN[i] = -CNT;
          FP[i] = CNT * (2. * FP[i] - sqrt(G * HVT[nn]) * FVHDXE[nn] * DXIV[nn] / HVT[nn]) / TMP;
        } else {
          double CNT;
          const double FP1G = FP[i] / G - HDRY2;
          if (FP1G < BELV[i] || FP1G < BELV[nn]) {
            FP[i] = (BELV[i] + HDRY2) * G;
            CNT = 0.;
            o->lopenbcdry[i] = 1;
          } else {
            CNT = 0.5 * D_LTD2 * G * HRVO[nn] * RCY[nn] * HVT[nn];
          }
          FP[nn] = FP[nn] + CNT * FP[i];
          FP[i] = DE_TI * DXYP[i] * FP[i];
        }
      } else { /* NORTH :186-205 */
        const size_t ss = HX(o->LSC[i]);
        if (rad) {
          const double CST = 0.5 * D_LTD2 * G * HRVO[i] * RCY[i] * HVT[i];
          const double TMP = D_LTD2 * sqrt(G * HVT[i]) * DYIV[i];
          CC[i] = CST * (1. + TMP) / TMP;
          CS[i] = -CST;
          FP[i] = CST * (2. * FP[i] + sqrt(G * HVT[i]) * FVHDXE[i] * DXIV[i] / HVT[i]) / TMP;
        } else {
          double CST;
          const double FP1G = FP[i] / G - HDRY2;
          if (FP1G < BELV[i] || FP1G < BELV[ss]) {
            FP[i] = (BELV[i] + HDRY2) * G;
            CST = 0.;
            o->lopenbcdry[i] = 1;
          } else {
            CST = 0.5 * D_LTD2 * G * HRVO[i] * RCY[i] * HVT[i];
          }
          FP[ss] = FP[ss] + CST * FP[i];
          FP[i] = DE_TI * DXYP[i] * FP[i];
        }
      }
    }
  }
}

/* SETOPENBC2: EFDC/setopenbc2.f90:17-45 */
static void orc_setopenbc2(orc_ctx *o) {
  for (int LL = 0; LL < o->pbc[0].n; ++LL)
    if (o->pbc[0].ispb[LL] == 0)
      A(EFDC_CW)[HX(o->LEC[HX(o->pbc[0].L[LL])])] = 0.;
  for (int LL = 0; LL < o->pbc[1].n; ++LL)
    if (o->pbc[1].ispb[LL] == 0)
      A(EFDC_CE)[HX(o->LWC[HX(o->pbc[1].L[LL])])] = 0.;
  for (int LL = 0; LL < o->pbc[2].n; ++LL)
    if (o->pbc[2].ispb[LL] == 0)
      A(EFDC_CS)[HX(o->LNC[HX(o->pbc[2].L[LL])])] = 0.;
  for (int LL = 0; LL < o->pbc[3].n; ++LL)
    if (o->pbc[3].ispb[LL] == 0)
      A(EFDC_CN)[HX(o->LSC[HX(o->pbc[3].L[LL])])] = 0.;
}

/* ===================================================================== */
/* CALPUV2C assembly: EFDC/calpuv2c.f90:128-136,164-398 (BSC=0, ISITB=0,  */
/* ISVEG=0, MDCHH=0, ISGWIE=0; the drying passes follow further down)      */
/* ===================================================================== */
/* The part of CALPUV2C before label 1000: the explicit stage, the
   time-level shifts and FP1 (EFDC/calpuv2c.f90:214-290). */
static int explicit_stage(orc_ctx *o, double dt, int64_t n) {
  (void)n;
  const int LA = o->LA, KC = o->KC;
  const double G = o->G;
  const double DELT = dt, DELTD2 = 0.5 * dt, DELTI = 1. / DELT; /* :128-131 */
  const int nopt = noptimal(o);
  double *UHDYE = A(EFDC_UHDYE), *VHDXE = A(EFDC_VHDXE), *P = A(EFDC_P),
         *SUB = A(EFDC_SUB), *SVB = A(EFDC_SVB), *HRUO = A(EFDC_HRUO),
         *HRVO = A(EFDC_HRVO), *HU = A(EFDC_HU), *HV = A(EFDC_HV),
         *DXIU = A(EFDC_DXIU), *DYIV = A(EFDC_DYIV), *DXYU = A(EFDC_DXYU),
         *DXYV = A(EFDC_DXYV), *TSX = A(EFDC_TSX), *TSY = A(EFDC_TSY),
         *TBX = A(EFDC_TBX), *TBY = A(EFDC_TBY), *FCAXE = A(EFDC_FCAXE),
         *FCAYE = A(EFDC_FCAYE), *FPGXE = A(EFDC_FPGXE),
         *FPGYE = A(EFDC_FPGYE), *FXE = A(EFDC_FXE), *FYE = A(EFDC_FYE),
         *FUHDYE = A(EFDC_FUHDYE), *FVHDXE = A(EFDC_FVHDXE),
         *DXYP = A(EFDC_DXYP), *FP1 = A(EFDC_FP1);
  const double RITB1 = o->RITB1, SNLT = o->SNLT;

  /* explicit stage, :214-220 */
#pragma omp parallel for num_threads(nopt) schedule(static)
  for (int L = 2; L <= LA; ++L) {
    const size_t i = HX(L), s = HX(o->LSC[i]), w = HX(o->LWC[i]);
    FUHDYE[i] = UHDYE[i] - DELTD2 * SUB[i] * HRUO[i] * HU[i] * (P[i] - P[w]) +
                SUB[i] * DELT * DXIU[i] *
                    (DXYU[i] * PLANT(7, TSX[i] + RITB1 * TBX[i], TSX[i] - RITB1 * TBX[i]) +
                     FCAXE[i] + FPGXE[i] - SNLT * FXE[i]);
    FVHDXE[i] = VHDXE[i] - DELTD2 * SVB[i] * HRVO[i] * HV[i] * (P[i] - P[s]) +
                SVB[i] * DELT * DYIV[i] *
                    (DXYV[i] * PLANT(7, TSY[i] + RITB1 * TBY[i], TSY[i] - RITB1 * TBY[i]) -
                     FCAYE[i] + FPGYE[i] - SNLT * FYE[i]);
  }
  /* implicit bottom and vegetation drag, :223-243 (RCX/RCY formed by the
     host) */
  if (o->IMPLICIT_DRAG) {
    const double *RCX = A(EFDC_RCX), *RCY = A(EFDC_RCY);
    for (int L = 2; L <= LA; ++L) {
      FUHDYE[HX(L)] = FUHDYE[HX(L)] * RCX[HX(L)];
      FVHDXE[HX(L)] = FVHDXE[HX(L)] * RCY[HX(L)];
    }
  }
  /* faces reopened with drying on, :246-253 */
  if (o->ISDRY > 0) {
    const double *SUBO = A(EFDC_SUBO), *SVBO = A(EFDC_SVBO);
    for (int L = 2; L <= LA; ++L) {
      SUB[HX(L)] = SUBO[HX(L)];
      SVB[HX(L)] = SVBO[HX(L)];
    }
  }
  /* time-level shifts, :256-274 */
  {
    double *UHDY1E = A(EFDC_UHDY1E), *VHDX1E = A(EFDC_VHDX1E),
           *P1 = A(EFDC_P1), *H1U = A(EFDC_H1U), *H1V = A(EFDC_H1V),
           *H1UI = A(EFDC_H1UI), *H1VI = A(EFDC_H1VI), *HUI = A(EFDC_HUI),
           *HVI = A(EFDC_HVI), *H2P = A(EFDC_H2P), *H1P = A(EFDC_H1P),
           *HP = A(EFDC_HP), *H1PK = A(EFDC_H1PK), *HPK = A(EFDC_HPK);
#pragma omp parallel for num_threads(nopt) schedule(static)
    for (int L = 2; L <= LA; ++L) {
      const size_t i = HX(L);
      UHDY1E[i] = UHDYE[i];
      VHDX1E[i] = VHDXE[i];
      P1[i] = P[i];
      H1U[i] = HU[i];
      H1V[i] = HV[i];
      H1UI[i] = HUI[i];
      H1VI[i] = HVI[i];
      H2P[i] = H1P[i];
      H1P[i] = HP[i];
      for (int K = 1; K <= KC; ++K) H1PK[KX(L, K)] = HPK[KX(L, K)];
    }
    /* UHDY1EK = UHDYEK, VHDX1EK = VHDXEK (:270-272), kept only when the
       internal-mode arrays are bound */
    double *UHDYEK = A(EFDC_UHDYEK), *UHDY1EK = A(EFDC_UHDY1EK),
           *VHDXEK = A(EFDC_VHDXEK), *VHDX1EK = A(EFDC_VHDX1EK);
    if (UHDYEK && UHDY1EK && VHDXEK && VHDX1EK)
      for (int K = 1; K <= KC; ++K)
        for (int L = 2; L <= LA; ++L) {
          UHDY1EK[KX(L, K)] = UHDYEK[KX(L, K)];
          VHDX1EK[KX(L, K)] = VHDXEK[KX(L, K)];
        }
  }
  /* FP1, :285-290 */
  {
    const double C1 = PLANT(2, G, 0.5 * G);
#pragma omp parallel for num_threads(nopt) schedule(static)
    for (int L = 2; L <= LA; ++L) {
      const size_t i = HX(L), e = HX(o->LEC[i]), nn = HX(o->LNC[i]);
      FP1[i] = DELTI * DXYP[i] * P[i] -
               C1 * (UHDYE[e] - UHDYE[i] + VHDXE[nn] - VHDXE[i]);
    }
  }
  return EFDC_OK;
}

/* The part of CALPUV2C after label 1000 (EFDC/calpuv2c.f90:302): FP and
   the coefficients (:315-342), SETOPENBC, the minimum CCMNM, SETOPENBC2 and
   the scaling (:349-398).  The drying loop re-enters here. */
static double system_stage(orc_ctx *o, double dt, int64_t n) {
  const int LA = o->LA;
  const double G = o->G;
  const double DELT = dt, DELTD2 = 0.5 * dt, DELTI = 1. / DELT;
  const int nopt = noptimal(o);
  double *SUB = A(EFDC_SUB), *SVB = A(EFDC_SVB), *HRUO = A(EFDC_HRUO),
         *HRVO = A(EFDC_HRVO), *HU = A(EFDC_HU), *HV = A(EFDC_HV),
         *FUHDYE = A(EFDC_FUHDYE), *FVHDXE = A(EFDC_FVHDXE),
         *DXYP = A(EFDC_DXYP), *FP1 = A(EFDC_FP1), *FP = A(EFDC_FP),
         *QSUME = A(EFDC_QSUME), *RCX = A(EFDC_RCX), *RCY = A(EFDC_RCY),
         *CC = A(EFDC_CC), *CS = A(EFDC_CS), *CW = A(EFDC_CW),
         *CE = A(EFDC_CE), *CN = A(EFDC_CN);
  /* FP, CS, CW, CE, CN, CC: :315-342 */
#pragma omp parallel for num_threads(nopt) schedule(static)
  for (int L = 2; L <= LA; ++L) {
    const size_t i = HX(L), e = HX(o->LEC[i]), nn = HX(o->LNC[i]);
    const double C1 = PLANT(3, G, 0.5 * G);
    FP[i] = FP1[i] - C1 * (SUB[e] * FUHDYE[e] - SUB[i] * FUHDYE[i] +
                           SVB[nn] * FVHDXE[nn] - SVB[i] * FVHDXE[i] -
                           PLANT(8, 1.0, 2.0) * QSUME[i]);
  }
#pragma omp parallel for num_threads(nopt) schedule(static)
  for (int L = 2; L <= LA; ++L) {
    const size_t i = HX(L), e = HX(o->LEC[i]), nn = HX(o->LNC[i]);
    const double C1 = PLANT(1, -DELTD2 * G, -0.5 * DELTD2 * G);
    CS[i] = C1 * SVB[i] * HRVO[i] * RCY[i] * HV[i];
    CW[i] = C1 * SUB[i] * HRUO[i] * RCX[i] * HU[i];
    CE[i] = C1 * SUB[e] * HRUO[e] * RCX[e] * HU[e];
    CN[i] = C1 * SVB[nn] * HRVO[nn] * RCY[nn] * HV[nn];
  }
#pragma omp parallel for num_threads(nopt) schedule(static)
  for (int L = 2; L <= LA; ++L) {
    const size_t i = HX(L);
    CC[i] = DELTI * DXYP[i] - CS[i] - CW[i] - CE[i] - CN[i];
  }
  /* open boundaries, :349 */
  orc_setopenbc(o, dt, n, DELTD2, DELTI);
  /* min diagonal, :360-376 */
  double ccmnm = 1.E+18;
  double *FPTMP = A(EFDC_FPTMP);
  for (int L = 2; L <= LA; ++L) {
    ccmnm = dmin(ccmnm, CC[HX(L)]);
    FPTMP[HX(L)] = FP[HX(L)];
  }
  const double CCMNMI = 1. / ccmnm;
  orc_setopenbc2(o); /* :379 */
  /* scaling, :384-398 */
  {
    double *CCS = A(EFDC_CCS), *CCW = A(EFDC_CCW), *CCE = A(EFDC_CCE),
           *CCN = A(EFDC_CCN), *CCC = A(EFDC_CCC), *CCCI = A(EFDC_CCCI);
#pragma omp parallel for num_threads(nopt) schedule(static)
    for (int L = 2; L <= LA; ++L) {
      const size_t i = HX(L);
      CCS[i] = CS[i] * CCMNMI;
      CCW[i] = CW[i] * CCMNMI;
      CCE[i] = CE[i] * CCMNMI;
      CCN[i] = CN[i] * CCMNMI;
      CCC[i] = CC[i] * CCMNMI;
      FPTMP[i] = FPTMP[i] * CCMNMI;
      CCCI[i] = 1. / CCC[i];
    }
  }
  return ccmnm;
}

int orc_assemble(orc_ctx *o, double dt, int64_t n, double *ccmnm_out) {
  explicit_stage(o, dt, n);
  const double ccmnm = system_stage(o, dt, n);
  if (ccmnm_out) *ccmnm_out = ccmnm;
  return EFDC_OK;
}

/* ===================================================================== */
/* CALPUV2C: assembly + CONGRAD + post-solve updates                       */
/* ===================================================================== */
/* Open-boundary cells that dry, ISDRY > 0: EFDC/calpuv2c.f90:497-530 (the
   depth is reset, the cell and its faces are closed for the rest of the
   step).  Run after the open-boundary depths are set. */
static void open_bc_drying(orc_ctx *o, double dt) {
  const double G = o->G, HDRY = o->HDRY, DELTI = 1. / dt;
  double *HP = A(EFDC_HP), *H1P = A(EFDC_H1P), *P = A(EFDC_P), *BELV = A(EFDC_BELV),
         *SUB = A(EFDC_SUB), *SVB = A(EFDC_SVB), *UHDYE = A(EFDC_UHDYE),
         *VHDXE = A(EFDC_VHDXE), *CC = A(EFDC_CC), *FP1 = A(EFDC_FP1),
         *DXYP = A(EFDC_DXYP);
  static const int order[4] = {2, 0, 1, 3};
  for (int q = 0; q < 4; ++q) {
    const orc_pbc_side *sd = &o->pbc[order[q]];
    for (int LL = 0; LL < sd->n; ++LL) {
      const size_t i = HX(sd->L[LL]);
      const int dry = o->lopenbcdry[i] != 0;
      if (!((dry || (H1P[i] < 0.9 * HDRY && HP[i] < H1P[i]) || HP[i] < 0.) &&
            o->iscdry[i] == 0))
        continue;
      if (HP[i] <= 0. && dry) HP[i] = 0.2 * HDRY;
      else HP[i] = dmin(dmax(H1P[i], 0.1 * HDRY), 0.9 * HDRY);
      o->iscdry[i] = 1;
      const size_t e = HX(o->LEC[i]), nn = HX(o->LNC[i]);
      SUB[i] = 0.;
      SVB[i] = 0.;
      SUB[e] = 0.;
      SVB[nn] = 0.;
      UHDYE[i] = 0.;
      UHDYE[e] = 0.;
      VHDXE[i] = 0.;
      VHDXE[nn] = 0.;
      o->lopenbcdry[i] = 1;
      CC[i] = DELTI * DXYP[i];
      P[i] = (HP[i] + BELV[i]) * G;
      FP1[i] = DELTI * DXYP[i] * P[i];
    }
  }
}

/* Cells that dry, 0 < ISDRY < 98: EFDC/calpuv2c.f90:574-602.  Returns 1
   when a cell became dry in this pass (ICORDRY = 1: solve again). */
static int cell_drying(orc_ctx *o) {
  const double HDRY = o->HDRY;
  double *HP = A(EFDC_HP), *SUB = A(EFDC_SUB), *SVB = A(EFDC_SVB);
  int icordry = 0;
  for (int L = 2; L <= o->LA; ++L) {
    const size_t i = HX(L);
    if (HP[i] <= HDRY) {
      if (o->iscdry[i] == 0) {
        o->iscdry[i] = 1;
        icordry = 1;
      }
      const size_t e = HX(o->LEC[i]), nn = HX(o->LNC[i]);
      SUB[i] = 0.;
      SVB[i] = 0.;
      SUB[e] = 0.;
      SVB[nn] = 0.;
    }
  }
  return icordry;
}

/* Drying and rewetting, ISDRY = 99: EFDC/calpuv2c.f90:604-716, one pass
   over L = 2..LA in order, switches updated in place (a cell reads the
   faces its west/south neighbours already set).  Returns ICORDRY. */
static int rewet_99(orc_ctx *o) {
  const double HDRY = o->HDRY, HDRY2 = 2. * HDRY, HDRY90 = 0.9 * HDRY;
  double *HP = A(EFDC_HP), *H1P = A(EFDC_H1P), *SUB = A(EFDC_SUB), *SVB = A(EFDC_SVB),
         *SUBO = A(EFDC_SUBO), *SVBO = A(EFDC_SVBO), *UHDYE = A(EFDC_UHDYE),
         *VHDXE = A(EFDC_VHDXE);
  int icordry = 0;
  for (int L = 2; L <= o->LA; ++L) {
    const size_t i = HX(L), w = HX(o->LWC[i]), e = HX(o->LEC[i]), s = HX(o->LSC[i]),
                 nn = HX(o->LNC[i]);
    if (!(HP[i] <= HDRY)) continue;
    const double SUBW = SUB[i], SUBE = SUB[e], SVBS = SVB[i], SVBN = SVB[nn];
    const double DHPDT = HP[i] - H1P[i];
    if (DHPDT > 0.0) { /* re-wetting allowed, :626-700 */
      SUB[i] = 0.;
      SUB[e] = 0.;
      SVB[i] = 0.;
      SVB[nn] = 0.;
      const double DHPDT2 = (HDRY - H1P[i]) * 0.01;
      if (DHPDT > DHPDT2) { /* rising fast enough to stay wet */
        if (SUBO[i] > 0.5 && UHDYE[i] > 0.0 && HP[w] > HDRY2) SUB[i] = SUBO[i];
        if (SUBO[e] > 0.5 && UHDYE[e] < 0.0 && HP[e] > HDRY2) SUB[e] = SUBO[e];
        if (SVBO[i] > 0.5 && VHDXE[i] > 0.0 && HP[s] > HDRY2) SVB[i] = SVBO[i];
        if (SVBO[nn] > 0.5 && VHDXE[nn] < 0.0 && HP[nn] > HDRY2) SVB[nn] = SVBO[nn];
        const double RDRY = SUB[i] + SUB[e] + SVB[i] + SVB[nn];
        o->iscdry[i] = RDRY < 0.5 ? 1 : 0;
        if (fabs(SUB[i] - SUBW) > 0.5 || fabs(SUB[e] - SUBE) > 0.5 ||
            fabs(SVB[i] - SVBS) > 0.5 || fabs(SVB[nn] - SVBN) > 0.5)
          icordry = 1;
      } else if (o->iscdry[i] == 0) { /* rising, not fast enough */
        o->iscdry[i] = 1;
        icordry = 1;
      }
    } else if (HP[i] < HDRY90 || H1P[i] < HDRY) { /* :700-714 */
      SUB[i] = 0.;
      SUB[e] = 0.;
      SVB[i] = 0.;
      SVB[nn] = 0.;
      if (o->iscdry[i] == 0) {
        o->iscdry[i] = 1;
        icordry = 1;
      }
    }
  }
  return icordry;
}

int orc_calpuv2c(orc_ctx *o, double dt, int64_t n, orc_cg_out *out) {
  const int LA = o->LA;
  const double G = o->G, GI = 1. / o->G; /* EFDC/aaefdc.f90:648 */
  const double DELTD2 = 0.5 * dt;
  const int nopt = noptimal(o);
  const int isdry = o->ISDRY;
  double *UHDYE = A(EFDC_UHDYE), *VHDXE = A(EFDC_VHDXE), *P = A(EFDC_P),
         *SUB = A(EFDC_SUB), *SVB = A(EFDC_SVB), *HRUO = A(EFDC_HRUO),
         *HRVO = A(EFDC_HRVO), *HU = A(EFDC_HU), *HV = A(EFDC_HV),
         *RCX = A(EFDC_RCX), *RCY = A(EFDC_RCY), *FUHDYE = A(EFDC_FUHDYE),
         *FVHDXE = A(EFDC_FVHDXE), *UHE = A(EFDC_UHE), *VHE = A(EFDC_VHE),
         *DYIU = A(EFDC_DYIU), *DXIV = A(EFDC_DXIV), *HP = A(EFDC_HP),
         *H1P = A(EFDC_H1P), *DXYIP = A(EFDC_DXYIP), *QSUME = A(EFDC_QSUME),
         *UHDY1E = A(EFDC_UHDY1E), *VHDX1E = A(EFDC_VHDX1E),
         *BELV = A(EFDC_BELV), *DXYP = A(EFDC_DXYP), *DXYU = A(EFDC_DXYU),
         *DXYV = A(EFDC_DXYV), *HPI = A(EFDC_HPI), *HUI = A(EFDC_HUI),
         *HVI = A(EFDC_HVI);
  /* step start with drying on: :153-179 */
  if (isdry > 0) {
    for (int s = 0; s < 4; ++s)
      for (int LL = 0; LL < o->pbc[s].n; ++LL) o->lopenbcdry[HX(o->pbc[s].L[LL])] = 0;
    for (int L = 2; L <= LA; ++L) {
      const size_t i = HX(L);
      o->iscdry[i] = 0;
      o->sub1[i] = SUB[i];
      o->svb1[i] = SVB[i];
      if (o->NDRYSTP < 0) o->qdwaste[i] = 0.; /* :181-188 (ISBAL > 0) */
    }
  }
  explicit_stage(o, dt, n);
  int st = EFDC_OK;
  o->ncordry = 0;
  for (;;) {
    /* label 1000, :302 */
    const double ccmnm = system_stage(o, dt, n);
    st = orc_congrad(o, out); /* :409 */
    out->ccmnm = ccmnm;
    /* face transports, :423-434 */
#pragma omp parallel for num_threads(nopt) schedule(static)
    for (int L = 2; L <= LA; ++L) {
      const size_t i = HX(L), s = HX(o->LSC[i]), w = HX(o->LWC[i]);
      const double DTU = PLANT(4, 2.0 * DELTD2, DELTD2);
      UHDYE[i] = SUB[i] * (FUHDYE[i] - DTU * HRUO[i] * RCX[i] * HU[i] * (P[i] - P[w]));
      VHDXE[i] = SVB[i] * (FVHDXE[i] - DTU * HRVO[i] * RCY[i] * HV[i] * (P[i] - P[s]));
      UHE[i] = UHDYE[i] * DYIU[i];
      VHE[i] = VHDXE[i] * DXIV[i];
    }
    /* depths, :472-477 */
#pragma omp parallel for num_threads(nopt) schedule(static)
    for (int L = 2; L <= LA; ++L) {
      const size_t i = HX(L), e = HX(o->LEC[i]), nn = HX(o->LNC[i]);
      HP[i] = H1P[i] + PLANT(9, 2.0 * DELTD2, DELTD2) * DXYIP[i] *
                           (2. * QSUME[i] -
                            (UHDYE[e] + UHDY1E[e] - UHDYE[i] - UHDY1E[i] +
                             VHDXE[nn] + VHDX1E[nn] - VHDXE[i] - VHDX1E[i]));
    }
    /* open boundary depths, :490-492 (LOBCS order S, W, E, N,
       EFDC/setbcs.f90:884-938) */
    {
      static const int order[4] = {2, 0, 1, 3};
      for (int q = 0; q < 4; ++q) {
        const orc_pbc_side *sd = &o->pbc[order[q]];
        for (int LL = 0; LL < sd->n; ++LL) {
          const size_t i = HX(sd->L[LL]);
          HP[i] = GI * P[i] - BELV[i];
        }
      }
    }
    if (isdry > 0) {
      open_bc_drying(o, dt);
      /* intermediate update of P, :559-569 (the next pass starts from it) */
      for (int L = 2; L <= LA; ++L) {
        const size_t i = HX(L);
        P[i] = G * (HP[i] + BELV[i]);
      }
    }
    const int again = (isdry > 0 && isdry < 98) ? cell_drying(o)
                      : isdry == 99            ? rewet_99(o)
                                               : 0;
    if (again && st == EFDC_OK) {
      if (++o->ncordry < 500) continue; /* GOTO 1000 */
      st = EFDC_ERR_STATE;               /* '*** NCORDRY > 500' */
    }
    break;
  }
  /* volume wasting, NDRYSTP < 0 (:838-893, then :1010-1018): a wet cell
     whose four faces are closed counts its isolated steps; past |NDRYSTP|,
     if an originally open face leads to a lower bed or HP < 2 HDRY, its
     depth goes to 0.9 HDRY (H1P and the layer depths too; P stays) and the
     volume taken is booked in QDWASTE (rate) and VDWASTE (total) */
  if (isdry > 0 && o->NDRYSTP < 0) {
    const int ntmp = -o->NDRYSTP;
    const double HDRY = o->HDRY, DELTI = 1. / dt;
    const double *SUBO = A(EFDC_SUBO), *SVBO = A(EFDC_SVBO), *DZC = A(EFDC_DZC);
    double *H1PK = A(EFDC_H1PK), *HPK = A(EFDC_HPK), *HPKI = A(EFDC_HPKI);
    for (int L = 2; L <= LA; ++L) {
      const size_t i = HX(L);
      if (!(HP[i] >= HDRY)) continue;
      const size_t w = HX(o->LWC[i]), e = HX(o->LEC[i]), s = HX(o->LSC[i]),
                   nn = HX(o->LNC[i]);
      const double RDRY = SUB[i] + SUB[e] + SVB[i] + SVB[nn];
      if (RDRY > 0.5) {
        o->natdry[i] = 0;
        continue;
      }
      o->natdry[i] = o->natdry[i] + 1;
      if (o->natdry[i] <= ntmp) continue;
      double RVAL = 0.;
      if (SUBO[e] > 0.5 && BELV[e] < BELV[i]) RVAL = RVAL + 1.;
      if (SUBO[i] > 0.5 && BELV[w] < BELV[i]) RVAL = RVAL + 1.;
      if (SVBO[nn] > 0.5 && BELV[nn] < BELV[i]) RVAL = RVAL + 1.;
      if (SVBO[i] > 0.5 && BELV[s] < BELV[i]) RVAL = RVAL + 1.;
      if (!(RVAL > 0. || HP[i] < HDRY * 2.)) continue;
      const double HOLDTMP = HP[i];
      HP[i] = 0.90 * HDRY;
      H1P[i] = 0.90 * HDRY; /* :1012-1016 */
      for (int K = 1; K <= o->KC; ++K) {
        HPK[KX(L, K)] = HP[i] * DZC[KX(L, K)];
        HPKI[KX(L, K)] = 1. / HPK[KX(L, K)];
        H1PK[KX(L, K)] = H1P[i] * DZC[KX(L, K)];
      }
      o->natdry[i] = 0;
      o->qdwaste[i] = DELTI * DXYP[i] * (HOLDTMP - HP[i]);
      o->vdwaste[i] = o->vdwaste[i] + DXYP[i] * (HOLDTMP - HP[i]);
    }
  }
  /* final updates, :896-914 (FSGZUDXYPI = 0.5/DXYU, :110-117); with
     drying on, P was updated in the loop (:896-900) */
  if (isdry == 0) {
#pragma omp parallel for num_threads(nopt) schedule(static)
    for (int L = 2; L <= LA; ++L) {
      const size_t i = HX(L);
      P[i] = G * (HP[i] + BELV[i]);
    }
  }
#pragma omp parallel for num_threads(nopt) schedule(static)
  for (int L = 2; L <= LA; ++L) {
    const size_t i = HX(L), s = HX(o->LSC[i]), w = HX(PLANT(5, o->LEC[i], o->LWC[i]));
    HU[i] = (DXYP[i] * HP[i] + DXYP[w] * HP[w]) * (0.5 / DXYU[i]);
    HV[i] = (DXYP[i] * HP[i] + DXYP[s] * HP[s]) * (PLANT(6, 1.0, 0.5) / DXYV[i]);
  }
#pragma omp parallel for num_threads(nopt) schedule(static)
  for (int L = 2; L <= LA; ++L) {
    const size_t i = HX(L);
    HPI[i] = 1. / HP[i];
    HUI[i] = 1. / HU[i];
    HVI[i] = 1. / HV[i];
  }
  /* transport mask of dry cells, :916-939: dry when HP < HDRY and all four
     faces were closed at the step start and are closed now */
  if (isdry > 0) {
    for (int L = 2; L <= LA; ++L) {
      const size_t i = HX(L), e = HX(o->LEC[i]), nn = HX(o->LNC[i]);
      int m = 1;
      if (HP[i] < o->HDRY) {
        const int iue = o->sub1[i] < 0.5 && SUB[i] < 0.5;
        const int iuw = o->sub1[e] < 0.5 && SUB[e] < 0.5;
        const int ivs = o->svb1[i] < 0.5 && SVB[i] < 0.5;
        const int ivn = o->svb1[nn] < 0.5 && SVB[nn] < 0.5;
        if (iue + iuw + ivs + ivn == 4) m = 0;
      }
      o->lmaskdry[i] = m;
    }
  }
  /* NEGDEP, EFDC/negdep.f90:34-45 */
  if (st == EFDC_OK) {
    for (int L = 2; L <= LA; ++L)
      if (HP[HX(L)] <= 0.) {
        st = EFDC_ERR_NEGDEPTH;
        break;
      }
  }
  out->status = st;
  return st;
}

/* ===================================================================== */
/* CALTRAN (2TL: ISTL=2, IS2TL=1, ISCDCA=0) with CALFQC                     */
/* ===================================================================== */
typedef struct {
  double *FUHUD, *FVHUD, *FWUU, *FQC, *CD, *POS, *UUUU, *VVVV, *WWWW,
      *FUHVD, *FVHVD, *FWVV, *DUU, *DVV, *CMAX, *CMIN, *CONTMX, *CONTMN,
      *CON2, *WQBCCON;
} tran_scratch;

/* The first five arrays are always needed; the anti-diffusion/FCT set
   only when ISADAC>0.  Zero-initialised like the reference's allocations,
   so ghost entries (L=1, L=LC) read as 0. */
static void scratch_alloc(const orc_ctx *o, tran_scratch *s, int adac) {
  const size_t nk = (size_t)o->LCM * (size_t)(o->KC + 2);
  double **p = (double **)s;
  const size_t nall = sizeof(tran_scratch) / sizeof(double *);
  for (size_t j = 0; j < nall; ++j)
    p[j] = (j < 5 || adac) ? calloc(nk, sizeof(double)) : NULL;
}
static void scratch_free(tran_scratch *s) {
  double **p = (double **)s;
  for (size_t j = 0; j < sizeof(tran_scratch) / sizeof(double *); ++j)
    free(p[j]);
}

/* arrays with a K=0 plane (FWUU, FWVV, WWWW): index (L,K), K in 0..KC+1 */
#define ZX(L, K) ((size_t)(L) - 1 + (size_t)(K) * (size_t)o->LCM)

static void orc_caltran(orc_ctx *o, int c, double dt, int64_t niter,
                        tran_scratch *s) {
  const int LA = o->LA, KC = o->KC, KS = KC - 1, LC = LA + 1;
  orc_const *cc = &o->c[c];
  double *CON = cc->con, *CON1 = cc->con1;
  const int M = cc->mo, nc = o->nconst;
  const double DDELT = dt, DDELTA = dt, BSMALL = 1.0E-6; /* :36-50 */
  double *UHDY2 = A(EFDC_UHDY2), *VHDX2 = A(EFDC_VHDX2), *W2 = A(EFDC_W2),
         *H1PK = A(EFDC_H1PK), *HPKI = A(EFDC_HPKI), *DXYIP = A(EFDC_DXYIP),
         *HPK = A(EFDC_HPK);
  const int32_t *LWC = o->LWC, *LEC = o->LEC, *LSC = o->LSC, *LNC = o->LNC;
  (void)LC;

  /* CON1 = CON, :59-65 */
  for (int K = 1; K <= KC; ++K)
    for (int L = 1; L <= o->LCM; ++L) CON1[KX(L, K)] = CON[KX(L, K)];

  /* CALFQC, EFDC/calfqc.f90:42-225 (2TL, MVAR!=8, no GW, no structures):
     zero FQC at the flow cells, CONQ = 0.5*(3*CON-CON1), accumulate. */
  for (int ns = 0; ns < o->nqsij; ++ns)
    for (int K = 1; K <= KC; ++K) s->FQC[KX(o->lqs[ns], K)] = 0.;
  for (int ns = 0; ns < o->nqsij; ++ns) {
    const int L = o->lqs[ns];
    for (int K = 1; K <= KC; ++K) {
      const double conq = 0.5 * (3. * CON[KX(L, K)] - CON1[KX(L, K)]);
      const double q = o->qss[ns * KC + K - 1];
      const double qs = o->qsercell[ns * KC + K - 1];
      const double cq = o->cqs[(ns * KC + K - 1) * nc + c];
      const double cs = o->csert_q[(ns * KC + K - 1) * nc + c];
      s->FQC[KX(L, K)] = s->FQC[KX(L, K)] + dmax(q, 0.) * cq +
                         dmin(q, 0.) * conq + dmax(qs, 0.) * cs +
                         dmin(qs, 0.) * conq;
    }
  }

  /* upwind fluxes with the donor maps of CALCONC, EFDC/calconc.f90:89-121,
     EFDC/caltran.f90:95-111 */
  for (int K = 1; K <= KC; ++K)
    for (int L = 2; L <= LA; ++L) {
      const size_t i = HX(L);
      const int LUPU = UHDY2[KX(L, K)] >= 0.0 ? LWC[i] : L;
      const int LUPV = VHDX2[KX(L, K)] >= 0.0 ? LSC[i] : L;
      s->FUHUD[KX(L, K)] = UHDY2[KX(L, K)] * CON1[KX(LUPU, K)];
      s->FVHUD[KX(L, K)] = VHDX2[KX(L, K)] * CON1[KX(LUPV, K)];
    }
  if (KC > 1)
    for (int K = 1; K <= KS; ++K)
      for (int L = 2; L <= LA; ++L) {
        const int KUPW = W2[WX(L, K)] >= 0. ? K : K + 1;
        s->FWUU[ZX(L, K)] = W2[WX(L, K)] * CON1[KX(L, KUPW)];
      }

  /* 2TL update, :190-199, over the wet list LKWET (:95-300): a cell left
     out by LMASKDRY keeps CON = CON1.  Its faces were closed at both time
     levels, so its UHDY2/VHDX2 and the fluxes above are 0 (the reference
     zeroes them for the dry cells, EFDC/calconc.f90:127-150). */
  const int32_t *wet = o->ISDRY > 0 ? o->lmaskdry : NULL;
  for (int K = 1; K <= KC; ++K)
    for (int L = 2; L <= LA; ++L) {
      const size_t i = HX(L);
      if (wet && !wet[i]) continue;
      s->CD[KX(L, K)] =
          CON1[KX(L, K)] * H1PK[KX(L, K)] +
          DDELT * ((s->FQC[KX(L, K)] + s->FUHUD[KX(L, K)] -
                    s->FUHUD[KX(LEC[i], K)] + s->FVHUD[KX(L, K)] -
                    s->FVHUD[KX(LNC[i], K)]) *
                       DXYIP[i] +
                   (s->FWUU[ZX(L, K - 1)] - s->FWUU[ZX(L, K)]));
    }
  if (cc->isfct >= 1 && cc->isadac > 0) /* :200-207 */
    for (int K = 1; K <= KC; ++K)
      for (int L = 2; L <= LA; ++L)
        s->CON2[KX(L, K)] = dmax(CON1[KX(L, K)], 0.0);
  for (int K = 1; K <= KC; ++K) /* :240-245 */
    for (int L = 2; L <= LA; ++L)
      if (!wet || wet[HX(L)]) CON[KX(L, K)] = s->CD[KX(L, K)] * HPKI[KX(L, K)];

  /* open boundary concentrations S, W, E, N: :316-471 (ISTL=2, ISUD=1) */
  for (int side = 0; side < 4; ++side) {
    orc_cbc_side *b = &o->cbc[side];
    for (int K = 1; K <= KC; ++K)
      for (int LL = 0; LL < b->n; ++LL) {
        const int L = b->L[LL];
        const size_t i = HX(L);
        if (wet && !wet[i]) continue; /* :324,367,404,441 */
        const size_t m = ((size_t)LL * KC + K - 1) * nc + c;
        const double cb1 = b->cb[((size_t)LL * 2 + 0) * nc + c];
        const double cb2 = b->cb[((size_t)LL * 2 + 1) * nc + c];
        const double cs1 = b->csert[((size_t)LL * KC + 0) * nc + c];
        const double csk = b->csert[m];
        double flow, flux;
        int outflow;
        if (side == 0) { /* south: face LN */
          const int LN = LNC[i];
          flow = VHDX2[KX(LN, K)];
          flux = s->FVHUD[KX(LN, K)];
          outflow = flow <= 0.;
        } else if (side == 1) { /* west: face LEC */
          const int LE = LEC[i];
          flow = UHDY2[KX(LE, K)];
          flux = s->FUHUD[KX(LE, K)];
          outflow = flow <= 0.;
        } else if (side == 2) { /* east: own face */
          flow = UHDY2[KX(L, K)];
          flux = s->FUHUD[KX(L, K)];
          outflow = flow >= 0.;
        } else { /* north: own face */
          flow = VHDX2[KX(L, K)];
          flux = s->FVHUD[KX(L, K)];
          outflow = flow >= 0.;
        }
        if (outflow) {
          double CTMP;
          if (side <= 1)
            CTMP = CON1[KX(L, K)] + DDELT * (flow * CON1[KX(L, K)] - flux) *
                                        DXYIP[i] * HPKI[KX(L, K)];
          else
            CTMP = CON1[KX(L, K)] + DDELT * (flux - flow * CON1[KX(L, K)]) *
                                        DXYIP[i] * HPKI[KX(L, K)];
          CON[KX(L, K)] = dmax(CTMP, 0.);
          if (M == 1) {
            const double cbtmp = cb1 + cs1;
            if (CON[KX(L, K)] > cbtmp) CON[KX(L, K)] = cbtmp;
          }
          b->clo[m] = CON[KX(L, K)];
          b->nlo[m] = niter;
        } else {
          const double CBT = o->wtci[(K - 1) * 2 + 0] * cb1 +
                             o->wtci[(K - 1) * 2 + 1] * cb2 + csk;
          const int64_t nmnlod = niter - b->nlo[m];
          if (nmnlod >= b->ntscr[LL]) {
            CON[KX(L, K)] = CBT;
          } else {
            const double cbtmp =
                b->clo[m] + (CBT - b->clo[m]) * (double)nmnlod / (double)b->ntscr[LL];
            CON[KX(L, K)] = dmax(cbtmp, 0.);
          }
        }
        CON1[KX(L, K)] = CON[KX(L, K)]; /* ISUD=1 */
      }
  }

  /* anti-diffusion, :475-786 */
  if (cc->isadac == 0) return;
  double *U2 = A(EFDC_U2), *V2 = A(EFDC_V2), *DXIU = A(EFDC_DXIU),
         *DYIV = A(EFDC_DYIV), *HPI = A(EFDC_HPI), *DZIG = A(EFDC_DZIG),
         *QSUM = A(EFDC_QSUM), *SUB3D = A(EFDC_SUB3D), *SVB3D = A(EFDC_SVB3D);
  const int32_t *LNWC = o->LNWC, *LSEC = o->LSEC;
  /* open-BC cell list LOBCS in the order S, W, E, N (EFDC/setbcs.f90) */
  int nob = 0, nob2 = 0;
  int *lob = malloc(sizeof(int) * (size_t)(o->pbc[0].n + o->pbc[1].n + o->pbc[2].n + o->pbc[3].n + 1));
  int *lob2 = malloc(sizeof(int) * (size_t)(o->pbc[0].n + o->pbc[1].n + o->pbc[2].n + o->pbc[3].n + 1));
  {
    static const int order[4] = {2, 0, 1, 3};
    for (int q = 0; q < 4; ++q) {
      const orc_pbc_side *ps = &o->pbc[order[q]];
      for (int LL = 0; LL < ps->n; ++LL) {
        const int L = ps->L[LL];
        lob[nob++] = L;
        if (ps->ispb[LL] <= 1) {
          int nb = order[q] == 2 ? LNC[HX(L)] : order[q] == 0 ? LEC[HX(L)]
                 : order[q] == 1 ? LWC[HX(L)] : LSC[HX(L)];
          lob2[nob2++] = nb;
        }
      }
    }
  }
  for (int q = 0; q < nob; ++q) /* :482-487 */
    for (int K = 1; K <= KC; ++K)
      s->WQBCCON[(size_t)q * KC + K - 1] = CON[KX(lob[q], K)];
  for (int K = 1; K <= KC; ++K) /* :490-495 */
    for (int L = 2; L <= LA; ++L) s->POS[KX(L, K)] = dmax(CON[KX(L, K)], 0.);
  for (int K = 1; K <= KC; ++K) /* :497-503 */
    for (int L = 2; L <= LA; ++L) {
      const size_t i = HX(L);
      s->UUUU[KX(L, K)] = U2[KX(L, K)] * (s->POS[KX(L, K)] - s->POS[KX(LWC[i], K)]) * DXIU[i];
      s->VVVV[KX(L, K)] = V2[KX(L, K)] * (s->POS[KX(L, K)] - s->POS[KX(LSC[i], K)]) * DYIV[i];
    }
  if (KC > 1) /* :505-511 */
    for (int K = 1; K <= KS; ++K)
      for (int L = 2; L <= LA; ++L)
        s->WWWW[ZX(L, K)] = W2[WX(L, K)] * (s->POS[KX(L, K + 1)] - s->POS[KX(L, K)]) *
                            HPI[HX(L)] * DZIG[WX(L, K)];
  for (int K = 1; K <= KC; ++K) /* :522-558 */
    for (int L = 2; L <= LA; ++L) {
      const size_t i = HX(L);
      const int LE = LEC[i], LW = LWC[i], LN = LNC[i], LS = LSC[i];
      const int LNW = LNWC[i], LSE = LSEC[i];
      const double uh = UHDY2[KX(L, K)];
      double UTERM = fabs(uh) * (s->POS[KX(L, K)] - s->POS[KX(LW, K)]);
      if (uh >= 0.0)
        UTERM = UTERM - 0.5 * DDELTA * uh *
                            (s->VVVV[KX(LNW, K)] + s->VVVV[KX(LW, K)] +
                             s->WWWW[ZX(LW, K)] + s->WWWW[ZX(LW, K - 1)] +
                             s->UUUU[KX(L, K)] + s->UUUU[KX(LW, K)]);
      else
        UTERM = UTERM - 0.5 * DDELTA * uh *
                            (s->VVVV[KX(LN, K)] + s->VVVV[KX(L, K)] +
                             s->WWWW[ZX(L, K)] + s->WWWW[ZX(L, K - 1)] +
                             s->UUUU[KX(L, K)] + s->UUUU[KX(LE, K)]);
      const double UHU = UTERM / (s->POS[KX(L, K)] + s->POS[KX(LW, K)] + BSMALL);
      s->FUHUD[KX(L, K)] = dmax(UHU, 0.) * s->POS[KX(LW, K)] + dmin(UHU, 0.) * s->POS[KX(L, K)];
      const double vh = VHDX2[KX(L, K)];
      double VTERM = fabs(vh) * (s->POS[KX(L, K)] - s->POS[KX(LS, K)]);
      if (vh >= 0.0)
        VTERM = VTERM - 0.5 * DDELTA * vh *
                            (s->UUUU[KX(LS, K)] + s->UUUU[KX(LSE, K)] +
                             s->WWWW[ZX(LS, K)] + s->WWWW[ZX(LS, K - 1)] +
                             s->VVVV[KX(LS, K)] + s->VVVV[KX(L, K)]);
      else
        VTERM = VTERM - 0.5 * DDELTA * vh *
                            (s->UUUU[KX(L, K)] + s->UUUU[KX(LE, K)] +
                             s->WWWW[ZX(L, K)] + s->WWWW[ZX(L, K - 1)] +
                             s->VVVV[KX(LN, K)] + s->VVVV[KX(L, K)]);
      const double VHV = VTERM / (s->POS[KX(L, K)] + s->POS[KX(LS, K)] + BSMALL);
      s->FVHUD[KX(L, K)] = dmax(VHV, 0.) * s->POS[KX(LS, K)] + dmin(VHV, 0.) * s->POS[KX(L, K)];
    }
  if (KC > 1) /* :560-577 */
    for (int K = 1; K <= KS; ++K)
      for (int L = 2; L <= LA; ++L) {
        const size_t i = HX(L);
        const int LE = LEC[i], LN = LNC[i];
        const double ww = W2[WX(L, K)];
        double WTERM = fabs(ww) * (s->POS[KX(L, K + 1)] - s->POS[KX(L, K)]);
        if (ww >= 0.0)
          WTERM = WTERM - 0.5 * DDELTA * ww *
                              (s->UUUU[KX(L, K)] + s->UUUU[KX(LE, K)] +
                               s->VVVV[KX(L, K)] + s->VVVV[KX(LN, K)] +
                               s->WWWW[ZX(L, K)] + s->WWWW[ZX(L, K - 1)]);
        else
          WTERM = WTERM - 0.5 * DDELTA * ww *
                              (s->UUUU[KX(L, K + 1)] + s->UUUU[KX(LE, K + 1)] +
                               s->VVVV[KX(L, K + 1)] + s->VVVV[KX(LN, K + 1)] +
                               s->WWWW[ZX(L, K)] + s->WWWW[ZX(L, K + 1)]);
        const double WW = WTERM / (s->POS[KX(L, K + 1)] + s->POS[KX(L, K)] + BSMALL);
        s->FWUU[ZX(L, K)] = dmax(WW, 0.) * s->POS[KX(L, K)] + dmin(WW, 0.) * s->POS[KX(L, K + 1)];
      }
  if (cc->isadac == 1) /* :583-599 */
    for (int K = 1; K <= KC; ++K)
      for (int L = 2; L <= LA; ++L)
        if (fabs(QSUM[KX(L, K)]) > 1.E-8) {
          const size_t i = HX(L);
          s->FUHUD[KX(L, K)] = 0.;
          s->FUHUD[KX(LEC[i], K)] = 0.;
          s->FVHUD[KX(L, K)] = 0.;
          s->FVHUD[KX(LNC[i], K)] = 0.;
          s->FWUU[ZX(L, K)] = 0.;
          s->FWVV[ZX(L, K - 1)] = 0.;
        }
  for (int K = 1; K <= KC; ++K) { /* :602-620 */
    for (int LL = 0; LL < o->cbc[0].n; ++LL)
      s->FVHUD[KX(LNC[HX(o->cbc[0].L[LL])], K)] = 0.0;
    for (int LL = 0; LL < o->cbc[1].n; ++LL)
      s->FUHUD[KX(LEC[HX(o->cbc[1].L[LL])], K)] = 0.0;
    for (int LL = 0; LL < o->cbc[2].n; ++LL)
      s->FUHUD[KX(o->cbc[2].L[LL], K)] = 0.0;
    for (int LL = 0; LL < o->cbc[3].n; ++LL)
      s->FVHUD[KX(o->cbc[3].L[LL], K)] = 0.0;
  }
  if (cc->isfct != 0) { /* :624-768 */
    for (int K = 1; K <= KC; ++K)
      for (int L = 2; L <= LA; ++L) {
        s->CONTMX[KX(L, K)] = dmax(s->POS[KX(L, K)], fabs(s->CON2[KX(L, K)]));
        s->CONTMN[KX(L, K)] = dmin(s->POS[KX(L, K)], fabs(s->CON2[KX(L, K)]));
      }
    for (int L = 2; L <= LA; ++L) { /* KSZ = 1 (sigma grid) */
      s->CMAX[KX(L, 1)] = dmax(s->CONTMX[KX(L, 1)], s->CONTMX[KX(L, 2)]);
      s->CMAX[KX(L, KC)] = dmax(s->CONTMX[KX(L, KS)], s->CONTMX[KX(L, KC)]);
      s->CMIN[KX(L, 1)] = dmin(s->CONTMN[KX(L, 1)], s->CONTMN[KX(L, 2)]);
      s->CMIN[KX(L, KC)] = dmin(s->CONTMN[KX(L, KS)], s->CONTMN[KX(L, KC)]);
    }
    for (int K = 2; K <= KS; ++K)
      for (int L = 2; L <= LA; ++L) {
        const double CMAXT = dmax(s->CONTMX[KX(L, K - 1)], s->CONTMX[KX(L, K + 1)]);
        s->CMAX[KX(L, K)] = dmax(s->CONTMX[KX(L, K)], CMAXT);
        const double CMINT = dmin(s->CONTMN[KX(L, K - 1)], s->CONTMN[KX(L, K + 1)]);
        s->CMIN[KX(L, K)] = dmin(s->CONTMN[KX(L, K)], CMINT);
      }
    /* with drying, SUB3D = SUB*SUB3DO (EFDC/calpuv2c.f90:982-990): the
       SUB3D/SVB3D arrays hold SUB3DO/SVB3DO and are scaled here */
    const double *SUBh = A(EFDC_SUB), *SVBh = A(EFDC_SVB);
    const int dry3d = o->ISDRY > 0;
#define SB3(L_, K_) (dry3d ? SUBh[HX(L_)] * SUB3D[KX(L_, K_)] : SUB3D[KX(L_, K_)])
#define SV3(L_, K_) (dry3d ? SVBh[HX(L_)] * SVB3D[KX(L_, K_)] : SVB3D[KX(L_, K_)])
    for (int K = 1; K <= KC; ++K)
      for (int L = 2; L <= LA; ++L) {
        const size_t i = HX(L);
        const int LE = LEC[i], LS = LSC[i], LN = LNC[i], LW = LWC[i];
        const double sbi = SB3(L, K), sbe = SB3(LE, K), svi = SV3(L, K), svn = SV3(LN, K);
        const double CWMAX = sbi * s->CONTMX[KX(LW, K)];
        const double CEMAX = sbe * s->CONTMX[KX(LE, K)];
        const double CSMAX = svi * s->CONTMX[KX(LS, K)];
        const double CNMAX = svn * s->CONTMX[KX(LN, K)];
        double CMAXT = dmax(CNMAX, CEMAX);
        CMAXT = dmax(CMAXT, CSMAX);
        CMAXT = dmax(CMAXT, CWMAX);
        s->CMAX[KX(L, K)] = dmax(s->CMAX[KX(L, K)], CMAXT);
        const double CWMIN = sbi * s->CONTMN[KX(LW, K)] + 1.E+6 * (1. - sbi);
        const double CEMIN = sbe * s->CONTMN[KX(LE, K)] + 1.E+6 * (1. - sbe);
        const double CSMIN = svi * s->CONTMN[KX(LS, K)] + 1.E+6 * (1. - svi);
        const double CNMIN = svn * s->CONTMN[KX(LN, K)] + 1.E+6 * (1. - svn);
        double CMINT = dmin(CNMIN, CEMIN);
        CMINT = dmin(CMINT, CSMIN);
        CMINT = dmin(CMINT, CWMIN);
        s->CMIN[KX(L, K)] = dmin(s->CMIN[KX(L, K)], CMINT);
      }
#undef SB3
#undef SV3
    for (int K = 1; K <= KC; ++K) /* :683-691 */
      for (int L = 2; L <= LA; ++L) {
        s->FUHVD[KX(L, K)] = dmin(s->FUHUD[KX(L, K)], 0.);
        s->FUHUD[KX(L, K)] = dmax(s->FUHUD[KX(L, K)], 0.);
        s->FVHVD[KX(L, K)] = dmin(s->FVHUD[KX(L, K)], 0.);
        s->FVHUD[KX(L, K)] = dmax(s->FVHUD[KX(L, K)], 0.);
      }
    if (KC > 1) /* :693-701 */
      for (int K = 1; K <= KS; ++K)
        for (int L = 2; L <= LA; ++L) {
          s->FWVV[ZX(L, K)] = dmin(s->FWUU[ZX(L, K)], 0.);
          s->FWUU[ZX(L, K)] = dmax(s->FWUU[ZX(L, K)], 0.);
        }
    for (int K = 1; K <= KC; ++K) /* :705-714 */
      for (int L = 2; L <= LA; ++L) {
        const size_t i = HX(L);
        const int LE = LEC[i], LN = LNC[i];
        s->DUU[KX(L, K)] = DDELT *
            (DXYIP[i] * (s->FUHUD[KX(L, K)] - s->FUHVD[KX(LE, K)] + s->FVHUD[KX(L, K)] - s->FVHVD[KX(LN, K)]) +
             (s->FWUU[ZX(L, K - 1)] - s->FWVV[ZX(L, K)])) * HPKI[KX(L, K)];
        s->DVV[KX(L, K)] = DDELT *
            (DXYIP[i] * (s->FUHUD[KX(LE, K)] - s->FUHVD[KX(L, K)] + s->FVHUD[KX(LN, K)] - s->FVHVD[KX(L, K)]) +
             (s->FWUU[ZX(L, K)] - s->FWVV[ZX(L, K - 1)])) * HPKI[KX(L, K)];
      }
    for (int K = 1; K <= KC; ++K) /* :717-723 */
      for (int q = 0; q < nob; ++q) {
        s->DUU[KX(lob[q], K)] = 0.;
        s->DVV[KX(lob[q], K)] = 0.;
      }
    for (int K = 1; K <= KC; ++K) /* :726-734 */
      for (int L = 2; L <= LA; ++L) {
        const size_t j = KX(L, K);
        if (s->DUU[j] > 0.) s->DUU[j] = (s->CMAX[j] - s->POS[j]) / (s->DUU[j] + BSMALL);
        s->DUU[j] = dmin(s->DUU[j], 1.);
        if (s->DVV[j] > 0.) s->DVV[j] = (s->POS[j] - s->CMIN[j]) / (s->DVV[j] + BSMALL);
        s->DVV[j] = dmin(s->DVV[j], 1.);
      }
    for (int K = 1; K <= KC; ++K) /* :737-745 */
      for (int L = 2; L <= LA; ++L) {
        const size_t i = HX(L);
        const int LS = LSC[i], LW = LWC[i];
        s->FUHUD[KX(L, K)] = dmin(s->DVV[KX(LW, K)], s->DUU[KX(L, K)]) * s->FUHUD[KX(L, K)] +
                             dmin(s->DUU[KX(LW, K)], s->DVV[KX(L, K)]) * s->FUHVD[KX(L, K)];
        s->FVHUD[KX(L, K)] = dmin(s->DVV[KX(LS, K)], s->DUU[KX(L, K)]) * s->FVHUD[KX(L, K)] +
                             dmin(s->DUU[KX(LS, K)], s->DVV[KX(L, K)]) * s->FVHVD[KX(L, K)];
      }
    for (int K = 1; K <= KS; ++K) /* :747-752 */
      for (int L = 2; L <= LA; ++L)
        s->FWUU[ZX(L, K)] = dmin(s->DVV[KX(L, K)], s->DUU[KX(L, K + 1)]) * s->FWUU[ZX(L, K)] +
                            dmin(s->DUU[KX(L, K)], s->DVV[KX(L, K + 1)]) * s->FWVV[ZX(L, K)];
    for (int q = 0; q < nob; ++q) /* :755-760 */
      for (int K = 1; K <= KS; ++K) s->FWUU[ZX(lob[q], K)] = 0.0;
    for (int q = 0; q < nob2; ++q) /* :761-766 */
      for (int K = 1; K <= KS; ++K) s->FWUU[ZX(lob2[q], K)] = 0.0;
  }
  for (int K = 1; K <= KC; ++K) /* :771-778, wet list */
    for (int L = 2; L <= LA; ++L) {
      const size_t i = HX(L);
      if (wet && !wet[i]) continue;
      s->CD[KX(L, K)] = CON[KX(L, K)] * HPK[KX(L, K)] +
          DDELT * ((s->FUHUD[KX(L, K)] - s->FUHUD[KX(LEC[i], K)] + s->FVHUD[KX(L, K)] -
                    s->FVHUD[KX(LNC[i], K)]) * DXYIP[i] +
                   (s->FWUU[ZX(L, K - 1)] - s->FWUU[ZX(L, K)]));
      CON[KX(L, K)] = s->CD[KX(L, K)] * HPKI[KX(L, K)];
    }
  for (int K = 1; K <= KC; ++K) /* :781-786 */
    for (int q = 0; q < nob; ++q)
      CON[KX(lob[q], K)] = s->WQBCCON[(size_t)q * KC + K - 1];
  free(lob);
  free(lob2);
}

/* ===================================================================== */
/* CALCONC vertical diffusion: EFDC/calconc.f90:239-473                    */
/* ===================================================================== */
int orc_vertical_diffusion(orc_ctx *o, double dt) {
  const int LA = o->LA, KC = o->KC, KS = KC - 1;
  if (KC == 1) return EFDC_OK; /* :239 */
  const double DELTD2 = dt;    /* :65-73 */
  double *CDZKK = A(EFDC_CDZKK), *CDZKMK = A(EFDC_CDZKMK), *AB = A(EFDC_AB),
         *HPI = A(EFDC_HPI);
  const int nc = o->nconst;
  const int nt = o->nthreads > 0 ? o->nthreads : 1;
#pragma omp parallel num_threads(nt)
  {
    double *VCU = malloc(sizeof(double) * (size_t)(KC + 1));
#pragma omp for schedule(dynamic, 256)
    for (int L = 2; L <= LA; ++L) {
      const size_t i = HX(L);
      if (o->ISDRY > 0 && !o->lmaskdry[i]) continue; /* LKWET loops */
      double EEB[64], CCLB[64];
      /* bottom layer, :265-272 (KSZ=1) */
      {
        const double RCDZKK = -DELTD2 * CDZKK[KX(L, 1)];
        const double CCUBTMP = RCDZKK * HPI[i] * AB[KX(L, 1)];
        const double CCMBTMP = 1.0 - CCUBTMP;
        EEB[1] = 1.0 / CCMBTMP;
        VCU[1] = CCUBTMP * EEB[1];
      }
      /* middle layers, :317-327 */
      for (int K = 2; K <= KS; ++K) {
        const double RCDZKMK = -DELTD2 * CDZKMK[KX(L, K)];
        const double RCDZKK = -DELTD2 * CDZKK[KX(L, K)];
        CCLB[K] = RCDZKMK * HPI[i] * AB[KX(L, K - 1)];
        const double CCUBTMP = RCDZKK * HPI[i] * AB[KX(L, K)];
        const double CCMBTMP = 1.0 - CCLB[K] - CCUBTMP;
        EEB[K] = 1.0 / (CCMBTMP - CCLB[K] * VCU[K - 1]);
        VCU[K] = CCUBTMP * EEB[K];
      }
      /* top layer, :373-380 */
      {
        const int K = KC;
        const double RCDZKMK = -DELTD2 * CDZKMK[KX(L, K)];
        CCLB[K] = RCDZKMK * HPI[i] * AB[KX(L, K - 1)];
        const double CCMBTMP = 1.0 - CCLB[K];
        EEB[K] = 1.0 / (CCMBTMP - CCLB[K] * VCU[K - 1]);
      }
      for (int c = 0; c < nc; ++c) {
        double *C = o->c[c].con;
        C[KX(L, 1)] = C[KX(L, 1)] * EEB[1];                       /* :276 */
        for (int K = 2; K <= KC; ++K)                             /* :331,384 */
          C[KX(L, K)] = (C[KX(L, K)] - CCLB[K] * C[KX(L, K - 1)]) * EEB[K];
        for (int K = KS; K >= 1; --K)                             /* :425-468 */
          C[KX(L, K)] = C[KX(L, K)] - VCU[K] * C[KX(L, K + 1)];
      }
    }
    free(VCU);
  }
  return EFDC_OK;
}

/* Cohesive settling of the water column after the transport, SSEDTOX ->
   CALSED with DTSED = DT (EFDC/calconc.f90:621-630 when SEDSTEP <= DT) and
   WSETA = WSEDO (ISEDVW = 0, EFDC/calsed.f90:39-49): the implicit upwind
   sweep from the top layer down, EFDC/calsed.f90:232-256, then the bottom
   layer with the deposition probability PROBDEP and the resuspension flux
   WESE given (the bed model that sets them is not restated),
   EFDC/calsed.f90:308-312. */
void orc_settle(orc_ctx *o, double dtsed) {
  const int KC = o->KC;
  const double *HPKI = A(EFDC_HPKI);
  for (int c = 0; c < o->nconst; ++c) {
    const double ws = o->c[c].wset;
    if (!(ws > 0.0)) continue;
    double *SED = o->c[c].con;
    const double WSETMP = o->c[c].probdep * ws, WESE = o->c[c].wese;
    for (int L = 2; L <= o->LA; ++L) {
      if (o->ISDRY > 0 && !o->lmaskdry[HX(L)]) continue; /* LWET loops */
      double sedf = 0.0; /* SEDF(L,KC) = 0 */
      if (KC > 1) {
        {
          const int K = KC;
          const double WVEL = dtsed * HPKI[KX(L, K)];
          const double CLEFT = 1. + ws * WVEL;
          const double CRIGHT = dmax(SED[KX(L, K)], 0.);
          SED[KX(L, K)] = CRIGHT / CLEFT;
          sedf = -ws * SED[KX(L, K)];
        }
        for (int K = KC - 1; K >= 2; --K) {
          const double WVEL = dtsed * HPKI[KX(L, K)];
          const double CLEFT = 1. + ws * WVEL;
          const double CRIGHT = dmax(SED[KX(L, K)], 0.) - sedf * WVEL;
          SED[KX(L, K)] = CRIGHT / CLEFT;
          sedf = -ws * SED[KX(L, K)];
        }
      }
      const double WVEL = dtsed * HPKI[KX(L, 1)];
      const double CLEFT = 1. + WSETMP * WVEL;
      const double CRIGHT = dmax(SED[KX(L, 1)], 0.) + (WESE - sedf) * WVEL;
      SED[KX(L, 1)] = CRIGHT / CLEFT;
    }
  }
}

static void bal2t_sources(orc_ctx *o, double dt);
static void bal2t_open(orc_ctx *o, double dt);

int orc_calconc(orc_ctx *o, double dt, int64_t niter) {
  const int nc = o->nconst;
  const int nt = o->nthreads > 0 ? o->nthreads : 1;
  if (o->isbal) bal2t_sources(o, dt);  /* EFDC/calconc.f90:76-80 */
  /* one constituent per thread, EFDC/calconc.f90:159-229 */
#pragma omp parallel for num_threads(nt < nc ? nt : (nc > 0 ? nc : 1)) schedule(static, 1)
  for (int c = 0; c < nc; ++c) {
    tran_scratch s;
    scratch_alloc(o, &s, o->c[c].isadac > 0);
    orc_caltran(o, c, dt, niter, &s);
    scratch_free(&s);
  }
  orc_vertical_diffusion(o, dt);
  /* dye decay / age, EFDC/calconc.f90:581-616 */
  for (int c = 0; c < nc; ++c) {
    if (!o->c[c].is_dye) continue;
    const double rk = o->c[c].rkdye;
    double *C = o->c[c].con;
    if (rk == 1000.0) {
      const double DAGE = dt / 86400.;
      for (int K = 1; K <= o->KC; ++K)
        for (int L = 2; L <= o->LA; ++L)
          if (o->ISDRY == 0 || o->lmaskdry[HX(L)]) C[KX(L, K)] = C[KX(L, K)] + DAGE;
    } else if (rk != 0.0) {
      const double f = rk < 0.0 ? exp(-rk * dt) : 1. / (1. + dt * rk);
      for (int K = 1; K <= o->KC; ++K)
        for (int L = 2; L <= o->LA; ++L)
          if (o->ISDRY == 0 || o->lmaskdry[HX(L)]) C[KX(L, K)] = f * C[KX(L, K)];
    }
  }
  orc_settle(o, dt);
  if (o->isbal) bal2t_open(o, dt);  /* EFDC/calconc.f90:659-664 */
  return EFDC_OK;
}

/* FP64 totals, EFDC/bal2t.f90:90 (volume) and :103-104 (mass) with unit
   masks */
int orc_balance(orc_ctx *o, double *volume, double *mass) {
  double v = 0.;
  for (int L = 2; L <= o->LA; ++L) v = v + A(EFDC_DXYP)[HX(L)] * A(EFDC_HP)[HX(L)];
  *volume = v;
  for (int c = 0; c < o->nconst; ++c) {
    double m = 0.;
    for (int K = 1; K <= o->KC; ++K)
      for (int L = 2; L <= o->LA; ++L)
        m = m + A(EFDC_DXYP)[HX(L)] * A(EFDC_HP)[HX(L)] *
                    o->c[c].con[KX(L, K)] * A(EFDC_DZC)[KX(L, K)];
    mass[c] = m;
  }
  return EFDC_OK;
}

/* SETOPENBC followed by SETOPENBC2 on the arrays as they are (no
   min-diagonal reduction); used to compare with the reference units. */
int orc_setopenbc_apply(orc_ctx *o, double dt, int64_t n) {
  orc_setopenbc(o, dt, n, 0.5 * dt, 1. / dt);
  orc_setopenbc2(o);
  return EFDC_OK;
}

/* CALTRAN of one constituent alone (with CALFQC), no vertical diffusion */
int orc_caltran_only(orc_ctx *o, int c, double dt, int64_t niter) {
  tran_scratch s;
  scratch_alloc(o, &s, o->c[c].isadac > 0);
  orc_caltran(o, c, dt, niter, &s);
  scratch_free(&s);
  return EFDC_OK;
}

/* Depth-uniform closure standing in for the internal mode: Manning drag
   on the external velocity before CALPUV2C; layer transports and layer
   depths after it. */
int orc_closure_pre(orc_ctx *o, const double *nman) {
  const double G = o->G;
  double *TBX = A(EFDC_TBX), *TBY = A(EFDC_TBY), *UHE = A(EFDC_UHE),
         *VHE = A(EFDC_VHE), *HU = A(EFDC_HU), *HV = A(EFDC_HV);
  for (int L = 2; L <= o->LA; ++L) {
    const size_t i = HX(L);
    const double n2 = nman[i] * nman[i];
    const double u = UHE[i] / HU[i], v = VHE[i] / HV[i];
    TBX[i] = G * n2 * u * fabs(u) / cbrt(HU[i]);
    TBY[i] = G * n2 * v * fabs(v) / cbrt(HV[i]);
  }
  return EFDC_OK;
}

int orc_closure_post(orc_ctx *o) {
  double *UHDYE = A(EFDC_UHDYE), *UHDY1E = A(EFDC_UHDY1E),
         *VHDXE = A(EFDC_VHDXE), *VHDX1E = A(EFDC_VHDX1E), *HP = A(EFDC_HP),
         *DZC = A(EFDC_DZC), *UHDY2 = A(EFDC_UHDY2), *VHDX2 = A(EFDC_VHDX2),
         *HPK = A(EFDC_HPK), *HPKI = A(EFDC_HPKI), *U2 = A(EFDC_U2),
         *V2 = A(EFDC_V2), *HU = A(EFDC_HU), *HV = A(EFDC_HV),
         *DYIU = A(EFDC_DYIU), *DXIV = A(EFDC_DXIV);
  for (int K = 1; K <= o->KC; ++K)
    for (int L = 2; L <= o->LA; ++L) {
      const size_t i = HX(L), j = KX(L, K);
      const double ub = 0.5 * (UHDYE[i] + UHDY1E[i]);
      const double vb = 0.5 * (VHDXE[i] + VHDX1E[i]);
      UHDY2[j] = ub * DZC[j];
      VHDX2[j] = vb * DZC[j];
      U2[j] = ub * DYIU[i] / HU[i];
      V2[j] = vb * DXIV[i] / HV[i];
      HPK[j] = HP[i] * DZC[j];
      HPKI[j] = 1. / HPK[j];
    }
  return EFDC_OK;
}

/* ===================================================================== */
/* CALUVW(ISTL = 2), sigma layers (IGRIDV = 0, KSZ = 1), with the steps of   */
/* HDMT2T around it: EFDC/hdmt2t.f90:513-527 (advance the internal            */
/* variables), EFDC/caluvw.f90:176-1172, EFDC/hdmt2t.f90:638-655 (U at V and  */
/* V at U after the step, when uv_after is set).  No waves, no subgrid        */
/* channels, no groundwater, 2TL with ISDYNSTP = 0 (DELT = DT).               */
/* ===================================================================== */

/* Face constants of the U faces (v = 0) or V faces (v = 1) of cell L, as
   EFDC/aaefdc.f90:2001-2126 sets SGZU, DZGU, CDZFU, CDZUU, CDZLU, CDZRU,
   CDZDU, CDZMU and EFDC/caluvw.f90:82-113 sets CERRU (each array KC long,
   index K-1). */
typedef struct {
  double sg[64], dzg[64], cdzf[64], cdzu[64], cdzl[64], cdzr[64], cdzd[64],
      cdzm[64], cerr[64];
} face_consts;

static void face_constants(const orc_ctx *o, int L, int v, face_consts *f) {
  const int KC = o->KC, KS = KC - 1;
  const size_t i = HX(L);
  const double *DZC = A(EFDC_DZC);
  const double open = v ? A(EFDC_SVBO)[i] : A(EFDC_SUBO)[i];
  const int Lb = v ? o->LSC[i] : o->LWC[i];
  memset(f, 0, sizeof(*f));
  for (int K = 1; K <= KC; ++K)
    if (open > 0.) f->sg[K - 1] = dmax(DZC[KX(Lb, K)], DZC[KX(L, K)]);
  for (int K = 1; K <= KS; ++K) {
    const double a = f->sg[K - 1], b = f->sg[K];
    f->dzg[K - 1] = 0.5 * (a + b);
    if (open > 0.) {
      f->cdzf[K - 1] = a * b / (a + b);
      f->cdzu[K - 1] = -a / (a + b);
      f->cdzl[K - 1] = -b / (a + b);
    }
  }
  if (open > 0.) {
    f->cdzr[0] = f->sg[0] - 1.;
    f->cdzd[0] = f->sg[0];
    for (int K = 2; K <= KS; ++K) {
      f->cdzr[K - 1] = f->cdzr[K - 2] + f->sg[K - 1];
      f->cdzd[K - 1] = f->cdzd[K - 2] + f->sg[K - 1];
    }
  }
  for (int K = 1; K <= KS; ++K)
    if (open > 0.) {
      f->cdzr[K - 1] = f->cdzr[K - 1] * f->dzg[K - 1] * f->cdzl[0];
      f->cdzm[K - 1] = 0.5 * f->sg[K - 1] * f->sg[K];
    }
  if (KC > 1) {
    double q1 = 0.;
    for (int K = 1; K <= KC; ++K) q1 = q1 + 1. / (1. - f->sg[K - 1]);
    if (q1 > 0.)
      for (int K = 1; K <= KC; ++K) {
        f->cerr[K - 1] = (1. / (1. - f->sg[K - 1])) / q1;
        f->cerr[K - 1] = f->cerr[K - 1] * (double)KC;
      }
  } else {
    f->cerr[KC - 1] = 1.;
  }
}

/* is L one of the pressure open-boundary cells (LOBCS) */
static int is_lobcs(const orc_ctx *o, int L) {
  for (int s = 0; s < 4; ++s)
    for (int q = 0; q < o->pbc[s].n; ++q)
      if (o->pbc[s].L[q] == L) return 1;
  return 0;
}

/* ISPB of index LL of a pressure side (0 past its end) */
static int ispb_at(const orc_ctx *o, int side, int LL) {
  return LL < o->pbc[side].n ? o->pbc[side].ispb[LL] : 0;
}

int orc_caluvw(orc_ctx *o, double dt, const efdc_uvw_opts_t *opt) {
  const int LA = o->LA, LCM = o->LCM, KC = o->KC, KS = KC - 1;
  const size_t LC = (size_t)LA + 1;
  const double DELT = dt, DELTI = 1. / DELT;
  double *UHDYF = A(EFDC_UHDYF), *VHDXF = A(EFDC_VHDXF), *UHDY = A(EFDC_UHDY),
         *VHDX = A(EFDC_VHDX), *UHDYF1 = A(EFDC_UHDYF1), *VHDXF1 = A(EFDC_VHDXF1),
         *UHDY1 = A(EFDC_UHDY1), *VHDX1 = A(EFDC_VHDX1), *U = A(EFDC_U), *V = A(EFDC_V),
         *U1 = A(EFDC_U1), *V1 = A(EFDC_V1), *UHDYF2 = A(EFDC_UHDYF2),
         *VHDXF2 = A(EFDC_VHDXF2), *UHDYEK = A(EFDC_UHDYEK), *VHDXEK = A(EFDC_VHDXEK),
         *UHDY1EK = A(EFDC_UHDY1EK), *VHDX1EK = A(EFDC_VHDX1EK), *DU = A(EFDC_DU),
         *DV = A(EFDC_DV), *AVUI = A(EFDC_AVUI), *AVVI = A(EFDC_AVVI), *W = A(EFDC_W),
         *W1 = A(EFDC_W1), *STBX = A(EFDC_STBX), *STBY = A(EFDC_STBY), *VU = A(EFDC_VU),
         *UV = A(EFDC_UV), *V1U = A(EFDC_V1U), *U1V = A(EFDC_U1V), *DYU = A(EFDC_DYU),
         *DXV = A(EFDC_DXV);
  double *UHDYE = A(EFDC_UHDYE), *VHDXE = A(EFDC_VHDXE), *UHDY1E = A(EFDC_UHDY1E),
         *VHDX1E = A(EFDC_VHDX1E), *UHE = A(EFDC_UHE), *VHE = A(EFDC_VHE),
         *HU = A(EFDC_HU), *HV = A(EFDC_HV), *HUI = A(EFDC_HUI), *HVI = A(EFDC_HVI),
         *H1U = A(EFDC_H1U), *H1V = A(EFDC_H1V), *SUB = A(EFDC_SUB), *SVB = A(EFDC_SVB),
         *DYIU = A(EFDC_DYIU), *DXIV = A(EFDC_DXIV), *DXYIP = A(EFDC_DXYIP),
         *QSUME = A(EFDC_QSUME), *HP = A(EFDC_HP), *H1P = A(EFDC_H1P), *HPI = A(EFDC_HPI),
         *P = A(EFDC_P), *BELV = A(EFDC_BELV), *TSX = A(EFDC_TSX), *TSY = A(EFDC_TSY);
  double *QSUM = A(EFDC_QSUM), *DZC = A(EFDC_DZC), *SUB3D = A(EFDC_SUB3D),
         *SVB3D = A(EFDC_SVB3D), *UHDY2 = A(EFDC_UHDY2), *VHDX2 = A(EFDC_VHDX2),
         *U2 = A(EFDC_U2), *V2 = A(EFDC_V2), *HPK = A(EFDC_HPK), *HPKI = A(EFDC_HPKI),
         *W2 = A(EFDC_W2);
  const int32_t *LWC = o->LWC, *LEC = o->LEC, *LSC = o->LSC, *LNC = o->LNC;
  int status = EFDC_OK;
  if (KC > 64) return EFDC_ERR_ARG;
  /* wet cells: LKWET (all L = 2..LA with ISDRY = 0; LMASKDRY otherwise,
     EFDC/calpuv2c.f90:1061-1079) */
#define WETC(L) (o->ISDRY == 0 || o->lmaskdry[HX(L)])
  /* SUB3D after the drying of CALPUV2C: SUB*SUB3DO (:982-990) */
#define S3U(L, K) (o->ISDRY > 0 ? SUB[HX(L)] * SUB3D[KX(L, K)] : SUB3D[KX(L, K)])
#define S3V(L, K) (o->ISDRY > 0 ? SVB[HX(L)] * SVB3D[KX(L, K)] : SVB3D[KX(L, K)])

  /* the internal variables advance, EFDC/hdmt2t.f90:513-527 */
  for (int K = 1; K <= KC; ++K)
    for (int L = 2; L <= LA; ++L) {
      if (!WETC(L)) continue;
      const size_t j = KX(L, K);
      UHDYF1[j] = UHDYF[j];
      VHDXF1[j] = VHDXF[j];
      UHDY1[j] = UHDY[j];
      VHDX1[j] = VHDX[j];
      U1[j] = U[j];
      V1[j] = V[j];
      W1[WX(L, K)] = W[WX(L, K)];
    }
  /* CALEXP2T's internal shears with FCAX = FBBX = FX = 0, from the levels
     of the step start (EFDC/calexp2t.f90:1332-1343) */
  if (opt->shear && KC > 1) {
    for (int L = 2; L <= LA; ++L) {
      if (!WETC(L)) continue;
      face_consts fu, fv;
      face_constants(o, L, 0, &fu);
      face_constants(o, L, 1, &fv);
      const size_t i = HX(L);
      for (int K = 1; K <= KS; ++K) {
        DU[KX(L, K)] = fu.cdzf[K - 1] * (H1U[i] * (U1[KX(L, K + 1)] - U1[KX(L, K)]) * DELTI +
                                         0.0);
        DV[KX(L, K)] = fv.cdzf[K - 1] * (H1V[i] * (V1[KX(L, K + 1)] - V1[KX(L, K)]) * DELTI +
                                         0.0);
      }
      if (opt->wind) {
        DU[KX(L, KS)] = DU[KX(L, KS)] - fu.cdzu[KS - 1] * TSX[i];
        DV[KX(L, KS)] = DV[KX(L, KS)] - fv.cdzu[KS - 1] * TSY[i];
      }
    }
  }
  /* KC = 1: HDMT2T sets the single layer from the external flow
     (EFDC/hdmt2t.f90:537-553) */
  if (KC == 1)
    for (int L = 2; L <= LA; ++L) {
      const size_t i = HX(L), j = KX(L, 1);
      UHDYF[j] = UHDYE[i];
      UHDY[j] = UHDYE[i];
      U[j] = UHDYE[i] * HUI[i] * DYIU[i];
      VHDXF[j] = VHDXE[i];
      VHDX[j] = VHDXE[i];
      V[j] = VHDXE[i] * HVI[i] * DXIV[i];
      W[WX(L, 1)] = 0.;
    }
  /* zero the dry cells (:186-263).  The reference zeroes the newly dry
     ones; a cell that stays dry keeps the zeros (no loop writes it), so
     zeroing every dry cell gives the same state. */
  if (o->ISDRY > 0)
    for (int L = 2; L <= LA; ++L) {
      if (WETC(L)) continue;
      const size_t i = HX(L), e = HX(LEC[i]), n = HX(LNC[i]);
      for (int K = 1; K <= KC; ++K) {
        const size_t j = KX(L, K), je = e + (size_t)(K - 1) * LCM,
                     jn = n + (size_t)(K - 1) * LCM;
        UHDY[j] = VHDX[j] = UHDYF[j] = VHDXF[j] = U[j] = V[j] = 0.;
        W[WX(L, K)] = 0.;
        UHDY1[j] = VHDX1[j] = UHDYF1[j] = VHDXF1[j] = U1[j] = V1[j] = 0.;
        W1[WX(L, K)] = 0.;
        UHDY2[j] = VHDX2[j] = UHDYF2[j] = VHDXF2[j] = U2[j] = V2[j] = 0.;
        W2[WX(L, K)] = 0.;
        UHDY[je] = 0.;
        VHDX[jn] = 0.;
        UHDYF[je] = 0.;
        VHDXF[jn] = 0.;
        V[jn] = 0.;
        U[je] = 0.;
        UHDY1[je] = 0.;
        VHDX1[jn] = 0.;
        UHDYF1[je] = 0.;
        VHDXF1[jn] = 0.;
        U1[je] = 0.;
        V1[jn] = 0.;
        UHDY2[je] = 0.;
        VHDX2[jn] = 0.;
        UHDYF2[je] = 0.;
        VHDXF2[jn] = 0.;
        U2[je] = 0.;
        V2[jn] = 0.;
      }
    }
  double *TVARE = calloc(LCM, sizeof(double)), *TVARN = calloc(LCM, sizeof(double));
  if (KC > 1) {
    double *RCXX = calloc(LCM, sizeof(double)), *RCYY = calloc(LCM, sizeof(double));
    double *CU1 = calloc((size_t)LCM * KC, sizeof(double)),
           *CU2 = calloc((size_t)LCM * KC, sizeof(double)),
           *UUU = calloc((size_t)LCM * KC, sizeof(double)),
           *VVV = calloc((size_t)LCM * KC, sizeof(double));
    double *AAU = calloc(LCM, sizeof(double)), *AAV = calloc(LCM, sizeof(double)),
           *BBU = calloc(LCM, sizeof(double)), *BBV = calloc(LCM, sizeof(double));
    const double AVCON1 = opt->AVCON1;
    /* bottom drag coefficients, ISTL = 2 (:296-330) */
    for (int L = 2; L <= LA; ++L) {
      if (!WETC(L)) continue;
      const size_t i = HX(L), j1 = KX(L, 1);
      double Q1 = sqrt(U1[j1] * U1[j1] + V1U[i] * V1U[i]);
      double Q2 = sqrt(U[j1] * U[j1] + VU[i] * VU[i]);
      double qq = Q1 * Q2;
      if (AVCON1 < 0.00001)
        RCXX[i] = STBX[i] * sqrt(qq);
      else
        RCXX[i] = AVCON1 / sqrt(H1U[i] * HU[i]) + STBX[i] * sqrt(qq);
      Q1 = sqrt(V1[j1] * V1[j1] + U1V[i] * U1V[i]);
      Q2 = sqrt(V[j1] * V[j1] + UV[i] * UV[i]);
      qq = Q1 * Q2;
      if (AVCON1 < 0.00001)
        RCYY[i] = STBY[i] * sqrt(qq);
      else
        RCYY[i] = AVCON1 / sqrt(H1V[i] * HV[i]) + STBY[i] * sqrt(qq);
    }
    /* the external flows split by layer (:332-345) */
    for (int K = 1; K <= KS; ++K)
      for (int L = 2; L <= LA; ++L) {
        if (!WETC(L)) continue;
        const size_t i = HX(L), j = KX(L, K);
        UHDYEK[j] = DZC[j] * (UHDYE[i] - UHDYE[HX(LEC[i])]);
        VHDXEK[j] = DZC[j] * (VHDXE[i] - VHDXE[HX(LNC[i])]);
      }
    /* the U and V shears (:415-524), then the layer flows (:527-567),
       their correction to the external solution (:571-611) and the
       velocities (:613-647), one wet cell at a time */
    for (int L = 2; L <= LA; ++L) {
      if (!WETC(L)) continue;
      const size_t i = HX(L);
      face_consts fu, fv;
      face_constants(o, L, 0, &fu);
      face_constants(o, L, 1, &fv);
      for (int K = 1; K <= KS; ++K) {
        const size_t j = KX(L, K), jm = KX(L, K - 1 > 0 ? K - 1 : 1);
        const double cl_u = fu.cdzl[K - 1], cu_u = fu.cdzu[K - 1];
        DU[j] = S3U(L, K) * DU[j];
        if (K > 1) {
          const double RCDZL = cl_u, RCDZM = fu.cdzm[K - 1] * DELTI, RCDZU = cu_u;
          const double CMU = 1. + RCDZM * HU[i] * AVUI[j];
          const double EU = 1. / (CMU - RCDZL * CU1[jm]);
          CU1[j] = RCDZU * EU;
          DU[j] = (DU[j] - RCDZL * DU[jm]) * EU;
          UUU[j] = -RCDZL * UUU[jm] * EU;
        } else {
          const double RCDZL = cl_u, RCDZM = fu.cdzm[K - 1] * DELTI, RCDZU = cu_u;
          const double CMU = 1. + RCDZM * HU[i] * AVUI[j];
          const double EU = 1. / CMU;
          CU1[j] = RCDZU * EU;
          DU[j] = (DU[j] - RCDZL * RCXX[i] * UHE[i] * HUI[i]) * EU;
          UUU[j] = EU;
        }
        DV[j] = S3V(L, K) * DV[j];
        if (K > 1) {
          const double RCDZL = fv.cdzl[K - 1], RCDZM = fv.cdzm[K - 1] * DELTI,
                       RCDZU = fv.cdzu[K - 1];
          const double CMV = 1. + RCDZM * HV[i] * AVVI[j];
          const double EV = 1. / (CMV - RCDZL * CU2[jm]);
          CU2[j] = RCDZU * EV;
          DV[j] = (DV[j] - RCDZL * DV[jm]) * EV;
          VVV[j] = -RCDZL * VVV[jm] * EV;
        } else {
          const double RCDZL = fv.cdzl[K - 1], RCDZM = fv.cdzm[K - 1] * DELTI,
                       RCDZU = fv.cdzu[K - 1];
          const double CMV = 1. + RCDZM * HV[i] * AVVI[j];
          const double EV = 1. / CMV;
          CU2[j] = RCDZU * EV;
          DV[j] = (DV[j] - RCDZL * RCYY[i] * VHE[i] * HVI[i]) * EV;
          VVV[j] = EV;
        }
      }
      for (int K = KS - 1; K >= 1; --K) {
        const size_t j = KX(L, K), jp = KX(L, K + 1);
        DU[j] = DU[j] - CU1[j] * DU[jp];
        DV[j] = DV[j] - CU2[j] * DV[jp];
        UUU[j] = UUU[j] - CU1[j] * UUU[jp];
        VVV[j] = VVV[j] - CU2[j] * VVV[jp];
      }
      AAU[i] = 0.;
      AAV[i] = 0.;
      BBU[i] = 1.;
      BBV[i] = 1.;
      for (int K = 1; K <= KS; ++K) {
        const size_t j = KX(L, K);
        const double CRU = fu.cdzr[K - 1] * RCXX[i] * AVUI[j];
        AAU[i] = AAU[i] + CRU * DU[j];
        BBU[i] = BBU[i] + CRU * UUU[j];
        const double CRV = fv.cdzr[K - 1] * RCYY[i] * AVVI[j];
        AAV[i] = AAV[i] + CRV * DV[j];
        BBV[i] = BBV[i] + CRV * VVV[j];
      }
      AAU[i] = PLANT(21, 0., AAU[i] / BBU[i]);
      AAV[i] = PLANT(21, 0., AAV[i] / BBV[i]);
      for (int K = 1; K <= KS; ++K) {
        const size_t j = KX(L, K);
        DU[j] = S3U(L, K) * fu.dzg[K - 1] * HU[i] * AVUI[j] * (DU[j] - AAU[i] * UUU[j]);
        DV[j] = S3V(L, K) * fv.dzg[K - 1] * HV[i] * AVVI[j] * (DV[j] - AAV[i] * VVV[j]);
      }
      /* layer flows (:527-567) */
      for (int K = 1; K <= KS; ++K) {
        UHE[i] = UHE[i] + fu.cdzd[K - 1] * DU[KX(L, K)];
        VHE[i] = VHE[i] + fv.cdzd[K - 1] * DV[KX(L, K)];
      }
      UHDYF[KX(L, KC)] = UHE[i] * SUB[i];
      VHDXF[KX(L, KC)] = VHE[i] * SVB[i];
      for (int K = KS; K >= 1; --K) {
        UHDYF[KX(L, K)] = S3U(L, K) * (UHDYF[KX(L, K + 1)] - DU[KX(L, K)]);
        VHDXF[KX(L, K)] = S3V(L, K) * (VHDXF[KX(L, K + 1)] - DV[KX(L, K)]);
      }
      for (int K = 1; K <= KC; ++K) {
        UHDYF[KX(L, K)] = UHDYF[KX(L, K)] * DYU[i];
        VHDXF[KX(L, K)] = VHDXF[KX(L, K)] * DXV[i];
      }
      /* correction to the external solution (:571-611) */
      TVARE[i] = 0.;
      TVARN[i] = 0.;
      for (int K = 1; K <= KC; ++K) {
        TVARE[i] = TVARE[i] + UHDYF[KX(L, K)] * fu.sg[K - 1];
        TVARN[i] = TVARN[i] + VHDXF[KX(L, K)] * fv.sg[K - 1];
      }
      TVARE[i] = TVARE[i] - UHDYE[i];
      TVARN[i] = TVARN[i] - VHDXE[i];
      for (int K = 1; K <= KC; ++K) {
        const size_t j = KX(L, K);
        UHDYF[j] = PLANT(20, UHDYF[j], UHDYF[j] - TVARE[i] * fu.cerr[K - 1]);
        UHDY[j] = UHDYF[j] * fu.sg[K - 1];
        VHDXF[j] = PLANT(20, VHDXF[j], VHDXF[j] - TVARN[i] * fv.cerr[K - 1]);
        VHDX[j] = VHDXF[j] * fv.sg[K - 1];
      }
      /* velocities (:613-647) */
      UHE[i] = 0.;
      VHE[i] = 0.;
      for (int K = 1; K <= KC; ++K) {
        const size_t j = KX(L, K);
        UHE[i] = UHE[i] + UHDYF[j] * fu.sg[K - 1];
        VHE[i] = VHE[i] + VHDXF[j] * fv.sg[K - 1];
        U[j] = UHDYF[j] * HUI[i];
        V[j] = VHDXF[j] * HVI[i];
      }
      for (int K = 1; K <= KC; ++K) {
        U[KX(L, K)] = U[KX(L, K)] * DYIU[i];
        V[KX(L, K)] = V[KX(L, K)] * DXIV[i];
      }
      UHE[i] = UHE[i] * DYIU[i];
      VHE[i] = VHE[i] * DXIV[i];
    }
    /* W from the layer continuity, ISTL = 2 (:665-683) */
    for (int K = 1; K <= KS; ++K)
      for (int L = 2; L <= LA; ++L) {
        if (!WETC(L)) continue;
        const size_t i = HX(L), j = KX(L, K), jw = WX(L, K), jwm = WX(L, K - 1);
        const size_t je = HX(LEC[i]) + (size_t)(K - 1) * LCM,
                     jn = HX(LNC[i]) + (size_t)(K - 1) * LCM;
        const double ek = PLANT(22, 0., UHDYEK[j]), e1k = UHDY1EK[j];
        const double vk = PLANT(22, 0., VHDXEK[j]), v1k = VHDX1EK[j];
        W[jw] = W[jwm] -
                PLANT(23, DXYIP[i], 0.5 * DXYIP[i]) *
                    (UHDY[je] - UHDY[j] + ek + UHDY1[je] - UHDY1[j] + e1k + VHDX[jn] -
                     VHDX[j] + vk + VHDX1[jn] - VHDX1[j] + v1k) +
                (QSUM[j] - DZC[j] * QSUME[i]) * DXYIP[i];
      }
    /* W on the open-boundary cells (:687-694) */
    for (int s = 0; s < 4; ++s)
      for (int q = 0; q < o->pbc[s].n; ++q)
        for (int K = 1; K <= KS; ++K) W[WX(o->pbc[s].L[q], K)] = 0.;
    free(RCXX);
    free(RCYY);
    free(CU1);
    free(CU2);
    free(UUU);
    free(VVV);
    free(AAU);
    free(AAV);
    free(BBU);
    free(BBV);
  }
  /* label 30: U and V on the open boundaries (:700-767), sides S, W, E, N
     with the concentration-BC cell lists */
  {
    face_consts f;
    for (int K = 1; K <= KC; ++K) {
      const orc_cbc_side *cs = &o->cbc[0];   /* south */
      for (int LL = 0; LL < cs->n; ++LL) {
        if (ispb_at(o, 2, LL) == 2) continue;
        const int L = cs->L[LL], LN = LNC[HX(L)], LNN = LNC[HX(LN)];
        if ((size_t)LN != LC) {
          face_constants(o, LN, 1, &f);
          const size_t jn = KX(LN, K);
          VHDXF[jn] = VHDXF[KX(LNN, K)] - VHDXE[HX(LNN)] + VHDXE[HX(LN)];
          VHDX[jn] = VHDXF[jn] * f.sg[K - 1];
          V[jn] = VHDXF[jn] / (HV[HX(LN)] * DXV[HX(LN)]);
        }
        W[WX(LN, K)] = 0.;
      }
    }
    for (int K = 1; K <= KC; ++K) {
      const orc_cbc_side *cs = &o->cbc[1];   /* west */
      for (int LL = 0; LL < cs->n; ++LL) {
        if (ispb_at(o, 0, LL) == 2) continue;
        const int L = cs->L[LL], LE = LEC[HX(L)], L1P = LEC[HX(LE)];
        if ((size_t)LE != LC) {
          face_constants(o, LE, 0, &f);
          const size_t je = KX(LE, K);
          UHDYF[je] = UHDYF[KX(L1P, K)] - UHDYE[HX(L1P)] + UHDYE[HX(LE)];
          UHDY[je] = UHDYF[je] * f.sg[K - 1];
          U[je] = UHDYF[je] / (HU[HX(LE)] * DYU[HX(LE)]);
        }
        W[WX(LE, K)] = 0.;
      }
    }
    for (int K = 1; K <= KC; ++K) {
      const orc_cbc_side *cs = &o->cbc[2];   /* east */
      for (int LL = 0; LL < cs->n; ++LL) {
        const int L = cs->L[LL], LW = LWC[HX(L)];
        if (ispb_at(o, 1, LL) != 2) {
          face_constants(o, L, 0, &f);
          const size_t j = KX(L, K);
          UHDYF[j] = UHDYF[KX(LW, K)] - UHDYE[HX(LW)] + UHDYE[HX(L)];
          UHDY[j] = UHDYF[j] * f.sg[K - 1];
          U[j] = UHDYF[j] / (HU[HX(L)] * DYU[HX(L)]);
        }
        W[WX(L, K)] = 0.;
      }
    }
    for (int K = 1; K <= KC; ++K) {
      const orc_cbc_side *cs = &o->cbc[3];   /* north */
      for (int LL = 0; LL < cs->n; ++LL) {
        const int L = cs->L[LL];
        if (ispb_at(o, 3, LL) != 2) {
          const int LS = LSC[HX(L)];
          face_constants(o, L, 1, &f);
          const size_t j = KX(L, K);
          VHDXF[j] = VHDXF[KX(LS, K)] - VHDXE[HX(LS)] + VHDXE[HX(L)];
          VHDX[j] = VHDXF[j] * f.sg[K - 1];
          V[j] = VHDXF[j] / (HV[HX(L)] * DXV[HX(L)]);
        }
        W[WX(L, K)] = 0.;
      }
    }
  }
  /* averages of the two levels, ISTL = 2 (:770-791) */
  for (int K = 1; K <= KC; ++K)
    for (int L = 2; L <= LA; ++L) {
      if (!WETC(L)) continue;
      const size_t j = KX(L, K);
      UHDYF2[j] = 0.5 * (UHDYF[j] + UHDYF1[j]);
      VHDXF2[j] = 0.5 * (VHDXF[j] + VHDXF1[j]);
      UHDY2[j] = 0.5 * (UHDY[j] + UHDY1[j]);
      VHDX2[j] = 0.5 * (VHDX[j] + VHDX1[j]);
      U2[j] = 0.5 * (U[j] + U1[j]);
      V2[j] = 0.5 * (V[j] + V1[j]);
      W2[WX(L, K)] = 0.5 * (W[WX(L, K)] + W1[WX(L, K)]);
    }
  /* the depth from the layer flows, 2TL (:813-1040) */
  if (KC > 1) {
    for (int L = 2; L <= LA; ++L) {
      if (!WETC(L)) continue;
      const size_t i = HX(L);
      TVARE[i] = 0.;
      TVARN[i] = 0.;
      for (int K = 1; K <= KC; ++K) {
        TVARE[i] = TVARE[i] + UHDY2[KX(L, K)];
        TVARN[i] = TVARN[i] + VHDX2[KX(L, K)];
      }
    }
    for (int L = 2; L <= LA; ++L) {
      if (!WETC(L)) continue;
      const size_t i = HX(L), e = HX(LEC[i]), n = HX(LNC[i]);
      double HPPTMP = H1P[i] + DELT * DXYIP[i] *
                                   (QSUME[i] + TVARE[i] + PLANT(24, TVARE[e], -TVARE[e]) +
                                    TVARN[i] - TVARN[n]);
      if (HPPTMP <= 0. && HP[i] > 0.) {
        if (is_lobcs(o, L)) continue;
        if (HP[i] <= o->HDRY && o->ISDRY > 0) {
          HPPTMP = HP[i];
          UHDYE[i] = 0.;
          VHDXE[i] = 0.;
          for (int K = 1; K <= KC; ++K) {
            const size_t j = KX(L, K), je = e + (size_t)(K - 1) * LCM,
                         jn = n + (size_t)(K - 1) * LCM;
            UHDYF[j] = VHDXF[j] = UHDY[j] = VHDX[j] = U[j] = V[j] = 0.;
            W[WX(L, K)] = 0.;
            UHDYF[je] = 0.;
            VHDXF[jn] = 0.;
            UHDY[je] = 0.;
            VHDX[jn] = 0.;
            U[je] = 0.;
            V[jn] = 0.;
            UHDY2[j] = VHDX2[j] = U2[j] = V2[j] = 0.;
            W2[WX(L, K)] = 0.;
            UHDY2[je] = 0.;
            VHDX2[jn] = 0.;
            U2[je] = 0.;
            V2[jn] = 0.;
          }
        } else {
          status = EFDC_ERR_NEGDEPTH;
        }
      }
      HP[i] = HPPTMP;
      HPI[i] = 1. / HP[i];
    }
  }
  /* open-boundary cells take HP from P (:1130-1134) */
  for (int s = 0; s < 4; ++s)
    for (int q = 0; q < o->pbc[s].n; ++q) {
      const size_t i = HX(o->pbc[s].L[q]);
      HP[i] = (1. / o->G) * P[i] - BELV[i];
      HPI[i] = 1. / HP[i];
    }
  /* layer thicknesses (:1161-1172) */
  for (int K = 1; K <= KC; ++K)
    for (int L = 2; L <= LA; ++L) {
      if (!WETC(L)) continue;
      const size_t j = KX(L, K);
      HPK[j] = HP[HX(L)] * DZC[j];
      HPKI[j] = 1. / HPK[j];
    }
  /* U at V and V at U at the new level, EFDC/hdmt2t.f90:638-655 */
  if (opt->uv_after) {
    for (int L = 2; L <= LA; ++L) {
      if (!WETC(L)) continue;
      const size_t i = HX(L);
      const int LSE = o->LSEC[i], LS = LSC[i], LE = LEC[i], LNW = o->LNWC[i], LW = LWC[i],
                LN = LNC[i];
      U1V[i] = UV[i];
      V1U[i] = VU[i];
      UV[i] = ((HP[HX(LSE)] * U[KX(LSE, 1)] + HP[HX(LS)] * U[KX(LS, 1)]) +
               (HP[HX(LE)] * U[KX(LE, 1)] + HP[i] * U[KX(L, 1)])) /
              (1.0 + SUB[HX(LE)] + SUB[HX(LS)] + SUB[HX(LSE)]) * HVI[i];
      VU[i] = ((HP[HX(LNW)] * V[KX(LNW, 1)] + HP[HX(LW)] * V[KX(LW, 1)]) +
               (HP[HX(LN)] * V[KX(LN, 1)] + HP[i] * V[KX(L, 1)])) /
              (1.0 + SVB[HX(LN)] + SVB[HX(LW)] + SVB[HX(LNW)]) * HUI[i];
    }
  }
  free(TVARE);
  free(TVARN);
#undef WETC
#undef S3U
#undef S3V
  (void)UHDY1E;
  (void)VHDX1E;
  return status;
}

/* ===================================================================== */
/* Two-level balances (ISBAL >= 1): BAL2T1 totals, the source terms of      */
/* BAL2T3A and the open-boundary fluxes of BAL2T2 for the volume and every  */
/* constituent, the totals of BAL2T5 (EFDC/bal2t.f90).  The balance region  */
/* leaves out the open-boundary cells: SPB = 0 on pressure-BC cells and     */
/* SCB = 0 on concentration-BC cells (EFDC/setbcs.f90:180-239,952).         */
/* ===================================================================== */
static void bal2t_masks(const orc_ctx *o, double *spb, double *scb) {
  for (int L = 1; L <= o->LCM; ++L) spb[HX(L)] = scb[HX(L)] = 1.0;
  for (int s = 0; s < 4; ++s)
    for (int q = 0; q < o->pbc[s].n; ++q) spb[HX(o->pbc[s].L[q])] = PLANT(32, 1.0, 0.0);
  for (int s = 0; s < 4; ++s)
    for (int q = 0; q < o->cbc[s].n; ++q) scb[HX(o->cbc[s].L[q])] = 0.0;
}

/* VOLBEG2T / VOLEND2T (:86,939) and the constituent masses (:93-96,
   945-948): K outer, L inner, in the reference's order */
int orc_bal2t_totals(orc_ctx *o, double *volume, double *mass) {
  double *spb = malloc(sizeof(double) * o->LCM), *scb = malloc(sizeof(double) * o->LCM);
  bal2t_masks(o, spb, scb);
  const double *DXYP = A(EFDC_DXYP), *HP = A(EFDC_HP), *DZC = A(EFDC_DZC);
  double v = 0.;
  for (int L = 2; L <= o->LA; ++L) v = v + spb[HX(L)] * DXYP[HX(L)] * HP[HX(L)];
  *volume = v;
  for (int c = 0; c < o->nconst; ++c) {
    const double *C = o->c[c].con;
    double m = 0.;
    for (int K = 1; K <= o->KC; ++K)
      for (int L = 2; L <= o->LA; ++L)
        m = m + scb[HX(L)] * DXYP[HX(L)] * HP[HX(L)] * C[KX(L, K)] * DZC[KX(L, K)];
    mass[c] = m;
  }
  free(spb);
  free(scb);
  return EFDC_OK;
}

/* BAL2T1 (:8-176): the totals, and the outflow totals zeroed */
int orc_bal2t_begin(orc_ctx *o, double *volume, double *mass) {
  o->bal_volout = 0.;
  for (int c = 0; c < ORC_MAXC; ++c) o->bal_massout[c] = 0.;
  return orc_bal2t_totals(o, volume, mass);
}

/* BAL2T3A (:316-692): the volume the sources add, QSUME - QDWASTE over
   every cell (:332-336), and the mass of each constituent they add, the
   inflow at CQS / CSERT and the withdrawal at the cell's concentration
   (:360-366 for salinity, the same for every constituent); called before
   the transport, so C is the level at the step start */
static void bal2t_sources(orc_ctx *o, double dt) {
  const double DELT = dt;
  const double *QSUME = A(EFDC_QSUME);
  for (int L = 2; L <= o->LA; ++L) {
    const double qd = (o->NDRYSTP < 0 && o->qdwaste) ? o->qdwaste[HX(L)] : 0.;
    o->bal_volout = o->bal_volout - DELT * (QSUME[HX(L)] - qd);
  }
  const int KC = o->KC, nc = o->nconst;
  for (int c = 0; c < nc; ++c) {
    const double *C = o->c[c].con;
    double out = o->bal_massout[c];
    for (int q = 0; q < o->nqsij; ++q) {
      const int L = o->lqs[q];
      for (int K = 1; K <= KC; ++K) {
        const double qss = o->qss[q * KC + K - 1], qser = o->qsercell[q * KC + K - 1];
        const double cqs = o->cqs[(size_t)(q * KC + K - 1) * nc + c];
        const double cser = o->csert_q[(size_t)(q * KC + K - 1) * nc + c];
        const double ck = C[KX(L, K)];
        out = out - DELT * dmax(qss, 0.) * cqs - PLANT(31, 0., DELT * dmin(qss, 0.) * ck) -
              DELT * dmax(qser, 0.) * cser - DELT * dmin(qser, 0.) * ck;
      }
    }
    o->bal_massout[c] = out;
  }
}

/* BAL2T2 (:179-315): the volume and constituent mass across the faces of
   the concentration-BC cells, sides S, W, E, N, layer by layer; the mass
   is carried at the upwind cell's concentration after the transport */
static void bal2t_open(orc_ctx *o, double dt) {
  const double DELT = dt;
  const int KC = o->KC, nc = o->nconst;
  const double *UHDY2 = A(EFDC_UHDY2), *VHDX2 = A(EFDC_VHDX2);
  double vol = o->bal_volout;
  double mo[ORC_MAXC];
  for (int c = 0; c < nc; ++c) mo[c] = o->bal_massout[c];
  for (int K = 1; K <= KC; ++K)       /* south: the face of LN (:193-214) */
    for (int q = 0; q < o->cbc[0].n; ++q) {
      const int L = o->cbc[0].L[q], LN = o->LNC[HX(L)];
      const double f = VHDX2[KX(LN, K)];
      vol = vol - DELT * f;
      for (int c = 0; c < nc; ++c)
        mo[c] = mo[c] - DELT * dmin(f, 0.) * o->c[c].con[KX(LN, K)] -
                DELT * dmax(f, 0.) * o->c[c].con[KX(L, K)];
    }
  for (int K = 1; K <= KC; ++K)       /* west: the face of LEC(L) (:217-238) */
    for (int q = 0; q < o->cbc[1].n; ++q) {
      const int L = o->cbc[1].L[q], LE = o->LEC[HX(L)];
      const double f = UHDY2[KX(LE, K)];
      vol = PLANT(30, vol + DELT * f, vol - DELT * f);
      for (int c = 0; c < nc; ++c)
        mo[c] = mo[c] - DELT * dmin(f, 0.) * o->c[c].con[KX(LE, K)] -
                DELT * dmax(f, 0.) * o->c[c].con[KX(L, K)];
    }
  for (int K = 1; K <= KC; ++K)       /* east: the face of L (:241-262) */
    for (int q = 0; q < o->cbc[2].n; ++q) {
      const int L = o->cbc[2].L[q], LW = o->LWC[HX(L)];
      const double f = UHDY2[KX(L, K)];
      vol = vol + DELT * f;
      for (int c = 0; c < nc; ++c)
        mo[c] = mo[c] + DELT * dmin(f, 0.) * o->c[c].con[KX(L, K)] +
                DELT * dmax(f, 0.) * o->c[c].con[KX(LW, K)];
    }
  for (int K = 1; K <= KC; ++K)       /* north: the face of L (:265-287) */
    for (int q = 0; q < o->cbc[3].n; ++q) {
      const int L = o->cbc[3].L[q], LS = o->LSC[HX(L)];
      const double f = VHDX2[KX(L, K)];
      vol = vol + DELT * f;
      for (int c = 0; c < nc; ++c)
        mo[c] = mo[c] + DELT * dmin(f, 0.) * o->c[c].con[KX(L, K)] +
                DELT * dmax(f, 0.) * o->c[c].con[KX(LS, K)];
    }
  o->bal_volout = vol;
  for (int c = 0; c < nc; ++c) o->bal_massout[c] = mo[c];
}
