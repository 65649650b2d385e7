/*
 * efdc_oracle.h -- CPU restatement (FP64, plain C + OpenMP) of the EFDC+
 * 8.5 hot path, used ONLY as the checker in tests/, as the smoke check and
 * as the CPU baseline of bench.py.  Nothing in the product links it.
 *
 * Arrays are the var_global_mod arrays in Fortran layout (see
 * include/efdc_hip.h); they are addressed through the same efdc_array_id
 * values as the C-ABI.
 */
#ifndef EFDC_ORACLE_H
#define EFDC_ORACLE_H
#include <stdint.h>
#include "../include/efdc_hip.h"

#define ORC_NARR 256
#define ORC_MAXC 16

typedef struct {
  int32_t n;
  const int32_t *L, *ispb;
  const double *pser, *pcb, *psb;
} orc_pbc_side;

typedef struct {
  int32_t n;
  const int32_t *L;
  const double *cb;      /* [(LL*2+j)*nc + c] */
  const double *csert;   /* [(LL*KC+K-1)*nc + c] */
  const int32_t *ntscr;
  double *clo;           /* CLO*(LL,K,M) memory: [(LL*KC+K-1)*nc + c] */
  int64_t *nlo;          /* NLO*(LL,K,M) memory, same indexing           */
} orc_cbc_side;

typedef struct {
  int32_t mo, isadac, isfct, is_dye;
  double rkdye;
  double *con, *con1;    /* (LCM,KCM) */
  double wset, probdep, wese;  /* cohesive settling (0: none) */
} orc_const;

typedef struct {
  int32_t LA, LCM, KC, KCM;
  const int32_t *LWC, *LEC, *LSC, *LNC, *LNWC, *LSEC;
  double *a[ORC_NARR];   /* indexed by efdc_array_id */
  /* params */
  double G, RITB1, SNLT, HDRY, TCON_TBEGIN, RSQM;
  double CG_REL_TOL;     /* > 0: RSQM = CG_REL_TOL^2 * sum FPTMP^2 */
  int32_t ITERM;
  int32_t nthreads;      /* OpenMP threads (reference threading structure) */
  /* pressure open BCs: W, E, S, N */
  orc_pbc_side pbc[4];
  int32_t mtide;
  const double *tcp;
  int32_t *lopenbcdry;   /* (LCM) LOPENBCDRY */
  /* concentration open BCs: S, W, E, N */
  orc_cbc_side cbc[4];
  const double *wtci;    /* [(K-1)*2 + j] */
  /* sources */
  int32_t nqsij;
  const int32_t *lqs;
  const double *qss, *qsercell, *cqs, *csert_q;
  /* constituents */
  int32_t nconst;
  orc_const c[ORC_MAXC];
  /* wetting and drying (ISDRY = 0: off; 1..97: EFDC/calpuv2c.f90:574-602;
     98: open-boundary drying only; 99: rewetting, :604-716) */
  int32_t ISDRY;
  int32_t *iscdry;       /* (LCM) ISCDRY */
  int32_t *lmaskdry;     /* (LCM) LMASKDRY, 1 = wet */
  double *sub1, *svb1;   /* (LCM) SUB1, SVB1: face switches at the step start */
  int32_t ncordry;       /* re-solves of the last CALPUV2C */
  /* volume wasting of isolated wet cells, NDRYSTP < 0 (:838-893) */
  int32_t NDRYSTP;
  int32_t *natdry;       /* (LCM) NATDRY: steps a wet cell has been isolated */
  double *qdwaste;       /* (LCM) QDWASTE of the last step (m^3/s) */
  double *vdwaste;       /* (LCM) VDWASTE accumulated (m^3) */
  /* two-level balances of a period (ISBAL >= 1, BAL2T1..BAL2T5): while
     isbal is set CALCONC adds BAL2T3A before and BAL2T2 after its
     transport to the outflow totals */
  int32_t isbal;
  double bal_volout;
  double bal_massout[ORC_MAXC];
  /* ISITB >= 1 or ISVEG > 0: FUHDYE/FVHDXE scaled by RCX/RCY after the
     explicit terms (EFDC/calpuv2c.f90:223-243) */
  int32_t IMPLICIT_DRAG;
} orc_ctx;

typedef struct {
  int32_t iters;
  double rsq;
  double ccmnm;
  int32_t status;
} orc_cg_out;

/* CONGRAD on the scaled system held in a[]: EFDC/congrad.f90:54-187 */
int orc_congrad(orc_ctx *o, orc_cg_out *out);
/* CALPUV2C, 2TL: EFDC/calpuv2c.f90:128-571,574-939 (drying with ISDRY > 0) */
int orc_calpuv2c(orc_ctx *o, double dt, int64_t n, orc_cg_out *out);
/* Only the assembly part (explicit stage .. scaling), for unit tests */
int orc_assemble(orc_ctx *o, double dt, int64_t n, double *ccmnm);
/* CALCONC(2,1): EFDC/calconc.f90:65-473,581-616 with CALTRAN/CALFQC */
int orc_calconc(orc_ctx *o, double dt, int64_t niter);
/* cohesive settling of the water column (part of orc_calconc),
   EFDC/calsed.f90:232-256,308-312 */
void orc_settle(orc_ctx *o, double dtsed);
/* only the vertical diffusion pass, EFDC/calconc.f90:239-473 */
int orc_vertical_diffusion(orc_ctx *o, double dt);
/* balance totals: EFDC/bal2t.f90:90,103-104 */
int orc_balance(orc_ctx *o, double *volume, double *mass);
/* BAL2T1 totals with the SPB/SCB masks (EFDC/bal2t.f90:82-105), used at the
   start (BAL2T1, which also zeroes the outflow totals) and the end
   (BAL2T5, :935-958) of a balance period */
int orc_bal2t_totals(orc_ctx *o, double *volume, double *mass);
int orc_bal2t_begin(orc_ctx *o, double *volume, double *mass);
/* SETOPENBC + SETOPENBC2 alone, and CALTRAN (+CALFQC) of one constituent
   alone: the units compared with the reference build in oracle/_ref */
int orc_setopenbc_apply(orc_ctx *o, double dt, int64_t n);
int orc_caltran_only(orc_ctx *o, int c, double dt, int64_t niter);
/* CALUVW(ISTL=2) on sigma layers with the HDMT2T steps around it
   (EFDC/hdmt2t.f90:513-527,638-655, EFDC/caluvw.f90:176-1172); the
   internal-mode arrays must be bound */
int orc_caluvw(orc_ctx *o, double dt, const efdc_uvw_opts_t *opt);
/* depth-uniform closure standing in for the internal mode */
/* On-disk formats (efdc_oracle_io.c): RESTOUT's text into a malloc'ed
   buffer (released with orc_free), HEADEROUT / WSOUT / VELOUT's EE_WS.OUT
   and EE_VEL.OUT, and the E17.8 edit descriptor */
int orc_restart_text(orc_ctx *o, const efdc_restart_t *r, char **text, int64_t *bytes);
void orc_free(void *p);
int orc_ee_open(orc_ctx *o, const char *dir, const efdc_ee_grid_t *g);
int orc_ee_append(orc_ctx *o, const char *dir, const efdc_ee_grid_t *g, int32_t n,
                  double eetime, double delt, int32_t lminstep);
void orc_fmt_e17_8(double v, char out[17]);
int orc_closure_pre(orc_ctx *o, const double *manning_n);
int orc_closure_post(orc_ctx *o);

#endif
