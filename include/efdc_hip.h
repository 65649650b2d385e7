/*
 * efdc_hip.h -- C-ABI of the MI355X (gfx950) implementation of EFDC+'s
 * per-timestep external-mode solve and constituent transport.
 *
 * Host arrays follow the var_global_mod layout of EFDC+ 8.5:
 *   - horizontal arrays are (LCM), element L (1-based) at offset L-1;
 *   - layer arrays are (LCM,KCM) column-major, element (L,K) at
 *     (L-1) + (K-1)*LCM;
 *   - W2 and DZIG are (LCM,0:KCM), element (L,K) at (L-1) + K*LCM;
 *   - wet cells are L = 2..LA, ghosts are L = 1 and L = LC = LA+1;
 *   - neighbour maps hold 1-based L values and point to LC for land.
 *   (EFDC/var_global_mod.f90:1083-1114, EFDC/varalloc.f90:54,839,1993-2002,
 *    EFDC/cellmap.f90:12-148)
 *
 * All host arrays are FP64 (the ReleaseDP64 build: /real_size:64,
 * EFDC/EFDCPlus_085_OMP.vfproj:48).  The library keeps its own device copy
 * of every array in HBM; efdc_set_* / efdc_get_* move data across the
 * boundary, every compute entry point works on the device copy only.
 *
 * Every entry point returns an efdc_status (0 = success).  The text of the
 * last error of a context is available from efdc_last_error().  A context
 * owns one HIP stream; calls on one context must not overlap.
 */
#ifndef EFDC_HIP_H
#define EFDC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EFDC_ABI_VERSION 1

/* ---------------------------------------------------------------- status */
typedef enum {
  EFDC_OK = 0,
  EFDC_ERR_ARG = 1,        /* invalid argument or shape                       */
  EFDC_ERR_HIP = 2,        /* a HIP runtime call failed                        */
  EFDC_ERR_CG_MAXITER = 3, /* CONGRAD: RSQ > RSQM after ITERM iterations
                              (EFDC/congrad.f90:171-184)                       */
  EFDC_ERR_NEGDEPTH = 4,   /* NEGDEP: HP <= 0 after the solve
                              (EFDC/negdep.f90:34-45)                          */
  EFDC_ERR_STATE = 5,      /* call out of order (e.g. no open-BC table yet)    */
  EFDC_ERR_NOMEM = 6,      /* device allocation failed                         */
  EFDC_ERR_NODEVICE = 7,   /* no HIP device visible                            */
  EFDC_ERR_IO = 8          /* a file could not be opened or written            */
} efdc_status;

/* ------------------------------------------------------------- precision */
typedef enum {
  /* Plain FP64 Jacobi-PCG, the method of EFDC/congrad.f90:1-191. */
  EFDC_CG_FP64 = 0,
  /* FP32 coefficients and SpMV, FP64 dot products and alpha/beta, wrapped
     in an FP64 residual-based iterative refinement of P.  No reference
     counterpart (EFDC/congrad.f90:20-21 runs the whole CG in the build's
     REAL kind). */
  EFDC_CG_MIXED = 1
} efdc_cg_mode;

typedef enum {
  EFDC_CG_KERNEL_AUTO = 0,       /* persistent solve when the rows fit one
                                    co-resident grid, otherwise multi-kernel */
  EFDC_CG_KERNEL_MULTI = 1,      /* two launches per iteration               */
  EFDC_CG_KERNEL_PERSISTENT = 2  /* one launch per solve of a co-resident
                                    grid (efdc_set_cg_launch); an error
                                    when the rows do not fit                 */
} efdc_cg_kernel;

typedef enum {
  EFDC_STORE_FP64 = 0, /* constituents stored as FP64 on the device          */
  EFDC_STORE_FP32 = 1  /* constituents stored as FP32, tridiagonal in FP64   */
} efdc_store;

/* ------------------------------------------------------------------ grid */
typedef struct {
  int32_t LA;   /* last wet cell index; wet cells are L = 2..LA             */
  int32_t LCM;  /* leading dimension of host (LCM) arrays, >= LA+1          */
  int32_t KC;   /* number of sigma layers (K = 1 bottom .. KC top)          */
  int32_t KCM;  /* second dimension of host (LCM,KCM) arrays, >= KC         */
  const int32_t *LWC, *LEC, *LSC, *LNC; /* (LCM) neighbour maps             */
  const int32_t *LNWC, *LSEC;           /* (LCM) diagonal maps; may be NULL
                                           when no constituent uses
                                           anti-diffusion (ISADAC=0)          */
} efdc_grid_t;

/* ---------------------------------------------------------------- params */
typedef struct {
  double G;        /* gravity (EFDC GLOBAL G)                               */
  double RITB1;    /* explicit bottom-stress weight (EFDC/caltbxy.f90:94-106)*/
  double SNLT;     /* nonlinear-term switch multiplying FXE/FYE              */
  double HDRY;     /* drying depth: the open-boundary inactivation test
                      (EFDC/setopenbc.f90:62-74) and, with ISDRY > 0, the
                      drying of open-boundary cells (EFDC/calpuv2c.f90:
                      497-530), cell drying (ISDRY 1..97, :574-602),
                      drying and rewetting (99, :604-716), the draining of
                      isolated cells (NDRYSTP < 0, :838-893) and the
                      transport mask LMASKDRY (:916-939)                   */
  double TCON_TBEGIN; /* TCON*TBEGIN, offset of the tidal clock             */
  int32_t ITERM;   /* CG iteration cap                                      */
  double RSQM;     /* CG stop: sum r^2 <= RSQM on the scaled system          */
  double cg_rel_tol; /* if > 0: stop when ||r||_2 <= cg_rel_tol*||FPTMP||_2   */
  int32_t cg_mode;   /* efdc_cg_mode                                        */
  int32_t store;     /* efdc_store, transport storage precision             */
  double mixed_inner_rtol; /* EFDC_CG_MIXED: inner FP32 relative reduction  */
  int32_t mixed_max_outer; /* EFDC_CG_MIXED: refinement step cap            */
  int32_t cg_kernel;       /* efdc_cg_kernel: how the solve is launched     */
  int32_t mixed_precond_degree; /* EFDC_CG_MIXED, persistent solve: terms of
                              the workgroup-local Neumann preconditioner
                              (1 = none; 0 = the library's default)          */
  int32_t tran_shape;      /* CALCONC column kernel launch shape: 0 = the
                              library's default or the autotuner's choice;
                              1, 2 or 4 = that many constituents per thread
                              (needs mutual neighbour maps), + 16 = at most
                              168 VGPRs, + 32 (1 or 2 only) = that and the
                              transport fields loaded two layers ahead;
                              -1, -2 or -4 = one constituent per thread at a
                              time, that many waves per workgroup.  Every
                              shape gives the same bits.                      */
  int32_t ISDRY;           /* wetting and drying: 0 off; 1..97 the drying of
                              EFDC/calpuv2c.f90:497-530,574-602 with the
                              re-solve loop and the transport mask LMASKDRY
                              (:916-939); 98 open-boundary drying only; 99
                              drying and rewetting (:604-716, needs mutual
                              neighbour maps)                               */
  int32_t NDRYSTP;         /* with ISDRY > 0 and NDRYSTP < 0: a wet cell
                              isolated by dry neighbours for more than
                              -NDRYSTP steps is drained to 0.9 HDRY when
                              it is shallow (< 2 HDRY) or above a
                              neighbour's bed (:838-893, 1010-1018); >= 0:
                              no draining                                   */
  int32_t implicit_drag;   /* 1: FUHDYE and FVHDXE are scaled by the host's
                              RCX and RCY after the explicit terms, the
                              implicit bottom and vegetation drag of ISITB
                              >= 1 or ISVEG > 0 (EFDC/calpuv2c.f90:223-243);
                              0: RCX/RCY enter only the coupling and the
                              transport update (:329-336, :426)             */
} efdc_params_t;

/* Fill *p with the values EFDC uses by default: G=9.81, RITB1=1, SNLT=0,
   HDRY=0.2, TCON_TBEGIN=0, ITERM=200, RSQM=1e-9 (the sample decks,
   SampleModels/Lake_Washington/efdc.inp:92), cg_rel_tol=0, FP64 CG, FP64
   storage, mixed_inner_rtol=1e-5, mixed_max_outer=30, cg_kernel=AUTO,
   mixed_precond_degree=0, tran_shape=0, ISDRY=0, NDRYSTP=0. */
void efdc_params_default(efdc_params_t *p);

/* ------------------------------------------------------------- array ids */
typedef enum {
  /* (LCM) arrays: geometry (EFDC/aaefdc.f90:2172-2185) */
  EFDC_DXYP = 0, EFDC_DXYIP, EFDC_DXYU, EFDC_DXYV,
  EFDC_DXIU, EFDC_DYIU, EFDC_DXIV, EFDC_DYIV,
  EFDC_HRUO, EFDC_HRVO, EFDC_SUB, EFDC_SVB, EFDC_BELV,
  /* (LCM) external-mode state (EFDC/calpuv2c.f90) */
  EFDC_P, EFDC_P1, EFDC_HP, EFDC_H1P, EFDC_H2P,
  EFDC_HU, EFDC_HV, EFDC_HUI, EFDC_HVI, EFDC_HPI,
  EFDC_H1U, EFDC_H1V, EFDC_H1UI, EFDC_H1VI,
  EFDC_UHDYE, EFDC_VHDXE, EFDC_UHDY1E, EFDC_VHDX1E, EFDC_UHE, EFDC_VHE,
  /* (LCM) forcings produced upstream of CALPUV2C */
  EFDC_TSX, EFDC_TSY, EFDC_TBX, EFDC_TBY, EFDC_FCAXE, EFDC_FCAYE,
  EFDC_FPGXE, EFDC_FPGYE, EFDC_FXE, EFDC_FYE, EFDC_QSUME, EFDC_RCX, EFDC_RCY,
  /* (LCM) linear system (EFDC/calpuv2c.f90:214-398) */
  EFDC_FUHDYE, EFDC_FVHDXE, EFDC_FP, EFDC_FP1,
  EFDC_CC, EFDC_CS, EFDC_CW, EFDC_CE, EFDC_CN,
  EFDC_CCC, EFDC_CCS, EFDC_CCW, EFDC_CCE, EFDC_CCN, EFDC_CCCI, EFDC_FPTMP,
  /* (LCM) face switches of the full wet grid, SUBO/SVBO: with ISDRY > 0,
     SUB/SVB are reset to them at each step start after FUHDYE/FVHDXE
     (EFDC/calpuv2c.f90:246-253), which reopens the faces that dried */
  EFDC_SUBO, EFDC_SVBO,
  EFDC_NUM_H,
  /* (LCM,KCM) arrays (EFDC/varalloc.f90:54,118-120,611,773,803,1959-1996) */
  EFDC_HPK = 64, EFDC_H1PK, EFDC_HPKI, EFDC_UHDY2, EFDC_VHDX2,
  EFDC_CDZKK, EFDC_CDZKMK, EFDC_AB, EFDC_QSUM, EFDC_U2, EFDC_V2,
  EFDC_SUB3D, EFDC_SVB3D, EFDC_DZC,
  EFDC_NUM_K_END,
  /* (LCM,0:KCM) arrays (EFDC/varalloc.f90:199,839) */
  EFDC_W2 = 96, EFDC_DZIG,
  EFDC_NUM_W_END,
  /* Internal mode (CALUVW, EFDC/caluvw.f90, with the steps of HDMT2T around
     it, EFDC/hdmt2t.f90:513-527,638-655).  FP64 on the device in every
     storage mode; allocated by the first efdc_set_array/efdc_caluvw that
     needs them (zero until set).  (LCM) arrays (EFDC/varalloc.f90:196,
     693-694,834-835,...): bottom-drag coefficients, V at U points and U at
     V points at the two time levels, face widths DYU and DXV: */
  EFDC_STBX = 128, EFDC_STBY, EFDC_VU, EFDC_UV, EFDC_V1U, EFDC_U1V,
  EFDC_DYU, EFDC_DXV,
  EFDC_NUM_IH_END,
  /* (LCM,KCM) arrays (EFDC/varalloc.f90:76,170,769-790,2001-2003; KSM =
     KCM): layer flows, velocities, their previous level and averages,
     the external flow split per layer, the shear forcing DU/DV of
     CALEXP2T and the inverse face viscosities AVUI/AVVI of CALAVB: */
  EFDC_UHDYF = 160, EFDC_VHDXF, EFDC_UHDY, EFDC_VHDX, EFDC_UHDYF1, EFDC_VHDXF1,
  EFDC_UHDY1, EFDC_VHDX1, EFDC_U, EFDC_V, EFDC_U1, EFDC_V1, EFDC_UHDYF2,
  EFDC_VHDXF2, EFDC_UHDYEK, EFDC_VHDXEK, EFDC_UHDY1EK, EFDC_VHDX1EK,
  EFDC_DU, EFDC_DV, EFDC_AVUI, EFDC_AVVI,
  EFDC_NUM_IK_END,
  /* (LCM,0:KCM) arrays (EFDC/varalloc.f90:837-838): vertical velocity */
  EFDC_W = 192, EFDC_W1,
  EFDC_NUM_IW_END
} efdc_array_id;

/* ------------------------------------------------------------ open BCs  */
/* Pressure open boundaries, one table per side in the order W, E, S, N
   (LPBW/LPBE/LPBS/LPBN, EFDC/setopenbc.f90:35-206).  For cell LL of a side:
     ispb[LL]        ISPB* (0 = clamped; 1,2 = radiation)
     pser[LL]        PSERT+0.5*PSERZDF+PSERST+0.5*PSERZDS, already blended
                     with the second series when NPFORT>=1
                     (EFDC/setopenbc.f90:44-48); the time-series part of the
                     boundary value, in P units (m^2/s^2)
     pcb[LL*mtide+M], psb[...]  PCB*(LL,M), PSB*(LL,M) (P units)            */
typedef struct {
  int32_t n;
  const int32_t *L;
  const int32_t *ispb;
  const double *pser;
  const double *pcb;
  const double *psb;
} efdc_pbc_side_t;

typedef struct {
  efdc_pbc_side_t side[4]; /* 0=W 1=E 2=S 3=N */
  int32_t mtide;           /* MTIDE */
  const double *tcp;       /* TCP(M), tidal periods in seconds */
} efdc_open_bc_t;

/* Concentration open boundaries for transport (LCBS/LCBW/LCBE/LCBN,
   EFDC/caltran.f90:316-471), sides in the order S, W, E, N as the
   reference applies them.  For each side and each transported constituent
   c (0..nconst-1, as registered by efdc_add_constituent):
     cb[(LL*2 + j)*nconst + c]          CB*(LL,j+1,M), j = 0,1
     csert[(LL*KC + (K-1))*nconst + c]  CSERT(K,NCSER*(LL,M),M) this step
     ntscr[LL]                          NTSCR*(LL)                          */
typedef struct {
  int32_t n;
  const int32_t *L;
  const double *cb;
  const double *csert;
  const int32_t *ntscr;
} efdc_cbc_side_t;

typedef struct {
  efdc_cbc_side_t side[4]; /* 0=S 1=W 2=E 3=N */
  const double *wtci;      /* WTCI(K,2): wtci[(K-1)*2 + j] */
} efdc_conc_bc_t;

/* Flow sources (CALFQC, EFDC/calfqc.f90:209-225): for source NS
   (0-based s), at cell LQS:
     qss[s*KC + K-1], qsercell[s*KC + K-1]          QSS(K,NS), QSERCELL(K,NS)
     cqs[(s*KC + K-1)*nconst + c], csert[...]      CQS(K,NS,M),
                                                   CSERT(K,NCSERQ(NS,M),M)  */
typedef struct {
  int32_t n;
  const int32_t *LQS;
  const double *qss;
  const double *qsercell;
  const double *cqs;
  const double *csert;
} efdc_sources_t;

/* ---------------------------------------------------------- constituents */
typedef struct {
  int32_t mo;      /* reference index MO: 1=SAL, 2=TEM, 3=DYE, >=4 others
                      (only MO==1 gets the bottom-value outflow clamp,
                      EFDC/caltran.f90:336-342)                             */
  int32_t isadac;  /* ISADAC(MVAR): 0 off, 1 anti-diffusion                  */
  int32_t isfct;   /* ISFCT(MVAR):  0 off, 1 flux-corrected limiter         */
  double rkdye;    /* dye decay RKDYE (0 none, 1000 age of water), only
                      applied when is_dye != 0 (EFDC/calconc.f90:581-616)    */
  int32_t is_dye;
  /* Cohesive settling (CALSED water column, EFDC/calsed.f90:232-256 and the
     bottom-layer update :308-316), applied after the transport of every
     CALCONC with DTSED = DT (SEDSTEP <= DT) and a constant settling speed
     (ISEDVW = 0).  wset = WSEDO(NS) in m/s (0: no settling).  The bed model
     that sets the deposition probability and the resuspension flux from the
     bed shear is not part of this library: probdep (0..1) and wese (mass
     per m^2 per s, >= 0) are given, and the bed never runs out.            */
  double wset;
  double probdep;
  double wese;
} efdc_constituent_t;

/* ------------------------------------------------------------------ step */
typedef struct {
  double dt;      /* DT (2TL, ISDYNSTP=0)                                   */
  int64_t n;      /* step counter N (tidal phase TN = DT*N + TCON*TBEGIN)   */
  int64_t niter;  /* NITER, the clock of the open-BC recovery ramps         */
} efdc_step_t;

typedef struct {
  int32_t iters;        /* CG iterations of the last solve (IDRYTBP)        */
  int32_t outer;        /* refinement steps (EFDC_CG_MIXED), else 1         */
  double rsq;           /* final sum r^2 on the scaled system (FP64)        */
  double rsqm_used;     /* the threshold that was applied                   */
  double ccmnm;         /* min diagonal CCMNM used for scaling              */
} efdc_cg_stats_t;

typedef struct {
  double volume;          /* sum DXYP*HP over L=2..LA (BAL2T1 VOLBEG2T form,
                             EFDC/bal2t.f90:90)                              */
  double mass[16];        /* per constituent: sum DXYP*HP*C*DZC
                             (EFDC/bal2t.f90:103-104)                        */
  int32_t nconst;
} efdc_balance_t;

typedef struct efdc_ctx efdc_ctx_t;

/* --------------------------------------------------------------- entries */
int efdc_abi_version(void);

/* Create a context on the current HIP device; copies and validates the
   maps (replaces the SAVE allocation blocks of EFDC/calpuv2c.f90:59-126 and
   EFDC/congrad.f90:34-49). */
int efdc_ctx_create(const efdc_grid_t *grid, const efdc_params_t *params,
                    efdc_ctx_t **out);
int efdc_ctx_destroy(efdc_ctx_t *ctx);
/* Text of the last error of a context; with ctx == NULL, the reason the
   last efdc_ctx_create or efdc_congrad_noalloc on the calling thread
   failed. */
const char *efdc_last_error(const efdc_ctx_t *ctx);
int efdc_set_params(efdc_ctx_t *ctx, const efdc_params_t *params);

/* Host <-> device copies of one var_global array (full extent, ghosts
   included: L = 1..LA+1, K = 1..KC or 0..KC). */
int efdc_set_array(efdc_ctx_t *ctx, int32_t id, const double *host);
int efdc_get_array(efdc_ctx_t *ctx, int32_t id, double *host);

int efdc_set_open_bc(efdc_ctx_t *ctx, const efdc_open_bc_t *bc);
int efdc_set_conc_bc(efdc_ctx_t *ctx, const efdc_conc_bc_t *bc);
int efdc_set_sources(efdc_ctx_t *ctx, const efdc_sources_t *src);
/* Replace only the source concentrations CQS of the current table (same
   layout as efdc_sources_t.cqs), e.g. a load series advanced every step
   the way CALFQC interpolates CQSER (EFDC/calfqc.f90:42-90). */
int efdc_set_source_cqs(efdc_ctx_t *ctx, const double *cqs);

/* Register a transported constituent (the CALTRAN calls of
   EFDC/calconc.f90:159-228); returns its slot in *slot. */
int efdc_add_constituent(efdc_ctx_t *ctx, const efdc_constituent_t *desc,
                         int32_t *slot);
/* CON and CON1 of one constituent, (LCM,KCM) host layout; con1 may be NULL */
int efdc_set_constituent(efdc_ctx_t *ctx, int32_t slot, const double *con,
                         const double *con1);
int efdc_get_constituent(efdc_ctx_t *ctx, int32_t slot, double *con,
                         double *con1);
/* CON of one constituent at n cells (L in 2..LA, K in 1..KC, Fortran
   indices) into out[n], FP64: the observation operator of a host that
   reads a few cells every few steps (no counterpart in the reference). */
int efdc_get_constituent_at(efdc_ctx_t *ctx, int32_t slot, int32_t n, const int32_t *L,
                            const int32_t *K, double *out);

/* Drying state after the last efdc_calpuv2c (ISDRY > 0): LMASKDRY (1 wet,
   0 dry: left out of the transport, EFDC/caltran.f90 LKWET loops) and
   ISCDRY per cell (L order, LCM entries; either may be NULL) and the
   number of re-solves NCORDRY of that step. */
int efdc_get_dry_state(efdc_ctx_t *ctx, int32_t *lmaskdry, int32_t *iscdry, int32_t *ncordry);

/* Drained volume with NDRYSTP < 0 (L order, LCM entries, any may be NULL):
   QDWASTE, the rate of the last step (m^3/s, 0 where nothing was drained;
   the reference keeps it only with ISBAL > 0), VDWASTE, the volume drained
   since the context was created (m^3), and NATDRY, the steps each wet cell
   has been isolated (EFDC/calpuv2c.f90:838-893). */
int efdc_get_dry_waste(efdc_ctx_t *ctx, double *qdwaste, double *vdwaste, int32_t *natdry);
/* Sets VDWASTE and NATDRY (L order, LCM entries; either may be NULL), e.g.
   from a restart, so that the drained volume and the isolation counters
   (EFDC/calpuv2c.f90:838-893) continue from the host's values instead of
   starting at 0 when the context is created. */
int efdc_set_dry_counters(efdc_ctx_t *ctx, const double *vdwaste, const int32_t *natdry);

/* CONGRAD (EFDC/congrad.f90:1-191) on the scaled system already held in
   CCC, CCS, CCW, CCE, CCN, CCCI, FPTMP, starting from and updating P. */
int efdc_congrad(efdc_ctx_t *ctx, efdc_cg_stats_t *stats);

/* CONGRAD without any device allocation.  No counterpart in the reference:
   it is for callers that cannot allocate device memory (code under HIP
   graph capture, tight serving loops, harnesses that forbid hipMalloc).
   The same FP64 Jacobi-PCG as efdc_congrad with EFDC_CG_KERNEL_MULTI (the
   same kernels, so the same bits), on arrays the caller already holds in
   device memory: the neighbour maps LWC/LEC/LSC/LNC (1-based L values as
   the host keeps them, LC for land) and CCC, CCS, CCW, CCE, CCN, CCCI,
   FPTMP, P, each with at least LA+1 entries in the (LCM) layout; P is the
   initial guess and is overwritten by the solution.  RSQM and ITERM are
   the stop test and the cap (sum r^2 <= RSQM, EFDC/congrad.f90:165).
   Work goes to `stream` (a hipStream_t; NULL: the null stream) and the
   call returns after the solve has finished.
   Scratch: one static device array of EFDC_NOALLOC_SCRATCH_BYTES per GPU
   (held until the process ends), which limits LA+1 to EFDC_NOALLOC_MAX_LCM
   (EFDC_ERR_ARG beyond).  Every call on a GPU shares that GPU's array:
   calls that overlap on one GPU race.  A wet row whose map points outside
   2..LA+1 is EFDC_ERR_ARG; efdc_last_error(NULL) gives the reason of a
   failed call on the calling thread. */
#define EFDC_NOALLOC_MAX_LCM 1048576
#define EFDC_NOALLOC_SCRATCH_BYTES (48u * 1048576u + 65536u)
int efdc_congrad_noalloc(int32_t LA, const int32_t *LWC, const int32_t *LEC,
                         const int32_t *LSC, const int32_t *LNC, const double *CCC,
                         const double *CCS, const double *CCW, const double *CCE,
                         const double *CCN, const double *CCCI, const double *FPTMP,
                         double *P, double RSQM, int32_t ITERM, void *stream,
                         efdc_cg_stats_t *stats);

/* CALPUV2C for the 2TL scheme (EFDC/calpuv2c.f90:128-571,896-939):
   explicit stage, assembly, open boundaries, min-diagonal scaling,
   CONGRAD, face transports, depths, P, HU/HV and inverses; with ISDRY > 0
   the drying passes and the re-solve loop from label 1000 (:302, at most
   500 re-solves, then EFDC_ERR_STATE) and LMASKDRY. */
int efdc_calpuv2c(efdc_ctx_t *ctx, const efdc_step_t *step,
                  efdc_cg_stats_t *stats);

/* CALCONC(ISTL=2, IS2TL=1) for every registered constituent
   (EFDC/calconc.f90:65-473,581-616, EFDC/caltran.f90:1-813,
   EFDC/calfqc.f90:1-225): upwind advection with sources and open
   boundaries, optional anti-diffusion with FCT, implicit vertical diffusion
   and dye decay. */
int efdc_calconc(efdc_ctx_t *ctx, const efdc_step_t *step);

/* FP64 volume and constituent-mass totals of the device state. */
int efdc_balance(efdc_ctx_t *ctx, efdc_balance_t *out);

/* ------------------------------------------------- two-level balances
   The balances of a period (ISBAL >= 1, EFDC/bal2t.f90) for the volume and
   every registered constituent, in FP64 on the device, over the balance
   region the reference uses (the open-boundary cells left out: SPB = 0 on
   pressure-BC cells, SCB = 0 on concentration-BC cells,
   EFDC/setbcs.f90:180-239,952).  efdc_bal2t_begin is BAL2T1 (the totals at
   the start, EFDC/bal2t.f90:8-176); while the period is open every
   efdc_calconc adds BAL2T3A before its transport (the volume of the sources,
   QSUME - QDWASTE, and the mass they bring in or take out, :316-692) and
   BAL2T2 after it (the volume and mass across the faces of the
   concentration-BC cells, at the upwind concentration after the transport,
   :179-315); efdc_bal2t_end is BAL2T5's error accounting (:866-1040):
   the totals at the end, the outflow totals and END - (BEG - OUT).  Not
   accumulated: the momentum and energy terms (UMO, UUE, PPE, BBE), the
   sediment-bed, toxics and hydraulic-structure terms, which need state
   outside this library. */
typedef struct {
  int32_t nconst;         /* constituents reported                           */
  int64_t nsteps;         /* CALCONC calls in the period                     */
  double volbeg, volend;  /* VOLBEG2T, VOLEND2T (m^3)                       */
  double volout;          /* VOLOUT: net volume out of the region            */
  double volerr;          /* VOLERR2T = VOLEND2T - (VOLBEG2T - VOLOUT)       */
  double massbeg[16], massend[16], massout[16], masserr[16]; /* per constituent */
} efdc_bal2t_t;
int efdc_bal2t_begin(efdc_ctx_t *ctx);
int efdc_bal2t_end(efdc_ctx_t *ctx, efdc_bal2t_t *out);

/* ------------------------------------------------------ restart output
   RESTART.OUT as RESTOUT writes it (EFDC/restartmod.f90:17-297, VER 720):
   the line N, TIMEDAY, VER (I20,4X,F12.4,I10), ISCO(0:8) (9I6), then per
   cell L = 2..LA the records of FORMAT 906 (50E17.8): HP H1P HWQ H2WQ BELV;
   UHDYE UHDY1E VHDXE VHDX1E; U, U1, V, V1 (K = 1..KC); QQ, QQ1, QQL, QQL1,
   DML (K = 0..KC); the ice pair when ISTRAN(2) >= 1 and ISICE > 2; SAL and
   SAL1; TEMB with TEM, and TEM1; DYE and DYE1 (each when ISTRAN(M) >= 1 and
   ISCO(M) == 1); per cohesive class SEDB, SED, SEDB1, SED1, then HBED,
   HBED1, VDRBED, VDRBED1; then for SAL, TEM, DYE and every sediment class
   the concentration-BC cells of sides S, W, E, N: NLO (908 = 50I10) after
   NLO = MAX(NLO - N, -NTSCR), as RESTOUT leaves it in memory, and CLO; then
   QSUME with QSUM(K = 1..KC) per cell.  The records the host owns that
   follow (MDCHH channels, groundwater AGWELV, rating-curve controls, :273-
   295) are appended by the host after this call.  The values the library
   holds come from the device (HP, H1P, BELV, the external flows, the
   internal-mode U/U1/V/V1, the constituents CON/CON1, NLO/CLO, QSUME/QSUM);
   the others are host arrays in the var_global layout (L = 1..LCM fastest,
   then K).  Shellfish larvae, toxics and non-cohesive sediment
   (ISTRAN(4), (5), (7) with ISCO == 1) are refused: this library does not
   transport them.  Numbers are formatted as the Fortran runtime writes
   them (correctly rounded, ties to even; exponents of three digits drop
   the E; NaN and Inf right-justified). */
typedef struct {
  int64_t n;                 /* N                                              */
  double timeday;            /* TIMEDAY                                        */
  int32_t isco[9];           /* ISCO(0:8)                                      */
  int32_t istran[9];         /* ISTRAN(0:8), 1..7 read                         */
  int32_t isice;             /* ISICE                                          */
  int32_t slot[3];           /* constituent slots of SAL, TEM, DYE (when written) */
  int32_t nsed;              /* NSCM cohesive classes written (ISTRAN(6))      */
  const int32_t *sed_slot;   /* their constituent slots                        */
  int32_t kb, kbm;           /* KB and the declared bed dimension KBM          */
  int32_t keep_nlo;          /* 1: leave NLO on the device unchanged           */
  const double *hwq, *h2wq;                     /* (LCM)                       */
  const double *qq, *qq1, *qql, *qql1, *dml;    /* (LCM, 0:KC)                 */
  const double *ice_thick;   /* (LCM) the first ice value, ICETHICK + ICEVOL*999.8426/RHOI*DXYIP
                                (+ the frazil ice with ISICE == 4)           */
  const double *ice_temp;    /* (LCM) ICETEMP                                  */
  const double *temb;        /* (LCM) TEMB                                     */
  const double *sedb, *sedb1;                   /* (LCM, KBM, NSCM)            */
  const double *hbed, *hbed1, *vdrbed, *vdrbed1; /* (LCM, KBM)                 */
} efdc_restart_t;
/* The restart text into buf (cap bytes); *bytes = its length (buf == NULL:
   only the length, nothing changes on the device). */
int efdc_restart_text(efdc_ctx_t *ctx, const efdc_restart_t *r, char *buf, int64_t cap,
                      int64_t *bytes);
/* RESTOUT(0): the restart text written to path. */
int efdc_restart_write(efdc_ctx_t *ctx, const efdc_restart_t *r, const char *path);

/* ---------------------------------------------------------- EE linkage
   EE_WS.OUT and EE_VEL.OUT, unformatted stream files (FORM='BINARY').
   efdc_ee_open is HEADEROUT's part for them (EFDC/efdcout.f90:73-113):
   EE_WS.OUT gets VER = 8400, HSIZE, BSIZE, IC, JC, NACTIVE = LA-1 (INT*4);
   EE_VEL.OUT gets VER, HSIZE, BSIZE, IGRIDV, CELL3D, IC, JC, KC, NACTIVE
   and RSSBCE, RSSBCW, RSSBCS, RSSBCN as REAL*4 over L = 1..LCM.
   efdc_ee_append is WSOUT and VELOUT (:371-497) for a run that does not
   continue an earlier file: EE_WS.OUT gets N, EETIME (REAL*8), DELT
   (REAL*4), LMINSTEP and HP(L = 2..LA) as REAL*4; EE_VEL.OUT gets N, EETIME,
   DELT and U, V, W (REAL*4) over K = KSZ(L)..KC for L = 2..LA, each array
   in turn.  N is the caller's N+NRESTART; EETIME = (DT*N + TCON*TBEGIN)/86400
   (:44-45). */
typedef struct {
  int32_t ic, jc, igridv;
  const int32_t *ksz;        /* (LCM) KSZ(L); NULL: 1 in every cell            */
  const double *rssbce, *rssbcw, *rssbcs, *rssbcn;  /* (LCM)                   */
} efdc_ee_grid_t;
int efdc_ee_open(efdc_ctx_t *ctx, const char *dir, const efdc_ee_grid_t *grid);
int efdc_ee_append(efdc_ctx_t *ctx, const char *dir, int32_t n, double eetime, double delt,
                   int32_t lminstep);

/* ---------------------------------------------------------- internal mode
   CALUVW(ISTL=2) on sigma layers (IGRIDV = 0, KSZ = 1), 2TL, no waves and
   no subgrid channels (EFDC/caluvw.f90:176-1172), with the two steps HDMT2T
   takes around it: the advance of the internal variables UHDYF1 = UHDYF,
   UHDY1 = UHDY, U1 = U, W1 = W (and V's) before (EFDC/hdmt2t.f90:513-527)
   and, when uv_after is set, the update of U1V, V1U, UV, VU after
   (:638-655).  Per wet cell: the bottom-drag coefficients RCXX/RCYY
   (caluvw.f90:285-330), the split of the external flow per layer
   UHDYEK/VHDXEK (:332-345), the implicit shear solve (tridiagonal sweep
   plus the Sherman-Morrison correction for the bottom drag, :415-524),
   the layer flows UHDYF/VHDXF and their correction to the external
   solution with CERRU/CERRV (:527-611), U and V (:613-647), W from the
   layer continuity (:649-685) zeroed on open-boundary cells (:687-694),
   the layer flows of the open-boundary faces (:700-767), the averages
   UHDY2/VHDX2/UHDYF2/VHDXF2/U2/V2/W2 (:770-811), the depth from the layer
   flows (:813-1040) with the negative-depth branches, HP of open-boundary
   cells from P (:1130-1134) and HPK/HPKI (:1161-1172).  The face constants
   SGZU, CDZFU/CDZUU/CDZLU/CDZRU/CDZDU/CDZMU, DZGU, CERRU (and V's) follow
   from DZC and SUBO/SVBO as EFDC/aaefdc.f90:2001-2126 and
   EFDC/caluvw.f90:82-113 set them.  The open-boundary faces use the
   concentration-BC cell lists (LCBS/LCBW/LCBE/LCBN) with the radiation type
   ISPB* of the pressure-BC cell of the same index (0 past the end of the
   pressure side), as the reference does.  The next efdc_calconc reads the
   averages and HPK/HPKI this leaves.  Requirements: DZC the same in every
   wet cell of a layer; with ISDRY > 0 the transport mask of the last
   efdc_calpuv2c decides the wet cells.  A non-open-boundary cell whose
   depth from the layer flows is <= 0 while HP > 0 is EFDC_ERR_NEGDEPTH
   (the reference logs it, :1010-1035), unless ISDRY > 0 and HP <= HDRY,
   where the cell keeps HP and its flows are zeroed (:957-1009). */
typedef struct {
  double AVCON1;     /* 2*(1-AVCON)*DZI*AVO (EFDC/aaefdc.f90:645), the
                        molecular part of RCXX/RCYY; below 1e-5 it is left
                        out as the reference does (caluvw.f90:299-330)      */
  int32_t shear;     /* 1: form DU/DV on the device as CALEXP2T does before
                        CALPUV2C when its advective, Coriolis, baroclinic and
                        horizontal-diffusion terms FCAX/FBBX/FX are zero
                        (EFDC/calexp2t.f90:1332-1333): DU(L,K) = CDZFU(L,K)*
                        (H1U(L)*(U1(L,K+1)-U1(L,K))/DT) from the velocities
                        of the step start; 0: DU/DV as the host set them    */
  int32_t wind;      /* with shear = 1: the wind-stress term of NWSER > 0,
                        DU(L,KS) -= CDZUU(L,KS)*TSX(L) (:1337-1343)          */
  int32_t uv_after;  /* 1: update U1V, V1U, UV, VU after the step
                        (EFDC/hdmt2t.f90:638-655; needs LNWC and LSEC)     */
} efdc_uvw_opts_t;
int efdc_caluvw(efdc_ctx_t *ctx, const efdc_step_t *step, const efdc_uvw_opts_t *opts);

/* ------------------------------------------------------------ ensembles */
/* Declare that the wet cells form nseg independent basins: basin s holds
   the cells L = first_L[s] .. first_L[s+1]-1 (the last one up to LA),
   first_L[0] = 2, ascending.  No neighbour map may cross a basin.  From then
   on every entry point advances all basins at once (the pointwise kernels
   run once over all cells), while everything the reference computes per
   model stays per basin: the min-diagonal scaling CCMNM of CALPUV2C
   (EFDC/calpuv2c.f90:360-376), the CG recurrences, stop test and iteration
   cap of CONGRAD (EFDC/congrad.f90:52-187), and the balances.  With the
   FP64 CG, each basin's results are bit for bit those of the same basin in
   a context of its own solved with EFDC_CG_KERNEL_MULTI.  nseg = 1 returns
   to a single model.  No reference counterpart: the batched-ensemble
   addition, for hosts that run many basins. */
int efdc_set_segments(efdc_ctx_t *ctx, int32_t nseg, const int32_t *first_L);

/* CG tile of the basins: each basin's cells are cut into tiles of `cells`
   (256, 512, 1024 or 2048) whose partial sums the basin's reduction adds in
   order.  512, the default, is the tile of the single-model solve, so each
   basin's results are bit for bit those of the basin alone; other sizes
   sum in another order (results within round-off).  cells = 0 times the
   four sizes on the next solve (each from the same P) and keeps the
   fastest; efdc_get_segment_tile gives the size in use and
   efdc_tune_report lists the trial times. */
int efdc_set_segment_tile(efdc_ctx_t *ctx, int32_t cells);
int efdc_get_segment_tile(efdc_ctx_t *ctx, int32_t *cells);

/* How the FP64 multi-kernel solves (one model and basins) issue their
   iterations: 0 launch by launch (the default), 1 by replaying captured
   hipGraphs of 8 and 1 iterations, 2 time both on the next solve from the
   same P and keep the faster (efdc_autotune(ctx, 1) does the same).  The
   two give the same bits; with the library's timers on the solves launch
   kernel by kernel. */
int efdc_set_cg_replay(efdc_ctx_t *ctx, int32_t mode);
int efdc_get_cg_replay(efdc_ctx_t *ctx, int32_t *mode);
/* Form of the mixed persistent inner solve (EFDC_CG_MIXED on one
   co-resident grid; no counterpart in the reference, whose CONGRAD is the
   FP64 Jacobi-PCG): 0 (default) the autotuner chooses, the pipelined
   form until it has run; 1 the Chronopoulos-Gear form, one exchange
   between workgroups per iteration that carries the dot products and three
   halo words; 2 its pipelined form, which waits for the dot products one
   SpMV later and for one halo word from the neighbouring workgroups only.
   Both give the same iterates in exact arithmetic.  get returns 1 or 2,
   the form the next solve runs. */
int efdc_set_cg_pipeline(efdc_ctx_t *ctx, int32_t mode);
int efdc_get_cg_pipeline(efdc_ctx_t *ctx, int32_t *mode);
/* Launch of the persistent solves (no counterpart in the reference).  The
   grid is sized to be co-resident on the GPU.  mode 0 (default): a plain
   launch, which is co-resident whenever no other work holds the CUs; if it
   was not, its bounded waits give up and the solve is repeated with a
   cooperative launch from the P the first one reached (in a CALPUV2C step,
   together with the passes after the solve), so the result is a converged
   solve either way; steps with drying or adjoint recording launch
   cooperatively.  mode 1: always the cooperative launch (guaranteed
   co-residency, about 15 us more per solve).  spin_limit > 0 gives the
   plain launches' waits that many polling rounds instead of the full limit,
   which makes them give up whenever a wait is not met at once (to exercise
   the repetition); 0 restores the full limit.  get returns the mode and
   the number of solves repeated so far. */
int efdc_set_cg_launch(efdc_ctx_t *ctx, int32_t mode, uint32_t spin_limit);
/* Form of the FP64 persistent solve (EFDC_CG_FP64 on one co-resident grid):
   0 (default) the reference's recurrences (EFDC/congrad.f90:91-169), two
   exchanges between workgroups per iteration; 1 the same Jacobi-PCG in the
   pipelined order of Ghysels and Vanroose, one exchange per iteration whose
   wait comes one SpMV later, and one halo value from the neighbouring
   workgroups only.  Same iterates in exact arithmetic and the same stop
   test ((r,r) <= RSQM); the rounding differs, so form 1 is not the
   reference's bits.  No counterpart in the reference. */
int efdc_set_cg64_pipeline(efdc_ctx_t *ctx, int32_t on);
int efdc_get_cg64_pipeline(efdc_ctx_t *ctx, int32_t *on);
int efdc_get_cg_launch(efdc_ctx_t *ctx, int32_t *mode, int32_t *retries);

/* Per-basin results of the last efdc_calpuv2c / efdc_congrad with
   segments: iters[s] and status[s] (0 or EFDC_ERR_CG_MAXITER) for s <
   min(nseg, cap).  The stats of the call itself report the largest
   iteration count and the first failure. */
int efdc_segment_stats(efdc_ctx_t *ctx, int32_t cap, int32_t *iters, int32_t *status);

/* FP64 volume and constituent-mass totals per basin: volume[s] and
   mass[s*nconst + c] for s < min(nseg, cap). */
int efdc_balance_segments(efdc_ctx_t *ctx, int32_t cap, double *volume, double *mass);

/* Achievable HBM bandwidth on this device: best of `reps` timed runs of a
   plain streaming copy of `bytes` (read + write bytes / time, GB/s).  No
   reference counterpart (the roofline reference of the measurements). */
int efdc_stream_copy_gbs(efdc_ctx_t *ctx, size_t bytes, int32_t reps, double *gbs);

/* Autotuner (no reference counterpart).  efdc_autotune(ctx, 1) makes the
   next efdc_calconc time every launch shape of the column transport kernel
   (constituent lanes per workgroup 4/2/1, two register budgets) on that
   step's own data and keep the fastest for the following steps; every shape
   gives the same bits.  efdc_autotune(ctx, 0) returns to the default shape.
   With the mixed CG and the persistent solve, the next efdc_calpuv2c also
   times the solve with preconditioner degrees 1, 2, 3, 4, 6 from the same P
   and keeps the fastest (mixed_precond_degree = 0 only).
   efdc_tune_report lists the candidates timed and the ones kept. */
typedef struct {
  int32_t n;
  struct {
    char name[32];
    float ms;        /* per launch (transport) or per solve (CG)            */
    int32_t chosen;  /* 1 for the candidate kept                            */
  } v[8];
} efdc_tune_report_t;
int efdc_autotune(efdc_ctx_t *ctx, int32_t on);
int efdc_tune_report(efdc_ctx_t *ctx, efdc_tune_report_t *out);

/* Wait for all work queued on the context's stream. */
int efdc_synchronize(efdc_ctx_t *ctx);

/* Per-kernel timers, the device-side counterpart of the reference's phase
   timers TPUV/TCONG/TSADV/TVDIF (DSTIME, EFDC/dstime.f90:1-56,
   EFDC/congrad.f90:52,186, EFDC/calconc.f90:85,230,240,473).  When
   enabled, every kernel launch of the context is bracketed by HIP events
   on the context's stream; efdc_timers_read() waits for the stream,
   returns the count and total duration per kernel and clears the record. */
#define EFDC_TIMER_MAX 64
typedef struct {
  int32_t n;
  struct {
    char name[32];
    int64_t count;
    double total_ms;
  } k[EFDC_TIMER_MAX];
} efdc_timers_t;
int efdc_timers_enable(efdc_ctx_t *ctx, int32_t on);
int efdc_timers_read(efdc_ctx_t *ctx, efdc_timers_t *out);

/* Device address of an array of the context (FP64 device copies have the
   device pitch returned in *pitch elements for (LCM,KCM) arrays). */
int efdc_device_ptr(efdc_ctx_t *ctx, int32_t id, void **ptr, int64_t *pitch);

/* ------------------------------------------ internal-mode stand-in
   The momentum closure and the layer transports that EFDC+ computes
   between CALPUV2C and CALCONC (CALTBXY, CALUVW: EFDC/caltbxy.f90,
   EFDC/caluvw.f90:668-680,770-790,1163-1172) are outside this library.
   For device-resident time stepping without them, a depth-uniform
   stand-in is provided:
     pre : TBX = G n^2 u|u| / HU^(1/3), u = UHE/HU (likewise TBY);
     post: UHDY2 = 0.5*(UHDYE+UHDY1E)*DZC, VHDX2 likewise,
           U2 = 0.5*(UHDYE+UHDY1E)*DYIU/HU, V2 likewise,
           HPK = HP*DZC, HPKI = 1/HPK (W2 untouched).
   These have no reference counterpart. */
int efdc_closure_set_manning(efdc_ctx_t *ctx, const double *n_lcm);
int efdc_closure_pre(efdc_ctx_t *ctx);
int efdc_closure_post(efdc_ctx_t *ctx);

/* ------------------------------------------------------ state snapshots
   No reference counterpart (checkpointing for the adjoint; the reference
   restarts from RESTART.OUT files).  A snapshot holds every array,
   constituent level, open-BC recovery memory and dry-boundary flag of the
   context, so that loading it and repeating a step reproduces the step
   bit for bit.  The buffer is device memory owned by the caller, at least
   efdc_state_bytes() long. */
int efdc_state_bytes(efdc_ctx_t *ctx, size_t *bytes);
int efdc_state_save(efdc_ctx_t *ctx, void *dev_buf);
int efdc_state_load(efdc_ctx_t *ctx, const void *dev_buf);
/* One device buffer of efdc_state_bytes() bytes, for hosts without a
   device allocator of their own (Fortran, Python without torch); while
   the adjoint is enabled, efdc_adj_record_bytes() more, so that a step's
   recording can follow its state at offset efdc_state_bytes(). */
int efdc_state_alloc(efdc_ctx_t *ctx, void **dev_buf);
int efdc_state_free(efdc_ctx_t *ctx, void *dev_buf);

/* --------------------------------------------------- discrete adjoint
   No reference counterpart (gradient addition).  The reverse of one
   hot-path step with the depth-uniform closure, in this order forward:
     efdc_closure_pre (when a Manning field is set), efdc_calpuv2c,
     efdc_closure_post, efdc_calconc.
   The adjoint state is lambda = dJ/d(state after the step) for the
   constituents named in efdc_adj_opts_t and for HP, UHDYE, VHDXE, P, HU,
   HV; efdc_adj_step maps it to dJ/d(state before the step), accumulates
   dJ/dn (Manning, per cell) and leaves dJ/dCQS of that step.  The solve
   A P = FP is reversed by a second CG on the same operator (A is symmetric
   on the cells that are not clamped open boundaries); the settling sweep
   of a constituent with wset > 0 is reversed before its transport, and
   the anti-diffusion / FCT passes (ISADAC, ISFCT) of one before its upwind
   step (their atomic sums can vary in the last bits from run to run).
   Requirements: clamped pressure boundaries (ISPB=0), one basin, mutual
   neighbour maps.
   Use: enable; run the forward step from the state before it (the context
   records the pieces of the step the adjoint reads: the CG solution and
   the old levels of the adjoint constituents); add the misfit gradient of
   that step's state with efdc_adj_seed_con; call efdc_adj_step with the
   same step; read efdc_adj_source_grad; load the state before the
   previous step, and so on backwards.
   A driver that keeps the states after each step of a checkpoint segment
   can skip the repeated forward step: save the recording of every step
   with efdc_adj_record_save next to the state after it, and before
   efdc_adj_step load both (efdc_state_load, efdc_adj_record_load). */
typedef struct {
  int32_t nadj;                 /* constituents differentiated            */
  int32_t slot[16];             /* their slots                            */
  double cg_rel_tol;            /* stop of the adjoint CG: ||r|| <= tol*||b|| (0: 1e-12) */
  int32_t cg_itmax;             /* 0: ITERM                               */
} efdc_adj_opts_t;

typedef struct {
  int32_t cg_iters;             /* iterations of the adjoint solve        */
  double cg_rsq;                /* its final squared residual (scaled)    */
} efdc_adj_stats_t;

typedef enum {
  EFDC_ADJ_MANNING = 0,         /* (LCM) accumulated dJ/dn                */
  EFDC_ADJ_HP, EFDC_ADJ_UHDYE, EFDC_ADJ_VHDXE, EFDC_ADJ_P, EFDC_ADJ_HU, EFDC_ADJ_HV,
  EFDC_ADJ_CON = 16             /* + slot: (KCM,LCM) lambda of CON        */
} efdc_adj_field;

int efdc_adj_enable(efdc_ctx_t *ctx, const efdc_adj_opts_t *opts);
int efdc_adj_disable(efdc_ctx_t *ctx);
/* zero every lambda and the Manning gradient (tables must be set) */
int efdc_adj_reset(efdc_ctx_t *ctx);
/* lambda of CON(L,K) of constituent `slot` += val, for n entries */
int efdc_adj_seed_con(efdc_ctx_t *ctx, int32_t slot, int32_t n, const int32_t *L,
                      const int32_t *K, const double *val);
int efdc_adj_step(efdc_ctx_t *ctx, const efdc_step_t *step, efdc_adj_stats_t *stats);
/* dJ/dCQS(q,K,c) of the last adjoint step: out[(q*KC + K-1)*nconst + c],
   zero for constituents that are not differentiated */
int efdc_adj_source_grad(efdc_ctx_t *ctx, double *out);
int efdc_adj_get(efdc_ctx_t *ctx, int32_t field, double *out);
/* The recording of the last forward step (its CG solution, its scaling
   minimum CCMNM, its source concentrations CQS and the old levels of the
   adjoint constituents) to and
   from a device buffer of efdc_adj_record_bytes bytes (no counterpart in
   the reference: a checkpointed adjoint driver's storage). */
int efdc_adj_record_bytes(efdc_ctx_t *ctx, size_t *bytes);
int efdc_adj_record_save(efdc_ctx_t *ctx, void *dev_buf);
int efdc_adj_record_load(efdc_ctx_t *ctx, const void *dev_buf);

#ifdef __cplusplus
}
#endif
#endif /* EFDC_HIP_H */
