// efdc_adj.hip -- state snapshots and the discrete adjoint of one step.
//
// The step is the hot path with the depth-uniform closure:
//   closure_pre  TBX = G n^2 u|u| / HU^(1/3), u = UHE/HU        (Manning)
//   CALPUV2C     explicit stage (EFDC/calpuv2c.f90:214-289), assembly
//                (:315-342), clamped open boundaries (EFDC/setopenbc.f90,
//                EFDC/setopenbc2.f90), CONGRAD, face transports, depths,
//                HU/HV (:416-571,896-914)
//   closure_post UHDY2 = 0.5 (UHDYE+UHDY1E) DZC, HPK = HP DZC, HPKI = 1/HPK
//   CALCONC      upwind 2TL transport with sources and open-boundary
//                concentrations, implicit vertical diffusion, dye decay
//                (EFDC/caltran.f90, EFDC/calfqc.f90, EFDC/calconc.f90)
// and this file runs its exact reverse: every kernel below is the
// transpose of one forward kernel's Jacobian, written as a gather over the
// cells that read a value (so there are no scattered atomics and the
// result does not depend on scheduling).  The solve A P = FP is reversed
// by a second FP64 CG with the same operator: on the cells that are not
// clamped boundaries A is symmetric, the clamped rows are diagonal and
// their adjoint right-hand side is zero.  The upwind donor choice, the
// max(.,0) clips and the boundary inflow/outflow switches are taken as
// the forward step took them (the derivative of the branch taken).
//
// The context records during a forward step while the adjoint is enabled:
// the CG solution (the forward step replaces P by G (HP + BELV)
// afterwards), the old level of every differentiated constituent (the
// forward step overwrites CON1 on boundary columns) and, for a settling
// one, its column after the transport (the settling overwrites it).
//
// Settling (SSEDTOX -> CALSED, EFDC/calsed.f90:232-256,308-312) is an
// implicit upwind sweep from the top layer down; its reverse runs from the
// bottom layer up and passes each layer's lambda to the layer above (whose
// settled value fed the flux into it) and to HPKI of the layer.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "efdc_internal.h"

struct AdjState {
  bool recording = false;
  int nadj = 0;
  int slot[EFDC_MAXC] = {};
  double rel_tol = 1e-12;
  int itmax = 0;
  int cur = 0;
  // hydrodynamic lambdas: [cur] for the state after the step being
  // reversed, [cur^1] accumulates the state before it
  double *hp[2] = {}, *u[2] = {}, *v[2] = {}, *p[2] = {}, *hu[2] = {}, *hv[2] = {};
  double *ps = nullptr;    // recorded CG solution of the forward step
  double *mu = nullptr;    // adjoint CG solution
  double *rhs = nullptr;   // its right-hand side (scaled)
  double *fu = nullptr, *fv = nullptr, *fp1 = nullptr;
  double *nman = nullptr;  // dJ/dn
  double *lam[EFDC_MAXC] = {};  // lambda of CON, KC planes (FP64)
  void *c0[EFDC_MAXC] = {};     // recorded CON before the step (storage type)
  void *cset[EFDC_MAXC] = {};   // recorded CON before the settling (storage type)
  void *cadac[EFDC_MAXC] = {};  // recorded CON before the anti-diffusion (storage type)
  // anti-diffusion reverse: lambda of CON after the upwind step and of
  // CON1, and the work planes of the reverse passes (FP64)
  double *ad[11] = {};
  double *lcd = nullptr, *own = nullptr;
  double *wk[5] = {};
  double *bcx = nullptr;
  size_t bcx_n = 0;
  double *lamclo[4] = {};
  size_t lamclo_n[4] = {};
  double *srcg = nullptr;
  size_t srcg_n = 0;
  unsigned char *z2 = nullptr;  // couplings zeroed by SETOPENBC2: 1 W, 2 E, 4 S, 8 N
  // seed lists (efdc_adj_seed_con): L, K as ints and the values
  int *seed_i = nullptr;
  double *seed_v = nullptr;
  int seed_cap = 0;
  CgState *st = nullptr, *st_host = nullptr;
};

namespace {

template <typename T>
__device__ __forceinline__ double ld(const T *p, int64_t j) {
  return (double)p[j];
}
__device__ __forceinline__ double dmx(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double dmn(double a, double b) { return a < b ? a : b; }

__global__ void k_fill0(double *p, size_t n) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0.0;
}

__global__ void k_add(double *p, const double *q, size_t n) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] += q[t];
}

int zero(efdc_ctx *c, double *p, size_t n) {
  if (!p || n == 0) return EFDC_OK;
  k_fill0<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(p, n);
  HIP_TRY(c, hipGetLastError());
  return EFDC_OK;
}

// ------------------------------------------------------------ transport
template <typename T>
struct TaArgs {
  const T *uhdy2, *vhdx2, *w2, *h1pk, *hpki, *ab, *cdzkk, *cdzkmk, *dzc, *c0;
  const double *dxyip, *hpi, *cdzkk_k, *cdzkmk_k;
  int dz_uniform;
  const int *lw, *le, *ls, *ln;
  int nwet, lc0, KC, nc, c, mo, dye_mode;
  int64_t pitch;
  double dt, dye_f;
  long long niter;
  const int *srcrec, *srcnext;
  const double *qss, *qsercell, *cqs, *csertq;
  const int *bcrec;
  const int4 *bcinfo;
  CbcDev cbc[4];
  const double *wtci;
  double *lamclo[4];
  double *lam, *lcd, *own, *wcd, *wv, *wvcu, *weeb, *ww;
  const double *zin;  // lambda of the column's update given (anti-diffusion), or null
  double *bcx;
  double *srcg;
  double *ahp_cur, *ahp_prev, *au_cur, *au_prev, *av_cur, *av_prev;
};

// One layer of the forward column update (the expressions of the forward
// transport kernel, EFDC/caltran.f90:95-111,188-245 and
// EFDC/calfqc.f90:211-224).
struct LayerVals {
  double c1, u_i, u_e, v_i, v_n, fu_i, fu_e, fv_i, fv_n, fw, cd, h1pk, hpki, qneg;
};

struct Nb {
  int w, s, e, n, w_of_e, s_of_n;
  bool eg, ng;
};

template <typename T>
__device__ __forceinline__ double fw_of(const TaArgs<T> &a, int i, int k) {
  // FWUU between layers k and k+1 (0-based), zero at the surface
  if (k < 0 || k >= a.KC - 1) return 0.0;
  const int64_t o = (int64_t)k * a.pitch;
  const double ww = ld(a.w2, i + (int64_t)(k + 1) * a.pitch);
  return ww * (ww >= 0.0 ? ld(a.c0, i + o) : ld(a.c0, i + o + a.pitch));
}

template <typename T>
__device__ __forceinline__ void layer_fwd(const TaArgs<T> &a, int i, int k, const Nb &q,
                                          LayerVals &L) {
  const int64_t o = (int64_t)k * a.pitch;
  const T *C1 = a.c0;
  L.u_i = ld(a.uhdy2, i + o);
  L.u_e = q.eg ? 0.0 : ld(a.uhdy2, q.e + o);
  L.v_i = ld(a.vhdx2, i + o);
  L.v_n = q.ng ? 0.0 : ld(a.vhdx2, q.n + o);
  L.h1pk = ld(a.h1pk, i + o);
  L.hpki = ld(a.hpki, i + o);
  const double c1 = ld(C1, i + o);
  L.c1 = c1;
  const double cw = ld(C1, q.w + o), cs = ld(C1, q.s + o);
  L.fu_i = L.u_i * (L.u_i >= 0.0 ? cw : c1);
  L.fv_i = L.v_i * (L.v_i >= 0.0 ? cs : c1);
  L.fu_e = 0.0;
  L.fv_n = 0.0;
  if (!q.eg) L.fu_e = L.u_e * (L.u_e >= 0.0 ? ld(C1, q.w_of_e + o) : ld(C1, q.e + o));
  if (!q.ng) L.fv_n = L.v_n * (L.v_n >= 0.0 ? ld(C1, q.s_of_n + o) : ld(C1, q.n + o));
  L.fw = fw_of(a, i, k);
  double fqc = 0.0, qneg = 0.0;
  for (int r = a.srcrec[i]; r >= 0; r = a.srcnext[r]) {
    const double conq = 0.5 * (3. * c1 - c1);
    const size_t j = (size_t)r * a.KC + k;
    const double qv = a.qss[j], qs = a.qsercell[j];
    fqc = fqc + dmx(qv, 0.) * a.cqs[j * a.nc + a.c] + dmn(qv, 0.) * conq +
          dmx(qs, 0.) * a.csertq[j * a.nc + a.c] + dmn(qs, 0.) * conq;
    qneg += dmn(qv, 0.) + dmn(qs, 0.);
  }
  L.qneg = qneg;
  const double fwprev = fw_of(a, i, k - 1);
  L.cd = c1 * L.h1pk + a.dt * ((fqc + L.fu_i - L.fu_e + L.fv_i - L.fv_n) * a.dxyip[i] +
                               (fwprev - L.fw));
}

// Open-boundary sequence of one boundary column at one layer
// (EFDC/caltran.f90:319-471), with what its reverse needs.
struct BcSide {
  double cv1, ctmp, flow, flux, r;
  size_t m;
  int present, out, gate, ramp;
};

template <typename T>
__device__ __forceinline__ double bc_fwd(const TaArgs<T> &a, int4 bi, int k, double D,
                                         const LayerVals &L, BcSide (&bs)[4]) {
  const int LLs[4] = {bi.x, bi.y, bi.z, bi.w};
  const int KC = a.KC, nc = a.nc, c = a.c;
  double cv1 = L.c1, con = 0.0;
#pragma unroll
  for (int side = 0; side < 4; ++side) {
    BcSide &b = bs[side];
    b.present = 0;
    const int LL = LLs[side];
    if (LL < 0) continue;
    const CbcDev &t = a.cbc[side];
    b.present = 1;
    b.cv1 = cv1;
    double flow, flux;
    bool out;
    if (side == 0) { flow = L.v_n; flux = L.fv_n; out = flow <= 0.; }
    else if (side == 1) { flow = L.u_e; flux = L.fu_e; out = flow <= 0.; }
    else if (side == 2) { flow = L.u_i; flux = L.fu_i; out = flow >= 0.; }
    else { flow = L.v_i; flux = L.fv_i; out = flow >= 0.; }
    b.flow = flow;
    b.flux = flux;
    b.out = out ? 1 : 0;
    const size_t m = ((size_t)LL * KC + k) * nc + c;
    b.m = m;
    const double cb1 = t.cb[((size_t)LL * 2 + 0) * nc + c];
    const double cb2 = t.cb[((size_t)LL * 2 + 1) * nc + c];
    if (out) {
      double CTMP;
      if (side <= 1) CTMP = cv1 + a.dt * (flow * cv1 - flux) * D * L.hpki;
      else CTMP = cv1 + a.dt * (flux - flow * cv1) * D * L.hpki;
      b.ctmp = CTMP;
      con = dmx(CTMP, 0.);
      b.gate = CTMP > 0. ? 1 : 0;
      if (a.mo == 1) {
        const double cbtmp = cb1 + t.csert[((size_t)LL * KC + 0) * nc + c];
        if (con > cbtmp) {
          con = cbtmp;
          b.gate = 0;
        }
      }
    } else {
      const double CBT = a.wtci[k * 2 + 0] * cb1 + a.wtci[k * 2 + 1] * cb2 + t.csert[m];
      // NLO is unchanged by an inflow step, CLO too
      const long long nmnlod = a.niter - t.nlo[m];
      b.ramp = 0;
      b.gate = 0;
      if (nmnlod >= t.ntscr[LL]) {
        con = CBT;
      } else {
        const double r = (double)nmnlod / (double)t.ntscr[LL];
        const double cbtmp = t.clo[m] + (CBT - t.clo[m]) * (double)nmnlod / (double)t.ntscr[LL];
        con = dmx(cbtmp, 0.);
        b.ramp = 1;
        b.r = r;
        b.gate = cbtmp > 0. ? 1 : 0;
      }
    }
    cv1 = con;
  }
  return con;
}

// Tridiagonal coefficients of layer k (EFDC/calconc.f90:264-422): lower
// l_k, upper u_k and the hpi-derivative factors of each.
template <typename T>
__device__ __forceinline__ void vd_coef(const TaArgs<T> &a, int i, int k, double hpi,
                                        double &l, double &u, double &dl, double &du) {
  const int64_t o = (int64_t)k * a.pitch;
  const double ckk = a.dz_uniform ? a.cdzkk_k[k] : ld(a.cdzkk, i + o);
  const double ckm = a.dz_uniform ? a.cdzkmk_k[k] : ld(a.cdzkmk, i + o);
  const double DELTD2 = a.dt;
  l = u = dl = du = 0.0;
  if (k > 0) {
    const double RCDZKMK = -DELTD2 * ckm;
    dl = RCDZKMK * ld(a.ab, i + o - a.pitch);
    l = RCDZKMK * hpi * ld(a.ab, i + o - a.pitch);
  }
  if (k < a.KC - 1) {
    const double RCDZKK = -DELTD2 * ckk;
    du = RCDZKK * ld(a.ab, i + o);
    u = RCDZKK * hpi * ld(a.ab, i + o);
  }
}

template <typename T>
__device__ __forceinline__ Nb nb_of(const TaArgs<T> &a, int i) {
  Nb q;
  q.w = a.lw[i];
  q.s = a.ls[i];
  q.e = a.le[i];
  q.n = a.ln[i];
  q.eg = (q.e == a.lc0);
  q.ng = (q.n == a.lc0);
  q.w_of_e = q.eg ? i : a.lw[q.e];
  q.s_of_n = q.ng ? i : a.ls[q.n];
  return q;
}

// Column pass: recompute the forward column, solve the transposed
// tridiagonal system for lambda of the pre-diffusion concentrations, and
// reverse the boundary sequence.  Leaves lambda of the interior update
// (lcd), the direct lambda of the column's own old value (own), the
// boundary face terms (bcx), dJ/dCQS and the depth terms.
template <typename T>
__global__ __launch_bounds__(EFDC_BLOCK) void k_adj_tran_col(TaArgs<T> a) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (i > a.nwet) return;
  const int KC = a.KC;
  const int64_t P = a.pitch;
  const Nb q = nb_of(a, i);
  const double D = a.dxyip[i], hpi = a.hpi[i];
  const int rec = a.bcrec ? a.bcrec[i] : -1;
  int4 bi = make_int4(-1, -1, -1, -1);
  if (rec >= 0) bi = a.bcinfo[rec];
  // pass 1: forward values, tridiagonal forward elimination (with
  // anti-diffusion the column finish's reverse has given lambda of the
  // update, zin, and passes 1-3 are skipped)
  double vcu_prev = 0.0, y_prev = 0.0;
  for (int k = 0; k < (a.zin ? 0 : KC); ++k) {
    LayerVals L;
    layer_fwd(a, i, k, q, L);
    double con = L.cd * L.hpki;
    if (rec >= 0) {
      BcSide bs[4];
      con = bc_fwd(a, bi, k, D, L, bs);
    }
    const int64_t o = i + (int64_t)k * P;
    a.wcd[o] = con;
    if (KC == 1) {
      a.wv[o] = con;
      continue;
    }
    double l, u, dl, du;
    vd_coef(a, i, k, hpi, l, u, dl, du);
    double eeb;
    if (k == 0) eeb = 1.0 / (1.0 - u);
    else if (k < KC - 1) eeb = 1.0 / ((1.0 - l - u) - l * vcu_prev);
    else eeb = 1.0 / ((1.0 - l) - l * vcu_prev);
    const double y = (k == 0) ? con * eeb : (con - l * y_prev) * eeb;
    const double vcu = u * eeb;
    a.wv[o] = y;
    a.wvcu[o] = vcu;
    a.weeb[o] = eeb;
    vcu_prev = vcu;
    y_prev = y;
  }
  // pass 2: back substitution -> solution v (before the decay factor)
  if (KC > 1 && !a.zin) {
    double up = a.wv[i + (int64_t)(KC - 1) * P];
    for (int k = KC - 2; k >= 0; --k) {
      const int64_t o = i + (int64_t)k * P;
      const double v = a.wv[o] - a.wvcu[o] * up;
      a.wv[o] = v;
      up = v;
    }
  }
  // pass 3: lambda of v, then U^T w = g
  const double fdec = a.dye_mode == 1 ? a.dye_f : 1.0;
  if (KC > 1 && !a.zin) {
    double wprev = 0.0;
    for (int k = 0; k < KC; ++k) {
      const int64_t o = i + (int64_t)k * P;
      const double g = fdec * a.lam[o];
      const double wk = (k == 0) ? g : g - a.wvcu[o - P] * wprev;
      a.ww[o] = wk;
      wprev = wk;
    }
  }
  // pass 4: L^T z = w, then every per-layer term
  double lam_hpi = 0.0, lam_hp_cur = 0.0, lam_hp_prev = 0.0;
  double z_next = 0.0, l_next = 0.0;
  for (int k = KC - 1; k >= 0; --k) {
    const int64_t o = i + (int64_t)k * P;
    double z;
    double l = 0.0, u = 0.0, dl = 0.0, du = 0.0;
    if (a.zin) {
      z = a.zin[o];
    } else if (KC == 1) {
      z = fdec * a.lam[o];
    } else {
      vd_coef(a, i, k, hpi, l, u, dl, du);
      z = (a.ww[o] - l_next * z_next) * a.weeb[o];
      const double vk = a.wv[o];
      const double lam_d = -z * vk;
      if (k > 0) lam_hpi += (-z * a.wv[o - P] - lam_d) * dl;
      if (k < KC - 1) lam_hpi += (-z * a.wv[o + P] - lam_d) * du;
    }
    z_next = z;
    l_next = l;
    LayerVals L;
    layer_fwd(a, i, k, q, L);
    const double dz = ld(a.dzc, o);
    double lam_hpki = 0.0, own = 0.0, lcd = 0.0;
    if (rec < 0) {
      lcd = z * L.hpki;
      lam_hpki = z * L.cd;
      own = lcd * L.h1pk + lcd * a.dt * D * L.qneg;
      lam_hp_prev += dz * lcd * L.c1;
      for (int r = a.srcrec[i]; r >= 0; r = a.srcnext[r]) {
        const size_t j = (size_t)r * KC + k;
        a.srcg[j * a.nc + a.c] = lcd * a.dt * D * dmx(a.qss[j], 0.);
      }
    } else {
      BcSide bs[4];
      (void)bc_fwd(a, bi, k, D, L, bs);
      double lm = z;
#pragma unroll
      for (int side = 3; side >= 0; --side) {
        const BcSide &b = bs[side];
        if (!b.present) continue;
        double lcv1 = 0.0;
        double *lclo = a.lamclo[side];
        double lflow = 0.0, lflux = 0.0;
        if (b.out) {
          // CLO = CON of this step: its lambda (from later inflow ramps)
          // joins, and CLO before the step no longer matters
          const double lc = lm + lclo[b.m];
          lclo[b.m] = 0.0;
          const double lct = b.gate ? lc : 0.0;
          const double sg = side <= 1 ? 1.0 : -1.0;
          const double f = a.dt * D * L.hpki;
          lcv1 = lct * (1.0 + sg * f * b.flow);
          lflow = lct * sg * f * b.cv1;
          lflux = -lct * sg * f;
          lam_hpki += lct * sg * a.dt * (b.flow * b.cv1 - b.flux) * D;
        } else if (b.ramp && b.gate) {
          lclo[b.m] += lm * (1.0 - b.r);
        }
        double *x = a.bcx + (((size_t)rec * KC + k) * 4 + side) * 2;
        x[0] = lflow;
        x[1] = lflux;
        lm = lcv1;
      }
      own = lm;
    }
    lam_hp_cur += dz * (-lam_hpki * L.hpki * L.hpki);
    a.lcd[o] = lcd;
    a.own[o] = own;
  }
  lam_hp_cur += -lam_hpi * hpi * hpi;
  a.ahp_cur[i] += lam_hp_cur;
  a.ahp_prev[i] += lam_hp_prev;
}

// lambda of the flux through the west (u) or south (v) face of cell f
template <typename T>
__device__ __forceinline__ double face_lam(const TaArgs<T> &a, int f, int k, bool vface,
                                           double &lflow) {
  const int64_t P = a.pitch;
  const int64_t o = (int64_t)k * P;
  double val = a.dt * a.dxyip[f] * a.lcd[f + o];
  const int b = vface ? a.ls[f] : a.lw[f];
  const bool back = (b != a.lc0) && ((vface ? a.ln[b] : a.le[b]) == f);
  if (back) val -= a.dt * a.dxyip[b] * a.lcd[b + o];
  lflow = 0.0;
  if (a.bcrec) {
    // side 2 (u_i) / side 3 (v_i) of the column itself, side 1 (u_e) /
    // side 0 (v_n) of the column behind the face
    const int r = a.bcrec[f];
    if (r >= 0) {
      const int4 bi = a.bcinfo[r];
      const int sd = vface ? 3 : 2;
      if ((vface ? bi.w : bi.z) >= 0) {
        const double *x = a.bcx + (((size_t)r * a.KC + k) * 4 + sd) * 2;
        lflow += x[0];
        val += x[1];
      }
    }
    if (back) {
      const int r2 = a.bcrec[b];
      if (r2 >= 0) {
        const int4 bi = a.bcinfo[r2];
        const int sd = vface ? 0 : 1;
        if ((vface ? bi.x : bi.y) >= 0) {
          const double *x = a.bcx + (((size_t)r2 * a.KC + k) * 4 + sd) * 2;
          lflow += x[0];
          val += x[1];
        }
      }
    }
  }
  return val;
}

// Gather pass: lambda of the old concentrations and of the face
// transports (and through the closure of UHDYE/VHDXE of both levels).
template <typename T>
__global__ __launch_bounds__(EFDC_BLOCK) void k_adj_tran_gather(TaArgs<T> a) {
  const int j = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (j > a.nwet) return;
  const int KC = a.KC;
  const int64_t P = a.pitch;
  const int w = a.lw[j], s = a.ls[j], e = a.le[j], n = a.ln[j];
  double aub = 0.0, avb = 0.0;
  for (int k = 0; k < KC; ++k) {
    const int64_t o = (int64_t)k * P;
    double lc = a.own[j + o];
    double lflow;
    // west face of j
    {
      const double u = ld(a.uhdy2, j + o);
      const double lf = face_lam(a, j, k, false, lflow);
      if (u < 0.0) lc += lf * u;
      const int dn = u >= 0.0 ? w : j;
      aub += ld(a.dzc, j + o) * (lf * ld(a.c0, dn + o) + lflow);
    }
    if (e != a.lc0) {
      const double u = ld(a.uhdy2, e + o);
      if (u >= 0.0) lc += face_lam(a, e, k, false, lflow) * u;
    }
    // south face of j
    {
      const double v = ld(a.vhdx2, j + o);
      const double lf = face_lam(a, j, k, true, lflow);
      if (v < 0.0) lc += lf * v;
      const int dn = v >= 0.0 ? s : j;
      avb += ld(a.dzc, j + o) * (lf * ld(a.c0, dn + o) + lflow);
    }
    if (n != a.lc0) {
      const double v = ld(a.vhdx2, n + o);
      if (v >= 0.0) lc += face_lam(a, n, k, true, lflow) * v;
    }
    // vertical fluxes FWUU(k) (between k and k+1) and FWUU(k-1)
    if (k < KC - 1) {
      const double ww = ld(a.w2, j + (int64_t)(k + 1) * P);
      if (ww >= 0.0) lc += a.dt * (a.lcd[j + o + P] - a.lcd[j + o]) * ww;
    }
    if (k > 0) {
      const double ww = ld(a.w2, j + (int64_t)k * P);
      if (ww < 0.0) lc += a.dt * (a.lcd[j + o] - a.lcd[j + o - P]) * ww;
    }
    a.lam[j + o] = lc;
  }
  a.au_cur[j] += 0.5 * aub;
  a.au_prev[j] += 0.5 * aub;
  a.av_cur[j] += 0.5 * avb;
  a.av_prev[j] += 0.5 * avb;
}

// Reverse of the settling sweep of one column.  Forward (top K = KC down):
//   v_KC = m_KC / (1 + ws W_KC),
//   v_K  = (m_K + ws v_(K+1) W_K) / (1 + ws W_K),             1 < K < KC,
//   v_1  = (m_1 + (wese + ws v_2) W_1) / (1 + wsetmp W_1),
// with m_K = max(c_K, 0), W_K = DTSED HPKI_K (v_2 absent when KC = 1).
// lambda of v becomes lambda of c in place; lambda of HPKI goes to HP
// (HPKI = 1 / (HP DZC)).  c are the recorded values before the sweep, v
// the values it stored.
template <typename T>
__global__ __launch_bounds__(EFDC_BLOCK) void k_adj_settle(
    double *lam, const T *cpre, const T *cpost, const T *hpki, const T *dzc, int64_t pitch,
    int KC, int nwet, double dtsed, double ws, double wsetmp, double wese, double *ahp) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (i > nwet) return;
  double carry = 0.0, lam_hp = 0.0;
  for (int k = 0; k < KC; ++k) {
    const int64_t o = i + (int64_t)k * pitch;
    const double lv = lam[o] + carry;
    const double h = ld(hpki, o), W = dtsed * h;
    const double v = ld(cpost, o), c = ld(cpre, o);
    const double vup = (k + 1 < KC) ? ld(cpost, o + pitch) : 0.0;
    double CL, dW;
    if (k == 0) {
      CL = 1. + wsetmp * W;
      dW = (wese + ws * vup) - wsetmp * v;
    } else if (k + 1 < KC) {
      CL = 1. + ws * W;
      dW = ws * vup - ws * v;
    } else {
      CL = 1. + ws * W;
      dW = -ws * v;
    }
    carry = (k + 1 < KC) ? lv * ws * W / CL : 0.0;
    lam[o] = c > 0.0 ? lv / CL : 0.0;
    const double lam_hpki = lv * dW / CL * dtsed;
    lam_hp += lam_hpki * (-h * h * ld(dzc, o));
  }
  ahp[i] += lam_hp;
}

// ---------------------------------------------------------- hydrodynamics
struct HPtrA {
  double *a[EFDC_NUM_H];
};
#define HA(id) h.a[EFDC_##id]

struct HyArgs {
  const int *lw, *le, *ls, *ln;
  const unsigned char *pbc, *z2;
  int nwet, lc0;
  double dt, G, RITB1;
  double *hp_c, *hp_p, *u_c, *u_p, *v_c, *v_p, *p_c, *p_p, *hu_c, *hu_p, *hv_c, *hv_p;
  double *ps, *mu, *rhs, *fu, *fv, *fp1, *an;
  const double *nman;
  const CgState *st;
};

__device__ __forceinline__ bool wet(const HyArgs &a, int j) { return j != a.lc0 && j > 0; }

// HP of the step: everything that read it (HU/HV, P = G(HP+BELV), the
// transport terms already added) gives lambda of HP; clamped boundary
// cells hold a fixed depth.
__global__ __launch_bounds__(EFDC_BLOCK) void k_adj_h1(HPtrA h, HyArgs a) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (i > a.nwet) return;
  double t = 0.0;
  if (!a.pbc[i]) {
    const double *DXYP = HA(DXYP), *DXYU = HA(DXYU), *DXYV = HA(DXYV);
    t = a.hp_c[i] + a.G * a.p_c[i];
    t += 0.5 * DXYP[i] * a.hu_c[i] / DXYU[i];
    const int e = a.le[i], n = a.ln[i];
    if (wet(a, e) && a.lw[e] == i) t += 0.5 * DXYP[i] * a.hu_c[e] / DXYU[e];
    t += 0.5 * DXYP[i] * a.hv_c[i] / DXYV[i];
    if (wet(a, n) && a.ls[n] == i) t += 0.5 * DXYP[i] * a.hv_c[n] / DXYV[n];
  }
  a.hp_c[i] = t;
}

// continuity (EFDC/calpuv2c.f90:536-548)
__global__ __launch_bounds__(EFDC_BLOCK) void k_adj_h2(HPtrA h, HyArgs a) {
  const int j = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (j > a.nwet) return;
  const double DELTD2 = 0.5 * a.dt;
  const double *D = HA(DXYIP);
  const double lj = a.hp_c[j];
  const int w = a.lw[j], s = a.ls[j];
  const double lw_ = (wet(a, w) && a.le[w] == j) ? a.hp_c[w] * D[w] : 0.0;
  const double ls_ = (wet(a, s) && a.ln[s] == j) ? a.hp_c[s] * D[s] : 0.0;
  const double du = DELTD2 * (D[j] * lj - lw_);
  const double dv = DELTD2 * (D[j] * lj - ls_);
  a.u_c[j] += du;
  a.u_p[j] += du;
  a.v_c[j] += dv;
  a.v_p[j] += dv;
  a.hp_p[j] += lj;
}

// face transports UHDYE/VHDXE from FUHDYE and the solution (:416-434);
// writes the adjoint CG's right-hand side
__global__ __launch_bounds__(EFDC_BLOCK) void k_adj_h3(HPtrA h, HyArgs a) {
  const int j = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (j > a.nwet) return;
  const double DELTD2 = 0.5 * a.dt;
  const double *SUB = HA(SUB), *SVB = HA(SVB), *HRUO = HA(HRUO), *HRVO = HA(HRVO),
               *RCX = HA(RCX), *RCY = HA(RCY), *H1U = HA(H1U), *H1V = HA(H1V);
  const double lu = a.u_c[j], lv = a.v_c[j];
  const int w = a.lw[j], s = a.ls[j], e = a.le[j], n = a.ln[j];
  a.fu[j] = SUB[j] * lu;
  a.fv[j] = SVB[j] * lv;
  a.hu_p[j] += -SUB[j] * DELTD2 * HRUO[j] * RCX[j] * (a.ps[j] - a.ps[w]) * lu;
  a.hv_p[j] += -SVB[j] * DELTD2 * HRVO[j] * RCY[j] * (a.ps[j] - a.ps[s]) * lv;
  double lp = -SUB[j] * DELTD2 * HRUO[j] * RCX[j] * H1U[j] * lu -
              SVB[j] * DELTD2 * HRVO[j] * RCY[j] * H1V[j] * lv;
  if (wet(a, e) && a.lw[e] == j) lp += SUB[e] * DELTD2 * HRUO[e] * RCX[e] * H1U[e] * a.u_c[e];
  if (wet(a, n) && a.ls[n] == j) lp += SVB[n] * DELTD2 * HRVO[n] * RCY[n] * H1V[n] * a.v_c[n];
  const double CCMNMI = 1. / a.st->ccmnm;
  a.rhs[j] = a.pbc[j] ? 0.0 : lp * CCMNMI;
  a.mu[j] = 0.0;
}

// right-hand side FP (:320) and coefficients (:325-342) of the solve
__global__ __launch_bounds__(EFDC_BLOCK) void k_adj_h4(HPtrA h, HyArgs a) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (i > a.nwet) return;
  const double DELTD2 = 0.5 * a.dt, C2 = -0.5 * DELTD2 * a.G;
  const double *SUB = HA(SUB), *SVB = HA(SVB), *HRUO = HA(HRUO), *HRVO = HA(HRVO),
               *RCX = HA(RCX), *RCY = HA(RCY);
  const int w = a.lw[i], s = a.ls[i];
  const double mi = a.pbc[i] ? 0.0 : a.mu[i];
  const bool bw = wet(a, w) && a.le[w] == i && !a.pbc[w];
  const bool bs = wet(a, s) && a.ln[s] == i && !a.pbc[s];
  const double mw = bw ? a.mu[w] : 0.0, ms = bs ? a.mu[s] : 0.0;
  a.fu[i] += 0.5 * a.G * SUB[i] * (mi - mw);
  a.fv[i] += 0.5 * a.G * SVB[i] * (mi - ms);
  const double *P = a.ps;
  const double lcc_i = -mi * P[i];
  // west face of i: CW(i) and CE(w), inside CC(i) and CC(w)
  const double lcw = (a.z2[i] & 1) ? 0.0 : -mi * P[w];
  const double lce = (bw && !(a.z2[w] & 2)) ? -mw * P[i] : 0.0;
  const double lccw = bw ? -mw * P[w] : 0.0;
  a.hu_p[i] += C2 * SUB[i] * HRUO[i] * RCX[i] * (lcw + lce - lcc_i - lccw);
  // south face of i: CS(i) and CN(s)
  const double lcs = (a.z2[i] & 4) ? 0.0 : -mi * P[s];
  const double lcn = (bs && !(a.z2[s] & 8)) ? -ms * P[i] : 0.0;
  const double lccs = bs ? -ms * P[s] : 0.0;
  a.hv_p[i] += C2 * SVB[i] * HRVO[i] * RCY[i] * (lcs + lcn - lcc_i - lccs);
  a.fp1[i] = mi;
}

// clamped open boundaries: FP(neighbour) += C*T * FP(boundary)
// (EFDC/setopenbc.f90:62-74 and the other sides)
__global__ void k_adj_h4bc(HPtrA h, HyArgs a, PbcDev s, int side, const int *lopenbcdry) {
  const int LL = blockIdx.x * blockDim.x + threadIdx.x;
  if (LL >= s.n) return;
  const int b = s.L[LL];
  if (lopenbcdry[b]) return;
  const double D_LTD2 = 0.5 * a.dt, f = 0.5 * D_LTD2 * a.G * a.ps[b];
  const double *HRUO = HA(HRUO), *HRVO = HA(HRVO), *RCX = HA(RCX), *RCY = HA(RCY);
  int row, face;
  if (side == 0) { row = a.le[b]; face = row; }
  else if (side == 1) { row = a.lw[b]; face = b; }
  else if (side == 2) { row = a.ln[b]; face = row; }
  else { row = a.ls[b]; face = b; }
  if (!wet(a, row) || a.pbc[row]) return;
  const double lfp = a.mu[row];
  if (side <= 1) atomicAdd(&a.hu_p[face], f * HRUO[face] * RCX[face] * lfp);
  else atomicAdd(&a.hv_p[face], f * HRVO[face] * RCY[face] * lfp);
}

// explicit stage (:214-231), FP1 (:289) and the Manning closure
__global__ __launch_bounds__(EFDC_BLOCK) void k_adj_h5(HPtrA h, HyArgs a) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (i > a.nwet) return;
  const double DELT = a.dt, DELTD2 = 0.5 * a.dt, DELTI = 1. / a.dt, G = a.G;
  const double *SUB = HA(SUB), *SVB = HA(SVB), *HRUO = HA(HRUO), *HRVO = HA(HRVO),
               *H1U = HA(H1U), *H1V = HA(H1V), *P1 = HA(P1), *DXYP = HA(DXYP),
               *DXIU = HA(DXIU), *DYIV = HA(DYIV), *DXYU = HA(DXYU), *DXYV = HA(DXYV),
               *DYIU = HA(DYIU), *DXIV = HA(DXIV), *U1 = HA(UHDY1E), *V1 = HA(VHDX1E);
  const int w = a.lw[i], s = a.ls[i], e = a.le[i], n = a.ln[i];
  const double f1 = a.fp1[i];
  const double f1w = (wet(a, w) && a.le[w] == i) ? a.fp1[w] : 0.0;
  const double f1s = (wet(a, s) && a.ln[s] == i) ? a.fp1[s] : 0.0;
  const double lfu = a.fu[i], lfv = a.fv[i];
  a.u_p[i] += lfu + 0.5 * G * (f1 - f1w);
  a.v_p[i] += lfv + 0.5 * G * (f1 - f1s);
  double lp = DELTI * DXYP[i] * f1;
  lp += -DELTD2 * SUB[i] * HRUO[i] * H1U[i] * lfu - DELTD2 * SVB[i] * HRVO[i] * H1V[i] * lfv;
  if (wet(a, e) && a.lw[e] == i) lp += DELTD2 * SUB[e] * HRUO[e] * H1U[e] * a.fu[e];
  if (wet(a, n) && a.ls[n] == i) lp += DELTD2 * SVB[n] * HRVO[n] * H1V[n] * a.fv[n];
  a.p_p[i] += lp;
  double lhu = -DELTD2 * SUB[i] * HRUO[i] * (P1[i] - P1[w]) * lfu;
  double lhv = -DELTD2 * SVB[i] * HRVO[i] * (P1[i] - P1[s]) * lfv;
  if (a.nman) {
    const double ltbx = -SUB[i] * DELT * DXIU[i] * DXYU[i] * a.RITB1 * lfu;
    const double ltby = -SVB[i] * DELT * DYIV[i] * DXYV[i] * a.RITB1 * lfv;
    const double nn = a.nman[i];
    double ln_ = 0.0;
    {
      const double hu = H1U[i], cb = cbrt(hu);
      const double u = U1[i] * DYIU[i] / hu;
      const double tb = G * nn * nn * u * fabs(u) / cb;
      const double lu = ltbx * G * nn * nn * 2.0 * fabs(u) / cb;
      a.u_p[i] += lu * DYIU[i] / hu;
      lhu += -lu * u / hu - ltbx * tb / (3.0 * hu);
      ln_ += ltbx * 2.0 * G * nn * u * fabs(u) / cb;
    }
    {
      const double hv = H1V[i], cb = cbrt(hv);
      const double v = V1[i] * DXIV[i] / hv;
      const double tb = G * nn * nn * v * fabs(v) / cb;
      const double lv = ltby * G * nn * nn * 2.0 * fabs(v) / cb;
      a.v_p[i] += lv * DXIV[i] / hv;
      lhv += -lv * v / hv - ltby * tb / (3.0 * hv);
      ln_ += ltby * 2.0 * G * nn * v * fabs(v) / cb;
    }
    a.an[i] += ln_;
  }
  a.hu_p[i] += lhu;
  a.hv_p[i] += lhv;
}

__global__ void k_adj_seed(double *lam, int64_t pitch, int n, const int *L, const int *K,
                           const double *val) {
  // one thread, in list order: repeated entries add up deterministically
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  for (int t = 0; t < n; ++t) lam[(int64_t)(L[t] - 1) + (int64_t)(K[t] - 1) * pitch] += val[t];
}

// ---------------------------------------------------------------- host
int alloc_d(efdc_ctx *c, double **p, size_t n) { return dev_alloc(c, (void **)p, sizeof(double) * n); }

void dfree_d(void *&p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

template <typename T>
TaArgs<T> ta_args(efdc_ctx *c, AdjState *A, int slot, const efdc_step_t *step) {
  TaArgs<T> a;
  std::memset(&a, 0, sizeof(a));
  auto K = [&](int id) { return (const T *)c->dK[id - EFDC_HPK]; };
  a.uhdy2 = K(EFDC_UHDY2);
  a.vhdx2 = K(EFDC_VHDX2);
  a.w2 = (const T *)c->dW[0];
  a.h1pk = K(EFDC_H1PK);
  a.hpki = K(EFDC_HPKI);
  a.ab = K(EFDC_AB);
  a.cdzkk = K(EFDC_CDZKK);
  a.cdzkmk = K(EFDC_CDZKMK);
  a.dzc = K(EFDC_DZC);
  a.c0 = (const T *)A->c0[slot];
  a.dxyip = c->dH[EFDC_DXYIP];
  a.hpi = c->dH[EFDC_HPI];
  a.cdzkk_k = c->cdzkk_k;
  a.cdzkmk_k = c->cdzkmk_k;
  a.dz_uniform = c->dz_uniform() ? 1 : 0;
  a.lw = c->dLW;
  a.le = c->dLE;
  a.ls = c->dLS;
  a.ln = c->dLN;
  a.nwet = c->n_wet();
  a.lc0 = c->LC - 1;
  a.KC = c->KC;
  a.nc = c->nconst;
  a.c = slot;
  const efdc_constituent_t &d = c->con[slot].desc;
  a.mo = d.mo;
  a.dye_mode = 0;
  a.dye_f = 0.0;
  if (d.is_dye) {
    const double rk = d.rkdye;
    if (rk == 1000.0) {
      a.dye_mode = 2;
      a.dye_f = step->dt / 86400.;
    } else if (rk != 0.0) {
      a.dye_mode = 1;
      a.dye_f = rk < 0.0 ? std::exp(-rk * step->dt) : 1. / (1. + step->dt * rk);
    }
  }
  a.pitch = c->pitch;
  a.dt = step->dt;
  a.niter = step->niter;
  a.srcrec = c->srcrec;
  a.srcnext = c->srcnext;
  a.qss = c->qss;
  a.qsercell = c->qsercell;
  a.cqs = c->cqs;
  a.csertq = c->csertq;
  a.bcrec = c->nbcrec > 0 ? c->bcrec : nullptr;
  a.bcinfo = c->bcinfo;
  for (int s = 0; s < 4; ++s) {
    a.cbc[s] = c->cbc[s];
    a.lamclo[s] = A->lamclo[s];
  }
  a.wtci = c->wtci;
  a.lam = A->lam[slot];
  a.lcd = A->lcd;
  a.own = A->own;
  a.wcd = A->wk[0];
  a.wv = A->wk[1];
  a.wvcu = A->wk[2];
  a.weeb = A->wk[3];
  a.ww = A->wk[4];
  a.bcx = A->bcx;
  a.srcg = A->srcg;
  const int cu = A->cur, pv = A->cur ^ 1;
  a.ahp_cur = A->hp[cu];
  a.ahp_prev = A->hp[pv];
  a.au_cur = A->u[cu];
  a.au_prev = A->u[pv];
  a.av_cur = A->v[cu];
  a.av_prev = A->v[pv];
  return a;
}

HyArgs hy_args(efdc_ctx *c, AdjState *A, const efdc_step_t *step) {
  HyArgs a;
  a.lw = c->dLW;
  a.le = c->dLE;
  a.ls = c->dLS;
  a.ln = c->dLN;
  a.pbc = c->pbcflag;
  a.z2 = A->z2;
  a.nwet = c->n_wet();
  a.lc0 = c->LC - 1;
  a.dt = step->dt;
  a.G = c->prm.G;
  a.RITB1 = c->prm.RITB1;
  const int cu = A->cur, pv = A->cur ^ 1;
  a.hp_c = A->hp[cu]; a.hp_p = A->hp[pv];
  a.u_c = A->u[cu]; a.u_p = A->u[pv];
  a.v_c = A->v[cu]; a.v_p = A->v[pv];
  a.p_c = A->p[cu]; a.p_p = A->p[pv];
  a.hu_c = A->hu[cu]; a.hu_p = A->hu[pv];
  a.hv_c = A->hv[cu]; a.hv_p = A->hv[pv];
  a.ps = A->ps;
  a.mu = A->mu;
  a.rhs = A->rhs;
  a.fu = A->fu;
  a.fv = A->fv;
  a.fp1 = A->fp1;
  a.an = A->nman;
  a.nman = c->nman;
  a.st = c->st;
  return a;
}

// (re)allocate the arrays whose size follows the boundary and source tables
int adj_tables(efdc_ctx *c, AdjState *A) {
  int s;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  const size_t nbx = (size_t)c->nbcrec * c->KC * 8;
  if (nbx != A->bcx_n) {
    dfree_d((void *&)A->bcx);
    if ((s = alloc_d(c, &A->bcx, nbx))) return s;
    A->bcx_n = nbx;
  }
  for (int sd = 0; sd < 4; ++sd) {
    const size_t n = (size_t)c->cbc[sd].n * c->KC * c->nconst;
    if (n != A->lamclo_n[sd]) {
      dfree_d((void *&)A->lamclo[sd]);
      if ((s = alloc_d(c, &A->lamclo[sd], n))) return s;
      A->lamclo_n[sd] = n;
    }
  }
  const size_t ns = (size_t)c->nqsij * c->KC * c->nconst;
  if (ns != A->srcg_n) {
    dfree_d((void *&)A->srcg);
    if ((s = alloc_d(c, &A->srcg, ns))) return s;
    A->srcg_n = ns;
  }
  // couplings removed by SETOPENBC2 (EFDC/setopenbc2.f90:17-45)
  std::vector<unsigned char> z2(c->LCM, 0);
  for (int sd = 0; sd < 4; ++sd) {
    const PbcDev &p = c->pbc[sd];
    if (p.n == 0) continue;
    std::vector<int> L(p.n), ispb(p.n), nbm(c->LCM);
    HIP_TRY(c, hipMemcpy(L.data(), p.L, sizeof(int) * p.n, hipMemcpyDeviceToHost));
    HIP_TRY(c, hipMemcpy(ispb.data(), p.ispb, sizeof(int) * p.n, hipMemcpyDeviceToHost));
    const int *map = sd == 0 ? c->dLE : sd == 1 ? c->dLW : sd == 2 ? c->dLN : c->dLS;
    HIP_TRY(c, hipMemcpy(nbm.data(), map, sizeof(int) * c->LCM, hipMemcpyDeviceToHost));
    const unsigned char bit = sd == 0 ? 1 : sd == 1 ? 2 : sd == 2 ? 4 : 8;
    for (int t = 0; t < p.n; ++t) {
      if (ispb[t] != 0) {
        c->err = "adjoint: radiation open boundaries (ISPB=1,2) are not differentiated";
        return EFDC_ERR_ARG;
      }
      const int j = nbm[L[t]];
      if (j != c->LC - 1) z2[j] |= bit;
    }
  }
  HIP_TRY(c, hipMemcpyAsync(A->z2, z2.data(), c->LCM, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return EFDC_OK;
}

template <typename T>
int adj_tran(efdc_ctx *c, AdjState *A, int slot, const efdc_step_t *step,
             const double *zin = nullptr) {
  TaArgs<T> a = ta_args<T>(c, A, slot, step);
  a.zin = zin;
  const int nb = (c->n_wet() + EFDC_BLOCK - 1) / EFDC_BLOCK;
  int s;
  if ((s = zero(c, A->bcx, A->bcx_n))) return s;
  TIMED(c, TK_ADJ_TCOL, k_adj_tran_col<T><<<nb, EFDC_BLOCK, 0, c->stream>>>(a));
  TIMED(c, TK_ADJ_TGATHER, k_adj_tran_gather<T><<<nb, EFDC_BLOCK, 0, c->stream>>>(a));
  HIP_TRY(c, hipGetLastError());
  return EFDC_OK;
}

template <typename T>
int adj_settle(efdc_ctx *c, AdjState *A, int slot, const efdc_step_t *step) {
  const efdc_constituent_t &d = c->con[slot].desc;
  if (!(d.wset > 0.0) || !A->cset[slot]) return EFDC_OK;
  const int nb = (c->n_wet() + EFDC_BLOCK - 1) / EFDC_BLOCK;
  TIMED(c, TK_ADJ_TCOL, k_adj_settle<T><<<nb, EFDC_BLOCK, 0, c->stream>>>(
      A->lam[slot], (const T *)A->cset[slot], (const T *)c->con[slot].buf[c->con[slot].cur],
      (const T *)c->dK[EFDC_HPKI - EFDC_HPK], (const T *)c->dK[EFDC_DZC - EFDC_HPK],
      c->pitch, c->KC, c->n_wet(), step->dt, d.wset, d.probdep * d.wset, d.wese,
      A->hp[A->cur]));
  HIP_TRY(c, hipGetLastError());
  return EFDC_OK;
}

// Reverse of the anti-diffusion of one constituent: lambda of its CON
// after the upwind step into ad[0], of CON1 into ad[1].
int adj_adac(efdc_ctx *c, AdjState *A, int slot, const efdc_step_t *step) {
  const size_t n = (size_t)c->pitch * (c->KC + 1);
  int s;
  for (double *p : A->ad)
    if (!p) {
      c->err = "adjoint: anti-diffusion work planes missing";
      return EFDC_ERR_STATE;
    }
  if (!A->cadac[slot] || !A->c0[slot] || !A->lam[slot]) {
    c->err = "adjoint: anti-diffusion recording missing";
    return EFDC_ERR_STATE;
  }
  for (double *p : A->ad)
    if ((s = zero(c, p, n))) return s;
  AdacAdj io;
  std::memset(&io, 0, sizeof(io));
  io.conu = A->cadac[slot];
  io.con1 = A->c0[slot];
  io.lam = A->lam[slot];
  io.lconu = A->ad[0];
  io.lcon1 = A->ad[1];
  io.lcd = A->ad[2];
  io.lfu = A->ad[3];
  io.lfv = A->ad[4];
  io.lfw = A->ad[5];
  io.lduu = A->ad[6];
  io.ldvv = A->ad[7];
  io.lu = A->ad[8];
  io.lv = A->ad[9];
  io.lww = A->ad[10];
  io.wv = A->wk[1];
  io.wvcu = A->wk[2];
  io.weeb = A->wk[3];
  io.ww = A->wk[4];
  const int cu = A->cur, pv = A->cur ^ 1;
  io.ahp = A->hp[cu];
  io.au_c = A->u[cu];
  io.au_p = A->u[pv];
  io.av_c = A->v[cu];
  io.av_p = A->v[pv];
  io.ahu = A->hu[cu];
  io.ahv = A->hv[cu];
  io.hu = c->dH[EFDC_HU];
  io.hv = c->dH[EFDC_HV];
  io.uhdye = c->dH[EFDC_UHDYE];
  io.uhdy1e = c->dH[EFDC_UHDY1E];
  io.vhdxe = c->dH[EFDC_VHDXE];
  io.vhdx1e = c->dH[EFDC_VHDX1E];
  io.dyiu = c->dH[EFDC_DYIU];
  io.dxiv = c->dH[EFDC_DXIV];
  {
    // the decay of the column finish, as in ta_args
    const TaArgs<double> ta = ta_args<double>(c, A, slot, step);
    io.dye_mode = ta.dye_mode;
    io.dye_f = ta.dye_f;
  }
  return adj_adac_run(c, slot, io, step);
}

// KC planes of pitch doubles to a host (KCM, LCM) array
int get_planes(efdc_ctx *c, const double *d, double *out) {
  std::memset(out, 0, sizeof(double) * (size_t)c->KCM * c->LCM);
  HIP_TRY(c, hipMemcpy2DAsync(out, sizeof(double) * c->LCM, d, sizeof(double) * c->pitch,
                              sizeof(double) * c->LCM, c->KC, hipMemcpyDeviceToHost,
                              c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return EFDC_OK;
}

// ------------------------------------------------------------ snapshots
struct Piece {
  void *p;
  size_t bytes;
};

std::vector<Piece> state_pieces(efdc_ctx *c) {
  std::vector<Piece> v;
  const size_t kel = (size_t)c->pitch * c->KC, wel = (size_t)c->pitch * (c->KC + 1);
  for (int id = 0; id < EFDC_NUM_H; ++id) v.push_back({c->dH[id], sizeof(double) * c->LCM});
  for (int j = 0; j < EFDC_NK; ++j) v.push_back({c->dK[j], kel * c->tsz});
  for (int j = 0; j < EFDC_NW; ++j) v.push_back({c->dW[j], wel * c->tsz});
  for (int j = 0; j < c->nconst; ++j) {
    v.push_back({c->con[j].buf[c->con[j].cur], kel * c->tsz});
    v.push_back({c->con[j].buf[c->con[j].cur ^ 1], kel * c->tsz});
  }
  for (int s = 0; s < 4; ++s) {
    const size_t n = (size_t)c->cbc[s].n * c->KC * c->cbc_nc;
    if (n == 0 || !c->cbc[s].clo) continue;
    v.push_back({c->cbc[s].clo, sizeof(double) * n});
    v.push_back({c->cbc[s].nlo, sizeof(long long) * n});
  }
  v.push_back({c->lopenbcdry, sizeof(int) * c->LCM});
  std::vector<std::pair<void *, size_t>> iv;
  uvw_state_pieces(c, iv);
  for (auto &q : iv) v.push_back({q.first, q.second});
  return v;
}

// Copy the pieces to (save) or from (load) consecutive 256-byte aligned
// slots of buf on the context's stream.
int copy_pieces(efdc_ctx *c, const std::vector<Piece> &pcs, char *buf, bool save) {
  for (const Piece &p : pcs) {
    if (save)
      HIP_TRY(c, hipMemcpyAsync(buf, p.p, p.bytes, hipMemcpyDeviceToDevice, c->stream));
    else
      HIP_TRY(c, hipMemcpyAsync(p.p, buf, p.bytes, hipMemcpyDeviceToDevice, c->stream));
    buf += (p.bytes + 255) & ~(size_t)255;
  }
  return EFDC_OK;
}

// The recording of one forward step: the CG solution, the forward CgState
// (its CCMNM scales the adjoint right-hand side), the step's source
// concentrations (efdc_set_source_cqs may change them every step; the
// source term enters CD, whose derivative with respect to HPKI the reverse
// needs) and the old levels of the adjoint constituents, in 256-byte
// aligned pieces.
std::vector<Piece> record_pieces(efdc_ctx *c) {
  std::vector<Piece> v;
  AdjState *A = c->adj;
  v.push_back({A->ps, sizeof(double) * c->LCM});
  v.push_back({c->st, sizeof(CgState)});
  if (c->nqsij > 0 && c->cqs)
    v.push_back({c->cqs, sizeof(double) * (size_t)c->nqsij * c->KC * c->src_nc});
  const size_t kel = (size_t)c->pitch * c->KC;
  for (int j = 0; j < A->nadj; ++j) v.push_back({A->c0[A->slot[j]], kel * c->tsz});
  for (int j = 0; j < A->nadj; ++j)
    if (A->cset[A->slot[j]]) v.push_back({A->cset[A->slot[j]], kel * c->tsz});
  for (int j = 0; j < A->nadj; ++j)
    if (A->cadac[A->slot[j]]) v.push_back({A->cadac[A->slot[j]], kel * c->tsz});
  return v;
}


}  // namespace

bool adj_recording(const efdc_ctx *c) { return c->adj && c->adj->recording; }

int adj_record_ps(efdc_ctx *c) {
  HIP_TRY(c, hipMemcpyAsync(c->adj->ps, c->dH[EFDC_P], sizeof(double) * c->LCM,
                            hipMemcpyDeviceToDevice, c->stream));
  return EFDC_OK;
}

int adj_record_c0(efdc_ctx *c) {
  AdjState *A = c->adj;
  const size_t bytes = (size_t)c->pitch * c->KC * c->tsz;
  for (int j = 0; j < A->nadj; ++j) {
    const int sl = A->slot[j];
    HIP_TRY(c, hipMemcpyAsync(A->c0[sl], c->con[sl].buf[c->con[sl].cur], bytes,
                              hipMemcpyDeviceToDevice, c->stream));
  }
  return EFDC_OK;
}

int adj_record_settle(efdc_ctx *c, int slot, const void *con) {
  AdjState *A = c->adj;
  if (!A || !A->cset[slot]) return EFDC_OK;
  HIP_TRY(c, hipMemcpyAsync(A->cset[slot], con, (size_t)c->pitch * c->KC * c->tsz,
                            hipMemcpyDeviceToDevice, c->stream));
  return EFDC_OK;
}

int adj_record_adac(efdc_ctx *c, int slot, const void *con) {
  AdjState *A = c->adj;
  if (!A || !A->cadac[slot]) return EFDC_OK;
  HIP_TRY(c, hipMemcpyAsync(A->cadac[slot], con, (size_t)c->pitch * c->KC * c->tsz,
                            hipMemcpyDeviceToDevice, c->stream));
  return EFDC_OK;
}

void adj_free(efdc_ctx *c) {
  AdjState *A = c->adj;
  if (!A) return;
  (void)hipStreamSynchronize(c->stream);
  for (int q = 0; q < 2; ++q) {
    dfree_d((void *&)A->hp[q]); dfree_d((void *&)A->u[q]); dfree_d((void *&)A->v[q]);
    dfree_d((void *&)A->p[q]); dfree_d((void *&)A->hu[q]); dfree_d((void *&)A->hv[q]);
  }
  dfree_d((void *&)A->ps); dfree_d((void *&)A->mu); dfree_d((void *&)A->rhs);
  dfree_d((void *&)A->fu); dfree_d((void *&)A->fv); dfree_d((void *&)A->fp1);
  dfree_d((void *&)A->nman); dfree_d((void *&)A->lcd); dfree_d((void *&)A->own);
  for (auto &p : A->wk) dfree_d((void *&)p);
  for (int j = 0; j < EFDC_MAXC; ++j) {
    dfree_d((void *&)A->lam[j]);
    dfree_d(A->c0[j]);
    dfree_d(A->cset[j]);
    dfree_d(A->cadac[j]);
  }
  for (auto &p : A->ad) dfree_d((void *&)p);
  dfree_d((void *&)A->bcx);
  for (auto &p : A->lamclo) dfree_d((void *&)p);
  dfree_d((void *&)A->srcg);
  dfree_d((void *&)A->z2);
  dfree_d((void *&)A->seed_i);
  dfree_d((void *&)A->seed_v);
  dfree_d((void *&)A->st);
  if (A->st_host) (void)hipHostFree(A->st_host);
  delete A;
  c->adj = nullptr;
}

extern "C" {

int efdc_state_bytes(efdc_ctx_t *c, size_t *bytes) {
  if (!c || !bytes) return EFDC_ERR_ARG;
  size_t n = 0;
  for (const Piece &p : state_pieces(c)) n += (p.bytes + 255) & ~(size_t)255;
  *bytes = n;
  return EFDC_OK;
}

int efdc_state_save(efdc_ctx_t *c, void *dev_buf) {
  if (!c || !dev_buf) return EFDC_ERR_ARG;
  return copy_pieces(c, state_pieces(c), (char *)dev_buf, true);
}

int efdc_state_load(efdc_ctx_t *c, const void *dev_buf) {
  if (!c || !dev_buf) return EFDC_ERR_ARG;
  return copy_pieces(c, state_pieces(c), (char *)const_cast<void *>(dev_buf), false);
}

int efdc_state_alloc(efdc_ctx_t *c, void **dev_buf) {
  if (!c || !dev_buf) return EFDC_ERR_ARG;
  size_t n = 0;
  int s;
  if ((s = efdc_state_bytes(c, &n))) return s;
  if (c->adj) {
    // room for a step's adjoint recording after the state
    size_t r = 0;
    for (const Piece &p : record_pieces(c)) r += (p.bytes + 255) & ~(size_t)255;
    n += r;
  }
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (hipMalloc(dev_buf, n) != hipSuccess) {
    *dev_buf = nullptr;
    c->err = "efdc_state_alloc: hipMalloc failed";
    return EFDC_ERR_NOMEM;
  }
  return EFDC_OK;
}

int efdc_state_free(efdc_ctx_t *c, void *dev_buf) {
  if (!c) return EFDC_ERR_ARG;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (dev_buf) HIP_TRY(c, hipFree(dev_buf));
  return EFDC_OK;
}

int efdc_adj_enable(efdc_ctx_t *c, const efdc_adj_opts_t *o) {
  if (!c || !o || o->nadj < 0 || o->nadj > c->nconst || o->nadj > 16) return EFDC_ERR_ARG;
  if (c->nseg > 1) {
    c->err = "adjoint: one basin per context";
    return EFDC_ERR_ARG;
  }
  if (c->prm.ISDRY > 0) {
    c->err = "adjoint: drying (ISDRY > 0) is not differentiated";
    return EFDC_ERR_ARG;
  }
  if (!c->mutual) {
    c->err = "adjoint: the neighbour maps must be mutual (LEC(LWC(L)) = L, LNC(LSC(L)) = L)";
    return EFDC_ERR_ARG;
  }
  for (int j = 0; j < o->nadj; ++j) {
    const int sl = o->slot[j];
    if (sl < 0 || sl >= c->nconst) return EFDC_ERR_ARG;
  }
  if (c->KC > 64) {
    c->err = "adjoint: KC > 64";
    return EFDC_ERR_ARG;
  }
  adj_free(c);
  AdjState *A = new AdjState();
  c->adj = A;
  A->nadj = o->nadj;
  for (int j = 0; j < o->nadj; ++j) A->slot[j] = o->slot[j];
  A->rel_tol = o->cg_rel_tol > 0 ? o->cg_rel_tol : 1e-12;
  A->itmax = o->cg_itmax > 0 ? o->cg_itmax : c->prm.ITERM;
  const size_t L = c->LCM, KP = (size_t)c->pitch * c->KC;
  int s = 0;
  auto fail = [&](int st) {
    adj_free(c);
    return st;
  };
  for (int q = 0; q < 2; ++q)
    if ((s = alloc_d(c, &A->hp[q], L)) || (s = alloc_d(c, &A->u[q], L)) ||
        (s = alloc_d(c, &A->v[q], L)) || (s = alloc_d(c, &A->p[q], L)) ||
        (s = alloc_d(c, &A->hu[q], L)) || (s = alloc_d(c, &A->hv[q], L)))
      return fail(s);
  if ((s = alloc_d(c, &A->ps, L)) || (s = alloc_d(c, &A->mu, L)) ||
      (s = alloc_d(c, &A->rhs, L)) || (s = alloc_d(c, &A->fu, L)) ||
      (s = alloc_d(c, &A->fv, L)) || (s = alloc_d(c, &A->fp1, L)) ||
      (s = alloc_d(c, &A->nman, L)) || (s = alloc_d(c, &A->lcd, KP)) ||
      (s = alloc_d(c, &A->own, KP)))
    return fail(s);
  for (auto &p : A->wk)
    if ((s = alloc_d(c, &p, KP))) return fail(s);
  for (int j = 0; j < o->nadj; ++j) {
    const int sl = o->slot[j];
    if ((s = alloc_d(c, &A->lam[sl], KP)) || (s = dev_alloc(c, &A->c0[sl], KP * c->tsz)))
      return fail(s);
    if (c->con[sl].desc.wset > 0.0 && (s = dev_alloc(c, &A->cset[sl], KP * c->tsz)))
      return fail(s);
    if (c->con[sl].desc.isadac > 0) {
      if ((s = dev_alloc(c, &A->cadac[sl], KP * c->tsz))) return fail(s);
      if (!A->ad[0])
        for (auto &p : A->ad)
          if ((s = alloc_d(c, &p, KP + (size_t)c->pitch))) return fail(s);
    }
  }
  if ((s = dev_alloc(c, (void **)&A->z2, L)) ||
      (s = dev_alloc(c, (void **)&A->st, sizeof(CgState))))
    return fail(s);
  if (hipHostMalloc((void **)&A->st_host, sizeof(CgState)) != hipSuccess) {
    c->err = "hipHostMalloc failed";
    return fail(EFDC_ERR_NOMEM);
  }
  A->recording = true;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return EFDC_OK;
}

int efdc_adj_disable(efdc_ctx_t *c) {
  if (!c) return EFDC_ERR_ARG;
  adj_free(c);
  return EFDC_OK;
}

int efdc_adj_reset(efdc_ctx_t *c) {
  if (!c) return EFDC_ERR_ARG;
  AdjState *A = c->adj;
  if (!A) {
    c->err = "adjoint not enabled (efdc_adj_enable)";
    return EFDC_ERR_STATE;
  }
  int s;
  if ((s = adj_tables(c, A))) return s;
  const size_t L = c->LCM, KP = (size_t)c->pitch * c->KC;
  for (int q = 0; q < 2; ++q)
    for (double *p : {A->hp[q], A->u[q], A->v[q], A->p[q], A->hu[q], A->hv[q]})
      if ((s = zero(c, p, L))) return s;
  if ((s = zero(c, A->nman, L))) return s;
  for (int j = 0; j < A->nadj; ++j)
    if ((s = zero(c, A->lam[A->slot[j]], KP))) return s;
  for (int sd = 0; sd < 4; ++sd)
    if ((s = zero(c, A->lamclo[sd], A->lamclo_n[sd]))) return s;
  if ((s = zero(c, A->srcg, A->srcg_n))) return s;
  A->cur = 0;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return EFDC_OK;
}

int efdc_adj_seed_con(efdc_ctx_t *c, int32_t slot, int32_t n, const int32_t *L,
                      const int32_t *K, const double *val) {
  if (!c || n < 0 || (n > 0 && (!L || !K || !val))) return EFDC_ERR_ARG;
  AdjState *A = c->adj;
  if (!A || slot < 0 || slot >= c->nconst || !A->lam[slot]) {
    c->err = "efdc_adj_seed_con: constituent not differentiated";
    return EFDC_ERR_ARG;
  }
  if (n == 0) return EFDC_OK;
  for (int t = 0; t < n; ++t)
    if (L[t] < 2 || L[t] > c->LA || K[t] < 1 || K[t] > c->KC) return EFDC_ERR_ARG;
  // the context's stream is non-blocking: the scratch is (re)allocated only
  // once that stream is idle, and every copy into it is queued on that
  // stream, so nothing queued earlier (a zero fill, a kernel still reading
  // the previous list) can run after the copies
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (n > A->seed_cap) {
    dfree_d((void *&)A->seed_i);
    dfree_d((void *&)A->seed_v);
    A->seed_cap = 0;
    int s;
    if ((s = dev_alloc(c, (void **)&A->seed_i, sizeof(int) * 2 * (size_t)n)) ||
        (s = dev_alloc(c, (void **)&A->seed_v, sizeof(double) * (size_t)n)))
      return s;
    A->seed_cap = n;
  }
  int *dL = A->seed_i, *dK = A->seed_i + A->seed_cap;
  HIP_TRY(c, hipMemcpyAsync(dL, L, sizeof(int) * n, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(c, hipMemcpyAsync(dK, K, sizeof(int) * n, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(c, hipMemcpyAsync(A->seed_v, val, sizeof(double) * n, hipMemcpyHostToDevice,
                            c->stream));
  k_adj_seed<<<1, 64, 0, c->stream>>>(A->lam[slot], c->pitch, n, dL, dK, A->seed_v);
  HIP_TRY(c, hipGetLastError());
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return EFDC_OK;
}

int efdc_adj_step(efdc_ctx_t *c, const efdc_step_t *step, efdc_adj_stats_t *stats) {
  if (!c || !step || step->dt <= 0) return EFDC_ERR_ARG;
  AdjState *A = c->adj;
  if (!A) {
    c->err = "adjoint not enabled (efdc_adj_enable)";
    return EFDC_ERR_STATE;
  }
  if (A->bcx_n != (size_t)c->nbcrec * c->KC * 8 ||
      A->srcg_n != (size_t)c->nqsij * c->KC * c->nconst) {
    c->err = "adjoint: boundary or source tables changed since efdc_adj_reset";
    return EFDC_ERR_STATE;
  }
  int s;
  if ((s = zero(c, A->srcg, A->srcg_n))) return s;
  // settling, then transport, one differentiated constituent after the other
  for (int j = 0; j < A->nadj; ++j) {
    const int sl = A->slot[j];
    s = c->store == EFDC_STORE_FP64 ? adj_settle<double>(c, A, sl, step)
                                    : adj_settle<float>(c, A, sl, step);
    if (s) return s;
    if (c->con[sl].desc.isadac > 0) {
      if ((s = adj_adac(c, A, sl, step))) return s;
      s = c->store == EFDC_STORE_FP64 ? adj_tran<double>(c, A, sl, step, A->ad[0])
                                      : adj_tran<float>(c, A, sl, step, A->ad[0]);
      if (s) return s;
      // lambda of CON1 through the FCT bounds joins that of the old level
      const size_t n = (size_t)c->pitch * c->KC;
      k_add<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(A->lam[sl], A->ad[1], n);
      HIP_TRY(c, hipGetLastError());
      continue;
    }
    s = c->store == EFDC_STORE_FP64 ? adj_tran<double>(c, A, sl, step)
                                    : adj_tran<float>(c, A, sl, step);
    if (s) return s;
  }
  // hydrodynamics
  HPtrA h;
  for (int j = 0; j < EFDC_NUM_H; ++j) h.a[j] = c->dH[j];
  HyArgs a = hy_args(c, A, step);
  const int nb = (c->n_wet() + EFDC_BLOCK - 1) / EFDC_BLOCK;
  TIMED(c, TK_ADJ_HYDRO, k_adj_h1<<<nb, EFDC_BLOCK, 0, c->stream>>>(h, a));
  TIMED(c, TK_ADJ_HYDRO, k_adj_h2<<<nb, EFDC_BLOCK, 0, c->stream>>>(h, a));
  TIMED(c, TK_ADJ_HYDRO, k_adj_h3<<<nb, EFDC_BLOCK, 0, c->stream>>>(h, a));
  HIP_TRY(c, hipGetLastError());
  int iters = 0;
  double rsq = 0.0;
  if ((s = cg_fp64_vec(c, A->mu, A->rhs, A->st, A->st_host, A->rel_tol, A->itmax, &iters,
                       &rsq)))
    return s;
  TIMED(c, TK_ADJ_HYDRO, k_adj_h4<<<nb, EFDC_BLOCK, 0, c->stream>>>(h, a));
  for (int sd = 0; sd < 4; ++sd)
    if (c->pbc[sd].n > 0)
      k_adj_h4bc<<<(c->pbc[sd].n + 255) / 256, 256, 0, c->stream>>>(h, a, c->pbc[sd], sd,
                                                                    c->lopenbcdry);
  TIMED(c, TK_ADJ_HYDRO, k_adj_h5<<<nb, EFDC_BLOCK, 0, c->stream>>>(h, a));
  HIP_TRY(c, hipGetLastError());
  // the state before the step becomes the state to reverse next
  const int cu = A->cur;
  A->cur ^= 1;
  const size_t L = c->LCM;
  for (double *p : {A->hp[cu], A->u[cu], A->v[cu], A->p[cu], A->hu[cu], A->hv[cu]})
    if ((s = zero(c, p, L))) return s;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (stats) {
    stats->cg_iters = iters;
    stats->cg_rsq = rsq;
  }
  return EFDC_OK;
}

int efdc_adj_record_bytes(efdc_ctx_t *c, size_t *bytes) {
  if (!c || !bytes) return EFDC_ERR_ARG;
  if (!c->adj) {
    c->err = "adjoint not enabled (efdc_adj_enable)";
    return EFDC_ERR_STATE;
  }
  size_t n = 0;
  for (const Piece &p : record_pieces(c)) n += (p.bytes + 255) & ~(size_t)255;
  *bytes = n;
  return EFDC_OK;
}

int efdc_adj_record_save(efdc_ctx_t *c, void *dev_buf) {
  if (!c || !dev_buf) return EFDC_ERR_ARG;
  if (!c->adj) {
    c->err = "adjoint not enabled (efdc_adj_enable)";
    return EFDC_ERR_STATE;
  }
  return copy_pieces(c, record_pieces(c), (char *)dev_buf, true);
}

int efdc_adj_record_load(efdc_ctx_t *c, const void *dev_buf) {
  if (!c || !dev_buf) return EFDC_ERR_ARG;
  if (!c->adj) {
    c->err = "adjoint not enabled (efdc_adj_enable)";
    return EFDC_ERR_STATE;
  }
  return copy_pieces(c, record_pieces(c), (char *)const_cast<void *>(dev_buf), false);
}

int efdc_adj_source_grad(efdc_ctx_t *c, double *out) {
  if (!c || !out) return EFDC_ERR_ARG;
  AdjState *A = c->adj;
  if (!A) return EFDC_ERR_STATE;
  if (A->srcg_n == 0) return EFDC_OK;
  HIP_TRY(c, hipMemcpyAsync(out, A->srcg, sizeof(double) * A->srcg_n, hipMemcpyDeviceToHost,
                            c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return EFDC_OK;
}

int efdc_adj_get(efdc_ctx_t *c, int32_t field, double *out) {
  if (!c || !out) return EFDC_ERR_ARG;
  AdjState *A = c->adj;
  if (!A) return EFDC_ERR_STATE;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  const int cu = A->cur;
  const double *src = nullptr;
  switch (field) {
    case EFDC_ADJ_MANNING: src = A->nman; break;
    case EFDC_ADJ_HP: src = A->hp[cu]; break;
    case EFDC_ADJ_UHDYE: src = A->u[cu]; break;
    case EFDC_ADJ_VHDXE: src = A->v[cu]; break;
    case EFDC_ADJ_P: src = A->p[cu]; break;
    case EFDC_ADJ_HU: src = A->hu[cu]; break;
    case EFDC_ADJ_HV: src = A->hv[cu]; break;
    default: break;
  }
  if (src) {
    HIP_TRY(c, hipMemcpyAsync(out, src, sizeof(double) * c->LCM, hipMemcpyDeviceToHost,
                              c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return EFDC_OK;
  }
  const int sl = field - EFDC_ADJ_CON;
  if (sl < 0 || sl >= c->nconst || !A->lam[sl]) return EFDC_ERR_ARG;
  return get_planes(c, A->lam[sl], out);
}

}  // extern "C"
