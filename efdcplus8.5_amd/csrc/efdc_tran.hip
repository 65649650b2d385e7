// efdc_tran.hip -- CALCONC(ISTL=2, IS2TL=1) on gfx950.
//
// The transport step of every constituent in one pass over the water
// columns: at every layer the upwind update of CALTRAN
// (EFDC/caltran.f90:95-111,188-245) with the CALFQC sources
// (EFDC/calfqc.f90:209-225), the open-boundary sequence
// (EFDC/caltran.f90:316-471) on boundary columns, and the forward sweep of
// the implicit vertical diffusion (EFDC/calconc.f90:264-422) in FP64; the
// back substitution (EFDC/calconc.f90:425-468) and the dye decay
// (EFDC/calconc.f90:581-616) then write the new concentrations.  CON and
// CON1 are two buffers whose roles swap every step (CON1 = CON of
// EFDC/caltran.f90:59-65 costs nothing).
//
// k_tran_colm (the default with mutual neighbour maps) runs one thread per
// column and 1, 2 or 4 constituents: the transport fields and the
// tridiagonal factors are loaded and computed once per layer for all of
// them.  k_tran_col (any maps) runs one thread per (column, constituent):
// a workgroup is 64 columns x up to 4 constituent waves.  In both, each
// layer's loads are issued before the previous layer is computed.
// k_tran_upwind_vdiff (one thread per column, all constituents, LDS
// staging) remains for more than 64 layers.  HBM traffic per column-step is
//   KC*s*(3 flows + H1PK + HPKI + AB + 2*NC) + maps/metrics,
// the roofline of this step.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "efdc_internal.h"

namespace {

#define TBLOCK 128

template <typename T>
struct TranArgs {
  const T *uhdy2, *vhdx2, *w2, *h1pk, *hpki, *ab, *cdzkk, *cdzkmk;
  const double *dxyip, *hpi;
  const double *cdzkk_k, *cdzkmk_k;  // per-K constants when uniform
  int dz_uniform;
  const int *lw, *le, *ls, *ln;
  int nwet, lc0, KC, nc, rowmajor;
  int64_t pitch;
  double dt;
  long long niter;
  const T *old[EFDC_MAXC];
  T *nw[EFDC_MAXC];
  int mo[EFDC_MAXC];
  int dye_mode[EFDC_MAXC];   // 0 none, 1 decay factor, 2 age
  double dye_f[EFDC_MAXC];   // factor or age increment

  // sources
  const int *srcrec, *srcnext;
  const double *qss, *qsercell, *cqs, *csertq;
  // open boundaries
  const int *bcrec;
  const int4 *bcinfo;
  CbcDev cbc[4];
  const double *wtci;
  double *bc_con1;
  int vd;  // 0: stop after the open boundaries (anti-diffusion follows)
  const int *lmask;  // LMASKDRY (0: column left out), or null without drying
};

template <typename T>
__device__ __forceinline__ double ld(const T *p, int64_t j) {
  return (double)p[j];
}

// MAX/MIN with the semantics of the restatement the kernels are compared
// with bit for bit
__device__ __forceinline__ double dmx(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double dmn(double a, double b) { return a < b ? a : b; }

// A column whose cell is left out of the transport (LMASKDRY false,
// EFDC/calpuv2c.f90:916-939: every CALCONC/CALTRAN loop runs over the wet
// list LKWET, EFDC/caltran.f90:95-300,324-441, EFDC/calconc.f90:89-121,
// 259-473) keeps CON for constituents c0, c0+cstep, .. < cend.  The faces
// of such a cell were closed at both time levels, so UHDY2/VHDX2 are 0
// there and the fluxes its wet neighbours see are 0 with or without it.
template <typename T>
__device__ __forceinline__ bool col_keep(const TranArgs<T> &a, int i, int c0, int cstep,
                                         int cend) {
  if (!a.lmask || a.lmask[i] != 0) return false;
  for (int c = c0; c < cend; c += cstep)
    for (int k = 0; k < a.KC; ++k) a.nw[c][i + (int64_t)k * a.pitch] = a.old[c][i + (int64_t)k * a.pitch];
  return true;
}

template <typename T>
__global__ __launch_bounds__(TBLOCK) void k_tran_upwind_vdiff(TranArgs<T> a) {
  extern __shared__ unsigned char smem[];
  const int TB = blockDim.x;  // 128, or 64 when the LDS staging needs it
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int i = 1 + tile * TB + (int)threadIdx.x;
  if (i > a.nwet) return;
  if (col_keep(a, i, 0, 1, a.nc)) return;
  const int tid = threadIdx.x;
  const int KC = a.KC, nc = a.nc;
  const int64_t pitch = a.pitch;
  // LDS: VCU[K][tid], FWUU(K-1)[c][tid] (double), forward values
  // y[K][c][tid] (T)
  double *vcu = (double *)smem;
  double *fwp = vcu + (size_t)KC * TB;
  T *ys = (T *)(fwp + (size_t)nc * TB);
  auto Y = [&](int k, int c) -> T & { return ys[((size_t)k * nc + c) * TB + tid]; };

  const int w = a.lw[i], s = a.ls[i];
  const int e = a.rowmajor ? i + 1 : a.le[i];
  const int n = a.ln[i];
  const bool eg = a.rowmajor ? (a.le[i] == a.lc0) : (e == a.lc0);
  const bool ng = (n == a.lc0);
  const int w_of_e = eg ? i : (a.rowmajor ? i : a.lw[e]);
  const int s_of_n = ng ? i : a.ls[n];
  const double DDELT = a.dt, DELTD2 = a.dt;
  const double dxyip = a.dxyip[i], hpi = a.hpi[i];
  const int src0 = a.srcrec[i];
  const int rec = a.bcrec ? a.bcrec[i] : -1;
  int4 bi = make_int4(-1, -1, -1, -1);
  if (rec >= 0) bi = a.bcinfo[rec];

  for (int c = 0; c < nc; ++c) fwp[c * TB + tid] = 0.0;  // FWUU(L,0)=0
  double vcu_prev = 0.0;

  for (int k = 0; k < KC; ++k) {
    const int K = k + 1;
    const int64_t o = (int64_t)k * pitch;
    const double u_i = ld(a.uhdy2, i + o);
    const double u_e = eg ? 0.0 : ld(a.uhdy2, e + o);
    const double v_i = ld(a.vhdx2, i + o);
    const double v_n = ng ? 0.0 : ld(a.vhdx2, n + o);
    const double ww = (K < KC) ? ld(a.w2, i + (int64_t)K * pitch) : 0.0;
    const double h1pk = ld(a.h1pk, i + o), hpki = ld(a.hpki, i + o);
    // vertical diffusion coefficients of this layer (FP64)
    double cclb = 0.0, eeb = 1.0, vcu_k = 0.0;
    if (KC > 1) {
      const double ckk = a.dz_uniform ? a.cdzkk_k[k] : ld(a.cdzkk, i + o);
      const double ckm = a.dz_uniform ? a.cdzkmk_k[k] : ld(a.cdzkmk, i + o);
      if (K == 1) {
        const double RCDZKK = -DELTD2 * ckk;
        const double CCUBTMP = RCDZKK * hpi * ld(a.ab, i + o);
        const double CCMBTMP = 1.0 - CCUBTMP;
        eeb = 1.0 / CCMBTMP;
        vcu_k = CCUBTMP * eeb;
      } else if (K < KC) {
        const double RCDZKMK = -DELTD2 * ckm;
        const double RCDZKK = -DELTD2 * ckk;
        cclb = RCDZKMK * hpi * ld(a.ab, i + o - pitch);
        const double CCUBTMP = RCDZKK * hpi * ld(a.ab, i + o);
        const double CCMBTMP = 1.0 - cclb - CCUBTMP;
        eeb = 1.0 / (CCMBTMP - cclb * vcu_prev);
        vcu_k = CCUBTMP * eeb;
      } else {
        const double RCDZKMK = -DELTD2 * ckm;
        cclb = RCDZKMK * hpi * ld(a.ab, i + o - pitch);
        const double CCMBTMP = 1.0 - cclb;
        eeb = 1.0 / (CCMBTMP - cclb * vcu_prev);
      }
      vcu[(size_t)k * TB + tid] = vcu_k;
    }
    for (int c = 0; c < nc; ++c) {
      const T *C1 = a.old[c];
      const double c1 = ld(C1, i + o);
      const double cw = ld(C1, w + o);
      const double cs = ld(C1, s + o);
      // upwind fluxes (LUPU/LUPV/KUPW of EFDC/calconc.f90:95-118)
      const double fu_i = u_i * (u_i >= 0.0 ? cw : c1);
      const double fv_i = v_i * (v_i >= 0.0 ? cs : c1);
      double fu_e = 0.0, fv_n = 0.0;
      if (!eg) fu_e = u_e * (u_e >= 0.0 ? ld(C1, w_of_e + o) : ld(C1, e + o));
      if (!ng) fv_n = v_n * (v_n >= 0.0 ? ld(C1, s_of_n + o) : ld(C1, n + o));
      double fw = 0.0;
      if (K < KC) fw = ww * (ww >= 0.0 ? c1 : ld(C1, i + o + pitch));
      // sources, EFDC/calfqc.f90:211-224
      double fqc = 0.0;
      for (int q = src0; q >= 0; q = a.srcnext[q]) {
        const double conq = 0.5 * (3. * c1 - c1);
        const size_t j = (size_t)q * KC + k;
        const double qv = a.qss[j], qs = a.qsercell[j];
        fqc = fqc + dmx(qv, 0.) * a.cqs[j * nc + c] + dmn(qv, 0.) * conq +
              dmx(qs, 0.) * a.csertq[j * nc + c] + dmn(qs, 0.) * conq;
      }
      // EFDC/caltran.f90:194-196,243
      const double cd = c1 * h1pk + DDELT * ((fqc + fu_i - fu_e + fv_i - fv_n) * dxyip +
                                             (fwp[c * TB + tid] - fw));
      double con = cd * hpki;
      fwp[c * TB + tid] = fw;
      if (rec >= 0) {
        // open boundaries S, W, E, N (EFDC/caltran.f90:319-471)
        double cv1 = c1;
        const int LLs[4] = {bi.x, bi.y, bi.z, bi.w};
        for (int side = 0; side < 4; ++side) {
          const int LL = LLs[side];
          if (LL < 0) continue;
          const CbcDev &b = a.cbc[side];
          double flow, flux;
          bool out;
          if (side == 0) { flow = v_n; flux = fv_n; out = flow <= 0.; }
          else if (side == 1) { flow = u_e; flux = fu_e; out = flow <= 0.; }
          else if (side == 2) { flow = u_i; flux = fu_i; out = flow >= 0.; }
          else { flow = v_i; flux = fv_i; out = flow >= 0.; }
          const size_t m = ((size_t)LL * KC + k) * nc + c;
          const double cb1 = b.cb[((size_t)LL * 2 + 0) * nc + c];
          const double cb2 = b.cb[((size_t)LL * 2 + 1) * nc + c];
          if (out) {
            double CTMP;
            if (side <= 1) CTMP = cv1 + DDELT * (flow * cv1 - flux) * dxyip * hpki;
            else CTMP = cv1 + DDELT * (flux - flow * cv1) * dxyip * hpki;
            con = dmx(CTMP, 0.);
            if (a.mo[c] == 1) {
              const double cbtmp = cb1 + b.csert[((size_t)LL * KC + 0) * nc + c];
              if (con > cbtmp) con = cbtmp;
            }
            b.clo[m] = con;
            b.nlo[m] = a.niter;
          } else {
            const double CBT = a.wtci[k * 2 + 0] * cb1 + a.wtci[k * 2 + 1] * cb2 + b.csert[m];
            const long long nmnlod = a.niter - b.nlo[m];
            if (nmnlod >= b.ntscr[LL]) {
              con = CBT;
            } else {
              const double cbtmp =
                  b.clo[m] + (CBT - b.clo[m]) * (double)nmnlod / (double)b.ntscr[LL];
              con = dmx(cbtmp, 0.);
            }
          }
          cv1 = con;
        }
        a.bc_con1[((size_t)rec * KC + k) * nc + c] = con;
      }
      if (!a.vd) {
        a.nw[c][i + o] = (T)con;
        continue;
      }
      // forward sweep of the tridiagonal solve (FP64)
      double y;
      if (KC == 1) y = con;
      else if (K == 1) y = con * eeb;
      else y = (con - cclb * (double)Y(k - 1, c)) * eeb;
      Y(k, c) = (T)y;
    }
    vcu_prev = vcu_k;
  }
  if (!a.vd) return;
  // back substitution and dye decay, write CON
  for (int c = 0; c < nc; ++c) {
    T *C = a.nw[c];
    double up = (double)Y(KC - 1, c);
    const int dm = a.dye_mode[c];
    const double f = a.dye_f[c];
    auto fin = [&](double v) { return dm == 1 ? f * v : (dm == 2 ? v + f : v); };
    C[i + (int64_t)(KC - 1) * pitch] = (T)fin(up);
    for (int k = KC - 2; k >= 0; --k) {
      const double v = (double)Y(k, c) - vcu[(size_t)k * TB + tid] * up;
      C[i + (int64_t)k * pitch] = (T)fin(v);
      up = v;
    }
  }
}

// Same step as k_tran_upwind_vdiff with one thread per (column,
// constituent): the workgroup is 64 columns x NCG constituent lanes (a
// wave per lane, so every wave streams one constituent's plane with
// coalesced loads while the transport fields it shares with the other
// lanes come from L1/L2).  VCU of the tridiagonal solve (KC doubles per
// column, shared by the constituents) and the forward values (KC values of
// the storage type per thread, as the LDS staging of k_tran_upwind_vdiff
// keeps them) stay in LDS, so the new concentration plane is written once.
// Arithmetic is expression for expression that of k_tran_upwind_vdiff, so
// both kernels give identical bits.
// Inputs of one layer of one column for one constituent.  The upwind
// donor values are loaded for both flow signs and selected afterwards, so
// that all loads of a layer are independent of each other and of the
// previous layer (the kernel issues layer K+1's loads before computing
// layer K).
template <typename T>
struct LayerIn {
  T u_i, u_e, v_i, v_n, ww, h1pk, hpki, ab;  // storage type: exact
  double ckk, ckm;
  T c1, cw, cs, ce_w, ce, cn_s, cn, cup;
};

// Source term FQC of one layer and constituent, EFDC/calfqc.f90:211-224
// (a separate function keeps the column loop compact).
template <typename T>
__device__ __forceinline__ double col_sources(const TranArgs<T> &a, int src0, int k, int c,
                                           double c1) {
  const int KC = a.KC, nc = a.nc;
  double fqc = 0.0;
  for (int q = src0; q >= 0; q = a.srcnext[q]) {
    const double conq = 0.5 * (3. * c1 - c1);
    const size_t j = (size_t)q * KC + k;
    const double qv = a.qss[j], qs = a.qsercell[j];
    fqc = fqc + dmx(qv, 0.) * a.cqs[j * nc + c] + dmn(qv, 0.) * conq +
          dmx(qs, 0.) * a.csertq[j * nc + c] + dmn(qs, 0.) * conq;
  }
  return fqc;
}

// Open-boundary sequence S, W, E, N of one layer and constituent,
// EFDC/caltran.f90:319-471 (a separate function keeps the column loop
// compact).  Returns the new concentration.
template <typename T>
__device__ __forceinline__ double col_open_bc(const TranArgs<T> &a, int4 bi, int rec, int k,
                                           int c, double c1, double con, double u_i,
                                           double u_e, double v_i, double v_n, double fu_i,
                                           double fu_e, double fv_i, double fv_n,
                                           double dxyip, double hpki) {
  const int KC = a.KC, nc = a.nc;
  const double DDELT = a.dt;
  double cv1 = c1;
  const int LLs[4] = {bi.x, bi.y, bi.z, bi.w};
  for (int side = 0; side < 4; ++side) {
    const int LL = LLs[side];
    if (LL < 0) continue;
    const CbcDev &b = a.cbc[side];
    double flow, flux;
    bool out;
    if (side == 0) { flow = v_n; flux = fv_n; out = flow <= 0.; }
    else if (side == 1) { flow = u_e; flux = fu_e; out = flow <= 0.; }
    else if (side == 2) { flow = u_i; flux = fu_i; out = flow >= 0.; }
    else { flow = v_i; flux = fv_i; out = flow >= 0.; }
    const size_t m = ((size_t)LL * KC + k) * nc + c;
    const double cb1 = b.cb[((size_t)LL * 2 + 0) * nc + c];
    const double cb2 = b.cb[((size_t)LL * 2 + 1) * nc + c];
    if (out) {
      double CTMP;
      if (side <= 1) CTMP = cv1 + DDELT * (flow * cv1 - flux) * dxyip * hpki;
      else CTMP = cv1 + DDELT * (flux - flow * cv1) * dxyip * hpki;
      con = dmx(CTMP, 0.);
      if (a.mo[c] == 1) {
        const double cbtmp = cb1 + b.csert[((size_t)LL * KC + 0) * nc + c];
        if (con > cbtmp) con = cbtmp;
      }
      b.clo[m] = con;
      b.nlo[m] = a.niter;
    } else {
      const double CBT = a.wtci[k * 2 + 0] * cb1 + a.wtci[k * 2 + 1] * cb2 + b.csert[m];
      const long long nmnlod = a.niter - b.nlo[m];
      if (nmnlod >= b.ntscr[LL]) {
        con = CBT;
      } else {
        const double cbtmp =
            b.clo[m] + (CBT - b.clo[m]) * (double)nmnlod / (double)b.ntscr[LL];
        con = dmx(cbtmp, 0.);
      }
    }
    cv1 = con;
  }
  a.bc_con1[((size_t)rec * KC + k) * nc + c] = con;
  return con;
}

template <typename T>
__device__ __forceinline__ void tran_col_body(const TranArgs<T> &a, double *vcus) {
  // LDS: VCU of the tridiagonal solve [KC][64] (FP64, the same for every
  // constituent of a column: each lane computes it from the same inputs,
  // so the waves of a block write identical values), then the forward
  // values [blockDim.y][KC][64] in the storage type
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int i = 1 + tile * 64 + (int)threadIdx.x;
  if (i > a.nwet) return;
  if (col_keep(a, i, (int)threadIdx.y, (int)blockDim.y, a.nc)) return;
  const int KC = a.KC, nc = a.nc;
  // 32-bit element offsets (the host checks that every layer array has
  // fewer than 2^31 elements), so the loads use a scalar base plus a
  // 32-bit vector offset instead of a 64-bit address per pointer
  const unsigned pitch = (unsigned)a.pitch;
  const int w = a.lw[i], s = a.ls[i];
  const int e = a.rowmajor ? i + 1 : a.le[i];
  const int n = a.ln[i];
  const bool eg = a.rowmajor ? (a.le[i] == a.lc0) : (e == a.lc0);
  const bool ng = (n == a.lc0);
  const int w_of_e = eg ? i : (a.rowmajor ? i : a.lw[e]);
  const int s_of_n = ng ? i : a.ls[n];
  const unsigned ui = (unsigned)i, uw = (unsigned)w, us = (unsigned)s, ue = (unsigned)e,
                 un = (unsigned)n, uwe = (unsigned)w_of_e, usn = (unsigned)s_of_n;
  const double DDELT = a.dt, DELTD2 = a.dt;
  const double dxyip = a.dxyip[i], hpi = a.hpi[i];
  const int src0 = a.srcrec[i];
  const int rec = a.bcrec ? a.bcrec[i] : -1;
  int4 bi = make_int4(-1, -1, -1, -1);
  if (rec >= 0) bi = a.bcinfo[rec];
  const bool vdk = a.vd && KC > 1;

  double *vcu = vcus + threadIdx.x;  // stride 64
  T *ys = (T *)(vcus + (size_t)KC * 64) + (size_t)threadIdx.y * KC * 64 + threadIdx.x;
  for (int c = threadIdx.y; c < nc; c += blockDim.y) {
    const T *__restrict__ C1 = a.old[c];
    T *__restrict__ C = a.nw[c];
    auto load = [&](int k, LayerIn<T> &L) {
      const unsigned o = (unsigned)k * pitch;
      const int K = k + 1;
      L.u_i = a.uhdy2[ui + o];
      L.u_e = eg ? (T)0 : a.uhdy2[ue + o];
      L.v_i = a.vhdx2[ui + o];
      L.v_n = ng ? (T)0 : a.vhdx2[un + o];
      L.ww = (K < KC) ? a.w2[ui + o + pitch] : (T)0;
      L.h1pk = a.h1pk[ui + o];
      L.hpki = a.hpki[ui + o];
      L.ab = vdk && K < KC ? a.ab[ui + o] : (T)0;
      L.ckk = vdk ? (a.dz_uniform ? a.cdzkk_k[k] : (double)a.cdzkk[ui + o]) : 0.0;
      L.ckm = vdk ? (a.dz_uniform ? a.cdzkmk_k[k] : (double)a.cdzkmk[ui + o]) : 0.0;
      L.c1 = C1[ui + o];
      L.cw = C1[uw + o];
      L.cs = C1[us + o];
      L.ce_w = eg ? (T)0 : C1[uwe + o];
      L.ce = eg ? (T)0 : C1[ue + o];
      L.cn_s = ng ? (T)0 : C1[usn + o];
      L.cn = ng ? (T)0 : C1[un + o];
      L.cup = (K < KC) ? C1[ui + o + pitch] : (T)0;
    };
    LayerIn<T> cur, nxt;
    load(0, cur);
    double fwp = 0.0;  // FWUU(L,K-1), FWUU(L,0)=0
    double vcu_prev = 0.0, yprev = 0.0, ab_prev = 0.0;
    for (int k = 0; k < KC; ++k) {
      const int K = k + 1;
      const unsigned o = (unsigned)k * pitch;
      if (K < KC) load(k + 1, nxt);
      const double u_i = (double)cur.u_i, u_e = (double)cur.u_e, v_i = (double)cur.v_i,
                   v_n = (double)cur.v_n;
      const double ww = (double)cur.ww, h1pk = (double)cur.h1pk, hpki = (double)cur.hpki;
      double cclb = 0.0, eeb = 1.0, vcu_k = 0.0;
      if (vdk) {
        const double ckk = cur.ckk, ckm = cur.ckm;
        if (K == 1) {
          const double RCDZKK = -DELTD2 * ckk;
          const double CCUBTMP = RCDZKK * hpi * (double)cur.ab;
          const double CCMBTMP = 1.0 - CCUBTMP;
          eeb = 1.0 / CCMBTMP;
          vcu_k = CCUBTMP * eeb;
        } else if (K < KC) {
          const double RCDZKMK = -DELTD2 * ckm;
          const double RCDZKK = -DELTD2 * ckk;
          cclb = RCDZKMK * hpi * ab_prev;
          const double CCUBTMP = RCDZKK * hpi * (double)cur.ab;
          const double CCMBTMP = 1.0 - cclb - CCUBTMP;
          eeb = 1.0 / (CCMBTMP - cclb * vcu_prev);
          vcu_k = CCUBTMP * eeb;
        } else {
          const double RCDZKMK = -DELTD2 * ckm;
          cclb = RCDZKMK * hpi * ab_prev;
          const double CCMBTMP = 1.0 - cclb;
          eeb = 1.0 / (CCMBTMP - cclb * vcu_prev);
        }
        vcu[k * 64] = vcu_k;
      }
      const double c1 = (double)cur.c1;
      // upwind fluxes (LUPU/LUPV/KUPW of EFDC/calconc.f90:95-118)
      const double fu_i = u_i * (u_i >= 0.0 ? (double)cur.cw : c1);
      const double fv_i = v_i * (v_i >= 0.0 ? (double)cur.cs : c1);
      double fu_e = 0.0, fv_n = 0.0;
      if (!eg) fu_e = u_e * (u_e >= 0.0 ? (double)cur.ce_w : (double)cur.ce);
      if (!ng) fv_n = v_n * (v_n >= 0.0 ? (double)cur.cn_s : (double)cur.cn);
      double fw = 0.0;
      if (K < KC) fw = ww * (ww >= 0.0 ? c1 : (double)cur.cup);
      // sources, EFDC/calfqc.f90:211-224
      const double fqc = src0 >= 0 ? col_sources(a, src0, k, c, c1) : 0.0;
      // EFDC/caltran.f90:194-196,243
      const double cd = c1 * h1pk + DDELT * ((fqc + fu_i - fu_e + fv_i - fv_n) * dxyip +
                                             (fwp - fw));
      double con = cd * hpki;
      fwp = fw;
      if (rec >= 0)
        con = col_open_bc(a, bi, rec, k, c, c1, con, u_i, u_e, v_i, v_n, fu_i, fu_e, fv_i,
                          fv_n, dxyip, hpki);
      ab_prev = (double)cur.ab;
      cur = nxt;
      if (!a.vd) {
        C[ui + o] = (T)con;
        continue;
      }
      // forward sweep of the tridiagonal solve (FP64), forward values
      // kept in the storage type
      double y;
      if (KC == 1) y = con;
      else if (K == 1) y = con * eeb;
      else y = (con - cclb * yprev) * eeb;
      const T yt = (T)y;
      ys[k * 64] = yt;
      yprev = (double)yt;
      vcu_prev = vcu_k;
    }
    if (!a.vd) continue;
    // back substitution and dye decay, write CON
    const int dm = a.dye_mode[c];
    const double f = a.dye_f[c];
    auto fin = [&](double v) { return dm == 1 ? f * v : (dm == 2 ? v + f : v); };
    double up = (double)ys[(KC - 1) * 64];
    C[ui + (unsigned)(KC - 1) * pitch] = (T)fin(up);
    for (int k = KC - 2; k >= 0; --k) {
      const unsigned o = (unsigned)k * pitch;
      const double v = (double)ys[k * 64] - vcu[k * 64] * up;
      C[ui + o] = (T)fin(v);
      up = v;
    }
  }
}

// Two register budgets of the same column kernel: the compiler's choice,
// and at most 168 VGPRs (three waves per SIMD, a few spills for FP64).
// Which is faster depends on KC, the constituent count and the storage
// type; the autotuner times both.
template <typename T>
__global__ __launch_bounds__(256) void k_tran_col(TranArgs<T> a) {
  extern __shared__ double vcus[];
  tran_col_body<T>(a, vcus);
}

template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_tran_col_w3(
    TranArgs<T> a) {
  extern __shared__ double vcus[];
  tran_col_body<T>(a, vcus);
}

// ---------------------------------------------------------------------
// k_tran_colm: the same column step with NCW constituents per thread.
//
// One wave owns 64 columns and NCW constituents and walks the layers once:
// the transport fields (flows, H1PK, HPKI, AB) are loaded and converted and
// the tridiagonal factors of the vertical solve (EEB, CCLB, VCU, with their
// FP64 divide) are computed once per layer for all NCW constituents, where
// the one-constituent kernel repeats them per constituent.  The kernel
// relies on mutual neighbour maps (LWC(LEC(L)) = L and LSC(LNC(L)) = L for
// wet neighbours), so the donor cells of the east and north faces that lie
// in the column itself are CON1 of the column: six CON1 loads per
// constituent and layer (the column, W, S, E, N, and the layer above).
//
// The loads of layer k+1 are issued before layer k is computed, and the
// loop body issues the same loads on every path (a missing neighbour or
// layer reads a valid address whose value is then not used), so the
// compiler's wait counts stay exact and the prefetch is never drained
// early.  Waves without a source or open-boundary column (all but a few)
// run a copy of the loop without those branches (RARE = false).
// Every value and operation order is that of tran_col_body, so both
// kernels give the same bits.  LDS: VCU [KC][64] (FP64, shared by the
// waves of a block, which all compute the same values) and the forward
// values [waves][NCW][KC][64].
// ---------------------------------------------------------------------
// The loads of one layer: the transport fields (shared by the thread's
// constituents) and six CON1 values per constituent.
template <typename T>
struct FlowL {
  T u_i, u_e, v_i, v_n, ww, h1pk, hpki, ab;
  T ckk, ckm;  // CDZKK/CDZKMK (not used when they are uniform)
};
template <typename T, int NCW>
struct ConL {
  T c1[NCW], cw[NCW], cs[NCW], ce[NCW], cn[NCW], cup[NCW];
};

template <typename T>
__device__ __forceinline__ T ldb(const T *base, unsigned boff) {
  return *(const T *)((const char *)base + boff);
}
template <typename T>
__device__ __forceinline__ void stb(T x, T *base, unsigned boff) {
  *(T *)((char *)base + boff) = x;
}

// FPF: how many layers ahead the transport fields are loaded (1 or 2; the
// CON1 values are loaded one layer ahead).  With FPF = 2 a layer's CON1
// loads are issued before the transport fields of the layer after next, so
// that waiting for them does not wait for those.
template <typename T, int NCW, bool RARE, int FPF>
__device__ __forceinline__ void tran_colm_run(const TranArgs<T> &a, double *vcus, int i, int c0,
                                              int src0, int rec) {
  const int KC = a.KC, nc = a.nc;
  const int w = a.lw[i], s = a.ls[i];
  const int e = a.rowmajor ? i + 1 : a.le[i];
  const int n = a.ln[i];
  const bool eg = a.rowmajor ? (a.le[i] == a.lc0) : (e == a.lc0);
  const bool ng = (n == a.lc0);
  // the east / north values of a column without that neighbour are read
  // from the column itself and then not used (flux 0, face flow 0)
  const unsigned bi = (unsigned)i * sizeof(T), bw = (unsigned)w * sizeof(T),
                 bs = (unsigned)s * sizeof(T), be = eg ? bi : (unsigned)e * sizeof(T),
                 bn = ng ? bi : (unsigned)n * sizeof(T);
  const size_t pitch = (size_t)a.pitch;
  const double DDELT = a.dt, DELTD2 = a.dt;
  const double dxyip = a.dxyip[i], hpi = a.hpi[i];
  int4 bc = make_int4(-1, -1, -1, -1);
  if (RARE && rec >= 0) bc = a.bcinfo[rec];
  const bool vdk = a.vd && KC > 1;
  // CDZKK/CDZKMK planes when they vary by column, else any plane that is
  // read anyway (the values are then taken from the per-layer table)
  const bool colz = vdk && !a.dz_uniform;
  const T *pkk = colz ? a.cdzkk : a.hpki, *pkm = colz ? a.cdzkmk : a.hpki;
  bool on[NCW];
  const T *C1[NCW];
#pragma unroll
  for (int j = 0; j < NCW; ++j) {
    // the launch has ceil(nc/NCW) waves, so the first slot is always used
    on[j] = j == 0 || c0 + j < nc;
    C1[j] = a.old[on[j] ? c0 + j : 0];
  }

  double *vcu = vcus + threadIdx.x;  // stride 64
  T *ys = (T *)(vcus + (size_t)KC * 64) + (size_t)threadIdx.y * NCW * KC * 64 + threadIdx.x;
  // layer k (clamped to the top layer, so that every trip issues the same
  // loads); the upward neighbour plane is k+1 below the top, else k
  auto clampk = [&](int k) { return k < KC ? k : KC - 1; };
  auto loadF = [&](int k, FlowL<T> &L) {
    const size_t o = (size_t)k * pitch;
    const T *U = a.uhdy2 + o, *V = a.vhdx2 + o;
    L.u_i = ldb(U, bi);
    L.u_e = ldb(U, be);
    L.v_i = ldb(V, bi);
    L.v_n = ldb(V, bn);
    L.ww = ldb(a.w2 + o + pitch, bi);  // W2 has KC+1 planes
    L.h1pk = ldb(a.h1pk + o, bi);
    L.hpki = ldb(a.hpki + o, bi);
    L.ab = ldb(a.ab + o, bi);
    L.ckk = ldb(pkk + o, bi);
    L.ckm = ldb(pkm + o, bi);
  };
  auto loadC = [&](int k, ConL<T, NCW> &L) {
    const size_t o = (size_t)k * pitch;
    const size_t ou = (size_t)(k + 1 < KC ? k + 1 : k) * pitch;
#pragma unroll
    for (int j = 0; j < NCW; ++j) {
      if (!on[j]) continue;
      const T *P = C1[j] + o;
      L.c1[j] = ldb(P, bi);
      L.cw[j] = ldb(P, bw);
      L.cs[j] = ldb(P, bs);
      L.ce[j] = ldb(P, be);
      L.cn[j] = ldb(P, bn);
      L.cup[j] = ldb(C1[j] + ou, bi);
    }
  };
  FlowL<T> F0, F1, F2;
  ConL<T, NCW> Q0, Q1;
  loadC(0, Q0);
  loadF(0, F0);
  if (FPF == 2) loadF(clampk(1), F1);
  double fwp[NCW], yprev[NCW];
#pragma unroll
  for (int j = 0; j < NCW; ++j) fwp[j] = yprev[j] = 0.0;
  double vcu_prev = 0.0, ab_prev = 0.0;
  for (int k = 0; k < KC; ++k) {
    const int K = k + 1;
    const size_t o = (size_t)k * pitch;
    loadC(clampk(k + 1), Q1);
    if (FPF == 2) loadF(clampk(k + 2), F2);
    else loadF(clampk(k + 1), F1);
    const bool below_top = K < KC;
    const double u_i = (double)F0.u_i, u_e = eg ? 0.0 : (double)F0.u_e,
                 v_i = (double)F0.v_i, v_n = ng ? 0.0 : (double)F0.v_n;
    const double ww = below_top ? (double)F0.ww : 0.0;
    const double h1pk = (double)F0.h1pk, hpki = (double)F0.hpki;
    const double abk = (vdk && below_top) ? (double)F0.ab : 0.0;
    double cclb = 0.0, eeb = 1.0, vcu_k = 0.0;
    if (vdk) {
      const double ckk = a.dz_uniform ? a.cdzkk_k[k] : (double)F0.ckk;
      const double ckm = a.dz_uniform ? a.cdzkmk_k[k] : (double)F0.ckm;
      if (K == 1) {
        const double RCDZKK = -DELTD2 * ckk;
        const double CCUBTMP = RCDZKK * hpi * abk;
        const double CCMBTMP = 1.0 - CCUBTMP;
        eeb = 1.0 / CCMBTMP;
        vcu_k = CCUBTMP * eeb;
      } else if (K < KC) {
        const double RCDZKMK = -DELTD2 * ckm;
        const double RCDZKK = -DELTD2 * ckk;
        cclb = RCDZKMK * hpi * ab_prev;
        const double CCUBTMP = RCDZKK * hpi * abk;
        const double CCMBTMP = 1.0 - cclb - CCUBTMP;
        eeb = 1.0 / (CCMBTMP - cclb * vcu_prev);
        vcu_k = CCUBTMP * eeb;
      } else {
        const double RCDZKMK = -DELTD2 * ckm;
        cclb = RCDZKMK * hpi * ab_prev;
        const double CCMBTMP = 1.0 - cclb;
        eeb = 1.0 / (CCMBTMP - cclb * vcu_prev);
      }
      vcu[k * 64] = vcu_k;
    }
    const bool pu = F0.u_i >= (T)0, pv = F0.v_i >= (T)0, pue = F0.u_e >= (T)0,
               pvn = F0.v_n >= (T)0, pw = F0.ww >= (T)0;
#pragma unroll
    for (int j = 0; j < NCW; ++j) {
      if (!on[j]) continue;
      const int cc = c0 + j;
      const double c1 = (double)Q0.c1[j];
      // upwind fluxes (LUPU/LUPV/KUPW of EFDC/calconc.f90:95-118); the
      // donor is chosen in the storage type, where the sign test and the
      // value are those of the FP64 conversion
      const double fu_i = u_i * (double)(pu ? Q0.cw[j] : Q0.c1[j]);
      const double fv_i = v_i * (double)(pv ? Q0.cs[j] : Q0.c1[j]);
      const double fu_e = eg ? 0.0 : u_e * (double)(pue ? Q0.c1[j] : Q0.ce[j]);
      const double fv_n = ng ? 0.0 : v_n * (double)(pvn ? Q0.c1[j] : Q0.cn[j]);
      const double fw = below_top ? ww * (double)(pw ? Q0.c1[j] : Q0.cup[j]) : 0.0;
      const double fqc = (RARE && src0 >= 0) ? col_sources(a, src0, k, cc, c1) : 0.0;
      // EFDC/caltran.f90:194-196,243
      const double cd = c1 * h1pk + DDELT * ((fqc + fu_i - fu_e + fv_i - fv_n) * dxyip +
                                             (fwp[j] - fw));
      double con = cd * hpki;
      fwp[j] = fw;
      if (RARE && rec >= 0)
        con = col_open_bc(a, bc, rec, k, cc, c1, con, u_i, u_e, v_i, v_n, fu_i, fu_e, fv_i,
                          fv_n, dxyip, hpki);
      if (!a.vd) {
        stb((T)con, a.nw[cc] + o, bi);
        continue;
      }
      double y;
      if (KC == 1) y = con;
      else if (K == 1) y = con * eeb;
      else y = (con - cclb * yprev[j]) * eeb;
      const T yt = (T)y;
      ys[(j * KC + k) * 64] = yt;
      yprev[j] = (double)yt;
    }
    ab_prev = abk;
    vcu_prev = vcu_k;
    Q0 = Q1;
    F0 = F1;
    if (FPF == 2) F1 = F2;
  }
  if (!a.vd) return;
  // back substitution and dye decay, write CON
  double up[NCW];
  int dm[NCW];
  double f[NCW];
#pragma unroll
  for (int j = 0; j < NCW; ++j) {
    if (!on[j]) continue;
    dm[j] = a.dye_mode[c0 + j];
    f[j] = a.dye_f[c0 + j];
    up[j] = (double)ys[(j * KC + KC - 1) * 64];
    const double v = up[j];
    stb((T)(dm[j] == 1 ? f[j] * v : (dm[j] == 2 ? v + f[j] : v)),
        a.nw[c0 + j] + (size_t)(KC - 1) * pitch, bi);
  }
  for (int kk = KC - 2; kk >= 0; --kk) {
    const double vk = vcu[kk * 64];
#pragma unroll
    for (int j = 0; j < NCW; ++j) {
      if (!on[j]) continue;
      const double v = (double)ys[(j * KC + kk) * 64] - vk * up[j];
      stb((T)(dm[j] == 1 ? f[j] * v : (dm[j] == 2 ? v + f[j] : v)),
          a.nw[c0 + j] + (size_t)kk * pitch, bi);
      up[j] = v;
    }
  }
}

template <typename T, int NCW, int FPF>
__device__ __forceinline__ void tran_colm_body(const TranArgs<T> &a, double *vcus) {
  // tiles of 64 columns start at L-1 = 0 mod 64, so every plane access of
  // a wave covers whole 256-byte segments (the dummy cell 1 is masked)
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int i = tile * 64 + (int)threadIdx.x;
  if (i < 1 || i > a.nwet) return;
  const int c0 = __builtin_amdgcn_readfirstlane((int)threadIdx.y) * NCW;
  if (col_keep(a, i, c0, 1, c0 + NCW < a.nc ? c0 + NCW : a.nc)) return;
  const int src0 = a.srcrec[i];
  const int rec = a.bcrec ? a.bcrec[i] : -1;
  if (__ballot(src0 >= 0 || rec >= 0) != 0)
    tran_colm_run<T, NCW, true, FPF>(a, vcus, i, c0, src0, rec);
  else
    tran_colm_run<T, NCW, false, FPF>(a, vcus, i, c0, src0, rec);
}

template <typename T, int NCW>
__global__ __launch_bounds__(256) void k_tran_colm(TranArgs<T> a) {
  extern __shared__ double vcus[];
  tran_colm_body<T, NCW, 1>(a, vcus);
}

// at most 168 VGPRs (three waves per SIMD), transport fields FPF layers ahead
template <typename T, int NCW, int FPF>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_tran_colm_w3(
    TranArgs<T> a) {
  extern __shared__ double vcus[];
  tran_colm_body<T, NCW, FPF>(a, vcus);
}

// Cohesive settling after the transport (SSEDTOX -> CALSED with DTSED = DT,
// EFDC/calconc.f90:621-630, WSETA = WSEDO): the implicit upwind sweep from
// the top layer down, EFDC/calsed.f90:232-256, then the bottom layer with
// the given deposition probability and resuspension flux,
// EFDC/calsed.f90:308-312.  One thread per column; each layer's new value
// is stored in the storage type and the flux to the layer below is taken
// from the stored value, as the reference reads SED back.
template <typename T>
__global__ __launch_bounds__(EFDC_BLOCK) void k_settle(T *__restrict__ sed,
                                                       const T *__restrict__ hpki,
                                                       int64_t pitch, int KC, int nwet,
                                                       double dtsed, double ws, double wsetmp,
                                                       double wese, const int *lmask) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;  // 256-aligned tiles
  if (i < 1 || i > nwet) return;
  if (lmask && lmask[i] == 0) return;  // CALSED runs over LWET (calsed.f90:232-256)
  double sedf = 0.0;  // SEDF(L,KC) = 0
  if (KC > 1) {
    {
      const int64_t j = i + (int64_t)(KC - 1) * pitch;
      const double WVEL = dtsed * (double)hpki[j];
      const double CLEFT = 1. + ws * WVEL;
      const double CRIGHT = dmx((double)sed[j], 0.);
      const T v = (T)(CRIGHT / CLEFT);
      sed[j] = v;
      sedf = -ws * (double)v;
    }
    for (int K = KC - 1; K >= 2; --K) {
      const int64_t j = i + (int64_t)(K - 1) * pitch;
      const double WVEL = dtsed * (double)hpki[j];
      const double CLEFT = 1. + ws * WVEL;
      const double CRIGHT = dmx((double)sed[j], 0.) - sedf * WVEL;
      const T v = (T)(CRIGHT / CLEFT);
      sed[j] = v;
      sedf = -ws * (double)v;
    }
  }
  const double WVEL = dtsed * (double)hpki[i];
  const double CLEFT = 1. + wsetmp * WVEL;
  const double CRIGHT = dmx((double)sed[i], 0.) + (wese - sedf) * WVEL;
  sed[i] = (T)(CRIGHT / CLEFT);
}

// CON1 at open-boundary columns takes the boundary value (ISUD=1,
// EFDC/caltran.f90:356,393,430,468).
template <typename T>
__global__ void k_tran_bc_con1(TranArgs<T> a, const int *bccell, int nrec) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int tot = nrec * a.KC;
  if (t >= tot) return;
  const int rec = t / a.KC, k = t % a.KC;
  const int i = bccell[rec];
  if (a.lmask && a.lmask[i] == 0) return;  // skipped, EFDC/caltran.f90:324-441
  for (int c = 0; c < a.nc; ++c)
    ((T *)a.old[c])[i + (int64_t)k * a.pitch] =
        (T)a.bc_con1[((size_t)rec * a.KC + k) * a.nc + c];
}

// ------------------------------------------------------------ balance
template <typename T>
__global__ __launch_bounds__(EFDC_BLOCK) void k_balance(
    const double *dxyp, const double *hp, const T *con, const T *dzc,
    int64_t pitch, int KC, int nwet, double *part, unsigned *ticket,
    double *out) {
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int i = 1 + tile * EFDC_BLOCK + (int)threadIdx.x;
  double v[1] = {0.0};
  if (i <= nwet) {
    const double base = dxyp[i] * hp[i];
    if (con == nullptr) {
      v[0] = base;
    } else {
      double m = 0.0;
      for (int k = 0; k < KC; ++k)
        m += base * (double)con[i + k * pitch] * (double)dzc[i + k * pitch];
      v[0] = m;
    }
  }
  double s[1];
  if (grid_sum<1>(v, part, ticket, gridDim.x, tile, s) && threadIdx.x == 0)
    *out = s[0];
}

// ------------------------------------------------ internal-mode stand-in
__global__ void k_closure_pre(double *tbx, double *tby, const double *uhe,
                              const double *vhe, const double *hu,
                              const double *hv, const double *nman, double G,
                              int nwet) {
  // workgroups start at L-1 = 0 mod 256 (whole 256-byte segments); the
  // dummy cell 1 is masked
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 1 || i > nwet) return;
  const double n2 = nman[i] * nman[i];
  const double u = uhe[i] / hu[i], v = vhe[i] / hv[i];
  tbx[i] = G * n2 * u * fabs(u) / cbrt(hu[i]);
  tby[i] = G * n2 * v * fabs(v) / cbrt(hv[i]);
}

template <typename T>
__global__ void k_closure_post(const double *__restrict__ uhdye, const double *__restrict__ uhdy1e,
                               const double *__restrict__ vhdxe, const double *__restrict__ vhdx1e,
                               const double *__restrict__ hp, const double *__restrict__ hu,
                               const double *__restrict__ hv, const double *__restrict__ dyiu,
                               const double *__restrict__ dxiv, const T *__restrict__ dzc,
                               T *__restrict__ uhdy2, T *__restrict__ vhdx2, T *__restrict__ u2,
                               T *__restrict__ v2, T *__restrict__ hpk, T *__restrict__ hpki,
                               int64_t pitch, int KC, int nwet) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;  // as k_closure_pre
  if (i < 1 || i > nwet) return;
  const double ub = 0.5 * (uhdye[i] + uhdy1e[i]);
  const double vb = 0.5 * (vhdxe[i] + vhdx1e[i]);
  // the same for every layer of the column
  const T uk = (T)(ub * dyiu[i] / hu[i]);
  const T vk = (T)(vb * dxiv[i] / hv[i]);
  const double hpi = hp[i];
  for (int k = 0; k < KC; ++k) {
    const int64_t j = i + k * pitch;
    const double dz = (double)dzc[j];
    uhdy2[j] = (T)(ub * dz);
    vhdx2[j] = (T)(vb * dz);
    u2[j] = uk;
    v2[j] = vk;
    const double hk = hpi * dz;
    hpk[j] = (T)hk;
    hpki[j] = (T)(1. / hk);
  }
}


// ---------------------------------------------------------------------
// Anti-diffusion and flux-corrected transport, EFDC/caltran.f90:473-786.
// The fields of one constituent go through three (L,K)-parallel passes
// because every pass needs neighbour values of the previous one:
//   k_adac_grad : POS = max(CON,0); UUUU, VVVV, WWWW (:490-511)
//   k_adac_flux : antidiffusive fluxes (:522-577) with the zeroing at flow
//                 sources (:583-599) and open boundaries (:602-620)
//   k_fct_beta  : CMAX/CMIN (:627-679), DUU/DVV and the beta limiters
//                 (:705-734)
// and k_col_finish applies the (limited) fluxes per column (:737-786),
// then the vertical diffusion and decay of CALCONC.
// Ghost cells (L=1, LC) are zero in every field, as the never-written
// entries of the reference's zero-initialised arrays are.
// ---------------------------------------------------------------------
template <typename T>
struct AdacArgs {
  const T *con, *con1;  // CON after the open boundaries, CON1 before them
  const T *uhdy2, *vhdx2, *w2, *u2, *v2, *dzig, *qsum, *sub3d, *svb3d, *hpk, *hpki;
  const double *dxiu, *dyiv, *hpi, *dxyip;
  const int *lw, *le, *ls, *ln, *lnw, *lse;
  const int *bcrec;
  const int4 *bcinfo;
  const unsigned char *lobflag;  // bit 0: LOBCS, bit 1: LOBCS2
  int nwet, lc0, KC, isadac, isfct;
  int64_t pitch;
  double dt;
  T *uuuu, *vvvv, *wwww, *fu, *fv, *fw, *duu, *dvv;
  // drying: SUB/SVB scale SUB3D/SVB3D (which hold SUB3DO/SVB3DO,
  // EFDC/calpuv2c.f90:982-990), LMASKDRY leaves columns out; null without
  const double *sub, *svb;
  const int *lmask;
};

template <typename T>
__device__ __forceinline__ double posv(const AdacArgs<T> &a, int j, int64_t o) {
  return (j == 0 || j == a.lc0) ? 0.0 : dmx((double)a.con[j + o], 0.);
}

template <typename T>
__global__ __launch_bounds__(EFDC_BLOCK) void k_adac_grad(AdacArgs<T> a) {
  const int i = 1 + xcd_tile(blockIdx.x, gridDim.x) * EFDC_BLOCK + threadIdx.x;
  if (i > a.nwet) return;
  const int k = blockIdx.y;
  const int64_t o = (int64_t)k * a.pitch, p = a.pitch;
  const double pc = posv(a, i, o);
  a.uuuu[i + o] = (T)((double)a.u2[i + o] * (pc - posv(a, a.lw[i], o)) * a.dxiu[i]);
  a.vvvv[i + o] = (T)((double)a.v2[i + o] * (pc - posv(a, a.ls[i], o)) * a.dyiv[i]);
  if (k < a.KC - 1) {  // WWWW(L,K), K = k+1 <= KS, plane K
    const int64_t oK = (int64_t)(k + 1) * p;
    a.wwww[i + oK] = (T)((double)a.w2[i + oK] * (posv(a, i, o + p) - pc) * a.hpi[i] *
                         (double)a.dzig[i + oK]);
  }
}

template <typename T>
__device__ __forceinline__ bool bc_side(const AdacArgs<T> &a, int j, int side) {
  if (!a.bcrec || j == 0 || j == a.lc0) return false;
  const int r = a.bcrec[j];
  if (r < 0) return false;
  const int4 q = a.bcinfo[r];
  return (side == 0 ? q.x : side == 1 ? q.y : side == 2 ? q.z : q.w) >= 0;
}

template <typename T>
__device__ __forceinline__ bool src_at(const AdacArgs<T> &a, int j, int64_t o) {
  return j != 0 && j != a.lc0 && fabs((double)a.qsum[j + o]) > 1.E-8;
}

template <typename T>
__global__ __launch_bounds__(EFDC_BLOCK) void k_adac_flux(AdacArgs<T> a) {
  const int i = 1 + xcd_tile(blockIdx.x, gridDim.x) * EFDC_BLOCK + threadIdx.x;
  if (i > a.nwet) return;
  const int k = blockIdx.y, K = k + 1;
  const int64_t o = (int64_t)k * a.pitch, p = a.pitch, oK = (int64_t)K * p;
  const double DDELTA = a.dt, BSMALL = 1.0E-6;
  const int LE = a.le[i], LW = a.lw[i], LN = a.ln[i], LS = a.ls[i];
  const int LNW = a.lnw[i], LSE = a.lse[i];
  const T *U = a.uuuu, *V = a.vvvv, *Wz = a.wwww;
  auto u = [&](int j, int64_t off) { return (double)U[j + off]; };
  auto v = [&](int j, int64_t off) { return (double)V[j + off]; };
  auto w = [&](int j, int KK) { return (double)Wz[j + (int64_t)KK * p]; };
  const double pc = posv(a, i, o), pw = posv(a, LW, o), ps = posv(a, LS, o);
  // U components (:533-543)
  const double uh = (double)a.uhdy2[i + o];
  double UTERM = fabs(uh) * (pc - pw);
  if (uh >= 0.0)
    UTERM = UTERM - 0.5 * DDELTA * uh *
                        (v(LNW, o) + v(LW, o) + w(LW, K) + w(LW, K - 1) + u(i, o) + u(LW, o));
  else
    UTERM = UTERM - 0.5 * DDELTA * uh *
                        (v(LN, o) + v(i, o) + w(i, K) + w(i, K - 1) + u(i, o) + u(LE, o));
  const double UHU = UTERM / (pc + pw + BSMALL);
  double fu = dmx(UHU, 0.) * pw + dmn(UHU, 0.) * pc;
  // V components (:546-556)
  const double vh = (double)a.vhdx2[i + o];
  double VTERM = fabs(vh) * (pc - ps);
  if (vh >= 0.0)
    VTERM = VTERM - 0.5 * DDELTA * vh *
                        (u(LS, o) + u(LSE, o) + w(LS, K) + w(LS, K - 1) + v(LS, o) + v(i, o));
  else
    VTERM = VTERM - 0.5 * DDELTA * vh *
                        (u(i, o) + u(LE, o) + w(i, K) + w(i, K - 1) + v(LN, o) + v(i, o));
  const double VHV = VTERM / (pc + ps + BSMALL);
  double fv = dmx(VHV, 0.) * ps + dmn(VHV, 0.) * pc;
  // W component (:560-577), FWUU(L,K) for K <= KS
  double fw = 0.0;
  if (K < a.KC) {
    const double pu = posv(a, i, o + p);
    const double ww = (double)a.w2[i + oK];
    double WTERM = fabs(ww) * (pu - pc);
    if (ww >= 0.0)
      WTERM = WTERM - 0.5 * DDELTA * ww *
                          (u(i, o) + u(LE, o) + v(i, o) + v(LN, o) + w(i, K) + w(i, K - 1));
    else
      WTERM = WTERM - 0.5 * DDELTA * ww *
                          (u(i, o + p) + u(LE, o + p) + v(i, o + p) + v(LN, o + p) + w(i, K) +
                           w(i, K + 1));
    const double WW = WTERM / (pu + pc + BSMALL);
    fw = dmx(WW, 0.) * pc + dmn(WW, 0.) * pu;
  }
  // zero at flow sources (ISADAC=1) and at concentration open boundaries
  if (a.isadac == 1) {
    if (src_at(a, i, o) || src_at(a, LW, o)) fu = 0.;
    if (src_at(a, i, o) || src_at(a, LS, o)) fv = 0.;
    if (src_at(a, i, o)) fw = 0.;
  }
  if (bc_side(a, LS, 0) || bc_side(a, i, 3)) fv = 0.0;
  if (bc_side(a, LW, 1) || bc_side(a, i, 2)) fu = 0.0;
  a.fu[i + o] = (T)fu;
  a.fv[i + o] = (T)fv;
  if (K < a.KC) a.fw[i + oK] = (T)fw;
}

template <typename T>
__device__ __forceinline__ void cont_mx_mn(const AdacArgs<T> &a, int j, int64_t o,
                                           double &mx, double &mn) {
  if (j == 0 || j == a.lc0) {
    mx = 0.0;
    mn = 0.0;
    return;
  }
  const double pp = dmx((double)a.con[j + o], 0.);
  const double c2 = fabs(dmx((double)a.con1[j + o], 0.));  // |CON2|
  mx = dmx(pp, c2);
  mn = dmn(pp, c2);
}

template <typename T>
__global__ __launch_bounds__(EFDC_BLOCK) void k_fct_beta(AdacArgs<T> a) {
  const int i = 1 + xcd_tile(blockIdx.x, gridDim.x) * EFDC_BLOCK + threadIdx.x;
  if (i > a.nwet) return;
  const int k = blockIdx.y, K = k + 1, KC = a.KC;
  const int64_t o = (int64_t)k * a.pitch, p = a.pitch;
  const double DDELT = a.dt, BSMALL = 1.0E-6;
  const int LE = a.le[i], LW = a.lw[i], LN = a.ln[i], LS = a.ls[i];
  double mx0, mn0, mxa, mna, mxb, mnb;
  cont_mx_mn(a, i, o, mx0, mn0);
  double cmax, cmin;
  if (K == 1) {  // KSZ = 1
    cont_mx_mn(a, i, o + p, mxa, mna);
    cmax = dmx(mx0, mxa);
    cmin = dmn(mn0, mna);
  } else if (K == KC) {
    cont_mx_mn(a, i, o - p, mxa, mna);
    cmax = dmx(mxa, mx0);
    cmin = dmn(mna, mn0);
  } else {
    cont_mx_mn(a, i, o - p, mxa, mna);
    cont_mx_mn(a, i, o + p, mxb, mnb);
    cmax = dmx(mx0, dmx(mxa, mxb));
    cmin = dmn(mn0, dmn(mna, mnb));
  }
  double mxW, mnW, mxE, mnE, mxS, mnS, mxN, mnN;
  cont_mx_mn(a, LW, o, mxW, mnW);
  cont_mx_mn(a, LE, o, mxE, mnE);
  cont_mx_mn(a, LS, o, mxS, mnS);
  cont_mx_mn(a, LN, o, mxN, mnN);
  double sbi = (double)a.sub3d[i + o], sbe = (double)a.sub3d[LE + o];
  double svi = (double)a.svb3d[i + o], svn = (double)a.svb3d[LN + o];
  if (a.sub) {
    sbi = a.sub[i] * sbi;
    sbe = a.sub[LE] * sbe;
    svi = a.svb[i] * svi;
    svn = a.svb[LN] * svn;
  }
  const double CWMAX = sbi * mxW, CEMAX = sbe * mxE, CSMAX = svi * mxS, CNMAX = svn * mxN;
  double CMAXT = dmx(CNMAX, CEMAX);
  CMAXT = dmx(CMAXT, CSMAX);
  CMAXT = dmx(CMAXT, CWMAX);
  cmax = dmx(cmax, CMAXT);
  const double CWMIN = sbi * mnW + 1.E+6 * (1. - sbi);
  const double CEMIN = sbe * mnE + 1.E+6 * (1. - sbe);
  const double CSMIN = svi * mnS + 1.E+6 * (1. - svi);
  const double CNMIN = svn * mnN + 1.E+6 * (1. - svn);
  double CMINT = dmn(CNMIN, CEMIN);
  CMINT = dmn(CMINT, CSMIN);
  CMINT = dmn(CMINT, CWMIN);
  cmin = dmn(cmin, CMINT);
  // influx / outflux in concentration units (:705-714)
  const double fui = (double)a.fu[i + o], fue = (double)a.fu[LE + o];
  const double fvi = (double)a.fv[i + o], fvn = (double)a.fv[LN + o];
  const double fwb = (double)a.fw[i + (int64_t)(K - 1) * p], fwt = (double)a.fw[i + (int64_t)K * p];
  const double dxyip = a.dxyip[i], hpki = (double)a.hpki[i + o];
  double duu = DDELT * (dxyip * (dmx(fui, 0.) - dmn(fue, 0.) + dmx(fvi, 0.) - dmn(fvn, 0.)) +
                        (dmx(fwb, 0.) - dmn(fwt, 0.))) * hpki;
  double dvv = DDELT * (dxyip * (dmx(fue, 0.) - dmn(fui, 0.) + dmx(fvn, 0.) - dmn(fvi, 0.)) +
                        (dmx(fwt, 0.) - dmn(fwb, 0.))) * hpki;
  if (a.lobflag[i] & 1) {  // :717-723
    duu = 0.;
    dvv = 0.;
  }
  const double pc = dmx((double)a.con[i + o], 0.);
  if (duu > 0.) duu = (cmax - pc) / (duu + BSMALL);  // :729-732
  duu = dmn(duu, 1.);
  if (dvv > 0.) dvv = (pc - cmin) / (dvv + BSMALL);
  dvv = dmn(dvv, 1.);
  a.duu[i + o] = (T)duu;
  a.dvv[i + o] = (T)dvv;
}

// Limited (or raw) antidiffusive flux across the face whose downstream
// cell is `d` and upstream neighbour `up` (the W/S neighbour of d).
template <typename T>
__device__ __forceinline__ double lim_h(const AdacArgs<T> &a, const T *f, int d, int up,
                                        int64_t o) {
  const double fl = (double)f[d + o];
  if (!a.isfct) return fl;
  const double duu_d = (double)a.duu[d + o], dvv_d = (double)a.dvv[d + o];
  const double duu_u = (double)a.duu[up + o], dvv_u = (double)a.dvv[up + o];
  return dmn(dvv_u, duu_d) * dmx(fl, 0.) + dmn(duu_u, dvv_d) * dmn(fl, 0.);
}

template <typename T>
__device__ __forceinline__ double lim_w(const AdacArgs<T> &a, int i, int K) {
  // FWUU(L,K), K = 1..KS, limited (:747-752), zero at LOBCS/LOBCS2 (:755-766)
  if (K <= 0 || K >= a.KC) return 0.0;
  const int64_t p = a.pitch;
  const double fl = (double)a.fw[i + (int64_t)K * p];
  if (!a.isfct) return fl;
  if (a.lobflag[i] & 3) return 0.0;
  const int64_t o = (int64_t)(K - 1) * p;  // layer K
  const double duuK = (double)a.duu[i + o], dvvK = (double)a.dvv[i + o];
  const double duuK1 = (double)a.duu[i + o + p], dvvK1 = (double)a.dvv[i + o + p];
  return dmn(dvvK, duuK1) * dmx(fl, 0.) + dmn(duuK, dvvK1) * dmn(fl, 0.);
}

template <typename T>
struct FinishArgs {
  AdacArgs<T> ad;
  T *con;          // in: CON after the open boundaries; out: final CON
  const T *ab, *cdzkk, *cdzkmk;
  const double *cdzkk_k, *cdzkmk_k;
  int dz_uniform;
  const unsigned char *pbcflag;
  int adac;        // apply the antidiffusive correction
  int dye_mode;
  double dye_f;
};

template <typename T>
__global__ __launch_bounds__(TBLOCK) void k_col_finish(FinishArgs<T> f) {
  extern __shared__ unsigned char smem[];
  const AdacArgs<T> &a = f.ad;
  const int TB = blockDim.x;
  const int i = 1 + xcd_tile(blockIdx.x, gridDim.x) * TB + (int)threadIdx.x;
  if (i > a.nwet) return;
  if (a.lmask && a.lmask[i] == 0) return;  // CON = CON1 from the column kernel
  const int tid = threadIdx.x, KC = a.KC;
  const int64_t p = a.pitch;
  double *vcu = (double *)smem;
  T *ys = (T *)(vcu + (size_t)KC * TB);
  const double DDELT = a.dt, DELTD2 = a.dt;
  const double dxyip = a.dxyip[i], hpi = a.hpi[i];
  const int LE = a.le[i], LN = a.ln[i], LW = a.lw[i], LS = a.ls[i];
  const bool eg = (LE == a.lc0), ng = (LN == a.lc0);
  double vcu_prev = 0.0, yprev = 0.0;
  for (int k = 0; k < KC; ++k) {
    const int K = k + 1;
    const int64_t o = (int64_t)k * p;
    double con = (double)f.con[i + o];
    if (f.adac && !(f.pbcflag[i])) {
      // EFDC/caltran.f90:771-778 (LOBCS cells keep CON, :781-786)
      const double fui = lim_h(a, a.fu, i, LW, o);
      const double fue = eg ? 0.0 : lim_h(a, a.fu, LE, a.lw[LE], o);
      const double fvi = lim_h(a, a.fv, i, LS, o);
      const double fvn = ng ? 0.0 : lim_h(a, a.fv, LN, a.ls[LN], o);
      const double cd = con * (double)a.hpk[i + o] +
                        DDELT * ((fui - fue + fvi - fvn) * dxyip + (lim_w(a, i, K - 1) - lim_w(a, i, K)));
      con = cd * (double)a.hpki[i + o];
    }
    // vertical diffusion, forward sweep (EFDC/calconc.f90:264-422)
    double y = con;
    if (KC > 1) {
      const double ckk = f.dz_uniform ? f.cdzkk_k[k] : (double)f.cdzkk[i + o];
      const double ckm = f.dz_uniform ? f.cdzkmk_k[k] : (double)f.cdzkmk[i + o];
      double vcu_k = 0.0;
      if (K == 1) {
        const double CCUBTMP = (-DELTD2 * ckk) * hpi * (double)f.ab[i + o];
        const double eeb = 1.0 / (1.0 - CCUBTMP);
        vcu_k = CCUBTMP * eeb;
        y = con * eeb;
      } else if (K < KC) {
        const double cclb = (-DELTD2 * ckm) * hpi * (double)f.ab[i + o - p];
        const double CCUBTMP = (-DELTD2 * ckk) * hpi * (double)f.ab[i + o];
        const double CCMBTMP = 1.0 - cclb - CCUBTMP;
        const double eeb = 1.0 / (CCMBTMP - cclb * vcu_prev);
        vcu_k = CCUBTMP * eeb;
        y = (con - cclb * yprev) * eeb;
      } else {
        const double cclb = (-DELTD2 * ckm) * hpi * (double)f.ab[i + o - p];
        const double eeb = 1.0 / ((1.0 - cclb) - cclb * vcu_prev);
        y = (con - cclb * yprev) * eeb;
      }
      vcu[(size_t)k * TB + tid] = vcu_k;
      vcu_prev = vcu_k;
    }
    ys[(size_t)k * TB + tid] = (T)y;
    yprev = (double)ys[(size_t)k * TB + tid];
  }
  auto fin = [&](double v) {
    return f.dye_mode == 1 ? f.dye_f * v : (f.dye_mode == 2 ? v + f.dye_f : v);
  };
  double up = (double)ys[(size_t)(KC - 1) * TB + tid];
  f.con[i + (int64_t)(KC - 1) * p] = (T)fin(up);
  for (int k = KC - 2; k >= 0; --k) {
    const double v = (double)ys[(size_t)k * TB + tid] - vcu[(size_t)k * TB + tid] * up;
    f.con[i + (int64_t)k * p] = (T)fin(v);
    up = v;
  }
}

template <typename T>
T *kp(efdc_ctx *c, int id) {
  return (T *)c->dK[id - EFDC_HPK];
}


// The fields, maps and scratch planes of the anti-diffusion passes (the
// constituent's CON/CON1 and switches are set per constituent).
template <typename T>
AdacArgs<T> adac_args(efdc_ctx *c, const T *uhdy2, const T *vhdx2, const T *w2,
                      const T *hpki, const int *bcrec, double dt) {
  AdacArgs<T> a;
  std::memset(&a, 0, sizeof(a));
  a.uhdy2 = uhdy2;
  a.vhdx2 = vhdx2;
  a.w2 = w2;
  a.u2 = kp<T>(c, EFDC_U2);
  a.v2 = kp<T>(c, EFDC_V2);
  a.dzig = (const T *)c->dW[EFDC_DZIG - EFDC_W2];
  a.qsum = kp<T>(c, EFDC_QSUM);
  a.sub3d = kp<T>(c, EFDC_SUB3D);
  a.svb3d = kp<T>(c, EFDC_SVB3D);
  const bool dry = c->prm.ISDRY > 0;
  a.sub = dry ? c->dH[EFDC_SUB] : nullptr;
  a.svb = dry ? c->dH[EFDC_SVB] : nullptr;
  a.lmask = dry ? c->lmaskdry : nullptr;
  a.hpk = kp<T>(c, EFDC_HPK);
  a.hpki = hpki;
  a.dxiu = c->dH[EFDC_DXIU];
  a.dyiv = c->dH[EFDC_DYIV];
  a.hpi = c->dH[EFDC_HPI];
  a.dxyip = c->dH[EFDC_DXYIP];
  a.lw = c->dLW;
  a.le = c->dLE;
  a.ls = c->dLS;
  a.ln = c->dLN;
  a.lnw = c->dLNW;
  a.lse = c->dLSE;
  a.bcrec = bcrec;
  a.bcinfo = c->bcinfo;
  a.lobflag = c->lobflag;
  a.nwet = c->n_wet();
  a.lc0 = c->LC - 1;
  a.KC = c->KC;
  a.pitch = c->pitch;
  a.dt = dt;
  a.uuuu = (T *)c->adac[0];
  a.vvvv = (T *)c->adac[1];
  a.wwww = (T *)c->adac[2];
  a.fu = (T *)c->adac[3];
  a.fv = (T *)c->adac[4];
  a.fw = (T *)c->adac[5];
  a.duu = (T *)c->adac[6];
  a.dvv = (T *)c->adac[7];
  return a;
}

// Anti-diffusion/FCT for each constituent that asks for it, then the
// vertical diffusion and decay of every constituent.
template <typename T>
int adac_pipeline(efdc_ctx *c, TranArgs<T> &ta, const efdc_step_t *step) {
  const int KC = c->KC, nwet = c->n_wet();
  const size_t plane = (size_t)c->pitch * (KC + 1);
  if (!c->adac_alloc) {
    for (int j = 0; j < 8; ++j) {
      char *q = nullptr;
      HIP_TRY(c, hipMalloc((void **)&q, plane * sizeof(T)));
      HIP_TRY(c, hipMemsetAsync(q, 0, plane * sizeof(T), c->stream));
      c->adac[j] = q;
    }
    c->adac_alloc = true;
  }
  AdacArgs<T> a = adac_args<T>(c, ta.uhdy2, ta.vhdx2, ta.w2, ta.hpki, ta.bcrec, step->dt);
  const dim3 grid((nwet + EFDC_BLOCK - 1) / EFDC_BLOCK, KC);
  auto lds_for = [&](int tb) { return (sizeof(double) + sizeof(T)) * (size_t)KC * tb; };
  int tb = TBLOCK;
  while (tb > 32 && lds_for(tb) > 64 * 1024) tb /= 2;
  const int nbc = (nwet + tb - 1) / tb;
  for (int j = 0; j < c->nconst; ++j) {
    const efdc_constituent_t &d = c->con[j].desc;
    a.con = ta.nw[j];
    a.con1 = ta.old[j];
    a.isadac = d.isadac;
    a.isfct = d.isfct;
    if (d.isadac > 0) {
      if (adj_recording(c)) {
        const int s = adj_record_adac(c, j, a.con);
        if (s) return s;
      }
      TIMED(c, TK_ADAC_GRAD, k_adac_grad<T><<<grid, EFDC_BLOCK, 0, c->stream>>>(a));
      TIMED(c, TK_ADAC_FLUX, k_adac_flux<T><<<grid, EFDC_BLOCK, 0, c->stream>>>(a));
      if (d.isfct != 0)
        TIMED(c, TK_FCT, k_fct_beta<T><<<grid, EFDC_BLOCK, 0, c->stream>>>(a));
    }
    FinishArgs<T> f;
    f.ad = a;
    f.con = ta.nw[j];
    f.ab = ta.ab;
    f.cdzkk = ta.cdzkk;
    f.cdzkmk = ta.cdzkmk;
    f.cdzkk_k = ta.cdzkk_k;
    f.cdzkmk_k = ta.cdzkmk_k;
    f.dz_uniform = ta.dz_uniform;
    f.pbcflag = c->pbcflag;
    f.adac = d.isadac > 0 ? 1 : 0;
    f.dye_mode = ta.dye_mode[j];
    f.dye_f = ta.dye_f[j];
    TIMED(c, TK_COL_FINISH, k_col_finish<T><<<nbc, tb, lds_for(tb), c->stream>>>(f));
  }
  return EFDC_OK;
}

// CON1 of the open-boundary columns, then the CON/CON1 buffer swap.
template <typename T>
int finish_tran(efdc_ctx *c, TranArgs<T> &a, const efdc_step_t *step) {
  for (int j = 0; j < c->nconst; ++j) {
    const efdc_constituent_t &d = c->con[j].desc;
    if (!(d.wset > 0.0)) continue;
    if (adj_recording(c)) {
      const int s = adj_record_settle(c, j, a.nw[j]);
      if (s) return s;
    }
    TIMED(c, TK_SETTLE, k_settle<T><<<(a.nwet + 256) / 256, 256, 0, c->stream>>>(
        a.nw[j], a.hpki, a.pitch, a.KC, a.nwet, step->dt, d.wset, d.probdep * d.wset,
        d.wese, a.lmask));
  }
  if (c->nbcrec > 0) {
    const int tot = c->nbcrec * c->KC;
    TIMED(c, TK_TRAN_BC, k_tran_bc_con1<T><<<(tot + 255) / 256, 256, 0, c->stream>>>(a, c->bccell,
                                                                c->nbcrec));
  }
  for (int j = 0; j < c->nconst; ++j) c->con[j].cur ^= 1;
  return EFDC_OK;
}

template <typename T>
int tran_impl(efdc_ctx *c, const efdc_step_t *step) {
  const int nc = c->nconst, KC = c->KC;
  if (nc == 0) return EFDC_OK;
  if ((c->nqsij > 0 && c->src_nc != nc) || (c->nbcrec > 0 && c->cbc_nc != nc)) {
    c->err = "source / boundary tables were set before all constituents were added";
    return EFDC_ERR_STATE;
  }
  bool any_adac = false;
  for (int j = 0; j < nc; ++j) {
    if (c->con[j].desc.isadac > 0) any_adac = true;
    if (c->con[j].desc.isadac > 0 && c->con[j].desc.isfct != 0 && KC < 2) {
      c->err = "FCT needs KC >= 2 (EFDC/caltran.f90:637-640 reads layer 2)";
      return EFDC_ERR_ARG;
    }
  }
  TranArgs<T> a;
  std::memset(&a, 0, sizeof(a));
  a.uhdy2 = kp<T>(c, EFDC_UHDY2);
  a.vhdx2 = kp<T>(c, EFDC_VHDX2);
  a.w2 = (const T *)c->dW[EFDC_W2 - EFDC_W2];
  a.h1pk = kp<T>(c, EFDC_H1PK);
  a.hpki = kp<T>(c, EFDC_HPKI);
  a.ab = kp<T>(c, EFDC_AB);
  a.cdzkk = kp<T>(c, EFDC_CDZKK);
  a.cdzkmk = kp<T>(c, EFDC_CDZKMK);
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
  a.KC = KC;
  a.nc = nc;
  a.rowmajor = c->rowmajor ? 1 : 0;
  a.pitch = c->pitch;
  a.dt = step->dt;
  a.niter = step->niter;
  for (int j = 0; j < nc; ++j) {
    ConstDev &k = c->con[j];
    a.old[j] = (const T *)k.buf[k.cur];
    a.nw[j] = (T *)k.buf[k.cur ^ 1];
    a.mo[j] = k.desc.mo;
    a.dye_mode[j] = 0;
    a.dye_f[j] = 0.0;
    if (k.desc.is_dye) {
      const double rk = k.desc.rkdye;
      if (rk == 1000.0) {
        a.dye_mode[j] = 2;
        a.dye_f[j] = step->dt / 86400.;
      } else if (rk != 0.0) {
        a.dye_mode[j] = 1;
        a.dye_f[j] = rk < 0.0 ? std::exp(-rk * step->dt) : 1. / (1. + step->dt * rk);
      }
    }
  }
  a.srcrec = c->srcrec;
  a.srcnext = c->srcnext;
  a.qss = c->qss;
  a.qsercell = c->qsercell;
  a.cqs = c->cqs;
  a.csertq = c->csertq;
  a.bcrec = c->nbcrec > 0 ? c->bcrec : nullptr;
  a.bcinfo = c->bcinfo;
  for (int s = 0; s < 4; ++s) a.cbc[s] = c->cbc[s];
  a.wtci = c->wtci;
  a.bc_con1 = c->bc_con1;
  a.vd = any_adac ? 0 : 1;
  a.lmask = c->prm.ISDRY > 0 ? c->lmaskdry : nullptr;

  if (adj_recording(c)) {
    const int s = adj_record_c0(c);
    if (s) return s;
  }
  const bool off32 = (int64_t)c->pitch * (c->KC + 1) < ((int64_t)1 << 31);
  // Column kernels, 64 columns per workgroup.  With mutual neighbour maps:
  // k_tran_colm, NCW constituents per wave (the shape parameter q = NCW),
  // ceil(nc/NCW) waves; else k_tran_col, q waves each walking its
  // constituents one after the other.  Every shape gives the same bits.
  // requested shape: sign and count of constituents, variant in bits 4-5
  const int req = c->prm.tran_shape > 0 ? (c->prm.tran_shape & 15) : c->prm.tran_shape;
  const int req_var = c->prm.tran_shape > 0 ? (c->prm.tran_shape >> 4) : 0;
  if (req > 0 && !c->mutual) {
    c->err = "tran_shape > 0 needs mutual neighbour maps";
    return EFDC_ERR_ARG;
  }
  const bool colm = c->mutual && req >= 0;
  auto waves_of = [&](int q) { return colm ? (nc + q - 1) / q : q; };
  auto lds_of = [&](int q) {
    const size_t per_wave = colm ? (size_t)q : 1;
    return (sizeof(double) + sizeof(T) * per_wave * waves_of(q)) * KC * 64;
  };
  auto shape_ok = [&](int q) {
    return q >= 1 && waves_of(q) <= 4 && lds_of(q) <= 64 * 1024 &&
           (colm ? (q == 1 || q < 2 * nc) : q <= nc);
  };
  int qs[3], nq = 0;
  for (int q : {4, 2, 1})
    if (shape_ok(q)) qs[nq++] = q;
  if (KC <= 64 && off32 && nq > 0) {
    const int nbc = colm ? (a.nwet + 64) / 64 : (a.nwet + 63) / 64;
    auto launch = [&](int q, int w3) {
      const dim3 blk(64, waves_of(q));
      const size_t lds = lds_of(q);
      if (!colm) {
        if (w3) k_tran_col_w3<T><<<nbc, blk, lds, c->stream>>>(a);
        else k_tran_col<T><<<nbc, blk, lds, c->stream>>>(a);
      } else if (q == 4) {
        if (w3) k_tran_colm_w3<T, 4, 1><<<nbc, blk, lds, c->stream>>>(a);
        else k_tran_colm<T, 4><<<nbc, blk, lds, c->stream>>>(a);
      } else if (q == 2) {
        if (w3 == 2) k_tran_colm_w3<T, 2, 2><<<nbc, blk, lds, c->stream>>>(a);
        else if (w3) k_tran_colm_w3<T, 2, 1><<<nbc, blk, lds, c->stream>>>(a);
        else k_tran_colm<T, 2><<<nbc, blk, lds, c->stream>>>(a);
      } else {
        if (w3 == 2) k_tran_colm_w3<T, 1, 2><<<nbc, blk, lds, c->stream>>>(a);
        else if (w3) k_tran_colm_w3<T, 1, 1><<<nbc, blk, lds, c->stream>>>(a);
        else k_tran_colm<T, 1><<<nbc, blk, lds, c->stream>>>(a);
      }
    };
    if (req != 0 && !shape_ok(req < 0 ? -req : req)) {
      c->err = "tran_shape does not fit this grid (waves per workgroup <= 4, LDS <= 64 KiB)";
      return EFDC_ERR_ARG;
    }
    if (c->tune_pending && !any_adac && req == 0) {
      // every shape computes the same bits, and a repeated launch
      // rewrites the same outputs (CON from CON1, the recovery memory of
      // outflow boundary layers), so the shapes are timed on the step
      // itself; the chosen one runs last
      c->tune_pending = false;
      int best = -1;
      float best_ms = 1e30f;
      hipEvent_t e0 = timer_event(c), e1 = timer_event(c);
      for (int qi = 0; qi < nq; ++qi) {
        // variants: 0 the compiler's register budget, 1 at most 168 VGPRs,
        // 2 (k_tran_colm with 1 or 2 constituents per thread) at most 168
        // VGPRs with the transport fields two layers ahead
        const int nvar = (colm && qs[qi] <= 2) ? 3 : 2;
        for (int w3 = 0; w3 < nvar; ++w3) {
          float tot = 0.f;
          for (int r = 0; r < 4; ++r) {
            HIP_TRY(c, hipEventRecord(e0, c->stream));
            launch(qs[qi], w3);
            HIP_TRY(c, hipEventRecord(e1, c->stream));
            HIP_TRY(c, hipEventSynchronize(e1));
            float ms = 0.f;
            HIP_TRY(c, hipEventElapsedTime(&ms, e0, e1));
            if (r > 0) tot += ms;  // the first run is a warm-up
          }
          const float avg = tot / 3.f;
          int slot = -1;
          if (c->tune_n < 8) {
            std::snprintf(c->tune_name[c->tune_n], 32, "%s=%d%s",
                          colm ? "tran_colm ncw" : "tran_col lanes", qs[qi],
                          w3 == 2 ? " vgpr<=168 pf2" : (w3 ? " vgpr<=168" : ""));
            slot = c->tune_n;
            c->tune_ms[c->tune_n++] = avg;
          }
          if (avg < best_ms) {
            best_ms = avg;
            best = slot;
            c->tran_ncg = qs[qi];
            c->tran_w3 = w3;
          }
        }
      }
      if (best >= 0) c->tune_chosen[best] = 1;
      c->ev_pool.push_back(e0);
      c->ev_pool.push_back(e1);
    }
    // untuned default: 2 constituents per wave (column kernel with mutual
    // maps), else as many lanes as fit
    int q = req != 0 ? (req < 0 ? -req : req) : c->tran_ncg;
    if (q <= 0 || !shape_ok(q)) q = (colm && shape_ok(2)) ? 2 : qs[0];
    TIMED(c, TK_TRAN, launch(q, req != 0 ? req_var : (c->tran_ncg > 0 ? c->tran_w3 : 0)));
    if (any_adac) {
      int s = adac_pipeline<T>(c, a, step);
      if (s) return s;
    }
    return finish_tran(c, a, step);
  }
  if (req != 0) {
    c->err = "tran_shape names a column kernel, which does not fit this grid";
    return EFDC_ERR_ARG;
  }
  // LDS per block: VCU and the FWUU carry (FP64) plus the forward values
  // (storage type); 128 columns per block when that fits in 64 KiB
  auto lds_for = [&](int tb) {
    return a.vd ? sizeof(double) * (size_t)(KC + nc) * tb + sizeof(T) * (size_t)KC * nc * tb
                : sizeof(double) * (size_t)(KC + nc) * tb;
  };
  int tb = TBLOCK;
  while (tb > 32 && lds_for(tb) > 64 * 1024) tb /= 2;
  if (lds_for(tb) > 64 * 1024) {
    c->err = "KC x constituents too large for the LDS staging of one column block";
    return EFDC_ERR_ARG;
  }
  const int nb = (a.nwet + tb - 1) / tb;
  TIMED(c, TK_TRAN, k_tran_upwind_vdiff<T><<<nb, tb, lds_for(tb), c->stream>>>(a));
  if (any_adac) {
    int s = adac_pipeline<T>(c, a, step);
    if (s) return s;
  }
  return finish_tran(c, a, step);
}

// ---------------------------------------------------------------------
// Reverse of the anti-diffusion / FCT passes of one differentiated
// constituent (EFDC/caltran.f90:473-786) and of the column finish that
// applies them.  The forward fields (UUUU..DVV) are recomputed from the
// recorded CON after the upwind step and CON1 by the forward kernels; the
// passes below run in the reverse order of the forward ones:
//   k_adj_adac_col  : per column, the vertical diffusion and decay (the
//                     transposed Thomas solve) and the update
//                     CON = (CON*HPK + DT*(flux divergence))*HPKI, leaving
//                     lambda of that update per cell (lcd)
//   k_adj_adac_lim  : per face, lambda of the limited flux from the two
//                     cells it enters, through the limiter to the raw flux
//                     and to DUU/DVV of the cells whose factors it took
//   k_adj_fct       : DUU/DVV -> CMAX/CMIN (to the CON/CON1 value the
//                     MAX/MIN chain selected), the raw fluxes and HPKI
//   k_adj_adac_flux : raw fluxes -> POS, UUUU/VVVV/WWWW, UHDY2/VHDX2
//   k_adj_adac_grad : UUUU/VVVV/WWWW -> POS, U2/V2, HPI
// Every MAX/MIN and sign switch takes the branch the forward step took.
// A value read by several cells' formulas receives their terms by FP64
// atomic adds (the sums' order, and so their last bits, can vary from run
// to run); the upwind step's reverse then starts from lambda of CON after
// the upwind step.
// ---------------------------------------------------------------------
template <typename T>
struct AdjAdacK {
  AdacArgs<T> a;
  const T *conu, *con1, *ab, *cdzkk, *cdzkmk, *dzc;
  const double *cdzkk_k, *cdzkmk_k;
  int dz_uniform;
  const unsigned char *pbcflag;
  AdacAdj io;
};

__device__ __forceinline__ void aadd(double *p, double v) {
  if (v != 0.0) atomicAdd(p, v);
}

// lambda of POS(j) = max(CON(j), 0) (zero at the ghost cells)
template <typename T>
__device__ __forceinline__ void lpos(const AdjAdacK<T> &q, int j, int64_t o, double l) {
  const AdacArgs<T> &a = q.a;
  if (j == 0 || j == a.lc0) return;
  if ((double)q.conu[j + o] > 0.0) aadd(q.io.lconu + j + o, l);
}

// lambda of the limited flux across a horizontal face (d: the cell whose
// W/S face it is, up: its W/S neighbour): to the raw flux (returned) and
// to the DUU/DVV factors the limiter took
template <typename T>
__device__ __forceinline__ double lim_h_rev(const AdjAdacK<T> &q, const T *f, int d, int up,
                                           int64_t o, double lF) {
  const AdacArgs<T> &a = q.a;
  if (!a.isfct) return lF;
  const double fl = (double)f[d + o];
  const double duu_d = (double)a.duu[d + o], dvv_d = (double)a.dvv[d + o];
  const double duu_u = (double)a.duu[up + o], dvv_u = (double)a.dvv[up + o];
  const double m1 = dmn(dvv_u, duu_d), m2 = dmn(duu_u, dvv_d);
  const double l1 = lF * dmx(fl, 0.), l2 = lF * dmn(fl, 0.);
  if (dvv_u < duu_d) aadd(q.io.ldvv + up + o, l1); else aadd(q.io.lduu + d + o, l1);
  if (duu_u < dvv_d) aadd(q.io.lduu + up + o, l2); else aadd(q.io.ldvv + d + o, l2);
  return lF * ((fl > 0. ? m1 : 0.) + (fl < 0. ? m2 : 0.));
}

template <typename T>
__global__ __launch_bounds__(EFDC_BLOCK) void k_adj_adac_col(AdjAdacK<T> q) {
  const AdacArgs<T> &a = q.a;
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (i > a.nwet) return;
  const int KC = a.KC;
  const int64_t P = a.pitch;
  const bool adac = !q.pbcflag[i];
  const double DDELT = a.dt, DELTD2 = a.dt;
  const double D = a.dxyip[i], hpi = a.hpi[i];
  const int LE = a.le[i], LN = a.ln[i], LW = a.lw[i], LS = a.ls[i];
  const bool eg = (LE == a.lc0), ng = (LN == a.lc0);
  double *wv = q.io.wv, *wvcu = q.io.wvcu, *weeb = q.io.weeb, *ww = q.io.ww;
  auto cd_of = [&](int k, double &cd) {
    const int64_t o = (int64_t)k * P;
    const double con = (double)q.conu[i + o];
    if (!adac) {
      cd = con;
      return con;
    }
    const double fui = lim_h(a, a.fu, i, LW, o);
    const double fue = eg ? 0.0 : lim_h(a, a.fu, LE, a.lw[LE], o);
    const double fvi = lim_h(a, a.fv, i, LS, o);
    const double fvn = ng ? 0.0 : lim_h(a, a.fv, LN, a.ls[LN], o);
    cd = con * (double)a.hpk[i + o] +
         DDELT * ((fui - fue + fvi - fvn) * D + (lim_w(a, i, k) - lim_w(a, i, k + 1)));
    return cd * (double)a.hpki[i + o];
  };
  auto coef = [&](int k, double &l, double &u, double &dl, double &du) {
    const int64_t o = (int64_t)k * P;
    const double ckk = q.dz_uniform ? q.cdzkk_k[k] : (double)q.cdzkk[i + o];
    const double ckm = q.dz_uniform ? q.cdzkmk_k[k] : (double)q.cdzkmk[i + o];
    l = u = dl = du = 0.0;
    if (k > 0) {
      dl = (-DELTD2 * ckm) * (double)q.ab[i + o - P];
      l = (-DELTD2 * ckm) * hpi * (double)q.ab[i + o - P];
    }
    if (k < KC - 1) {
      du = (-DELTD2 * ckk) * (double)q.ab[i + o];
      u = (-DELTD2 * ckk) * hpi * (double)q.ab[i + o];
    }
  };
  // forward: the updated CON of every layer and the Thomas solve
  double vcu_prev = 0.0, y_prev = 0.0;
  for (int k = 0; k < KC; ++k) {
    const int64_t o = i + (int64_t)k * P;
    double cd;
    const double con = cd_of(k, cd);
    if (KC == 1) {
      wv[o] = con;
      continue;
    }
    double l, u, dl, du;
    coef(k, l, u, dl, du);
    double eeb;
    if (k == 0) eeb = 1.0 / (1.0 - u);
    else if (k < KC - 1) eeb = 1.0 / ((1.0 - l - u) - l * vcu_prev);
    else eeb = 1.0 / ((1.0 - l) - l * vcu_prev);
    const double y = (k == 0) ? con * eeb : (con - l * y_prev) * eeb;
    wv[o] = y;
    wvcu[o] = u * eeb;
    weeb[o] = eeb;
    vcu_prev = u * eeb;
    y_prev = (double)(T)y;
  }
  if (KC > 1) {
    double up = (double)(T)wv[i + (int64_t)(KC - 1) * P];
    wv[i + (int64_t)(KC - 1) * P] = up;
    for (int k = KC - 2; k >= 0; --k) {
      const int64_t o = i + (int64_t)k * P;
      const double v = (double)(T)wv[o] - wvcu[o] * up;
      wv[o] = v;
      up = v;
    }
  }
  // reverse: decay, U^T w = g, L^T z = w
  const double fdec = q.io.dye_mode == 1 ? q.io.dye_f : 1.0;
  if (KC > 1) {
    double wprev = 0.0;
    for (int k = 0; k < KC; ++k) {
      const int64_t o = i + (int64_t)k * P;
      const double g = fdec * q.io.lam[o];
      const double wk = (k == 0) ? g : g - wvcu[o - P] * wprev;
      ww[o] = wk;
      wprev = wk;
    }
  }
  double lam_hpi = 0.0, lam_hp = 0.0, z_next = 0.0, l_next = 0.0;
  for (int k = KC - 1; k >= 0; --k) {
    const int64_t o = i + (int64_t)k * P;
    double z, l = 0.0;
    if (KC == 1) {
      z = fdec * q.io.lam[o];
    } else {
      double u, dl, du;
      coef(k, l, u, dl, du);
      z = (ww[o] - l_next * z_next) * weeb[o];
      const double lam_d = -z * wv[o];
      if (k > 0) lam_hpi += (-z * wv[o - P] - lam_d) * dl;
      if (k < KC - 1) lam_hpi += (-z * wv[o + P] - lam_d) * du;
    }
    z_next = z;
    l_next = l;
    if (!adac) {
      q.io.lconu[o] = z;
      q.io.lcd[o] = 0.0;
      continue;
    }
    double cd;
    (void)cd_of(k, cd);
    const double hpk = (double)a.hpk[o], hpki = (double)a.hpki[o], dz = (double)q.dzc[o];
    const double lcd = z * hpki;
    q.io.lcd[o] = lcd;
    q.io.lconu[o] = lcd * hpk;
    // HPK = HP DZC, HPKI = 1 / (HP DZC)
    lam_hp += lcd * (double)q.conu[o] * dz + z * cd * (-hpki * hpki * dz);
  }
  lam_hp += -lam_hpi * hpi * hpi;
  q.io.ahp[i] += lam_hp;
}

// Faces owned by cell i at layer k: its W and S faces and the face above
// layer k (K = k+1 < KC).
template <typename T>
__global__ __launch_bounds__(EFDC_BLOCK) void k_adj_adac_lim(AdjAdacK<T> q) {
  const AdacArgs<T> &a = q.a;
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (i > a.nwet) return;
  const int k = blockIdx.y, KC = a.KC;
  const int64_t P = a.pitch, o = (int64_t)k * P;
  const double dt = a.dt;
  const double *lcd = q.io.lcd;
  auto lcdD = [&](int j) { return (j == 0 || j == a.lc0) ? 0.0 : lcd[j + o] * a.dxyip[j]; };
  {
    const int LW = a.lw[i];
    const double lF = dt * (lcdD(i) - lcdD(LW));
    q.io.lfu[i + o] = lim_h_rev(q, a.fu, i, LW, o, lF);
  }
  {
    const int LS = a.ls[i];
    const double lF = dt * (lcdD(i) - lcdD(LS));
    q.io.lfv[i + o] = lim_h_rev(q, a.fv, i, LS, o, lF);
  }
  const int K = k + 1;  // face between layers k and k+1 (FWUU plane K)
  if (K < KC) {
    const double lF = dt * (lcd[i + (int64_t)K * P] - lcd[i + o]);
    double lfl = lF;
    if (a.isfct) {
      if (a.lobflag[i] & 3) {
        lfl = 0.0;
      } else {
        const double fl = (double)a.fw[i + (int64_t)K * P];
        const int64_t oa = o, ob = o + P;  // layers K-1 and K (1-based K, K+1)
        const double duuK = (double)a.duu[i + oa], dvvK = (double)a.dvv[i + oa];
        const double duuK1 = (double)a.duu[i + ob], dvvK1 = (double)a.dvv[i + ob];
        const double m1 = dmn(dvvK, duuK1), m2 = dmn(duuK, dvvK1);
        const double l1 = lF * dmx(fl, 0.), l2 = lF * dmn(fl, 0.);
        if (dvvK < duuK1) aadd(q.io.ldvv + i + oa, l1); else aadd(q.io.lduu + i + ob, l1);
        if (duuK < dvvK1) aadd(q.io.lduu + i + oa, l2); else aadd(q.io.ldvv + i + ob, l2);
        lfl = lF * ((fl > 0. ? m1 : 0.) + (fl < 0. ? m2 : 0.));
      }
    }
    aadd(q.io.lfw + i + (int64_t)K * P, lfl);
  }
}

// One candidate of the CMAX/CMIN chains: its value and where its lambda
// goes (kind 1: CON after the upwind step, 2: CON1, 0: a constant).
struct Cand {
  double v, s;   // value, factor of the lambda
  int kind;
  int64_t at;
};

template <typename T>
__device__ __forceinline__ void cand_mx_mn(const AdjAdacK<T> &q, int j, int64_t o, Cand &mx,
                                           Cand &mn) {
  const AdacArgs<T> &a = q.a;
  if (j == 0 || j == a.lc0) {
    mx = Cand{0.0, 0.0, 0, 0};
    mn = mx;
    return;
  }
  const double cu = (double)q.conu[j + o], c1 = (double)q.con1[j + o];
  const Cand pp = cu > 0. ? Cand{cu, 1.0, 1, j + o} : Cand{0.0, 0.0, 0, 0};
  const Cand c2 = c1 > 0. ? Cand{c1, 1.0, 2, j + o} : Cand{0.0, 0.0, 0, 0};
  mx = pp.v > c2.v ? pp : c2;   // dmx(pp, c2)
  mn = pp.v < c2.v ? pp : c2;   // dmn(pp, c2)
}
__device__ __forceinline__ Cand cmx(const Cand &x, const Cand &y) { return x.v > y.v ? x : y; }
__device__ __forceinline__ Cand cmn(const Cand &x, const Cand &y) { return x.v < y.v ? x : y; }
__device__ __forceinline__ Cand cscale(const Cand &x, double s, double add) {
  return Cand{s * x.v + add, s * x.s, x.kind, x.at};
}

template <typename T>
__device__ __forceinline__ void lcand(const AdjAdacK<T> &q, const Cand &c, double l) {
  if (c.kind == 1) aadd(q.io.lconu + c.at, l * c.s);
  else if (c.kind == 2) aadd(q.io.lcon1 + c.at, l * c.s);
}

template <typename T>
__global__ __launch_bounds__(EFDC_BLOCK) void k_adj_fct(AdjAdacK<T> q) {
  const AdacArgs<T> &a = q.a;
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (i > a.nwet) return;
  const int k = blockIdx.y, K = k + 1, KC = a.KC;
  const int64_t P = a.pitch, o = (int64_t)k * P;
  if (a.lobflag[i] & 1) return;  // DUU = DVV = 0
  const double lduu = q.io.lduu[i + o], ldvv = q.io.ldvv[i + o];
  if (lduu == 0.0 && ldvv == 0.0) return;
  const double DDELT = a.dt, BSMALL = 1.0E-6;
  const int LE = a.le[i], LW = a.lw[i], LN = a.ln[i], LS = a.ls[i];
  // CMAX / CMIN with the candidates they took
  Cand mx0, mn0, mxa, mna, mxb, mnb, cmax, cmin;
  cand_mx_mn(q, i, o, mx0, mn0);
  if (K == 1) {
    cand_mx_mn(q, i, o + P, mxa, mna);
    cmax = cmx(mx0, mxa);
    cmin = cmn(mn0, mna);
  } else if (K == KC) {
    cand_mx_mn(q, i, o - P, mxa, mna);
    cmax = cmx(mxa, mx0);
    cmin = cmn(mna, mn0);
  } else {
    cand_mx_mn(q, i, o - P, mxa, mna);
    cand_mx_mn(q, i, o + P, mxb, mnb);
    cmax = cmx(mx0, cmx(mxa, mxb));
    cmin = cmn(mn0, cmn(mna, mnb));
  }
  Cand mxW, mnW, mxE, mnE, mxS, mnS, mxN, mnN;
  cand_mx_mn(q, LW, o, mxW, mnW);
  cand_mx_mn(q, LE, o, mxE, mnE);
  cand_mx_mn(q, LS, o, mxS, mnS);
  cand_mx_mn(q, LN, o, mxN, mnN);
  const double sbi = (double)a.sub3d[i + o], sbe = (double)a.sub3d[LE + o];
  const double svi = (double)a.svb3d[i + o], svn = (double)a.svb3d[LN + o];
  Cand CMAXT = cmx(cscale(mxN, svn, 0.0), cscale(mxE, sbe, 0.0));
  CMAXT = cmx(CMAXT, cscale(mxS, svi, 0.0));
  CMAXT = cmx(CMAXT, cscale(mxW, sbi, 0.0));
  cmax = cmx(cmax, CMAXT);
  Cand CMINT = cmn(cscale(mnN, svn, 1.E+6 * (1. - svn)), cscale(mnE, sbe, 1.E+6 * (1. - sbe)));
  CMINT = cmn(CMINT, cscale(mnS, svi, 1.E+6 * (1. - svi)));
  CMINT = cmn(CMINT, cscale(mnW, sbi, 1.E+6 * (1. - sbi)));
  cmin = cmn(cmin, CMINT);
  const double fui = (double)a.fu[i + o], fue = (double)a.fu[LE + o];
  const double fvi = (double)a.fv[i + o], fvn = (double)a.fv[LN + o];
  const double fwb = (double)a.fw[i + (int64_t)(K - 1) * P], fwt = (double)a.fw[i + (int64_t)K * P];
  const double dxyip = a.dxyip[i], hpki = (double)a.hpki[i + o];
  const double XU = DDELT * (dxyip * (dmx(fui, 0.) - dmn(fue, 0.) + dmx(fvi, 0.) - dmn(fvn, 0.)) +
                             (dmx(fwb, 0.) - dmn(fwt, 0.)));
  const double XV = DDELT * (dxyip * (dmx(fue, 0.) - dmn(fui, 0.) + dmx(fvn, 0.) - dmn(fvi, 0.)) +
                             (dmx(fwt, 0.) - dmn(fwb, 0.)));
  const double duu_raw = XU * hpki, dvv_raw = XV * hpki;
  const double cu = (double)q.conu[i + o], pc = dmx(cu, 0.);
  double l_pc = 0.0, l_cmax = 0.0, l_cmin = 0.0, l_duu = 0.0, l_dvv = 0.0;
  if (duu_raw > 0.) {
    const double den = duu_raw + BSMALL, r = (cmax.v - pc) / den;
    if (r < 1.) {
      l_cmax += lduu / den;
      l_pc -= lduu / den;
      l_duu -= lduu * r / den;
    }
  } else {
    l_duu += lduu;  // DUU <= 0 < 1
  }
  if (dvv_raw > 0.) {
    const double den = dvv_raw + BSMALL, r = (pc - cmin.v) / den;
    if (r < 1.) {
      l_pc += ldvv / den;
      l_cmin -= ldvv / den;
      l_dvv -= ldvv * r / den;
    }
  } else {
    l_dvv += ldvv;
  }
  lcand(q, cmax, l_cmax);
  lcand(q, cmin, l_cmin);
  if (cu > 0.) aadd(q.io.lconu + i + o, l_pc);
  // the raw fluxes and HPKI
  const double lxu = l_duu * hpki, lxv = l_dvv * hpki;
  const double gu = lxu * DDELT, gv = lxv * DDELT;
  aadd(q.io.lfu + i + o, gu * dxyip * (fui > 0.) - gv * dxyip * (fui < 0.));
  aadd(q.io.lfu + LE + o, -gu * dxyip * (fue < 0.) + gv * dxyip * (fue > 0.));
  aadd(q.io.lfv + i + o, gu * dxyip * (fvi > 0.) - gv * dxyip * (fvi < 0.));
  aadd(q.io.lfv + LN + o, -gu * dxyip * (fvn < 0.) + gv * dxyip * (fvn > 0.));
  aadd(q.io.lfw + i + (int64_t)(K - 1) * P, gu * (fwb > 0.) - gv * (fwb < 0.));
  aadd(q.io.lfw + i + (int64_t)K * P, -gu * (fwt < 0.) + gv * (fwt > 0.));
  const double lhpki = l_duu * XU + l_dvv * XV;
  aadd(q.io.ahp + i, lhpki * (-hpki * hpki * (double)q.dzc[i + o]));
}

template <typename T>
__global__ __launch_bounds__(EFDC_BLOCK) void k_adj_adac_flux(AdjAdacK<T> q) {
  const AdacArgs<T> &a = q.a;
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (i > a.nwet) return;
  const int k = blockIdx.y, K = k + 1;
  const int64_t o = (int64_t)k * a.pitch, p = a.pitch, oK = (int64_t)K * p;
  const double DDELTA = a.dt, BSMALL = 1.0E-6;
  const int LE = a.le[i], LW = a.lw[i], LN = a.ln[i], LS = a.ls[i];
  const int LNW = a.lnw[i], LSE = a.lse[i];
  const T *U = a.uuuu, *V = a.vvvv, *Wz = a.wwww;
  auto u = [&](int j, int64_t off) { return (double)U[j + off]; };
  auto v = [&](int j, int64_t off) { return (double)V[j + off]; };
  auto w = [&](int j, int KK) { return (double)Wz[j + (int64_t)KK * p]; };
  auto lu = [&](int j, int64_t off, double l) { aadd(q.io.lu + j + off, l); };
  auto lv = [&](int j, int64_t off, double l) { aadd(q.io.lv + j + off, l); };
  auto lw = [&](int j, int KK, double l) { aadd(q.io.lww + j + (int64_t)KK * p, l); };
  const double pc = posv(a, i, o), pw = posv(a, LW, o), ps = posv(a, LS, o);
  // the zeroing of the forward flux (sources, open boundaries)
  bool zu = false, zv = false, zw = false;
  if (a.isadac == 1) {
    if (src_at(a, i, o) || src_at(a, LW, o)) zu = true;
    if (src_at(a, i, o) || src_at(a, LS, o)) zv = true;
    if (src_at(a, i, o)) zw = true;
  }
  if (bc_side(a, LS, 0) || bc_side(a, i, 3)) zv = true;
  if (bc_side(a, LW, 1) || bc_side(a, i, 2)) zu = true;
  const double dzc = (double)q.dzc[i + o];
  // U component
  const double lfu = zu ? 0.0 : q.io.lfu[i + o];
  if (lfu != 0.0) {
    const double uh = (double)a.uhdy2[i + o];
    double S;
    if (uh >= 0.0) S = v(LNW, o) + v(LW, o) + w(LW, K) + w(LW, K - 1) + u(i, o) + u(LW, o);
    else S = v(LN, o) + v(i, o) + w(i, K) + w(i, K - 1) + u(i, o) + u(LE, o);
    const double UTERM = fabs(uh) * (pc - pw) - 0.5 * DDELTA * uh * S;
    const double den = pc + pw + BSMALL, UHU = UTERM / den;
    const double lUHU = lfu * ((UHU > 0. ? pw : 0.) + (UHU < 0. ? pc : 0.));
    double l_pw = lfu * dmx(UHU, 0.), l_pc = lfu * dmn(UHU, 0.);
    const double lT = lUHU / den;
    l_pc += -lUHU * UHU / den + lT * fabs(uh);
    l_pw += -lUHU * UHU / den - lT * fabs(uh);
    const double luh = lT * ((uh >= 0.0 ? 1.0 : -1.0) * (pc - pw) - 0.5 * DDELTA * S);
    const double lS = -0.5 * DDELTA * uh * lT;
    if (uh >= 0.0) {
      lv(LNW, o, lS); lv(LW, o, lS); lw(LW, K, lS); lw(LW, K - 1, lS); lu(i, o, lS); lu(LW, o, lS);
    } else {
      lv(LN, o, lS); lv(i, o, lS); lw(i, K, lS); lw(i, K - 1, lS); lu(i, o, lS); lu(LE, o, lS);
    }
    lpos(q, i, o, l_pc);
    lpos(q, LW, o, l_pw);
    // UHDY2 = (UHDYE + UHDY1E) DZC / 2
    aadd(q.io.au_c + i, 0.5 * dzc * luh);
    aadd(q.io.au_p + i, 0.5 * dzc * luh);
  }
  // V component
  const double lfv = zv ? 0.0 : q.io.lfv[i + o];
  if (lfv != 0.0) {
    const double vh = (double)a.vhdx2[i + o];
    double S;
    if (vh >= 0.0) S = u(LS, o) + u(LSE, o) + w(LS, K) + w(LS, K - 1) + v(LS, o) + v(i, o);
    else S = u(i, o) + u(LE, o) + w(i, K) + w(i, K - 1) + v(LN, o) + v(i, o);
    const double VTERM = fabs(vh) * (pc - ps) - 0.5 * DDELTA * vh * S;
    const double den = pc + ps + BSMALL, VHV = VTERM / den;
    const double lVHV = lfv * ((VHV > 0. ? ps : 0.) + (VHV < 0. ? pc : 0.));
    double l_ps = lfv * dmx(VHV, 0.), l_pc = lfv * dmn(VHV, 0.);
    const double lT = lVHV / den;
    l_pc += -lVHV * VHV / den + lT * fabs(vh);
    l_ps += -lVHV * VHV / den - lT * fabs(vh);
    const double lvh = lT * ((vh >= 0.0 ? 1.0 : -1.0) * (pc - ps) - 0.5 * DDELTA * S);
    const double lS = -0.5 * DDELTA * vh * lT;
    if (vh >= 0.0) {
      lu(LS, o, lS); lu(LSE, o, lS); lw(LS, K, lS); lw(LS, K - 1, lS); lv(LS, o, lS); lv(i, o, lS);
    } else {
      lu(i, o, lS); lu(LE, o, lS); lw(i, K, lS); lw(i, K - 1, lS); lv(LN, o, lS); lv(i, o, lS);
    }
    lpos(q, i, o, l_pc);
    lpos(q, LS, o, l_ps);
    aadd(q.io.av_c + i, 0.5 * dzc * lvh);
    aadd(q.io.av_p + i, 0.5 * dzc * lvh);
  }
  // W component (FWUU plane K, K < KC); W2 is an input field here
  if (K < a.KC && !zw) {
    const double lfw = q.io.lfw[i + oK];
    if (lfw != 0.0) {
      const double pu = posv(a, i, o + p);
      const double ww = (double)a.w2[i + oK];
      double S;
      if (ww >= 0.0) S = u(i, o) + u(LE, o) + v(i, o) + v(LN, o) + w(i, K) + w(i, K - 1);
      else S = u(i, o + p) + u(LE, o + p) + v(i, o + p) + v(LN, o + p) + w(i, K) + w(i, K + 1);
      const double WTERM = fabs(ww) * (pu - pc) - 0.5 * DDELTA * ww * S;
      const double den = pu + pc + BSMALL, WW = WTERM / den;
      const double lWW = lfw * ((WW > 0. ? pc : 0.) + (WW < 0. ? pu : 0.));
      double l_pc = lfw * dmx(WW, 0.), l_pu = lfw * dmn(WW, 0.);
      const double lT = lWW / den;
      l_pu += -lWW * WW / den + lT * fabs(ww);
      l_pc += -lWW * WW / den - lT * fabs(ww);
      const double lS = -0.5 * DDELTA * ww * lT;
      if (ww >= 0.0) {
        lu(i, o, lS); lu(LE, o, lS); lv(i, o, lS); lv(LN, o, lS); lw(i, K, lS); lw(i, K - 1, lS);
      } else {
        lu(i, o + p, lS); lu(LE, o + p, lS); lv(i, o + p, lS); lv(LN, o + p, lS);
        lw(i, K, lS); lw(i, K + 1, lS);
      }
      lpos(q, i, o, l_pc);
      lpos(q, i, o + p, l_pu);
    }
  }
}

template <typename T>
__global__ __launch_bounds__(EFDC_BLOCK) void k_adj_adac_grad(AdjAdacK<T> q) {
  const AdacArgs<T> &a = q.a;
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (i > a.nwet) return;
  const int k = blockIdx.y;
  const int64_t o = (int64_t)k * a.pitch, p = a.pitch;
  const double pc = posv(a, i, o);
  const int LW = a.lw[i], LS = a.ls[i];
  double l_pc = 0.0;
  const double lU = q.io.lu[i + o], lV = q.io.lv[i + o];
  if (lU != 0.0) {
    const double u2 = (double)a.u2[i + o], pw = posv(a, LW, o);
    const double g = lU * u2 * a.dxiu[i];
    l_pc += g;
    lpos(q, LW, o, -g);
    // U2 = (UHDYE + UHDY1E)/2 * DYIU / HU, the same in every layer
    const double lu2 = lU * (pc - pw) * a.dxiu[i];
    const double hu = q.io.hu[i], ub = q.io.uhdye[i] * 0.5 + q.io.uhdy1e[i] * 0.5;
    aadd(q.io.au_c + i, 0.5 * lu2 * q.io.dyiu[i] / hu);
    aadd(q.io.au_p + i, 0.5 * lu2 * q.io.dyiu[i] / hu);
    aadd(q.io.ahu + i, -lu2 * ub * q.io.dyiu[i] / (hu * hu));
  }
  if (lV != 0.0) {
    const double v2 = (double)a.v2[i + o], ps = posv(a, LS, o);
    const double g = lV * v2 * a.dyiv[i];
    l_pc += g;
    lpos(q, LS, o, -g);
    const double lv2 = lV * (pc - ps) * a.dyiv[i];
    const double hv = q.io.hv[i], vb = q.io.vhdxe[i] * 0.5 + q.io.vhdx1e[i] * 0.5;
    aadd(q.io.av_c + i, 0.5 * lv2 * q.io.dxiv[i] / hv);
    aadd(q.io.av_p + i, 0.5 * lv2 * q.io.dxiv[i] / hv);
    aadd(q.io.ahv + i, -lv2 * vb * q.io.dxiv[i] / (hv * hv));
  }
  if (k < a.KC - 1) {  // WWWW plane K = k+1
    const int64_t oK = (int64_t)(k + 1) * p;
    const double lW = q.io.lww[i + oK];
    if (lW != 0.0) {
      const double w2 = (double)a.w2[i + oK], dzig = (double)a.dzig[i + oK];
      const double pu = posv(a, i, o + p), hpi = a.hpi[i];
      const double g = lW * w2 * hpi * dzig;
      lpos(q, i, o + p, g);
      l_pc -= g;
      // HPI = 1/HP
      aadd(q.io.ahp + i, -lW * w2 * (pu - pc) * dzig * hpi * hpi);
    }
  }
  lpos(q, i, o, l_pc);
}


template <typename T>
int adj_adac_impl(efdc_ctx *c, int slot, const AdacAdj &io, const efdc_step_t *step) {
  const efdc_constituent_t &d = c->con[slot].desc;
  const int KC = c->KC, nwet = c->n_wet();
  const int *bcrec = c->nbcrec > 0 ? c->bcrec : nullptr;
  AdjAdacK<T> q;
  std::memset(&q, 0, sizeof(q));
  q.a = adac_args<T>(c, kp<T>(c, EFDC_UHDY2), kp<T>(c, EFDC_VHDX2), (const T *)c->dW[0],
                     kp<T>(c, EFDC_HPKI), bcrec, step->dt);
  q.a.con = (const T *)io.conu;
  q.a.con1 = (const T *)io.con1;
  q.a.isadac = d.isadac;
  q.a.isfct = d.isfct;
  q.conu = (const T *)io.conu;
  q.con1 = (const T *)io.con1;
  q.ab = kp<T>(c, EFDC_AB);
  q.cdzkk = kp<T>(c, EFDC_CDZKK);
  q.cdzkmk = kp<T>(c, EFDC_CDZKMK);
  q.dzc = kp<T>(c, EFDC_DZC);
  q.cdzkk_k = c->cdzkk_k;
  q.cdzkmk_k = c->cdzkmk_k;
  q.dz_uniform = c->dz_uniform() ? 1 : 0;
  q.pbcflag = c->pbcflag;
  q.io = io;
  const dim3 grid((nwet + EFDC_BLOCK - 1) / EFDC_BLOCK, KC);
  const int nb = (nwet + EFDC_BLOCK - 1) / EFDC_BLOCK;
  // the forward fields from the recorded levels
  k_adac_grad<T><<<grid, EFDC_BLOCK, 0, c->stream>>>(q.a);
  k_adac_flux<T><<<grid, EFDC_BLOCK, 0, c->stream>>>(q.a);
  if (d.isfct != 0) k_fct_beta<T><<<grid, EFDC_BLOCK, 0, c->stream>>>(q.a);
  TIMED(c, TK_ADJ_TCOL, k_adj_adac_col<T><<<nb, EFDC_BLOCK, 0, c->stream>>>(q));
  TIMED(c, TK_ADJ_TCOL, k_adj_adac_lim<T><<<grid, EFDC_BLOCK, 0, c->stream>>>(q));
  if (d.isfct != 0) TIMED(c, TK_ADJ_TCOL, k_adj_fct<T><<<grid, EFDC_BLOCK, 0, c->stream>>>(q));
  TIMED(c, TK_ADJ_TCOL, k_adj_adac_flux<T><<<grid, EFDC_BLOCK, 0, c->stream>>>(q));
  TIMED(c, TK_ADJ_TCOL, k_adj_adac_grad<T><<<grid, EFDC_BLOCK, 0, c->stream>>>(q));
  HIP_TRY(c, hipGetLastError());
  return EFDC_OK;
}

}  // namespace

int tran_run(efdc_ctx *c, const efdc_step_t *step) {
  int s;
  if ((s = bal_before_tran(c, step))) return s;  // BAL2T3A, EFDC/calconc.f90:76-80
  s = c->store == EFDC_STORE_FP64 ? tran_impl<double>(c, step) : tran_impl<float>(c, step);
  if (s) return s;
  if ((s = bal_after_tran(c, step))) return s;   // BAL2T2, EFDC/calconc.f90:659-664
  HIP_TRY(c, hipGetLastError());
  return EFDC_OK;
}

int balance_run(efdc_ctx *c, efdc_balance_t *out) {
  const int nb = efdc_grid_blocks(c->n_wet());
  std::memset(out, 0, sizeof(*out));
  double *res = c->scal;
  for (int j = -1; j < c->nconst; ++j) {
    if (c->store == EFDC_STORE_FP64)
      TIMED(c, TK_BALANCE, k_balance<double><<<nb, EFDC_BLOCK, 0, c->stream>>>(
          c->dH[EFDC_DXYP], c->dH[EFDC_HP],
          j < 0 ? nullptr : (const double *)c->con[j].buf[c->con[j].cur],
          (const double *)c->dK[EFDC_DZC - EFDC_HPK], c->pitch, c->KC,
          c->n_wet(), c->part, c->tickets + TS_BALANCE * TICKET_SITE, res + (j + 1)));
    else
      TIMED(c, TK_BALANCE, k_balance<float><<<nb, EFDC_BLOCK, 0, c->stream>>>(
          c->dH[EFDC_DXYP], c->dH[EFDC_HP],
          j < 0 ? nullptr : (const float *)c->con[j].buf[c->con[j].cur],
          (const float *)c->dK[EFDC_DZC - EFDC_HPK], c->pitch, c->KC,
          c->n_wet(), c->part, c->tickets + TS_BALANCE * TICKET_SITE, res + (j + 1)));
  }
  std::vector<double> h(c->nconst + 1);
  HIP_TRY(c, hipMemcpyAsync(h.data(), res, sizeof(double) * (c->nconst + 1),
                            hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  out->volume = h[0];
  out->nconst = c->nconst;
  for (int j = 0; j < c->nconst && j < 16; ++j) out->mass[j] = h[j + 1];
  return EFDC_OK;
}

int closure_pre_run(efdc_ctx *c, const double *nman) {
  const int nwet = c->n_wet();
  TIMED(c, TK_CLOSURE_PRE, k_closure_pre<<<(nwet + 256) / 256, 256, 0, c->stream>>>(
      c->dH[EFDC_TBX], c->dH[EFDC_TBY], c->dH[EFDC_UHE], c->dH[EFDC_VHE],
      c->dH[EFDC_HU], c->dH[EFDC_HV], nman, c->prm.G, nwet));
  HIP_TRY(c, hipGetLastError());
  return EFDC_OK;
}

int closure_post_run(efdc_ctx *c) {
  const int nwet = c->n_wet();
  const int nb = (nwet + 256) / 256;
  if (c->store == EFDC_STORE_FP64)
    TIMED(c, TK_CLOSURE_POST, k_closure_post<double><<<nb, 256, 0, c->stream>>>(
        c->dH[EFDC_UHDYE], c->dH[EFDC_UHDY1E], c->dH[EFDC_VHDXE],
        c->dH[EFDC_VHDX1E], c->dH[EFDC_HP], c->dH[EFDC_HU], c->dH[EFDC_HV],
        c->dH[EFDC_DYIU], c->dH[EFDC_DXIV], kp<double>(c, EFDC_DZC),
        kp<double>(c, EFDC_UHDY2), kp<double>(c, EFDC_VHDX2),
        kp<double>(c, EFDC_U2), kp<double>(c, EFDC_V2), kp<double>(c, EFDC_HPK),
        kp<double>(c, EFDC_HPKI), c->pitch, c->KC, nwet));
  else
    TIMED(c, TK_CLOSURE_POST, k_closure_post<float><<<nb, 256, 0, c->stream>>>(
        c->dH[EFDC_UHDYE], c->dH[EFDC_UHDY1E], c->dH[EFDC_VHDXE],
        c->dH[EFDC_VHDX1E], c->dH[EFDC_HP], c->dH[EFDC_HU], c->dH[EFDC_HV],
        c->dH[EFDC_DYIU], c->dH[EFDC_DXIV], kp<float>(c, EFDC_DZC),
        kp<float>(c, EFDC_UHDY2), kp<float>(c, EFDC_VHDX2), kp<float>(c, EFDC_U2),
        kp<float>(c, EFDC_V2), kp<float>(c, EFDC_HPK), kp<float>(c, EFDC_HPKI),
        c->pitch, c->KC, nwet));
  HIP_TRY(c, hipGetLastError());
  return EFDC_OK;
}

int adj_adac_run(efdc_ctx *c, int slot, const AdacAdj &io, const efdc_step_t *step) {
  if (!c->adac_alloc) {
    c->err = "adjoint: anti-diffusion fields were never formed (no forward step with ISADAC)";
    return EFDC_ERR_STATE;
  }
  return c->store == EFDC_STORE_FP64 ? adj_adac_impl<double>(c, slot, io, step)
                                     : adj_adac_impl<float>(c, slot, io, step);
}
