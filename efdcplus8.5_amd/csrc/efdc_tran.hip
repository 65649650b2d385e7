// efdc_tran.hip -- CALCONC(ISTL=2, IS2TL=1) on gfx950.
//
// One thread per water column.  The kernel walks K = 1..KC once; at every
// layer it loads the shared transport fields (UHDY2, VHDX2, W2, H1PK,
// HPKI, AB) once for all constituents, forms the upwind update of
// CALTRAN (EFDC/caltran.f90:95-111,188-245) with the CALFQC sources
// (EFDC/calfqc.f90:209-225), applies the open-boundary sequence
// (EFDC/caltran.f90:316-471) on boundary columns, and runs the forward
// sweep of the implicit vertical diffusion (EFDC/calconc.f90:264-422) in
// FP64.  The forward values are staged in LDS; the back substitution
// (EFDC/calconc.f90:425-468) and the dye decay (EFDC/calconc.f90:581-616)
// then write the new concentrations once.  CON and CON1 are two buffers
// whose roles swap every step (CON1 = CON of EFDC/caltran.f90:59-65 costs
// nothing).  HBM traffic per column-step is
//   KC*s*(5 flows + 1 AB + 2*NC) + maps/metrics,
// the roofline of this kernel.
#include <cmath>
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
};

template <typename T>
__device__ __forceinline__ double ld(const T *p, int64_t j) {
  return (double)p[j];
}

template <typename T>
__global__ __launch_bounds__(TBLOCK) void k_tran_upwind_vdiff(TranArgs<T> a) {
  extern __shared__ unsigned char smem[];
  const int TB = blockDim.x;  // 128, or 64 when the LDS staging needs it
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int i = 1 + tile * TB + (int)threadIdx.x;
  if (i > a.nwet) return;
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
        fqc = fqc + fmax(qv, 0.) * a.cqs[j * nc + c] + fmin(qv, 0.) * conq +
              fmax(qs, 0.) * a.csertq[j * nc + c] + fmin(qs, 0.) * conq;
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
            con = fmax(CTMP, 0.);
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
              con = fmax(cbtmp, 0.);
            }
          }
          cv1 = con;
        }
        a.bc_con1[((size_t)rec * KC + k) * nc + c] = con;
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

// CON1 at open-boundary columns takes the boundary value (ISUD=1,
// EFDC/caltran.f90:356,393,430,468).
template <typename T>
__global__ void k_tran_bc_con1(TranArgs<T> a, const int *bccell, int nrec) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int tot = nrec * a.KC;
  if (t >= tot) return;
  const int rec = t / a.KC, k = t % a.KC;
  const int i = bccell[rec];
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
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (i > nwet) return;
  const double n2 = nman[i] * nman[i];
  const double u = uhe[i] / hu[i], v = vhe[i] / hv[i];
  tbx[i] = G * n2 * u * fabs(u) / cbrt(hu[i]);
  tby[i] = G * n2 * v * fabs(v) / cbrt(hv[i]);
}

template <typename T>
__global__ void k_closure_post(const double *uhdye, const double *uhdy1e,
                               const double *vhdxe, const double *vhdx1e,
                               const double *hp, const double *hu,
                               const double *hv, const double *dyiu,
                               const double *dxiv, const T *dzc, T *uhdy2,
                               T *vhdx2, T *u2, T *v2, T *hpk, T *hpki,
                               int64_t pitch, int KC, int nwet) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (i > nwet) return;
  const double ub = 0.5 * (uhdye[i] + uhdy1e[i]);
  const double vb = 0.5 * (vhdxe[i] + vhdx1e[i]);
  for (int k = 0; k < KC; ++k) {
    const int64_t j = i + k * pitch;
    const double dz = (double)dzc[j];
    uhdy2[j] = (T)(ub * dz);
    vhdx2[j] = (T)(vb * dz);
    u2[j] = (T)(ub * dyiu[i] / hu[i]);
    v2[j] = (T)(vb * dxiv[i] / hv[i]);
    const double hk = hp[i] * dz;
    hpk[j] = (T)hk;
    hpki[j] = (T)(1. / hk);
  }
}

template <typename T>
T *kp(efdc_ctx *c, int id) {
  return (T *)c->dK[id - EFDC_HPK];
}

template <typename T>
int tran_impl(efdc_ctx *c, const efdc_step_t *step) {
  const int nc = c->nconst, KC = c->KC;
  if (nc == 0) return EFDC_OK;
  if ((c->nqsij > 0 && c->src_nc != nc) || (c->nbcrec > 0 && c->cbc_nc != nc)) {
    c->err = "source / boundary tables were set before all constituents were added";
    return EFDC_ERR_STATE;
  }
  for (int j = 0; j < nc; ++j)
    if (c->con[j].desc.isadac > 0) {
      c->err = "anti-diffusion (ISADAC>0) is not available in this build";
      return EFDC_ERR_ARG;
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

  // LDS per block: VCU and the FWUU carry (FP64) plus the forward values
  // (storage type); 128 columns per block when that fits in 64 KiB
  auto lds_for = [&](int tb) {
    return sizeof(double) * (size_t)(KC + nc) * tb + sizeof(T) * (size_t)KC * nc * tb;
  };
  int tb = TBLOCK;
  while (tb > 32 && lds_for(tb) > 64 * 1024) tb /= 2;
  if (lds_for(tb) > 64 * 1024) {
    c->err = "KC x constituents too large for the LDS staging of one column block";
    return EFDC_ERR_ARG;
  }
  const int nb = (a.nwet + tb - 1) / tb;
  TIMED(c, TK_TRAN, k_tran_upwind_vdiff<T><<<nb, tb, lds_for(tb), c->stream>>>(a));
  if (c->nbcrec > 0) {
    const int tot = c->nbcrec * KC;
    TIMED(c, TK_TRAN_BC, k_tran_bc_con1<T><<<(tot + 255) / 256, 256, 0, c->stream>>>(a, c->bccell,
                                                                c->nbcrec));
  }
  for (int j = 0; j < nc; ++j) c->con[j].cur ^= 1;
  return EFDC_OK;
}

}  // namespace

int tran_run(efdc_ctx *c, const efdc_step_t *step) {
  int s = c->store == EFDC_STORE_FP64 ? tran_impl<double>(c, step)
                                      : tran_impl<float>(c, step);
  if (s) return s;
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
  TIMED(c, TK_CLOSURE_PRE, k_closure_pre<<<(nwet + 255) / 256, 256, 0, c->stream>>>(
      c->dH[EFDC_TBX], c->dH[EFDC_TBY], c->dH[EFDC_UHE], c->dH[EFDC_VHE],
      c->dH[EFDC_HU], c->dH[EFDC_HV], nman, c->prm.G, nwet));
  HIP_TRY(c, hipGetLastError());
  return EFDC_OK;
}

int closure_post_run(efdc_ctx *c) {
  const int nwet = c->n_wet();
  const int nb = (nwet + 255) / 256;
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
