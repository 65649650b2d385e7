// efdc_uvw.hip -- the internal mode, CALUVW(ISTL=2) on sigma layers, on gfx950.
//
// EFDC/caluvw.f90:176-1172 with the two steps HDMT2T takes around it
// (EFDC/hdmt2t.f90:513-527 before, :638-655 after).  Five kernels per step:
//   k_uvw_col  one thread per water column: both face columns (U face and
//              V face of the cell) are solved in registers -- the bottom
//              drag coefficient, the tridiagonal sweep of the internal
//              shears with the Sherman-Morrison step for the drag, the layer
//              flows, their correction to the external solution and the
//              velocities -- plus the external flow split per layer;
//   k_uvw_w    W from the layer continuity (it reads the east and north
//              neighbours' new layer flows), the two-level averages and the
//              column sums of the averaged layer flows;
//   k_uvw_obc_side, k_uvw_obc_avg  the layer flows of the open-boundary
//              faces, one launch per side in the reference's order, then the
//              averages and sums of the faces they changed;
//   k_uvw_hp   the depth from the layer flows, HP of open-boundary cells,
//              HPK and HPKI;
//   k_uvw_uv   (optional) U at V points and V at U points of the new level.
// The time-level advance of the internal variables is a swap of buffer
// pointers.  All internal arrays are FP64; the transport-side arrays
// (DZC, QSUM, SUB3D/SVB3D, the averages, HPK, HPKI) are stored in the
// context's storage type T.  Every kernel streams plane-major (LCM,KCM)
// arrays, so the roofline is HBM.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "efdc_internal.h"

namespace {

constexpr int NIH = EFDC_NUM_IH_END - EFDC_STBX;
constexpr int NIK = EFDC_NUM_IK_END - EFDC_UHDYF;
constexpr int NIW = EFDC_NUM_IW_END - EFDC_W;

// Face constants of an open face per layer (K-1), from DZC of a sigma grid:
// SGZU, DZGU, CDZFU, CDZUU, CDZLU, CDZRU, CDZDU, CDZMU (EFDC/aaefdc.f90:
// 2001-2126) and CERRU (EFDC/caluvw.f90:82-113); a closed face (SUBO = 0)
// has all of them 0 except CERRU = cerr_closed.
struct LayerConsts {
  double sg[64], dzg[64], cdzf[64], cdzu[64], cdzl[64], cdzr[64], cdzd[64], cdzm[64],
      cerr[64], cerr_closed[64];
};

// Open-boundary faces of the internal mode (caluvw.f90:700-767), device.
struct UvwObcDev {
  int n[4];            // S, W, E, N: concentration-BC cells
  int serial[4];       // 1: a cell of the side reads a face another writes
  int *cell[4];        // 0-based cell index
  int *ispb[4];        // ISPB of the pressure side at the same index (0 past it)
  int navg_u, navg_v, nwz;
  int *avg_u, *avg_v, *wz;  // cells whose U face / V face / W the sides change
};

}  // namespace

struct UvwState {
  double *h[NIH] = {};
  double *k[NIK] = {};
  double *w[NIW] = {};
  double *tvare = nullptr, *tvarn = nullptr;
  LayerConsts *lc = nullptr;   // device
  LayerConsts lc_host{};
  int lc_kc = -1;              // KC the table was built for
  UvwObcDev obc{};
  bool obc_ready = false;
  int *flag = nullptr;         // device: 1 = negative depth
  int *flag_host = nullptr;    // pinned
};

namespace {

template <typename T>
void dfree_dev(T *&p) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

int upload_dev(efdc_ctx *c, int **dst, const std::vector<int> &v) {
  dfree_dev(*dst);
  if (v.empty()) return EFDC_OK;
  int s;
  if ((s = dev_alloc(c, (void **)dst, v.size() * sizeof(int)))) return s;
  HIP_TRY(c, hipMemcpyAsync(*dst, v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice,
                            c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return EFDC_OK;
}

__global__ void k_uvw_flag_reset(int *f) { *f = 0; }

#define IH(id) (u->h[EFDC_##id - EFDC_STBX])
#define IK(id) (u->k[EFDC_##id - EFDC_UHDYF])
#define IW(id) (u->w[EFDC_##id - EFDC_W])

template <typename T>
struct UvwArgs {
  // (LCM) FP64 arrays of the external mode
  const double *uhdye, *vhdxe, *hu, *hv, *hui, *hvi, *h1u, *h1v, *sub, *svb, *subo, *svbo;
  const double *dyiu, *dxiv, *dxyip, *qsume, *h1p, *p, *belv, *tsx, *tsy;
  double *uhe, *vhe, *hp, *hpi, *uhdye_w, *vhdxe_w;
  // (LCM) FP64 arrays of the internal mode
  const double *stbx, *stby, *dyu, *dxv;
  double *vu, *uv, *v1u, *u1v;
  // (LCM,KCM) FP64, plane K-1 at (K-1)*pitch
  double *uhdyf, *vhdxf, *uhdy, *vhdx, *uhdyf1, *vhdxf1, *uhdy1, *vhdx1, *u, *v, *u1, *v1;
  double *uhdyf2, *vhdxf2, *uhdyek, *vhdxek, *uhdy1ek, *vhdx1ek;
  const double *du, *dv, *avui, *avvi;
  double *w, *w1;  // (LCM,0:KCM): plane K at K*pitch
  // transport side, storage type T
  const T *dzc, *qsum, *sub3d, *svb3d;
  T *uhdy2, *vhdx2, *u2, *v2, *w2, *hpk, *hpki;
  const int *lw, *le, *ls, *ln, *lnw, *lse;
  const int *lmask;                 // LMASKDRY (1 wet), nullptr with ISDRY = 0
  const unsigned char *lobflag;     // bit 0: LOBCS
  double *tvare, *tvarn;
  const LayerConsts *lc;
  int *flag;
  int64_t P;
  int KC, nwet, lc0, isdry, shear, wind;
  double dt, dti, avcon1, gi, hdry;
};

__device__ __forceinline__ int uvw_cell(int nwet) {
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int i = t * EFDC_BLOCK + (int)threadIdx.x;
  return i == 0 ? nwet + 1 : i;  // wet iff <= nwet
}

template <typename T>
__device__ __forceinline__ bool wet_at(const UvwArgs<T> &a, int i) {
  return a.lmask == nullptr || a.lmask[i] != 0;
}

// One face column (the U face of cell i when v = false, its V face when
// v = true): the shear solve of EFDC/caluvw.f90:415-524 (with the internal
// shears of CALEXP2T formed here when shear is set, EFDC/calexp2t.f90:
// 1332-1343), the layer flows (:527-567), their correction to the external
// solution (:571-611) and the velocities (:613-647).  Returns the new
// UHE (VHE).  Arrays hold only the first KS (shears) or KC (flows) entries.
template <typename T, int KM, bool EXACT>
__device__ __forceinline__ double face_column(const UvwArgs<T> &a, int i, bool vface,
                                              double rc) {
  // with EXACT the layer count is the template's, so every layer loop
  // below is straight-line code without predicates
  const int KC = EXACT ? KM : a.KC, KS = KC - 1;
  const int64_t P = a.P;
  const LayerConsts &C = *a.lc;
  const bool open = (vface ? a.svbo[i] : a.subo[i]) > 0.0;
  const double hh = vface ? a.hv[i] : a.hu[i];
  const double hhi = vface ? a.hvi[i] : a.hui[i];
  const double sb = vface ? a.svb[i] : a.sub[i];
  const double h1 = vface ? a.h1v[i] : a.h1u[i];
  const double ts = vface ? a.tsy[i] : a.tsx[i];
  const double width = vface ? a.dxv[i] : a.dyu[i];
  const double widthi = vface ? a.dxiv[i] : a.dyiu[i];
  const double ext = vface ? a.vhdxe[i] : a.uhdye[i];
  double uhe = vface ? a.vhe[i] : a.uhe[i];
  const double *din = vface ? a.dv : a.du;
  const double *avi = vface ? a.avvi : a.avui;
  const double *vel1 = vface ? a.v1 : a.u1;
  const T *s3d = vface ? a.svb3d : a.sub3d;
  double *flowf = vface ? a.vhdxf : a.uhdyf;
  double *flow = vface ? a.vhdx : a.uhdy;
  double *vel = vface ? a.v : a.u;
  const bool dscale = a.isdry > 0;
  const double dti = a.dti;
  // Loads are issued for every unrolled layer at a valid address (layers
  // past the column's read its last shear layer and are not used), so the
  // compiler can issue them ahead of the dependent FP64 chain.
  auto lay = [&](int k) -> int64_t { return i + (int64_t)(k < KS ? k : KS - 1) * P; };

  double du[KM], uu[KM], cu[KM], s3[KM];
  // forward sweep (:419-486)
#pragma unroll
  for (int k = 0; k < KM; ++k) {
    // EXACT: loads at a valid address on every layer (issued early); else
    // only on the column's layers (fewer registers live at once)
    double av = 0.0, s3k = 0.0, d = 0.0;
    if (EXACT || k < KS) {
      const int64_t j = lay(k);
      av = avi[j];
      s3k = dscale ? sb * (double)s3d[j] : (double)s3d[j];
      if (a.shear) {
        const double v0 = vel1[j], v1 = vel1[j + P];
        const double cf = open ? C.cdzf[k] : 0.0;
        d = cf * (h1 * (v1 - v0) * dti + 0.0);
        if (a.wind && k == KS - 1) d = d - (open ? C.cdzu[k] : 0.0) * ts;
      } else {
        d = din[j];
      }
    }
    if (k < KS) {
      s3[k] = s3k;
      d = s3k * d;
      const double cl = open ? C.cdzl[k] : 0.0;
      const double cm = (open ? C.cdzm[k] : 0.0) * dti;
      const double cuu = open ? C.cdzu[k] : 0.0;
      const double cmu = 1. + cm * hh * av;
      if (k > 0) {
        const double eu = 1. / (cmu - cl * cu[k - 1]);
        cu[k] = cuu * eu;
        du[k] = (d - cl * du[k - 1]) * eu;
        uu[k] = -cl * uu[k - 1] * eu;
      } else {
        const double eu = 1. / cmu;
        cu[k] = cuu * eu;
        du[k] = (d - cl * rc * uhe * hhi) * eu;
        uu[k] = eu;
      }
    }
  }
  // back substitution (:477-486)
#pragma unroll
  for (int k = KM - 2; k >= 0; --k) {
    if (k < KS - 1) {
      du[k] = du[k] - cu[k] * du[k + 1];
      uu[k] = uu[k] - cu[k] * uu[k + 1];
    }
  }
  // Sherman-Morrison step for the bottom drag (:489-524)
  double aa = 0., bb = 1.;
#pragma unroll
  for (int k = 0; k < KM; ++k) {
    const double av = (EXACT || k < KS) ? avi[lay(k)] : 0.0;
    if (k < KS) {
      const double cr = (open ? C.cdzr[k] : 0.0) * rc * av;
      aa = aa + cr * du[k];
      bb = bb + cr * uu[k];
    }
  }
  aa = aa / bb;
#pragma unroll
  for (int k = 0; k < KM; ++k) {
    const double av = (EXACT || k < KS) ? avi[lay(k)] : 0.0;
    if (k < KS) du[k] = s3[k] * (open ? C.dzg[k] : 0.0) * hh * av * (du[k] - aa * uu[k]);
  }
  // layer flows (:527-567), kept in uu from here on (KC entries)
#pragma unroll
  for (int k = 0; k < KM; ++k)
    if (k < KS) uhe = uhe + (open ? C.cdzd[k] : 0.0) * du[k];
#pragma unroll
  for (int k = KM - 1; k >= 0; --k) {
    if (k == KC - 1)
      uu[k] = uhe * sb;
    else if (k < KC - 1)
      uu[k] = s3[k] * (uu[k + 1] - du[k]);
  }
#pragma unroll
  for (int k = 0; k < KM; ++k)
    if (k < KC) uu[k] = uu[k] * width;
  // correction to the external solution (:571-611)
  double tv = 0.;
#pragma unroll
  for (int k = 0; k < KM; ++k)
    if (k < KC) tv = tv + uu[k] * (open ? C.sg[k] : 0.0);
  tv = tv - ext;
  double un = 0.;
#pragma unroll
  for (int k = 0; k < KM; ++k) {
    if (k < KC) {
      const int64_t j = i + (int64_t)k * P;
      const double sg = open ? C.sg[k] : 0.0;
      const double f = uu[k] - tv * (open ? C.cerr[k] : C.cerr_closed[k]);
      flowf[j] = f;
      flow[j] = f * sg;
      un = un + f * sg;
      // velocities (:613-647): two roundings as in the reference
      const double ve = f * hhi;
      vel[j] = ve * widthi;
    }
  }
  return un * widthi;
}

// Zero the column of a dry cell at both levels and its averages
// (caluvw.f90:186-263 for a newly dry cell; a cell that stays dry keeps the
// zeros, so every dry cell is zeroed); keep its UHDYEK as the reference's
// copy UHDY1EK = UHDYEK at the step start leaves it.
template <typename T>
__device__ void zero_dry_column(const UvwArgs<T> &a, int i) {
  const int KC = a.KC;
  const int64_t P = a.P;
  for (int k = 0; k < KC; ++k) {
    const int64_t j = i + k * P;
    a.uhdyf[j] = a.vhdxf[j] = a.uhdy[j] = a.vhdx[j] = a.u[j] = a.v[j] = 0.0;
    a.uhdyf1[j] = a.vhdxf1[j] = a.uhdy1[j] = a.vhdx1[j] = a.u1[j] = a.v1[j] = 0.0;
    a.uhdyf2[j] = a.vhdxf2[j] = 0.0;
    a.uhdy2[j] = a.vhdx2[j] = a.u2[j] = a.v2[j] = (T)0;
    a.w[j + P] = a.w1[j + P] = 0.0;
    a.w2[j + P] = (T)0;
    a.uhdyek[j] = a.uhdy1ek[j];
    a.vhdxek[j] = a.vhdx1ek[j];
  }
  a.tvare[i] = 0.0;
  a.tvarn[i] = 0.0;
}

template <typename T, int KM, bool EXACT>
__global__ __launch_bounds__(EFDC_BLOCK) void k_uvw_col(UvwArgs<T> a) {
  const int i = uvw_cell(a.nwet);
  if (i > a.nwet) return;
  const int KC = a.KC;
  const int64_t P = a.P;
  if (!wet_at(a, i)) {
    zero_dry_column(a, i);
    return;
  }
  if (a.lmask != nullptr) {
    // the old level of a face shared with a dry cell is zero (the
    // reference zeroes the faces of a newly dry cell, :230-262)
    if (a.lmask[a.lw[i]] == 0)
      for (int k = 0; k < KC; ++k) {
        const int64_t j = i + k * P;
        a.uhdy1[j] = a.uhdyf1[j] = a.u1[j] = 0.0;
      }
    if (a.lmask[a.ls[i]] == 0)
      for (int k = 0; k < KC; ++k) {
        const int64_t j = i + k * P;
        a.vhdx1[j] = a.vhdxf1[j] = a.v1[j] = 0.0;
      }
  }
  if (KC == 1) {
    // KC = 1: HDMT2T sets the single layer from the external flow
    // (EFDC/hdmt2t.f90:537-553)
    const double ue = a.uhdye[i], ve = a.vhdxe[i];
    a.uhdyf[i] = ue;
    a.uhdy[i] = ue;
    a.vhdxf[i] = ve;
    a.vhdx[i] = ve;
    a.u[i] = ue * a.hui[i] * a.dyiu[i];
    a.v[i] = ve * a.hvi[i] * a.dxiv[i];
    a.w[i + P] = 0.0;
    return;
  }
  // bottom drag coefficients, ISTL = 2 (:296-330); U = U1 here (the
  // advance has just copied them), so both speeds come from U1
  const double u1b = a.u1[i], v1b = a.v1[i];
  const double v1u = a.v1u[i], vu = a.vu[i], u1v = a.u1v[i], uv = a.uv[i];
  double q1 = sqrt(u1b * u1b + v1u * v1u);
  double q2 = sqrt(u1b * u1b + vu * vu);
  double rcx, rcy;
  if (a.avcon1 < 0.00001)
    rcx = a.stbx[i] * sqrt(q1 * q2);
  else
    rcx = a.avcon1 / sqrt(a.h1u[i] * a.hu[i]) + a.stbx[i] * sqrt(q1 * q2);
  q1 = sqrt(v1b * v1b + u1v * u1v);
  q2 = sqrt(v1b * v1b + uv * uv);
  if (a.avcon1 < 0.00001)
    rcy = a.stby[i] * sqrt(q1 * q2);
  else
    rcy = a.avcon1 / sqrt(a.h1v[i] * a.hv[i]) + a.stby[i] * sqrt(q1 * q2);
  // the external flows split by layer (:332-345)
  {
    const double ue = a.uhdye[i] - a.uhdye[a.le[i]];
    const double ve = a.vhdxe[i] - a.vhdxe[a.ln[i]];
    for (int k = 0; k < KC - 1; ++k) {
      const int64_t j = i + k * P;
      const double dz = (double)a.dzc[j];
      a.uhdyek[j] = dz * ue;
      a.vhdxek[j] = dz * ve;
    }
  }
  a.uhe[i] = face_column<T, KM, EXACT>(a, i, false, rcx);
  a.vhe[i] = face_column<T, KM, EXACT>(a, i, true, rcy);
}

// W from the layer continuity, ISTL = 2 (:665-683), zero on LOBCS cells
// (:687-694); the averages of the two levels (:770-791) and the column sums
// of the averaged layer flows (:813-836).
template <typename T>
__global__ __launch_bounds__(EFDC_BLOCK) void k_uvw_w(UvwArgs<T> a) {
  const int i = uvw_cell(a.nwet);
  if (i > a.nwet || !wet_at(a, i)) return;
  const int KC = a.KC;
  const int64_t P = a.P;
  const int e = a.le[i], n = a.ln[i];
  const bool obc = (a.lobflag[i] & 1) != 0;
  const double dxyip = a.dxyip[i], qsume = a.qsume[i], half = 0.5 * dxyip;
  double wprev = a.w[i];
  for (int k = 0; k < KC - 1; ++k) {
    const int64_t j = i + k * P, je = e + k * P, jn = n + k * P;
    const double wk =
        wprev -
        half * (a.uhdy[je] - a.uhdy[j] + a.uhdyek[j] + a.uhdy1[je] - a.uhdy1[j] + a.uhdy1ek[j] +
                a.vhdx[jn] - a.vhdx[j] + a.vhdxek[j] + a.vhdx1[jn] - a.vhdx1[j] + a.vhdx1ek[j]) +
        ((double)a.qsum[j] - (double)a.dzc[j] * qsume) * dxyip;
    a.w[j + P] = obc ? 0.0 : wk;
    wprev = wk;
  }
  double te = 0.0, tn = 0.0;
  for (int k = 0; k < KC; ++k) {
    const int64_t j = i + k * P;
    a.uhdyf2[j] = 0.5 * (a.uhdyf[j] + a.uhdyf1[j]);
    a.vhdxf2[j] = 0.5 * (a.vhdxf[j] + a.vhdxf1[j]);
    const double uh2 = 0.5 * (a.uhdy[j] + a.uhdy1[j]);
    const double vh2 = 0.5 * (a.vhdx[j] + a.vhdx1[j]);
    a.uhdy2[j] = (T)uh2;
    a.vhdx2[j] = (T)vh2;
    a.u2[j] = (T)(0.5 * (a.u[j] + a.u1[j]));
    a.v2[j] = (T)(0.5 * (a.v[j] + a.v1[j]));
    a.w2[j + P] = (T)(0.5 * (a.w[j + P] + a.w1[j + P]));
    te = te + uh2;
    tn = tn + vh2;
  }
  a.tvare[i] = te;
  a.tvarn[i] = tn;
}

// One side of the open-boundary layer flows (caluvw.f90:700-767) for cell
// LL at layer k.  side: 0 S, 1 W, 2 E, 3 N.
template <typename T>
__device__ void obc_face(const UvwArgs<T> &a, const UvwObcDev &b, int side, int LL, int k) {
  const int64_t P = a.P;
  const LayerConsts &C = *a.lc;
  const int L = b.cell[side][LL];
  const int isp = b.ispb[side][LL];
  if (side == 0) {
    if (isp == 2) return;
    const int LN = a.ln[L];
    if (LN == a.lc0) return;  // W(LC,K) = 0 only touches the ghost
    const int LNN = a.ln[LN];
    const int64_t j = LN + k * P;
    const double f = a.vhdxf[LNN + k * P] - a.vhdxe[LNN] + a.vhdxe[LN];
    a.vhdxf[j] = f;
    a.vhdx[j] = f * (a.svbo[LN] > 0.0 ? C.sg[k] : 0.0);
    a.v[j] = f / (a.hv[LN] * a.dxv[LN]);
    a.w[j + P] = 0.0;
  } else if (side == 1) {
    if (isp == 2) return;
    const int LE = a.le[L];
    if (LE == a.lc0) return;
    const int L1P = a.le[LE];
    const int64_t j = LE + k * P;
    const double f = a.uhdyf[L1P + k * P] - a.uhdye[L1P] + a.uhdye[LE];
    a.uhdyf[j] = f;
    a.uhdy[j] = f * (a.subo[LE] > 0.0 ? C.sg[k] : 0.0);
    a.u[j] = f / (a.hu[LE] * a.dyu[LE]);
    a.w[j + P] = 0.0;
  } else if (side == 2) {
    const int64_t j = L + k * P;
    if (isp != 2) {
      const int LW = a.lw[L];
      const double f = a.uhdyf[LW + k * P] - a.uhdye[LW] + a.uhdye[L];
      a.uhdyf[j] = f;
      a.uhdy[j] = f * (a.subo[L] > 0.0 ? C.sg[k] : 0.0);
      a.u[j] = f / (a.hu[L] * a.dyu[L]);
    }
    a.w[j + P] = 0.0;
  } else {
    const int64_t j = L + k * P;
    if (isp != 2) {
      const int LS = a.ls[L];
      const double f = a.vhdxf[LS + k * P] - a.vhdxe[LS] + a.vhdxe[L];
      a.vhdxf[j] = f;
      a.vhdx[j] = f * (a.svbo[L] > 0.0 ? C.sg[k] : 0.0);
      a.v[j] = f / (a.hv[L] * a.dxv[L]);
    }
    a.w[j + P] = 0.0;
  }
}

// One side of the open-boundary layer flows: one thread per (layer, cell)
// of the side (the sides run one launch after the other in the order S, W,
// E, N), or, when a cell of the side reads a face another cell of it
// writes, one thread walking the side layer by layer as the reference does.
template <typename T>
__global__ __launch_bounds__(EFDC_BLOCK) void k_uvw_obc_side(UvwArgs<T> a, UvwObcDev b,
                                                             int side) {
  const int n = b.n[side], KC = a.KC;
  if (b.serial[side]) {
    if (blockIdx.x == 0 && threadIdx.x == 0)
      for (int k = 0; k < KC; ++k)
        for (int LL = 0; LL < n; ++LL) obc_face(a, b, side, LL, k);
    return;
  }
  const int t = blockIdx.x * EFDC_BLOCK + (int)threadIdx.x;
  if (t >= n * KC) return;
  obc_face(a, b, side, t % n, t / n);
}

// The averages, W2 and column sums of the cells whose faces or W the sides
// changed: one thread per listed cell.
template <typename T>
__global__ __launch_bounds__(EFDC_BLOCK) void k_uvw_obc_avg(UvwArgs<T> a, UvwObcDev b) {
  const int KC = a.KC;
  const int64_t P = a.P;
  int q = blockIdx.x * EFDC_BLOCK + (int)threadIdx.x;
  if (q < b.navg_u) {
    const int x = b.avg_u[q];
    if (!wet_at(a, x)) return;
    double te = 0.0;
    for (int k = 0; k < KC; ++k) {
      const int64_t j = x + k * P;
      a.uhdyf2[j] = 0.5 * (a.uhdyf[j] + a.uhdyf1[j]);
      const double uh2 = 0.5 * (a.uhdy[j] + a.uhdy1[j]);
      a.uhdy2[j] = (T)uh2;
      a.u2[j] = (T)(0.5 * (a.u[j] + a.u1[j]));
      te = te + uh2;
    }
    a.tvare[x] = te;
    return;
  }
  q -= b.navg_u;
  if (q < b.navg_v) {
    const int x = b.avg_v[q];
    if (!wet_at(a, x)) return;
    double tn = 0.0;
    for (int k = 0; k < KC; ++k) {
      const int64_t j = x + k * P;
      a.vhdxf2[j] = 0.5 * (a.vhdxf[j] + a.vhdxf1[j]);
      const double vh2 = 0.5 * (a.vhdx[j] + a.vhdx1[j]);
      a.vhdx2[j] = (T)vh2;
      a.v2[j] = (T)(0.5 * (a.v[j] + a.v1[j]));
      tn = tn + vh2;
    }
    a.tvarn[x] = tn;
    return;
  }
  q -= b.navg_v;
  if (q < b.nwz) {
    const int x = b.wz[q];
    if (!wet_at(a, x)) return;
    for (int k = 0; k < KC; ++k) {
      const int64_t j = x + k * P;
      a.w2[j + P] = (T)(0.5 * (a.w[j + P] + a.w1[j + P]));
    }
  }
}

// The depth from the layer flows, 2TL (:949-1040), HP of open-boundary
// cells from P (:1130-1134), HPK and HPKI (:1161-1172).
template <typename T>
__global__ __launch_bounds__(EFDC_BLOCK) void k_uvw_hp(UvwArgs<T> a) {
  const int i = uvw_cell(a.nwet);
  if (i > a.nwet || !wet_at(a, i)) return;
  const int KC = a.KC;
  const int64_t P = a.P;
  const bool obc = (a.lobflag[i] & 1) != 0;
  double hp = a.hp[i];
  if (obc) {
    hp = a.gi * a.p[i] - a.belv[i];
  } else if (KC > 1) {
    const int e = a.le[i], n = a.ln[i];
    double hn = a.h1p[i] + a.dt * a.dxyip[i] *
                               (a.qsume[i] + a.tvare[i] + (-a.tvare[e]) + a.tvarn[i] - a.tvarn[n]);
    if (hn <= 0. && hp > 0.) {
      if (hp <= a.hdry && a.isdry > 0) {
        // the cell keeps its depth and its flows go (:957-1009)
        hn = hp;
        a.uhdye_w[i] = 0.;
        a.vhdxe_w[i] = 0.;
        for (int k = 0; k < KC; ++k) {
          const int64_t j = i + k * P, je = e + k * P, jn = n + k * P;
          a.uhdyf[j] = a.vhdxf[j] = a.uhdy[j] = a.vhdx[j] = a.u[j] = a.v[j] = 0.;
          a.w[j + P] = 0.;
          a.uhdyf[je] = 0.;
          a.vhdxf[jn] = 0.;
          a.uhdy[je] = 0.;
          a.vhdx[jn] = 0.;
          a.u[je] = 0.;
          a.v[jn] = 0.;
          a.uhdy2[j] = a.vhdx2[j] = a.u2[j] = a.v2[j] = (T)0;
          a.w2[j + P] = (T)0;
          a.uhdy2[je] = (T)0;
          a.vhdx2[jn] = (T)0;
          a.u2[je] = (T)0;
          a.v2[jn] = (T)0;
        }
      } else {
        __hip_atomic_store(a.flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    hp = hn;
  }
  a.hp[i] = hp;
  a.hpi[i] = 1. / hp;
  for (int k = 0; k < KC; ++k) {
    const int64_t j = i + k * P;
    const double hk = hp * (double)a.dzc[j];
    a.hpk[j] = (T)hk;
    a.hpki[j] = (T)(1. / hk);
  }
}

// U at V points and V at U points of the new level (EFDC/hdmt2t.f90:638-655).
template <typename T>
__global__ __launch_bounds__(EFDC_BLOCK) void k_uvw_uv(UvwArgs<T> a) {
  const int i = uvw_cell(a.nwet);
  if (i > a.nwet || !wet_at(a, i)) return;
  const int lse = a.lse[i], ls = a.ls[i], le = a.le[i], lnw = a.lnw[i], lw = a.lw[i],
            ln = a.ln[i];
  const double *hp = a.hp;
  a.u1v[i] = a.uv[i];
  a.v1u[i] = a.vu[i];
  a.uv[i] = ((hp[lse] * a.u[lse] + hp[ls] * a.u[ls]) + (hp[le] * a.u[le] + hp[i] * a.u[i])) /
            (1.0 + a.sub[le] + a.sub[ls] + a.sub[lse]) * a.hvi[i];
  a.vu[i] = ((hp[lnw] * a.v[lnw] + hp[lw] * a.v[lw]) + (hp[ln] * a.v[ln] + hp[i] * a.v[i])) /
            (1.0 + a.svb[ln] + a.svb[lw] + a.svb[lnw]) * a.hui[i];
}

template <typename T>
T *kpt(efdc_ctx *c, int id) {
  return (T *)c->dK[id - EFDC_HPK];
}

template <typename T>
UvwArgs<T> uvw_args(efdc_ctx *c, const efdc_step_t *st, const efdc_uvw_opts_t *o) {
  UvwState *u = c->uvw;
  UvwArgs<T> a;
  double **H = c->dH;
  a.uhdye = H[EFDC_UHDYE];
  a.vhdxe = H[EFDC_VHDXE];
  a.hu = H[EFDC_HU];
  a.hv = H[EFDC_HV];
  a.hui = H[EFDC_HUI];
  a.hvi = H[EFDC_HVI];
  a.h1u = H[EFDC_H1U];
  a.h1v = H[EFDC_H1V];
  a.sub = H[EFDC_SUB];
  a.svb = H[EFDC_SVB];
  a.subo = H[EFDC_SUBO];
  a.svbo = H[EFDC_SVBO];
  a.dyiu = H[EFDC_DYIU];
  a.dxiv = H[EFDC_DXIV];
  a.dxyip = H[EFDC_DXYIP];
  a.qsume = H[EFDC_QSUME];
  a.h1p = H[EFDC_H1P];
  a.p = H[EFDC_P];
  a.belv = H[EFDC_BELV];
  a.tsx = H[EFDC_TSX];
  a.tsy = H[EFDC_TSY];
  a.uhe = H[EFDC_UHE];
  a.vhe = H[EFDC_VHE];
  a.hp = H[EFDC_HP];
  a.hpi = H[EFDC_HPI];
  a.uhdye_w = H[EFDC_UHDYE];
  a.vhdxe_w = H[EFDC_VHDXE];
  a.stbx = IH(STBX);
  a.stby = IH(STBY);
  a.dyu = IH(DYU);
  a.dxv = IH(DXV);
  a.vu = IH(VU);
  a.uv = IH(UV);
  a.v1u = IH(V1U);
  a.u1v = IH(U1V);
  a.uhdyf = IK(UHDYF);
  a.vhdxf = IK(VHDXF);
  a.uhdy = IK(UHDY);
  a.vhdx = IK(VHDX);
  a.uhdyf1 = IK(UHDYF1);
  a.vhdxf1 = IK(VHDXF1);
  a.uhdy1 = IK(UHDY1);
  a.vhdx1 = IK(VHDX1);
  a.u = IK(U);
  a.v = IK(V);
  a.u1 = IK(U1);
  a.v1 = IK(V1);
  a.uhdyf2 = IK(UHDYF2);
  a.vhdxf2 = IK(VHDXF2);
  a.uhdyek = IK(UHDYEK);
  a.vhdxek = IK(VHDXEK);
  a.uhdy1ek = IK(UHDY1EK);
  a.vhdx1ek = IK(VHDX1EK);
  a.du = IK(DU);
  a.dv = IK(DV);
  a.avui = IK(AVUI);
  a.avvi = IK(AVVI);
  a.w = IW(W);
  a.w1 = IW(W1);
  a.dzc = kpt<T>(c, EFDC_DZC);
  a.qsum = kpt<T>(c, EFDC_QSUM);
  a.sub3d = kpt<T>(c, EFDC_SUB3D);
  a.svb3d = kpt<T>(c, EFDC_SVB3D);
  a.uhdy2 = kpt<T>(c, EFDC_UHDY2);
  a.vhdx2 = kpt<T>(c, EFDC_VHDX2);
  a.u2 = kpt<T>(c, EFDC_U2);
  a.v2 = kpt<T>(c, EFDC_V2);
  a.w2 = (T *)c->dW[EFDC_W2 - EFDC_W2];
  a.hpk = kpt<T>(c, EFDC_HPK);
  a.hpki = kpt<T>(c, EFDC_HPKI);
  a.lw = c->dLW;
  a.le = c->dLE;
  a.ls = c->dLS;
  a.ln = c->dLN;
  a.lnw = c->dLNW;
  a.lse = c->dLSE;
  a.lmask = c->prm.ISDRY > 0 ? c->lmaskdry : nullptr;
  a.lobflag = c->lobflag;
  a.tvare = u->tvare;
  a.tvarn = u->tvarn;
  a.lc = u->lc;
  a.flag = u->flag;
  a.P = c->pitch;
  a.KC = c->KC;
  a.nwet = c->n_wet();
  a.lc0 = c->LC - 1;
  a.isdry = c->prm.ISDRY;
  a.shear = o->shear ? 1 : 0;
  a.wind = o->wind ? 1 : 0;
  a.dt = st->dt;
  a.dti = 1. / st->dt;
  a.avcon1 = o->AVCON1;
  a.gi = 1. / c->prm.G;
  a.hdry = c->prm.HDRY;
  return a;
}

// The face constants of a sigma grid with layer thicknesses dz[K-1], in
// the reference's order of operations (see LayerConsts).
void build_layer_consts(const std::vector<double> &dz, int KC, LayerConsts &t) {
  std::memset(&t, 0, sizeof(t));
  const int KS = KC - 1;
  for (int k = 0; k < KC; ++k) t.sg[k] = dz[k];
  for (int k = 0; k < KS; ++k) {
    const double a = t.sg[k], b = t.sg[k + 1];
    t.dzg[k] = 0.5 * (a + b);
    t.cdzf[k] = a * b / (a + b);
    t.cdzu[k] = -a / (a + b);
    t.cdzl[k] = -b / (a + b);
  }
  if (KS >= 1) {
    t.cdzr[0] = t.sg[0] - 1.;
    t.cdzd[0] = t.sg[0];
    for (int k = 1; k < KS; ++k) {
      t.cdzr[k] = t.cdzr[k - 1] + t.sg[k];
      t.cdzd[k] = t.cdzd[k - 1] + t.sg[k];
    }
  }
  for (int k = 0; k < KS; ++k) {
    t.cdzr[k] = t.cdzr[k] * t.dzg[k] * t.cdzl[0];
    t.cdzm[k] = 0.5 * t.sg[k] * t.sg[k + 1];
  }
  if (KC > 1) {
    double q1 = 0., qc = 0.;
    for (int k = 0; k < KC; ++k) {
      q1 = q1 + 1. / (1. - t.sg[k]);
      qc = qc + 1. / (1. - 0.);
    }
    for (int k = 0; k < KC; ++k) {
      t.cerr[k] = (1. / (1. - t.sg[k])) / q1;
      t.cerr[k] = t.cerr[k] * (double)KC;
      t.cerr_closed[k] = (1. / (1. - 0.)) / qc;
      t.cerr_closed[k] = t.cerr_closed[k] * (double)KC;
    }
  } else {
    t.cerr[0] = 1.;
    t.cerr_closed[0] = 1.;
  }
}

int obc_build(efdc_ctx *c) {
  UvwState *u = c->uvw;
  UvwObcDev &b = u->obc;
  for (int s = 0; s < 4; ++s) {
    dfree_dev(b.cell[s]);
    dfree_dev(b.ispb[s]);
  }
  dfree_dev(b.avg_u);
  dfree_dev(b.avg_v);
  dfree_dev(b.wz);
  std::memset(&b, 0, sizeof(b));
  // concentration-BC sides S, W, E, N and the pressure side (W=0, E=1,
  // S=2, N=3) whose ISPB they use
  static const int pside[4] = {2, 0, 1, 3};
  const int LC0 = c->LC - 1;
  std::vector<int> au, av, wz;
  const std::vector<int> &LE = c->host_le, &LN = c->host_ln, &LW = c->host_lw,
                         &LS = c->host_ls;
  for (int s = 0; s < 4; ++s) {
    const std::vector<int> &cells = c->cbc_cells_host[s];
    const std::vector<int> &isp = c->pbc_ispb_host[pside[s]];
    const int n = (int)cells.size();
    std::vector<int> ispb(n);
    std::vector<int> tgt, src;
    for (int LL = 0; LL < n; ++LL) {
      const int L = cells[LL];
      ispb[LL] = LL < (int)isp.size() ? isp[LL] : 0;
      const bool mod = ispb[LL] != 2;
      if (s == 0 && mod && LN[L] != LC0) {
        av.push_back(LN[L]);
        wz.push_back(LN[L]);
        tgt.push_back(LN[L]);
        src.push_back(LN[LN[L]]);
      } else if (s == 1 && mod && LE[L] != LC0) {
        au.push_back(LE[L]);
        wz.push_back(LE[L]);
        tgt.push_back(LE[L]);
        src.push_back(LE[LE[L]]);
      } else if (s == 2) {
        if (mod) {
          au.push_back(L);
          tgt.push_back(L);
          src.push_back(LW[L]);
        }
        wz.push_back(L);
      } else if (s == 3) {
        if (mod) {
          av.push_back(L);
          tgt.push_back(L);
          src.push_back(LS[L]);
        }
        wz.push_back(L);
      }
    }
    std::vector<int> st = tgt;
    std::sort(st.begin(), st.end());
    bool serial = false;
    for (int x : src)
      if (std::binary_search(st.begin(), st.end(), x)) serial = true;
    b.n[s] = n;
    b.serial[s] = serial ? 1 : 0;
    int e;
    if ((e = upload_dev(c, &b.cell[s], cells)) || (e = upload_dev(c, &b.ispb[s], ispb)))
      return e;
  }
  auto uniq = [](std::vector<int> &v) {
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
  };
  uniq(au);
  uniq(av);
  uniq(wz);
  b.navg_u = (int)au.size();
  b.navg_v = (int)av.size();
  b.nwz = (int)wz.size();
  int e;
  if ((e = upload_dev(c, &b.avg_u, au)) || (e = upload_dev(c, &b.avg_v, av)) ||
      (e = upload_dev(c, &b.wz, wz)))
    return e;
  u->obc_ready = true;
  return EFDC_OK;
}

int uvw_alloc(efdc_ctx *c) {
  if (c->uvw) return EFDC_OK;
  UvwState *u = new UvwState();
  c->uvw = u;
  const size_t L = (size_t)c->LCM;
  const size_t kel = (size_t)c->pitch * c->KC, wel = (size_t)c->pitch * (c->KC + 1);
  int s;
  for (int j = 0; j < NIH; ++j)
    if ((s = dev_alloc(c, (void **)&u->h[j], L * sizeof(double)))) return s;
  for (int j = 0; j < NIK; ++j)
    if ((s = dev_alloc(c, (void **)&u->k[j], kel * sizeof(double)))) return s;
  for (int j = 0; j < NIW; ++j)
    if ((s = dev_alloc(c, (void **)&u->w[j], wel * sizeof(double)))) return s;
  if ((s = dev_alloc(c, (void **)&u->tvare, L * sizeof(double))) ||
      (s = dev_alloc(c, (void **)&u->tvarn, L * sizeof(double))) ||
      (s = dev_alloc(c, (void **)&u->lc, sizeof(LayerConsts))) ||
      (s = dev_alloc(c, (void **)&u->flag, sizeof(int))))
    return s;
  if (hipHostMalloc((void **)&u->flag_host, sizeof(int)) != hipSuccess) {
    c->err = "hipHostMalloc failed (internal-mode flag)";
    return EFDC_ERR_NOMEM;
  }
  *u->flag_host = 0;
  return EFDC_OK;
}

template <typename T>
int uvw_launch(efdc_ctx *c, const efdc_step_t *st, const efdc_uvw_opts_t *o) {
  UvwState *u = c->uvw;
  const UvwArgs<T> a = uvw_args<T>(c, st, o);
  const int nb = efdc_grid_blocks(c->n_wet() + 1);
  const int KC = c->KC;
  switch (KC) {  // the common layer counts get a kernel of their own
#define UVW_EXACT(n)                                                                     \
  case n:                                                                                \
    TIMED(c, TK_UVW_COL, k_uvw_col<T, n, true><<<nb, EFDC_BLOCK, 0, c->stream>>>(a)); \
    break;
    UVW_EXACT(2) UVW_EXACT(3) UVW_EXACT(4) UVW_EXACT(5) UVW_EXACT(6) UVW_EXACT(8)
    UVW_EXACT(10) UVW_EXACT(12) UVW_EXACT(16) UVW_EXACT(20)
#undef UVW_EXACT
    default:
      if (KC <= 8)
        TIMED(c, TK_UVW_COL, k_uvw_col<T, 8, false><<<nb, EFDC_BLOCK, 0, c->stream>>>(a));
      else if (KC <= 16)
        TIMED(c, TK_UVW_COL, k_uvw_col<T, 16, false><<<nb, EFDC_BLOCK, 0, c->stream>>>(a));
      else if (KC <= 24)
        TIMED(c, TK_UVW_COL, k_uvw_col<T, 24, false><<<nb, EFDC_BLOCK, 0, c->stream>>>(a));
      else if (KC <= 32)
        TIMED(c, TK_UVW_COL, k_uvw_col<T, 32, false><<<nb, EFDC_BLOCK, 0, c->stream>>>(a));
      else
        TIMED(c, TK_UVW_COL, k_uvw_col<T, 64, false><<<nb, EFDC_BLOCK, 0, c->stream>>>(a));
  }
  TIMED(c, TK_UVW_W, k_uvw_w<T><<<nb, EFDC_BLOCK, 0, c->stream>>>(a));
  const UvwObcDev &b = u->obc;
  if (b.n[0] + b.n[1] + b.n[2] + b.n[3] > 0) {
    for (int side = 0; side < 4; ++side) {
      if (b.n[side] == 0) continue;
      const int nt = b.serial[side] ? 1 : (b.n[side] * KC + EFDC_BLOCK - 1) / EFDC_BLOCK;
      TIMED(c, TK_UVW_OBC,
            k_uvw_obc_side<T><<<nt, EFDC_BLOCK, 0, c->stream>>>(a, b, side));
    }
    const int na = b.navg_u + b.navg_v + b.nwz;
    if (na > 0)
      TIMED(c, TK_UVW_OBC, k_uvw_obc_avg<T><<<(na + EFDC_BLOCK - 1) / EFDC_BLOCK, EFDC_BLOCK,
                                               0, c->stream>>>(a, b));
  }
  TIMED(c, TK_UVW_HP, k_uvw_hp<T><<<nb, EFDC_BLOCK, 0, c->stream>>>(a));
  if (o->uv_after) TIMED(c, TK_UVW_UV, k_uvw_uv<T><<<nb, EFDC_BLOCK, 0, c->stream>>>(a));
  HIP_TRY(c, hipGetLastError());
  return EFDC_OK;
}

}  // namespace

// ------------------------------------------------------------ host entry
bool uvw_id(int id) {
  return (id >= EFDC_STBX && id < EFDC_NUM_IH_END) ||
         (id >= EFDC_UHDYF && id < EFDC_NUM_IK_END) || (id >= EFDC_W && id < EFDC_NUM_IW_END);
}

int uvw_ptr(efdc_ctx *c, int id, double **p, int *planes) {
  int s;
  if ((s = uvw_alloc(c))) return s;
  UvwState *u = c->uvw;
  if (id >= EFDC_STBX && id < EFDC_NUM_IH_END) {
    *p = u->h[id - EFDC_STBX];
    *planes = 0;
  } else if (id >= EFDC_UHDYF && id < EFDC_NUM_IK_END) {
    *p = u->k[id - EFDC_UHDYF];
    *planes = c->KC;
  } else {
    *p = u->w[id - EFDC_W];
    *planes = c->KC + 1;
  }
  return EFDC_OK;
}

void uvw_step_begin(efdc_ctx *c) {
  // UHDY1EK = UHDYEK, VHDX1EK = VHDXEK (EFDC/calpuv2c.f90:270-272); the
  // internal mode rewrites UHDYEK of every wet cell and copies it back for
  // the dry ones, so a swap of the two buffers is the copy
  UvwState *u = c->uvw;
  if (!u) return;
  std::swap(IK(UHDYEK), IK(UHDY1EK));
  std::swap(IK(VHDXEK), IK(VHDX1EK));
}

void uvw_state_pieces(efdc_ctx *c, std::vector<std::pair<void *, size_t>> &v) {
  UvwState *u = c->uvw;
  if (!u) return;
  const size_t kel = (size_t)c->pitch * c->KC, wel = (size_t)c->pitch * (c->KC + 1);
  for (int j = 0; j < NIH; ++j) v.push_back({u->h[j], sizeof(double) * c->LCM});
  for (int j = 0; j < NIK; ++j) v.push_back({u->k[j], sizeof(double) * kel});
  for (int j = 0; j < NIW; ++j) v.push_back({u->w[j], sizeof(double) * wel});
}

void uvw_obc_changed(efdc_ctx *c) {
  if (c->uvw) c->uvw->obc_ready = false;
}

void uvw_free(efdc_ctx *c) {
  UvwState *u = c->uvw;
  if (!u) return;
  for (auto &p : u->h) dfree_dev(p);
  for (auto &p : u->k) dfree_dev(p);
  for (auto &p : u->w) dfree_dev(p);
  dfree_dev(u->tvare);
  dfree_dev(u->tvarn);
  dfree_dev(u->lc);
  dfree_dev(u->flag);
  for (int s = 0; s < 4; ++s) {
    dfree_dev(u->obc.cell[s]);
    dfree_dev(u->obc.ispb[s]);
  }
  dfree_dev(u->obc.avg_u);
  dfree_dev(u->obc.avg_v);
  dfree_dev(u->obc.wz);
  if (u->flag_host) (void)hipHostFree(u->flag_host);
  delete u;
  c->uvw = nullptr;
}

int uvw_run(efdc_ctx *c, const efdc_step_t *st, const efdc_uvw_opts_t *o) {
  if (!c->dzc_uni) {
    c->err = "the internal mode needs DZC set and equal in every wet cell of a layer "
             "(sigma layers, IGRIDV = 0)";
    return EFDC_ERR_ARG;
  }
  if (o->uv_after && !c->has_diag) {
    c->err = "uv_after needs the diagonal maps LNWC and LSEC";
    return EFDC_ERR_ARG;
  }
  if (c->nseg > 1) {
    c->err = "the internal mode with several basins is not built";
    return EFDC_ERR_ARG;
  }
  int s;
  if ((s = uvw_alloc(c))) return s;
  UvwState *u = c->uvw;
  if (u->lc_kc != c->KC || c->dzc_changed) {
    build_layer_consts(c->dzc_k, c->KC, u->lc_host);
    HIP_TRY(c, hipMemcpyAsync(u->lc, &u->lc_host, sizeof(LayerConsts), hipMemcpyHostToDevice,
                              c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));  // the host copy may change later
    u->lc_kc = c->KC;
    c->dzc_changed = false;
  }
  if (!u->obc_ready && (s = obc_build(c))) return s;
  // advance the internal variables, EFDC/hdmt2t.f90:513-527
  std::swap(IK(UHDYF), IK(UHDYF1));
  std::swap(IK(VHDXF), IK(VHDXF1));
  std::swap(IK(UHDY), IK(UHDY1));
  std::swap(IK(VHDX), IK(VHDX1));
  std::swap(IK(U), IK(U1));
  std::swap(IK(V), IK(V1));
  std::swap(IW(W), IW(W1));
  k_uvw_flag_reset<<<1, 1, 0, c->stream>>>(u->flag);
  if (c->store == EFDC_STORE_FP64)
    s = uvw_launch<double>(c, st, o);
  else
    s = uvw_launch<float>(c, st, o);
  if (s) return s;
  HIP_TRY(c, hipMemcpyAsync(u->flag_host, u->flag, sizeof(int), hipMemcpyDeviceToHost,
                            c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (*u->flag_host) {
    c->err = "CALUVW: negative depth in the continuity check";
    return EFDC_ERR_NEGDEPTH;
  }
  return EFDC_OK;
}
