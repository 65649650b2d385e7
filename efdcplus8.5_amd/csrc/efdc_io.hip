// efdc_io.hip -- the on-disk formats EFDC_Explorer and GetEFDC read, written
// from the state on the device:
//
//   RESTART.OUT (RESTOUT, EFDC/restartmod.f90:17-297): Fortran formatted
//     text, every real as E17.8 (FORMAT 906 = 50E17.8) and every counter as
//     I10 (908 = 50I10), records of at most 50 values.  Each value has a
//     fixed width, so the byte offset of every value in the file is known
//     before anything is formatted: one thread formats one value
//     (efdc_fmt.h) into its slot of a device buffer that holds the whole
//     file body, and the host writes it out with one copy.  The few values
//     the device formatter hands back (ties, near-ties, magnitudes outside
//     [1e-280, 1e280]) are formatted on the host from their exact decimal
//     expansion and patched in before the write.
//
//   EE_WS.OUT, EE_VEL.OUT (HEADEROUT, WSOUT, VELOUT, EFDC/efdcout.f90:73-117,
//     371-497): unformatted stream records (FORM='BINARY') of INT*4, REAL*8
//     and REAL*4.  One kernel converts HP and the layer velocities U, V, W of
//     the cells' active layers to REAL*4 in file order; the host appends the
//     two records.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "efdc_internal.h"
#include "efdc_fmt.h"

namespace {

__device__ const double g_pow10[EFDC_POW10_MAX - EFDC_POW10_MIN + 1][2] = EFDC_POW10_INIT;

using efdc_fmt::E_W;
using efdc_fmt::I_W;
constexpr int PER_REC = 50;  // 50E17.8 / 50I10: values per record

__global__ void k_fill_u32(unsigned *p, int n, unsigned v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

enum { SRC_F64 = 0, SRC_F32 = 1, SRC_NLO = 2 };
enum { D_INT = 1, D_NL = 2 };

struct IoSrc {
  const void *p;
  int64_t ks;   // stride between layers (cell arrays)
  int32_t type;
  int32_t _pad;
};

struct IoDesc {
  int32_t src, k, off, flags;
};

// a descriptor with its source resolved (what the kernels read)
struct IoVal {
  const void *p;
  int64_t ks;
  int32_t k, off, flags, type;
};

// A run of items (cells L = 2..LA, or the cells of one concentration-BC
// side) that all have the same records.
struct IoSeg {
  int64_t file_off, val0, nitems, item_bytes;
  int64_t blk0;           // first workgroup of the segment (staged kernel)
  int32_t cper;           // items a workgroup stages (staged kernel)
  int32_t d0, nd;
  int32_t bc;             // 1: item = LL of a BC side, values at [(LL*KC+K)*nc+m]
  int32_t m, nc, kc;
  const int32_t *ntscr;   // NTSCR of the side (the NLO write-back)
};

struct FmtArgs {
  const IoSeg *seg;
  int nseg;
  const IoVal *desc;
  char *out;
  int64_t nvals;
  long long n;            // N of RESTOUT's NLO = MAX(NLO - N, -NTSCR)
  int update_nlo;
  unsigned *nflag;
  int64_t *flag_off;
  double *flag_val;
  unsigned flag_cap;
};

// One value of the restart into o (a byte of LDS or of the file image);
// gofs = its offset in the file body, for the values the host formats.
__device__ __forceinline__ int fmt_value(const FmtArgs &a, const IoSeg &sg, int64_t item, int j,
                                         char *o, int64_t gofs) {
  const IoVal d = a.desc[sg.d0 + j];
  const int64_t idx = sg.bc ? ((item * sg.kc + d.k) * sg.nc + sg.m) : (item + 1 + d.k * d.ks);
  int w;
  if (d.flags & D_INT) {
    long long *np = (long long *)d.p + idx;
    long long v = *np;
    if (d.type == SRC_NLO) {  // the value RESTOUT writes, and leaves in memory
      v = v - a.n;
      const long long floor_ = -(long long)sg.ntscr[item];
      if (v < floor_) v = floor_;
      if (a.update_nlo) *np = v;
    }
    efdc_fmt::fmt_i10(v, o);
    w = I_W;
  } else {
    const double v = d.type == SRC_F32 ? (double)((const float *)d.p)[idx]
                                       : ((const double *)d.p)[idx];
    if (!efdc_fmt::fmt_e17(g_pow10, v, o)) {
      const unsigned q = atomicAdd(a.nflag, 1u);
      if (q < a.flag_cap) {
        a.flag_off[q] = gofs;
        a.flag_val[q] = v;
      }
    }
    w = E_W;
  }
  if (d.flags & D_NL) o[w] = '\n';
  return w;
}

// file bytes a workgroup stages in LDS (19 KB: seven workgroups of 256
// lanes per CU, the most the 70 VGPRs of a lane allow; 1.90 ms against
// 1.97 ms with 39 KB on a 2.6 GB restart)
#ifndef EFDC_IO_STAGE_KB
#define EFDC_IO_STAGE_KB 19
#endif
constexpr int STAGE_BYTES = EFDC_IO_STAGE_KB * 1024;

// A workgroup formats C consecutive items of one segment (C = sg.cper,
// C * item_bytes <= STAGE_BYTES) into LDS, lanes over items first so that
// the loads of one record are contiguous, then writes the contiguous file
// range with 16-byte stores aligned as in the file (the partial 16-byte
// pieces at either end byte by byte).
__global__ __launch_bounds__(256) void k_fmt_staged(FmtArgs a) {
  extern __shared__ __align__(16) char lds[];
  // consecutive chunks on one XCD: the cache lines a chunk shares with its
  // neighbours (the next cells of every array) are then fetched into one L2
  const int64_t blk = xcd_tile((int)blockIdx.x, (int)gridDim.x);
  int s = 0;
  while (s + 1 < a.nseg && a.seg[s + 1].blk0 <= blk) ++s;
  const IoSeg sg = a.seg[s];
  const unsigned C = (unsigned)sg.cper;
  const int64_t item0 = (blk - sg.blk0) * C;
  const unsigned ce = (unsigned)(sg.nitems - item0 < (int64_t)C ? sg.nitems - item0 : C);
  const int64_t g0 = sg.file_off + item0 * sg.item_bytes;
  const int64_t g1 = g0 + (int64_t)ce * sg.item_bytes;
  const int64_t base = g0 & ~(int64_t)15;
  const int lead = (int)(g0 - base);
  const unsigned nv = (unsigned)sg.nd * ce;
#ifndef EFDC_IO_UNROLL
#define EFDC_IO_UNROLL 2
#endif
#pragma unroll EFDC_IO_UNROLL
  for (unsigned v = threadIdx.x; v < nv; v += 256) {
    const unsigned j = v / ce, cl = v - j * ce;
    const int64_t at = (int64_t)cl * sg.item_bytes + a.desc[sg.d0 + j].off;
    fmt_value(a, sg, item0 + cl, (int)j, lds + lead + at, g0 + at);
  }
  __syncthreads();
  const int nq = (int)((g1 - base + 15) >> 4);
  for (int q = threadIdx.x; q < nq; q += 256) {
    const int64_t ga = base + ((int64_t)q << 4);
    if (ga >= g0 && ga + 16 <= g1) {
      *(uint4 *)(a.out + ga) = *(const uint4 *)(lds + (q << 4));
    } else {
      for (int b = 0; b < 16; ++b)
        if (ga + b >= g0 && ga + b < g1) a.out[ga + b] = lds[(q << 4) + b];
    }
  }
}

// One thread per value, written straight into the file image: the path
// for records too long to stage (an item above STAGE_BYTES).
__global__ __launch_bounds__(256) void k_fmt_restart(FmtArgs a) {
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= a.nvals) return;
  int s = 0;
  while (s + 1 < a.nseg && a.seg[s + 1].val0 <= g) ++s;
  const IoSeg sg = a.seg[s];
  const int64_t r = g - sg.val0;
  const int64_t item = r / sg.nd;
  const int j = (int)(r - item * sg.nd);
  const int64_t at = sg.file_off + item * sg.item_bytes + a.desc[sg.d0 + j].off;
  fmt_value(a, sg, item, j, a.out + at, at);  // (k_fmt_restart)
}

// --- EE_WS / EE_VEL: REAL*4 of HP (L = 2..LA) and of U, V, W over the
// active layers K = KSZ(L)..KC of every cell, in file order
__global__ __launch_bounds__(256) void k_ee_pack(const double *hp, const double *u,
                                                 const double *v, const double *w,
                                                 const int *ksz, const int64_t *off, int KC,
                                                 int64_t pitch, int nwet, int64_t cell3d,
                                                 float *out) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= (int64_t)nwet * KC) return;
  const int c = (int)(t / KC), k = (int)(t - (int64_t)c * KC);  // k = K - 1
  const int L0 = c + 1;
  if (k == 0) out[c] = (float)hp[L0];
  const int k0 = ksz ? ksz[L0] - 1 : 0;
  if (k < k0) return;
  const int64_t o = nwet + (off ? off[c] : (int64_t)c * KC) + (k - k0);
  out[o] = (float)u[L0 + k * pitch];
  out[o + cell3d] = (float)v[L0 + k * pitch];
  out[o + 2 * cell3d] = (float)w[L0 + (k + 1) * pitch];
}

}  // namespace

struct IoState {
  char *dbuf = nullptr;       // device file body
  size_t dbuf_cap = 0;
  double *stage = nullptr;    // host arrays staged on the device
  size_t stage_cap = 0;
  void *meta = nullptr;       // segments, descriptors, sources
  size_t meta_cap = 0;
  unsigned *nflag = nullptr;
  int64_t *flag_off = nullptr;
  double *flag_val = nullptr;
  unsigned flag_cap = 0;
  char *hbuf = nullptr;       // pinned host copy of the file
  size_t hbuf_cap = 0;
  // EE files
  bool ee_open = false;
  int ic = 0, jc = 0, igridv = 0;
  int64_t cell3d = 0;
  int *ksz = nullptr;         // device KSZ(L) (0-based cell index), or null (all 1)
  int64_t *off = nullptr;     // device first value of each cell in the layer lists
  float *pack = nullptr;
  size_t pack_cap = 0;
};

namespace {

int grow(efdc_ctx *c, void **p, size_t *cap, size_t need) {
  if (*cap >= need) return EFDC_OK;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  int s = dev_alloc(c, p, need);
  if (s) return s;
  *cap = need;
  return EFDC_OK;
}

int grow_host(efdc_ctx *c, IoState *io, size_t need) {
  if (io->hbuf_cap >= need) return EFDC_OK;
  if (io->hbuf) (void)hipHostFree(io->hbuf);
  io->hbuf = nullptr;
  io->hbuf_cap = 0;
  if (hipHostMalloc((void **)&io->hbuf, need, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    c->err = "pinned host memory for the file could not be allocated";
    return EFDC_ERR_NOMEM;
  }
  io->hbuf_cap = need;
  return EFDC_OK;
}

IoState *io_state(efdc_ctx *c) {
  if (!c->io) c->io = new IoState();
  return c->io;
}

// The exact E17.8 text of v (for the values the device hands over): the
// shortest exact decimal expansion from the C library, rounded to eight
// significant digits with ties to even.
void host_e17(double v, char *o) {
  char buf[E_W + 1];
  if (std::isnan(v)) { std::snprintf(buf, sizeof buf, "%17s", "NaN"); std::memcpy(o, buf, E_W); return; }
  if (std::isinf(v)) { std::snprintf(buf, sizeof buf, "%17s", v < 0 ? "-Inf" : "Inf"); std::memcpy(o, buf, E_W); return; }
  // glibc's printf rounds the exact binary value to nearest, ties to even
  char s[40];
  std::snprintf(s, sizeof s, "%.7e", v);  // [-]d.ddddddde[+-]xx[x]
  const bool neg = s[0] == '-';
  const char *m = s + (neg ? 1 : 0);
  char dig[8];
  dig[0] = m[0];
  std::memcpy(dig + 1, m + 2, 7);
  int e10 = std::atoi(m + 10);
  int E = (v == 0.0) ? 0 : e10 + 1;
  std::string r = neg ? "-0." : "0.";
  r.append(dig, 8);
  const int ae = E < 0 ? -E : E;
  char ex[8];
  if (ae <= 99) std::snprintf(ex, sizeof ex, "E%c%02d", E < 0 ? '-' : '+', ae);
  else std::snprintf(ex, sizeof ex, "%c%03d", E < 0 ? '-' : '+', ae);
  r += ex;
  std::snprintf(buf, sizeof buf, "%17s", r.c_str());
  std::memcpy(o, buf, E_W);
}

struct Layout {
  std::vector<IoSrc> src;
  std::vector<IoDesc> desc;
  std::vector<IoSeg> seg;
  std::vector<std::pair<const double *, size_t>> host;  // staged host arrays (doubles)
  std::vector<size_t> host_src;                          // their src slots
  int64_t body = 0, nvals = 0;
};

struct Rec {  // one Fortran WRITE: values of (src, plane), E17.8 or I10
  std::vector<std::pair<int, int>> v;
  bool integer = false;
};

// lay out the records of one item; returns the item's byte size
int64_t lay_item(Layout &L, const std::vector<Rec> &recs) {
  int64_t off = 0;
  for (const Rec &r : recs) {
    const int w = r.integer ? I_W : E_W;
    for (size_t i = 0; i < r.v.size(); ++i) {
      const bool nl = (i + 1 == r.v.size()) || ((i + 1) % PER_REC == 0);
      L.desc.push_back({r.v[i].first, r.v[i].second, (int32_t)off,
                        (r.integer ? D_INT : 0) | (nl ? D_NL : 0)});
      off += w + (nl ? 1 : 0);
    }
  }
  return off;
}

void add_seg(Layout &L, int d0, int64_t nitems, int64_t item_bytes, int bc, int m, int nc,
             int kc, const int32_t *ntscr) {
  const int nd = (int)L.desc.size() - d0;
  if (nitems <= 0 || nd == 0) {
    L.desc.resize(d0);
    return;
  }
  IoSeg s{};
  s.file_off = L.body;
  s.val0 = L.nvals;
  s.nitems = nitems;
  s.item_bytes = item_bytes;
  s.d0 = d0;
  s.nd = nd;
  s.bc = bc;
  s.m = m;
  s.nc = nc;
  s.kc = kc;
  s.ntscr = ntscr;
  L.seg.push_back(s);
  L.body += nitems * item_bytes;
  L.nvals += nitems * nd;
}

bool writes(const efdc_restart_t *r, int m) { return r->istran[m] >= 1 && r->isco[m] == 1; }

int build_restart(efdc_ctx *c, const efdc_restart_t *r, Layout &L) {
  if (writes(r, 4) || writes(r, 5) || writes(r, 7)) {
    c->err = "restart output of shellfish larvae (ISTRAN(4)), toxics (ISTRAN(5)) or "
             "non-cohesive sediment (ISTRAN(7)) is not built: those classes are not "
             "transported by this library";
    return EFDC_ERR_ARG;
  }
  if (!r->hwq || !r->h2wq || !r->qq || !r->qq1 || !r->qql || !r->qql1 || !r->dml) {
    c->err = "the restart needs the host's HWQ, H2WQ, QQ, QQ1, QQL, QQL1 and DML";
    return EFDC_ERR_ARG;
  }
  const bool ice = r->istran[2] >= 1 && r->isice > 2;
  if (ice && (!r->ice_thick || !r->ice_temp)) {
    c->err = "ISICE > 2: the restart needs the ice thickness and temperature";
    return EFDC_ERR_ARG;
  }
  const int nc = c->nconst, KC = c->KC, LCM = c->LCM;
  auto slot_ok = [&](int s) { return s >= 0 && s < nc; };
  for (int m = 1; m <= 3; ++m)
    if (writes(r, m) && !slot_ok(r->slot[m - 1])) {
      c->err = "a constituent written to the restart has no valid slot";
      return EFDC_ERR_ARG;
    }
  if (writes(r, 2) && !r->temb) {
    c->err = "the restart of temperature needs TEMB";
    return EFDC_ERR_ARG;
  }
  const bool sed = writes(r, 6);
  if (sed) {
    if (r->nsed < 0 || (r->nsed > 0 && !r->sed_slot) || r->kb < 1 || r->kbm < r->kb ||
        !r->sedb || !r->sedb1 || !r->hbed || !r->hbed1 || !r->vdrbed || !r->vdrbed1) {
      c->err = "the restart of cohesive sediment needs KB, KBM, the class slots and the bed "
               "arrays SEDB, SEDB1, HBED, HBED1, VDRBED, VDRBED1";
      return EFDC_ERR_ARG;
    }
    for (int s = 0; s < r->nsed; ++s)
      if (!slot_ok(r->sed_slot[s])) {
        c->err = "a sediment class written to the restart has no valid slot";
        return EFDC_ERR_ARG;
      }
  }
  {  // RESTOUT writes each transported class once (distinct MSV indices)
    std::vector<int> used;
    for (int m = 1; m <= 3; ++m)
      if (writes(r, m)) used.push_back(r->slot[m - 1]);
    if (sed)
      for (int s = 0; s < r->nsed; ++s) used.push_back(r->sed_slot[s]);
    std::sort(used.begin(), used.end());
    if (std::adjacent_find(used.begin(), used.end()) != used.end()) {
      c->err = "a constituent slot is written to the restart twice";
      return EFDC_ERR_ARG;
    }
  }
  const bool f32 = c->store == EFDC_STORE_FP32;
  auto dsrc = [&](const void *p, int type) {
    L.src.push_back({p, c->pitch, type, 0});
    return (int)L.src.size() - 1;
  };
  auto hsrc = [&](const double *p, size_t n) {
    L.src.push_back({nullptr, (int64_t)LCM, SRC_F64, 0});
    L.host.push_back({p, n});
    L.host_src.push_back(L.src.size() - 1);
    return (int)L.src.size() - 1;
  };
  auto H = [&](int id) { return dsrc(c->dH[id], SRC_F64); };
  auto U = [&](int id) -> int {
    double *d = nullptr;
    int planes = 0;
    if (uvw_ptr(c, id, &d, &planes)) return -1;
    return dsrc(d, SRC_F64);
  };
  auto Kc = [&](int slot, int level) {
    const ConstDev &k = c->con[slot];
    return dsrc(k.buf[level == 0 ? k.cur : (k.cur ^ 1)], f32 ? SRC_F32 : SRC_F64);
  };
  auto layers = [](int s, int k0, int k1) {  // planes k0..k1 of source s
    Rec r;
    for (int k = k0; k <= k1; ++k) r.v.push_back({s, k});
    return r;
  };
  auto one = [](std::initializer_list<int> ss) {
    Rec r;
    for (int s : ss) r.v.push_back({s, 0});
    return r;
  };
  const size_t W = (size_t)LCM * (KC + 1);
  // --- the cells (EFDC/restartmod.f90:88-154)
  std::vector<Rec> recs;
  recs.push_back(one({H(EFDC_HP), H(EFDC_H1P), hsrc(r->hwq, LCM), hsrc(r->h2wq, LCM),
                      H(EFDC_BELV)}));
  recs.push_back(one({H(EFDC_UHDYE), H(EFDC_UHDY1E), H(EFDC_VHDXE), H(EFDC_VHDX1E)}));
  for (int id : {EFDC_U, EFDC_U1, EFDC_V, EFDC_V1}) {
    const int s = U(id);
    if (s < 0) return EFDC_ERR_NOMEM;
    recs.push_back(layers(s, 0, KC - 1));
  }
  for (const double *p : {r->qq, r->qq1, r->qql, r->qql1, r->dml})
    recs.push_back(layers(hsrc(p, W), 0, KC));
  if (ice) recs.push_back(one({hsrc(r->ice_thick, LCM), hsrc(r->ice_temp, LCM)}));
  if (writes(r, 1)) {
    recs.push_back(layers(Kc(r->slot[0], 0), 0, KC - 1));
    recs.push_back(layers(Kc(r->slot[0], 1), 0, KC - 1));
  }
  if (writes(r, 2)) {
    Rec t = one({hsrc(r->temb, LCM)});
    const Rec tl = layers(Kc(r->slot[1], 0), 0, KC - 1);
    t.v.insert(t.v.end(), tl.v.begin(), tl.v.end());
    recs.push_back(t);
    recs.push_back(layers(Kc(r->slot[1], 1), 0, KC - 1));
  }
  if (writes(r, 3)) {
    recs.push_back(layers(Kc(r->slot[2], 0), 0, KC - 1));
    recs.push_back(layers(Kc(r->slot[2], 1), 0, KC - 1));
  }
  if (sed) {
    const size_t B = (size_t)LCM * r->kbm;
    for (int s = 0; s < r->nsed; ++s) {
      // (LCM, KBM, NSCM): class s starts at s * LCM * KBM
      auto bed = [&](const double *p) {
        const int id = hsrc(p + s * B, (size_t)LCM * r->kb);
        return layers(id, 0, r->kb - 1);
      };
      recs.push_back(bed(r->sedb));
      recs.push_back(layers(Kc(r->sed_slot[s], 0), 0, KC - 1));
      recs.push_back(bed(r->sedb1));
      recs.push_back(layers(Kc(r->sed_slot[s], 1), 0, KC - 1));
    }
    for (const double *p : {r->hbed, r->hbed1, r->vdrbed, r->vdrbed1})
      recs.push_back(layers(hsrc(p, (size_t)LCM * r->kb), 0, r->kb - 1));
  }
  int d0 = (int)L.desc.size();
  int64_t ib = lay_item(L, recs);
  add_seg(L, d0, c->n_wet(), ib, 0, 0, 0, KC, nullptr);
  // --- the open-boundary concentrations (:156-269), sides S, W, E, N
  auto bc_group = [&](int slot) {
    for (int side = 0; side < 4; ++side) {
      const CbcDev &b = c->cbc[side];
      if (b.n == 0) continue;
      const int sn = dsrc(b.nlo, SRC_NLO), sc = dsrc(b.clo, SRC_F64);
      std::vector<Rec> br(2);
      br[0] = layers(sn, 0, KC - 1);
      br[0].integer = true;
      br[1] = layers(sc, 0, KC - 1);
      const int e0 = (int)L.desc.size();
      const int64_t bb = lay_item(L, br);
      add_seg(L, e0, b.n, bb, 1, slot, c->cbc_nc, KC, b.ntscr);
    }
  };
  if (c->cbc[0].n + c->cbc[1].n + c->cbc[2].n + c->cbc[3].n > 0) {
    for (int m = 1; m <= 3; ++m)
      if (writes(r, m)) {
        if (r->slot[m - 1] >= c->cbc_nc) {
          c->err = "the concentration-BC tables were set before all constituents were added";
          return EFDC_ERR_STATE;
        }
        bc_group(r->slot[m - 1]);
      }
    if (sed)
      for (int s = 0; s < r->nsed; ++s) {
        if (r->sed_slot[s] >= c->cbc_nc) {
          c->err = "the concentration-BC tables were set before all constituents were added";
          return EFDC_ERR_STATE;
        }
        bc_group(r->sed_slot[s]);
      }
  }
  // --- the source flows (:270-272)
  std::vector<Rec> qr(1);
  qr[0] = one({H(EFDC_QSUME)});
  const int sq = dsrc(c->dK[EFDC_QSUM - EFDC_HPK], f32 ? SRC_F32 : SRC_F64);
  for (int k = 0; k < KC; ++k) qr[0].v.push_back({sq, k});
  d0 = (int)L.desc.size();
  ib = lay_item(L, qr);
  add_seg(L, d0, c->n_wet(), ib, 0, 0, 0, KC, nullptr);
  return EFDC_OK;
}

// The two header lines: WRITE(99,'(I20,4X,F12.4,2I10)') N,TIMEDAY,VER and
// WRITE(99,'(9I6)') ISCO(0:8)
std::string restart_head(const efdc_restart_t *r) {
  char b[128];
  std::string h;
  std::snprintf(b, sizeof b, "%20lld    ", (long long)r->n);
  h += b;
  std::snprintf(b, sizeof b, "%12.4f", r->timeday);
  h += std::strlen(b) > 12 ? std::string(12, '*') : std::string(b);
  std::snprintf(b, sizeof b, "%10d\n", 720);
  h += b;
  for (int i = 0; i < 9; ++i) {
    std::snprintf(b, sizeof b, "%6d", r->isco[i]);
    h += std::strlen(b) > 6 ? std::string(6, '*') : std::string(b);
  }
  h += "\n";
  return h;
}

int restart_format(efdc_ctx *c, const efdc_restart_t *r, int64_t *bytes, const char **text) {
  IoState *io = io_state(c);
  Layout L;
  int s;
  if ((s = build_restart(c, r, L))) return s;
  const std::string head = restart_head(r);
  const int64_t total = (int64_t)head.size() + L.body;
  *bytes = total;
  if (!text) return EFDC_OK;
  // stage the host arrays
  size_t nstage = 0;
  for (auto &h : L.host) nstage += h.second;
  if ((s = grow(c, (void **)&io->stage, &io->stage_cap, sizeof(double) * std::max<size_t>(nstage, 1))))
    return s;
  size_t at = 0;
  for (size_t i = 0; i < L.host.size(); ++i) {
    HIP_TRY(c, hipMemcpyAsync(io->stage + at, L.host[i].first, sizeof(double) * L.host[i].second,
                              hipMemcpyHostToDevice, c->stream));
    L.src[L.host_src[i]].p = io->stage + at;
    at += L.host[i].second;
  }
  // staged launch: each segment in chunks of cper items of at most
  // STAGE_BYTES; records longer than that go through the direct kernel
  bool staged = true;
  int64_t nblk = 0;
  for (IoSeg &sg : L.seg) {
    if (sg.item_bytes > STAGE_BYTES) staged = false;
    const int64_t cmax = std::max<int64_t>(1, STAGE_BYTES / std::max<int64_t>(sg.item_bytes, 1));
    sg.cper = (int32_t)std::min<int64_t>(cmax, 1024);  // items per workgroup
    sg.blk0 = nblk;
    nblk += (sg.nitems + sg.cper - 1) / sg.cper;
  }
  if (nblk >= ((int64_t)1 << 31)) staged = false;
  std::vector<IoVal> vals(L.desc.size());
  for (size_t i = 0; i < L.desc.size(); ++i) {
    const IoDesc &d = L.desc[i];
    const IoSrc &sr = L.src[d.src];
    vals[i] = IoVal{sr.p, sr.ks, d.k, d.off, d.flags, sr.type};
  }
  const size_t o_desc = (sizeof(IoSeg) * L.seg.size() + 15) & ~(size_t)15;
  const size_t mb = o_desc + sizeof(IoVal) * vals.size() + 64;
  if ((s = grow(c, &io->meta, &io->meta_cap, mb))) return s;
  std::vector<char> meta(mb);
  char *mp = meta.data();
  std::memcpy(mp, L.seg.data(), sizeof(IoSeg) * L.seg.size());
  std::memcpy(mp + o_desc, vals.data(), sizeof(IoVal) * vals.size());
  HIP_TRY(c, hipMemcpyAsync(io->meta, meta.data(), mb, hipMemcpyHostToDevice, c->stream));
  if ((s = grow(c, (void **)&io->dbuf, &io->dbuf_cap, (size_t)std::max<int64_t>(L.body, 1))))
    return s;
  if (!io->nflag && (s = dev_alloc(c, (void **)&io->nflag, sizeof(unsigned)))) return s;
  const unsigned want = (unsigned)std::min<int64_t>(std::max<int64_t>(L.nvals, 1), 1 << 20);
  if (io->flag_cap < want) {
    if (io->flag_off) (void)hipFree(io->flag_off);
    if (io->flag_val) (void)hipFree(io->flag_val);
    io->flag_off = nullptr;
    io->flag_val = nullptr;
    io->flag_cap = 0;
    if ((s = dev_alloc(c, (void **)&io->flag_off, sizeof(int64_t) * want)) ||
        (s = dev_alloc(c, (void **)&io->flag_val, sizeof(double) * want)))
      return s;
    io->flag_cap = want;
  }
  if ((s = grow_host(c, io, (size_t)total))) return s;
  std::memcpy(io->hbuf, head.data(), head.size());
  FmtArgs a{};
  a.seg = (const IoSeg *)io->meta;
  a.nseg = (int)L.seg.size();
  a.desc = (const IoVal *)((char *)io->meta + o_desc);
  a.out = io->dbuf;
  a.nvals = L.nvals;
  a.n = (long long)r->n;
  a.update_nlo = r->keep_nlo ? 0 : 1;
  a.nflag = io->nflag;
  a.flag_off = io->flag_off;
  a.flag_val = io->flag_val;
  a.flag_cap = io->flag_cap;
  unsigned nflag = 0;
  if (L.nvals > 0) {
    // a counter cleared by a kernel of this stream (no memset node)
    k_fill_u32<<<1, 64, 0, c->stream>>>(io->nflag, 1, 0u);
    if (staged)
      TIMED(c, TK_IO_FMT, k_fmt_staged<<<(unsigned)nblk, 256, STAGE_BYTES + 16, c->stream>>>(a));
    else
      TIMED(c, TK_IO_FMT, k_fmt_restart<<<(unsigned)((L.nvals + 255) / 256), 256, 0,
                                          c->stream>>>(a));
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipMemcpyAsync(io->hbuf + head.size(), io->dbuf, (size_t)L.body,
                              hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipMemcpyAsync(&nflag, io->nflag, sizeof(unsigned), hipMemcpyDeviceToHost,
                              c->stream));
  }
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (nflag > io->flag_cap) {
    c->err = "more than 2^20 values of the restart need host formatting";
    return EFDC_ERR_STATE;
  }
  if (nflag > 0) {
    std::vector<int64_t> fo(nflag);
    std::vector<double> fv(nflag);
    HIP_TRY(c, hipMemcpy(fo.data(), io->flag_off, sizeof(int64_t) * nflag, hipMemcpyDeviceToHost));
    HIP_TRY(c, hipMemcpy(fv.data(), io->flag_val, sizeof(double) * nflag, hipMemcpyDeviceToHost));
    for (unsigned i = 0; i < nflag; ++i) host_e17(fv[i], io->hbuf + head.size() + fo[i]);
  }
  *text = io->hbuf;
  return EFDC_OK;
}

int write_all(efdc_ctx *c, const char *path, const char *mode,
              std::initializer_list<std::pair<const void *, size_t>> parts) {
  FILE *f = std::fopen(path, mode);
  if (!f) {
    c->err = std::string("cannot open ") + path;
    return EFDC_ERR_IO;
  }
  for (auto &p : parts)
    if (p.second && std::fwrite(p.first, 1, p.second, f) != p.second) {
      std::fclose(f);
      c->err = std::string("cannot write ") + path;
      return EFDC_ERR_IO;
    }
  if (std::fclose(f) != 0) {
    c->err = std::string("cannot write ") + path;
    return EFDC_ERR_IO;
  }
  return EFDC_OK;
}

std::string join(const char *dir, const char *name) {
  std::string p = dir ? dir : "";
  if (!p.empty() && p.back() != '/') p += '/';
  return p + name;
}

}  // namespace

void io_free(efdc_ctx *c) {
  IoState *io = c->io;
  if (!io) return;
  for (void *p : {(void *)io->dbuf, (void *)io->stage, io->meta, (void *)io->nflag,
                  (void *)io->flag_off, (void *)io->flag_val, (void *)io->ksz, (void *)io->off,
                  (void *)io->pack})
    if (p) (void)hipFree(p);
  if (io->hbuf) (void)hipHostFree(io->hbuf);
  delete io;
  c->io = nullptr;
}

extern "C" {

int efdc_restart_text(efdc_ctx_t *c, const efdc_restart_t *r, char *buf, int64_t cap,
                      int64_t *bytes) {
  if (!c || !r || !bytes) return EFDC_ERR_ARG;
  if (!buf) return restart_format(c, r, bytes, nullptr);
  const char *text = nullptr;
  int s = restart_format(c, r, bytes, &text);
  if (s) return s;
  if (*bytes > cap) {
    c->err = "the buffer is smaller than the restart text";
    return EFDC_ERR_ARG;
  }
  std::memcpy(buf, text, (size_t)*bytes);
  return EFDC_OK;
}

int efdc_restart_write(efdc_ctx_t *c, const efdc_restart_t *r, const char *path) {
  if (!c || !r || !path) return EFDC_ERR_ARG;
  const char *text = nullptr;
  int64_t bytes = 0;
  int s = restart_format(c, r, &bytes, &text);
  if (s) return s;
  return write_all(c, path, "wb", {{text, (size_t)bytes}});
}

int efdc_ee_open(efdc_ctx_t *c, const char *dir, const efdc_ee_grid_t *g) {
  if (!c || !g || !g->rssbce || !g->rssbcw || !g->rssbcs || !g->rssbcn) return EFDC_ERR_ARG;
  IoState *io = io_state(c);
  const int LA = c->LA, KC = c->KC, LCM = c->LCM, nwet = c->n_wet();
  // CELL3D and each cell's first value (EFDC/efdcout.f90:79-82)
  std::vector<int64_t> off(nwet);
  int64_t cell3d = 0;
  for (int L = 2; L <= LA; ++L) {
    const int ksz = g->ksz ? g->ksz[L - 1] : 1;
    if (ksz < 1 || ksz > KC) {
      c->err = "KSZ out of 1..KC";
      return EFDC_ERR_ARG;
    }
    off[L - 2] = cell3d;
    cell3d += KC - ksz + 1;
  }
  if (io->ksz) (void)hipFree(io->ksz);
  if (io->off) (void)hipFree(io->off);
  io->ksz = nullptr;
  io->off = nullptr;
  int s;
  if (g->ksz) {
    if ((s = dev_alloc(c, (void **)&io->ksz, sizeof(int) * LCM)) ||
        (s = dev_alloc(c, (void **)&io->off, sizeof(int64_t) * std::max(nwet, 1))))
      return s;
    HIP_TRY(c, hipMemcpy(io->ksz, g->ksz, sizeof(int) * LCM, hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(io->off, off.data(), sizeof(int64_t) * nwet, hipMemcpyHostToDevice));
  }
  io->ic = g->ic;
  io->jc = g->jc;
  io->igridv = g->igridv;
  io->cell3d = cell3d;
  const int32_t nact = LA - 1;
  // EE_WS.OUT: VER, HSIZE, BSIZE, IC, JC, NACTIVE (:85-96)
  const int32_t ws[6] = {8400, 6 * 4, (2 + 3 + nact) * 4, g->ic, g->jc, nact};
  if ((s = write_all(c, join(dir, "EE_WS.OUT").c_str(), "wb", {{ws, sizeof ws}}))) return s;
  // EE_VEL.OUT: VER, HSIZE, BSIZE, IGRIDV, CELL3D, IC, JC, KC, NACTIVE and
  // RSSBCE, RSSBCW, RSSBCS, RSSBCN as REAL*4 over L = 1..LCM (:98-113)
  const int32_t vh[9] = {8400, (9 + 4 * LCM) * 4, (int32_t)((2 + 2 + 3 * cell3d) * 4), g->igridv,
                         (int32_t)cell3d, g->ic, g->jc, KC, nact};
  std::vector<float> rs((size_t)4 * LCM);
  const double *arr[4] = {g->rssbce, g->rssbcw, g->rssbcs, g->rssbcn};
  for (int q = 0; q < 4; ++q)
    for (int L = 0; L < LCM; ++L) rs[(size_t)q * LCM + L] = (float)arr[q][L];
  if ((s = write_all(c, join(dir, "EE_VEL.OUT").c_str(), "wb",
                     {{vh, sizeof vh}, {rs.data(), sizeof(float) * rs.size()}})))
    return s;
  io->ee_open = true;
  return EFDC_OK;
}

int efdc_ee_append(efdc_ctx_t *c, const char *dir, int32_t n, double eetime, double delt,
                   int32_t lminstep) {
  if (!c) return EFDC_ERR_ARG;
  IoState *io = c->io;
  if (!io || !io->ee_open) {
    c->err = "efdc_ee_open has not written the EE headers";
    return EFDC_ERR_STATE;
  }
  const int nwet = c->n_wet(), KC = c->KC;
  const size_t nfl = (size_t)nwet + 3 * (size_t)io->cell3d;
  int s;
  if ((s = grow(c, (void **)&io->pack, &io->pack_cap, sizeof(float) * std::max<size_t>(nfl, 1))))
    return s;
  double *u = nullptr, *v = nullptr, *w = nullptr;
  int planes = 0;
  if ((s = uvw_ptr(c, EFDC_U, &u, &planes)) || (s = uvw_ptr(c, EFDC_V, &v, &planes)) ||
      (s = uvw_ptr(c, EFDC_W, &w, &planes)))
    return s;
  if ((s = grow_host(c, io, sizeof(float) * nfl))) return s;
  const int64_t nt = (int64_t)nwet * KC;
  if (nt > 0)
    TIMED(c, TK_IO_PACK, k_ee_pack<<<(unsigned)((nt + 255) / 256), 256, 0, c->stream>>>(
        c->dH[EFDC_HP], u, v, w, io->ksz, io->off, KC, c->pitch, nwet, io->cell3d, io->pack));
  HIP_TRY(c, hipGetLastError());
  HIP_TRY(c, hipMemcpyAsync(io->hbuf, io->pack, sizeof(float) * nfl, hipMemcpyDeviceToHost,
                            c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  // WSOUT (:418-426): N, EETIME, DELT, LMINSTEP, then HP
  const float d4 = (float)delt;
  if ((s = write_all(c, join(dir, "EE_WS.OUT").c_str(), "ab",
                     {{&n, 4}, {&eetime, 8}, {&d4, 4}, {&lminstep, 4},
                      {io->hbuf, sizeof(float) * nwet}})))
    return s;
  // VELOUT (:487-493): N, EETIME, DELT, then U, V, W over the active layers
  return write_all(c, join(dir, "EE_VEL.OUT").c_str(), "ab",
                   {{&n, 4}, {&eetime, 8}, {&d4, 4},
                    {io->hbuf + sizeof(float) * nwet, sizeof(float) * 3 * io->cell3d}});
}

}  // extern "C"
