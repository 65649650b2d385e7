// efdc_seg.hip -- ensembles of independent basins in one context.
//
// The host concatenates the basins' L ranges (each basin keeps its own
// neighbour maps, land points at the common LC) and declares the ranges
// with efdc_set_segments.  Pointwise kernels then run once over all cells;
// the per-model reductions of the hot path -- the CCMNM scaling minimum,
// the CG dot products with their stop test, the balances -- run per basin
// (efdc_puv.hip: k_seg_min, k_scg_*; here: k_seg_balance).  The CG tiles of
// a basin start at the basin's first cell, so a basin's partial sums and
// their reduction order are those of the same basin alone.
#include <algorithm>
#include <cstring>
#include <vector>

#include "efdc_internal.h"

namespace {


// FP64 volume and mass of one basin: thread t sums the basin's cells
// t, t+B, ... in order, then a fixed tree.  mass: con == nullptr -> volume.
template <typename T>
__global__ __launch_bounds__(EFDC_BLOCK) void k_seg_balance(
    const double *dxyp, const double *hp, const T *con, const T *dzc, int64_t pitch, int KC,
    const int *segfirst, double *out, int nseg, int slot, int nslot) {
  const int sgi = blockIdx.x;
  const int i0 = segfirst[sgi], i1 = segfirst[sgi + 1];
  double v[1] = {0.0};
  for (int i = i0 + (int)threadIdx.x; i < i1; i += blockDim.x) {
    const double base = dxyp[i] * hp[i];
    if (!con) {
      v[0] += base;
    } else {
      double m = 0.0;
      for (int k = 0; k < KC; ++k)
        m += base * (double)con[i + k * pitch] * (double)dzc[i + k * pitch];
      v[0] += m;
    }
  }
  block_sum<1>(v);
  if (threadIdx.x == 0) out[(size_t)sgi * nslot + slot] = v[0];
}

}  // namespace

void seg_select_tiles(efdc_ctx *c, int q) {
  c->seg_tset = q;
  c->d_tiles = c->d_tiles_s[q];
  c->d_segtile0 = c->d_segtile0_s[q];
  c->ntiles = c->ntiles_s[q];
}

void seg_free(efdc_ctx *c) {
  auto f = [](auto *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
  };
  for (int q = 0; q < 4; ++q) {
    f(c->d_tiles_s[q]);
    f(c->d_segtile0_s[q]);
    c->ntiles_s[q] = 0;
  }
  c->d_tiles = nullptr;
  c->d_segtile0 = nullptr;
  f(c->d_segid); f(c->d_segfirst); f(c->d_segst);
  f(c->d_segtk); f(c->d_segdone); f(c->d_segmin); f(c->d_segpart); f(c->d_segbal);
  f(c->d_segfin);
  if (c->h_segst) (void)hipHostFree(c->h_segst);
  c->h_segst = nullptr;
  c->nseg = 1;
  c->seg_first.clear();
  c->ntiles = 0;
}

extern "C" {

int efdc_set_segments(efdc_ctx_t *c, int32_t nseg, const int32_t *first_L) {
  if (!c || nseg < 1 || (nseg > 1 && !first_L)) return EFDC_ERR_ARG;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  seg_free(c);
  if (nseg == 1) return EFDC_OK;
  const int LA = c->LA, nwet = c->n_wet(), LCM = c->LCM;
  // validate the ranges (Fortran L) and convert to 0-based cell bounds
  std::vector<int> first(nseg + 1);
  for (int s = 0; s < nseg; ++s) {
    const int L = first_L[s];
    if ((s == 0 && L != 2) || L < 2 || L > LA || (s > 0 && L <= first_L[s - 1])) {
      c->err = "efdc_set_segments: first_L must start at 2 and increase within 2..LA";
      return EFDC_ERR_ARG;
    }
    first[s] = L - 1;
  }
  first[nseg] = nwet + 1;
  // no neighbour map may leave its basin
  std::vector<int> segid(LCM, 0), lw(LCM), le(LCM), ls(LCM), ln(LCM);
  for (int s = 0; s < nseg; ++s)
    for (int i = first[s]; i < first[s + 1]; ++i) segid[i] = s;
  HIP_TRY(c, hipMemcpy(lw.data(), c->dLW, sizeof(int) * LCM, hipMemcpyDeviceToHost));
  HIP_TRY(c, hipMemcpy(le.data(), c->dLE, sizeof(int) * LCM, hipMemcpyDeviceToHost));
  HIP_TRY(c, hipMemcpy(ls.data(), c->dLS, sizeof(int) * LCM, hipMemcpyDeviceToHost));
  HIP_TRY(c, hipMemcpy(ln.data(), c->dLN, sizeof(int) * LCM, hipMemcpyDeviceToHost));
  const int lc0 = c->LC - 1;
  for (int i = 1; i <= nwet; ++i)
    for (int j : {lw[i], le[i], ls[i], ln[i]})
      if (j != lc0 && (j < 1 || j > nwet || segid[j] != segid[i])) {
        c->err = "efdc_set_segments: a neighbour map crosses basins at L=" +
                 std::to_string(i + 1);
        return EFDC_ERR_ARG;
      }
  c->nseg = nseg;
  c->seg_first = first;
  int st = EFDC_OK;
  auto up = [&](auto **dst, const auto *src, size_t n) -> int {
    if ((st = dev_alloc(c, (void **)dst, sizeof(*src) * n))) return st;
    // on the context's (non-blocking) stream, after dev_alloc's zero fill
    HIP_TRY(c, hipMemcpyAsync(*dst, src, sizeof(*src) * n, hipMemcpyHostToDevice, c->stream));
    return EFDC_OK;
  };
  if (up(&c->d_segid, segid.data(), LCM) || up(&c->d_segfirst, first.data(), nseg + 1)) {
    seg_free(c);
    return st ? st : EFDC_ERR_HIP;
  }
  // CG tiles of 256 << q cells counted from each basin's first cell
  int maxtiles = 0;
  for (int q = 0; q < 4; ++q) {
    const int tile = EFDC_BLOCK << q;
    std::vector<int4> tiles;
    std::vector<int> tile0(nseg + 1);
    for (int s = 0; s < nseg; ++s) {
      tile0[s] = (int)tiles.size();
      int t = 0;
      for (int i0 = first[s]; i0 < first[s + 1]; i0 += tile, ++t)
        tiles.push_back(make_int4(i0, std::min(i0 + tile, first[s + 1]), s, t));
    }
    tile0[nseg] = (int)tiles.size();
    c->ntiles_s[q] = (int)tiles.size();
    maxtiles = std::max(maxtiles, c->ntiles_s[q]);
    if (up(&c->d_tiles_s[q], tiles.data(), tiles.size()) ||
        up(&c->d_segtile0_s[q], tile0.data(), nseg + 1)) {
      seg_free(c);
      return st ? st : EFDC_ERR_HIP;
    }
  }
  seg_select_tiles(c, c->seg_tset);
  if ((st = dev_alloc(c, (void **)&c->d_segst, sizeof(CgState) * nseg)) ||
      (st = dev_alloc(c, (void **)&c->d_segtk, sizeof(unsigned) * TICKET_PAD * nseg)) ||
      (st = dev_alloc(c, (void **)&c->d_segdone, sizeof(unsigned))) ||
      (st = dev_alloc(c, (void **)&c->d_segfin, sizeof(unsigned))) ||
      (st = dev_alloc(c, (void **)&c->d_segmin, sizeof(double) * nseg)) ||
      (st = dev_alloc(c, (void **)&c->d_segpart, sizeof(double) * 2 * maxtiles)) ||
      (st = dev_alloc(c, (void **)&c->d_segbal, sizeof(double) * nseg * (EFDC_MAXC + 1)))) {
    seg_free(c);
    return st;
  }
  if (hipHostMalloc((void **)&c->h_segst, sizeof(CgState) * nseg) != hipSuccess) {
    seg_free(c);
    c->err = "hipHostMalloc failed";
    return EFDC_ERR_NOMEM;
  }
  std::memset(c->h_segst, 0, sizeof(CgState) * nseg);
  c->last_iters = 0;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return EFDC_OK;
}

int efdc_set_segment_tile(efdc_ctx_t *c, int32_t cells) {
  if (!c) return EFDC_ERR_ARG;
  if (cells == 0) {
    c->seg_tune_pending = true;
    return EFDC_OK;
  }
  for (int q = 0; q < 4; ++q)
    if (cells == (EFDC_BLOCK << q)) {
      c->seg_tune_pending = false;
      if (c->nseg > 1) seg_select_tiles(c, q);
      else c->seg_tset = q;
      return EFDC_OK;
    }
  c->err = "efdc_set_segment_tile: cells must be 0, 256, 512, 1024 or 2048";
  return EFDC_ERR_ARG;
}

int efdc_get_segment_tile(efdc_ctx_t *c, int32_t *cells) {
  if (!c || !cells) return EFDC_ERR_ARG;
  *cells = EFDC_BLOCK << c->seg_tset;
  return EFDC_OK;
}

int efdc_segment_stats(efdc_ctx_t *c, int32_t cap, int32_t *iters, int32_t *status) {
  if (!c || cap < 0 || (cap > 0 && (!iters || !status))) return EFDC_ERR_ARG;
  if (c->nseg < 2) {
    c->err = "no basins declared (efdc_set_segments)";
    return EFDC_ERR_STATE;
  }
  for (int s = 0; s < std::min<int>(cap, c->nseg); ++s) {
    iters[s] = c->h_segst[s].iter;
    status[s] = c->h_segst[s].status ? EFDC_ERR_CG_MAXITER : EFDC_OK;
  }
  return EFDC_OK;
}

int efdc_balance_segments(efdc_ctx_t *c, int32_t cap, double *volume, double *mass) {
  if (!c || cap < 0 || (cap > 0 && (!volume || (c->nconst > 0 && !mass)))) return EFDC_ERR_ARG;
  if (c->nseg < 2) {
    c->err = "no basins declared (efdc_set_segments)";
    return EFDC_ERR_STATE;
  }
  const int nseg = c->nseg, nslot = c->nconst + 1;
  for (int j = -1; j < c->nconst; ++j) {
    if (c->store == EFDC_STORE_FP64)
      k_seg_balance<double><<<nseg, EFDC_BLOCK, 0, c->stream>>>(
          c->dH[EFDC_DXYP], c->dH[EFDC_HP],
          j < 0 ? nullptr : (const double *)c->con[j].buf[c->con[j].cur],
          (const double *)c->dK[EFDC_DZC - EFDC_HPK], c->pitch, c->KC, c->d_segfirst,
          c->d_segbal, nseg, j + 1, nslot);
    else
      k_seg_balance<float><<<nseg, EFDC_BLOCK, 0, c->stream>>>(
          c->dH[EFDC_DXYP], c->dH[EFDC_HP],
          j < 0 ? nullptr : (const float *)c->con[j].buf[c->con[j].cur],
          (const float *)c->dK[EFDC_DZC - EFDC_HPK], c->pitch, c->KC, c->d_segfirst,
          c->d_segbal, nseg, j + 1, nslot);
  }
  HIP_TRY(c, hipGetLastError());
  std::vector<double> h((size_t)nseg * nslot);
  HIP_TRY(c, hipMemcpyAsync(h.data(), c->d_segbal, sizeof(double) * h.size(),
                            hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  for (int s = 0; s < std::min<int>(cap, nseg); ++s) {
    volume[s] = h[(size_t)s * nslot];
    for (int j = 0; j < c->nconst; ++j) mass[(size_t)s * c->nconst + j] = h[(size_t)s * nslot + 1 + j];
  }
  return EFDC_OK;
}

}  // extern "C"
