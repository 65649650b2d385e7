// efdc_cgp.hip -- CONGRAD as one persistent cooperative launch per solve.
//
// The multi-kernel solve (efdc_puv.hip) pays two launches and two grid
// reductions per iteration; at the sizes of a real lake (10^4 .. 10^6 wet
// cells, a working set that lives in the 256 MB MALL) that is latency, not
// bandwidth.  Here every workgroup owns one contiguous range of wet cells
// (L order) and keeps the rows of its cells -- coefficients, neighbour
// indices and all CG vectors -- in registers for the whole solve.  Each
// exchange between iterations moves only
//   * the workgroup's own vector values, through an LDS window that also
//     holds the cells just outside the range that its rows reference, and
//   * the values that the neighbouring workgroups need (cells with a
//     neighbour outside the range) and the dot-product partials, as
//     tagged words in memory.
//
// Exchange protocol.  A workgroup writes the values its neighbours need
// (halo values, into a slot of the exchange's parity) and its dot-product
// partials with agent-scope stores, which are written through to the
// coherent memory side; every wave waits for its stores to complete
// (s_waitcnt), and then thread 0 stores the exchange number into the
// workgroup's flag word (one 128-byte line per workgroup).  A reader polls
// the flag words of all workgroups (thread q reads workgroup q's flag)
// until every one carries the exchange number, and then fetches the
// partials and the halo values it needs in a single round of loads.  Two
// slot parities suffice: a workgroup can only write exchange T+2 after it
// has seen every flag of T+1, and every workgroup raises T+1 only after it
// has consumed T.  Exchange numbers grow across launches (the host numbers
// launches), so flags never go backwards; when the numbering wraps the host
// clears them.
//
// FP64 (the reference's Jacobi-PCG, EFDC/congrad.f90:52-187): the same
// recurrences and stop test as the reference, two exchanges per iteration
// (alpha needs p.Ap, beta needs r.z).
//
// Mixed: the inner FP32 CG on the Jacobi-scaled unit-diagonal operator is
// written in the Chronopoulos-Gear form (one exchange per iteration
// carries (r,r), (Ar,r) and the halo triples r, w = Ar, s = Ap); FP64 dot
// products, FP64 alpha/beta, FP64 residual refinement of P inside the
// same launch.
//
// Every wait is bounded: a workgroup that waits too long raises the abort
// flag and leaves, the others see the flag and leave too, and the host
// reports the failure.  All workgroups compute the same sums from the
// same partials in the same order, so they take the same branches.
#include <algorithm>
#include <cstdio>

#include "efdc_internal.h"

namespace {

constexpr int PB = 512;           // threads per workgroup
constexpr int PW = PB / 64;       // waves per workgroup
constexpr unsigned SPIN_MAX = 1u << 21;
// An inner (FP32) solve stops at the smaller of two reductions of its
// residual: the set one (rtol), or what the outer FP64 test still needs,
// with a margin of 10 in the norm (squared here), so that a round that only
// has to finish the solve is not carried past the target.  The outer test
// is always made on the FP64 residual.
constexpr double INNER_MARGIN2 = 1e-2;
constexpr int WIN = 2048;         // LDS window reach beyond the own range
constexpr int HMAX = 2 * WIN;     // outside cells of one workgroup, at most
constexpr int NVMAX = 3;          // partials per exchange, at most
constexpr int HX_WORDS = 4;       // 4-byte words per cell and parity in the halo slots

// Halo slots are planes (one per value and parity), so that the stores of
// a wave and the loads of a run of outside cells are contiguous: FP32
// value q of cell i in parity p at word (p*HX_WORDS + q)*ncell + i, FP64
// value q (two words) at double (p*2 + q)*ncell + i.

typedef unsigned long long u64;

struct Bar {
  unsigned *ctr;                  // 9 counters, TICKET_PAD apart
  int *abort;
  unsigned spin_max;              // polling rounds before a wait gives up
};

// An index the compiler must treat as unknown at this point: keeps it from
// hoisting the 64-bit addresses of every neighbour access out of the
// iteration loop (which would pin two VGPRs per address for the whole
// solve and spill); the address is recomputed where it is used.
__device__ __forceinline__ int opaque(int v) {
  __asm__ volatile("" : "+v"(v));
  return v;
}

__device__ __forceinline__ void put_u32(unsigned *p, unsigned v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned get_u32(const unsigned *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void put_f64(unsigned *p, double v) {
  const u64 b = (u64)__double_as_longlong(v);
  __hip_atomic_store((u64 *)p, b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u64 get_u64(const unsigned *p) {
  return __hip_atomic_load((const u64 *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Grid barrier number `gen` (1, 2, ...) with counters, for the few places
// where a plain FP64 array (P) must become visible to every workgroup.
// Returns false if the barrier was abandoned.
__device__ bool grid_barrier(const Bar &b, unsigned gen) {
  __shared__ int s_ok;
  __builtin_amdgcn_s_waitcnt(0);  // this wave's stores have completed
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nb = (int)gridDim.x;
    const int x = blockIdx.x & 7;
    const unsigned nx = (unsigned)((nb >> 3) + (x < (nb & 7) ? 1 : 0));
    const unsigned ng = nb < 8 ? (unsigned)nb : 8u;
    unsigned *gx = b.ctr + x * TICKET_PAD, *gg = b.ctr + 8 * TICKET_PAD;
    if (atomic_inc(gx) == gen * nx - 1u) atomic_inc(gg);
    const unsigned target = gen * ng;
    int ok = 1;
    unsigned spins = 0;
    while (__hip_atomic_load(gg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (++spins > b.spin_max ||
          ((spins & 255u) == 0u &&
           __hip_atomic_load(b.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
        __hip_atomic_store(b.abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
    s_ok = ok;
  }
  __syncthreads();
  return s_ok != 0;
}

// Partials: [parity][NVMAX][PB workgroups] doubles.
__device__ __forceinline__ double *pslot(double *ps, unsigned tag, int j) {
  return ps + ((size_t)(tag & 1u) * NVMAX + j) * PB;
}

struct Xch {
  double *ps;          // partials
  unsigned *hx;        // halo slots, [2][HX_WORDS][ncell] words
  unsigned *flag;      // one word per workgroup, TICKET_PAD apart
  int ncell;
};

// Block sum of NV partials into this workgroup's slots of exchange `tag`
// (stored by thread 0, not yet announced).
template <int NV>
__device__ void post_partials(double (&v)[NV], const Xch &x, unsigned tag) {
  __shared__ double sh[NV][PW];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < NV; ++j) v[j] = wave_sum_dpp(v[j]);
  if (lane == 0) {
#pragma unroll
    for (int j = 0; j < NV; ++j) sh[j][w] = v[j];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    // all the loads first, so that they are in flight together; the sums
    // keep the wave order
    double pv[NV][PW];
#pragma unroll
    for (int j = 0; j < NV; ++j)
#pragma unroll
      for (int q = 0; q < PW; ++q) pv[j][q] = sh[j][q];
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      double s = 0.0;
#pragma unroll
      for (int q = 0; q < PW; ++q) s += pv[j][q];
      __hip_atomic_store(&pslot(x.ps, tag, j)[blockIdx.x], s, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// The flag of exchange `tag`: every wave's earlier stores (partials, halo
// values) have completed before thread 0 raises it.
__device__ void post_flag(const Xch &x, unsigned tag) {
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) put_u32(&x.flag[blockIdx.x * TICKET_PAD], tag);
}

template <int NV>
__device__ void post(double (&v)[NV], const Xch &x, unsigned tag) {
  post_partials<NV>(v, x, tag);
  post_flag(x, tag);
}

// Wait for exchange `tag`, then fetch the NV partials of every workgroup
// and NR words of each outside cell (hl[0..nhalo)) into raw[q*HMAX + h].
// Wave 0 alone polls the flags (lane l reads the flags of workgroups l,
// l+64, ...; a wave vote ends the loop), so the waiting costs no workgroup
// barrier per round; after one barrier every wave fetches the halo words
// while wave 0 adds the partials (lane l adds those of workgroups l, l+64,
// ... in that order, then a DPP wave sum).  Returns the sums (same bits in
// every workgroup), or false if the wait was abandoned.
template <int NV, int NR>
__device__ bool exchange(const Xch &x, unsigned tag, const int *hl, int nhalo,
                         unsigned *raw, const Bar &b, double (&sum)[NV]) {
  __shared__ double shs[NV];
  __shared__ int s_ok;
  const int t = threadIdx.x, lane = t & 63;
  const int nb = (int)gridDim.x;
  if (t < 64) {
    unsigned spins = 0;
    int ok = 1;
    for (;;) {
      // every lane loads (past the last workgroup, that one's flag again),
      // so that the loads of a round are in flight together: loads under a
      // condition would each wait for its own result
      // (the lane number made opaque: the flag addresses are formed here,
      // not held in registers across the caller's loops)
      const int ln = opaque(lane);
      unsigned f[PB / 64];
#pragma unroll
      for (int u = 0; u < PB / 64; ++u)
        f[u] = get_u32(&x.flag[min(ln + 64 * u, nb - 1) * TICKET_PAD]);
      bool ready = true;
#pragma unroll
      for (int u = 0; u < PB / 64; ++u) ready = ready && (ln + 64 * u >= nb || f[u] >= tag);
      if (__all(ready)) break;
      ++spins;
      const int ab = (spins & 255u) == 0u
                         ? __builtin_amdgcn_readfirstlane(
                               __hip_atomic_load(b.abort, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT))
                         : 0;
      if (spins > b.spin_max || ab) {
        if (lane == 0)
          __hip_atomic_store(b.abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    if (lane == 0) s_ok = ok;
  }
  __syncthreads();
  if (!s_ok) return false;
  if (t < 64) {
    double v[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) v[j] = 0.0;
    // four workgroups' partials per lane in flight at a time
    for (int u0 = 0; 64 * u0 < nb; u0 += 4) {
      double a[4][NV];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int q = min(opaque(lane) + 64 * (u0 + u), nb - 1);  // every lane loads
#pragma unroll
        for (int j = 0; j < NV; ++j)
          a[u][j] = __hip_atomic_load(&pslot(x.ps, tag, j)[q], __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int j = 0; j < NV; ++j) v[j] += lane + 64 * (u0 + u) < nb ? a[u][j] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < NV; ++j) v[j] = wave_sum_dpp(v[j]);
    if (lane == 0) {
#pragma unroll
      for (int j = 0; j < NV; ++j) shs[j] = v[j];
    }
  }
  if (NR > 0) {
    // all loads of a batch of HB outside cells are in flight before the
    // first result is stored (a store per load would serialise the round
    // trips)
    constexpr int HB = 4;
    const size_t par = tag & 1u, nce = (size_t)x.ncell;
    const unsigned *w32 = x.hx + par * HX_WORDS * nce;        // FP32 planes
    const unsigned *w64 = x.hx + par * 2 * 2 * nce;           // FP64 planes
    for (int h0 = t; h0 < nhalo; h0 += HB * PB) {
      u64 wv[HB][2];
      unsigned wz[HB];
#pragma unroll
      for (int u = 0; u < HB; ++u) {
        // every lane loads (past the list's end, its last cell again)
        const size_t j = (size_t)hl[min(h0 + u * PB, nhalo - 1)];
        wv[u][0] = wv[u][1] = 0;
        wz[u] = 0;
        if (NR == 4) {
          wv[u][0] = get_u64(w64 + 2 * j);
          wv[u][1] = get_u64(w64 + 2 * (nce + j));
        } else {
          wv[u][0] = get_u32(w32 + j);
          if (NR >= 2) wv[u][1] = get_u32(w32 + nce + j);
          if (NR >= 3) wz[u] = get_u32(w32 + 2 * nce + j);
        }
      }
#pragma unroll
      for (int u = 0; u < HB; ++u) {
        const int h = h0 + u * PB;
        if (h < nhalo) {
          if (NR == 4) {
            raw[h] = (unsigned)wv[u][0];
            raw[HMAX + h] = (unsigned)(wv[u][0] >> 32);
            raw[2 * HMAX + h] = (unsigned)wv[u][1];
            raw[3 * HMAX + h] = (unsigned)(wv[u][1] >> 32);
          } else {
            raw[h] = (unsigned)wv[u][0];
            if (NR >= 2) raw[HMAX + h] = (unsigned)wv[u][1];
            if (NR >= 3) raw[2 * HMAX + h] = wz[u];
          }
        }
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < NV; ++j) sum[j] = shs[j];
  __syncthreads();
  return true;
}

// neighbour slots of a row
enum { NS = 0, NN = 1, NW = 2, NE = 3 };


// Compact list of the outside cells a workgroup's rows reference (sorted
// by index).  `mark` holds 2*WIN flags on entry (slot g: cell b0-WIN+g for
// g < WIN, cell b1+g-WIN above) and the list on return; returns its length.
__device__ int compact_halo(int *mark, int b0, int b1) {
  __shared__ int wtot[PW];
  __shared__ int total;
  constexpr int PER = 2 * WIN / PB;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  int f[PER];
  int cnt = 0;
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    f[u] = mark[t * PER + u];
    cnt += f[u] ? 1 : 0;
  }
  int inc = cnt;  // inclusive scan over the workgroup
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int v = __shfl_up(inc, off, 64);
    if (lane >= off) inc += v;
  }
  if (lane == 63) wtot[w] = inc;
  __syncthreads();
  int base = 0;
  for (int q = 0; q < w; ++q) base += wtot[q];
  if (t == PB - 1) total = base + inc;
  int pos = base + inc - cnt;
  __syncthreads();   // every flag has been read
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    if (f[u]) {
      const int g = t * PER + u;
      mark[pos++] = g < WIN ? b0 - WIN + g : b1 + (g - WIN);
    }
  }
  __syncthreads();
  return total;
}

__device__ __forceinline__ int ghost_of(int j, int b0, int b1) {
  return j < b0 ? WIN - (b0 - j) : WIN + (j - b1);
}

// ---------------------------------------------------------------------
// FP64 Jacobi-PCG (EFDC/congrad.f90)
// ---------------------------------------------------------------------
struct P64Args {
  const double *ccc, *ccs, *ccw, *ccn, *cce, *ccci, *fptmp;
  double *x;                 // P, in/out
  const int *lw, *le, *ls, *ln;
  int nwet, lc0, rowmajor, mutual;
  Xch xc;                    // exchange slots; halo words: r (2), p (2)
  unsigned tag0;
  Bar bar;
  CgState *st;
  double rsqm, rel_tol;
  int iterm;
};

template <int CPT>
__global__ __launch_bounds__(PB) void k_pcg64(P64Args a) {
  // the window, then one word that stays zero (read in place of a
  // neighbour that is not a cell, so that every lane loads)
  __shared__ double lds[2 * WIN + CPT * PB + 1];
  __shared__ int hl[2 * WIN];
  __shared__ unsigned raw[4 * HMAX];
  constexpr int ZS = 2 * WIN + CPT * PB;
  const int b0 = 1 + (int)blockIdx.x * CPT * PB;
  const int b1 = min(b0 + CPT * PB, a.nwet + 1);
  const bool fast = a.rowmajor && (a.st->symok == 0);
  const int lc0 = a.lc0;

  double ccc[CPT], cs[CPT], cn[CPT], cw[CPT], ce[CPT], cci[CPT];
  double x[CPT], r[CPT], p[CPT], ap[CPT];
  int nb[CPT][4];
  unsigned expo = 0u;        // bit k: another workgroup reads cell k

  auto inside = [&](int j) { return j >= b0 && j < b1; };
  auto real = [&](int j) { return j > 0 && j != lc0; };
  auto slot = [&](int j) { return j - b0 + WIN; };
  auto publish_rp = [&](int i, unsigned tag, double rv, double pv) {
    const size_t nce = (size_t)a.xc.ncell, j = (size_t)opaque(i);
    unsigned *q = a.xc.hx + (size_t)(tag & 1u) * 2 * 2 * nce;
    put_f64(q + 2 * j, rv);
    put_f64(q + 2 * (nce + j), pv);
  };

  for (int g = threadIdx.x; g < 2 * WIN; g += PB) hl[g] = 0;
  if (threadIdx.x == 0) lds[ZS] = 0.0;
  __syncthreads();
  double v2[2] = {0.0, 0.0};
#pragma unroll
  for (int k = 0; k < CPT; ++k) {
    const int i = b0 + k * PB + (int)threadIdx.x;
    x[k] = r[k] = p[k] = ap[k] = 0.0;
    ccc[k] = 1.0; cs[k] = cn[k] = cw[k] = ce[k] = cci[k] = 0.0;
    nb[k][NS] = nb[k][NN] = nb[k][NW] = nb[k][NE] = lc0;
    if (i < b1) {
      // the row as in the multi-kernel solve (row_of)
      const int js = a.ls[i], jn = a.ln[i];
      int jw, je;
      double s_ = a.ccs[i], w_ = a.ccw[i], n_, e_;
      if (fast) {
        jw = i - 1;
        je = i + 1;
        n_ = (jn != lc0) ? a.ccs[jn] : 0.0;
        e_ = (je != lc0) ? a.ccw[je] : 0.0;
      } else {
        jw = a.lw[i];
        je = a.le[i];
        n_ = (jn != lc0) ? a.ccn[i] : 0.0;
        e_ = (je != lc0) ? a.cce[i] : 0.0;
        if (jw == lc0) w_ = 0.0;
      }
      if (js == lc0) s_ = 0.0;
      nb[k][NS] = js; nb[k][NN] = jn; nb[k][NW] = jw; nb[k][NE] = je;
      ccc[k] = a.ccc[i]; cs[k] = s_; cn[k] = n_; cw[k] = w_; ce[k] = e_;
      cci[k] = a.ccci[i];
      bool out = false;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int j = nb[k][q];
        if (real(j) && !inside(j)) {
          out = true;
          hl[ghost_of(j, b0, b1)] = 1;
        }
      }
      if (out || !a.mutual) expo |= 1u << k;
      const double *P = a.x;
      x[k] = P[i];
      // RCG = FPTMP - CCC*P - CCN*PN - CCS*PS - CCW*PW - CCE*PE (:73)
      const double rr = a.fptmp[i] - ccc[k] * x[k] - n_ * P[jn] - s_ * P[js] -
                        w_ * P[jw] - e_ * P[je];
      r[k] = rr;
      v2[0] += rr * (rr * cci[k]);   // RPCG = sum RCG*PCG (:80-85)
      v2[1] += a.fptmp[i] * a.fptmp[i];
    }
  }
  __syncthreads();
  const int nhalo = compact_halo(hl, b0, b1);

  unsigned tag = a.tag0;
  // exchange: z = CCCI*r and p (= 0) of the outside cells, RPCG and
  // |FPTMP|^2 (the owner forms z, the product its neighbours would form, so
  // they load no CCCI of an outside cell)
#pragma unroll
  for (int k = 0; k < CPT; ++k)
    if (expo & (1u << k))
      publish_rp(b0 + k * PB + (int)threadIdx.x, tag, cci[k] * r[k], 0.0);
  post<2>(v2, a.xc, tag);
  double s2[2];
  if (!exchange<2, 4>(a.xc, tag, hl, nhalo, raw, a.bar, s2)) return;
  double rpcg = s2[0];
  const double bnorm2 = s2[1];
  const double rsqm = a.rel_tol > 0.0 ? a.rel_tol * a.rel_tol * bnorm2 : a.rsqm;
  double beta = 0.0, alpha = 0.0, rsq = 0.0, papcg = 0.0;
  int iter = 0, status = 0;
  bool done = (rpcg == 0.0);  // :88

  while (!done) {
    // p = CCCI*r + beta*p (:134,158) on the own rows ...  (no branch per
    // row: the lanes past b1, the last workgroup's, hold zero vectors and
    // zero couplings to lc0, so their rows stay zero; their slots are read
    // by no row)
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
      p[k] = cci[k] * r[k] + beta * p[k];
      lds[slot(b0 + k * PB + (int)threadIdx.x)] = p[k];
    }
    // ... and on the outside cells, from their published z = CCCI*r and p
    for (int h = threadIdx.x; h < nhalo; h += PB) {
      const int j = hl[h];
      const double zj = __longlong_as_double(
          (long long)(((u64)raw[HMAX + h] << 32) | raw[h]));
      const double pj = __longlong_as_double(
          (long long)(((u64)raw[3 * HMAX + h] << 32) | raw[2 * HMAX + h]));
      lds[slot(j)] = zj + beta * pj;
    }
    __syncthreads();
    double v1[1] = {0.0};
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
      double pq[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int j = nb[k][q];
        pq[q] = lds[real(j) ? slot(j) : ZS];
      }
      // APCG = CCC*PCG + CCS*PS + CCN*PN + CCW*PW + CCE*PE (:113)
      ap[k] = ccc[k] * p[k] + cs[k] * pq[NS] + cn[k] * pq[NN] + cw[k] * pq[NW] +
              ce[k] * pq[NE];
      v1[0] += ap[k] * p[k];
    }
    ++tag;
    post<1>(v1, a.xc, tag);
    double s1[1];
    if (!exchange<1, 0>(a.xc, tag, hl, nhalo, raw, a.bar, s1)) return;
    papcg = s1[0];
    alpha = rpcg / papcg;  // :124
    double v[2] = {0.0, 0.0};
    ++tag;
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
      const int i = b0 + k * PB + (int)threadIdx.x;
      x[k] = x[k] + alpha * p[k];        // :132
      const double rr = r[k] - alpha * ap[k];  // :133
      r[k] = rr;
      const double z = cci[k] * rr;      // :134
      v[0] += rr * z;
      v[1] += rr * rr;
      if (expo & (1u << k)) publish_rp(i, tag, z, p[k]);
    }
    post<2>(v, a.xc, tag);
    if (!exchange<2, 4>(a.xc, tag, hl, nhalo, raw, a.bar, s2)) return;
    beta = s2[0] / rpcg;  // :147-148
    rpcg = s2[0];
    rsq = s2[1];
    iter += 1;
    if (rsq <= rsqm) done = true;  // :165
    else if (iter >= a.iterm) { done = true; status = 1; }
  }
#pragma unroll
  for (int k = 0; k < CPT; ++k) {
    const int i = b0 + k * PB + (int)threadIdx.x;
    if (i < b1) a.x[opaque(i)] = x[k];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    CgState *st = a.st;
    st->rpcg = rpcg;
    st->papcg = papcg;
    st->alpha = alpha;
    st->beta = beta;
    st->rsq = rsq;
    st->rsqm = rsqm;
    st->bnorm2 = bnorm2;
    st->iter = iter;
    st->done = 1;
    st->status = status;
  }
}

// ---------------------------------------------------------------------
// Mixed: FP32 Chronopoulos-Gear CG on A' = D^-1/2 A D^-1/2, FP64
// refinement (same operator and scaling as the multi-kernel mixed solve)
// ---------------------------------------------------------------------
struct PMxArgs {
  const double *ccc, *ccs, *ccw, *fptmp;
  const double *dis;         // 1/sqrt(CCC)
  const float *as, *aw;      // A'(i,S), A'(i,W) in FP32
  double *x;                 // P, in/out
  const int *ls, *ln;
  int nwet, lc0;
  Xch xc;                    // exchange slots; halo words: r, w = A r, s = A p
  unsigned tag0;
  Bar bar;
  CgState *st;
  double rsqm, rel_tol, rtol2;
  int iterm, max_outer;
  int pdeg;                  // terms of the workgroup-local Neumann preconditioner
};

template <int CPT>
__global__ __launch_bounds__(PB) void k_pcgmx(PMxArgs a) {
  // the window, then one word that stays zero: a row reads it in place of
  // a neighbour it does not couple to, so that every lane loads
  __shared__ float lds[2 * WIN + CPT * PB + 1];
  __shared__ int hl[2 * WIN];
  __shared__ unsigned raw[3 * HMAX];
  constexpr int ZS = 2 * WIN + CPT * PB;
  // y (the inner solution) and z = M^-1 w of the own rows live in LDS:
  // with them in registers the 6..8-cells-per-thread shapes would spill
  __shared__ float lys[CPT * PB], lzs[CPT * PB];
  const int b0 = 1 + (int)blockIdx.x * CPT * PB;
  const int b1 = min(b0 + CPT * PB, a.nwet + 1);
  const int lc0 = a.lc0;

  float as_[CPT], an_[CPT], aw_[CPT], ae_[CPT];
  float r[CPT], w[CPT], s[CPT], p[CPT];
  float u[CPT], q[CPT];  // M^-1 r, M^-1 s
  auto Y = [&](int k) -> float & { return lys[k * PB + threadIdx.x]; };
  auto Z = [&](int k) -> float & { return lzs[k * PB + threadIdx.x]; };
  int js[CPT], jn[CPT];
  unsigned expo = 0u;

  auto inside = [&](int j) { return j >= b0 && j < b1; };
  auto real = [&](int j) { return j > 0 && j != lc0; };
  auto slot = [&](int j) { return j - b0 + WIN; };
  auto rval = [&](int j) { return lds[real(j) ? slot(j) : ZS]; };
  // FP32 halo value q of cell i in the slots of exchange `tag`
  auto hxw = [&](int i, unsigned tag, int q) {
    return a.xc.hx + ((size_t)(tag & 1u) * HX_WORDS + q) * (size_t)a.xc.ncell +
           (size_t)opaque(i);
  };

  // M^-1 v: the first pdeg terms of the Neumann series of A' restricted to
  // the workgroup's own cells, sum_j N^j v with N = I - A' (couplings to
  // cells of other workgroups dropped), by Horner's rule t <- v + N t.  M^-1
  // is a polynomial in the symmetric local block whose eigenvalues lie in
  // (-1, 1), so it is symmetric positive definite and CG stays CG; it needs
  // no exchange between workgroups.  The own slots of the LDS window carry t.
  // No branch per row: the lanes past b1 (the last workgroup's) hold zero
  // coefficients and zero vectors, so their rows stay zero, and their
  // slots, which no row reads, hold those zeros.
  auto precond = [&](const float (&v)[CPT], float (&t)[CPT]) {
#pragma unroll
    for (int k = 0; k < CPT; ++k) t[k] = v[k];
    for (int d = 1; d < a.pdeg; ++d) {
      __syncthreads();
#pragma unroll
      for (int k = 0; k < CPT; ++k) lds[slot(b0 + k * PB + (int)threadIdx.x)] = t[k];
      __syncthreads();
      float tn[CPT];
#pragma unroll
      for (int k = 0; k < CPT; ++k) {
        // indices made opaque: the slot addresses are formed at each use
        // rather than held in registers across the whole solve
        const int i = opaque(b0 + k * PB + (int)threadIdx.x);
        const int jsk = opaque(js[k]), jnk = opaque(jn[k]);
        const float ts = lds[inside(jsk) ? slot(jsk) : ZS];
        const float tnn = lds[inside(jnk) ? slot(jnk) : ZS];
        const float tw = lds[i - 1 >= b0 ? slot(i - 1) : ZS];
        const float te = lds[i + 1 < b1 ? slot(i + 1) : ZS];
        tn[k] = v[k] - (as_[k] * ts + an_[k] * tnn + aw_[k] * tw + ae_[k] * te);
      }
#pragma unroll
      for (int k = 0; k < CPT; ++k) t[k] = tn[k];
    }
    __syncthreads();  // every read of the own slots is done
  };

  for (int g = threadIdx.x; g < 2 * WIN; g += PB) hl[g] = 0;
  if (threadIdx.x == 0) lds[ZS] = 0.0f;
  __syncthreads();
  double fp2 = 0.0;
#pragma unroll
  for (int k = 0; k < CPT; ++k) {
    const int i = b0 + k * PB + (int)threadIdx.x;
    as_[k] = an_[k] = aw_[k] = ae_[k] = 0.0f;
    r[k] = w[k] = s[k] = p[k] = 0.0f;
    u[k] = q[k] = 0.0f;
    Y(k) = 0.0f;
    Z(k) = 0.0f;
    js[k] = jn[k] = lc0;
    if (i < b1) {
      const int s_ = a.ls[i], n_ = a.ln[i];
      js[k] = s_;
      jn[k] = n_;
      // A'(i,S) = CCS(i) d(i) d(S), A'(i,W) = CCW(i) d(i) d(i-1), formed
      // once per solve by the scaling pass; the north and east entries are
      // the south / west entries of the neighbour
      as_[k] = a.as[i];
      aw_[k] = a.aw[i];
      if (n_ != lc0) an_[k] = a.as[n_];
      if (i + 1 != lc0) ae_[k] = a.aw[i + 1];
      const int nbq[4] = {s_, n_, i - 1, i + 1};
      bool out = false;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int j = nbq[q];
        if (real(j) && !inside(j)) {
          out = true;
          hl[ghost_of(j, b0, b1)] = 1;
        }
      }
      if (out) expo |= 1u << k;
      fp2 += a.fptmp[i] * a.fptmp[i];
    }
  }
  __syncthreads();
  const int nhalo = compact_halo(hl, b0, b1);

  unsigned tag = a.tag0, gen = 0;
  double rsqm = a.rsqm, bnorm2 = 0.0, rsq64 = 0.0, bn_in = 0.0;
  int total = 0, outer = 0, status = 0, inner = 0;

  for (;;) {
    // ---- FP64 residual r = FPTMP - A P, scaled to r' = D^-1/2 r (FP32)
    double v3[3] = {0.0, 0.0, fp2};
    // every lane loads (a lane past b1 from the last own cell; its js, jn
    // are lc0) and the own rows are selected
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
      const int i = b0 + k * PB + (int)threadIdx.x;
      const bool own = i < b1;
      const int s_ = opaque(js[k]), n_ = opaque(jn[k]);
      const int io = opaque(own ? i : b1 - 1);
      const int ie = io + 1;  // <= lc0
      const double *P = a.x;
      const double ccs_i = a.ccs[io], ccs_n = a.ccs[n_], ccw_i = a.ccw[io];
      const double ccw_e = a.ccw[ie], fpt = a.fptmp[io], cci = a.ccc[io], di = a.dis[io];
      const double p_i = fetch(&P[io]), p_n = fetch(&P[n_]), p_s = fetch(&P[s_]);
      const double p_w = fetch(&P[io - 1]), p_e = fetch(&P[ie]);
      const double cs = (s_ != lc0) ? ccs_i : 0.0;
      const double cn = (n_ != lc0) ? ccs_n : 0.0;
      const double cw = ccw_i;
      const double ce = (ie != lc0) ? ccw_e : 0.0;
      const double rr = fpt - cci * p_i - cn * p_n - cs * p_s - cw * p_w - ce * p_e;
      const float rs = (float)(di * rr);
      r[k] = own ? rs : 0.0f;
      Y(k) = 0.0f;
      v3[0] += own ? rr * rr : 0.0;
      v3[1] += own ? (double)rs * (double)rs : 0.0;
    }
    // the partials go out while the preconditioner runs
    post_partials<3>(v3, a.xc, tag);
    // u0 = M^-1 r0, exported for the neighbours' A u0
    precond(r, u);
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
      const int i = b0 + k * PB + (int)threadIdx.x;
      lds[slot(i)] = u[k];
      if (expo & (1u << k)) put_u32(hxw(i, tag, 0), __float_as_uint(u[k]));
    }
    fp2 = 0.0;
    post_flag(a.xc, tag);
    {
      double s3[3];
      if (!exchange<3, 1>(a.xc, tag, hl, nhalo, raw, a.bar, s3)) return;
      if (outer == 0) {
        bnorm2 = s3[2];
        rsqm = a.rel_tol > 0.0 ? a.rel_tol * a.rel_tol * bnorm2 : a.rsqm;
      }
      rsq64 = s3[0];
      bn_in = s3[1];
    }
    if (rsq64 <= rsqm) break;
    if (outer >= a.max_outer || bn_in == 0.0) { status = 1; break; }
    const double rt2 = fmax(a.rtol2, INNER_MARGIN2 * rsqm / rsq64);

    // ---- w0 = A' u0, z0 = M^-1 w0; gamma0 = (r,u), delta0 = (w,u),
    // rho0 = (r,r); s_-1 = p_-1 = q_-1 = 0
    for (int h = threadIdx.x; h < nhalo; h += PB) lds[slot(hl[h])] = __uint_as_float(raw[h]);
    __syncthreads();
    ++tag;
    double v2[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
      const int i = b0 + k * PB + (int)threadIdx.x;
      s[k] = 0.0f;
      p[k] = 0.0f;
      q[k] = 0.0f;
      // a lane past b1 may read a slot no one wrote: its row is zero
      const float wk = u[k] + as_[k] * rval(js[k]) + an_[k] * rval(jn[k]) +
                       aw_[k] * rval(i - 1) + ae_[k] * rval(i + 1);
      w[k] = i < b1 ? wk : 0.0f;
      v2[0] += (double)r[k] * (double)u[k];
      v2[1] += (double)w[k] * (double)u[k];
      v2[2] += (double)r[k] * (double)r[k];
    }
    post_partials<3>(v2, a.xc, tag);
    {
      float zt[CPT];
      precond(w, zt);
#pragma unroll
      for (int k = 0; k < CPT; ++k) {
        const int i = b0 + k * PB + (int)threadIdx.x;
        Z(k) = zt[k];
        if (i < b1 && (expo & (1u << k))) {
          put_u32(hxw(i, tag, 0), __float_as_uint(u[k]));
          put_u32(hxw(i, tag, 1), __float_as_uint(zt[k]));
          put_u32(hxw(i, tag, 2), __float_as_uint(0.0f));
        }
      }
    }
    post_flag(a.xc, tag);
    double gam, del;
    {
      double s3[3];
      if (!exchange<3, 3>(a.xc, tag, hl, nhalo, raw, a.bar, s3)) return;
      gam = s3[0];
      del = s3[1];
    }
    double gam_prev = 0.0, alpha_prev = 1.0;
    inner = 0;
    for (;;) {
      double beta, alpha;
      if (inner == 0) {
        beta = 0.0;
        alpha = gam / del;
      } else {
        beta = gam / gam_prev;
        alpha = gam / (del - beta * gam / alpha_prev);
      }
      // own rows: p = u + beta p, s = w + beta s, q = z + beta q,
      // y += alpha p, r -= alpha s, u -= alpha q
      // the vector updates in FP32 (one rounding each, fused multiply-add),
      // the same operations as a neighbour's recomputation below
      const float bf = (float)beta, af = (float)alpha;
#pragma unroll
      for (int k = 0; k < CPT; ++k) {
        const int i = b0 + k * PB + (int)threadIdx.x;
        p[k] = fmaf(bf, p[k], u[k]);
        s[k] = fmaf(bf, s[k], w[k]);
        q[k] = fmaf(bf, q[k], Z(k));
        Y(k) = fmaf(af, p[k], Y(k));
        r[k] = fmaf(-af, s[k], r[k]);
        u[k] = fmaf(-af, q[k], u[k]);
        lds[slot(i)] = u[k];
      }
      // the outside cells' u, from their published u, z, q
      for (int h = threadIdx.x; h < nhalo; h += PB) {
        const float uj = __uint_as_float(raw[h]);
        const float zj = __uint_as_float(raw[HMAX + h]);
        const float qj0 = __uint_as_float(raw[2 * HMAX + h]);
        const float qj = fmaf(bf, qj0, zj);
        lds[slot(hl[h])] = fmaf(-af, qj, uj);
      }
      __syncthreads();
      ++tag;
      double v[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int k = 0; k < CPT; ++k) {
        const int i = b0 + k * PB + (int)threadIdx.x;
        const float wk = u[k] + as_[k] * rval(js[k]) + an_[k] * rval(jn[k]) +
                         aw_[k] * rval(i - 1) + ae_[k] * rval(i + 1);
        w[k] = i < b1 ? wk : 0.0f;
        v[0] += (double)r[k] * (double)u[k];
        v[1] += (double)w[k] * (double)u[k];
        v[2] += (double)r[k] * (double)r[k];
      }
      post_partials<3>(v, a.xc, tag);
      {
        float zt[CPT];
        precond(w, zt);
#pragma unroll
        for (int k = 0; k < CPT; ++k) {
          const int i = b0 + k * PB + (int)threadIdx.x;
          Z(k) = zt[k];
          if (i < b1 && (expo & (1u << k))) {
            put_u32(hxw(i, tag, 0), __float_as_uint(u[k]));
            put_u32(hxw(i, tag, 1), __float_as_uint(zt[k]));
            put_u32(hxw(i, tag, 2), __float_as_uint(q[k]));
          }
        }
      }
      post_flag(a.xc, tag);
      double s3[3];
      if (!exchange<3, 3>(a.xc, tag, hl, nhalo, raw, a.bar, s3)) return;
      inner += 1;
      gam_prev = gam;
      alpha_prev = alpha;
      gam = s3[0];
      del = s3[1];
      if (s3[2] <= rt2 * bn_in || inner >= a.iterm) break;
    }
    total += inner;
    ++tag;
    // ---- P += D^-1/2 y (FP64), visible to every workgroup for the
    // next residual
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
      const int i = b0 + k * PB + (int)threadIdx.x;
      const bool own = i < b1;
      const int io = opaque(own ? i : b1 - 1);  // every lane loads
      double *P = a.x;
      const double pn = fetch(&P[io]) + a.dis[io] * (double)Y(k);
      if (own) publish(&P[io], pn);
    }
    ++outer;
    if (!grid_barrier(a.bar, ++gen)) return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    CgState *st = a.st;
    st->rsq64 = rsq64;
    st->rsq = rsq64;
    st->rsqm = rsqm;
    st->bnorm2 = bnorm2;
    st->bnorm2_in = bn_in;
    st->iter = total;
    st->inner_iter = inner;
    st->outer = outer;
    st->done = 1;
    st->status = status;
  }
}

// Wait until the workgroups lo..hi (the owners of this workgroup's outside
// cells) have raised the flag of exchange `tag`.  Wave 0 polls (lane l:
// workgroup lo+l); false if the wait was abandoned.
__device__ bool halo_poll(const Xch &x, unsigned tag, int lo, int hi, const Bar &b) {
  __shared__ int s_ok2;
  const int t = threadIdx.x, lane = t & 63;
  if (t < 64) {
    unsigned spins = 0;
    int ok = 1;
    for (;;) {
      bool ready = true;
      for (int q = lo + lane; q <= hi; q += 64) ready = ready && get_u32(&x.flag[q * TICKET_PAD]) >= tag;
      if (__all(ready)) break;
      ++spins;
      const int ab = (spins & 255u) == 0u
                         ? __builtin_amdgcn_readfirstlane(
                               __hip_atomic_load(b.abort, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT))
                         : 0;
      if (spins > b.spin_max || ab) {
        if (lane == 0)
          __hip_atomic_store(b.abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    if (lane == 0) s_ok2 = ok;
  }
  __syncthreads();
  return s_ok2 != 0;
}

// The same, then one FP32 word of each outside cell into raw[h].
__device__ bool halo_wait(const Xch &x, unsigned tag, int lo, int hi, const int *hl, int nhalo,
                          unsigned *raw, const Bar &b) {
  if (!halo_poll(x, tag, lo, hi, b)) return false;
  const int t = threadIdx.x;
  const unsigned *w32 = x.hx + (size_t)(tag & 1u) * HX_WORDS * (size_t)x.ncell;
  constexpr int HB = 4;
  for (int h0 = t; h0 < nhalo; h0 += HB * PB) {
    unsigned v[HB];
#pragma unroll
    for (int u = 0; u < HB; ++u)  // every lane loads
      v[u] = get_u32(w32 + (size_t)hl[min(h0 + u * PB, nhalo - 1)]);
#pragma unroll
    for (int u = 0; u < HB; ++u) {
      const int h = h0 + u * PB;
      if (h < nhalo) raw[h] = v[u];
    }
  }
  __syncthreads();
  return true;
}

// FP64 value 0 of each outside cell in the slots of exchange `tag` (posted
// with put_f64 at word 2*i of that parity's first FP64 plane) into raw[h]
// (low word) and raw[HMAX + h] (high word); the flags must already be up.
__device__ void fetch_halo64(const Xch &x, unsigned tag, const int *hl, int nhalo,
                             unsigned *raw) {
  const int t = threadIdx.x;
  const unsigned *w64 = x.hx + (size_t)(tag & 1u) * 2 * 2 * (size_t)x.ncell;
  constexpr int HB = 4;
  for (int h0 = t; h0 < nhalo; h0 += HB * PB) {
    u64 v[HB];
#pragma unroll
    for (int u = 0; u < HB; ++u)  // every lane loads
      v[u] = get_u64(w64 + 2 * (size_t)hl[min(h0 + u * PB, nhalo - 1)]);
#pragma unroll
    for (int u = 0; u < HB; ++u) {
      const int h = h0 + u * PB;
      if (h < nhalo) {
        raw[h] = (unsigned)v[u];
        raw[HMAX + h] = (unsigned)(v[u] >> 32);
      }
    }
  }
  __syncthreads();
}

// ---------------------------------------------------------------------
// Mixed, pipelined: the same FP32 inner solve in the pipelined form of the
// preconditioned Chronopoulos-Gear CG (Ghysels and Vanroose, 2014): each
// iteration posts the partial sums of (r,u), (w,u), (r,r) and one halo word
// (m = M^-1 w), waits only for the workgroups that own its outside cells,
// forms n = A' m, and then waits for the sums of every workgroup -- which
// were posted one SpMV earlier -- to take alpha and beta:
//   z = n + beta z, q = m + beta q, s = w + beta s, p = u + beta p,
//   y += alpha p, r -= alpha s, u -= alpha q, w -= alpha z.
// The iterates are those of k_pcgmx in exact arithmetic; the outer FP64
// refinement is the same.
// ---------------------------------------------------------------------
template <int CPT>
__global__ __launch_bounds__(PB) void k_pcgmx_pl(PMxArgs a) {
  // the window, then one word that stays zero: a row reads it in place of
  // a neighbour it does not couple to, so that every lane loads
  __shared__ float lds[2 * WIN + CPT * PB + 1];
  __shared__ int hl[2 * WIN];
  __shared__ unsigned raw[HMAX];
  constexpr int ZS = 2 * WIN + CPT * PB;
  // y, q = M^-1 s and z = A' q of the own rows live in LDS
  __shared__ float lys[CPT * PB], lqs[CPT * PB], lzs[CPT * PB];
  const int b0 = 1 + (int)blockIdx.x * CPT * PB;
  const int b1 = min(b0 + CPT * PB, a.nwet + 1);
  const int lc0 = a.lc0;

  float as_[CPT], an_[CPT], aw_[CPT], ae_[CPT];
  float r[CPT], u[CPT], w[CPT], s[CPT], p[CPT];
  auto Y = [&](int k) -> float & { return lys[k * PB + threadIdx.x]; };
  auto Q = [&](int k) -> float & { return lqs[k * PB + threadIdx.x]; };
  auto Z = [&](int k) -> float & { return lzs[k * PB + threadIdx.x]; };
  int js[CPT], jn[CPT];
  unsigned expo = 0u;

  auto inside = [&](int j) { return j >= b0 && j < b1; };
  auto real = [&](int j) { return j > 0 && j != lc0; };
  auto slot = [&](int j) { return j - b0 + WIN; };
  auto hxw = [&](int i, unsigned tag) {
    return a.xc.hx + (size_t)(tag & 1u) * HX_WORDS * (size_t)a.xc.ncell + (size_t)opaque(i);
  };
  // M^-1 v by the local Neumann polynomial, as in k_pcgmx.  No branch per
  // row: the lanes past b1 (the last workgroup's) hold zero coefficients
  // and zero vectors, so their rows stay zero, and their slots, which no
  // row reads, hold those zeros.
  auto precond = [&](const float (&v)[CPT], float (&t)[CPT]) {
#pragma unroll
    for (int k = 0; k < CPT; ++k) t[k] = v[k];
    for (int d = 1; d < a.pdeg; ++d) {
      __syncthreads();
#pragma unroll
      for (int k = 0; k < CPT; ++k) lds[slot(b0 + k * PB + (int)threadIdx.x)] = t[k];
      __syncthreads();
      float tn[CPT];
#pragma unroll
      for (int k = 0; k < CPT; ++k) {
        const int i = b0 + k * PB + (int)threadIdx.x;
        const float ts = lds[inside(js[k]) ? slot(js[k]) : ZS];
        const float tnn = lds[inside(jn[k]) ? slot(jn[k]) : ZS];
        const float tw = lds[i - 1 >= b0 ? slot(i - 1) : ZS];
        const float te = lds[i + 1 < b1 ? slot(i + 1) : ZS];
        tn[k] = v[k] - (as_[k] * ts + an_[k] * tnn + aw_[k] * tw + ae_[k] * te);
      }
#pragma unroll
      for (int k = 0; k < CPT; ++k) t[k] = tn[k];
    }
    __syncthreads();
  };
  // y = A' x on the own rows from the LDS window (own slots and outside
  // cells); a lane past b1 may read a slot no one wrote, so its row is
  // replaced by zero
  auto spmv = [&](const float (&x)[CPT], float (&y)[CPT]) {
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
      const int i = b0 + k * PB + (int)threadIdx.x;
      const float vs = lds[real(js[k]) ? slot(js[k]) : ZS];
      const float vn = lds[real(jn[k]) ? slot(jn[k]) : ZS];
      const float vw = lds[real(i - 1) ? slot(i - 1) : ZS];
      const float ve = lds[real(i + 1) ? slot(i + 1) : ZS];
      const float yk = x[k] + as_[k] * vs + an_[k] * vn + aw_[k] * vw + ae_[k] * ve;
      y[k] = i < b1 ? yk : 0.0f;
    }
  };

  for (int g = threadIdx.x; g < 2 * WIN; g += PB) hl[g] = 0;
  if (threadIdx.x == 0) lds[ZS] = 0.0f;
  __syncthreads();
  double fp2 = 0.0;
  // every lane loads (a lane past b1 from the last own cell) and keeps the
  // values of an own row only
#pragma unroll
  for (int k = 0; k < CPT; ++k) {
    const int i = b0 + k * PB + (int)threadIdx.x;
    const bool own = i < b1;
    const int ic = own ? i : b1 - 1;
    r[k] = u[k] = w[k] = s[k] = p[k] = 0.0f;
    Y(k) = Q(k) = Z(k) = 0.0f;
    const int s_ = a.ls[ic], n_ = a.ln[ic];
    const float as_i = a.as[ic], aw_i = a.aw[ic], as_n = a.as[n_], aw_e = a.aw[ic + 1];
    const double fpt = a.fptmp[ic];
    js[k] = own ? s_ : lc0;
    jn[k] = own ? n_ : lc0;
    as_[k] = own ? as_i : 0.0f;
    aw_[k] = own ? aw_i : 0.0f;
    an_[k] = (own && n_ != lc0) ? as_n : 0.0f;
    ae_[k] = (own && ic + 1 != lc0) ? aw_e : 0.0f;
    if (own) {
      const int nbq[4] = {s_, n_, i - 1, i + 1};
      bool out = false;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int j = nbq[q];
        if (real(j) && !inside(j)) {
          out = true;
          hl[ghost_of(j, b0, b1)] = 1;
        }
      }
      if (out) expo |= 1u << k;
    }
    fp2 += own ? fpt * fpt : 0.0;
  }
  __syncthreads();
  const int nhalo = compact_halo(hl, b0, b1);
  // the workgroups that own the outside cells (hl is sorted)
  const int span = CPT * PB;
  const int nlo = nhalo > 0 ? (hl[0] - 1) / span : 1;
  const int nhi = nhalo > 0 ? (hl[nhalo - 1] - 1) / span : 0;

  unsigned tag = a.tag0, gen = 0;
  double rsqm = a.rsqm, bnorm2 = 0.0, rsq64 = 0.0, bn_in = 0.0;
  int total = 0, outer = 0, status = 0, inner = 0;

  for (;;) {
    // ---- FP64 residual r = FPTMP - A P, scaled to r' = D^-1/2 r (FP32)
    double v3[3] = {0.0, 0.0, fp2};
    // every lane loads (a lane past b1 from the last own cell; its js, jn
    // are lc0) and the own rows are selected, so the loads of all the
    // thread's rows are in flight together
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
      const int i = b0 + k * PB + (int)threadIdx.x;
      const bool own = i < b1;
      const int s_ = opaque(js[k]), n_ = opaque(jn[k]);
      const int io = opaque(own ? i : b1 - 1);
      const int ie = io + 1;  // <= lc0
      const double *P = a.x;
      const double ccs_i = a.ccs[io], ccs_n = a.ccs[n_], ccw_i = a.ccw[io];
      const double ccw_e = a.ccw[ie], fpt = a.fptmp[io], cci = a.ccc[io], di = a.dis[io];
      const double p_i = fetch(&P[io]), p_n = fetch(&P[n_]), p_s = fetch(&P[s_]);
      const double p_w = fetch(&P[io - 1]), p_e = fetch(&P[ie]);
      const double cs = (s_ != lc0) ? ccs_i : 0.0;
      const double cn = (n_ != lc0) ? ccs_n : 0.0;
      const double cw = ccw_i;
      const double ce = (ie != lc0) ? ccw_e : 0.0;
      const double rr = fpt - cci * p_i - cn * p_n - cs * p_s - cw * p_w - ce * p_e;
      const float rs = (float)(di * rr);
      r[k] = own ? rs : 0.0f;
      Y(k) = 0.0f;
      v3[0] += own ? rr * rr : 0.0;
      v3[1] += own ? (double)rs * (double)rs : 0.0;
    }
    post_partials<3>(v3, a.xc, tag);
    precond(r, u);
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
      const int i = b0 + k * PB + (int)threadIdx.x;
      lds[slot(i)] = u[k];
      if (expo & (1u << k)) put_u32(hxw(i, tag), __float_as_uint(u[k]));
    }
    fp2 = 0.0;
    post_flag(a.xc, tag);
    {
      double s3[3];
      if (!exchange<3, 1>(a.xc, tag, hl, nhalo, raw, a.bar, s3)) return;
      if (outer == 0) {
        bnorm2 = s3[2];
        rsqm = a.rel_tol > 0.0 ? a.rel_tol * a.rel_tol * bnorm2 : a.rsqm;
      }
      rsq64 = s3[0];
      bn_in = s3[1];
    }
    if (rsq64 <= rsqm) break;
    if (outer >= a.max_outer || bn_in == 0.0) { status = 1; break; }
    const double rt2 = fmax(a.rtol2, INNER_MARGIN2 * rsqm / rsq64);

    // ---- w0 = A' u0; s, p, q, z start at 0
    for (int h = threadIdx.x; h < nhalo; h += PB) lds[slot(hl[h])] = __uint_as_float(raw[h]);
    __syncthreads();
    spmv(u, w);
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
      s[k] = p[k] = 0.0f;
      Q(k) = Z(k) = 0.0f;
    }
    double gam_prev = 0.0, alpha_prev = 1.0;
    inner = 0;
    for (;;) {
      ++tag;
      // the sums of this iteration go out first ...
      // (the lanes past b1 add zeros)
      double v[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int k = 0; k < CPT; ++k) {
        v[0] += (double)r[k] * (double)u[k];
        v[1] += (double)w[k] * (double)u[k];
        v[2] += (double)r[k] * (double)r[k];
      }
      post_partials<3>(v, a.xc, tag);
      // ... then m = M^-1 w, its outside values, the flag
      float m[CPT];
      precond(w, m);
#pragma unroll
      for (int k = 0; k < CPT; ++k) {
        const int i = b0 + k * PB + (int)threadIdx.x;
        lds[slot(i)] = m[k];
        if (expo & (1u << k)) put_u32(hxw(i, tag), __float_as_uint(m[k]));
      }
      post_flag(a.xc, tag);
      // n = A' m once the neighbouring workgroups' m has arrived
      if (nhalo > 0 && !halo_wait(a.xc, tag, nlo, nhi, hl, nhalo, raw, a.bar)) return;
      for (int h = threadIdx.x; h < nhalo; h += PB) lds[slot(hl[h])] = __uint_as_float(raw[h]);
      __syncthreads();
      float n[CPT];
      spmv(m, n);
      // the sums of every workgroup
      double s3[3];
      if (!exchange<3, 0>(a.xc, tag, hl, 0, raw, a.bar, s3)) return;
      const double gam = s3[0], del = s3[1];
      if (s3[2] <= rt2 * bn_in || inner >= a.iterm) break;
      double beta, alpha;
      if (inner == 0) {
        beta = 0.0;
        alpha = gam / del;
      } else {
        beta = gam / gam_prev;
        alpha = gam / (del - beta * gam / alpha_prev);
      }
      // the vector updates in FP32 (one rounding each, fused multiply-add);
      // the sums, alpha and beta stay FP64 and the outer FP64 residual
      // decides convergence
      const float bf = (float)beta, af = (float)alpha;
#pragma unroll
      for (int k = 0; k < CPT; ++k) {
        const float zk = fmaf(bf, Z(k), n[k]);
        const float qk = fmaf(bf, Q(k), m[k]);
        s[k] = fmaf(bf, s[k], w[k]);
        p[k] = fmaf(bf, p[k], u[k]);
        Y(k) = fmaf(af, p[k], Y(k));
        r[k] = fmaf(-af, s[k], r[k]);
        u[k] = fmaf(-af, qk, u[k]);
        w[k] = fmaf(-af, zk, w[k]);
        Z(k) = zk;
        Q(k) = qk;
      }
      inner += 1;
      gam_prev = gam;
      alpha_prev = alpha;
    }
    total += inner;
    ++tag;
    // ---- P += D^-1/2 y (FP64), visible to every workgroup for the next
    // residual
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
      const int i = b0 + k * PB + (int)threadIdx.x;
      const bool own = i < b1;
      const int io = opaque(own ? i : b1 - 1);  // every lane loads
      double *P = a.x;
      const double pn = fetch(&P[io]) + a.dis[io] * (double)Y(k);
      if (own) publish(&P[io], pn);
    }
    ++outer;
    if (!grid_barrier(a.bar, ++gen)) return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    CgState *st = a.st;
    st->rsq64 = rsq64;
    st->rsq = rsq64;
    st->rsqm = rsqm;
    st->bnorm2 = bnorm2;
    st->bnorm2_in = bn_in;
    st->iter = total;
    st->inner_iter = inner;
    st->outer = outer;
    st->done = 1;
    st->status = status;
  }
}

// ---------------------------------------------------------------------
// FP64 Jacobi-PCG, pipelined (opt-in: efdc_set_cg64_pipeline).  The
// reference's PCG (EFDC/congrad.f90:91-169) rewritten in the pipelined order
// of Ghysels and Vanroose (2014), as k_pcgmx_pl does for the mixed inner
// solve: an iteration posts its sums (r,u), (w,u), (r,r), the m = CCCI w
// of its outside-read cells and its flag; it waits only for the workgroups
// that own its outside cells, forms n = A m, and only then waits for every
// workgroup's sums to take alpha and beta.  The Jacobi preconditioner is
// diagonal, so u = CCCI r, m = CCCI w and q = CCCI s are formed where they
// are used; the state is x, r, w, s, p (registers) and z = A q (LDS):
//   z = n + beta z, s = w + beta s, p = u + beta p,
//   x += alpha p, r -= alpha s, w -= alpha z.
// Same iterates as the reference's recurrences in exact arithmetic, one
// exchange per iteration instead of two.  The stop test is the
// reference's, (r,r) <= RSQM (:165), made on the sums of the iteration
// that would update from r.
// ---------------------------------------------------------------------
template <int CPT>
__global__ __launch_bounds__(PB) void k_pcg64_pl(P64Args a) {
  // the window, then one word that stays zero (read in place of a
  // neighbour that is not a cell, so that every lane loads)
  __shared__ double lds[2 * WIN + CPT * PB + 1];
  __shared__ int hl[2 * WIN];
  __shared__ unsigned raw[2 * HMAX];
  __shared__ double lzs[CPT * PB];
  constexpr int ZS = 2 * WIN + CPT * PB;
  const int b0 = 1 + (int)blockIdx.x * CPT * PB;
  const int b1 = min(b0 + CPT * PB, a.nwet + 1);
  const bool fast = a.rowmajor && (a.st->symok == 0);
  const int lc0 = a.lc0;

  double ccc[CPT], cs[CPT], cn[CPT], cw[CPT], ce[CPT], cci[CPT];
  double x[CPT], r[CPT], w[CPT], s[CPT], p[CPT];
  int nb[CPT][4];
  unsigned expo = 0u;        // bit k: another workgroup reads cell k
  auto Z = [&](int k) -> double & { return lzs[k * PB + threadIdx.x]; };

  auto inside = [&](int j) { return j >= b0 && j < b1; };
  auto real = [&](int j) { return j > 0 && j != lc0; };
  auto slot = [&](int j) { return j - b0 + WIN; };
  auto w64 = [&](int i, unsigned tag) {
    return a.xc.hx + (size_t)(tag & 1u) * 2 * 2 * (size_t)a.xc.ncell +
           2 * (size_t)opaque(i);
  };
  // y = A v on the own rows from the window (own slots and outside cells);
  // the lanes past b1 have ccc = 1, zero couplings and v = 0
  auto spmv = [&](const double (&v)[CPT], double (&y)[CPT]) {
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
      double pq[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int j = nb[k][q];
        pq[q] = lds[real(j) ? slot(j) : ZS];
      }
      // APCG = CCC*PCG + CCS*PS + CCN*PN + CCW*PW + CCE*PE (:113)
      y[k] = ccc[k] * v[k] + cs[k] * pq[NS] + cn[k] * pq[NN] + cw[k] * pq[NW] +
             ce[k] * pq[NE];
    }
  };

  for (int g = threadIdx.x; g < 2 * WIN; g += PB) hl[g] = 0;
  if (threadIdx.x == 0) lds[ZS] = 0.0;
  __syncthreads();
  double fp2 = 0.0;
#pragma unroll
  for (int k = 0; k < CPT; ++k) {
    const int i = b0 + k * PB + (int)threadIdx.x;
    x[k] = r[k] = w[k] = s[k] = p[k] = 0.0;
    Z(k) = 0.0;
    ccc[k] = 1.0; cs[k] = cn[k] = cw[k] = ce[k] = cci[k] = 0.0;
    nb[k][NS] = nb[k][NN] = nb[k][NW] = nb[k][NE] = lc0;
    if (i < b1) {
      // the row as in the multi-kernel solve (row_of)
      const int js = a.ls[i], jn = a.ln[i];
      int jw, je;
      double s_ = a.ccs[i], w_ = a.ccw[i], n_, e_;
      if (fast) {
        jw = i - 1;
        je = i + 1;
        n_ = (jn != lc0) ? a.ccs[jn] : 0.0;
        e_ = (je != lc0) ? a.ccw[je] : 0.0;
      } else {
        jw = a.lw[i];
        je = a.le[i];
        n_ = (jn != lc0) ? a.ccn[i] : 0.0;
        e_ = (je != lc0) ? a.cce[i] : 0.0;
        if (jw == lc0) w_ = 0.0;
      }
      if (js == lc0) s_ = 0.0;
      nb[k][NS] = js; nb[k][NN] = jn; nb[k][NW] = jw; nb[k][NE] = je;
      ccc[k] = a.ccc[i]; cs[k] = s_; cn[k] = n_; cw[k] = w_; ce[k] = e_;
      cci[k] = a.ccci[i];
      bool out = false;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int j = nb[k][q];
        if (real(j) && !inside(j)) {
          out = true;
          hl[ghost_of(j, b0, b1)] = 1;
        }
      }
      if (out || !a.mutual) expo |= 1u << k;
      const double *P = a.x;
      x[k] = P[i];
      // RCG = FPTMP - CCC*P - CCN*PN - CCS*PS - CCW*PW - CCE*PE (:73)
      r[k] = a.fptmp[i] - ccc[k] * x[k] - n_ * P[jn] - s_ * P[js] - w_ * P[jw] - e_ * P[je];
      fp2 += a.fptmp[i] * a.fptmp[i];
    }
  }
  __syncthreads();
  const int nhalo = compact_halo(hl, b0, b1);
  const int span = CPT * PB;
  const int nlo = nhalo > 0 ? (hl[0] - 1) / span : 1;
  const int nhi = nhalo > 0 ? (hl[nhalo - 1] - 1) / span : 0;

  // set-up exchange: u = CCCI r of the outside-read cells and |FPTMP|^2;
  // then w = A u (the owners publish u, so no CCCI of an outside cell is
  // loaded)
  unsigned tag = a.tag0;
#pragma unroll
  for (int k = 0; k < CPT; ++k) {
    const int i = b0 + k * PB + (int)threadIdx.x;
    const double uk = cci[k] * r[k];
    lds[slot(i)] = uk;
    if (expo & (1u << k)) put_f64(w64(i, tag), uk);
  }
  {
    double v1[1] = {fp2};
    post<1>(v1, a.xc, tag);
  }
  double s1[1];
  if (!exchange<1, 0>(a.xc, tag, hl, 0, raw, a.bar, s1)) return;
  const double bnorm2 = s1[0];
  const double rsqm = a.rel_tol > 0.0 ? a.rel_tol * a.rel_tol * bnorm2 : a.rsqm;
  fetch_halo64(a.xc, tag, hl, nhalo, raw);
  for (int h = threadIdx.x; h < nhalo; h += PB)
    lds[slot(hl[h])] = __longlong_as_double((long long)(((u64)raw[HMAX + h] << 32) | raw[h]));
  __syncthreads();
  {
    double u0[CPT];
#pragma unroll
    for (int k = 0; k < CPT; ++k) u0[k] = cci[k] * r[k];
    spmv(u0, w);
  }

  double gam_prev = 0.0, alpha_prev = 1.0, gam = 0.0, alpha = 0.0, beta = 0.0, rsq = 0.0;
  int iter = 0, status = 0;
  for (;;) {
    ++tag;
    // the sums of this iteration go out first (the lanes past b1 add zeros) ...
    double v[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
      const double uk = cci[k] * r[k];
      v[0] += r[k] * uk;
      v[1] += w[k] * uk;
      v[2] += r[k] * r[k];
    }
    post_partials<3>(v, a.xc, tag);
    // ... then m = CCCI w on the window (after every read of the last SpMV),
    // m of the outside-read cells, the flag
    __syncthreads();
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
      const int i = b0 + k * PB + (int)threadIdx.x;
      const double mk = cci[k] * w[k];
      lds[slot(i)] = mk;
      if (expo & (1u << k)) put_f64(w64(i, tag), mk);
    }
    post_flag(a.xc, tag);
    // n = A m once the neighbouring workgroups' m has arrived
    if (nhalo > 0) {
      if (!halo_poll(a.xc, tag, nlo, nhi, a.bar)) return;
      fetch_halo64(a.xc, tag, hl, nhalo, raw);
      for (int h = threadIdx.x; h < nhalo; h += PB)
        lds[slot(hl[h])] =
            __longlong_as_double((long long)(((u64)raw[HMAX + h] << 32) | raw[h]));
    }
    __syncthreads();
    double m[CPT], n[CPT];
#pragma unroll
    for (int k = 0; k < CPT; ++k) m[k] = cci[k] * w[k];
    spmv(m, n);
    // the sums of every workgroup
    double s3[3];
    if (!exchange<3, 0>(a.xc, tag, hl, 0, raw, a.bar, s3)) return;
    gam = s3[0];
    const double del = s3[1];
    rsq = s3[2];
    if (rsq <= rsqm || gam == 0.0) break;  // :165 (and RPCG = 0, :88)
    if (iter >= a.iterm) { status = 1; break; }
    if (iter == 0) {
      beta = 0.0;
      alpha = gam / del;
    } else {
      beta = gam / gam_prev;
      alpha = gam / (del - beta * gam / alpha_prev);
    }
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
      const double zk = n[k] + beta * Z(k);
      s[k] = w[k] + beta * s[k];
      p[k] = cci[k] * r[k] + beta * p[k];
      x[k] = x[k] + alpha * p[k];
      r[k] = r[k] - alpha * s[k];
      w[k] = w[k] - alpha * zk;
      Z(k) = zk;
    }
    iter += 1;
    gam_prev = gam;
    alpha_prev = alpha;
  }
#pragma unroll
  for (int k = 0; k < CPT; ++k) {
    const int i = b0 + k * PB + (int)threadIdx.x;
    if (i < b1) a.x[opaque(i)] = x[k];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    CgState *st = a.st;
    st->rpcg = gam;
    st->papcg = 0.0;
    st->alpha = alpha;
    st->beta = beta;
    st->rsq = rsq;
    st->rsqm = rsqm;
    st->bnorm2 = bnorm2;
    st->iter = iter;
    st->done = 1;
    st->status = status;
  }
}

// Jacobi scaling of the mixed solve from a scaled system (CCC, CCS, CCW)
// given directly: dis = 1/sqrt(CCC), A'(i,S), A'(i,W) in FP32.
__global__ __launch_bounds__(256) void k_mx_coef(const double *ccc, const double *ccs,
                                                 const double *ccw, const int *ls, int nwet,
                                                 int lc0, double *dis, float *as, float *aw) {
  const int i = 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i > nwet) return;
  const int js = ls[i];
  const double di = 1.0 / sqrt(ccc[i]);
  const double ds = (js != lc0) ? 1.0 / sqrt(ccc[js]) : 0.0;
  const double dw = (i - 1 > 0) ? 1.0 / sqrt(ccc[i - 1]) : 0.0;
  const double cs = (js != lc0) ? ccs[i] : 0.0;
  dis[i] = di;
  as[i] = (float)(cs * di * ds);
  aw[i] = (float)(ccw[i] * di * dw);
}

__global__ void k_pcg_reset(unsigned *ctr, int *abort) {
  const int t = threadIdx.x;
  if (t < PCG_NCTR) ctr[t] = 0u;
  if (t == 0) *abort = 0;
}

template <typename K>
int occupancy(K kern) {
  int nbpc = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nbpc, kern, PB, 0) != hipSuccess)
    return 0;
  return nbpc;
}

}  // namespace

// Largest cells-per-thread instantiation of each solve.
constexpr int PCG64_CPT_MAX = 6;
constexpr int PCGMX_CPT_MAX = 8;
// Terms of the workgroup-local preconditioner of the mixed persistent solve
// when the caller leaves the choice to the library.
constexpr int PCG_PDEG_DEFAULT = 3;
// Exchange numbers: launch number (1 .. 4095) in the top 12 bits, exchange
// within the launch in the low 20 bits.
constexpr int PCG_TAG_BITS = 20;
constexpr long PCG_TAGS_PER_LAUNCH = 1L << PCG_TAG_BITS;
constexpr unsigned PCG_LAUNCH_CYCLE = (1u << (32 - PCG_TAG_BITS)) - 1u;

int pcg_prepare(efdc_ctx *c) {
  if (c->pcg_ready) return EFDC_OK;
  int dev = c->device, ncu = 0;
  HIP_TRY(c, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  int coop = 0;
  HIP_TRY(c, hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
  c->pcg_ncu = coop ? ncu : 0;
  const int occ64[PCG64_CPT_MAX + 1] = {
      0, occupancy(k_pcg64<1>), occupancy(k_pcg64<2>), occupancy(k_pcg64<3>),
      occupancy(k_pcg64<4>), occupancy(k_pcg64<5>), occupancy(k_pcg64<6>)};
  const int occmx[PCGMX_CPT_MAX + 1] = {
      0, occupancy(k_pcgmx<1>), occupancy(k_pcgmx<2>), occupancy(k_pcgmx<3>),
      occupancy(k_pcgmx<4>), occupancy(k_pcgmx<5>), occupancy(k_pcgmx<6>),
      occupancy(k_pcgmx<7>), occupancy(k_pcgmx<8>)};
  const int occpl[PCGMX_CPT_MAX + 1] = {
      0, occupancy(k_pcgmx_pl<1>), occupancy(k_pcgmx_pl<2>), occupancy(k_pcgmx_pl<3>),
      occupancy(k_pcgmx_pl<4>), occupancy(k_pcgmx_pl<5>), occupancy(k_pcgmx_pl<6>),
      occupancy(k_pcgmx_pl<7>), occupancy(k_pcgmx_pl<8>)};
  const int occ64p[PCG64_CPT_MAX + 1] = {
      0, occupancy(k_pcg64_pl<1>), occupancy(k_pcg64_pl<2>), occupancy(k_pcg64_pl<3>),
      occupancy(k_pcg64_pl<4>), occupancy(k_pcg64_pl<5>), occupancy(k_pcg64_pl<6>)};
  for (int k = 0; k <= PCG64_CPT_MAX; ++k) c->pcg_occ64[k] = std::min(occ64[k], occ64p[k]);
  for (int k = 0; k <= PCGMX_CPT_MAX; ++k) c->pcg_occmx[k] = std::min(occmx[k], occpl[k]);
  c->pcg_ncell = c->n_wet() + 2;
  int s;
  if ((s = dev_alloc(c, &c->pcg_hx,
                     sizeof(unsigned) * 2 * (size_t)c->pcg_ncell * HX_WORDS)))
    return s;
  if ((s = dev_alloc(c, &c->pcg_ps, sizeof(double) * 2 * NVMAX * PB))) return s;
  if ((s = dev_alloc(c, (void **)&c->pcg_flag, sizeof(unsigned) * PB * TICKET_PAD))) return s;
  if ((s = dev_alloc(c, (void **)&c->pcg_ctr, sizeof(unsigned) * PCG_NCTR))) return s;
  c->pcg_abort = &c->st->pabort;
  c->pcg_launch = 0;
  c->pcg_ready = true;
  return EFDC_OK;
}

// Cells per thread and workgroup count of a persistent solve, or 0 when
// the problem does not fit in one co-resident grid.
static int pcg_shape(const efdc_ctx *c, const int *occ, int cpt_max, int &nb) {
  const int n = c->n_wet();
  for (int cpt = 1; cpt <= cpt_max; ++cpt) {
    const int need = (n + cpt * PB - 1) / (cpt * PB);
    const int cap = occ[cpt] * c->pcg_ncu;
    if (occ[cpt] > 0 && need <= cap && need <= PB) {
      nb = need;
      return cpt;
    }
  }
  return 0;
}

bool pcg_applicable(efdc_ctx *c, bool mixed) {
  if (c->prm.cg_kernel == EFDC_CG_KERNEL_MULTI) return false;
  if (c->band > WIN) return false;  // a neighbour beyond the LDS window
  // the exchanges of one solve must fit one launch's tag range
  const long per_outer = 3L + c->prm.ITERM;
  const long outers = mixed ? (long)c->prm.mixed_max_outer + 1 : 1;
  if ((mixed ? per_outer * outers : 1L + 2L * c->prm.ITERM) >= PCG_TAGS_PER_LAUNCH)
    return false;
  if (pcg_prepare(c) != EFDC_OK || c->pcg_ncu == 0) return false;
  int nb = 0;
  return mixed ? pcg_shape(c, c->pcg_occmx, PCGMX_CPT_MAX, nb) > 0
               : pcg_shape(c, c->pcg_occ64, PCG64_CPT_MAX, nb) > 0;
}

// The persistent grid is sized to be co-resident (the occupancy table times
// the CUs).  A cooperative launch guarantees it, at the price of the
// runtime's cooperative dispatch (about 15 us per solve on MI355X); a plain
// launch of the same grid is co-resident whenever no other work holds the
// CUs, and otherwise every bounded wait gives up and the solve reports an
// abort, after which it is solved again with the cooperative launch
// (pcg_solve, pcg_retry_deferred).  pcg_launch_mode 1 always uses the
// cooperative launch.
template <typename Args>
static int coop_launch(efdc_ctx *c, const void *fn, int nb, Args &args) {
  void *kargs[] = {&args};
  if (c->pcg_launch_mode == 0 && !c->pcg_force_coop) {
    HIP_TRY(c, hipLaunchKernel(fn, dim3(nb), dim3(PB), kargs, 0, c->stream));
    c->pcg_last_plain = true;
  } else {
    HIP_TRY(c, hipLaunchCooperativeKernel(fn, dim3(nb), dim3(PB), kargs, 0, c->stream));
    c->pcg_last_plain = false;
  }
  return EFDC_OK;
}

// First exchange number of the next launch.  Launches are numbered
// 1 .. 4095, so within a cycle exchange numbers only grow and the flags never
// go backwards; when the numbering starts a new cycle the flags are cleared.
static int next_tag0(efdc_ctx *c, unsigned &tag0) {
  const bool wrap = c->pcg_launch == PCG_LAUNCH_CYCLE;
  c->pcg_launch = c->pcg_launch % PCG_LAUNCH_CYCLE + 1u;
  if (wrap)
    HIP_TRY(c, hipMemsetAsync(c->pcg_flag, 0, sizeof(unsigned) * PB * TICKET_PAD,
                              c->stream));
  tag0 = c->pcg_launch << PCG_TAG_BITS;
  return EFDC_OK;
}

int pcg_solve(efdc_ctx *c, bool mixed, bool defer) {
  int s;
  if ((s = pcg_prepare(c))) return s;
  if (!c->pcg_reset_done) {
    k_pcg_reset<<<1, PCG_NCTR, 0, c->stream>>>(c->pcg_ctr, c->pcg_abort);
    HIP_TRY(c, hipGetLastError());
  }
  c->pcg_reset_done = false;  // this launch uses them
  int nb = 0;
  // a plain launch may be given a small spin limit (to exercise the retry
  // from the host); the cooperative one always has the full limit
  const bool plain = c->pcg_launch_mode == 0 && !c->pcg_force_coop;
  Bar bar{c->pcg_ctr, c->pcg_abort,
          plain && c->pcg_spin_test > 0 ? c->pcg_spin_test : SPIN_MAX};
  if (!mixed) {
    const int cpt = pcg_shape(c, c->pcg_occ64, PCG64_CPT_MAX, nb);
    if (!cpt) { c->err = "persistent CG: problem too large"; return EFDC_ERR_ARG; }
    P64Args a;
    a.ccc = c->dH[EFDC_CCC];
    a.ccs = c->dH[EFDC_CCS];
    a.ccw = c->dH[EFDC_CCW];
    a.ccn = c->dH[EFDC_CCN];
    a.cce = c->dH[EFDC_CCE];
    a.ccci = c->dH[EFDC_CCCI];
    a.fptmp = c->dH[EFDC_FPTMP];
    a.x = c->dH[EFDC_P];
    a.lw = c->dLW;
    a.le = c->dLE;
    a.ls = c->dLS;
    a.ln = c->dLN;
    a.nwet = c->n_wet();
    a.lc0 = c->LC - 1;
    a.rowmajor = c->rowmajor ? 1 : 0;
    a.mutual = c->mutual ? 1 : 0;
    a.xc = Xch{(double *)c->pcg_ps, (unsigned *)c->pcg_hx, c->pcg_flag, c->pcg_ncell};
    if ((s = next_tag0(c, a.tag0))) return s;
    a.bar = bar;
    a.st = c->st;
    a.rsqm = c->prm.RSQM;
    a.rel_tol = c->prm.cg_rel_tol;
    a.iterm = c->prm.ITERM;
    const void *fn = nullptr;
    if (c->pcg64_pipe) {
      switch (cpt) {
        case 1: fn = (const void *)k_pcg64_pl<1>; break;
        case 2: fn = (const void *)k_pcg64_pl<2>; break;
        case 3: fn = (const void *)k_pcg64_pl<3>; break;
        case 4: fn = (const void *)k_pcg64_pl<4>; break;
        case 5: fn = (const void *)k_pcg64_pl<5>; break;
        default: fn = (const void *)k_pcg64_pl<6>; break;
      }
    } else {
      switch (cpt) {
        case 1: fn = (const void *)k_pcg64<1>; break;
        case 2: fn = (const void *)k_pcg64<2>; break;
        case 3: fn = (const void *)k_pcg64<3>; break;
        case 4: fn = (const void *)k_pcg64<4>; break;
        case 5: fn = (const void *)k_pcg64<5>; break;
        default: fn = (const void *)k_pcg64<6>; break;
      }
    }
    TIMED(c, TK_PCG64, s = coop_launch(c, fn, nb, a));
    if (s) return s;
  } else {
    const int cpt = pcg_shape(c, c->pcg_occmx, PCGMX_CPT_MAX, nb);
    if (!cpt) { c->err = "persistent CG: problem too large"; return EFDC_ERR_ARG; }
    if (!c->mx_coef_ok) {
      const int n = c->n_wet();
      k_mx_coef<<<(n + 255) / 256, 256, 0, c->stream>>>(
          c->dH[EFDC_CCC], c->dH[EFDC_CCS], c->dH[EFDC_CCW], c->dLS, n, c->LC - 1, c->dis,
          c->as32, c->aw32);
      HIP_TRY(c, hipGetLastError());
    }
    PMxArgs a;
    a.ccc = c->dH[EFDC_CCC];
    a.ccs = c->dH[EFDC_CCS];
    a.ccw = c->dH[EFDC_CCW];
    a.fptmp = c->dH[EFDC_FPTMP];
    a.dis = c->dis;
    a.as = c->as32;
    a.aw = c->aw32;
    a.x = c->dH[EFDC_P];
    a.ls = c->dLS;
    a.ln = c->dLN;
    a.nwet = c->n_wet();
    a.lc0 = c->LC - 1;
    a.xc = Xch{(double *)c->pcg_ps, (unsigned *)c->pcg_hx, c->pcg_flag, c->pcg_ncell};
    if ((s = next_tag0(c, a.tag0))) return s;
    a.bar = bar;
    a.st = c->st;
    a.rsqm = c->prm.RSQM;
    a.rel_tol = c->prm.cg_rel_tol;
    const double rtol = c->prm.mixed_inner_rtol > 0 ? c->prm.mixed_inner_rtol : 1e-5;
    a.rtol2 = rtol * rtol;
    a.iterm = c->prm.ITERM;
    a.max_outer = c->prm.mixed_max_outer;
    a.pdeg = c->prm.mixed_precond_degree > 0 ? c->prm.mixed_precond_degree
             : c->pcg_pdeg_tuned > 0         ? c->pcg_pdeg_tuned
                                             : PCG_PDEG_DEFAULT;
    const bool pl = c->pcg_pipe_mode == 2 || (c->pcg_pipe_mode == 0 && c->pcg_pipe_tuned);
    const void *fn = nullptr;
    if (pl) {
      switch (cpt) {
        case 1: fn = (const void *)k_pcgmx_pl<1>; break;
        case 2: fn = (const void *)k_pcgmx_pl<2>; break;
        case 3: fn = (const void *)k_pcgmx_pl<3>; break;
        case 4: fn = (const void *)k_pcgmx_pl<4>; break;
        case 5: fn = (const void *)k_pcgmx_pl<5>; break;
        case 6: fn = (const void *)k_pcgmx_pl<6>; break;
        case 7: fn = (const void *)k_pcgmx_pl<7>; break;
        default: fn = (const void *)k_pcgmx_pl<8>; break;
      }
    } else {
      switch (cpt) {
        case 1: fn = (const void *)k_pcgmx<1>; break;
        case 2: fn = (const void *)k_pcgmx<2>; break;
        case 3: fn = (const void *)k_pcgmx<3>; break;
        case 4: fn = (const void *)k_pcgmx<4>; break;
        case 5: fn = (const void *)k_pcgmx<5>; break;
        case 6: fn = (const void *)k_pcgmx<6>; break;
        case 7: fn = (const void *)k_pcgmx<7>; break;
        default: fn = (const void *)k_pcgmx<8>; break;
      }
    }
    TIMED(c, TK_PCGMX, s = coop_launch(c, fn, nb, a));
    if (s) return s;
  }
  if (defer) {
    // the caller reads the state and the abort flag after the step
    c->pcg_deferred = true;
    return EFDC_OK;
  }
  HIP_TRY(c, hipMemcpyAsync(c->st_host, c->st, sizeof(CgState), hipMemcpyDeviceToHost,
                            c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (c->st_host->pabort) {
    if (c->pcg_last_plain) {
      // the plain launch was not co-resident: solve again, cooperatively,
      // from the P it reached (an iterate of the same system)
      ++c->pcg_retries;
      c->pcg_force_coop = true;
      s = pcg_solve(c, mixed, false);
      c->pcg_force_coop = false;
      return s;
    }
    c->err = "persistent CG: an exchange between workgroups timed out";
    return EFDC_ERR_HIP;
  }
  return EFDC_OK;
}

// Autotuner of the mixed persistent solve: the degree of the local
// preconditioner.  The solve is repeated from the same P for every
// candidate (timed with events), P is restored, and the fastest degree is
// kept for the following solves; the caller then solves once more.
int pcg_tune_degree(efdc_ctx *c) {
  int s;
  if (!c->tune_pbuf && (s = dev_alloc(c, (void **)&c->tune_pbuf, sizeof(double) * c->LCM)))
    return s;
  HIP_TRY(c, hipMemcpyAsync(c->tune_pbuf, c->dH[EFDC_P], sizeof(double) * c->LCM,
                            hipMemcpyDeviceToDevice, c->stream));
  hipEvent_t e0 = timer_event(c), e1 = timer_event(c);
  const int cand[5] = {1, 2, 3, 4, 6};
  float best_ms = 1e30f;
  int best = PCG_PDEG_DEFAULT, best_slot = -1;
  // the degrees are timed in the Chronopoulos-Gear form when the caller
  // leaves the form to the tuner
  if (c->pcg_pipe_mode == 0) c->pcg_pipe_tuned = false;
  for (int j = 0; j < 5; ++j) {
    c->pcg_pdeg_tuned = cand[j];
    float tot = 0.f;
    for (int r = 0; r < 3; ++r) {
      HIP_TRY(c, hipMemcpyAsync(c->dH[EFDC_P], c->tune_pbuf, sizeof(double) * c->LCM,
                                hipMemcpyDeviceToDevice, c->stream));
      HIP_TRY(c, hipEventRecord(e0, c->stream));
      if ((s = pcg_solve(c, true))) return s;
      HIP_TRY(c, hipEventRecord(e1, c->stream));
      HIP_TRY(c, hipEventSynchronize(e1));
      float ms = 0.f;
      HIP_TRY(c, hipEventElapsedTime(&ms, e0, e1));
      if (r > 0) tot += ms;  // the first solve is a warm-up
    }
    const float avg = tot / 2.f;
    if (c->tune_n < 8) {
      std::snprintf(c->tune_name[c->tune_n], 32, "pcg_mixed degree=%d", cand[j]);
      c->tune_ms[c->tune_n] = avg;
      ++c->tune_n;
    }
    if (avg < best_ms) {
      best_ms = avg;
      best = cand[j];
      best_slot = c->tune_n - 1;
    }
  }
  c->pcg_pdeg_tuned = best;
  // the pipelined form at that degree, when the caller leaves the choice
  if (c->pcg_pipe_mode == 0) {
    c->pcg_pipe_tuned = true;
    float tot = 0.f;
    for (int r = 0; r < 3; ++r) {
      HIP_TRY(c, hipMemcpyAsync(c->dH[EFDC_P], c->tune_pbuf, sizeof(double) * c->LCM,
                                hipMemcpyDeviceToDevice, c->stream));
      HIP_TRY(c, hipEventRecord(e0, c->stream));
      if ((s = pcg_solve(c, true))) return s;
      HIP_TRY(c, hipEventRecord(e1, c->stream));
      HIP_TRY(c, hipEventSynchronize(e1));
      float ms = 0.f;
      HIP_TRY(c, hipEventElapsedTime(&ms, e0, e1));
      if (r > 0) tot += ms;
    }
    const float avg = tot / 2.f;
    c->pcg_pipe_tuned = avg < best_ms;
    if (c->tune_n < 8) {
      std::snprintf(c->tune_name[c->tune_n], 32, "pcg_mixed pipelined d=%d", best);
      c->tune_ms[c->tune_n] = avg;
      ++c->tune_n;
      if (c->pcg_pipe_tuned) best_slot = c->tune_n - 1;
    }
  }
  if (best_slot >= 0 && best_slot < 8) c->tune_chosen[best_slot] = 1;
  c->ev_pool.push_back(e0);
  c->ev_pool.push_back(e1);
  HIP_TRY(c, hipMemcpyAsync(c->dH[EFDC_P], c->tune_pbuf, sizeof(double) * c->LCM,
                            hipMemcpyDeviceToDevice, c->stream));
  return EFDC_OK;
}
