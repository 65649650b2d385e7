// efdc_internal.h -- device context and shared device helpers of the
// gfx950 implementation.  Layout: device index i = L-1 (L is the Fortran
// cell index), wet cells are i = 1 .. LA-1, ghosts are i = 0 (L=1) and
// i = LA (L=LC).  Horizontal arrays have LCM entries; layer arrays are
// planes of `pitch` entries per K (K = 1..KC at plane K-1); W-class arrays
// (LCM,0:KCM) have plane K at offset K*pitch.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "../../include/efdc_hip.h"

#define EFDC_MAXC 16
#define EFDC_NK (EFDC_NUM_K_END - EFDC_HPK)
#define EFDC_NW (EFDC_NUM_W_END - EFDC_W2)
#define EFDC_BLOCK 256

// ------------------------------------------------------------ error glue
#define HIP_TRY(ctx, expr)                                                   \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess) {                                                  \
      (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(e_);        \
      return EFDC_ERR_HIP;                                                   \
    }                                                                        \
  } while (0)

// ------------------------------------------------------- device scalars
// State of one CG solve / one CALPUV2C call, kept on the device so that
// no host round trip is needed between iterations.
struct CgState {
  double rpcg, papcg, alpha, beta, rsq, rsqm, bnorm2, ccmnm;
  double rsq64, bnorm2_in, rz_in;  // mixed mode: outer residual, inner rhs
  int iter, done, status, symok;   // symok: coefficients verified symmetric
  int negdep, outer, inner_iter, drychg;  // drychg: a cell dried in this pass
  int fin;                         // basins' mixed solve: outer rounds ended
  int pabort;                      // persistent solve: a wait was abandoned
};

// Arrival counters of the grid reductions: one site per reduction kernel,
// each with 8 per-XCD counters and one global counter, 128 bytes apart so
// that no two counters share a cache line.
#define TICKET_PAD 32
#define TICKET_SITE (9 * TICKET_PAD)
// words of the persistent solve's barrier counters (9 counters, TICKET_PAD apart)
#define PCG_NCTR (9 * TICKET_PAD)
// cells per workgroup of the multi-kernel CG (CG_CPT per thread)
#define CG_CPT 2
#define TICKET_SITES 16
enum {
  TS_PUV_MIN = 0, TS_CG_INIT, TS_CG_A, TS_CG_B, TS_MX_PREP, TS_MX_RESID,
  TS_MX_A, TS_MX_B, TS_BALANCE
};

struct PbcDev {           // one pressure open-BC side on the device
  int n;
  int serial;             // 1: process the side with one thread
  int *L;                 // 0-based device cell index
  int *ispb;
  double *pser, *pcb, *psb;
};

struct CbcDev {           // one concentration open-BC side on the device
  int n;
  int *L;
  double *cb, *csert;     // [(LL*2+j)*nc+c], [(LL*KC+K-1)*nc+c]
  int *ntscr;
  double *clo;            // CLO*(LL,K,M)
  long long *nlo;         // NLO*(LL,K,M)
};

struct ConstDev {
  efdc_constituent_t desc;
  void *buf[2];           // CON and CON1 alternate between the two buffers
  int cur;                // buf[cur] = CON (current), buf[cur^1] = CON1
};

struct efdc_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int LA = 0, LCM = 0, KC = 0, KCM = 0, LC = 0;
  int64_t pitch = 0;      // device plane pitch of layer arrays (elements)
  efdc_params_t prm{};
  std::string err;
  int store = EFDC_STORE_FP64;
  size_t tsz = 8;         // bytes per stored layer/constituent value

  // maps, 0-based device indices
  int *dLW = nullptr, *dLE = nullptr, *dLS = nullptr, *dLN = nullptr;
  int *dLNW = nullptr, *dLSE = nullptr;
  bool rowmajor = false;  // LWC in {L-1,LC}, LEC in {L+1,LC}, maps mutual
  bool mutual = false;    // LEC(LWC(L)) = L and LNC(LSC(L)) = L for wet L
  int band = 0;           // max |neighbour index - L| over wet neighbours

  double *dH[EFDC_NUM_H] = {};
  void *dK[EFDC_NK] = {};
  void *dW[EFDC_NW] = {};
  double *staging = nullptr;   // FP64 staging of one layer array
  size_t staging_elems = 0;

  // CG scratch
  double *r = nullptr, *pb[2] = {nullptr, nullptr}, *ap = nullptr;
  double *dis = nullptr;       // 1/sqrt(CCC) (mixed mode)
  bool mx_coef_ok = false;     // dis/as32/aw32 formed for the current CCC
  float *as32 = nullptr, *aw32 = nullptr, *r32 = nullptr, *y32 = nullptr;
  float *p32[2] = {nullptr, nullptr}, *ap32 = nullptr;
  double *part = nullptr;      // reduction partials, 4 x max_blocks
  int max_blocks = 0;
  CgState *st = nullptr;       // device
  unsigned *tickets = nullptr; // TICKET_SITES x TICKET_SITE counters
  CgState *st_host = nullptr;  // pinned mirror
  int last_iters = 0;
  // independent basins (efdc_set_segments)
  int nseg = 1;
  std::vector<int> seg_first;  // first cell (0-based index i) of every basin
  int *d_segid = nullptr;      // (LCM) basin of every cell
  int *d_segfirst = nullptr;   // nseg+1 cell bounds
  int4 *d_tiles = nullptr;     // CG tiles: first cell, end cell, basin, tile in basin
  int *d_segtile0 = nullptr;   // first tile of each basin, nseg+1 entries
  int ntiles = 0;
  // the tile sets of 256, 512, 1024 and 2048 cells (d_tiles, d_segtile0
  // and ntiles are those of set seg_tset); 1 (512 cells) is the set whose
  // reduction order is that of a basin alone
  int4 *d_tiles_s[4] = {};
  int *d_segtile0_s[4] = {};
  int ntiles_s[4] = {};
  int seg_tset = 1;
  bool seg_tune_pending = false;
  double *tune_psave = nullptr;  // P saved for the CG trials (LCM)
  // replay mode of the FP64 multi-kernel solves: captured graphs or
  // launch by launch (the default); with autotune on, the next solve times
  // both from the same P and keeps the faster
  bool cg_graph = false;
  bool cg_graph_tune = false;
  // captured runs of CG iterations replayed by the FP64 multi-kernel
  // solves (0, 1, 2: one model, 8 iterations and one from the even and
  // the odd direction buffer; 3, 4, 5: the same for basins); key = the
  // launch arguments
  hipGraphExec_t cg_gexec[6] = {};
  std::vector<unsigned char> cg_gkey[6];
  CgState *d_segst = nullptr;  // per-basin CG state
  CgState *h_segst = nullptr;  // pinned mirror
  unsigned *d_segtk = nullptr; // per-basin arrival counters, TICKET_PAD apart
  unsigned *d_segdone = nullptr;  // basins whose solve has ended
  unsigned *d_segfin = nullptr;   // basins whose FP32-refined solve has ended
  double *d_segmin = nullptr;  // per-basin CCMNM
  double *d_segpart = nullptr; // CG partials, 2 x ntiles
  double *d_segbal = nullptr;  // per-basin balance sums
  // persistent solve (efdc_cgp.hip), allocated on first use
  bool pcg_ready = false;
  int pcg_ncu = 0;             // CUs, 0 when cooperative launch is unsupported
  int pcg_occ64[7] = {}, pcg_occmx[9] = {};  // workgroups per CU by CPT
  void *pcg_hx = nullptr;      // halo slots [2][ncell][4] words
  void *pcg_ps = nullptr;      // partial slots
  unsigned *pcg_flag = nullptr;  // exchange flag per workgroup
  int pcg_ncell = 0;
  unsigned pcg_launch = 0;     // launch number, gives the first tag
  unsigned *pcg_ctr = nullptr; // barrier counters
  int *pcg_abort = nullptr;     // &st->pabort (read back with the state)
  // the persistent solve's counters and abort flag are already zero (the
  // scale pass before it cleared them); the next persistent launch skips
  // its own clearing
  bool pcg_reset_done = false;
  int pcg_launch_mode = 0;      // 0 plain launch with a cooperative retry, 1 cooperative
  bool pcg64_pipe = false;      // the FP64 persistent solve in the pipelined form
  bool pcg_force_coop = false;  // the retry (and drying steps) launch cooperatively
  bool pcg_last_plain = false;  // the last persistent launch was a plain one
  unsigned pcg_spin_test = 0;   // > 0: spin limit of the plain launches (tests)
  int pcg_retries = 0;          // plain launches solved again cooperatively

  // open boundaries
  PbcDev pbc[4] = {};
  int npbc_total = 0;
  int mtide = 0;
  double *tcp_host = nullptr;  // host copy of TCP
  double *tide_dev = nullptr;  // [cos(M)..., sin(M)...] for the current call
  double *tide_host = nullptr; // pinned
  int *lopenbcdry = nullptr;   // (LCM)
  // drying (ISDRY > 0): ISCDRY, LMASKDRY, SUB1/SVB1, re-solves of the step
  int *iscdry = nullptr, *lmaskdry = nullptr;
  double *sub1 = nullptr, *svb1 = nullptr;
  double *sub_n = nullptr, *svb_n = nullptr;  // ISDRY = 99: new face switches of a pass
  int *natdry = nullptr;                        // NDRYSTP < 0: isolated steps per cell
  double *qdwaste = nullptr, *vdwaste = nullptr;  // drained rate / volume per cell
  int ncordry = 0;
  unsigned char *pbcflag = nullptr;   // (LCM) 1 on pressure-BC cells
  unsigned char *lobflag = nullptr;   // (LCM) bit 0 LOBCS, bit 1 LOBCS2
  int *lobcs = nullptr;        // LOBCS order S, W, E, N (0-based)
  int *lobcs2 = nullptr;
  int nlobcs = 0, nlobcs2 = 0;

  CbcDev cbc[4] = {};
  double *wtci = nullptr;
  int *bcrec = nullptr;        // (LCM) -1 or record index into bcinfo
  int4 *bcinfo = nullptr;      // per BC column: LL per side (-1 if none)
  int nbcrec = 0;
  double *bc_con1 = nullptr;   // BC values of CON1 written after the step
  int *bccell = nullptr;       // column index of every BC record

  int cbc_nc = 0;              // constituent count the BC tables were built for
  // sources
  int src_nc = 0;              // constituent count the source tables were built for
  int nqsij = 0;
  int *lqs = nullptr;          // 0-based
  int *srcrec = nullptr;       // (LCM) -1 or first source index at the cell
  int *srcnext = nullptr;      // next source index at the same cell or -1
  double *qss = nullptr, *qsercell = nullptr, *cqs = nullptr, *csertq = nullptr;

  // constituents
  int nconst = 0;
  ConstDev con[EFDC_MAXC] = {};
  void *tran_tmp = nullptr;    // forward-sweep storage when LDS is too small
  // anti-diffusion scratch (allocated on first use)
  void *adac[16] = {};
  bool adac_alloc = false;
  // uniform layer constants (sigma grid): CDZKK/CDZKMK per K
  bool dz_uni_kk = true, dz_uni_kmk = true;
  bool dz_uniform() const { return dz_uni_kk && dz_uni_kmk; }
  double *cdzkk_k = nullptr, *cdzkmk_k = nullptr;  // device, KC entries
  double *scal = nullptr;      // small result buffer (balance totals)
  double *nman = nullptr;      // per-cell Manning n of the closure

  // kernel timers (efdc_timers_*)
  bool timing = false;
  int64_t cg_iter_count = 0;  // TK_CG_ITERS while timing
  std::vector<hipEvent_t> ev_pool;
  struct TimerRec { int id; hipEvent_t a, b; };
  std::vector<TimerRec> trec;

  // symmetry of the scaled system (k_puv_scale's flag) as last seen by the
  // host, -1 unknown: it is fixed by the grid, the face masks and the
  // open-boundary tables (the two coefficients of a face are the same
  // expression), so the setters of those reset it
  int symok_cache = -1;
  // the persistent solve of a CALPUV2C step leaves its checks to the
  // step's final read-back (one host round trip per step)
  bool pcg_deferred = false;
  // efdc_set_source_cqs: pinned staging ring, a slot reused once the copy
  // that last read it has completed
  static constexpr int CQS_RING = 8;
  double *cqs_stage[CQS_RING] = {};
  hipEvent_t cqs_ev[CQS_RING] = {};
  size_t cqs_stage_bytes = 0;
  int cqs_next = 0;
  // efdc_get_constituent_at: cell offsets and values (device), grown on demand
  int64_t *gat_off = nullptr;
  double *gat_val = nullptr;
  int gat_cap = 0;
  // autotuner (efdc_autotune): launch shape of the column transport and
  // the preconditioner degree of the mixed persistent solve
  bool tune_pending = false, tune_cg_pending = false;
  int pcg_pdeg_tuned = 0;         // 0: not tuned
  // mixed persistent solve: 0 the autotuner picks (pipelined until it has
  // run), 1 the Chronopoulos-Gear form, 2 its pipelined form
  int pcg_pipe_mode = 0;
  bool pcg_pipe_tuned = true;     // the pipelined form (the autotuner's choice)
  double *tune_pbuf = nullptr;    // P saved while solves are timed
  int tran_ncg = 0, tran_w3 = 0;  // 0: not tuned
  int tune_n = 0, tune_best = -1;
  char tune_name[8][32] = {};
  float tune_ms[8] = {};
  int tune_chosen[8] = {};
  // discrete adjoint (efdc_adj.hip), allocated by efdc_adj_enable
  struct AdjState *adj = nullptr;
  // internal mode (efdc_uvw.hip), allocated on first use
  struct UvwState *uvw = nullptr;
  // two-level balances (efdc_bal.hip), allocated by efdc_bal2t_begin
  struct BalState *bal = nullptr;
  // restart and EE output (efdc_io.hip), allocated on first use
  struct IoState *io = nullptr;
  bool dzc_uni = false;           // DZC set and equal in every wet cell of a layer
  bool dzc_changed = false;
  std::vector<double> dzc_k;      // DZC per layer when dzc_uni
  bool has_diag = false;          // LNWC and LSEC were given
  std::vector<int> host_lw, host_le, host_ls, host_ln;  // 0-based maps
  std::vector<int> cbc_cells_host[4];  // concentration-BC cells S, W, E, N (0-based)
  std::vector<int> pbc_ispb_host[4];   // ISPB of the pressure sides W, E, S, N

  int n_wet() const { return LA - 1; }
};

// kernel ids of the timers
enum {
  TK_PUV_EXPLICIT = 0, TK_PUV_ASSEMBLE, TK_PUV_BC, TK_PUV_SCALE,
  TK_CG_INIT, TK_CG_A, TK_CG_B, TK_MX_PREP, TK_MX_RESID, TK_MX_A, TK_MX_B,
  TK_MX_APPLY, TK_PUV_POST1, TK_PUV_POST2, TK_PUV_POST3, TK_TRAN,
  TK_TRAN_BC, TK_CLOSURE_PRE, TK_CLOSURE_POST, TK_BALANCE, TK_ADAC_GRAD,
  TK_ADAC_FLUX, TK_FCT, TK_COL_FINISH, TK_PCG64, TK_PCGMX, TK_ADJ_TCOL, TK_ADJ_TGATHER,
  TK_ADJ_HYDRO, TK_ADJ_CG, TK_SETTLE,
  TK_UVW_COL, TK_UVW_W, TK_UVW_OBC, TK_UVW_HP, TK_UVW_UV, TK_IO_FMT, TK_IO_PACK,
  TK_CG_ITERS,  // not a kernel: multi-kernel CG iterations performed
  TK_COUNT
};
extern const char *const efdc_timer_names[TK_COUNT];

hipEvent_t timer_event(efdc_ctx *c);
// bracket one launch: TIMED(c, id, kernel<<<...>>>(...));
#define TIMED(c, id, ...)                                     \
  do {                                                        \
    if ((c)->timing) {                                        \
      efdc_ctx::TimerRec r_{(id), timer_event(c), timer_event(c)}; \
      (void)hipEventRecord(r_.a, (c)->stream);                \
      __VA_ARGS__;                                            \
      (void)hipEventRecord(r_.b, (c)->stream);                \
      (c)->trec.push_back(r_);                                \
    } else {                                                  \
      __VA_ARGS__;                                            \
    }                                                         \
  } while (0)

// ---------------------------------------------------------- device utils
// Workgroups are dispatched round-robin over the 8 XCDs of the MI355X.
// Remapping the hardware block id so that every XCD owns one contiguous
// range of tiles keeps the N/S neighbour rows of a tile in the same XCD's
// L2 and keeps the tiles of one XCD in that L2 between dependent kernels.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  const int x = b & 7, q = b >> 3;
  const int per = nb >> 3, rem = nb & 7;
  return x * per + (x < rem ? x : rem) + q;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// The same sum through DPP lane moves, which stay in the VALU (a shuffle is
// an LDS permute per 32-bit half and step): pairs, quads, half rows and
// rows (row_shr with bound_ctrl reads 0 from lanes before the row start),
// then row 0 into row 1 and rows 0-1 into rows 2-3 (row_bcast:15/31);
// lane 63 then holds the total, which is broadcast.  Fixed order, so every
// wave that sums the same values gets the same bits.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_move(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROWS, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWS, 0xf, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double wave_sum_dpp(double v) {
  v += dpp_move<0xb1, 0xf>(v);   // quad_perm [1,0,3,2]
  v += dpp_move<0x4e, 0xf>(v);   // quad_perm [2,3,0,1]
  v += dpp_move<0x114, 0xf>(v);  // row_shr:4
  v += dpp_move<0x118, 0xf>(v);  // row_shr:8
  v += dpp_move<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_move<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, 63);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
  return v;
}

// Sum NV values over the block; result valid in thread 0.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
  __shared__ double sh[NV][EFDC_BLOCK / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < NV; ++j) v[j] = wave_sum(v[j]);
  if (lane == 0) {
#pragma unroll
    for (int j = 0; j < NV; ++j) sh[j][w] = v[j];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      double s = 0.0;
      for (int q = 0; q < (int)(blockDim.x >> 6); ++q) s += sh[j][q];
      v[j] = s;
    }
  }
  __syncthreads();
}

// Cross-workgroup publication without L2 write-backs.  On the multi-XCD
// gfx950 every XCD has its own L2; an agent-scope release fence would write
// back the whole L2 of the issuing XCD, once per workgroup.  Instead the
// partial results are published with agent-scope atomic stores (written
// through to the coherent memory side), the issuing wave waits for their
// completion (s_waitcnt), and only then takes the ticket with an
// agent-scope atomic add.  The last workgroup reads the partials with
// agent-scope atomic loads, which are served coherently.
__device__ __forceinline__ void publish(double *p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double fetch(const double *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned atomic_inc(unsigned *t) {
  return __hip_atomic_fetch_add(t, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void reset_ticket(unsigned *t) {
  __hip_atomic_store(t, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Two-level arrival: a block counts in on the counter of its XCD (blocks
// are dispatched round-robin, so XCD x runs the hardware blocks b with
// b % 8 == x); the last block of each XCD counts in on the global
// counter.  Returns true in the one block that arrives last overall.
// Spreading the same-address atomics over 9 cache lines removes their
// serialisation.
__device__ __forceinline__ bool last_arrival(unsigned *site, int nblocks) {
  __builtin_amdgcn_s_waitcnt(0);  // the published partials have completed
  const int x = blockIdx.x & 7;
  const int nx = (nblocks >> 3) + (x < (nblocks & 7) ? 1 : 0);
  const int ngroups = nblocks < 8 ? nblocks : 8;
  unsigned *gx = site + x * TICKET_PAD, *gg = site + 8 * TICKET_PAD;
  if (atomic_inc(gx) != (unsigned)nx - 1u) return false;
  reset_ticket(gx);
  __builtin_amdgcn_s_waitcnt(0);
  return atomic_inc(gg) == (unsigned)ngroups - 1u;
}

// Sum of the partials part[j*stride + q], q < nb, j < NV, by one block in
// a fixed order: thread t adds q = t, t+B, t+2B, ... (loads of 8 terms in
// flight), then a fixed tree over the block.  Result valid in thread 0.
template <int NV>
__device__ __forceinline__ void sum_parts(const double *part, int stride, int nb,
                                          double (&acc)[NV]) {
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    double s = 0.0;
    const double *pj = part + (size_t)j * stride;
    for (int q0 = threadIdx.x; q0 < nb; q0 += 8 * (int)blockDim.x) {
      double a[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int q = q0 + u * (int)blockDim.x;
        a[u] = q < nb ? fetch(&pj[q]) : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) s += a[u];
    }
    acc[j] = s;
  }
  block_sum<NV>(acc);
}

// Deterministic grid reduction: every block publishes its partials, the
// last block to arrive (atomic ticket) sums the partials in block order
// with a fixed tree, so the result does not depend on scheduling.
// Returns true in every thread of the last block; the sums are then in
// out[0..NV-1] (valid in all threads of that block).  The ticket is reset
// for reuse by the next launch.
template <int NV>
__device__ __forceinline__ bool grid_sum(double (&v)[NV], double *part,
                                         unsigned *ticket, int nblocks,
                                         int bid, double (&out)[NV]) {
  __shared__ bool am_last;
  block_sum<NV>(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < NV; ++j) publish(&part[(size_t)j * nblocks + bid], v[j]);
    am_last = last_arrival(ticket, nblocks);
  }
  __syncthreads();
  if (!am_last) return false;
  double acc[NV];
  sum_parts<NV>(part, nblocks, nblocks, acc);
  __shared__ double res[NV];
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < NV; ++j) res[j] = acc[j];
    reset_ticket(ticket + 8 * TICKET_PAD);
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < NV; ++j) out[j] = res[j];
  return true;
}

// Two-level deterministic grid reduction for launches of many blocks whose
// hardware block b runs tile xcd_tile(b, nblocks): the last block of each
// XCD to arrive sums that XCD's (contiguous) tiles in order and publishes
// the XCD sum, and the last XCD to finish sums the 8 XCD sums in order.
// The serial tail is 1/8 of grid_sum's and the 8 XCD sums run at the same
// time.  The XCD sums are kept after the partials, at part[NV*nblocks +
// j*8 + x].  Returns true in every thread of the block that finishes.
template <int NV>
__device__ __forceinline__ bool grid_sum_x(double (&v)[NV], double *part, unsigned *ticket,
                                           int nblocks, int tile, double (&out)[NV]) {
  __shared__ int stage;  // 0 not last, 1 last of its XCD, 2 last overall
  block_sum<NV>(v);
  const int x = blockIdx.x & 7;
  const int per = nblocks >> 3, rem = nblocks & 7;
  const int nx = per + (x < rem ? 1 : 0);
  const int ngroups = nblocks < 8 ? nblocks : 8;
  unsigned *gx = ticket + x * TICKET_PAD, *gg = ticket + 8 * TICKET_PAD;
  double *xpart = part + (size_t)NV * nblocks;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < NV; ++j) publish(&part[(size_t)j * nblocks + tile], v[j]);
    __builtin_amdgcn_s_waitcnt(0);
    stage = (atomic_inc(gx) == (unsigned)nx - 1u) ? 1 : 0;
    if (stage) reset_ticket(gx);
  }
  __syncthreads();
  if (!stage) return false;
  // this XCD's tiles: [t0, t0 + nx)
  const int t0 = x * per + (x < rem ? x : rem);
  double acc[NV];
  sum_parts<NV>(part + t0, nblocks, nx, acc);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < NV; ++j) publish(&xpart[j * 8 + x], acc[j]);
    __builtin_amdgcn_s_waitcnt(0);
    stage = (atomic_inc(gg) == (unsigned)ngroups - 1u) ? 2 : 0;
    if (stage == 2) reset_ticket(gg);
  }
  __syncthreads();
  if (stage != 2) return false;
  __shared__ double res[NV];
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      double t = 0.0;
      for (int q = 0; q < ngroups; ++q) t += fetch(&xpart[j * 8 + q]);
      res[j] = t;
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < NV; ++j) out[j] = res[j];
  return true;
}

// Same pattern for a minimum.
__device__ __forceinline__ bool grid_min(double v, double *part,
                                         unsigned *ticket, int nblocks,
                                         int bid, double &out) {
  __shared__ double sh[EFDC_BLOCK / 64];
  __shared__ bool am_last;
  __shared__ double res;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  v = wave_min(v);
  if (lane == 0) sh[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = sh[0];
    for (int q = 1; q < (int)(blockDim.x >> 6); ++q) m = fmin(m, sh[q]);
    publish(&part[bid], m);
    am_last = last_arrival(ticket, nblocks);
  }
  __syncthreads();
  if (!am_last) return false;
  double m = 1.0e18;
  for (int q0 = threadIdx.x; q0 < nblocks; q0 += 8 * (int)blockDim.x) {
    double a[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int q = q0 + u * (int)blockDim.x;
      a[u] = q < nblocks ? fetch(&part[q]) : 1.0e18;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) m = fmin(m, a[u]);
  }
  m = wave_min(m);
  if (lane == 0) sh[w] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double mm = sh[0];
    for (int q = 1; q < (int)(blockDim.x >> 6); ++q) mm = fmin(mm, sh[q]);
    res = mm;
    reset_ticket(ticket + 8 * TICKET_PAD);
  }
  __syncthreads();
  out = res;
  return true;
}

// ------------------------------------------------------------ host glue
// hipMalloc + zero fill on the context's stream
inline int dev_alloc(efdc_ctx *c, void **p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) return EFDC_OK;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    c->err = std::string("hipMalloc failed: ") + hipGetErrorString(e);
    *p = nullptr;
    return EFDC_ERR_NOMEM;
  }
  e = hipMemsetAsync(*p, 0, bytes, c->stream);
  if (e != hipSuccess) {
    c->err = hipGetErrorString(e);
    return EFDC_ERR_HIP;
  }
  return EFDC_OK;
}
int efdc_grid_blocks(int n);
// the reason of a failed context-free call (efdc_last_error(NULL))
void efdc_thread_error(const char *why);
// persistent CONGRAD (efdc_cgp.hip)
bool pcg_applicable(efdc_ctx *c, bool mixed);
int pcg_solve(efdc_ctx *c, bool mixed, bool defer = false);
int pcg_tune_degree(efdc_ctx *c);
// independent basins (efdc_seg.hip)
int seg_min_run(efdc_ctx *c);
int seg_cg_run(efdc_ctx *c, efdc_cg_stats_t *stats, bool *ran_mixed = nullptr);
void seg_free(efdc_ctx *c);
// discrete adjoint (efdc_adj.hip): forward-step records and clean-up
bool adj_recording(const efdc_ctx *c);
int adj_record_ps(efdc_ctx *c);
int adj_record_c0(efdc_ctx *c);
int adj_record_settle(efdc_ctx *c, int slot, const void *con);
int adj_record_adac(efdc_ctx *c, int slot, const void *con);
// Reverse of the anti-diffusion / FCT passes of one constituent
// (efdc_tran.hip); the planes are FP64 with the layer pitch (KC+1 planes
// for the vertical-face fields), the hydrodynamic lambdas those of the
// step's state.
struct AdacAdj {
  const void *conu, *con1;   // recorded CON after the upwind step, CON1
  const double *lam;         // lambda of CON after the column finish
  double *lconu, *lcon1;     // out: lambda of CON after the upwind step, of CON1
  double *lcd, *lfu, *lfv, *lfw, *lduu, *ldvv, *lu, *lv, *lww;  // work
  double *wv, *wvcu, *weeb, *ww;                                 // work (KC planes)
  double *ahp, *au_c, *au_p, *av_c, *av_p, *ahu, *ahv;
  const double *hu, *hv, *uhdye, *uhdy1e, *vhdxe, *vhdx1e, *dyiu, *dxiv;
  int dye_mode;
  double dye_f;
};
int adj_adac_run(efdc_ctx *c, int slot, const AdacAdj &io, const efdc_step_t *step);
void adj_free(efdc_ctx *c);
// FP64 CG on the context's scaled operator for another right-hand side
// and solution vector (x holds the initial guess), with its own state
int cg_fp64_vec(efdc_ctx *c, double *x, const double *rhs, CgState *st, CgState *st_host,
                double rel_tol, int itmax, int *iters, double *rsq);
int puv_run(efdc_ctx *c, const efdc_step_t *step, efdc_cg_stats_t *stats);
// internal mode (efdc_uvw.hip)
bool uvw_id(int id);
int uvw_ptr(efdc_ctx *c, int id, double **p, int *planes);  // planes 0: (LCM)
void uvw_step_begin(efdc_ctx *c);
void uvw_state_pieces(efdc_ctx *c, std::vector<std::pair<void *, size_t>> &v);
void uvw_obc_changed(efdc_ctx *c);
void uvw_free(efdc_ctx *c);
int uvw_run(efdc_ctx *c, const efdc_step_t *step, const efdc_uvw_opts_t *o);
// two-level balances (efdc_bal.hip): the terms CALCONC adds while a
// period is open
int bal_before_tran(efdc_ctx *c, const efdc_step_t *step);
int bal_after_tran(efdc_ctx *c, const efdc_step_t *step);
void bal_free(efdc_ctx *c);
// basins (efdc_seg.hip): make tile set q (256 << q cells) the current one
void seg_select_tiles(efdc_ctx *c, int q);
// captured CG iterations (efdc_puv.hip)
void cg_graphs_free(efdc_ctx *c);
// restart and EE output (efdc_io.hip)
void io_free(efdc_ctx *c);
int cg_run(efdc_ctx *c, bool check_sym, efdc_cg_stats_t *stats);
int tran_run(efdc_ctx *c, const efdc_step_t *step);
int balance_run(efdc_ctx *c, efdc_balance_t *out);
int closure_pre_run(efdc_ctx *c, const double *nman_dev);
int closure_post_run(efdc_ctx *c);
