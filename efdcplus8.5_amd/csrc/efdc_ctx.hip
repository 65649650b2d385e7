// efdc_ctx.hip -- C-ABI entry points: context lifetime, host<->device
// copies in the var_global layout, open-boundary / source / constituent
// tables.  The compute entry points are in efdc_puv.hip and efdc_tran.hip.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <cstdio>
#include <set>

#include "efdc_internal.h"

namespace {

template <typename T>
__global__ void k_fill(T *a, T v, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

template <typename D, typename S>
__global__ void k_convert(D *dst, const S *src, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (D)src[i];
}

// first-call initialisation of CALPUV2C: RCX=RCY=1 on wet cells, 0 on the
// ghosts, CC=1 (EFDC/calpuv2c.f90:88-108)
__global__ void k_init_puv(double *rcx, double *rcy, double *cc, int lcm,
                           int lc0) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= lcm) return;
  const bool ghost = (i == 0 || i == lc0);
  rcx[i] = ghost ? 0.0 : 1.0;
  rcy[i] = ghost ? 0.0 : 1.0;
  cc[i] = 1.0;
}

// Plain streaming copy (16-byte loads and stores, grid-stride): the
// achievable-bandwidth reference of the roofline figures.
__global__ __launch_bounds__(256) void k_copy4(const float4 *__restrict__ x,
                                               float4 *__restrict__ y, size_t n4) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
       i += (size_t)gridDim.x * blockDim.x)
    y[i] = x[i];
}

int blocks_for(int64_t n, int bs = EFDC_BLOCK) { return (int)((n + bs - 1) / bs); }

template <typename T>
__global__ void k_gather(const T *src, const int64_t *off, double *val, int n) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) val[j] = (double)src[off[j]];
}

template <typename T>
int dmalloc(efdc_ctx *c, T **p, size_t n) {
  *p = nullptr;
  if (n == 0) return EFDC_OK;
  hipError_t e = hipMalloc((void **)p, n * sizeof(T));
  if (e != hipSuccess) {
    c->err = std::string("hipMalloc failed: ") + hipGetErrorString(e);
    *p = nullptr;
    return EFDC_ERR_NOMEM;
  }
  e = hipMemsetAsync(*p, 0, n * sizeof(T), c->stream);
  if (e != hipSuccess) {
    c->err = hipGetErrorString(e);
    return EFDC_ERR_HIP;
  }
  return EFDC_OK;
}

template <typename T>
void dfree(T *&p) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

#define TRY(x)                 \
  do {                         \
    int s_ = (x);              \
    if (s_ != EFDC_OK) return s_; \
  } while (0)

int upload(efdc_ctx *c, void *dst, const void *src, size_t bytes) {
  if (bytes == 0) return EFDC_OK;
  HIP_TRY(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  return EFDC_OK;
}

template <typename T>
int upload_new(efdc_ctx *c, T **dst, const T *src, size_t n) {
  dfree(*dst);
  if (n == 0) return EFDC_OK;
  TRY(dmalloc(c, dst, n));
  return upload(c, *dst, src, n * sizeof(T));
}

bool valid_id(int id) {
  return (id >= 0 && id < EFDC_NUM_H) ||
         (id >= EFDC_HPK && id < EFDC_NUM_K_END) ||
         (id >= EFDC_W2 && id < EFDC_NUM_W_END) || uvw_id(id);
}

}  // namespace

int efdc_grid_blocks(int n) { return blocks_for(n); }

const char *const efdc_timer_names[TK_COUNT] = {
    "puv_explicit", "puv_assemble", "puv_openbc", "puv_scale", "cg_init",
    "cg_spmv_pupdate", "cg_axpy_dots", "mx_prep", "mx_residual64",
    "mx_spmv_pupdate32", "mx_axpy_dots32", "mx_apply", "puv_transports",
    "puv_depths", "puv_hu_hv", "tran_upwind_vdiff", "tran_bc_con1",
    "closure_pre", "closure_post", "balance", "adac_gradients",
    "adac_fluxes", "fct_limiters", "tran_finish_vdiff", "cg_persistent64",
    "mx_persistent", "adj_tran_col", "adj_tran_gather", "adj_hydro", "adj_cg", "settle",
    "uvw_columns", "uvw_w_averages", "uvw_openbc", "uvw_depths", "uvw_uv",
    "restart_format", "ee_pack", "cg_iterations"};

hipEvent_t timer_event(efdc_ctx *c) {
  if (c->ev_pool.empty()) {
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
  }
  hipEvent_t e = c->ev_pool.back();
  c->ev_pool.pop_back();
  return e;
}

extern "C" {

int efdc_abi_version(void) { return EFDC_ABI_VERSION; }

void efdc_params_default(efdc_params_t *p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->G = 9.81;
  p->RITB1 = 1.0;
  p->SNLT = 0.0;
  p->HDRY = 0.2;
  p->TCON_TBEGIN = 0.0;
  p->ITERM = 200;
  p->RSQM = 1e-9;
  p->cg_rel_tol = 0.0;
  p->cg_mode = EFDC_CG_FP64;
  p->store = EFDC_STORE_FP64;
  p->mixed_inner_rtol = 1e-5;
  p->mixed_max_outer = 30;
  p->cg_kernel = EFDC_CG_KERNEL_AUTO;
  p->mixed_precond_degree = 0;
  p->tran_shape = 0;
  p->ISDRY = 0;
  p->NDRYSTP = 0;
}

static bool tran_shape_valid(int q) {
  if (q == 0 || q == -1 || q == -2 || q == -4) return true;
  const int n = q & 15, v = q >> 4;
  return q > 0 && (n == 1 || n == 2 || n == 4) && (v == 0 || v == 1 || (v == 2 && n <= 2));
}

// Reason of the last failed efdc_ctx_create on this thread (there is no
// context to hold it); efdc_last_error(NULL) returns it.
static thread_local std::string g_create_err = "no error";

}  // extern "C"
void efdc_thread_error(const char *why) { g_create_err = why; }
extern "C" {

const char *efdc_last_error(const efdc_ctx_t *ctx) {
  return ctx ? ctx->err.c_str() : g_create_err.c_str();
}

int efdc_ctx_create(const efdc_grid_t *g, const efdc_params_t *prm,
                    efdc_ctx_t **out) {
  auto reject = [](int st, const char *why) {
    g_create_err = why;
    return st;
  };
  if (!out) return reject(EFDC_ERR_ARG, "null output pointer");
  *out = nullptr;
  if (!g || !prm) return reject(EFDC_ERR_ARG, "null grid or params");
  if (g->LA < 2 || g->LCM < g->LA + 1 || g->KC < 1 || g->KCM < g->KC ||
      g->KC > 64 || !g->LWC || !g->LEC || !g->LSC || !g->LNC)
    return reject(EFDC_ERR_ARG, "grid sizes or maps invalid (need LA>=2, "
                                "LCM>=LA+1, 1<=KC<=64, KCM>=KC, four maps)");
  if (prm->store != EFDC_STORE_FP64 && prm->store != EFDC_STORE_FP32)
    return reject(EFDC_ERR_ARG, "unknown storage precision");
  if (prm->cg_mode != EFDC_CG_FP64 && prm->cg_mode != EFDC_CG_MIXED)
    return reject(EFDC_ERR_ARG, "unknown CG mode");
  if (prm->cg_kernel < EFDC_CG_KERNEL_AUTO || prm->cg_kernel > EFDC_CG_KERNEL_PERSISTENT)
    return reject(EFDC_ERR_ARG, "unknown CG kernel choice");
  if (prm->mixed_precond_degree < 0 || prm->mixed_precond_degree > 16)
    return reject(EFDC_ERR_ARG, "mixed_precond_degree must be 0..16");
  if (prm->ISDRY < 0 || prm->ISDRY > 99)
    return reject(EFDC_ERR_ARG, "ISDRY must be 0..99");
  if (!tran_shape_valid(prm->tran_shape))
    return reject(EFDC_ERR_ARG, "tran_shape must be 0, -1, -2, -4, or 1, 2, 4 plus 0, 16 or 32 (32 for 1 and 2)");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return reject(EFDC_ERR_NODEVICE, "no HIP device visible");
  efdc_ctx *c = new efdc_ctx();
  (void)hipGetDevice(&c->device);
  c->LA = g->LA;
  c->LCM = g->LCM;
  c->KC = g->KC;
  c->KCM = g->KCM;
  c->LC = g->LA + 1;
  c->pitch = ((int64_t)g->LCM + 63) / 64 * 64;
  c->prm = *prm;
  c->store = prm->store;
  c->tsz = prm->store == EFDC_STORE_FP32 ? 4 : 8;
  auto fail = [&](int s) {
    g_create_err = c->err.empty() ? "device set-up failed" : c->err;
    efdc_ctx_destroy(c);
    return s;
  };
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return fail(EFDC_ERR_HIP);

  // maps: validate on the host, store 0-based
  const int LA = g->LA, LC = c->LC, LCM = g->LCM;
  std::vector<int> lw(LCM), le(LCM), ls(LCM), ln(LCM), lnw(LCM), lse(LCM);
  bool rowmajor = true, mutual = true;
  int band = 1;
  for (int L = 1; L <= LCM; ++L) {
    const int i = L - 1;
    auto chk = [&](int v) { return v >= 1 && v <= LCM; };
    int a = g->LWC[i], b = g->LEC[i], s = g->LSC[i], n = g->LNC[i];
    if (L >= 2 && L <= LA) {
      if (!chk(a) || !chk(b) || !chk(s) || !chk(n)) {
        c->err = "neighbour map value out of range at L=" + std::to_string(L);
        return fail(EFDC_ERR_ARG);
      }
      if (!(a == L - 1 || a == LC) || !(b == L + 1 || b == LC)) rowmajor = false;
      // mutual: a wet neighbour in one direction sees L in the opposite one
      // (and no neighbour is the dummy cell 1)
      if (a == 1 || b == 1 || s == 1 || n == 1) mutual = false;
      if (a != LC && a >= 2 && a <= LA && g->LEC[a - 1] != L) mutual = false;
      if (s != LC && s >= 2 && s <= LA && g->LNC[s - 1] != L) mutual = false;
      if (b != LC && b >= 2 && b <= LA && g->LWC[b - 1] != L) mutual = false;
      if (n != LC && n >= 2 && n <= LA && g->LSC[n - 1] != L) mutual = false;
      for (int v : {a, b, s, n})
        if (v >= 2 && v <= LA) band = std::max(band, std::abs(v - L));
    }
    lw[i] = chk(a) ? a - 1 : LC - 1;
    le[i] = chk(b) ? b - 1 : LC - 1;
    ls[i] = chk(s) ? s - 1 : LC - 1;
    ln[i] = chk(n) ? n - 1 : LC - 1;
    lnw[i] = (g->LNWC && chk(g->LNWC[i])) ? g->LNWC[i] - 1 : LC - 1;
    lse[i] = (g->LSEC && chk(g->LSEC[i])) ? g->LSEC[i] - 1 : LC - 1;
  }
  c->rowmajor = rowmajor && mutual;
  c->mutual = mutual;
  c->has_diag = g->LNWC != nullptr && g->LSEC != nullptr;
  c->host_lw = lw;
  c->host_le = le;
  c->host_ls = ls;
  c->host_ln = ln;
  c->band = band;
  int s;
  if ((s = upload_new(c, &c->dLW, lw.data(), LCM)) ||
      (s = upload_new(c, &c->dLE, le.data(), LCM)) ||
      (s = upload_new(c, &c->dLS, ls.data(), LCM)) ||
      (s = upload_new(c, &c->dLN, ln.data(), LCM)) ||
      (s = upload_new(c, &c->dLNW, lnw.data(), LCM)) ||
      (s = upload_new(c, &c->dLSE, lse.data(), LCM)))
    return fail(s);

  for (int id = 0; id < EFDC_NUM_H; ++id)
    if ((s = dmalloc(c, &c->dH[id], (size_t)LCM))) return fail(s);
  const size_t kel = (size_t)c->pitch * (size_t)c->KC;
  const size_t wel = (size_t)c->pitch * (size_t)(c->KC + 1);
  for (int j = 0; j < EFDC_NK; ++j) {
    char *p = nullptr;
    if ((s = dmalloc(c, &p, kel * c->tsz))) return fail(s);
    c->dK[j] = p;
  }
  for (int j = 0; j < EFDC_NW; ++j) {
    char *p = nullptr;
    if ((s = dmalloc(c, &p, wel * c->tsz))) return fail(s);
    c->dW[j] = p;
  }
  c->staging_elems = wel;
  if ((s = dmalloc(c, &c->staging, wel))) return fail(s);

  // CG scratch
  c->max_blocks = blocks_for(LCM) + 8;
  if ((s = dmalloc(c, &c->r, LCM)) || (s = dmalloc(c, &c->pb[0], LCM)) ||
      (s = dmalloc(c, &c->pb[1], LCM)) || (s = dmalloc(c, &c->ap, LCM)) ||
      (s = dmalloc(c, &c->part, (size_t)8 * c->max_blocks)) ||
      (s = dmalloc(c, &c->st, 1)) ||
      (s = dmalloc(c, &c->tickets, (size_t)TICKET_SITES * TICKET_SITE)))
    return fail(s);
  if (prm->cg_mode == EFDC_CG_MIXED) {
    if ((s = dmalloc(c, &c->dis, LCM)) || (s = dmalloc(c, &c->as32, LCM)) ||
        (s = dmalloc(c, &c->aw32, LCM)) || (s = dmalloc(c, &c->r32, LCM)) ||
        (s = dmalloc(c, &c->y32, LCM)) || (s = dmalloc(c, &c->p32[0], LCM)) ||
        (s = dmalloc(c, &c->p32[1], LCM)) || (s = dmalloc(c, &c->ap32, LCM)))
      return fail(s);
  }
  if (hipHostMalloc((void **)&c->st_host, sizeof(CgState)) != hipSuccess)
    return fail(EFDC_ERR_NOMEM);
  std::memset(c->st_host, 0, sizeof(CgState));
  if ((s = dmalloc(c, &c->iscdry, LCM)) || (s = dmalloc(c, &c->lmaskdry, LCM)) ||
      (s = dmalloc(c, &c->sub1, LCM)) || (s = dmalloc(c, &c->svb1, LCM)) ||
      (s = dmalloc(c, &c->lopenbcdry, LCM)) || (s = dmalloc(c, &c->natdry, LCM)) ||
      (s = dmalloc(c, &c->qdwaste, LCM)) || (s = dmalloc(c, &c->vdwaste, LCM)) ||
      (s = dmalloc(c, &c->pbcflag, LCM)) || (s = dmalloc(c, &c->lobflag, LCM)) ||
      (s = dmalloc(c, &c->bcrec, LCM)) ||
      (s = dmalloc(c, &c->srcrec, LCM)) || (s = dmalloc(c, &c->cdzkk_k, 64)) ||
      (s = dmalloc(c, &c->cdzkmk_k, 64)) || (s = dmalloc(c, &c->scal, 64)) ||
      (s = dmalloc(c, &c->nman, LCM)))
    return fail(s);
  const int nb = blocks_for(LCM);
  k_fill<int><<<nb, EFDC_BLOCK, 0, c->stream>>>(c->bcrec, -1, LCM);
  k_fill<int><<<nb, EFDC_BLOCK, 0, c->stream>>>(c->srcrec, -1, LCM);
  k_fill<int><<<nb, EFDC_BLOCK, 0, c->stream>>>(c->lmaskdry, 1, LCM);
  k_init_puv<<<nb, EFDC_BLOCK, 0, c->stream>>>(c->dH[EFDC_RCX], c->dH[EFDC_RCY],
                                               c->dH[EFDC_CC], LCM, LC - 1);
  if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(EFDC_ERR_HIP);
  *out = c;
  return EFDC_OK;
}

int efdc_ctx_destroy(efdc_ctx_t *c) {
  if (!c) return EFDC_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  dfree(c->dLW); dfree(c->dLE); dfree(c->dLS); dfree(c->dLN);
  dfree(c->dLNW); dfree(c->dLSE);
  for (auto &p : c->dH) dfree(p);
  for (auto &p : c->dK) dfree(p);
  for (auto &p : c->dW) dfree(p);
  dfree(c->staging);
  dfree(c->r); dfree(c->pb[0]); dfree(c->pb[1]); dfree(c->ap);
  dfree(c->dis); dfree(c->as32); dfree(c->aw32); dfree(c->r32); dfree(c->y32);
  dfree(c->p32[0]); dfree(c->p32[1]); dfree(c->ap32);
  dfree(c->part); dfree(c->st); dfree(c->tickets);
  { char *q = (char *)c->pcg_hx; dfree(q); }
  { char *q = (char *)c->pcg_ps; dfree(q); }
  dfree(c->pcg_ctr); dfree(c->pcg_flag);
  seg_free(c);
  adj_free(c);
  uvw_free(c);
  bal_free(c);
  io_free(c);
  cg_graphs_free(c);
  dfree(c->tune_psave);
  if (c->st_host) (void)hipHostFree(c->st_host);
  dfree(c->gat_off);
  dfree(c->gat_val);
  for (int q = 0; q < efdc_ctx::CQS_RING; ++q) {
    if (c->cqs_stage[q]) (void)hipHostFree(c->cqs_stage[q]);
    if (c->cqs_ev[q]) (void)hipEventDestroy(c->cqs_ev[q]);
  }
  dfree(c->tune_pbuf);
  if (c->tide_host) (void)hipHostFree(c->tide_host);
  for (auto &b : c->pbc) {
    dfree(b.L); dfree(b.ispb); dfree(b.pser); dfree(b.pcb); dfree(b.psb);
  }
  delete[] c->tcp_host;
  dfree(c->iscdry); dfree(c->lmaskdry); dfree(c->sub1); dfree(c->svb1);
  dfree(c->sub_n); dfree(c->svb_n);
  dfree(c->natdry); dfree(c->qdwaste); dfree(c->vdwaste);
  dfree(c->tide_dev); dfree(c->lopenbcdry); dfree(c->pbcflag); dfree(c->lobflag);
  dfree(c->lobcs); dfree(c->lobcs2);
  for (auto &b : c->cbc) {
    dfree(b.L); dfree(b.cb); dfree(b.csert); dfree(b.ntscr); dfree(b.clo);
    dfree(b.nlo);
  }
  dfree(c->wtci); dfree(c->bcrec); dfree(c->bcinfo); dfree(c->bc_con1);
  dfree(c->lqs); dfree(c->srcrec); dfree(c->srcnext); dfree(c->qss);
  dfree(c->qsercell); dfree(c->cqs); dfree(c->csertq);
  for (int j = 0; j < EFDC_MAXC; ++j) {
    char *a = (char *)c->con[j].buf[0], *b = (char *)c->con[j].buf[1];
    dfree(a); dfree(b);
  }
  { char *t = (char *)c->tran_tmp; dfree(t); }
  for (auto &p : c->adac) { char *q = (char *)p; dfree(q); p = nullptr; }
  dfree(c->cdzkk_k); dfree(c->cdzkmk_k); dfree(c->scal); dfree(c->nman);
  dfree(c->bccell);
  for (auto &r : c->trec) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
  for (auto &e : c->ev_pool) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return EFDC_OK;
}

int efdc_set_params(efdc_ctx_t *c, const efdc_params_t *p) {
  if (!c || !p) return EFDC_ERR_ARG;
  c->symok_cache = -1;
  if (p->store != c->store) {
    c->err = "the storage precision is fixed at context creation";
    return EFDC_ERR_ARG;
  }
  if (p->cg_mode == EFDC_CG_MIXED && !c->dis) {
    c->err = "EFDC_CG_MIXED needs a context created in that mode";
    return EFDC_ERR_ARG;
  }
  if (p->mixed_precond_degree < 0 || p->mixed_precond_degree > 16) {
    c->err = "mixed_precond_degree must be 0..16";
    return EFDC_ERR_ARG;
  }
  if (p->ISDRY < 0 || p->ISDRY > 99) {
    c->err = "ISDRY must be 0..99";
    return EFDC_ERR_ARG;
  }
  if (!tran_shape_valid(p->tran_shape)) {
    c->err = "tran_shape must be 0, -1, -2, -4, or 1, 2, 4 plus 0, 16 or 32 (32 for 1 and 2)";
    return EFDC_ERR_ARG;
  }
  if (p->cg_kernel < EFDC_CG_KERNEL_AUTO || p->cg_kernel > EFDC_CG_KERNEL_PERSISTENT) {
    c->err = "unknown CG kernel choice";
    return EFDC_ERR_ARG;
  }
  c->prm = *p;
  return EFDC_OK;
}

int efdc_set_array(efdc_ctx_t *c, int32_t id, const double *host) {
  if (!c || !host || !valid_id(id)) return EFDC_ERR_ARG;
  c->symok_cache = -1;
  if (uvw_id(id)) {
    double *d = nullptr;
    int planes = 0, s;
    if ((s = uvw_ptr(c, id, &d, &planes))) return s;
    if (planes == 0)
      HIP_TRY(c, hipMemcpyAsync(d, host, sizeof(double) * c->LCM, hipMemcpyHostToDevice,
                                c->stream));
    else
      HIP_TRY(c, hipMemcpy2DAsync(d, sizeof(double) * c->pitch, host, sizeof(double) * c->LCM,
                                  sizeof(double) * c->LCM, planes, hipMemcpyHostToDevice,
                                  c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return EFDC_OK;
  }
  if (id == EFDC_DZC) {
    // sigma layers: one thickness per layer in every wet cell (the internal
    // mode's face constants, EFDC/aaefdc.f90:2001-2126, are then per layer)
    bool uni = true;
    std::vector<double> perk(64, 0.0);
    for (int k = 0; k < c->KC && uni; ++k) {
      const double *row = host + (size_t)k * c->LCM;
      perk[k] = row[1];
      for (int L = 2; L <= c->LA; ++L)
        if (row[L - 1] != perk[k]) { uni = false; break; }
    }
    c->dzc_uni = uni;
    c->dzc_k = perk;
    c->dzc_changed = true;
  }
  if (id < EFDC_NUM_H) {
    HIP_TRY(c, hipMemcpyAsync(c->dH[id], host, sizeof(double) * c->LCM,
                              hipMemcpyHostToDevice, c->stream));
  } else {
    const bool isw = id >= EFDC_W2;
    const int planes = isw ? c->KC + 1 : c->KC;
    void *dst = isw ? c->dW[id - EFDC_W2] : c->dK[id - EFDC_HPK];
    double *d64 = c->store == EFDC_STORE_FP64 ? (double *)dst : c->staging;
    HIP_TRY(c, hipMemcpy2DAsync(d64, sizeof(double) * c->pitch, host,
                                sizeof(double) * c->LCM, sizeof(double) * c->LCM,
                                planes, hipMemcpyHostToDevice, c->stream));
    if (c->store == EFDC_STORE_FP32) {
      const int64_t n = c->pitch * planes;
      k_convert<float, double><<<blocks_for(n), EFDC_BLOCK, 0, c->stream>>>(
          (float *)dst, c->staging, n);
    }
    if (id == EFDC_CDZKK || id == EFDC_CDZKMK) {
      // sigma grids carry the same layer constants in every column; keep
      // one FP64 value per K then (EFDC/aaefdc.f90:1226-1232)
      bool uni = true;
      std::vector<double> perk(64, 0.0);
      for (int k = 0; k < c->KC && uni; ++k) {
        const double *row = host + (size_t)k * c->LCM;
        perk[k] = row[1];
        for (int L = 2; L <= c->LA; ++L)
          if (row[L - 1] != perk[k]) { uni = false; break; }
      }
      if (uni)
        HIP_TRY(c, hipMemcpyAsync(id == EFDC_CDZKK ? c->cdzkk_k : c->cdzkmk_k,
                                  perk.data(), sizeof(double) * 64,
                                  hipMemcpyHostToDevice, c->stream));
      (id == EFDC_CDZKK ? c->dz_uni_kk : c->dz_uni_kmk) = uni;
    }
  }
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return EFDC_OK;
}

int efdc_get_array(efdc_ctx_t *c, int32_t id, double *host) {
  if (!c || !host || !valid_id(id)) return EFDC_ERR_ARG;
  if (uvw_id(id)) {
    double *d = nullptr;
    int planes = 0, s;
    if ((s = uvw_ptr(c, id, &d, &planes))) return s;
    if (planes == 0)
      HIP_TRY(c, hipMemcpyAsync(host, d, sizeof(double) * c->LCM, hipMemcpyDeviceToHost,
                                c->stream));
    else
      HIP_TRY(c, hipMemcpy2DAsync(host, sizeof(double) * c->LCM, d, sizeof(double) * c->pitch,
                                  sizeof(double) * c->LCM, planes, hipMemcpyDeviceToHost,
                                  c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return EFDC_OK;
  }
  if (id < EFDC_NUM_H) {
    HIP_TRY(c, hipMemcpyAsync(host, c->dH[id], sizeof(double) * c->LCM,
                              hipMemcpyDeviceToHost, c->stream));
  } else {
    const bool isw = id >= EFDC_W2;
    const int planes = isw ? c->KC + 1 : c->KC;
    void *src = isw ? c->dW[id - EFDC_W2] : c->dK[id - EFDC_HPK];
    const double *s64 = (const double *)src;
    if (c->store == EFDC_STORE_FP32) {
      const int64_t n = c->pitch * planes;
      k_convert<double, float><<<blocks_for(n), EFDC_BLOCK, 0, c->stream>>>(
          c->staging, (const float *)src, n);
      s64 = c->staging;
    }
    HIP_TRY(c, hipMemcpy2DAsync(host, sizeof(double) * c->LCM, s64,
                                sizeof(double) * c->pitch, sizeof(double) * c->LCM,
                                planes, hipMemcpyDeviceToHost, c->stream));
  }
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return EFDC_OK;
}

int efdc_device_ptr(efdc_ctx_t *c, int32_t id, void **ptr, int64_t *pitch) {
  if (!c || !ptr || !valid_id(id)) return EFDC_ERR_ARG;
  if (uvw_id(id)) {
    double *d = nullptr;
    int planes = 0, s;
    if ((s = uvw_ptr(c, id, &d, &planes))) return s;
    *ptr = d;
  } else if (id < EFDC_NUM_H) *ptr = c->dH[id];
  else if (id >= EFDC_W2) *ptr = c->dW[id - EFDC_W2];
  else *ptr = c->dK[id - EFDC_HPK];
  if (pitch) *pitch = c->pitch;
  return EFDC_OK;
}

int efdc_set_open_bc(efdc_ctx_t *c, const efdc_open_bc_t *bc) {
  if (c) c->symok_cache = -1;
  if (!c || !bc || bc->mtide < 0 || (bc->mtide > 0 && !bc->tcp))
    return EFDC_ERR_ARG;
  const int LA = c->LA, LC = c->LC;
  std::vector<unsigned char> flag(c->LCM, 0);
  std::vector<int> lob, lob2;
  // LOBCS / LOBCS2 in the order S, W, E, N (EFDC/setbcs.f90:884-938)
  static const int order[4] = {2, 0, 1, 3};
  for (int side = 0; side < 4; ++side) {
    const efdc_pbc_side_t &s = bc->side[side];
    if (s.n < 0 || (s.n > 0 && (!s.L || !s.ispb || !s.pser ||
                                (bc->mtide > 0 && (!s.pcb || !s.psb)))))
      return EFDC_ERR_ARG;
    for (int LL = 0; LL < s.n; ++LL)
      if (s.L[LL] < 2 || s.L[LL] > LA) {
        c->err = "open-boundary cell out of range";
        return EFDC_ERR_ARG;
      }
  }
  c->npbc_total = 0;
  for (int side = 0; side < 4; ++side) {
    const efdc_pbc_side_t &s = bc->side[side];
    PbcDev &d = c->pbc[side];
    std::vector<int> L0(s.n);
    std::set<int> own, nbs;
    bool serial = false;
    std::vector<int> nbv;
    for (int LL = 0; LL < s.n; ++LL) {
      L0[LL] = s.L[LL] - 1;
      own.insert(L0[LL]);
      flag[L0[LL]] = 1;
    }
    std::vector<int> mapv(c->LCM);
    int *dmap = side == 0 ? c->dLE : side == 1 ? c->dLW : side == 2 ? c->dLN : c->dLS;
    HIP_TRY(c, hipMemcpy(mapv.data(), dmap, sizeof(int) * c->LCM,
                         hipMemcpyDeviceToHost));
    for (int LL = 0; LL < s.n; ++LL) {
      const int nb = mapv[L0[LL]];
      if (nb != LC - 1 && (own.count(nb) || !nbs.insert(nb).second)) serial = true;
    }
    d.n = s.n;
    d.serial = serial ? 1 : 0;
    int st;
    std::vector<double> pc((size_t)s.n * std::max(bc->mtide, 1), 0.0),
        ps((size_t)s.n * std::max(bc->mtide, 1), 0.0);
    if (bc->mtide > 0 && s.n > 0) {
      std::memcpy(pc.data(), s.pcb, sizeof(double) * s.n * bc->mtide);
      std::memcpy(ps.data(), s.psb, sizeof(double) * s.n * bc->mtide);
    }
    c->pbc_ispb_host[side].assign(s.ispb, s.ispb + s.n);
    if ((st = upload_new(c, &d.L, L0.data(), s.n)) ||
        (st = upload_new(c, &d.ispb, s.ispb, s.n)) ||
        (st = upload_new(c, &d.pser, s.pser, s.n)) ||
        (st = upload_new(c, &d.pcb, pc.data(), pc.size())) ||
        (st = upload_new(c, &d.psb, ps.data(), ps.size())))
      return st;
    c->npbc_total += s.n;
  }
  for (int q = 0; q < 4; ++q) {
    const efdc_pbc_side_t &s = bc->side[order[q]];
    for (int LL = 0; LL < s.n; ++LL) {
      const int L = s.L[LL];
      lob.push_back(L - 1);
      if (s.ispb[LL] <= 1) {
        const int side = order[q];
        std::vector<int> one(1);
        int *dmap = side == 0 ? c->dLE : side == 1 ? c->dLW : side == 2 ? c->dLN : c->dLS;
        HIP_TRY(c, hipMemcpy(one.data(), dmap + (L - 1), sizeof(int),
                             hipMemcpyDeviceToHost));
        lob2.push_back(one[0]);
      }
    }
  }
  int st;
  c->nlobcs = (int)lob.size();
  c->nlobcs2 = (int)lob2.size();
  if ((st = upload_new(c, &c->lobcs, lob.data(), lob.size())) ||
      (st = upload_new(c, &c->lobcs2, lob2.data(), lob2.size())))
    return st;
  HIP_TRY(c, hipMemcpyAsync(c->pbcflag, flag.data(), c->LCM,
                            hipMemcpyHostToDevice, c->stream));
  {
    std::vector<unsigned char> lf(c->LCM, 0);
    for (int j : lob) lf[j] |= 1;
    for (int j : lob2) if (j >= 0 && j < c->LCM) lf[j] |= 2;
    HIP_TRY(c, hipMemcpyAsync(c->lobflag, lf.data(), c->LCM, hipMemcpyHostToDevice,
                              c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
  }
  uvw_obc_changed(c);
  c->mtide = bc->mtide;
  delete[] c->tcp_host;
  c->tcp_host = new double[std::max(bc->mtide, 1)];
  for (int m = 0; m < bc->mtide; ++m) c->tcp_host[m] = bc->tcp[m];
  dfree(c->tide_dev);
  if (c->tide_host) (void)hipHostFree(c->tide_host);
  c->tide_host = nullptr;
  const int nt = 2 * std::max(bc->mtide, 1);
  if ((st = dmalloc(c, &c->tide_dev, nt))) return st;
  HIP_TRY(c, hipHostMalloc((void **)&c->tide_host, sizeof(double) * nt));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return EFDC_OK;
}

int efdc_add_constituent(efdc_ctx_t *c, const efdc_constituent_t *d,
                         int32_t *slot) {
  if (!c || !d || !slot) return EFDC_ERR_ARG;
  if (c->nconst >= EFDC_MAXC) {
    c->err = "too many constituents";
    return EFDC_ERR_ARG;
  }
  if (!(d->wset >= 0.0 && d->wset < 1e3) || !(d->probdep >= 0.0 && d->probdep <= 1.0) ||
      !(d->wese >= 0.0 && d->wese < 1e30)) {
    c->err = "settling: need 0 <= wset, 0 <= probdep <= 1, 0 <= wese (all finite)";
    return EFDC_ERR_ARG;
  }
  if (d->isadac > 0 && !c->mutual) {
    // the source / open-boundary zeroing of the antidiffusive fluxes is
    // applied on the face shared with the neighbour (LEC(LWC(L)) = L)
    c->err = "anti-diffusion needs mutual neighbour maps";
    return EFDC_ERR_ARG;
  }
  ConstDev &k = c->con[c->nconst];
  k.desc = *d;
  k.cur = 0;
  const size_t n = (size_t)c->pitch * c->KC * c->tsz;
  for (int j = 0; j < 2; ++j) {
    char *p = nullptr;
    int s = dmalloc(c, &p, n);
    if (s) return s;
    k.buf[j] = p;
  }
  *slot = c->nconst++;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return EFDC_OK;
}

static int put_layers(efdc_ctx *c, void *dst, const double *host) {
  double *d64 = c->store == EFDC_STORE_FP64 ? (double *)dst : c->staging;
  HIP_TRY(c, hipMemcpy2DAsync(d64, sizeof(double) * c->pitch, host,
                              sizeof(double) * c->LCM, sizeof(double) * c->LCM,
                              c->KC, hipMemcpyHostToDevice, c->stream));
  if (c->store == EFDC_STORE_FP32) {
    const int64_t n = c->pitch * c->KC;
    k_convert<float, double><<<blocks_for(n), EFDC_BLOCK, 0, c->stream>>>(
        (float *)dst, c->staging, n);
  }
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return EFDC_OK;
}

static int get_layers(efdc_ctx *c, const void *src, double *host) {
  const double *s64 = (const double *)src;
  if (c->store == EFDC_STORE_FP32) {
    const int64_t n = c->pitch * c->KC;
    k_convert<double, float><<<blocks_for(n), EFDC_BLOCK, 0, c->stream>>>(
        c->staging, (const float *)src, n);
    s64 = c->staging;
  }
  HIP_TRY(c, hipMemcpy2DAsync(host, sizeof(double) * c->LCM, s64,
                              sizeof(double) * c->pitch, sizeof(double) * c->LCM,
                              c->KC, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return EFDC_OK;
}

int efdc_set_constituent(efdc_ctx_t *c, int32_t slot, const double *con,
                         const double *con1) {
  if (!c || !con || slot < 0 || slot >= c->nconst) return EFDC_ERR_ARG;
  ConstDev &k = c->con[slot];
  int s = put_layers(c, k.buf[k.cur], con);
  if (s) return s;
  if (!con1) return put_layers(c, k.buf[k.cur ^ 1], con);
  // CALTRAN starts with CON1=CON over the whole array, ghosts included
  // (EFDC/caltran.f90:59-65): keep the ghost entries of both buffers equal.
  std::vector<double> tmp((size_t)c->LCM * c->KC);
  for (int kk = 0; kk < c->KC; ++kk) {
    const double *a = con + (size_t)kk * c->LCM;
    const double *b = con1 + (size_t)kk * c->LCM;
    double *t = tmp.data() + (size_t)kk * c->LCM;
    std::memcpy(t, b, sizeof(double) * c->LCM);
    t[0] = a[0];
    for (int i = c->LA; i < c->LCM; ++i) t[i] = a[i];
  }
  return put_layers(c, k.buf[k.cur ^ 1], tmp.data());
}

int efdc_get_constituent(efdc_ctx_t *c, int32_t slot, double *con,
                         double *con1) {
  if (!c || slot < 0 || slot >= c->nconst) return EFDC_ERR_ARG;
  ConstDev &k = c->con[slot];
  if (con) {
    int s = get_layers(c, k.buf[k.cur], con);
    if (s) return s;
  }
  if (con1) return get_layers(c, k.buf[k.cur ^ 1], con1);
  return EFDC_OK;
}

int efdc_get_dry_state(efdc_ctx_t *c, int32_t *lmaskdry, int32_t *iscdry, int32_t *ncordry) {
  if (!c) return EFDC_ERR_ARG;
  if (lmaskdry)
    HIP_TRY(c, hipMemcpyAsync(lmaskdry, c->lmaskdry, sizeof(int) * c->LCM, hipMemcpyDeviceToHost,
                              c->stream));
  if (iscdry)
    HIP_TRY(c, hipMemcpyAsync(iscdry, c->iscdry, sizeof(int) * c->LCM, hipMemcpyDeviceToHost,
                              c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (ncordry) *ncordry = c->ncordry;
  return EFDC_OK;
}

int efdc_set_dry_counters(efdc_ctx_t *c, const double *vdwaste, const int32_t *natdry) {
  if (!c) return EFDC_ERR_ARG;
  if (vdwaste)
    HIP_TRY(c, hipMemcpyAsync(c->vdwaste, vdwaste, sizeof(double) * c->LCM,
                              hipMemcpyHostToDevice, c->stream));
  if (natdry)
    HIP_TRY(c, hipMemcpyAsync(c->natdry, natdry, sizeof(int) * c->LCM, hipMemcpyHostToDevice,
                              c->stream));
  // the host arrays may be reused as soon as the call returns
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return EFDC_OK;
}

int efdc_get_dry_waste(efdc_ctx_t *c, double *qdwaste, double *vdwaste, int32_t *natdry) {
  if (!c) return EFDC_ERR_ARG;
  if (qdwaste)
    HIP_TRY(c, hipMemcpyAsync(qdwaste, c->qdwaste, sizeof(double) * c->LCM,
                              hipMemcpyDeviceToHost, c->stream));
  if (vdwaste)
    HIP_TRY(c, hipMemcpyAsync(vdwaste, c->vdwaste, sizeof(double) * c->LCM,
                              hipMemcpyDeviceToHost, c->stream));
  if (natdry)
    HIP_TRY(c, hipMemcpyAsync(natdry, c->natdry, sizeof(int) * c->LCM, hipMemcpyDeviceToHost,
                              c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return EFDC_OK;
}

int efdc_get_constituent_at(efdc_ctx_t *c, int32_t slot, int32_t n, const int32_t *L,
                            const int32_t *K, double *out) {
  if (!c || slot < 0 || slot >= c->nconst || n < 0 || (n > 0 && (!L || !K || !out)))
    return EFDC_ERR_ARG;
  if (n == 0) return EFDC_OK;
  std::vector<int64_t> off(n);
  for (int j = 0; j < n; ++j) {
    if (L[j] < 2 || L[j] > c->LA || K[j] < 1 || K[j] > c->KC) {
      c->err = "efdc_get_constituent_at: cell outside 2..LA x 1..KC";
      return EFDC_ERR_ARG;
    }
    off[j] = (int64_t)(L[j] - 1) + (int64_t)(K[j] - 1) * c->pitch;
  }
  if (n > c->gat_cap) {
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    dfree(c->gat_off);
    dfree(c->gat_val);
    c->gat_cap = 0;
    int s;
    if ((s = dmalloc(c, &c->gat_off, (size_t)n)) || (s = dmalloc(c, &c->gat_val, (size_t)n)))
      return s;
    c->gat_cap = n;
  }
  ConstDev &k = c->con[slot];
  HIP_TRY(c, hipMemcpyAsync(c->gat_off, off.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice,
                            c->stream));
  const int nb = blocks_for(n);
  if (c->store == EFDC_STORE_FP64)
    k_gather<double><<<nb, EFDC_BLOCK, 0, c->stream>>>((const double *)k.buf[k.cur],
                                                       c->gat_off, c->gat_val, n);
  else
    k_gather<float><<<nb, EFDC_BLOCK, 0, c->stream>>>((const float *)k.buf[k.cur],
                                                      c->gat_off, c->gat_val, n);
  HIP_TRY(c, hipGetLastError());
  HIP_TRY(c, hipMemcpyAsync(out, c->gat_val, sizeof(double) * n, hipMemcpyDeviceToHost,
                            c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return EFDC_OK;
}

int efdc_set_conc_bc(efdc_ctx_t *c, const efdc_conc_bc_t *bc) {
  if (!c || !bc) return EFDC_ERR_ARG;
  const int nc = c->nconst, KC = c->KC, LA = c->LA;
  if (nc < 1) {
    c->err = "register the constituents before their boundary tables";
    return EFDC_ERR_STATE;
  }
  std::vector<int> rec(c->LCM, -1);
  std::vector<int4> info;
  int total = 0;
  for (int side = 0; side < 4; ++side) {
    const efdc_cbc_side_t &s = bc->side[side];
    if (s.n < 0 || (s.n > 0 && (!s.L || !s.cb || !s.csert || !s.ntscr)))
      return EFDC_ERR_ARG;
    total += s.n;
  }
  if (total > 0 && !bc->wtci) return EFDC_ERR_ARG;
  for (int side = 0; side < 4; ++side) {
    const efdc_cbc_side_t &s = bc->side[side];
    CbcDev &d = c->cbc[side];
    const bool same = (d.n == s.n);
    std::vector<int> L0(s.n);
    for (int LL = 0; LL < s.n; ++LL) {
      if (s.L[LL] < 2 || s.L[LL] > LA) return EFDC_ERR_ARG;
      L0[LL] = s.L[LL] - 1;
      int &r = rec[L0[LL]];
      if (r < 0) {
        r = (int)info.size();
        info.push_back(make_int4(-1, -1, -1, -1));
      }
      int4 &q = info[r];
      if (side == 0) q.x = LL;
      else if (side == 1) q.y = LL;
      else if (side == 2) q.z = LL;
      else q.w = LL;
    }
    c->cbc_cells_host[side] = L0;
    int st;
    if ((st = upload_new(c, &d.L, L0.data(), s.n)) ||
        (st = upload_new(c, &d.cb, s.cb, (size_t)s.n * 2 * nc)) ||
        (st = upload_new(c, &d.csert, s.csert, (size_t)s.n * KC * nc)) ||
        (st = upload_new(c, &d.ntscr, s.ntscr, s.n)))
      return st;
    if (!same || !d.clo) {  // recovery memory survives table refreshes
      dfree(d.clo);
      dfree(d.nlo);
      if ((st = dmalloc(c, &d.clo, (size_t)s.n * KC * nc)) ||
          (st = dmalloc(c, &d.nlo, (size_t)s.n * KC * nc)))
        return st;
    }
    d.n = s.n;
  }
  uvw_obc_changed(c);
  int st;
  c->nbcrec = (int)info.size();
  c->cbc_nc = nc;
  if ((st = upload_new(c, &c->wtci, bc->wtci, (size_t)KC * 2)) ||
      (st = upload_new(c, &c->bcinfo, info.data(), info.size())))
    return st;
  dfree(c->bc_con1);
  if (c->nbcrec > 0 &&
      (st = dmalloc(c, &c->bc_con1, (size_t)c->nbcrec * KC * nc)))
    return st;
  {
    std::vector<int> cell(std::max(c->nbcrec, 1), 0);
    for (int i = 0; i < c->LCM; ++i)
      if (rec[i] >= 0) cell[rec[i]] = i;
    if ((st = upload_new(c, &c->bccell, cell.data(), (size_t)c->nbcrec))) return st;
  }
  HIP_TRY(c, hipMemcpyAsync(c->bcrec, rec.data(), sizeof(int) * c->LCM,
                            hipMemcpyHostToDevice, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return EFDC_OK;
}

int efdc_set_sources(efdc_ctx_t *c, const efdc_sources_t *src) {
  if (!c || !src || src->n < 0) return EFDC_ERR_ARG;
  const int n = src->n, KC = c->KC, nc = std::max(c->nconst, 1);
  if (n > 0 && (!src->LQS || !src->qss || !src->qsercell || !src->cqs ||
                !src->csert))
    return EFDC_ERR_ARG;
  std::vector<int> rec(c->LCM, -1), next(std::max(n, 1), -1), l0(n);
  // chain sources of one cell in increasing NS (accumulation order of
  // EFDC/calfqc.f90:211-224)
  std::vector<int> last(c->LCM, -1);
  for (int s = 0; s < n; ++s) {
    const int L = src->LQS[s];
    if (L < 2 || L > c->LA) return EFDC_ERR_ARG;
    l0[s] = L - 1;
    if (rec[L - 1] < 0) rec[L - 1] = s;
    else next[last[L - 1]] = s;
    last[L - 1] = s;
  }
  int st;
  c->nqsij = n;
  c->src_nc = nc;
  if ((st = upload_new(c, &c->lqs, l0.data(), n)) ||
      (st = upload_new(c, &c->srcnext, next.data(), next.size())) ||
      (st = upload_new(c, &c->qss, src->qss, (size_t)n * KC)) ||
      (st = upload_new(c, &c->qsercell, src->qsercell, (size_t)n * KC)) ||
      (st = upload_new(c, &c->cqs, src->cqs, (size_t)n * KC * nc)) ||
      (st = upload_new(c, &c->csertq, src->csert, (size_t)n * KC * nc)))
    return st;
  HIP_TRY(c, hipMemcpyAsync(c->srcrec, rec.data(), sizeof(int) * c->LCM,
                            hipMemcpyHostToDevice, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return EFDC_OK;
}

int efdc_set_cg_replay(efdc_ctx_t *c, int32_t mode) {
  if (!c || mode < 0 || mode > 2) return EFDC_ERR_ARG;
  c->cg_graph = mode == 1;
  c->cg_graph_tune = mode == 2;
  return EFDC_OK;
}

int efdc_get_cg_replay(efdc_ctx_t *c, int32_t *mode) {
  if (!c || !mode) return EFDC_ERR_ARG;
  *mode = c->cg_graph_tune ? 2 : (c->cg_graph ? 1 : 0);
  return EFDC_OK;
}

int efdc_set_cg_pipeline(efdc_ctx_t *c, int32_t mode) {
  if (!c || mode < 0 || mode > 2) return EFDC_ERR_ARG;
  c->pcg_pipe_mode = mode;
  return EFDC_OK;
}

int efdc_set_cg64_pipeline(efdc_ctx_t *c, int32_t on) {
  if (!c || on < 0 || on > 1) return EFDC_ERR_ARG;
  c->pcg64_pipe = on != 0;
  return EFDC_OK;
}

int efdc_get_cg64_pipeline(efdc_ctx_t *c, int32_t *on) {
  if (!c || !on) return EFDC_ERR_ARG;
  *on = c->pcg64_pipe ? 1 : 0;
  return EFDC_OK;
}

int efdc_set_cg_launch(efdc_ctx_t *c, int32_t mode, uint32_t spin_limit) {
  if (!c || mode < 0 || mode > 1) return EFDC_ERR_ARG;
  c->pcg_launch_mode = mode;
  c->pcg_spin_test = spin_limit;
  return EFDC_OK;
}

int efdc_get_cg_launch(efdc_ctx_t *c, int32_t *mode, int32_t *retries) {
  if (!c || !mode || !retries) return EFDC_ERR_ARG;
  *mode = c->pcg_launch_mode;
  *retries = c->pcg_retries;
  return EFDC_OK;
}

int efdc_get_cg_pipeline(efdc_ctx_t *c, int32_t *mode) {
  if (!c || !mode) return EFDC_ERR_ARG;
  *mode = c->pcg_pipe_mode != 0 ? c->pcg_pipe_mode : (c->pcg_pipe_tuned ? 2 : 1);
  return EFDC_OK;
}

int efdc_autotune(efdc_ctx_t *c, int32_t on) {
  if (!c) return EFDC_ERR_ARG;
  c->tune_pending = on != 0;
  c->tune_cg_pending = on != 0;
  c->cg_graph_tune = on != 0;
  if (!on) c->cg_graph = false;
  if (on) {
    c->tune_n = 0;
    c->tune_best = -1;
    for (int &f : c->tune_chosen) f = 0;
  } else {
    c->tran_ncg = 0;
    c->tran_w3 = 0;
    c->pcg_pdeg_tuned = 0;
    c->pcg_pipe_tuned = true;
  }
  return EFDC_OK;
}

int efdc_tune_report(efdc_ctx_t *c, efdc_tune_report_t *out) {
  if (!c || !out) return EFDC_ERR_ARG;
  std::memset(out, 0, sizeof(*out));
  out->n = c->tune_n;
  for (int j = 0; j < c->tune_n && j < 8; ++j) {
    std::memcpy(out->v[j].name, c->tune_name[j], 32);
    out->v[j].ms = c->tune_ms[j];
    out->v[j].chosen = c->tune_chosen[j];
  }
  return EFDC_OK;
}

int efdc_stream_copy_gbs(efdc_ctx_t *c, size_t bytes, int32_t reps, double *gbs) {
  if (!c || !gbs || bytes < 4096 || reps < 1) return EFDC_ERR_ARG;
  const size_t n4 = bytes / 16;
  float4 *x = nullptr, *y = nullptr;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (hipMalloc((void **)&x, n4 * 16) != hipSuccess ||
      hipMalloc((void **)&y, n4 * 16) != hipSuccess) {
    if (x) (void)hipFree(x);
    c->err = "efdc_stream_copy_gbs: hipMalloc failed";
    return EFDC_ERR_NOMEM;
  }
  hipEvent_t a = timer_event(c), b = timer_event(c);
  const int nb = 8 * 256;
  float best = 1e30f;
  int st = EFDC_OK;
  for (int r = 0; r <= reps && st == EFDC_OK; ++r) {
    (void)hipEventRecord(a, c->stream);
    k_copy4<<<nb, 256, 0, c->stream>>>(x, y, n4);
    (void)hipEventRecord(b, c->stream);
    if (hipEventSynchronize(b) != hipSuccess) st = EFDC_ERR_HIP;
    float ms = 0.f;
    if (st == EFDC_OK && hipEventElapsedTime(&ms, a, b) != hipSuccess) st = EFDC_ERR_HIP;
    if (r > 0 && ms < best) best = ms;  // the first copy is a warm-up
  }
  c->ev_pool.push_back(a);
  c->ev_pool.push_back(b);
  (void)hipFree(x);
  (void)hipFree(y);
  if (st) {
    c->err = "efdc_stream_copy_gbs: copy failed";
    return st;
  }
  *gbs = 2.0 * (double)(n4 * 16) / ((double)best * 1e-3) / 1e9;
  return EFDC_OK;
}

int efdc_set_source_cqs(efdc_ctx_t *c, const double *cqs) {
  if (!c || !cqs) return EFDC_ERR_ARG;
  if (c->nqsij == 0) {
    c->err = "efdc_set_source_cqs: no source table (efdc_set_sources)";
    return EFDC_ERR_STATE;
  }
  // staged through a pinned ring so that the call does not wait for the
  // work already queued (the caller's array is free again on return)
  const size_t bytes = sizeof(double) * (size_t)c->nqsij * c->KC * c->src_nc;
  if (bytes != c->cqs_stage_bytes) {
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    for (int q = 0; q < efdc_ctx::CQS_RING; ++q) {
      if (c->cqs_stage[q]) (void)hipHostFree(c->cqs_stage[q]);
      c->cqs_stage[q] = nullptr;
      if (!c->cqs_ev[q]) HIP_TRY(c, hipEventCreateWithFlags(&c->cqs_ev[q], hipEventDisableTiming));
    }
    c->cqs_stage_bytes = 0;
    for (int q = 0; q < efdc_ctx::CQS_RING; ++q)
      if (hipHostMalloc((void **)&c->cqs_stage[q], bytes, hipHostMallocDefault) != hipSuccess) {
        c->err = "efdc_set_source_cqs: pinned staging allocation failed";
        return EFDC_ERR_NOMEM;
      }
    c->cqs_stage_bytes = bytes;
    c->cqs_next = 0;
  }
  const int q = c->cqs_next;
  c->cqs_next = (q + 1) % efdc_ctx::CQS_RING;
  HIP_TRY(c, hipEventSynchronize(c->cqs_ev[q]));
  std::memcpy(c->cqs_stage[q], cqs, bytes);
  HIP_TRY(c, hipMemcpyAsync(c->cqs, c->cqs_stage[q], bytes, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(c, hipEventRecord(c->cqs_ev[q], c->stream));
  return EFDC_OK;
}

int efdc_timers_enable(efdc_ctx_t *c, int32_t on) {
  if (!c) return EFDC_ERR_ARG;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  for (auto &r : c->trec) {
    c->ev_pool.push_back(r.a);
    c->ev_pool.push_back(r.b);
  }
  c->trec.clear();
  c->timing = on != 0;
  return EFDC_OK;
}

static_assert(TK_COUNT <= EFDC_TIMER_MAX, "efdc_timers_t holds fewer entries than the timers");

int efdc_timers_read(efdc_ctx_t *c, efdc_timers_t *out) {
  if (!c || !out) return EFDC_ERR_ARG;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  std::memset(out, 0, sizeof(*out));
  out->n = TK_COUNT;
  for (int k = 0; k < TK_COUNT; ++k)
    std::snprintf(out->k[k].name, sizeof(out->k[k].name), "%s", efdc_timer_names[k]);
  for (auto &r : c->trec) {
    float ms = 0.f;
    HIP_TRY(c, hipEventElapsedTime(&ms, r.a, r.b));
    out->k[r.id].count += 1;
    out->k[r.id].total_ms += ms;
    c->ev_pool.push_back(r.a);
    c->ev_pool.push_back(r.b);
  }
  c->trec.clear();
  out->k[TK_CG_ITERS].count = c->cg_iter_count;  // iterations, no duration
  c->cg_iter_count = 0;
  return EFDC_OK;
}

int efdc_synchronize(efdc_ctx_t *c) {
  if (!c) return EFDC_ERR_ARG;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return EFDC_OK;
}

int efdc_closure_set_manning(efdc_ctx_t *c, const double *nman) {
  if (!c || !nman) return EFDC_ERR_ARG;
  HIP_TRY(c, hipMemcpyAsync(c->nman, nman, sizeof(double) * c->LCM,
                            hipMemcpyHostToDevice, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return EFDC_OK;
}

int efdc_closure_pre(efdc_ctx_t *c) {
  if (!c) return EFDC_ERR_ARG;
  return closure_pre_run(c, c->nman);
}

int efdc_closure_post(efdc_ctx_t *c) {
  if (!c) return EFDC_ERR_ARG;
  return closure_post_run(c);
}

int efdc_congrad(efdc_ctx_t *c, efdc_cg_stats_t *stats) {
  if (!c) return EFDC_ERR_ARG;
  return cg_run(c, true, stats);
}

int efdc_calpuv2c(efdc_ctx_t *c, const efdc_step_t *step,
                  efdc_cg_stats_t *stats) {
  if (!c || !step || !(step->dt > 0.0)) return EFDC_ERR_ARG;
  return puv_run(c, step, stats);
}

int efdc_caluvw(efdc_ctx_t *c, const efdc_step_t *step, const efdc_uvw_opts_t *o) {
  if (!c || !step || !o || !(step->dt > 0.0)) return EFDC_ERR_ARG;
  return uvw_run(c, step, o);
}

int efdc_calconc(efdc_ctx_t *c, const efdc_step_t *step) {
  if (!c || !step || !(step->dt > 0.0)) return EFDC_ERR_ARG;
  return tran_run(c, step);
}

int efdc_balance(efdc_ctx_t *c, efdc_balance_t *out) {
  if (!c || !out) return EFDC_ERR_ARG;
  return balance_run(c, out);
}

}  // extern "C"
