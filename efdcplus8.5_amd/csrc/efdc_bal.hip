// efdc_bal.hip -- the two-level balances of a period (ISBAL >= 1,
// EFDC/bal2t.f90) for the volume and every transported constituent, FP64 on
// the device:
//   BAL2T1  the totals at the start of the period over the balance region
//           (SPB = 0 on pressure-BC cells, SCB = 0 on concentration-BC cells,
//           EFDC/setbcs.f90:180-239,952), the outflow totals zeroed;
//   BAL2T3A at the start of every CALCONC: the volume the sources add,
//           QSUME - QDWASTE over every cell, and the constituent mass they
//           add (inflow at CQS / CSERT, withdrawal at the cell's
//           concentration);
//   BAL2T2  at the end of every CALCONC: the volume and mass across the
//           faces of the concentration-BC cells, at the upwind cell's
//           concentration after the transport;
//   BAL2T5  the totals at the end and the errors END - (BEG - OUT).
// The totals are grid reductions (grid_sum, deterministic order); the
// source and boundary terms, O(sources + boundary cells) x KC, are summed by
// one workgroup and added to FP64 accumulators on the device, so a period
// costs no host round trip until its end.
#include <cstring>
#include <vector>

#include "efdc_internal.h"

struct BalState {
  double *acc = nullptr;        // device: [0] VOLOUT, [1 + c] mass out of constituent c
  double *tot = nullptr;        // device: the totals of the last reduction
  unsigned char *scb = nullptr; // device: 1 on concentration-BC cells (SCB = 0)
  double volbeg = 0.0, massbeg[EFDC_MAXC] = {};
  int64_t nsteps = 0;
  bool open = false;
};

namespace {

constexpr int NB = 1 + EFDC_MAXC;

template <typename T>
struct ConPtrs {
  const T *p[EFDC_MAXC];
};

// Sum over the wet cells outside `skip` (1 = left out) of DXYP*HP
// (con == nullptr) or of DXYP*HP*CON*DZC over the layers
// (EFDC/bal2t.f90:86,93-96 and :939,945-948).
template <typename T>
__global__ __launch_bounds__(EFDC_BLOCK) void k_bal_tot(
    const double *dxyp, const double *hp, const T *con, const T *dzc,
    const unsigned char *skip, int64_t pitch, int KC, int nwet, double *part,
    unsigned *ticket, double *out) {
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int i = 1 + tile * EFDC_BLOCK + (int)threadIdx.x;
  double v[1] = {0.0};
  if (i <= nwet && skip[i] == 0) {
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
  if (grid_sum<1>(v, part, ticket, gridDim.x, tile, s) && threadIdx.x == 0) *out = s[0];
}

// VOLOUT -= DT * sum (QSUME - QDWASTE) (EFDC/bal2t.f90:332-336)
__global__ __launch_bounds__(EFDC_BLOCK) void k_bal_qvol(
    const double *qsume, const double *qdwaste, int nwet, double dt, double *part,
    unsigned *ticket, double *acc) {
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int i = 1 + tile * EFDC_BLOCK + (int)threadIdx.x;
  double v[1] = {0.0};
  if (i <= nwet) v[0] = qsume[i] - (qdwaste ? qdwaste[i] : 0.0);
  double s[1];
  if (grid_sum<1>(v, part, ticket, gridDim.x, tile, s) && threadIdx.x == 0)
    acc[0] = acc[0] - dt * s[0];
}

// The source mass of every constituent (EFDC/bal2t.f90:360-366 and the
// same lines for the other constituents): one workgroup over (source,
// layer).
template <typename T>
__global__ __launch_bounds__(EFDC_BLOCK) void k_bal_src(
    ConPtrs<T> con, int nc, const int *lqs, const double *qss, const double *qser,
    const double *cqs, const double *csert, int nq, int KC, int64_t pitch, double dt,
    double *acc) {
  double v[NB];
#pragma unroll
  for (int q = 0; q < NB; ++q) v[q] = 0.0;
  for (int t = threadIdx.x; t < nq * KC; t += EFDC_BLOCK) {
    const int s = t / KC, k = t % KC;
    const int64_t j = lqs[s] + k * pitch;
    const double qs = qss[s * KC + k], qr = qser[s * KC + k];
#pragma unroll
    for (int c = 0; c < EFDC_MAXC; ++c) {
      if (c < nc) {
        const size_t x = (size_t)(s * KC + k) * nc + c;
        const double ck = (double)con.p[c][j];
        v[1 + c] += dt * fmax(qs, 0.) * cqs[x] + dt * fmin(qs, 0.) * ck +
                    dt * fmax(qr, 0.) * csert[x] + dt * fmin(qr, 0.) * ck;
      }
    }
  }
  block_sum<NB>(v);
  if (threadIdx.x == 0)
    for (int c = 0; c < nc; ++c) acc[1 + c] = acc[1 + c] - v[1 + c];
}

// The volume and mass across the faces of the concentration-BC cells
// (EFDC/bal2t.f90:193-287), sides S, W, E, N: one workgroup over (side
// cell, layer), the mass at the upwind concentration.
template <typename T>
__global__ __launch_bounds__(EFDC_BLOCK) void k_bal_open(
    ConPtrs<T> con, int nc, CbcDev s0, CbcDev s1, CbcDev s2, CbcDev s3, const T *uhdy2,
    const T *vhdx2, const int *lw, const int *le, const int *ls, const int *ln, int KC,
    int64_t pitch, double dt, double *acc) {
  double v[NB];
#pragma unroll
  for (int q = 0; q < NB; ++q) v[q] = 0.0;
  const CbcDev sd[4] = {s0, s1, s2, s3};
  for (int side = 0; side < 4; ++side) {
    const CbcDev &b = sd[side];
    for (int t = threadIdx.x; t < b.n * KC; t += EFDC_BLOCK) {
      const int q = t / KC, k = t % KC;
      const int L = b.L[q];
      // the face, the sign of an outflow and the cells on either side
      int face, out_c, in_c;
      double sgn;
      if (side == 0) { face = ln[L]; sgn = -1.0; out_c = ln[L]; in_c = L; }
      else if (side == 1) { face = le[L]; sgn = -1.0; out_c = le[L]; in_c = L; }
      else if (side == 2) { face = L; sgn = 1.0; out_c = L; in_c = lw[L]; }
      else { face = L; sgn = 1.0; out_c = L; in_c = ls[L]; }
      const int64_t jf = face + k * pitch;
      const double f = (double)((side == 0 || side == 3) ? vhdx2[jf] : uhdy2[jf]);
      v[0] += sgn * dt * f;
#pragma unroll
      for (int c = 0; c < EFDC_MAXC; ++c) {
        if (c < nc) {
          const double co = (double)con.p[c][out_c + k * pitch];
          const double ci = (double)con.p[c][in_c + k * pitch];
          v[1 + c] += sgn * (dt * fmin(f, 0.) * co + dt * fmax(f, 0.) * ci);
        }
      }
    }
  }
  block_sum<NB>(v);
  if (threadIdx.x == 0)
    for (int q = 0; q < 1 + nc; ++q) acc[q] = acc[q] + v[q];
}

template <typename T>
ConPtrs<T> con_ptrs(efdc_ctx *c) {
  ConPtrs<T> p{};
  for (int j = 0; j < c->nconst; ++j) p.p[j] = (const T *)c->con[j].buf[c->con[j].cur];
  return p;
}

// totals[0] = volume over SPB, totals[1 + c] = mass of c over SCB
template <typename T>
int totals(efdc_ctx *c, std::vector<double> &h) {
  BalState *b = c->bal;
  const int nb = efdc_grid_blocks(c->n_wet());
  for (int j = -1; j < c->nconst; ++j)
    TIMED(c, TK_BALANCE, k_bal_tot<T><<<nb, EFDC_BLOCK, 0, c->stream>>>(
        c->dH[EFDC_DXYP], c->dH[EFDC_HP],
        j < 0 ? nullptr : (const T *)c->con[j].buf[c->con[j].cur],
        (const T *)c->dK[EFDC_DZC - EFDC_HPK], j < 0 ? c->pbcflag : b->scb, c->pitch, c->KC,
        c->n_wet(), c->part, c->tickets + TS_BALANCE * TICKET_SITE, b->tot + (j + 1)));
  h.assign(NB, 0.0);
  HIP_TRY(c, hipMemcpyAsync(h.data(), b->tot, sizeof(double) * NB, hipMemcpyDeviceToHost,
                            c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return EFDC_OK;
}

template <typename T>
int sources(efdc_ctx *c, double dt) {
  BalState *b = c->bal;
  const int nb = efdc_grid_blocks(c->n_wet());
  TIMED(c, TK_BALANCE, k_bal_qvol<<<nb, EFDC_BLOCK, 0, c->stream>>>(
      c->dH[EFDC_QSUME], c->prm.NDRYSTP < 0 ? c->qdwaste : nullptr, c->n_wet(), dt, c->part,
      c->tickets + TS_BALANCE * TICKET_SITE, b->acc));
  if (c->nqsij > 0 && c->nconst > 0)
    TIMED(c, TK_BALANCE, k_bal_src<T><<<1, EFDC_BLOCK, 0, c->stream>>>(
        con_ptrs<T>(c), c->nconst, c->lqs, c->qss, c->qsercell, c->cqs, c->csertq, c->nqsij,
        c->KC, c->pitch, dt, b->acc));
  HIP_TRY(c, hipGetLastError());
  return EFDC_OK;
}

template <typename T>
int open_faces(efdc_ctx *c, double dt) {
  BalState *b = c->bal;
  if (c->cbc[0].n + c->cbc[1].n + c->cbc[2].n + c->cbc[3].n == 0) return EFDC_OK;
  TIMED(c, TK_BALANCE, k_bal_open<T><<<1, EFDC_BLOCK, 0, c->stream>>>(
      con_ptrs<T>(c), c->nconst, c->cbc[0], c->cbc[1], c->cbc[2], c->cbc[3],
      (const T *)c->dK[EFDC_UHDY2 - EFDC_HPK], (const T *)c->dK[EFDC_VHDX2 - EFDC_HPK], c->dLW,
      c->dLE, c->dLS, c->dLN, c->KC, c->pitch, dt, b->acc));
  HIP_TRY(c, hipGetLastError());
  return EFDC_OK;
}

}  // namespace

// BAL2T3A before the transport of a CALCONC, when a period is open
int bal_before_tran(efdc_ctx *c, const efdc_step_t *step) {
  if (!c->bal || !c->bal->open) return EFDC_OK;
  // source tables built for another constituent count: the transport
  // refuses the call, so there is nothing to add (and the tables do not
  // have the layout read here)
  if (c->nqsij > 0 && c->src_nc != c->nconst) return EFDC_OK;
  return c->store == EFDC_STORE_FP64 ? sources<double>(c, step->dt)
                                     : sources<float>(c, step->dt);
}

// BAL2T2 after it
int bal_after_tran(efdc_ctx *c, const efdc_step_t *step) {
  if (!c->bal || !c->bal->open) return EFDC_OK;
  c->bal->nsteps += 1;
  return c->store == EFDC_STORE_FP64 ? open_faces<double>(c, step->dt)
                                     : open_faces<float>(c, step->dt);
}

void bal_free(efdc_ctx *c) {
  BalState *b = c->bal;
  if (!b) return;
  if (b->acc) (void)hipFree(b->acc);
  if (b->tot) (void)hipFree(b->tot);
  if (b->scb) (void)hipFree(b->scb);
  delete b;
  c->bal = nullptr;
}

extern "C" {

int efdc_bal2t_begin(efdc_ctx_t *c) {
  if (!c) return EFDC_ERR_ARG;
  if (c->nseg > 1) {
    c->err = "the two-level balances with several basins are not built (efdc_balance_segments "
             "gives the totals per basin)";
    return EFDC_ERR_ARG;
  }
  if (!c->bal) {
    c->bal = new BalState();
    int s;
    if ((s = dev_alloc(c, (void **)&c->bal->acc, sizeof(double) * NB)) ||
        (s = dev_alloc(c, (void **)&c->bal->tot, sizeof(double) * NB)) ||
        (s = dev_alloc(c, (void **)&c->bal->scb, (size_t)c->LCM)))
      return s;
  }
  BalState *b = c->bal;
  std::vector<unsigned char> scb(c->LCM, 0);
  for (int s = 0; s < 4; ++s)
    for (int L0 : c->cbc_cells_host[s]) scb[L0] = 1;
  HIP_TRY(c, hipMemcpyAsync(b->scb, scb.data(), c->LCM, hipMemcpyHostToDevice, c->stream));
  std::vector<double> zero(NB, 0.0), h;
  HIP_TRY(c, hipMemcpyAsync(b->acc, zero.data(), sizeof(double) * NB, hipMemcpyHostToDevice,
                            c->stream));
  int s = c->store == EFDC_STORE_FP64 ? totals<double>(c, h) : totals<float>(c, h);
  if (s) return s;
  b->volbeg = h[0];
  for (int j = 0; j < EFDC_MAXC; ++j) b->massbeg[j] = h[1 + j];
  b->nsteps = 0;
  b->open = true;
  return EFDC_OK;
}

int efdc_bal2t_end(efdc_ctx_t *c, efdc_bal2t_t *out) {
  if (!c || !out) return EFDC_ERR_ARG;
  if (!c->bal || !c->bal->open) {
    c->err = "no balance period is open (efdc_bal2t_begin)";
    return EFDC_ERR_STATE;
  }
  BalState *b = c->bal;
  std::vector<double> h, acc(NB);
  int s = c->store == EFDC_STORE_FP64 ? totals<double>(c, h) : totals<float>(c, h);
  if (s) return s;
  HIP_TRY(c, hipMemcpyAsync(acc.data(), b->acc, sizeof(double) * NB, hipMemcpyDeviceToHost,
                            c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  std::memset(out, 0, sizeof(*out));
  out->nconst = c->nconst;
  out->nsteps = b->nsteps;
  out->volbeg = b->volbeg;
  out->volend = h[0];
  out->volout = acc[0];
  out->volerr = out->volend - (out->volbeg - out->volout);
  for (int j = 0; j < c->nconst && j < 16; ++j) {
    out->massbeg[j] = b->massbeg[j];
    out->massend[j] = h[1 + j];
    out->massout[j] = acc[1 + j];
    out->masserr[j] = out->massend[j] - (out->massbeg[j] - out->massout[j]);
  }
  b->open = false;
  return EFDC_OK;
}

}  // extern "C"
