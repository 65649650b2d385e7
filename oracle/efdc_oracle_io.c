/*
 * efdc_oracle_io.c -- CPU restatement of the on-disk formats the EFDC+
 * post-processors read, used ONLY as the checker in tests/ and as the CPU
 * baseline of bench.py: RESTOUT's RESTART.OUT text
 * (EFDC/restartmod.f90:17-297) and HEADEROUT / WSOUT / VELOUT's EE_WS.OUT
 * and EE_VEL.OUT streams (EFDC/efdcout.f90:73-117,371-497), written in the
 * reference's order, statement by statement, from the oracle's state.
 *
 * Formatted output follows the Fortran edit descriptors the reference
 * uses: Ew.d prints 0.d1..dd and the exponent as E+ee / E-ee, or +eee /
 * -eee when it needs three digits, right-justified in w; Iw right-justifies
 * the integer; Fw.d prints d decimals; a field too narrow is w asterisks;
 * FORMAT(50E17.8) starts a new record after every 50 values.  Digits are
 * the exact binary value rounded to nearest, ties to even (the C library's
 * conversion), which is what the Fortran runtime writes (pinned against
 * flang's formatted WRITE in tests/test_oracle_io.py).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "efdc_oracle.h"

#define A(id) (o->a[id])
#define HX(L) ((size_t)(L) - 1)
#define KX(L, K) ((size_t)(L) - 1 + (size_t)((K) - 1) * (size_t)o->LCM)
#define WX(L, K) ((size_t)(L) - 1 + (size_t)(K) * (size_t)o->LCM)

/* Ew.d, w = 17, d = 8 */
void orc_fmt_e17_8(double v, char out[17]) {
  char r[32], s[48];
  if (isnan(v)) {
    snprintf(s, sizeof s, "%17s", "NaN");
  } else if (isinf(v)) {
    snprintf(s, sizeof s, "%17s", v < 0 ? "-Inf" : "Inf");
  } else {
    snprintf(r, sizeof r, "%.7e", v); /* [-]d.ddddddde[+-]xx[x] */
    const int neg = r[0] == '-';
    const char *m = r + neg;
    const int E = v == 0.0 ? 0 : atoi(m + 10) + 1; /* 0.d1d2.. x 10^E */
    const int ae = E < 0 ? -E : E;
    char ex[8];
    if (ae <= 99) snprintf(ex, sizeof ex, "E%c%02d", E < 0 ? '-' : '+', ae);
    else snprintf(ex, sizeof ex, "%c%03d", E < 0 ? '-' : '+', ae);
    char t[32];
    snprintf(t, sizeof t, "%s0.%c%.7s%s", neg ? "-" : "", m[0], m + 2, ex);
    snprintf(s, sizeof s, "%17s", t);
  }
  memcpy(out, s, 17);
}

/* Iw */
static void fmt_i(long long n, int w, char *out) {
  char s[40];
  int len = snprintf(s, sizeof s, "%*lld", w, n);
  if (len > w) memset(out, '*', (size_t)w);
  else memcpy(out, s, (size_t)w);
}

typedef struct {
  char *p;
  size_t n, cap;
} sbuf;

static void put(sbuf *b, const char *s, size_t n) {
  if (b->n + n > b->cap) {
    size_t c = b->cap ? b->cap : 4096;
    while (c < b->n + n) c *= 2;
    b->p = realloc(b->p, c);
    b->cap = c;
  }
  memcpy(b->p + b->n, s, n);
  b->n += n;
}

/* one WRITE(99,906) of n values: 50E17.8, a record per 50 values */
typedef struct {
  sbuf *b;
  int count;
} rec;

static void rec_e(rec *r, double v) {
  char f[17];
  if (r->count > 0 && r->count % 50 == 0) put(r->b, "\n", 1);
  orc_fmt_e17_8(v, f);
  put(r->b, f, 17);
  r->count++;
}

static void rec_i(rec *r, long long v) {
  char f[10];
  if (r->count > 0 && r->count % 50 == 0) put(r->b, "\n", 1);
  fmt_i(v, 10, f);
  put(r->b, f, 10);
  r->count++;
}

static void rec_end(rec *r) { put(r->b, "\n", 1); r->count = 0; }

static int writes(const efdc_restart_t *r, int m) { return r->istran[m] >= 1 && r->isco[m] == 1; }

/* NLO/CLO of one constituent on the four sides S, W, E, N
   (EFDC/restartmod.f90:158-185) */
static void bc_group(orc_ctx *o, const efdc_restart_t *r, int m, sbuf *b) {
  const int KC = o->KC, nc = o->nconst;
  rec q = {b, 0};
  for (int side = 0; side < 4; ++side) {
    orc_cbc_side *s = &o->cbc[side];
    for (int LL = 0; LL < s->n; ++LL) {
      for (int K = 1; K <= KC; ++K) {
        int64_t *nl = &s->nlo[(size_t)(LL * KC + K - 1) * nc + m];
        long long v = *nl - r->n;
        if (v < -(long long)s->ntscr[LL]) v = -(long long)s->ntscr[LL];
        if (!r->keep_nlo) *nl = v;
        rec_i(&q, v);
      }
      rec_end(&q);
      for (int K = 1; K <= KC; ++K) rec_e(&q, s->clo[(size_t)(LL * KC + K - 1) * nc + m]);
      rec_end(&q);
    }
  }
}

void orc_free(void *p) { free(p); }

/* RESTOUT(0) (EFDC/restartmod.f90:86-272), the records the library owns */
int orc_restart_text(orc_ctx *o, const efdc_restart_t *r, char **text, int64_t *bytes) {
  if (writes(r, 4) || writes(r, 5) || writes(r, 7)) return EFDC_ERR_ARG;
  const int LA = o->LA, KC = o->KC, LCM = o->LCM;
  sbuf b = {0};
  char line[128];
  int n = snprintf(line, sizeof line, "%20lld    ", (long long)r->n);
  put(&b, line, (size_t)n);
  n = snprintf(line, sizeof line, "%12.4f", r->timeday);
  if (n > 12) put(&b, "************", 12);
  else put(&b, line, 12);
  n = snprintf(line, sizeof line, "%10d\n", 720);
  put(&b, line, (size_t)n);
  for (int i = 0; i < 9; ++i) {
    char f[6];
    fmt_i(r->isco[i], 6, f);
    put(&b, f, 6);
  }
  put(&b, "\n", 1);
  const double *HP = A(EFDC_HP), *H1P = A(EFDC_H1P), *BELV = A(EFDC_BELV);
  const double *U = A(EFDC_U), *U1 = A(EFDC_U1), *V = A(EFDC_V), *V1 = A(EFDC_V1);
  rec q = {&b, 0};
  for (int L = 2; L <= LA; ++L) {
    rec_e(&q, HP[HX(L)]); rec_e(&q, H1P[HX(L)]); rec_e(&q, r->hwq[HX(L)]);
    rec_e(&q, r->h2wq[HX(L)]); rec_e(&q, BELV[HX(L)]);
    rec_end(&q);
    rec_e(&q, A(EFDC_UHDYE)[HX(L)]); rec_e(&q, A(EFDC_UHDY1E)[HX(L)]);
    rec_e(&q, A(EFDC_VHDXE)[HX(L)]); rec_e(&q, A(EFDC_VHDX1E)[HX(L)]);
    rec_end(&q);
    const double *lay[4] = {U, U1, V, V1};
    for (int a = 0; a < 4; ++a) {
      for (int K = 1; K <= KC; ++K) rec_e(&q, lay[a] ? lay[a][KX(L, K)] : 0.0);
      rec_end(&q);
    }
    const double *tk[5] = {r->qq, r->qq1, r->qql, r->qql1, r->dml};
    for (int a = 0; a < 5; ++a) {
      for (int K = 0; K <= KC; ++K) rec_e(&q, tk[a][WX(L, K)]);
      rec_end(&q);
    }
    if (r->istran[2] >= 1 && r->isice > 2) {
      rec_e(&q, r->ice_thick[HX(L)]); rec_e(&q, r->ice_temp[HX(L)]);
      rec_end(&q);
    }
    if (writes(r, 1)) {
      for (int K = 1; K <= KC; ++K) rec_e(&q, o->c[r->slot[0]].con[KX(L, K)]);
      rec_end(&q);
      for (int K = 1; K <= KC; ++K) rec_e(&q, o->c[r->slot[0]].con1[KX(L, K)]);
      rec_end(&q);
    }
    if (writes(r, 2)) {
      rec_e(&q, r->temb[HX(L)]);
      for (int K = 1; K <= KC; ++K) rec_e(&q, o->c[r->slot[1]].con[KX(L, K)]);
      rec_end(&q);
      for (int K = 1; K <= KC; ++K) rec_e(&q, o->c[r->slot[1]].con1[KX(L, K)]);
      rec_end(&q);
    }
    if (writes(r, 3)) {
      for (int K = 1; K <= KC; ++K) rec_e(&q, o->c[r->slot[2]].con[KX(L, K)]);
      rec_end(&q);
      for (int K = 1; K <= KC; ++K) rec_e(&q, o->c[r->slot[2]].con1[KX(L, K)]);
      rec_end(&q);
    }
    if (writes(r, 6)) {
      const size_t B = (size_t)LCM * r->kbm;
      for (int s = 0; s < r->nsed; ++s) {
        const orc_const *c = &o->c[r->sed_slot[s]];
        for (int K = 1; K <= r->kb; ++K) rec_e(&q, r->sedb[s * B + KX(L, K)]);
        rec_end(&q);
        for (int K = 1; K <= KC; ++K) rec_e(&q, c->con[KX(L, K)]);
        rec_end(&q);
        for (int K = 1; K <= r->kb; ++K) rec_e(&q, r->sedb1[s * B + KX(L, K)]);
        rec_end(&q);
        for (int K = 1; K <= KC; ++K) rec_e(&q, c->con1[KX(L, K)]);
        rec_end(&q);
      }
      const double *bed[4] = {r->hbed, r->hbed1, r->vdrbed, r->vdrbed1};
      for (int a = 0; a < 4; ++a) {
        for (int K = 1; K <= r->kb; ++K) rec_e(&q, bed[a][KX(L, K)]);
        rec_end(&q);
      }
    }
  }
  for (int m = 1; m <= 3; ++m)
    if (writes(r, m)) bc_group(o, r, r->slot[m - 1], &b);
  if (writes(r, 6))
    for (int s = 0; s < r->nsed; ++s) bc_group(o, r, r->sed_slot[s], &b);
  for (int L = 2; L <= LA; ++L) {
    rec_e(&q, A(EFDC_QSUME)[HX(L)]);
    for (int K = 1; K <= KC; ++K) rec_e(&q, A(EFDC_QSUM)[KX(L, K)]);
    rec_end(&q);
  }
  *bytes = (int64_t)b.n;
  *text = b.p;  /* released with orc_free */
  return EFDC_OK;
}

static int wfile(const char *dir, const char *name, const char *mode, const void *p, size_t n,
                 FILE **keep) {
  char path[4096];
  snprintf(path, sizeof path, "%s/%s", dir, name);
  FILE *f = fopen(path, mode);
  if (!f) return EFDC_ERR_IO;
  if (n && fwrite(p, 1, n, f) != n) { fclose(f); return EFDC_ERR_IO; }
  if (keep) { *keep = f; return EFDC_OK; }
  return fclose(f) == 0 ? EFDC_OK : EFDC_ERR_IO;
}

static int64_t cell3d_of(const orc_ctx *o, const efdc_ee_grid_t *g) {
  int64_t n = 0;
  for (int L = 2; L <= o->LA; ++L) n += o->KC - (g->ksz ? g->ksz[L - 1] : 1) + 1;
  return n;
}

/* HEADEROUT for EE_WS.OUT and EE_VEL.OUT (EFDC/efdcout.f90:73-113) */
int orc_ee_open(orc_ctx *o, const char *dir, const efdc_ee_grid_t *g) {
  const int32_t NACTIVE = o->LA - 1;
  const int64_t CELL3D = cell3d_of(o, g);
  int32_t ws[6] = {8400, 6 * 4, (2 + 3 + NACTIVE) * 4, g->ic, g->jc, NACTIVE};
  int st = wfile(dir, "EE_WS.OUT", "wb", ws, sizeof ws, NULL);
  if (st) return st;
  int32_t vh[9] = {8400, (9 + 4 * o->LCM) * 4, (int32_t)((2 + 2 + 3 * CELL3D) * 4), g->igridv,
                   (int32_t)CELL3D, g->ic, g->jc, o->KC, NACTIVE};
  FILE *f = NULL;
  if ((st = wfile(dir, "EE_VEL.OUT", "wb", vh, sizeof vh, &f))) return st;
  const double *rs[4] = {g->rssbce, g->rssbcw, g->rssbcs, g->rssbcn};
  for (int a = 0; a < 4; ++a)
    for (int L = 1; L <= o->LCM; ++L) {
      float x = (float)rs[a][HX(L)];
      fwrite(&x, 4, 1, f);
    }
  return fclose(f) == 0 ? EFDC_OK : EFDC_ERR_IO;
}

/* WSOUT and VELOUT, a file that does not continue an earlier run
   (EFDC/efdcout.f90:418-426,487-493) */
int orc_ee_append(orc_ctx *o, const char *dir, const efdc_ee_grid_t *g, int32_t n, double eetime,
                  double delt, int32_t lminstep) {
  FILE *f = NULL;
  int st;
  const float d4 = (float)delt;
  if ((st = wfile(dir, "EE_WS.OUT", "ab", &n, 4, &f))) return st;
  fwrite(&eetime, 8, 1, f);
  fwrite(&d4, 4, 1, f);
  fwrite(&lminstep, 4, 1, f);
  for (int L = 2; L <= o->LA; ++L) {
    float x = (float)A(EFDC_HP)[HX(L)];
    fwrite(&x, 4, 1, f);
  }
  if (fclose(f) != 0) return EFDC_ERR_IO;
  if ((st = wfile(dir, "EE_VEL.OUT", "ab", &n, 4, &f))) return st;
  fwrite(&eetime, 8, 1, f);
  fwrite(&d4, 4, 1, f);
  const double *lay[2] = {A(EFDC_U), A(EFDC_V)};
  for (int a = 0; a < 2; ++a)
    for (int L = 2; L <= o->LA; ++L)
      for (int K = g->ksz ? g->ksz[L - 1] : 1; K <= o->KC; ++K) {
        float x = lay[a] ? (float)lay[a][KX(L, K)] : 0.0f;
        fwrite(&x, 4, 1, f);
      }
  for (int L = 2; L <= o->LA; ++L)
    for (int K = g->ksz ? g->ksz[L - 1] : 1; K <= o->KC; ++K) {
      float x = A(EFDC_W) ? (float)A(EFDC_W)[WX(L, K)] : 0.0f;
      fwrite(&x, 4, 1, f);
    }
  return fclose(f) == 0 ? EFDC_OK : EFDC_ERR_IO;
}
