// The restart writer's E17.8 / I10 formatter (efdcplus8.5_amd/csrc/efdc_fmt.h)
// compiled for the host, so that the CPU tests can run the device code path
// over millions of values.  Built by tests/test_oracle_io.py.
#include <cstdint>

#include "efdc_fmt.h"

static const double h_pow10[EFDC_POW10_MAX - EFDC_POW10_MIN + 1][2] = EFDC_POW10_INIT;

// out: 17 bytes per value; handed[i] = 1 where the formatter hands the value
// to the exact host path (out is then left unspecified)
extern "C" int64_t fmt_e17_many(const double *v, int64_t n, char *out, unsigned char *handed) {
  int64_t k = 0;
  for (int64_t i = 0; i < n; ++i) {
    const bool ok = efdc_fmt::fmt_e17(h_pow10, v[i], out + 17 * i);
    handed[i] = ok ? 0 : 1;
    k += ok ? 0 : 1;
  }
  return k;
}

extern "C" void fmt_i10_many(const long long *v, int64_t n, char *out) {
  for (int64_t i = 0; i < n; ++i) efdc_fmt::fmt_i10(v[i], out + 10 * i);
}
