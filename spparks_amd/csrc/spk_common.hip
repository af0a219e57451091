// spk_common.hip -- host helpers of libspk_hip.so (see spk_host.h) and the
// host-only ABI entries spk_hip_last_error / spk_hip_metropolis_table.
#include <cmath>
#include <cstdarg>
#include <cstdio>

#include "spk_host.h"

namespace spk {

namespace {
thread_local char g_err[512] = "";
}

int set_error(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int check_hip(hipError_t e, const char *what) {
  if (e == hipSuccess) return SPK_HIP_OK;
  return set_error(SPK_HIP_EDEVICE, "%s: %s", what, hipGetErrorString(e));
}

int make_geom(const spk_hip_desc *d, Geom *g) {
  if (!d) return set_error(SPK_HIP_EINVAL, "null lattice description");
  if (d->app < SPK_HIP_POTTS || d->app > SPK_HIP_ISING_SINGLE) return set_error(SPK_HIP_EINVAL, "unknown app %d", d->app);
  const int lat = d->lattice;
  const bool is3 = lat == SPK_HIP_SC_6N || lat == SPK_HIP_SC_26N;
  if (lat < SPK_HIP_SQ_4N || lat > SPK_HIP_SC_26N) return set_error(SPK_HIP_EINVAL, "unknown lattice %d", lat);
  if (d->nx < 1 || d->ny < 1 || d->nz < 1) return set_error(SPK_HIP_EINVAL, "lattice edges must be >= 1");
  if (!is3 && d->nz != 1) return set_error(SPK_HIP_EINVAL, "2d lattice needs nz = 1");
  // Lattice::ncolors (src/lattice.cpp:199-220)
  if (d->nx % 2 || d->ny % 2 || (is3 && d->nz % 2))
    return set_error(SPK_HIP_EINVAL, "Color stencil is incommensurate with lattice size");
  const long long n = (long long)d->nx * d->ny * d->nz;
  if (n > 2147483647LL) return set_error(SPK_HIP_EINVAL, "lattice exceeds 2^31-1 sites");
  geom_init(*g, lat, d->nx, d->ny, d->nz);
  if ((long long)g->Mp * g->K > 2147483647LL) return set_error(SPK_HIP_EINVAL, "lattice exceeds 2^31-1 slots");
  return SPK_HIP_OK;
}

// smallest s in [1, 2^31-1] with kAM*s > x; 2^31-1 when no state qualifies
static int32_t first_state_above(double x) {
  int32_t lo = 1, hi = kIM;
  while (lo < hi) {
    int32_t mid = lo + (hi - lo) / 2;
    if (kAM * (double)mid > x) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

// smallest s with kAM*s >= x
static int32_t first_state_at_least(double x) {
  int32_t lo = 1, hi = kIM;
  while (lo < hi) {
    int32_t mid = lo + (hi - lo) / 2;
    if (kAM * (double)mid >= x) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

// exp is called through a volatile pointer so the compiler can neither fold
// nor vectorise it: the value must be this process's libm exp, the function
// the reference calls at src/app_potts.cpp:130.
static double (*volatile g_exp)(double) = ::exp;

int make_rule(int app, int nspins, int lattice, double temperature, Rule *r) {
  if (potts_family(app) && nspins <= 0) return set_error(SPK_HIP_EINVAL, "Potts needs nspins > 0");
  if (!(temperature >= 0.0) || std::isinf(temperature))
    return set_error(SPK_HIP_EINVAL, "temperature must be finite and >= 0");
  *r = Rule{};
  r->app = app;
  r->nspins = app_qmax(app, nspins);
  r->zero_t = temperature == 0.0;
  r->half = first_state_at_least(0.5);
  const int maxd = maxneigh_of(lattice);
  if (!r->zero_t) {
    const double t_inverse = 1.0 / temperature;   // set_temperature, app_lattice.cpp:897
    for (int d = 1; d <= maxd; d++) {
      const double ediff = (double)(-d);          // einitial - efinal
      r->thr[d] = first_state_above(g_exp(ediff * t_inverse));
    }
  }
  return SPK_HIP_OK;
}

}  // namespace spk

extern "C" {

const char *spk_hip_last_error(void) { return spk::g_err; }

int spk_hip_metropolis_table(double temperature, int maxd, double *tab, int32_t *thr) {
  using namespace spk;
  if (maxd < 0 || maxd > 26) return set_error(SPK_HIP_EINVAL, "maxd outside [0, 26]");
  if (!(temperature > 0.0) || std::isinf(temperature))
    return set_error(SPK_HIP_EINVAL, "table needs a finite temperature > 0");
  const double t_inverse = 1.0 / temperature;
  if (tab) tab[0] = 1.0;
  if (thr) thr[0] = first_state_at_least(0.5);
  for (int d = 1; d <= maxd; d++) {
    const double v = g_exp((double)(-d) * t_inverse);
    if (tab) tab[d] = v;
    if (thr) thr[d] = first_state_above(v);
  }
  return SPK_HIP_OK;
}

// Marsaglia-Zaman universal generator as seeded by the `seed` command
int spk_hip_master_draws(int seed, int n, double *out) {
  using namespace spk;
  if (seed <= 0) return set_error(SPK_HIP_EINVAL, "Illegal seed command");
  if (n < 0 || (n > 0 && !out)) return set_error(SPK_HIP_EINVAL, "bad draw buffer");
  MasterRng m(seed);
  for (int i = 0; i < n; i++) out[i] = m.uniform();
  return SPK_HIP_OK;
}

}  // extern "C"
