// spk_host.h -- host-side helpers shared by the translation units of
// libspk_hip.so: error reporting, lattice geometry and Metropolis rule setup.
#pragma once

#include <hip/hip_runtime.h>

#include "spk_device.h"
#include "spk_master_rng.h"

namespace spk {

// records a printf-style message as the thread's last error; returns code
int set_error(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
// SPK_HIP_OK for hipSuccess, else SPK_HIP_EDEVICE with the HIP message
int check_hip(hipError_t e, const char *what);
// validates a lattice description and fills the colour-blocked geometry
int make_geom(const spk_hip_desc *d, Geom *g);
// Metropolis thresholds and proposal parameters for one temperature
int make_rule(int app, int nspins, int lattice, double temperature, Rule *r);

enum SpinType { kU8 = 0, kU16 = 1, kI32 = 2 };

// narrowest storage holding 1..Q
// Largest spin value of an app: Q for the Potts styles, 2 for the Ising ones.
inline bool potts_family(int app) {
  return app == SPK_HIP_POTTS || app == SPK_HIP_POTTS_NEIGH || app == SPK_HIP_POTTS_NEIGHONLY;
}
inline int app_qmax(int app, int nspins) { return potts_family(app) ? nspins : 2; }
// potts and ising run the packed per-colour kernels; the proposal variants
// run the ordered one-site-per-thread kernel
inline bool variant_app(int app) { return app >= SPK_HIP_POTTS_NEIGH && app <= SPK_HIP_ISING_SINGLE; }

inline int spin_type_for(int app, int nspins) {
  const int qmax = app_qmax(app, nspins);
  return qmax <= 255 ? kU8 : qmax <= 65535 ? kU16 : kI32;
}
inline size_t spin_bytes(int stype) { return stype == kU8 ? 1 : stype == kU16 ? 2 : 4; }

// Calls f.template operator()<T>() with the spin storage type.
template <typename F>
void with_type(int stype, F &&f) {
  switch (stype) {
    case kU8: f.template operator()<uint8_t>(); break;
    case kU16: f.template operator()<uint16_t>(); break;
    default: f.template operator()<int32_t>(); break;
  }
}

// Calls f.template operator()<LAT>() for the lattice style.
template <typename F>
void with_lattice(int lat, F &&f) {
  switch (lat) {
    case SPK_HIP_SQ_4N: f.template operator()<SPK_HIP_SQ_4N>(); break;
    case SPK_HIP_SQ_8N: f.template operator()<SPK_HIP_SQ_8N>(); break;
    case SPK_HIP_SC_6N: f.template operator()<SPK_HIP_SC_6N>(); break;
    default: f.template operator()<SPK_HIP_SC_26N>(); break;
  }
}

// 16807^(2^b) mod m, b = 0..31 (jump-ahead for `set site range`)
struct PowTable { uint32_t p[32]; };
inline PowTable park_powers() {
  PowTable pw;
  uint32_t p = kIA;
  for (int b = 0; b < 32; b++) {
    pw.p[b] = p;
    p = pm_mulmod(p, p);
  }
  return pw;
}
// 16807^100 mod m: tagreset's 100 warm-up steps
inline uint32_t park_warmup100() {
  uint32_t a = 1;
  for (int w = 0; w < 100; w++) a = pm_mulmod(a, kIA);
  return a;
}

}  // namespace spk
