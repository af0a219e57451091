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

}  // namespace spk
