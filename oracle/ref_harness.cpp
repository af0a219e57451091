// ref_harness.cpp -- C entry points over the reference's own RandomPark
// (src/random_park.{h,cpp}, compiled unmodified from the reference tree by
// oracle/Makefile into oracle/_ref/).  TEST INFRASTRUCTURE ONLY: used by
// tests/ to pin the oracle's Park-Miller restatement against the real thing.
// random_park.cpp is the only reference translation unit on the hot path that
// builds from its own sources; the rest needs generated style headers.
#include <cstdint>
#include "random_park.h"

using SPPARKS_NS::RandomPark;

extern "C" {

// n successive uniform() draws from `seed`; writes each double and the
// generator state after it.
void ref_uniform_seq(int seed, int n, double *u, int *state)
{
  RandomPark r(seed);
  for (int i = 0; i < n; i++) {
    u[i] = r.uniform();
    state[i] = r.seed;
  }
}

// n successive irandom(range) draws from `seed`
void ref_irandom_seq(int seed, int range, int n, int *out)
{
  RandomPark r(seed);
  for (int i = 0; i < n; i++) out[i] = r.irandom(range);
}

// RandomPark(double) constructor result
int ref_from_double(double rseed)
{
  RandomPark r(rseed);
  return r.seed;
}

// tagreset(rseed, offset, warmup) for offsets[0..n)
void ref_tagreset(double rseed, const int64_t *offsets, int n, int warmup,
                  int *out)
{
  RandomPark r(1);
  for (int i = 0; i < n; i++) {
    r.tagreset(rseed, (SPPARKS_NS::tagint)offsets[i], warmup);
    out[i] = r.seed;
  }
}

}
