// spk_kernels.h -- kernels shared by the single-lattice and batched paths of
// libspk_hip.so: energy reduction, layout conversion, device-side initial
// state.  Each translation unit gets its own copies (anonymous namespace).
#pragma once

#include "spk_device.h"
#include "spk_host.h"

namespace spk {
namespace {

// Total energy: sum over all sites of #unlike neighbours.
template <int LAT, typename T>
__global__ __launch_bounds__(256) void k_energy(const T *__restrict__ spin, Geom g,
                                                unsigned long long *__restrict__ acc) {
  constexpr int NN = NeighCount<LAT>::value;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int eng = 0;
  if (t < (int64_t)g.M * g.K) {
    const int c = (int)(t / g.M);
    const int ti = (int)(t - (int64_t)c * g.M);
    const int a = ti % g.hx;
    const int rest = ti / g.hx;
    int32_t nb[NN];
    gather<LAT, T>(spin, g, c, a, rest % g.hy, rest / g.hy, nb);
    const int32_t me = spin[(int64_t)c * g.Mp + a + (int64_t)g.px * rest];
#pragma unroll
    for (int m = 0; m < NN; m++) eng += (nb[m] != me);
  }
  block_accumulate(eng, 0, acc, blockIdx.x & (kAccSlots - 1));
}

__device__ __forceinline__ void natural_ijk(const Geom &g, int64_t t, int &i, int &j, int &k) {
  i = (int)(t % g.nx);
  const int64_t r = t / g.nx;
  j = (int)(r % g.ny);
  k = (int)(r / g.ny);
}

// natural order <-> colour-blocked order (with the storage type change)
template <typename T>
__global__ void k_to_blocked(const int32_t *__restrict__ nat, T *__restrict__ blk, Geom g) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)g.M * g.K) return;
  int i, j, k;
  natural_ijk(g, t, i, j, k);
  blk[blocked_index(g, i, j, k)] = (T)nat[t];
}

template <typename T>
__global__ void k_to_natural(const T *__restrict__ blk, int32_t *__restrict__ nat, Geom g) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)g.M * g.K) return;
  int i, j, k;
  natural_ijk(g, t, i, j, k);
  nat[t] = (int32_t)blk[blocked_index(g, i, j, k)];
}

// Compacts accumulator rows: out[2r], out[2r+1] = the sums of row r's
// `slots` slots (row r at acc + r*row_words, slot k at +k*kSlotStride).
__global__ __launch_bounds__(64) void k_fold_rows(const unsigned long long *__restrict__ acc, int slots,
                                                  int row_words, unsigned long long *__restrict__ out) {
  // one wave per row: lane k holds slot k (and k+64, ...), so a row costs
  // one round trip to L2 instead of `slots` dependent ones
  const unsigned long long *p = acc + (size_t)blockIdx.x * row_words;
  unsigned long long x = 0, y = 0;
  for (int k = threadIdx.x; k < slots; k += 64) {
    x += p[(size_t)k * kSlotStride];
    y += p[(size_t)k * kSlotStride + 1];
  }
  x = wave_sum64(x);
  y = wave_sum64(y);
  if (threadIdx.x == 0) {
    out[2 * (size_t)blockIdx.x] = x;
    out[2 * (size_t)blockIdx.x + 1] = y;
  }
}

// counts values outside [lo, hi]
__global__ void k_count_outside(const int32_t *__restrict__ v, int64_t n, int lo, int hi,
                                unsigned long long *__restrict__ bad) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int out = 0;
  if (t < n) out = (v[t] < lo || v[t] > hi);
  out = wave_sum(out);
  if ((threadIdx.x & 63) == 0 && out) atomicAdd(bad, (unsigned long long)out);
}

// `set site range lo hi`, loop all: site id gid = t+1 draws
// irandom(range) from a generator advanced gid times from seed0
// (src/set.cpp:598-627); 16807^gid mod m from the binary powers in pw.

template <typename T>
__global__ void k_set_range(T *__restrict__ spin, Geom g, int seed0, int lo, int range, PowTable pw) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)g.M * g.K) return;
  const uint64_t gid = (uint64_t)t + 1;
  uint32_t s = (uint32_t)seed0;
#pragma unroll
  for (int bit = 0; bit < 32; bit++)
    if ((gid >> bit) & 1u) s = pm_mulmod(s, pw.p[bit]);
  const double u = kAM * (double)(int32_t)s;      // RandomPark::uniform
  int v = (int)(u * (double)range) + 1;           // irandom(range)
  if (v > range) v = range;
  int i, j, k;
  natural_ijk(g, t, i, j, k);
  spin[blocked_index(g, i, j, k)] = (T)(v - 1 + lo);
}

template <typename T>
__global__ void k_fill(T *__restrict__ v, int64_t n, int32_t value) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) v[t] = (T)value;
}

// siteseeds = tagreset(rseed, id, 100) (src/random_park.cpp:67-73).
// xbase = rseed*IM is formed on the host; x = xbase + id is one rounded
// addition as in the reference (no contraction: built with
// -ffp-contract=off).  For 0 <= x < 2*IM, fmod(x, IM) is x or the exact
// difference x-IM (Sterbenz).  The 100 warm-up steps are one multiply by
// 16807^100 mod m.
__global__ void k_init_seeds(int32_t *__restrict__ seed, Geom g, double xbase, uint32_t a100) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)g.M * g.K) return;
  const double im = (double)kIM;
  double x = xbase + (double)(t + 1);
  while (x >= im) x -= im;
  int32_t s = (int32_t)x;
  if (s == 0) s = 1;
  s = (int32_t)pm_mulmod((uint32_t)s, a100);
  int i, j, k;
  natural_ijk(g, t, i, j, k);
  seed[blocked_index(g, i, j, k)] = s;
}


// Layout-independent state checksum: sum over sites g (0-based, reference
// order) of mix64(g*C1 + spin*C2 + seed), wrapping mod 2^64; mix64 is the
// splitmix64 finaliser.  Equal (spin, seed) arrays give equal sums.
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

template <typename T>
__global__ void k_checksum(const T *__restrict__ spin, const int32_t *__restrict__ seed, Geom g,
                           unsigned long long *__restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t h = 0;
  if (t < (int64_t)g.M * g.K) {
    int i, j, k;
    natural_ijk(g, t, i, j, k);
    const int64_t p = blocked_index(g, i, j, k);
    h = mix64((uint64_t)t * 0xD1B54A32D192ED03ull + (uint64_t)(uint32_t)(int32_t)spin[p] * 0xABC98388FB8FAC03ull +
              (uint64_t)(uint32_t)seed[p]);
  }
  h = wave_sum64(h);
  if ((threadIdx.x & 63) == 0) atomicAdd(out, (unsigned long long)h);
}

}  // namespace
}  // namespace spk
