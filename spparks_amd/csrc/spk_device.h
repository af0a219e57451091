// spk_device.h -- device-side building blocks of the strict colour sweep:
// Park-Miller arithmetic, the colour-blocked lattice geometry, neighbour
// gathers for each lattice style and the Potts / Ising rejection rule.
//
// Included by the .hip translation units of libspk_hip.so only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "spk_hip.h"

namespace spk {

// ---------------------------------------------------------------- RNG
// Park-Miller minimal standard: s' = 16807*s mod (2^31-1).  The reference
// evaluates it with Schrage's decomposition (src/random_park.cpp:79-86); both
// give the exact residue, here via one 64-bit product and a Mersenne fold.
constexpr int32_t kIM = 2147483647;
constexpr uint32_t kIA = 16807u;
constexpr double kAM = 1.0 / 2147483647.0;   // same constant as AM*seed

__host__ __device__ __forceinline__ int32_t pm_step(int32_t s) {
  uint64_t x = (uint64_t)(uint32_t)s * kIA;            // < 2^46
  uint32_t r = (uint32_t)(x & 0x7fffffffu) + (uint32_t)(x >> 31);
  if (r >= (uint32_t)kIM) r -= (uint32_t)kIM;
  return (int32_t)r;
}

// a*b mod (2^31-1) for a,b in [0, 2^31-1)
__host__ __device__ __forceinline__ uint32_t pm_mulmod(uint32_t a, uint32_t b) {
  uint64_t x = (uint64_t)a * b;                        // < 2^62
  uint64_t t = (x & 0x7fffffffull) + (x >> 31);        // < 2^32
  t = (t & 0x7fffffffull) + (t >> 31);
  if (t >= (uint64_t)kIM) t -= (uint64_t)kIM;
  return (uint32_t)t;
}

// ------------------------------------------------------------ geometry
// Colour-blocked layout.  Colour c (0-based; the reference's colour c+1,
// src/lattice.cpp:234-283) occupies [c*M, (c+1)*M).  Within a colour the
// sites form a dense (hx, hy, hz) sub-grid, x fastest:
//   sq/4n : hx=nx/2 hy=ny   hz=1    site (2h + ((j+c)&1), j)
//   sc/6n : hx=nx/2 hy=ny   hz=nz   site (2h + ((j+k+c)&1), j, k)
//   sq/8n : hx=nx/2 hy=ny/2 hz=1    c = 2(j&1)+(i&1),          site (2a+pi, 2b+pj)
//   sc/26n: hx=nx/2 hy=ny/2 hz=nz/2 c = 4(k&1)+2(j&1)+(i&1),   site (2a+pi, 2b+pj, 2e+pk)
struct Geom {
  int lat;
  int nx, ny, nz;
  int hx, hy, hz;
  int K;       // colours
  int M;       // sites per colour
};

__host__ __device__ inline int maxneigh_of(int lat) {
  return lat == SPK_HIP_SQ_4N ? 4 : lat == SPK_HIP_SQ_8N ? 8 : lat == SPK_HIP_SC_6N ? 6 : 26;
}

// natural (i,j,k) -> colour-blocked offset
__host__ __device__ __forceinline__ int64_t blocked_index(const Geom &g, int i, int j, int k) {
  int c;
  int64_t local;
  switch (g.lat) {
    case SPK_HIP_SQ_4N:
      c = (i + j) & 1;
      local = (i >> 1) + (int64_t)g.hx * j;
      break;
    case SPK_HIP_SC_6N:
      c = (i + j + k) & 1;
      local = (i >> 1) + (int64_t)g.hx * (j + (int64_t)g.ny * k);
      break;
    case SPK_HIP_SQ_8N:
      c = 2 * (j & 1) + (i & 1);
      local = (i >> 1) + (int64_t)g.hx * (j >> 1);
      break;
    default:
      c = 4 * (k & 1) + 2 * (j & 1) + (i & 1);
      local = (i >> 1) + (int64_t)g.hx * ((j >> 1) + (int64_t)g.hy * (k >> 1));
      break;
  }
  return (int64_t)c * g.M + local;
}

__host__ __device__ __forceinline__ int wrap_dec(int a, int n) { return a == 0 ? n - 1 : a - 1; }
__host__ __device__ __forceinline__ int wrap_inc(int a, int n) { return a == n - 1 ? 0 : a + 1; }

// Gather the neighbour spins of sub-grid site (a,b,e) of colour c into nb[].
// Returns the neighbour count.  The neighbour multiset equals the
// reference's periodic Moore / von Neumann shell
// (src/create_sites.cpp:589-640, :1303-1371), duplicates included on edges
// of length 2.
template <int LAT>
__host__ __device__ __forceinline__ void gather(const int32_t *__restrict__ spin, const Geom &g,
                                       int c, int a, int b, int e, int32_t *nb) {
  if constexpr (LAT == SPK_HIP_SQ_4N) {
    const int32_t *o = spin + (int64_t)(1 - c) * g.M;
    int p = (b + c) & 1;
    int xo = p ? wrap_inc(a, g.hx) : wrap_dec(a, g.hx);
    int64_t row = (int64_t)g.hx * b;
    nb[0] = o[a + row];
    nb[1] = o[xo + row];
    nb[2] = o[a + (int64_t)g.hx * wrap_dec(b, g.hy)];
    nb[3] = o[a + (int64_t)g.hx * wrap_inc(b, g.hy)];
  } else if constexpr (LAT == SPK_HIP_SC_6N) {
    const int32_t *o = spin + (int64_t)(1 - c) * g.M;
    int p = (b + e + c) & 1;
    int xo = p ? wrap_inc(a, g.hx) : wrap_dec(a, g.hx);
    int64_t hxy = (int64_t)g.hx * g.hy;
    int64_t row = (int64_t)g.hx * b + hxy * e;
    nb[0] = o[a + row];
    nb[1] = o[xo + row];
    nb[2] = o[a + (int64_t)g.hx * wrap_dec(b, g.hy) + hxy * e];
    nb[3] = o[a + (int64_t)g.hx * wrap_inc(b, g.hy) + hxy * e];
    nb[4] = o[a + (int64_t)g.hx * b + hxy * wrap_dec(e, g.hz)];
    nb[5] = o[a + (int64_t)g.hx * b + hxy * wrap_inc(e, g.hz)];
  } else if constexpr (LAT == SPK_HIP_SQ_8N) {
    const int pi = c & 1, pj = c >> 1;
    // x index per offset: dx=0 -> a, dx=-1 -> a+pi-1, dx=+1 -> a+pi
    int xs[3] = {pi ? a : wrap_dec(a, g.hx), a, pi ? wrap_inc(a, g.hx) : a};
    int ys[3] = {pj ? b : wrap_dec(b, g.hy), b, pj ? wrap_inc(b, g.hy) : b};
    int n = 0;
#pragma unroll
    for (int dy = 0; dy < 3; dy++)
#pragma unroll
      for (int dx = 0; dx < 3; dx++) {
        if (dx == 1 && dy == 1) continue;
        int q = (pi ^ (dx != 1)) + 2 * (pj ^ (dy != 1));
        nb[n++] = spin[(int64_t)q * g.M + xs[dx] + (int64_t)g.hx * ys[dy]];
      }
  } else {
    const int pi = c & 1, pj = (c >> 1) & 1, pk = c >> 2;
    int xs[3] = {pi ? a : wrap_dec(a, g.hx), a, pi ? wrap_inc(a, g.hx) : a};
    int ys[3] = {pj ? b : wrap_dec(b, g.hy), b, pj ? wrap_inc(b, g.hy) : b};
    int zs[3] = {pk ? e : wrap_dec(e, g.hz), e, pk ? wrap_inc(e, g.hz) : e};
    int n = 0;
#pragma unroll
    for (int dz = 0; dz < 3; dz++)
#pragma unroll
      for (int dy = 0; dy < 3; dy++)
#pragma unroll
        for (int dx = 0; dx < 3; dx++) {
          if (dx == 1 && dy == 1 && dz == 1) continue;
          int q = (pi ^ (dx != 1)) + 2 * (pj ^ (dy != 1)) + 4 * (pk ^ (dz != 1));
          nb[n++] = spin[(int64_t)q * g.M + xs[dx] +
                         (int64_t)g.hx * (ys[dy] + (int64_t)g.hy * zs[dz])];
        }
  }
}

template <int LAT> struct NeighCount { static constexpr int value = 26; };
template <> struct NeighCount<SPK_HIP_SQ_4N> { static constexpr int value = 4; };
template <> struct NeighCount<SPK_HIP_SC_6N> { static constexpr int value = 6; };
template <> struct NeighCount<SPK_HIP_SQ_8N> { static constexpr int value = 8; };

// ------------------------------------------------------------ the rule
// Metropolis parameters of one lattice.  thr[d] is the smallest seed s with
// AM*s > exp(-d/kT) (the reference's `u2 > exp((ei-ef)*t_inverse)` rejection,
// src/app_potts.cpp:130), so "reject" == (s >= thr[d]) exactly; half is the
// smallest s with AM*s >= 0.5 (Ising proposal, src/app_ising.cpp:112).
struct Rule {
  int app;          // SPK_HIP_POTTS / SPK_HIP_ISING
  int nspins;       // Q
  int zero_t;       // temperature == 0.0
  int half;
  int thr[28];      // index 1..maxneigh
};

struct SiteOutcome {
  int32_t spin;     // new value (== old when rejected)
  int32_t seed;     // advanced generator state
  int changed;      // 1 when the spin changed (naccept++)
  int dE;           // ef - ei for a change, else 0
};

// One strict site event: AppPotts/AppIsing::site_event_rejection
// (src/app_potts.cpp:112-146, src/app_ising.cpp:104-136) on gathered
// neighbours; thr points at the Metropolis thresholds (LDS copy).
template <int NN>
__host__ __device__ __forceinline__ SiteOutcome site_event(const Rule &r, const int *thr,
                                                  int32_t old, int32_t seed,
                                                  const int32_t *nb) {
  int ei = 0;
#pragma unroll
  for (int m = 0; m < NN; m++) ei += (nb[m] != old);
  int32_t s = pm_step(seed);
  int32_t prop;
  if (r.app == SPK_HIP_POTTS) {
    double u = kAM * (double)s;
    prop = (int32_t)((double)r.nspins * u) + 1;
    if (prop > r.nspins) prop = r.nspins;
  } else {
    prop = s < r.half ? 1 : 2;
  }
  int ef = 0;
#pragma unroll
  for (int m = 0; m < NN; m++) ef += (nb[m] != prop);
  bool accept = true;
  if (ef > ei) {
    if (r.zero_t) {
      accept = false;
    } else {
      s = pm_step(s);
      accept = s < thr[ef - ei];
    }
  }
  SiteOutcome o;
  o.changed = accept && prop != old;
  o.spin = o.changed ? prop : old;
  o.seed = s;
  o.dE = o.changed ? ef - ei : 0;
  return o;
}

// --------------------------------------------------------- reductions
// Wavefront (64-lane) sum, then one value per workgroup via LDS.
__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}
__device__ __forceinline__ long long wave_sum64(long long v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Adds (x, y) of every thread of the block into acc[2*slot], acc[2*slot+1].
// All threads of the block must call it.  Needs blockDim.x % 64 == 0.
__device__ __forceinline__ void block_accumulate(long long x, long long y,
                                                 unsigned long long *acc, int slot) {
  __shared__ long long red[2][16];
  x = wave_sum64(x);
  y = wave_sum64(y);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) { red[0][wid] = x; red[1][wid] = y; }
  __syncthreads();
  if (threadIdx.x == 0) {
    long long sx = 0, sy = 0;
    const int nw = blockDim.x >> 6;
    for (int w = 0; w < nw; w++) { sx += red[0][w]; sy += red[1][w]; }
    if (sx) atomicAdd(&acc[2 * slot], (unsigned long long)sx);
    if (sy) atomicAdd(&acc[2 * slot + 1], (unsigned long long)sy);
  }
}

constexpr int kAccSlots = 64;   // spread of per-sweep accumulator slots

}  // namespace spk
