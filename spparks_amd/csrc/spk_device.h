// spk_device.h -- device-side building blocks of the strict colour sweep:
// Park-Miller arithmetic, the colour-blocked lattice geometry, neighbour
// access for each lattice style and the Potts / Ising rejection rule.
//
// Everything here that touches lattice memory is __host__ __device__ so the
// same code can be run on the CPU (tests/emul) and checked there.
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
  // y = 2x with x = 16807*s < 2^46: y's high word is x >> 31 and its low
  // word 2*(x mod 2^31), so the fold takes no 64-bit shift
  const uint64_t y = (uint64_t)(uint32_t)s * (2u * kIA);
  uint32_t r = (uint32_t)(y >> 32) + ((uint32_t)y >> 1);
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
// src/lattice.cpp:234-283) occupies [c*Mp, (c+1)*Mp).  Within a colour the
// sites form an (hx, hy, hz) sub-grid, x fastest, stored with a row pitch px
// (hx rounded up to a multiple of 4, so 4-site packs never straddle rows):
// local index a + px*(b + hy*e), Mp = px*hy*hz; the px-hx pad slots of a row
// hold no site and are never read as a neighbour.
//   sq/4n : hx=nx/2 hy=ny   hz=1    site (2a + ((b+c)&1), b)
//   sc/6n : hx=nx/2 hy=ny   hz=nz   site (2a + ((b+e+c)&1), b, e)
//   sq/8n : hx=nx/2 hy=ny/2 hz=1    c = 2(j&1)+(i&1),        site (2a+pi, 2b+pj)
//   sc/26n: hx=nx/2 hy=ny/2 hz=nz/2 c = 4(k&1)+2(j&1)+(i&1), site (2a+pi, 2b+pj, 2e+pk)
// Spins are stored as uint8 / uint16 / int32 (the narrowest type holding
// 1..Q); seeds as int32.
// Exact unsigned division by a fixed d with one multiply-high: m = 2^32/d
// (2^32-1 for d = 1) gives a quotient that is exact or one too small, and a
// single correction makes it exact for every 32-bit dividend.
struct FastDiv {
  uint32_t d, m;
};

__host__ __device__ inline FastDiv fastdiv_make(uint32_t d) {
  FastDiv f;
  f.d = d;
  f.m = d <= 1 ? 0xffffffffu : (uint32_t)(0x100000000ull / d);
  return f;
}

__host__ __device__ __forceinline__ uint32_t mulhi32(uint32_t a, uint32_t b) {
#ifdef __HIP_DEVICE_COMPILE__
  return __umulhi(a, b);
#else
  return (uint32_t)(((uint64_t)a * b) >> 32);
#endif
}

__host__ __device__ __forceinline__ void fastdiv(const FastDiv &f, uint32_t t, uint32_t &q, uint32_t &r) {
  q = mulhi32(t, f.m);
  r = t - q * f.d;
  if (r >= f.d) {
    q++;
    r -= f.d;
  }
}

struct Geom {
  int lat;
  int nx, ny, nz;
  int hx, hy, hz;
  int K;       // colours
  int M;       // sites per colour (hx*hy*hz)
  int px;      // row pitch: hx rounded up to a multiple of 4
  int Mp;      // stored slots per colour (px*hy*hz)
  FastDiv row4;  // division by px/4: 4-site packs per row
  FastDiv rows;  // division by hy
};

// Fills the colour sub-grid of a lattice style; edges must be even.
__host__ __device__ inline void geom_init(Geom &g, int lat, int nx, int ny, int nz) {
  g.lat = lat;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  switch (lat) {
    case SPK_HIP_SQ_4N: g.K = 2; g.hx = nx / 2; g.hy = ny; g.hz = 1; break;
    case SPK_HIP_SC_6N: g.K = 2; g.hx = nx / 2; g.hy = ny; g.hz = nz; break;
    case SPK_HIP_SQ_8N: g.K = 4; g.hx = nx / 2; g.hy = ny / 2; g.hz = 1; break;
    default: g.K = 8; g.hx = nx / 2; g.hy = ny / 2; g.hz = nz / 2; break;
  }
  g.M = g.hx * g.hy * g.hz;
  g.px = (g.hx + 3) & ~3;
  g.Mp = g.px * g.hy * g.hz;
  g.row4 = fastdiv_make((uint32_t)(g.px / 4));
  g.rows = fastdiv_make((uint32_t)g.hy);
}

// thread t of a colour pass -> its first site a0 and sub-grid row (b, e):
// px/4 threads per row, 4 consecutive sites each
__host__ __device__ __forceinline__ void pack_site(const Geom &g, uint32_t t, int &a0, int &b, int &e) {
  uint32_t row, a4, zz, yy;
  fastdiv(g.row4, t, row, a4);
  fastdiv(g.rows, row, zz, yy);
  a0 = (int)a4 * 4;
  b = (int)yy;
  e = (int)zz;
}

// Workgroup tiles of the per-colour kernel.  A workgroup of `block` threads
// covers a box of 2^lx threads (4 sites each) along x, 2^ly sub-grid rows
// along y and 2^lz planes along z, so the rows a site reads in y and z are
// mostly rows that other waves of the same workgroup own (L1 hits), rather
// than rows of other workgroups.  Tiles are numbered x-chunk fastest, then
// y-group, then z-group; threads of a tile likewise.  Slots past the edge of
// the sub-grid (partial tiles) are idle.
struct TileMap {
  int lx, ly, lz;      // log2 extents of a tile in threads / rows / planes
  int nxc, nyg, nzg;   // tiles along x, y, z
  FastDiv fxc, fyg;    // division by nxc, nyg
  int linear;          // 1: tile t is packs t*block .. t*block+block-1 in pack order
};

__host__ __device__ inline int ceil_log2(int v) {
  int l = 0;
  while ((1 << l) < v) l++;
  return l;
}

// block must be a power of two
__host__ __device__ inline TileMap make_tile_map(const Geom &g, int block) {
  TileMap m;
  const int lb = ceil_log2(block);
  m.lx = min(min(ceil_log2(g.px / 4), 6), lb);        // up to one wavefront along x
  int rest = lb - m.lx;
  const int cy = ceil_log2(g.hy), cz = ceil_log2(g.hz);
  m.lz = min(min(rest / 2, cz), rest);                  // the rows split evenly over y and z
  m.ly = min(rest - m.lz, cy);
  m.lz = min(rest - m.ly, cz);
  m.nxc = (g.px / 4 + (1 << m.lx) - 1) >> m.lx;
  m.nyg = (g.hy + (1 << m.ly) - 1) >> m.ly;
  m.nzg = (g.hz + (1 << m.lz) - 1) >> m.lz;
  m.fxc = fastdiv_make((uint32_t)m.nxc);
  m.fyg = fastdiv_make((uint32_t)m.nyg);
  // box tiles round the row to a power of two of threads and cut partial
  // tiles at the edges; where that idles more than a tenth of the launched
  // threads (short or odd rows), consecutive packs waste less
  const int64_t packs = (int64_t)g.Mp / 4;
  const int64_t launched = (int64_t)m.nxc * m.nyg * m.nzg * block;
  m.linear = launched * 9 > packs * 10;
  if (m.linear) {
    m.lx = lb;   // tile_site: tile t starts at pack t << lb
    m.ly = m.lz = 0;
    m.nxc = (int)((packs + block - 1) / block);
    m.nyg = m.nzg = 1;
  }
  return m;
}

// tile `tile`, thread `tid` -> first site a0 and row (b, e); false for a
// slot outside the sub-grid (or a thread beyond the tile's extent)
__host__ __device__ __forceinline__ bool tile_site(const Geom &g, const TileMap &m, uint32_t tile, uint32_t tid,
                                                   int &a0, int &b, int &e) {
  if (m.linear) {
    const uint32_t t = (tile << m.lx) + tid;
    if (t >= (uint32_t)(g.Mp / 4)) return false;
    pack_site(g, t, a0, b, e);
    return true;
  }
  uint32_t rest, xc, zg, yg;
  fastdiv(m.fxc, tile, rest, xc);
  fastdiv(m.fyg, rest, zg, yg);
  const uint32_t tx = tid & ((1u << m.lx) - 1), ty = (tid >> m.lx) & ((1u << m.ly) - 1),
                 tz = tid >> (m.lx + m.ly);
  a0 = (int)(((xc << m.lx) + tx) * 4);
  b = (int)((yg << m.ly) + ty);
  e = (int)((zg << m.lz) + tz);
  return tz < (1u << m.lz) && a0 < g.px && b < g.hy && e < g.hz;
}

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
      local = (i >> 1) + (int64_t)g.px * j;
      break;
    case SPK_HIP_SC_6N:
      c = (i + j + k) & 1;
      local = (i >> 1) + (int64_t)g.px * (j + (int64_t)g.ny * k);
      break;
    case SPK_HIP_SQ_8N:
      c = 2 * (j & 1) + (i & 1);
      local = (i >> 1) + (int64_t)g.px * (j >> 1);
      break;
    default:
      c = 4 * (k & 1) + 2 * (j & 1) + (i & 1);
      local = (i >> 1) + (int64_t)g.px * ((j >> 1) + (int64_t)g.hy * (k >> 1));
      break;
  }
  return (int64_t)c * g.Mp + local;
}

__host__ __device__ __forceinline__ int wrap_dec(int a, int n) { return a == 0 ? n - 1 : a - 1; }
__host__ __device__ __forceinline__ int wrap_inc(int a, int n) { return a == n - 1 ? 0 : a + 1; }

template <int LAT> struct NeighCount { static constexpr int value = 26; };
template <> struct NeighCount<SPK_HIP_SQ_4N> { static constexpr int value = 4; };
template <> struct NeighCount<SPK_HIP_SC_6N> { static constexpr int value = 6; };
template <> struct NeighCount<SPK_HIP_SQ_8N> { static constexpr int value = 8; };

// Gather the neighbour spins of sub-grid site (a,b,e) of colour c into nb[].
// The neighbour multiset equals the reference's periodic Moore / von Neumann
// shell (src/create_sites.cpp:589-640, :1303-1371), duplicates included on
// edges of length 2.  Used by the energy reduction.
template <int LAT, typename T>
__host__ __device__ __forceinline__ void gather(const T *__restrict__ spin, const Geom &g,
                                                int c, int a, int b, int e, int32_t *nb) {
  if constexpr (LAT == SPK_HIP_SQ_4N) {
    const T *o = spin + (int64_t)(1 - c) * g.Mp;
    int p = (b + c) & 1;
    int xo = p ? wrap_inc(a, g.hx) : wrap_dec(a, g.hx);
    int64_t row = (int64_t)g.px * b;
    nb[0] = o[a + row];
    nb[1] = o[xo + row];
    nb[2] = o[a + (int64_t)g.px * wrap_dec(b, g.hy)];
    nb[3] = o[a + (int64_t)g.px * wrap_inc(b, g.hy)];
  } else if constexpr (LAT == SPK_HIP_SC_6N) {
    const T *o = spin + (int64_t)(1 - c) * g.Mp;
    int p = (b + e + c) & 1;
    int xo = p ? wrap_inc(a, g.hx) : wrap_dec(a, g.hx);
    int64_t hxy = (int64_t)g.px * g.hy;
    int64_t row = (int64_t)g.px * b + hxy * e;
    nb[0] = o[a + row];
    nb[1] = o[xo + row];
    nb[2] = o[a + (int64_t)g.px * wrap_dec(b, g.hy) + hxy * e];
    nb[3] = o[a + (int64_t)g.px * wrap_inc(b, g.hy) + hxy * e];
    nb[4] = o[a + (int64_t)g.px * b + hxy * wrap_dec(e, g.hz)];
    nb[5] = o[a + (int64_t)g.px * b + hxy * wrap_inc(e, g.hz)];
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
        nb[n++] = spin[(int64_t)q * g.Mp + xs[dx] + (int64_t)g.px * ys[dy]];
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
          nb[n++] = spin[(int64_t)q * g.Mp + xs[dx] +
                         (int64_t)g.px * (ys[dy] + (int64_t)g.hy * zs[dz])];
        }
  }
}

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

// Proposed spin from the first draw s (already advanced):
// Potts `(int)(nspins*uniform())+1` clamped to Q (src/app_potts.cpp:119-120),
// Ising `uniform() < 0.5 ? 1 : 2` (src/app_ising.cpp:112-113).
__host__ __device__ __forceinline__ int32_t propose(const Rule &r, int32_t s) {
  if (r.app == SPK_HIP_POTTS) {
    double u = kAM * (double)s;
    int32_t p = (int32_t)((double)r.nspins * u) + 1;
    return p > r.nspins ? r.nspins : p;
  }
  return s < r.half ? 1 : 2;
}

// propose() for a kernel that knows the state count at compile time.  With
// two states Potts' (int)(2u)+1 is 1 exactly when u < 0.5 (2u is exact), so
// both apps reduce to Ising's test against the integer threshold `half`;
// otherwise the app is Potts.
template <bool TWO>
__host__ __device__ __forceinline__ int32_t propose_as(const Rule &r, int32_t s) {
  if constexpr (TWO) {
    return s < r.half ? 1 : 2;
  } else {
    double u = kAM * (double)s;
    int32_t p = (int32_t)((double)r.nspins * u) + 1;
    return p > r.nspins ? r.nspins : p;
  }
}

// The accept/reject tail of site_event_rejection (src/app_potts.cpp:125-133):
// ef <= ei accepts; T == 0 rejects an uphill move without a second draw;
// otherwise a second draw s2 rejects when AM*s2 > exp(-(ef-ei)/kT), i.e.
// when s2 >= thr[ef-ei].  Advances s in place; returns accept.
__host__ __device__ __forceinline__ bool metropolis(const Rule &r, const int *thr, int ei, int ef,
                                                    int32_t &s) {
  if (ef <= ei) return true;
  if (r.zero_t) return false;
  s = pm_step(s);
  return s < thr[ef - ei];
}

// metropolis() without branches, for the packed kernels: the second draw is
// always taken and kept only for an uphill move at T > 0 (a wavefront nearly
// always holds an uphill lane, so a branch around the draw saves nothing and
// costs scalar instructions per site).  Same result and seed as metropolis().
__host__ __device__ __forceinline__ bool metropolis_select(const Rule &r, const int *thr, int ei, int ef,
                                                           int32_t &s) {
  const int d = ef - ei;
  const bool up = d > 0;
  const bool draw = up && !r.zero_t;
  const int32_t s2 = pm_step(s);
  const bool pass = s2 < thr[up ? d : 0];
  s = draw ? s2 : s;
  return !up || (draw && pass);
}

struct SiteOutcome {
  int32_t spin;     // new value (== old when rejected)
  int32_t seed;     // advanced generator state
  int changed;      // 1 when the spin changed (naccept++)
  int dE;           // ef - ei for a change, else 0
};

// One strict site event on gathered neighbours
// (AppPotts/AppIsing::site_event_rejection, src/app_potts.cpp:112-146,
// src/app_ising.cpp:104-136).  The mask bookkeeping (:140-145) is omitted:
// in color/strict mode the mask never skips a site.
template <int NN>
__host__ __device__ __forceinline__ SiteOutcome site_event(const Rule &r, const int *thr,
                                                           int32_t old, int32_t seed,
                                                           const int32_t *nb) {
  int32_t s = pm_step(seed);
  const int32_t prop = propose(r, s);
  int ei = 0, ef = 0;
#pragma unroll
  for (int m = 0; m < NN; m++) {
    ei += (nb[m] != old);
    ef += (nb[m] != prop);
  }
  const bool accept = metropolis(r, thr, ei, ef, s);
  SiteOutcome o;
  o.changed = accept && prop != old;
  o.spin = o.changed ? prop : old;
  o.seed = s;
  o.dE = o.changed ? ef - ei : 0;
  return o;
}

struct Counts {
  int nacc;   // accepted flips
  int de;     // sum of (ef - ei) over accepted flips
};

// ------------------------------------------- proposal variants (ordered)
// potts/neigh, potts/neighonly and ising/single differ from potts / ising
// only in the proposed move; energy and the Metropolis test are shared.
// The first two pick among the distinct unlike neighbour values in order of
// first appearance along the reference's neighbour list, so these rules see
// the neighbours in that order: offsets (dx, dy, dz) with dx outermost, each
// -1..1, kept when within the lattice's shell (CreateSites::offsets_2d /
// offsets_3d, src/create_sites.cpp:1303-1371); the list is built in that
// order (:562-640).  One site per thread; every neighbour is located from
// the site's natural coordinates through blocked_index.

template <int LAT>
struct Shell {
  static constexpr bool is3 = LAT == SPK_HIP_SC_6N || LAT == SPK_HIP_SC_26N;
  static constexpr int far = (LAT == SPK_HIP_SQ_4N || LAT == SPK_HIP_SC_6N) ? 1 : is3 ? 3 : 2;   // max dx^2+dy^2+dz^2
  // offset m of the reference's order, m < NeighCount<LAT>::value
  __host__ __device__ static constexpr void offset(int m, int &dx, int &dy, int &dz) {
    int n = 0;
    for (int i = -1; i <= 1; i++)
      for (int j = -1; j <= 1; j++)
        for (int k = is3 ? -1 : 0; k <= (is3 ? 1 : 0); k++) {
          const int r2 = i * i + j * j + k * k;
          if (r2 < 1 || r2 > far) continue;
          if (n++ == m) {
            dx = i;
            dy = j;
            dz = k;
            return;
          }
        }
  }
};

// natural (x, y, z) of sub-grid site (a, b, e) of colour c (see Geom)
template <int LAT>
__host__ __device__ __forceinline__ void natural_of(int c, int a, int b, int e, int &x, int &y, int &z) {
  if constexpr (LAT == SPK_HIP_SQ_4N) {
    x = 2 * a + ((b + c) & 1), y = b, z = 0;
  } else if constexpr (LAT == SPK_HIP_SC_6N) {
    x = 2 * a + ((b + e + c) & 1), y = b, z = e;
  } else if constexpr (LAT == SPK_HIP_SQ_8N) {
    x = 2 * a + (c & 1), y = 2 * b + (c >> 1), z = 0;
  } else {
    x = 2 * a + (c & 1), y = 2 * b + ((c >> 1) & 1), z = 2 * e + (c >> 2);
  }
}

// neighbour values of natural site (x, y, z) in the reference's order.
// The colour-blocked slot (blocked_index) is a sum of one part per axis, so
// the three periodic coordinates v-1, v, v+1 of each axis are turned into
// their parts once and a neighbour's slot is two additions:
//   sq/8n, sc/26n: the colour is a sum of per-axis parity bits, so its
//     Mp-multiple rides in the axis parts;
//   sq/4n, sc/6n: every nearest neighbour has the other colour (edges are
//     even), so one colour base serves all of them.
template <int LAT, typename T>
__host__ __device__ __forceinline__ void gather_ordered(const T *__restrict__ spin, const Geom &g, int x, int y,
                                                        int z, int32_t (&nb)[NeighCount<LAT>::value]) {
  constexpr bool is3 = LAT == SPK_HIP_SC_6N || LAT == SPK_HIP_SC_26N;
  constexpr bool checker = LAT == SPK_HIP_SQ_4N || LAT == SPK_HIP_SC_6N;
  const int xs[3] = {wrap_dec(x, g.nx), x, wrap_inc(x, g.nx)};
  const int ys[3] = {wrap_dec(y, g.ny), y, wrap_inc(y, g.ny)};
  const int zs[3] = {is3 ? wrap_dec(z, g.nz) : 0, z, is3 ? wrap_inc(z, g.nz) : 0};
  const int64_t Mp = g.Mp, px = g.px, pxy = (int64_t)g.px * g.hy;
  int64_t lx[3], ly[3], lz[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    if constexpr (checker) {
      lx[d] = xs[d] >> 1;
      ly[d] = px * ys[d];
      lz[d] = is3 ? pxy * zs[d] : 0;
    } else {
      lx[d] = (xs[d] >> 1) + (xs[d] & 1) * Mp;
      ly[d] = px * (ys[d] >> 1) + 2 * (ys[d] & 1) * Mp;
      lz[d] = is3 ? pxy * (zs[d] >> 1) + 4 * (zs[d] & 1) * Mp : 0;
    }
  }
  // checkerboards: colour (x + y + z) & 1 of the site, flipped
  const T *base = checker ? spin + (int64_t)(((x + y + z) & 1) ^ 1) * Mp : spin;
#pragma unroll
  for (int m = 0; m < NeighCount<LAT>::value; m++) {
    int dx = 0, dy = 0, dz = 0;
    Shell<LAT>::offset(m, dx, dy, dz);
    nb[m] = (int32_t)base[lx[dx + 1] + ly[dy + 1] + lz[dz + 1]];
  }
}

// One strict event of a proposal variant, on ordered neighbour values:
//   potts/neigh     AppPottsNeigh::site_event_rejection
//                   (src/app_potts_neigh.cpp:71-116): draw
//                   iran = (int)(maxneigh*u); no event if iran >= nevent
//   potts/neighonly AppPottsNeighOnly::site_event_rejection
//                   (src/app_potts_neighonly.cpp:71-118): no draw when no
//                   neighbour is unlike; iran = (int)(nevent*u), clamped
//   ising/single    AppIsingSingle::site_event_rejection
//                   (src/app_ising_single.cpp:39-77): propose the flip
// then the shared Metropolis tail (metropolis).  unique[iran] is found
// without a list: neighbour m is a new distinct value when it is unlike the
// old spin and unlike every earlier neighbour.
template <int APP, int NN>
__host__ __device__ __forceinline__ SiteOutcome variant_event(const Rule &r, const int *thr, int32_t old,
                                                              int32_t seed, const int32_t (&nb)[NN]) {
  int32_t s = seed;
  int32_t prop = old;
  bool event = true;
  if constexpr (APP == SPK_HIP_ISING_SINGLE) {
    prop = old == 1 ? 2 : 1;
  } else {
    bool fresh[NN];
    int nevent = 0;
#pragma unroll
    for (int m = 0; m < NN; m++) {
      bool f = nb[m] != old;
#pragma unroll
      for (int i = 0; i < m; i++) f = f && nb[i] != nb[m];
      fresh[m] = f;
      nevent += f;
    }
    int iran;
    if constexpr (APP == SPK_HIP_POTTS_NEIGH) {
      s = pm_step(s);
      iran = (int)((double)NN * (kAM * (double)s));
      event = iran < nevent;
    } else {
      event = nevent > 0;
      if (event) {
        s = pm_step(s);
        iran = (int)((double)nevent * (kAM * (double)s));
        if (iran >= nevent) iran = nevent - 1;
      }
    }
    if (event) {
      int n = 0;
#pragma unroll
      for (int m = 0; m < NN; m++) {
        if (fresh[m] && n == iran) prop = nb[m];
        n += fresh[m];
      }
    }
  }
  SiteOutcome o{old, s, 0, 0};
  if (!event) return o;
  int ei = 0, ef = 0;
#pragma unroll
  for (int m = 0; m < NN; m++) {
    ei += nb[m] != old;
    ef += nb[m] != prop;
  }
  const bool accept = metropolis(r, thr, ei, ef, s);
  o.seed = s;
  o.changed = accept && prop != old;
  o.spin = o.changed ? prop : old;
  o.dE = o.changed ? ef - ei : 0;
  return o;
}

// Slot t of colour c (pad slots do nothing): one variant event.
template <int LAT, typename T, int APP>
__host__ __device__ __forceinline__ Counts variant_site(T *__restrict__ spin, int32_t *__restrict__ seed,
                                                        const Geom &g, const Rule &r, const int *thr, int c,
                                                        uint32_t t) {
  Counts out{0, 0};
  int a0, b, e;
  pack_site(g, t >> 2, a0, b, e);
  const int a = a0 + (int)(t & 3);
  if (a >= g.hx) return out;
  int x, y, z;
  natural_of<LAT>(c, a, b, e, x, y, z);
  int32_t nb[NeighCount<LAT>::value];
  gather_ordered<LAT, T>(spin, g, x, y, z, nb);
  const int64_t self = (int64_t)c * g.Mp + a + (int64_t)g.px * (b + (int64_t)g.hy * e);
  const SiteOutcome o = variant_event<APP, NeighCount<LAT>::value>(r, thr, (int32_t)spin[self], seed[self], nb);
  seed[self] = o.seed;
  if (o.changed) spin[self] = (T)o.spin;
  out.nacc = o.changed;
  out.de = o.dE;
  return out;
}

// ------------------------------------------------ vectorised colour pass
// Each thread owns 4 consecutive sites of one sub-grid row (px % 4 == 0, so
// a thread's sites never straddle a row).  Their spins, and each of their
// neighbour values, are handled as 4 lanes packed into 32-bit words: one
// word of bytes (u8 storage), two words of halfwords (u16) or four words
// (i32).  Neighbours in y and z sit at the same lane positions of another
// row; the two x-neighbours of a lane are that lane and the one beside it
// in the other x-parity's row, so they come from a word-wide funnel shift.
// The last thread of a row may own pad lanes (a0+v >= hx): they are
// computed on pad data and discarded, and their junk cannot reach a real
// lane (see Tally).

// True when every element offset of the lattice's spin and seed arrays
// (K*Mp elements, at most 4 bytes each) fits a 32-bit byte offset, and the
// row pitch and row count are below 2^24 (24-bit row products, RowOffset).
__host__ __device__ inline bool offsets_fit_32(const Geom &g) {
  return (uint64_t)g.K * (uint64_t)g.Mp * 4u < (1ull << 32) && g.px < (1 << 24) && g.hy < (1 << 24);
}

// 32-bit or 64-bit element offsets.  With 32-bit offsets the byte offset is
// formed in 32 bits, so loads take the (scalar base + 32-bit offset) form;
// valid while the array holds fewer than 2^32 bytes.
template <typename I, typename T>
__host__ __device__ __forceinline__ T *elem(T *base, I idx) {
  if constexpr (sizeof(I) == 4)
    return reinterpret_cast<T *>(reinterpret_cast<char *>(base) + (uint32_t)(idx * (uint32_t)sizeof(T)));
  else
    return base + idx;
}
template <typename I, typename T>
__host__ __device__ __forceinline__ const T *elem(const T *base, I idx) {
  if constexpr (sizeof(I) == 4)
    return reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) +
                                       (uint32_t)(idx * (uint32_t)sizeof(T)));
  else
    return base + idx;
}

template <int NW>
struct alignas(4 * NW) Words {
  uint32_t w[NW];
};

// 4 lanes of storage type T: sizeof(T) words, lane v in bits
// [8*sizeof(T)*(v % (4/sizeof(T))), ...) of word v / (4/sizeof(T)).
template <typename T>
struct Lanes {
  static constexpr int NW = (int)sizeof(T);
  static constexpr int PER = 4 / NW;           // lanes per word
  static constexpr int BITS = 8 * NW;          // bits per lane
  static constexpr uint32_t MASK = NW == 4 ? 0xffffffffu : (1u << BITS) - 1;
  uint32_t w[NW];

  __host__ __device__ __forceinline__ int32_t get(int v) const {
    if constexpr (NW == 4) return (int32_t)w[v];
    else return (int32_t)((w[v / PER] >> (BITS * (v % PER))) & MASK);
  }
  __host__ __device__ __forceinline__ void set(int v, uint32_t x) {
    if constexpr (NW == 4) w[v] = x;
    else {
      const int sh = BITS * (v % PER);
      w[v / PER] = (w[v / PER] & ~(MASK << sh)) | ((x & MASK) << sh);
    }
  }
  __host__ __device__ __forceinline__ static Lanes pack(const int32_t (&x)[4]) {
    Lanes l;
    if constexpr (NW == 1) l.w[0] = (uint32_t)x[0] | (uint32_t)x[1] << 8 | (uint32_t)x[2] << 16 | (uint32_t)x[3] << 24;
    else if constexpr (NW == 2) {
      l.w[0] = (uint32_t)x[0] | (uint32_t)x[1] << 16;
      l.w[1] = (uint32_t)x[2] | (uint32_t)x[3] << 16;
    } else
      for (int v = 0; v < 4; v++) l.w[v] = (uint32_t)x[v];
    return l;
  }
};

template <typename T, typename I>
__host__ __device__ __forceinline__ Lanes<T> load_lanes(const T *base, I idx) {
  const Words<Lanes<T>::NW> x = *reinterpret_cast<const Words<Lanes<T>::NW> *>(elem<I>(base, idx));
  Lanes<T> l;
#pragma unroll
  for (int i = 0; i < Lanes<T>::NW; i++) l.w[i] = x.w[i];
  return l;
}

template <typename T, typename I>
__host__ __device__ __forceinline__ void store_lanes(T *base, I idx, const Lanes<T> &l) {
  Words<Lanes<T>::NW> x;
#pragma unroll
  for (int i = 0; i < Lanes<T>::NW; i++) x.w[i] = l.w[i];
  *reinterpret_cast<Words<Lanes<T>::NW> *>(elem<I>(base, idx)) = x;
}

// Byte i of the result is byte sel.byte[i] (0..7) of the 8 bytes hi:lo,
// bytes 0-3 being lo's: the v_perm_b32 instruction.
__host__ __device__ __forceinline__ uint32_t byte_perm(uint32_t hi, uint32_t lo, uint32_t sel) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_amdgcn_perm(hi, lo, sel);
#else
  const uint64_t v = (uint64_t)hi << 32 | lo;
  uint32_t r = 0;
  for (int i = 0; i < 4; i++) r |= (uint32_t)(v >> (8 * ((sel >> (8 * i)) & 7)) & 0xffu) << (8 * i);
  return r;
#endif
}

// Lanes of positions p-1 (off 0) or p+1 (off 1), p = a0+v, from the lanes P
// of positions a0..a0+3 and the element E one beyond them on that side (the
// periodic image: a0-1 or hx-1; a0+4 or 0), held in E's low bits (higher
// bits are ignored).  Narrow lanes take byte permutes whose selector is the
// only thing that depends on the side, so a side that is the same for the
// whole launch costs one permute per word.
template <typename T>
__host__ __device__ __forceinline__ Lanes<T> shifted(const Lanes<T> &P, uint32_t E, int off) {
  Lanes<T> S;
  if constexpr (Lanes<T>::NW == 1) {
    // off 1: P bytes 1, 2, 3 then E's byte 0; off 0: E's byte 0 then P bytes 0, 1, 2
    S.w[0] = byte_perm(E, P.w[0], off ? 0x04030201u : 0x02010004u);
  } else if constexpr (Lanes<T>::NW == 2) {
    // off 1: (P0 hi, P1 lo) and (P1 hi, E); off 0: (E, P0 lo) and (P0 hi, P1 lo)
    S.w[0] = byte_perm(off ? P.w[1] : E, P.w[0], off ? 0x05040302u : 0x01000504u);
    S.w[1] = byte_perm(off ? E : P.w[0], P.w[1], off ? 0x05040302u : 0x01000706u);
  } else {
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int src = off ? v + 1 : v - 1;
      S.w[v] = (src < 0 || src > 3) ? E : P.w[src];
    }
  }
  return S;
}

// The thread whose row ends inside its 4 lanes (k = hx-a0 < 4 real lanes;
// only when hx % 4 != 0): for off 1, lane k-1 must read position hx, i.e.
// the wrap to position 0, which E holds.  keep[i] marks the lanes of word i
// below k-1; the others take E (lanes from k on are pad lanes anyway).
template <typename T>
struct RowEnd {
  uint32_t keep[Lanes<T>::NW];
  __host__ __device__ __forceinline__ explicit RowEnd(int k) {
#pragma unroll
    for (int i = 0; i < Lanes<T>::NW; i++) {
      const int n = min(max(k - 1 - i * Lanes<T>::PER, 0), Lanes<T>::PER);
      keep[i] = n == Lanes<T>::PER ? 0xffffffffu : (1u << (n * Lanes<T>::BITS)) - 1u;
    }
  }
  __host__ __device__ __forceinline__ void apply(Lanes<T> &S, uint32_t E) const {
    E &= Lanes<T>::MASK;
    uint32_t Ew = E;
    if constexpr (Lanes<T>::NW == 1) Ew = E * 0x01010101u;
    else if constexpr (Lanes<T>::NW == 2) Ew = E | (E << 16);
#pragma unroll
    for (int i = 0; i < Lanes<T>::NW; i++) S.w[i] = (S.w[i] & keep[i]) | (Ew & ~keep[i]);
  }
};

// Keeps the compiler's instruction scheduler from moving code across this
// point (no instruction is emitted).
__host__ __device__ __forceinline__ void fence_schedule() {
#ifdef __HIP_DEVICE_COMPILE__
  __builtin_amdgcn_sched_barrier(0);
#endif
}

// acc += per-halfword "differs" flags of d (1 where the halfword of d is
// nonzero), and of d and e for the two-word form.  op_sel_hi:[1,0] makes
// the high lane read the constant's low half (1) as well; by default it
// would read its high half (0).  The sum is formed in the
// same asm statement, so the flags die at once instead of being gathered
// into one late sum tree that keeps every one of them live.
__host__ __device__ __forceinline__ void add_ne_halves(uint32_t &acc, uint32_t d) {
#ifdef __HIP_DEVICE_COMPILE__
  uint32_t t;
  asm("v_pk_min_u16 %1, %2, 1 op_sel_hi:[1,0]\n\tv_add_u32 %0, %0, %1" : "+v"(acc), "=&v"(t) : "v"(d));
#else
  acc += ((d & 0xffffu) ? 1u : 0u) | ((d >> 16) ? 0x10000u : 0u);
#endif
}

__host__ __device__ __forceinline__ void add_ne_halves2(uint32_t &acc, uint32_t d, uint32_t e) {
#ifdef __HIP_DEVICE_COMPILE__
  uint32_t t, u;
  asm("v_pk_min_u16 %1, %3, 1 op_sel_hi:[1,0]\n\tv_pk_min_u16 %2, %4, 1 op_sel_hi:[1,0]\n\t"
      "v_add3_u32 %0, %0, %1, %2"
      : "+v"(acc), "=&v"(t), "=&v"(u)
      : "v"(d), "v"(e));
#else
  add_ne_halves(acc, d);
  add_ne_halves(acc, e);
#endif
}

// Pins a running sum in a register at this point, so later terms are added
// one by one rather than gathered into a late sum tree.
__host__ __device__ __forceinline__ void keep_sum(uint32_t &acc) {
#ifdef __HIP_DEVICE_COMPILE__
  asm("" : "+v"(acc));
#endif
}

// acc += (d != 0) + (e != 0) for whole 32-bit lanes, in the same way
__host__ __device__ __forceinline__ void add_ne_words2(uint32_t &acc, uint32_t d, uint32_t e) {
#ifdef __HIP_DEVICE_COMPILE__
  uint32_t t, u;
  asm("v_min_u32 %1, %3, 1\n\tv_min_u32 %2, %4, 1\n\tv_add3_u32 %0, %0, %1, %2"
      : "+v"(acc), "=&v"(t), "=&v"(u)
      : "v"(d), "v"(e));
#else
  acc += (d != 0) + (e != 0);
#endif
}

__host__ __device__ __forceinline__ uint32_t ne_bytes(uint32_t d) {
  // bit 7 of each byte of ((d & 0x7f..) + 0x7f..) | d is set iff the byte is
  // nonzero; no carry crosses a byte (each byte sum is at most 0xfe)
  return ((((d & 0x7f7f7f7fu) + 0x7f7f7f7fu) | d) >> 7) & 0x01010101u;
}

// Per-lane neighbour counts: ei = #neighbours unlike the old spin, ef =
// #unlike the proposal.  Lane flags of 0/1 are summed by plain 32-bit adds:
// a lane count never exceeds 26, so no byte or halfword field carries.  TWO (two states, values 1 and 2):
// x ^ old is 0 or 3 per byte, so one word sum carries 3*ei per byte, and ef
// follows from ei.  Pad lanes may hold other values; their sums can carry
// only into higher lanes, which are pad lanes as well.
template <typename T, bool TWO>
struct Tally {
  static constexpr bool kXorSum = TWO && sizeof(T) == 1;
  static constexpr int NW = Lanes<T>::NW;
  uint32_t o[NW], p[NW];

  __host__ __device__ __forceinline__ Tally() {
#pragma unroll
    for (int i = 0; i < NW; i++) o[i] = p[i] = 0;
  }
  __host__ __device__ __forceinline__ void add(const Lanes<T> &x, const Lanes<T> &old, const Lanes<T> &prop) {
    if constexpr (kXorSum) {
      o[0] += x.w[0] ^ old.w[0];
      keep_sum(o[0]);
    } else if constexpr (NW == 1) {
      o[0] += ne_bytes(x.w[0] ^ old.w[0]);
      keep_sum(o[0]);
      if constexpr (!TWO) {
        p[0] += ne_bytes(x.w[0] ^ prop.w[0]);
        keep_sum(p[0]);
      }
    } else if constexpr (NW == 2) {
#pragma unroll
      for (int i = 0; i < 2; i++) {
        add_ne_halves(o[i], x.w[i] ^ old.w[i]);
        if constexpr (!TWO) add_ne_halves(p[i], x.w[i] ^ prop.w[i]);
      }
    } else {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        o[i] += x.w[i] != old.w[i];
        if constexpr (!TWO) p[i] += x.w[i] != prop.w[i];
      }
    }
  }
  // two neighbour words at once (the two x-neighbours of a segment)
  __host__ __device__ __forceinline__ void add2(const Lanes<T> &x, const Lanes<T> &y, const Lanes<T> &old,
                                                const Lanes<T> &prop) {
    if constexpr (NW == 2 && !kXorSum) {
#pragma unroll
      for (int i = 0; i < 2; i++) {
        add_ne_halves2(o[i], x.w[i] ^ old.w[i], y.w[i] ^ old.w[i]);
        if constexpr (!TWO) add_ne_halves2(p[i], x.w[i] ^ prop.w[i], y.w[i] ^ prop.w[i]);
      }
    } else if constexpr (NW == 4) {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        add_ne_words2(o[i], x.w[i] ^ old.w[i], y.w[i] ^ old.w[i]);
        if constexpr (!TWO) add_ne_words2(p[i], x.w[i] ^ prop.w[i], y.w[i] ^ prop.w[i]);
      }
    } else if constexpr (kXorSum) {
      o[0] += (x.w[0] ^ old.w[0]) + (y.w[0] ^ old.w[0]);
      keep_sum(o[0]);
    } else {
      o[0] += ne_bytes(x.w[0] ^ old.w[0]) + ne_bytes(y.w[0] ^ old.w[0]);
      keep_sum(o[0]);
      if constexpr (!TWO) {
        p[0] += ne_bytes(x.w[0] ^ prop.w[0]) + ne_bytes(y.w[0] ^ prop.w[0]);
        keep_sum(p[0]);
      }
    }
  }
  __host__ __device__ __forceinline__ int ei(int v) const {
    if constexpr (kXorSum) return (int)((((o[0] >> (8 * v)) & 0xffu) * 171u) >> 9);   // byte / 3
    else if constexpr (NW == 4) return (int)o[v];
    else return (int)((o[v / Lanes<T>::PER] >> (Lanes<T>::BITS * (v % Lanes<T>::PER))) & Lanes<T>::MASK);
  }
  __host__ __device__ __forceinline__ int ef(int v) const {
    if constexpr (NW == 4) return (int)p[v];
    else return (int)((p[v / Lanes<T>::PER] >> (Lanes<T>::BITS * (v % Lanes<T>::PER))) & Lanes<T>::MASK);
  }
};

// Strict events for sites a0..a0+3 of sub-grid row (b, e) of colour c, in
// two phases.  pack_count loads the pack's spins and seeds, takes the first
// draws and the proposals, and streams every neighbour word into the counts
// of unlike neighbours (old and proposed lanes) as it arrives.  pack_decide
// applies the Metropolis rule and stores the pack.  Within a colour pass the
// neighbour values are read-only and each pack touches only its own slots,
// so a thread may count several packs before deciding any of them.
// TWO: the lattice has two states (Ising, or Potts Q=2), so every neighbour
// equals either the old spin or the proposal and ef follows from ei.  I:
// element offset type (uint32_t while the seed array holds fewer than 2^32
// bytes).
template <typename T, bool TWO, typename I>
struct PackWork {
  I self;             // the pack's offset within its colour's sub-grid
  int k;              // real lanes (hx - a0; >= 4 except at a row's end)
  Lanes<T> old;
  int32_t s[4], pr[4];
  Tally<T, TWO> tally;
};

// The row pitch px times a row index.  For 32-bit offsets the product is a
// 24-bit multiply (v_mul_u32_u24): exact while both factors are below 2^24,
// which offsets_fit_32 requires of px and of the row count hy.
template <typename I>
struct RowOffset {
  I pitch;
  __host__ __device__ __forceinline__ explicit RowOffset(I p) : pitch(p) {}
  __host__ __device__ __forceinline__ I operator*(I idx) const {
#ifdef __HIP_DEVICE_COMPILE__
    if constexpr (sizeof(I) == 4) return (I)__umul24((unsigned)pitch, (unsigned)idx);
#endif
    return pitch * idx;
  }
};

// The neighbour words of a pack: NSEG row segments (a pack plus the element
// beyond it on the side `off` names) and NPACK plain packs.
template <int LAT>
struct StencilWords {
  static constexpr int NSEG = LAT == SPK_HIP_SQ_8N ? 3 : LAT == SPK_HIP_SC_26N ? 9 : 1;
  static constexpr int NPACK = LAT == SPK_HIP_SC_6N ? 4 : LAT == SPK_HIP_SC_26N ? 8 : 2;
};

template <typename T, int NSEG, int NPACK>
struct RawWords {
  Lanes<T> seg[NSEG];
  uint32_t extra[NSEG];
  int off[NSEG];
  Lanes<T> pk[NPACK];
};

// XDIR: the side (0, 1) every segment shifts towards, when the caller knows
// it for the whole launch (sq/8n and sc/26n: the colour's x parity); -1
// takes each segment's own.
template <int LAT, typename T, bool TWO, typename I, int XDIR = -1>
__host__ __device__ __forceinline__ PackWork<T, TWO, I> pack_count(const T *__restrict__ spin,
                                                                   const int32_t *__restrict__ seed,
                                                                   const Geom &g, const Rule &r, int c, int a0,
                                                                   int b, int e) {
  using SW = StencilWords<LAT>;
  PackWork<T, TWO, I> w;
  const I Mp = (I)g.Mp;
  // row offsets px*y + pxy*z; px*y as a 24-bit multiply (see RowOffset)
  const RowOffset<I> px((I)g.px);
  const I pxy = (I)g.px * (I)g.hy;
  // colour q's sub-grid: a base that is the same for the whole launch, so
  // loads of one row from several colours share one per-thread offset
  auto colour = [&](int q) { return elem<I>(spin, (I)q * Mp); };
  w.self = (I)a0 + px * (I)b + pxy * (I)e;
  w.k = g.hx - a0;

  // ---- load phase: the offsets of every word the pack needs first, then
  // every load, before any word is used: one memory round trip per pack
  RawWords<T, SW::NSEG, SW::NPACK> raw;
  const T *sbase[SW::NSEG], *pbase[SW::NPACK];
  I srow[SW::NSEG], prow[SW::NPACK];
  int ns = 0, np = 0;
  const int right = a0 + 4 >= g.hx ? 0 : a0 + 4, left = a0 == 0 ? g.hx - 1 : a0 - 1;
  auto seg = [&](const T *base, I row, int off) {
    sbase[ns] = base;
    srow[ns] = row;
    raw.off[ns] = off;
    ns++;
  };
  auto pack = [&](const T *base, I row) {
    pbase[np] = base;
    prow[np] = row;
    np++;
  };
  if constexpr (LAT == SPK_HIP_SQ_4N) {
    const T *o = colour(1 - c);
    seg(o, px * (I)b, (b + c) & 1);
    pack(o, px * (I)wrap_dec(b, g.hy));
    pack(o, px * (I)wrap_inc(b, g.hy));
  } else if constexpr (LAT == SPK_HIP_SC_6N) {
    const T *o = colour(1 - c);
    seg(o, px * (I)b + pxy * (I)e, (b + e + c) & 1);
    pack(o, px * (I)wrap_dec(b, g.hy) + pxy * (I)e);
    pack(o, px * (I)wrap_inc(b, g.hy) + pxy * (I)e);
    pack(o, px * (I)b + pxy * (I)wrap_dec(e, g.hz));
    pack(o, px * (I)b + pxy * (I)wrap_inc(e, g.hz));
  } else if constexpr (LAT == SPK_HIP_SQ_8N) {
    const int pi = c & 1, pj = c >> 1;
    const int ys[3] = {pj ? b : wrap_dec(b, g.hy), b, pj ? wrap_inc(b, g.hy) : b};
#pragma unroll
    for (int dy = 0; dy < 3; dy++) {
      const int qy = pj ^ (dy != 1);
      const I row = px * (I)ys[dy];
      if (dy != 1) pack(colour(pi + 2 * qy), row);
      seg(colour((1 - pi) + 2 * qy), row, pi);
    }
  } else {
    const int pi = c & 1, pj = (c >> 1) & 1, pk = c >> 2;
    const int ys[3] = {pj ? b : wrap_dec(b, g.hy), b, pj ? wrap_inc(b, g.hy) : b};
    const int zs[3] = {pk ? e : wrap_dec(e, g.hz), e, pk ? wrap_inc(e, g.hz) : e};
#pragma unroll
    for (int dz = 0; dz < 3; dz++)
#pragma unroll
      for (int dy = 0; dy < 3; dy++) {
        const int qyz = 2 * (pj ^ (dy != 1)) + 4 * (pk ^ (dz != 1));
        const I row = px * (I)ys[dy] + pxy * (I)zs[dz];
        if (dy != 1 || dz != 1) pack(colour(pi + qyz), row);
        seg(colour((1 - pi) + qyz), row, pi);
      }
  }
  w.old = load_lanes<T, I>(colour(c), w.self);
  const Words<4> sd = *reinterpret_cast<const Words<4> *>(elem<I>(elem<I>(seed, (I)c * Mp), w.self));
#pragma unroll
  for (int m = 0; m < SW::NSEG; m++) {
    raw.seg[m] = load_lanes<T, I>(sbase[m], srow[m] + (I)a0);
    const int off = XDIR >= 0 ? XDIR : raw.off[m];
    raw.extra[m] = (uint32_t)(*elem<I>(sbase[m], srow[m] + (I)(off ? right : left)));
  }
#pragma unroll
  for (int m = 0; m < SW::NPACK; m++) raw.pk[m] = load_lanes<T, I>(pbase[m], prow[m] + (I)a0);
  fence_schedule();   // no counting work may be scheduled between the loads

  // ---- count phase: first draws and proposals, then every word compared
  // with the old and the proposed lanes
#pragma unroll
  for (int v = 0; v < 4; v++) {
    w.s[v] = pm_step((int32_t)sd.w[v]);
    w.pr[v] = propose_as<TWO>(r, w.s[v]);
  }
  const Lanes<T> prop = Lanes<T>::pack(w.pr);
  const bool ragged = (g.hx & 3) != 0;   // same for the whole launch
  // segment m: its pack and the pack shifted one site towards side `off`
  auto segment = [&](int m, int off) {
    Lanes<T> S = shifted<T>(raw.seg[m], raw.extra[m], off);
    if (ragged) {
      const RowEnd<T> rowend(off ? w.k : 5);   // only the +x side wraps inside a pack
      rowend.apply(S, raw.extra[m]);
    }
    w.tally.add2(raw.seg[m], S, w.old, prop);
  };
#pragma unroll
  for (int m = 0; m < SW::NSEG; m++) segment(m, XDIR >= 0 ? XDIR : raw.off[m]);
  static_assert(SW::NPACK % 2 == 0, "plain packs are counted in pairs");
#pragma unroll
  for (int m = 0; m < SW::NPACK; m += 2) w.tally.add2(raw.pk[m], raw.pk[m + 1], w.old, prop);
  return w;
}

template <int LAT, typename T, bool TWO, typename I>
__host__ __device__ __forceinline__ Counts pack_decide(PackWork<T, TWO, I> &w, T *__restrict__ spin,
                                                      int32_t *__restrict__ seed, const Geom &g, const Rule &r,
                                                      const int *thr, int c) {
  const I Mp = (I)g.Mp;
  T *const spin_c = elem<I>(spin, (I)c * Mp);
  int32_t *const seed_c = elem<I>(seed, (I)c * Mp);
  const int k = w.k;
  const I self = w.self;
  const Lanes<T> &old = w.old;
  int32_t *s = w.s;
  const int32_t *pr = w.pr;
  const Tally<T, TWO> &tally = w.tally;
  constexpr int NN = NeighCount<LAT>::value;
  Counts out{0, 0};
  Lanes<T> nsp = old;
  Words<4> nsd;
  bool any = false;
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int o_v = old.get(v);
    const int ei = tally.ei(v);
    const int ef = TWO ? (pr[v] == o_v ? ei : NN - ei) : tally.ef(v);
    const bool accept = metropolis_select(r, thr, ei, ef, s[v]);
    const bool changed = v < k && accept && pr[v] != o_v;   // pad lanes count nothing
    if (changed) nsp.set(v, (uint32_t)pr[v]);
    nsd.w[v] = (uint32_t)s[v];
    any |= changed;
    out.nacc += changed;
    out.de += changed ? ef - ei : 0;
  }
  *reinterpret_cast<Words<4> *>(elem<I>(seed_c, self)) = nsd;
  if (any) store_lanes<T, I>(spin_c, self, nsp);
  return out;
}


template <int LAT, typename T, int V, bool TWO, typename I = int64_t>
__host__ __device__ __forceinline__ Counts sweep_sites(T *__restrict__ spin, int32_t *__restrict__ seed,
                                                       const Geom &g, const Rule &r, const int *thr,
                                                       int c, int a0, int b, int e) {
  static_assert(V == 4, "4 sites per thread");
  PackWork<T, TWO, I> w = pack_count<LAT, T, TWO, I>(spin, seed, g, r, c, a0, b, e);
  return pack_decide<LAT, T, TWO, I>(w, spin, seed, g, r, thr, c);
}

// --------------------------------------------------------- reductions
__device__ __forceinline__ int wave_sum(int v) { return __ockl_wfred_add_i32(v); }
// 64-bit wrapping sum (mod 2^64) for checksums; not on a hot path
__device__ __forceinline__ unsigned long long wave_sum64(unsigned long long v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Accumulator rows: `slots` slots of (x, y) uint64 sums, each slot on its
// own 128-byte line (kSlotStride words apart).  Device-scope atomics from
// many workgroups serialise per cache line, so slots that shared a line
// cost ~10% of a colour pass; one line each removes that.
constexpr int kSlotStride = 16;

// Adds (x, y) of every thread of the block into slot `slot` of the row at
// acc: a wavefront sum (DPP, 32-bit: per-wave totals stay far below 2^31),
// one LDS slot per wave, one atomic per block and value.  All threads of the
// block must call it; blockDim.x % 64 == 0.  Callers that reduce again put a
// barrier between consecutive calls (the LDS slots are reused).
__device__ __forceinline__ void block_accumulate(int x, int y, unsigned long long *acc, int slot) {
  __shared__ int red[2][16];
  x = wave_sum(x);
  y = wave_sum(y);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wid] = x;
    red[1][wid] = y;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int sx = 0, sy = 0;   // a block's totals stay far below 2^31 as well
    const int nw = blockDim.x >> 6;
    for (int w = 0; w < nw; w++) {
      sx += red[0][w];
      sy += red[1][w];
    }
    unsigned long long *a = acc + (size_t)slot * kSlotStride;
    if (sx) atomicAdd(&a[0], (unsigned long long)(long long)sx);
    if (sy) atomicAdd(&a[1], (unsigned long long)(long long)sy);
  }
}

// block_accumulate for small per-thread counts: 0 <= x, 64 * x < 2^16 and
// 64 * |y| < 2^15 (one pack of 4 sites: x <= 4, |y| <= 4 * 26).  A wave's
// two sums then fit one 32-bit word, x in the low half and y in the high
// half, so one wave reduction carries both.
__device__ __forceinline__ void block_accumulate_pack(int x, int y, unsigned long long *acc, int slot) {
  __shared__ int red[2][16];
  const int v = wave_sum((int)((uint32_t)x + ((uint32_t)y << 16)));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wid] = v & 0xffff;
    red[1][wid] = v >> 16;   // arithmetic: the high half is the signed sum of y
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int sx = 0, sy = 0;
    const int nw = blockDim.x >> 6;
    for (int w = 0; w < nw; w++) {
      sx += red[0][w];
      sy += red[1][w];
    }
    unsigned long long *a = acc + (size_t)slot * kSlotStride;
    if (sx) atomicAdd(&a[0], (unsigned long long)(long long)sx);
    if (sy) atomicAdd(&a[1], (unsigned long long)(long long)sy);
  }
}

// block_accumulate of two (x, y) pairs into slot `slot` of two rows (at
// acc0 and acc1), with one barrier.
__device__ __forceinline__ void block_accumulate2(int x0, int y0, int x1, int y1, unsigned long long *acc0,
                                                  unsigned long long *acc1, int slot) {
  __shared__ int red[4][16];
  const int v[4] = {wave_sum(x0), wave_sum(y0), wave_sum(x1), wave_sum(y1)};
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0)
    for (int k = 0; k < 4; k++) red[k][wid] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    int s[4] = {0, 0, 0, 0};   // a block's totals stay far below 2^31
    const int nw = blockDim.x >> 6;
    for (int w = 0; w < nw; w++)
      for (int k = 0; k < 4; k++) s[k] += red[k][w];
    unsigned long long *a0 = acc0 + (size_t)slot * kSlotStride, *a1 = acc1 + (size_t)slot * kSlotStride;
    if (s[0]) atomicAdd(&a0[0], (unsigned long long)(long long)s[0]);
    if (s[1]) atomicAdd(&a0[1], (unsigned long long)(long long)s[1]);
    if (s[2]) atomicAdd(&a1[0], (unsigned long long)(long long)s[2]);
    if (s[3]) atomicAdd(&a1[1], (unsigned long long)(long long)s[3]);
  }
}

// Makes LDS values one wavefront wrote visible to all of its lanes: LDS
// operations of a wavefront complete in order, so a wavefront-scope fence
// and barrier suffice (no workgroup barrier).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// block_accumulate for one wavefront alone: one wavefront reduction per
// value (32-bit: a wavefront's sums stay far below 2^31), then one atomic per
// value from lane 0.  Uses no LDS and waits for no other wavefront.
__device__ __forceinline__ void wave_accumulate(int x, int y, unsigned long long *acc, int slot) {
  x = wave_sum(x);
  y = wave_sum(y);
  if ((threadIdx.x & 63) == 0) {
    unsigned long long *a = acc + (size_t)slot * kSlotStride;
    if (x) atomicAdd(&a[0], (unsigned long long)(long long)x);
    if (y) atomicAdd(&a[1], (unsigned long long)(long long)y);
  }
}

// block_accumulate_pack for one wavefront alone (same bounds on x and y):
// one wavefront reduction, then one atomic per value from lane 0.  Uses no
// LDS and waits for no other wavefront.
__device__ __forceinline__ void wave_accumulate_pack(int x, int y, unsigned long long *acc, int slot) {
  const int v = wave_sum((int)((uint32_t)x + ((uint32_t)y << 16)));
  if ((threadIdx.x & 63) == 0) {
    const int sx = v & 0xffff, sy = v >> 16;   // arithmetic shift: the signed sum of y
    unsigned long long *a = acc + (size_t)slot * kSlotStride;
    if (sx) atomicAdd(&a[0], (unsigned long long)(long long)sx);
    if (sy) atomicAdd(&a[1], (unsigned long long)(long long)sy);
  }
}

constexpr int kAccSlots = 64;                       // slots per row of the single-lattice path
constexpr int kAccRow = kAccSlots * kSlotStride;     // uint64 words per row
constexpr int kNumXcd = 8;      // MI355X: workgroups are dealt round-robin to 8 XCDs

// Logical tile of hardware workgroup `bid` in a grid of `nb` workgroups
// (nb a multiple of 8): consecutive logical tiles land on the same XCD, so
// tiles that share halo rows share that XCD's L2.
__host__ __device__ __forceinline__ int xcd_tile(int bid, int nb) {
  return (bid % kNumXcd) * (nb / kNumXcd) + bid / kNumXcd;
}

}  // namespace spk
