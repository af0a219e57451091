// spk_plain.hip -- libspk_hip.so, single-lattice path: the per-colour strict
// sweep kernels, the fused energy / Naccept reductions, device-side initial
// state (set site range, tagreset seeds via Park-Miller jump-ahead), layout
// conversion, and the C ABI declared in include/spk_hip.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <algorithm>
#include <utility>
#include <vector>

#include "spk_bits.h"
#include "spk_device.h"
#include "spk_host.h"
#include "spk_kernels.h"

using namespace spk;

// ==================================================================== kernels

// One colour pass of a two-state sq/4n lattice on bit planes: thread t owns
// SPT sites of one plane word (see spk_bits.h).
template <int SPT, bool SINGLE>
__global__ __launch_bounds__(1024) void k_sweep_bits(uint32_t *__restrict__ bits, int32_t *__restrict__ sseed, Geom g,
                                                    BitGeom bg, Rule r, int color,
                                                    unsigned long long *__restrict__ acc) {
  const uint32_t t = (uint32_t)xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  Counts cnt{0, 0};
  if ((int64_t)t < bg.threads) cnt = bits_thread<SPT, SINGLE>(bits, sseed, g, bg, r, color, t);
  block_accumulate_pack(cnt.nacc, cnt.de, acc, blockIdx.x & (kAccSlots - 1));
}

// Two sweeps of a lattice that pair_fits, one band of kPairBand rows per
// workgroup (spk_bits.h): planes and seeds from (bits_in, seed_in) to
// (bits_out, seed_out); sweep 1's counts to the row at acc, sweep 2's to
// the next row.  The grid is the bands rounded up to a multiple of 8.
// Every thread issues the loads of its band's planes first, then the loads
// of its seed slots in the order it uses them, two slots ahead of their use:
// each is waited for just before that use while the later ones are in
// flight.  The seed loads are unconditional (a slot outside every pass, or
// an idle lane, loads a row that exists and is not used) so that the count
// of loads in flight is the same on every path and those waits stay exact.  FULL: rows of kPairMaxPw
// whole words (sub-grid rows of 2048 sites: nx = 4096) and an even number
// of them, so every lane works and the row arithmetic is simpler
// (pair_side, pair_thread).
template <bool SINGLE, bool FULL>
__global__ __launch_bounds__(kPairThreads) void k_sweep_bits_pair(const uint32_t *__restrict__ bits_in,
                                                                 const int32_t *__restrict__ seed_in,
                                                                 uint32_t *__restrict__ bits_out,
                                                                 int32_t *__restrict__ seed_out, Geom g, BitGeom bg,
                                                                 Rule r, unsigned long long *__restrict__ acc) {
  __shared__ PairLds L;
  const int y0 = xcd_tile(blockIdx.x, gridDim.x) * kPairBand;
  if (y0 >= g.hy) return;   // a padding workgroup: all its threads leave together
  // the group is the same on all lanes of a wavefront (kPairStage is a
  // multiple of 64); saying so keeps it and the rows in scalar registers
  const int grp = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kPairStage));
  const uint32_t tau = threadIdx.x % kPairStage;
  const bool lane_on = FULL || tau < 4u * (uint32_t)bg.pw;
  const uint32_t ta = lane_on ? tau : 0u;
  // the planes: at most 3 words per thread (2 * kPairRows * 64 <= 3 * 1024)
  constexpr int kPlaneLoads = (2 * kPairRows * kPairMaxPw + kPairThreads - 1) / kPairThreads;
  const int nplane = 2 * kPairRows * bg.pw;
  uint32_t pv[kPlaneLoads];
#pragma unroll
  for (int k = 0; k < kPlaneLoads; k++) {
    int c, lr, w;
    pv[k] = bits_in[pair_plane_src(g, bg, y0, min((int)threadIdx.x + k * kPairThreads, nplane - 1), c, lr, w)];
  }
  // (a thread past the last word writes that word again, with its value:
  // no condition, so the compiler keeps the loads where they were issued)
#pragma unroll
  for (int k = 0; k < kPlaneLoads; k++) {
    int c, lr, w;
    (void)pair_plane_src(g, bg, y0, min((int)threadIdx.x + k * kPairThreads, nplane - 1), c, lr, w);
    L.planes[c][lr][w] = pv[k];
  }
  // The seeds of the thread's rows, colour 0 (passes 0, 2) and colour 1
  // (passes 1, 3), go out a few at a time, two slots ahead of their use:
  // colour 0's first two slots now, the others during pass 0 (below).
  // Issuing all of them at the start had every workgroup's plane loads wait
  // behind every other's seed loads (7 us of a 48 us launch).
  Words<4> s0[kPairSlots0][2], s1[kPairSlots1][2];
  auto load = [&](int c, int j, Words<4>(&dst)[2]) {
    const int rho = min(pair_slot_row(y0, c, grp, j), y0 + kPairBand + 2 - c);   // within pass c's rows
    const int prow = pair_row(rho, g.hy);
#pragma unroll
    for (int h = 0; h < 2; h++)
      dst[h] = *pair_seeds<FULL>(seed_in, bg, c, prow, ta, h);
  };
  constexpr int kAhead = 2;
#pragma unroll
  for (int j = 0; j < kAhead; j++) load(0, j, s0[j]);
  __syncthreads();
  PairCounts a1, a2;   // sweep 1 (passes 0, 1) and sweep 2
  // full rows: the thread's LDS offsets are the same on every row; opaque,
  // so the compiler keeps them rather than recomputing them a row at a time
  PairLane lane_full{};
  if constexpr (FULL) {
    lane_full = pair_lane<true>(g, bg, 0, 0, grp, tau);
    asm volatile("" : "+v"(lane_full.w4), "+v"(lane_full.wn4), "+v"(lane_full.b));
  }
  // a pass over the thread's slots; before(j) runs before slot j
  auto pass = [&](int p, auto &seeds, int slots, auto &&before) {
#pragma unroll
    for (int j = 0; j < slots; j++) {
      before(j);
      const int rho = pair_slot_row(y0, p & 1, grp, j);
      if (lane_on && pair_pass_row(y0, p, rho)) {
        const PairLane ln = FULL ? lane_full : pair_lane<false>(g, bg, rho, p & 1, grp, tau);
        const PairWords pwd = pair_words(L, y0, p, rho, grp, ln);
        pair_thread<SINGLE, FULL>(L, g, bg, r, y0, p, rho, grp, tau, ln, pwd, seeds[j], seed_out, p < 2 ? a1 : a2);
      }
    }
    __syncthreads();
  };
  // colour 1's first kAhead slots go out during pass 0's last slots and the
  // rest during pass 1, two slots ahead of use, so that each pass reads
  // about one colour's seeds rather than pass 0 reading both
  static_assert(kPairSlots0 - kAhead >= kAhead, "colour 1's first slots go out after colour 0's");
  auto loads_ahead0 = [&](int j) {
    if (j + kAhead < kPairSlots0) load(0, j + kAhead, s0[j + kAhead]);
    if (j >= kPairSlots0 - kAhead) load(1, j - (kPairSlots0 - kAhead), s1[j - (kPairSlots0 - kAhead)]);
  };
  auto loads_ahead1 = [&](int j) {
    if (j + kAhead < kPairSlots1) load(1, j + kAhead, s1[j + kAhead]);
  };
  auto nothing = [](int) {};
  pass(0, s0, kPairSlots0, loads_ahead0);
  // colour 0's last slot that pass 2 uses waits in LDS through pass 1, when
  // both colours' seeds are held (its registers go to pass 1's work), and
  // comes back one slot before pass 2 needs it; the single-flip events need
  // fewer registers and keep it
  constexpr int kParked = kPairSlots0 - 2;
  if constexpr (!SINGLE) {
#pragma unroll
    for (int h = 0; h < 2; h++) L.parked[h][threadIdx.x] = s0[kParked][h];
  }
  pass(1, s1, kPairSlots1, loads_ahead1);
  // the last colour-0 slot's rows (y0+13+g+4 >= y0+kPairBand+1) lie past
  // pass 2's span: its seeds are not kept beyond pass 0
  static_assert(-3 + kPairGroups * (kPairSlots0 - 1) >= kPairBand + 1, "last colour-0 slot outside pass 2");
  auto unpark = [&](int j) {
    if (!SINGLE && j == kParked - 1) {
#pragma unroll
      for (int h = 0; h < 2; h++) s0[kParked][h] = L.parked[h][threadIdx.x];
    }
  };
  pass(2, s0, kPairSlots0 - 1, unpark);
  pass(3, s1, kPairSlots1, nothing);
  for (int i = threadIdx.x; i < 2 * kPairBand * bg.pw; i += kPairThreads) pair_store_plane(L, bits_out, g, bg, y0, i);
  const Counts c1 = pair_counts(a1), c2 = pair_counts(a2);
  block_accumulate2(c1.nacc, c1.de, c2.nacc, c2.de, acc, acc + kAccRow, blockIdx.x & (kAccSlots - 1));
}

// u8 state -> planes, one thread per plane word of both colours
__global__ void k_bits_pack(const uint8_t *__restrict__ spin, Geom g, BitGeom bg, uint32_t *__restrict__ bits) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t words = (int64_t)g.hy * bg.pw;
  if (q >= 2 * words) return;
  const int c = q >= words;
  const int64_t t = q - (c ? words : 0);
  const int b = (int)(t / bg.pw), w = (int)(t % bg.pw);
  bits[q] = bits_pack_word(spin, g, c, b, w);
}

__global__ void k_bits_unpack(const uint32_t *__restrict__ bits, Geom g, BitGeom bg, uint8_t *__restrict__ spin) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t words = (int64_t)g.hy * bg.pw;
  if (q >= 2 * words) return;
  const int c = q >= words;
  const int64_t t = q - (c ? words : 0);
  const int b = (int)(t / bg.pw), w = (int)(t % bg.pw);
  bits_unpack_word(spin, g, c, b, w, bits[q]);
}

// seeds <-> shadow seeds, one thread per 16-byte group (bits_seed_group_in
// / _out): each wavefront reads or writes 1 KiB of consecutive canonical
// bytes and 128-byte runs of the shadow
__global__ void k_seeds_to_shadow(const int32_t *__restrict__ seed, Geom g, BitGeom bg, int32_t *__restrict__ sseed) {
  const int64_t X = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (X < bg.seedc / 2) bits_seed_group_in(seed, g, bg, X, sseed);
}

__global__ void k_seeds_from_shadow(const int32_t *__restrict__ sseed, Geom g, BitGeom bg, int32_t *__restrict__ seed) {
  const int64_t X = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (X < bg.seedc / 2) bits_seed_group_out(sseed, g, bg, X, seed);
}

// One colour pass: every site of colour `color` performs one strict event.
// A pack is 4 consecutive sites of one sub-grid row (x fastest), so its own
// spins/seeds are one aligned load and a wavefront's accesses are
// unit-stride.  Each workgroup covers one TileMap tile (rows of its own
// waves supply most of the y/z neighbour rows); tiles are remapped so that
// neighbouring tiles share an XCD.
template <int LAT, typename T, int V, bool TWO, typename I, int XDIR>
__global__ __launch_bounds__(1024) void k_sweep_color(T *__restrict__ spin, int32_t *__restrict__ seed,
                                                      Geom g, Rule r, TileMap m, int color,
                                                      unsigned long long *__restrict__ acc) {
  static_assert(V == 4, "tiles map 4 sites per thread");
  __shared__ int s_thr[32];
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  int a0, b, e;
  const bool live = tile < m.nxc * m.nyg * m.nzg && tile_site(g, m, (uint32_t)tile, threadIdx.x, a0, b, e);
  // the thresholds are staged in LDS between counting and deciding, so the
  // pack's loads are not held behind a barrier
  const int thr_v = threadIdx.x < 28 ? r.thr[threadIdx.x] : 0;
  PackWork<T, TWO, I> w;
  if (live) w = pack_count<LAT, T, TWO, I, XDIR>(spin, seed, g, r, color, a0, b, e);
  if (threadIdx.x < 28) s_thr[threadIdx.x] = thr_v;
  __syncthreads();
  Counts cnt{0, 0};
  if (live) cnt = pack_decide<LAT, T, TWO, I>(w, spin, seed, g, r, s_thr, color);
  __syncthreads();   // block_accumulate reuses LDS after the decide step's reads
  block_accumulate_pack(cnt.nacc, cnt.de, acc, blockIdx.x & (kAccSlots - 1));
}

// Every sweep of a lattice small enough to live in one workgroup's LDS
// (the reference's CPU-sized cases): spins and seeds are read once, all
// colour passes run on the LDS copy with a barrier between passes, and the
// state is written back once.  The same per-pack events as k_sweep_color,
// on LDS addresses; sweep s's (Naccept, dE) go to slot 0 of acc row s.
template <int LAT, typename T, bool TWO>
__global__ __launch_bounds__(1024) void k_sweep_resident(T *__restrict__ spin, int32_t *__restrict__ seed, Geom g,
                                                         Rule r, int nsweeps, unsigned long long *__restrict__ acc) {
  extern __shared__ __align__(16) unsigned char lds[];
  __shared__ int s_thr[32];
  const uint32_t slots = (uint32_t)g.K * (uint32_t)g.Mp;
  const uint32_t spin_words = slots * (uint32_t)sizeof(T) / 4;   // Mp is a multiple of 4
  T *ls = reinterpret_cast<T *>(lds);
  int32_t *lk = reinterpret_cast<int32_t *>(lds + (spin_words * 4 + 15) / 16 * 16);
  for (uint32_t i = threadIdx.x; i < spin_words; i += blockDim.x)
    reinterpret_cast<uint32_t *>(ls)[i] = reinterpret_cast<const uint32_t *>(spin)[i];
  for (uint32_t i = threadIdx.x; i < slots; i += blockDim.x) lk[i] = seed[i];
  if (threadIdx.x < 28) s_thr[threadIdx.x] = r.thr[threadIdx.x];
  __syncthreads();
  const uint32_t npack = (uint32_t)g.Mp / 4;
  for (int sw = 0; sw < nsweeps; sw++) {
    int na = 0, de = 0;
    for (int c = 0; c < g.K; c++) {
      for (uint32_t p = threadIdx.x; p < npack; p += blockDim.x) {
        int a0, b, e;
        pack_site(g, p, a0, b, e);
        const Counts cnt = sweep_sites<LAT, T, 4, TWO, uint32_t>(ls, lk, g, r, s_thr, c, a0, b, e);
        na += cnt.nacc;
        de += cnt.de;
      }
      __syncthreads();   // the next colour reads this colour's new spins
    }
    block_accumulate(na, de, acc + (size_t)sw * kAccRow, 0);
    __syncthreads();
  }
  for (uint32_t i = threadIdx.x; i < spin_words; i += blockDim.x)
    reinterpret_cast<uint32_t *>(spin)[i] = reinterpret_cast<const uint32_t *>(ls)[i];
  for (uint32_t i = threadIdx.x; i < slots; i += blockDim.x) seed[i] = lk[i];
}

// One colour pass of a proposal variant (potts/neigh, potts/neighonly,
// ising/single): one slot per thread (4 consecutive slots per pack, so a
// wavefront's own-site accesses stay unit-stride), neighbours read in the
// reference's order.
template <int LAT, typename T, int APP>
__global__ __launch_bounds__(1024) void k_sweep_variant(T *__restrict__ spin, int32_t *__restrict__ seed, Geom g,
                                                        Rule r, int color, unsigned long long *__restrict__ acc) {
  __shared__ int s_thr[32];
  if (threadIdx.x < 28) s_thr[threadIdx.x] = r.thr[threadIdx.x];
  __syncthreads();
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int64_t t = (int64_t)tile * blockDim.x + threadIdx.x;
  Counts cnt{0, 0};
  if (t < (int64_t)g.Mp) cnt = variant_site<LAT, T, APP>(spin, seed, g, r, s_thr, color, (uint32_t)t);
  block_accumulate_pack(cnt.nacc, cnt.de, acc, blockIdx.x & (kAccSlots - 1));
}

// ===================================================================== host

constexpr int kBitSites = 32;   // default sites per thread on bit planes

struct spk_hip_lattice {
  spk_hip_desc d;
  Geom g;
  Rule rule;
  int64_t N = 0;
  int device = 0;
  int stype = kI32;                    // storage type of spins on the device
  int64_t slots = 0;                   // K*Mp stored values (sites + row pads)
  hipStream_t stream = nullptr;
  void *spin = nullptr;                // slots spins, colour-blocked, stype
  int32_t *seed = nullptr;             // slots seeds, colour-blocked
  // bit planes of a two-state sq/4n lattice (spk_bits.h).  Once a sweeps
  // call has packed them they hold the state (bits_active) until a call
  // that needs the byte layout unpacks them (bits_leave).
  // bitplanes: 0 planes off; 1 the default (two sweeps per launch where
  // pair_fits, else kBitSites sites per thread); 8, 16, 32 sites per thread
  // of the per-colour kernel
  int bitplanes = 1;
  bool bits_active = false;
  bool pairs = false;                  // planes in the spt-8 layout, sweeps two per launch
  BitGeom bg{};
  uint32_t *bits = nullptr;
  int32_t *sseed = nullptr;
  int64_t sseed_cap = 0;               // shadow seed slots allocated
  uint32_t *bits2 = nullptr;           // the second planes and seeds of two-sweep launches
  int32_t *sseed2 = nullptr;
  int64_t sseed2_cap = 0;
  // lattices that fit one workgroup's LDS run every sweep in one launch
  int resident = 1;
  size_t max_lds = 0;                  // LDS bytes a workgroup may use
  int32_t *stage = nullptr;            // N int32, natural order (transfers)
  unsigned long long *acc = nullptr;   // [cap][kAccRow]: per-sweep slots
  unsigned long long *accc = nullptr;  // [cap][2]: rows folded to (Naccept, dE)
  int acc_cap = 0;
  int block = 256;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
  int64_t last_nlaunch = 0;
  bool energy_valid = false;
  int64_t energy = 0;                  // total energy, exact integer
  std::vector<unsigned long long> hacc;
  // graph replay: a call of n sweeps replays n / chunk graphs of `chunk`
  // sweeps, then one graph of the remainder
  int chunk = 32;                      // sweeps per full graph (0: eager launches only)
  struct Piece {
    int sweeps;                        // sweeps the graph runs
    bool bits;                         // captured on the bit planes
    hipGraphExec_t exec;
  };
  std::vector<Piece> graphs;
  unsigned long long *pinned = nullptr;   // per-sweep rows of a call, host pinned
  int pinned_rows = 0;
};

namespace {

constexpr int kInitialAccRows = 1024;
constexpr int kMaxChunk = 1024;   // sweeps per captured graph
constexpr int kVec = 4;   // sites per thread in the sweep kernel

unsigned grid_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

int launch_variant(spk_hip_lattice *h, int color, unsigned long long *acc) {
  const int bs = h->block;
  unsigned grid = grid_for(h->g.Mp, bs);
  grid = (grid + kNumXcd - 1) / kNumXcd * kNumXcd;   // xcd_tile needs a multiple of 8
  const int app = h->d.app;
  with_lattice(h->g.lat, [&]<int LAT>() {
    with_type(h->stype, [&]<typename T>() {
      auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(grid), dim3(bs), 0, h->stream, static_cast<T *>(h->spin), h->seed, h->g,
                           h->rule, color, acc);
      };
      if (app == SPK_HIP_POTTS_NEIGH) go(k_sweep_variant<LAT, T, SPK_HIP_POTTS_NEIGH>);
      else if (app == SPK_HIP_POTTS_NEIGHONLY) go(k_sweep_variant<LAT, T, SPK_HIP_POTTS_NEIGHONLY>);
      else if constexpr (sizeof(T) == 1) go(k_sweep_variant<LAT, T, SPK_HIP_ISING_SINGLE>);   // two states: u8
    });
  });
  return check_hip(hipGetLastError(), "sweep kernel launch");
}

int launch_sweep(spk_hip_lattice *h, int color, unsigned long long *acc) {
  if (h->bits_active) {
    const int bs = h->block;
    unsigned grid = grid_for(h->bg.threads, bs);
    grid = (grid + kNumXcd - 1) / kNumXcd * kNumXcd;   // xcd_tile needs a multiple of 8
    auto go = [&](auto kern) {
      hipLaunchKernelGGL(kern, dim3(grid), dim3(bs), 0, h->stream, h->bits, h->sseed, h->g, h->bg, h->rule, color,
                         acc);
    };
    const bool single = h->d.app == SPK_HIP_ISING_SINGLE;
    if (h->bg.spt == 8) single ? go(k_sweep_bits<8, true>) : go(k_sweep_bits<8, false>);
    else if (h->bg.spt == 16) single ? go(k_sweep_bits<16, true>) : go(k_sweep_bits<16, false>);
    else single ? go(k_sweep_bits<32, true>) : go(k_sweep_bits<32, false>);
    return check_hip(hipGetLastError(), "bit-plane sweep launch");
  }
  if (variant_app(h->d.app)) return launch_variant(h, color, acc);
  const int bs = h->block;
  const TileMap m = make_tile_map(h->g, bs);
  unsigned grid = (unsigned)m.nxc * m.nyg * m.nzg;
  grid = (grid + kNumXcd - 1) / kNumXcd * kNumXcd;   // xcd_tile needs a multiple of 8
  const bool two = h->rule.nspins == 2;   // two states are always stored as u8
  const bool off32 = offsets_fit_32(h->g);
  with_lattice(h->g.lat, [&]<int LAT>() {
    with_type(h->stype, [&]<typename T>() {
      auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(grid), dim3(bs), 0, h->stream, static_cast<T *>(h->spin), h->seed, h->g,
                           h->rule, m, color, acc);
      };
      // sq/8n and sc/26n: the x-shift side is the colour's x parity, a
      // constant of the launch, so it is a template argument
      auto with_x = [&]<bool TWO, typename I>() {
        if constexpr (LAT == SPK_HIP_SQ_8N || LAT == SPK_HIP_SC_26N) {
          if (color & 1) go(k_sweep_color<LAT, T, kVec, TWO, I, 1>);
          else go(k_sweep_color<LAT, T, kVec, TWO, I, 0>);
        } else {
          go(k_sweep_color<LAT, T, kVec, TWO, I, -1>);
        }
      };
      if constexpr (sizeof(T) == 1) {
        if (two) {
          if (off32) with_x.template operator()<true, uint32_t>();
          else with_x.template operator()<true, int64_t>();
          return;
        }
      }
      if (off32) with_x.template operator()<false, uint32_t>();
      else with_x.template operator()<false, int64_t>();
    });
  });
  return check_hip(hipGetLastError(), "sweep kernel launch");
}

void drop_graphs(spk_hip_lattice *h) {
  for (auto &c : h->graphs) (void)hipGraphExecDestroy(c.exec);
  h->graphs.clear();
}

// Two-state sq/4n lattices on the per-colour path sweep on bit planes.
// (bits_events takes the proposal from bit 30 of the first draw, so half
// must be 2^30, which RandomPark's AM*s >= 0.5 gives)
bool bits_eligible(const spk_hip_lattice *h) {
  return h->bitplanes && h->g.lat == SPK_HIP_SQ_4N && h->rule.nspins == 2 && h->stype == kU8 &&
         h->rule.half == (1 << 30) &&
         (!variant_app(h->d.app) || h->d.app == SPK_HIP_ISING_SINGLE);
}

// Packs the state into the planes (allocated on first use) and routes the
// colour passes there until bits_leave.
// (Re)allocates a shadow seed array to hold `slots` slots.
int ensure_shadow(spk_hip_lattice *h, int32_t *&p, int64_t &cap, int64_t slots) {
  if (slots <= cap) return SPK_HIP_OK;
  (void)hipStreamSynchronize(h->stream);
  if (p) (void)hipFree(p);
  cap = 0;
  if (hipMalloc(&p, (size_t)slots * sizeof(int32_t)) != hipSuccess) {
    p = nullptr;
    return set_error(SPK_HIP_ENOMEM, "bit-plane shadow seeds not available");
  }
  cap = slots;
  return SPK_HIP_OK;
}

int ensure_planes(uint32_t *&p, const spk_hip_lattice *h) {
  if (p) return SPK_HIP_OK;   // 2*hy*pw words whatever the sites per thread
  if (hipMalloc(&p, 2 * (size_t)h->g.hy * h->bg.pw * sizeof(uint32_t)) != hipSuccess) {
    p = nullptr;
    return set_error(SPK_HIP_ENOMEM, "bit-plane shadow state not available");
  }
  return SPK_HIP_OK;
}

int bits_enter(spk_hip_lattice *h) {
  h->pairs = h->bitplanes == 1 && pair_fits(h->g);
  h->bg = bit_geom(h->g, h->pairs ? 8 : h->bitplanes == 1 ? kBitSites : h->bitplanes);
  int rc;
  if ((rc = ensure_planes(h->bits, h)) || (rc = ensure_shadow(h, h->sseed, h->sseed_cap, 2 * h->bg.seedc)))
    return rc;
  if (h->pairs &&
      ((rc = ensure_planes(h->bits2, h)) || (rc = ensure_shadow(h, h->sseed2, h->sseed2_cap, 2 * h->bg.seedc))))
    return rc;
  const int64_t words = 2 * (int64_t)h->g.hy * h->bg.pw;
  hipLaunchKernelGGL(k_bits_pack, dim3(grid_for(words, 256)), dim3(256), 0, h->stream,
                     static_cast<const uint8_t *>(h->spin), h->g, h->bg, h->bits);
  hipLaunchKernelGGL(k_seeds_to_shadow, dim3(grid_for(h->bg.seedc / 2, 256)), dim3(256), 0, h->stream, h->seed,
                     h->g, h->bg, h->sseed);
  rc = check_hip(hipGetLastError(), "bit-plane pack");
  if (!rc) h->bits_active = true;
  return rc;
}

int bits_leave(spk_hip_lattice *h) {
  if (!h->bits_active) return SPK_HIP_OK;
  h->bits_active = false;
  const int64_t words = 2 * (int64_t)h->g.hy * h->bg.pw;
  hipLaunchKernelGGL(k_bits_unpack, dim3(grid_for(words, 256)), dim3(256), 0, h->stream, h->bits, h->g, h->bg,
                     static_cast<uint8_t *>(h->spin));
  hipLaunchKernelGGL(k_seeds_from_shadow, dim3(grid_for(h->bg.seedc / 2, 256)), dim3(256), 0, h->stream,
                     h->sseed, h->g, h->bg, h->seed);
  return check_hip(hipGetLastError(), "bit-plane unpack");
}

__global__ void k_zero_u64(unsigned long long *__restrict__ v, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) v[t] = 0ull;
}

int ensure_pinned(spk_hip_lattice *h, int rows) {
  if (rows <= h->pinned_rows) return SPK_HIP_OK;
  int cap = std::max(rows, 2 * h->pinned_rows);
  unsigned long long *p = nullptr;
  if (hipHostMalloc(&p, (size_t)cap * 2 * sizeof(unsigned long long), hipHostMallocDefault) !=
      hipSuccess)
    return set_error(SPK_HIP_ENOMEM, "pinned host allocation failed");
  if (h->pinned) {
    (void)hipStreamSynchronize(h->stream);
    (void)hipHostFree(h->pinned);
  }
  h->pinned = p;
  h->pinned_rows = cap;
  return SPK_HIP_OK;
}

// Two sweeps in one launch (k_sweep_bits_pair), from the planes and seeds
// at `from` (0: bits/sseed, 1: bits2/sseed2) to the others.
int launch_pair(spk_hip_lattice *h, int from, unsigned long long *acc) {
  const int bands = (h->g.hy + kPairBand - 1) / kPairBand;
  const unsigned grid = (unsigned)((bands + kNumXcd - 1) / kNumXcd * kNumXcd);   // xcd_tile: a multiple of 8
  const uint32_t *bi = from ? h->bits2 : h->bits;
  const int32_t *si = from ? h->sseed2 : h->sseed;
  uint32_t *bo = from ? h->bits : h->bits2;
  int32_t *so = from ? h->sseed : h->sseed2;
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kPairThreads), 0, h->stream, bi, si, bo, so, h->g, h->bg, h->rule, acc);
  };
  // 64 words, the last one whole, and an even number of rows
  const bool full = h->bg.pw == kPairMaxPw && h->bg.r == 31 && h->g.hy % 2 == 0;
  if (h->d.app == SPK_HIP_ISING_SINGLE) full ? go(k_sweep_bits_pair<true, true>) : go(k_sweep_bits_pair<true, false>);
  else full ? go(k_sweep_bits_pair<false, true>) : go(k_sweep_bits_pair<false, false>);
  return check_hip(hipGetLastError(), "two-sweep launch");
}

// Sweep-kernel launches that launch_sweeps makes for n sweeps.
int64_t sweep_launches(const spk_hip_lattice *h, int n) {
  if (h->bits_active && h->pairs) return 2 * (n / 4) + (int64_t)h->g.K * (n % 4);
  return (int64_t)h->g.K * n;
}

// launches of sweeps whose accumulator rows start at acc; returns launches.
// On planes in the pair layout, every four sweeps are two two-sweep
// launches, there and back, so the state ends in bits/sseed; the last n % 4
// sweeps take the per-colour kernel.
int launch_sweeps(spk_hip_lattice *h, int n, unsigned long long *acc, int64_t *nlaunch) {
  int rc;
  int s = 0;
  if (h->bits_active && h->pairs)
    for (; s + 4 <= n; s += 4, *nlaunch += 2)
      if ((rc = launch_pair(h, 0, acc + (size_t)s * kAccRow)) ||
          (rc = launch_pair(h, 1, acc + (size_t)(s + 2) * kAccRow)))
        return rc;
  for (; s < n; s++)
    for (int c = 0; c < h->g.K; c++, (*nlaunch)++)
      if ((rc = launch_sweep(h, c, acc + (size_t)s * kAccRow))) return rc;
  return SPK_HIP_OK;
}

// folds `rows` accumulator rows at acc into compact (Naccept, dE) pairs at out
int launch_fold(spk_hip_lattice *h, const unsigned long long *acc, int rows, unsigned long long *out) {
  hipLaunchKernelGGL(k_fold_rows, dim3(rows), dim3(64), 0, h->stream, acc, kAccSlots, kAccRow,
                     out);
  return check_hip(hipGetLastError(), "fold launch");
}

// LDS bytes of the resident kernel's copy of the state (spins, then seeds
// at a 16-byte boundary), or 0 when the lattice does not fit one workgroup.
size_t resident_bytes(const spk_hip_lattice *h) {
  if (variant_app(h->d.app) || !h->resident) return 0;
  const size_t slots = (size_t)h->g.K * h->g.Mp;
  const size_t bytes = (slots * spin_bytes(h->stype) + 15) / 16 * 16 + slots * sizeof(int32_t);
  return bytes <= h->max_lds ? bytes : 0;
}

// nsweeps sweeps through the resident kernel, in pieces of at most acc_cap
// sweeps; each piece's rows are folded and copied to pinned rows [done, ..).
int run_resident(spk_hip_lattice *h, int nsweeps, size_t lds, int64_t *nlaunch) {
  const size_t row = 2 * sizeof(unsigned long long);
  int rc;
  if ((rc = check_hip(hipEventRecord(h->ev0, h->stream), "event"))) return rc;
  for (int done = 0; done < nsweeps;) {
    const int n = std::min(nsweeps - done, h->acc_cap);
    hipLaunchKernelGGL(k_zero_u64, dim3(grid_for((int64_t)n * kAccRow, 256)), dim3(256), 0, h->stream, h->acc,
                       (int64_t)n * kAccRow);
    const bool two = h->rule.nspins == 2;
    with_lattice(h->g.lat, [&]<int LAT>() {
      with_type(h->stype, [&]<typename T>() {
        auto go = [&](auto kern) {
          (void)hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
          hipLaunchKernelGGL(kern, dim3(1), dim3(1024), lds, h->stream, static_cast<T *>(h->spin), h->seed, h->g,
                             h->rule, n, h->acc);
        };
        if constexpr (sizeof(T) == 1) {
          if (two) {
            go(k_sweep_resident<LAT, T, true>);
            return;
          }
        }
        go(k_sweep_resident<LAT, T, false>);
      });
    });
    if ((rc = check_hip(hipGetLastError(), "resident sweep launch"))) return rc;
    (*nlaunch)++;
    if ((rc = launch_fold(h, h->acc, n, h->accc)) ||
        (rc = check_hip(hipMemcpyAsync(h->pinned + (size_t)done * 2, h->accc, row * n, hipMemcpyDeviceToHost,
                                       h->stream),
                        "accumulator readback")))
      return rc;
    done += n;
  }
  return check_hip(hipEventRecord(h->ev1, h->stream), "event");
}

void row_sum(const spk_hip_lattice *h, int row, int64_t *x, int64_t *y);

// After the launches of a sweeps call: waits, takes the timing, and turns
// the folded rows in pinned memory into cumulative Naccept and energies.
int finish_sweeps(spk_hip_lattice *h, int nsweeps, int64_t nlaunch, int64_t e0, int64_t *naccept_total,
                  int64_t *naccept_per_sweep, double *energy_per_sweep) {
  const size_t row = 2 * sizeof(unsigned long long);
  int rc;
  if ((rc = check_hip(hipStreamSynchronize(h->stream), "sweeps sync"))) return rc;
  float ms = 0.f;
  if ((rc = check_hip(hipEventElapsedTime(&ms, h->ev0, h->ev1), "event time"))) return rc;
  h->last_ms = ms;
  h->last_nlaunch = nlaunch;
  if (h->hacc.size() < (size_t)nsweeps * 2) h->hacc.resize((size_t)nsweeps * 2);
  std::memcpy(h->hacc.data(), h->pinned, row * nsweeps);

  int64_t nacc = 0, eng = e0;
  const bool track = h->energy_valid || energy_per_sweep;
  if (track && !energy_per_sweep) eng = h->energy;
  for (int s = 0; s < nsweeps; s++) {
    int64_t a, de;
    row_sum(h, s, &a, &de);
    nacc += a;
    eng += 2 * de;   // a flip changes site i's and its neighbours' counts
    if (naccept_per_sweep) naccept_per_sweep[s] = nacc;
    if (energy_per_sweep) energy_per_sweep[s] = (double)eng;
  }
  if (naccept_total) *naccept_total += nacc;
  if (track) h->energy = eng;
  h->energy_valid = track;
  return SPK_HIP_OK;
}


// The graph of `n` sweeps on the current layout (bit planes or bytes):
// a kernel zeroes accumulator rows [0, n) (no memset node), then the n
// sweeps' launches, then the fold of those rows into accc[0, n).  Cached
// per (n, layout) until a settings change drops the cache.
int piece_graph(spk_hip_lattice *h, int n, hipGraphExec_t *out) {
  for (auto &p : h->graphs)
    if (p.sweeps == n && p.bits == h->bits_active) {
      *out = p.exec;
      return SPK_HIP_OK;
    }
  int rc;
  hipGraph_t graph = nullptr;
  if ((rc = check_hip(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal), "begin capture")))
    return rc;
  const int64_t words = (int64_t)n * kAccRow;
  hipLaunchKernelGGL(k_zero_u64, dim3(grid_for(words, 256)), dim3(256), 0, h->stream, h->acc, words);
  int64_t nl = 0;
  rc = check_hip(hipGetLastError(), "zero launch");
  if (!rc) rc = launch_sweeps(h, n, h->acc, &nl);
  if (!rc) rc = launch_fold(h, h->acc, n, h->accc);
  const hipError_t e = hipStreamEndCapture(h->stream, &graph);
  if (!rc) rc = check_hip(e, "end capture");
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  hipGraphExec_t exec;
  rc = check_hip(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "graph instantiate");
  (void)hipGraphDestroy(graph);
  if (rc) return rc;
  h->graphs.push_back({n, h->bits_active, exec});
  *out = exec;
  return SPK_HIP_OK;
}

// Sweeps per graph replay of a call of n sweeps: n / chunk full graphs,
// then one of the remainder (so at most chunk + 1 graphs are ever cached,
// and a call shorter than chunk is a single replay).
std::vector<int> graph_pieces(int chunk, int n) {
  std::vector<int> p(chunk > 0 ? n / chunk : 0, chunk);
  if (chunk > 0 && n % chunk) p.push_back(n % chunk);
  return p;
}

int launch_energy(spk_hip_lattice *h, unsigned long long *acc) {
  with_lattice(h->g.lat, [&]<int LAT>() {
    with_type(h->stype, [&]<typename T>() {
      hipLaunchKernelGGL((k_energy<LAT, T>), dim3(grid_for(h->N, 256)), dim3(256), 0, h->stream,
                         static_cast<const T *>(h->spin), h->g, acc);
    });
  });
  return check_hip(hipGetLastError(), "energy kernel launch");
}

int launch_fill_spins(spk_hip_lattice *h, int32_t value) {
  with_type(h->stype, [&]<typename T>() {
    hipLaunchKernelGGL((k_fill<T>), dim3(grid_for(h->slots, 256)), dim3(256), 0, h->stream,
                       static_cast<T *>(h->spin), h->slots, value);
  });
  return check_hip(hipGetLastError(), "fill kernel launch");
}

int ensure_acc(spk_hip_lattice *h, int nsweeps) {
  if (nsweeps <= h->acc_cap) return SPK_HIP_OK;
  int cap = h->acc_cap ? h->acc_cap : kInitialAccRows;
  while (cap < nsweeps) cap *= 2;
  unsigned long long *p = nullptr, *q = nullptr;
  if (hipMalloc(&p, (size_t)cap * kAccRow * sizeof(unsigned long long)) != hipSuccess)
    return set_error(SPK_HIP_ENOMEM, "device allocation of %d accumulator rows failed", cap);
  if (hipMalloc(&q, (size_t)cap * 2 * sizeof(unsigned long long)) != hipSuccess) {
    (void)hipFree(p);
    return set_error(SPK_HIP_ENOMEM, "device allocation of %d accumulator rows failed", cap);
  }
  if (h->acc) {
    (void)hipStreamSynchronize(h->stream);
    (void)hipFree(h->acc);
    (void)hipFree(h->accc);
  }
  drop_graphs(h);   // the accumulator addresses are baked into captured launches
  h->acc = p;
  h->accc = q;
  h->acc_cap = cap;
  h->hacc.resize((size_t)cap * 2);
  return SPK_HIP_OK;
}

// (Naccept, dE) of accumulator row `row`, already folded and copied to h->hacc
void row_sum(const spk_hip_lattice *h, int row, int64_t *x, int64_t *y) {
  *x = (int64_t)h->hacc[2 * (size_t)row];
  *y = (int64_t)h->hacc[2 * (size_t)row + 1];
}

int compute_energy(spk_hip_lattice *h, int64_t *e) {
  int rc;
  if ((rc = bits_leave(h)) || (rc = ensure_acc(h, 1))) return rc;
  if ((rc = check_hip(hipMemsetAsync(h->acc, 0, kAccRow * sizeof(unsigned long long), h->stream), "memset")))
    return rc;
  if ((rc = launch_energy(h, h->acc)) || (rc = launch_fold(h, h->acc, 1, h->accc))) return rc;
  if ((rc = check_hip(hipMemcpyAsync(h->hacc.data(), h->accc, 2 * sizeof(unsigned long long),
                                     hipMemcpyDeviceToHost, h->stream),
                      "energy readback")))
    return rc;
  if ((rc = check_hip(hipStreamSynchronize(h->stream), "energy sync"))) return rc;
  int64_t y;
  row_sum(h, 0, e, &y);
  h->energy = *e;
  h->energy_valid = true;
  return SPK_HIP_OK;
}

// ---------------------------------------------- sweeps without allocation
// Accumulator rows of spk_hip_sweeps_noalloc live in a file-scope device
// array, one copy per GPU, held until the process ends.  Host code uses the
// address hipGetSymbolAddress gives on each device (looked up once per
// device, when the first handle is created there), never the array's name,
// which in host code is only a stand-in.
constexpr int kNoAllocSweeps = 256;   // sweeps per call (8 KiB of rows each: 2 MiB)
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned long long scratch[(size_t)kNoAllocSweeps * kAccRow];
std::mutex g_scratch_mu;
unsigned long long *g_scratch[kMaxDevices] = {};

// Looks up the scratch array's address on the current device `device`.
int scratch_lookup(int device) {
  if (device < 0 || device >= kMaxDevices)
    return set_error(SPK_HIP_EDEVICE, "device %d beyond the %d this library serves", device, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[device]) return SPK_HIP_OK;
  void *p = nullptr;
  const int rc = check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "scratch symbol address");
  if (rc) return rc;
  g_scratch[device] = static_cast<unsigned long long *>(p);
  return SPK_HIP_OK;
}

int validate_on_device(spk_hip_lattice *h, const int32_t *buf, int lo, int hi, int64_t *nbad) {
  int rc = ensure_acc(h, 1);
  if (rc) return rc;
  if ((rc = check_hip(hipMemsetAsync(h->acc, 0, sizeof(unsigned long long), h->stream), "memset")))
    return rc;
  hipLaunchKernelGGL(k_count_outside, dim3(grid_for(h->N, 256)), dim3(256), 0, h->stream, buf, h->N, lo,
                     hi, h->acc);
  if ((rc = check_hip(hipGetLastError(), "validate launch"))) return rc;
  unsigned long long v = 0;
  if ((rc = check_hip(hipMemcpyAsync(&v, h->acc, sizeof(v), hipMemcpyDeviceToHost, h->stream), "readback")))
    return rc;
  if ((rc = check_hip(hipStreamSynchronize(h->stream), "validate sync"))) return rc;
  *nbad = (int64_t)v;
  return SPK_HIP_OK;
}

}  // namespace

// ====================================================================== ABI

extern "C" {

int spk_hip_ncolors(const spk_hip_desc *d) {
  Geom g;
  return make_geom(d, &g) ? SPK_HIP_EINVAL : g.K;
}

int spk_hip_create(const spk_hip_desc *d, spk_hip_lattice **out) {
  if (!d || !out) return set_error(SPK_HIP_EINVAL, "null argument");
  *out = nullptr;
  Geom g;
  int rc = make_geom(d, &g);
  if (rc) return rc;
  Rule r;
  if ((rc = make_rule(d->app, d->nspins, d->lattice, d->temperature, &r))) return rc;
  spk_hip_lattice *h = new (std::nothrow) spk_hip_lattice();
  if (!h) return set_error(SPK_HIP_ENOMEM, "host allocation failed");
  h->d = *d;
  h->g = g;
  h->rule = r;
  h->N = (int64_t)g.M * g.K;
  h->slots = (int64_t)g.Mp * g.K;
  h->stype = spin_type_for(d->app, d->nspins);
  // measured best on MI355X: 512 for 4096^2 sq/4n, 256 for 256^3 sc/26n
  h->block = (g.lat == SPK_HIP_SQ_4N || g.lat == SPK_HIP_SQ_8N) ? 512 : 256;
  if ((rc = check_hip(hipGetDevice(&h->device), "hipGetDevice")) || (rc = scratch_lookup(h->device)) ||
      (rc = check_hip(hipStreamCreate(&h->stream), "hipStreamCreate")) ||
      (rc = check_hip(hipEventCreate(&h->ev0), "hipEventCreate")) ||
      (rc = check_hip(hipEventCreate(&h->ev1), "hipEventCreate"))) {
    spk_hip_destroy(h);
    return rc;
  }
  {
    // LDS a workgroup may take, less room for the resident kernel's static arrays
    int lds = 0, optin = 0;
    (void)hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, h->device);
    (void)hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, h->device);
    const int most = std::max(lds, optin);
    h->max_lds = most > 1024 ? (size_t)most - 1024 : 0;
  }
  if (hipMalloc(&h->spin, (size_t)h->slots * spin_bytes(h->stype)) != hipSuccess ||
      hipMalloc(&h->seed, (size_t)h->slots * sizeof(int32_t)) != hipSuccess ||
      hipMalloc(&h->stage, (size_t)h->N * sizeof(int32_t)) != hipSuccess) {
    spk_hip_destroy(h);
    return set_error(SPK_HIP_ENOMEM, "device allocation of %lld sites failed", (long long)h->N);
  }
  // accumulator rows and their pinned host copy for calls of up to 1024
  // sweeps, so that such calls allocate nothing
  if ((rc = ensure_acc(h, kInitialAccRows)) || (rc = ensure_pinned(h, kInitialAccRows)) ||
      (rc = launch_fill_spins(h, 1))) {
    spk_hip_destroy(h);
    return rc;
  }
  hipLaunchKernelGGL(k_fill<int32_t>, dim3(grid_for(h->slots, 256)), dim3(256), 0, h->stream, h->seed, h->slots,
                     1);
  if ((rc = check_hip(hipStreamSynchronize(h->stream), "create sync"))) {
    spk_hip_destroy(h);
    return rc;
  }
  *out = h;
  return SPK_HIP_OK;
}

void spk_hip_destroy(spk_hip_lattice *h) {
  if (!h) return;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  drop_graphs(h);
  if (h->pinned) (void)hipHostFree(h->pinned);
  if (h->spin) (void)hipFree(h->spin);
  if (h->seed) (void)hipFree(h->seed);
  if (h->stage) (void)hipFree(h->stage);
  if (h->bits) (void)hipFree(h->bits);
  if (h->sseed) (void)hipFree(h->sseed);
  if (h->bits2) (void)hipFree(h->bits2);
  if (h->sseed2) (void)hipFree(h->sseed2);
  if (h->acc) (void)hipFree(h->acc);
  if (h->accc) (void)hipFree(h->accc);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

int spk_hip_set_temperature(spk_hip_lattice *h, double temperature) {
  if (!h) return set_error(SPK_HIP_EINVAL, "null handle");
  Rule r;
  int rc = make_rule(h->d.app, h->d.nspins, h->d.lattice, temperature, &r);
  if (rc) return rc;
  h->rule = r;
  h->d.temperature = temperature;
  drop_graphs(h);   // thresholds are kernel arguments of the captured launches
  return SPK_HIP_OK;
}

int spk_hip_upload(spk_hip_lattice *h, const int32_t *spin, const int32_t *seed) {
  if (!h) return set_error(SPK_HIP_EINVAL, "null handle");
  const size_t bytes = (size_t)h->N * sizeof(int32_t);
  int rc;
  if ((rc = bits_leave(h))) return rc;   // the planes hold the state: back to bytes first
  const int qmax = app_qmax(h->d.app, h->d.nspins);
  if (spin) {
    if ((rc = check_hip(hipMemcpyAsync(h->stage, spin, bytes, hipMemcpyHostToDevice, h->stream), "upload")))
      return rc;
    int64_t nbad = 0;
    if ((rc = validate_on_device(h, h->stage, 1, qmax, &nbad))) return rc;
    if (nbad) return set_error(SPK_HIP_EVALUE, "One or more sites have invalid values (%lld)", (long long)nbad);
    with_type(h->stype, [&]<typename T>() {
      hipLaunchKernelGGL((k_to_blocked<T>), dim3(grid_for(h->N, 256)), dim3(256), 0, h->stream, h->stage,
                         static_cast<T *>(h->spin), h->g);
    });
    if ((rc = check_hip(hipGetLastError(), "layout kernel"))) return rc;
    h->energy_valid = false;
  }
  if (seed) {
    if ((rc = check_hip(hipMemcpyAsync(h->stage, seed, bytes, hipMemcpyHostToDevice, h->stream), "upload")))
      return rc;
    int64_t nbad = 0;
    if ((rc = validate_on_device(h, h->stage, 1, kIM - 1, &nbad))) return rc;
    if (nbad) return set_error(SPK_HIP_EVALUE, "%lld seeds outside [1, 2^31-2]", (long long)nbad);
    hipLaunchKernelGGL(k_to_blocked<int32_t>, dim3(grid_for(h->N, 256)), dim3(256), 0, h->stream, h->stage,
                       h->seed, h->g);
    if ((rc = check_hip(hipGetLastError(), "layout kernel"))) return rc;
  }
  return check_hip(hipStreamSynchronize(h->stream), "upload sync");
}

int spk_hip_download(spk_hip_lattice *h, int32_t *spin, int32_t *seed) {
  if (!h) return set_error(SPK_HIP_EINVAL, "null handle");
  const size_t bytes = (size_t)h->N * sizeof(int32_t);
  int rc;
  if ((rc = bits_leave(h))) return rc;
  if (spin) {
    with_type(h->stype, [&]<typename T>() {
      hipLaunchKernelGGL((k_to_natural<T>), dim3(grid_for(h->N, 256)), dim3(256), 0, h->stream,
                         static_cast<const T *>(h->spin), h->stage, h->g);
    });
    if ((rc = check_hip(hipGetLastError(), "layout kernel"))) return rc;
    if ((rc = check_hip(hipMemcpyAsync(spin, h->stage, bytes, hipMemcpyDeviceToHost, h->stream), "download")))
      return rc;
    if ((rc = check_hip(hipStreamSynchronize(h->stream), "download sync"))) return rc;
  }
  if (seed) {
    hipLaunchKernelGGL(k_to_natural<int32_t>, dim3(grid_for(h->N, 256)), dim3(256), 0, h->stream, h->seed,
                       h->stage, h->g);
    if ((rc = check_hip(hipGetLastError(), "layout kernel"))) return rc;
    if ((rc = check_hip(hipMemcpyAsync(seed, h->stage, bytes, hipMemcpyDeviceToHost, h->stream), "download")))
      return rc;
    if ((rc = check_hip(hipStreamSynchronize(h->stream), "download sync"))) return rc;
  }
  return SPK_HIP_OK;
}

int spk_hip_set_range(spk_hip_lattice *h, int seed0, int lo, int hi) {
  if (!h) return set_error(SPK_HIP_EINVAL, "null handle");
  if (hi < lo) return set_error(SPK_HIP_EINVAL, "set range: hi < lo");
  if (seed0 < 1 || seed0 >= kIM) return set_error(SPK_HIP_EINVAL, "set range: seed0 outside [1, 2^31-2]");
  const int qmax = app_qmax(h->d.app, h->d.nspins);
  if (lo < 1 || hi > qmax)
    return set_error(SPK_HIP_EVALUE, "One or more sites have invalid values (set range %d %d, Q = %d)", lo, hi,
                     qmax);
  int rc = bits_leave(h);   // the seeds may live on the planes
  if (rc) return rc;
  const PowTable pw = park_powers();
  with_type(h->stype, [&]<typename T>() {
    hipLaunchKernelGGL((k_set_range<T>), dim3(grid_for(h->N, 256)), dim3(256), 0, h->stream,
                       static_cast<T *>(h->spin), h->g, seed0, lo, hi - lo + 1, pw);
  });
  rc = check_hip(hipGetLastError(), "set_range launch");
  if (rc) return rc;
  h->energy_valid = false;
  return check_hip(hipStreamSynchronize(h->stream), "set_range sync");
}

int spk_hip_set_value(spk_hip_lattice *h, int value) {
  if (!h) return set_error(SPK_HIP_EINVAL, "null handle");
  const int qmax = app_qmax(h->d.app, h->d.nspins);
  if (value < 1 || value > qmax)
    return set_error(SPK_HIP_EVALUE, "One or more sites have invalid values (set value %d, Q = %d)", value, qmax);
  int rc;
  if ((rc = bits_leave(h)) || (rc = launch_fill_spins(h, value))) return rc;
  h->energy_valid = false;
  return check_hip(hipStreamSynchronize(h->stream), "set_value sync");
}

int spk_hip_init_seeds(spk_hip_lattice *h, double rseed) {
  if (!h) return set_error(SPK_HIP_EINVAL, "null handle");
  if (!(rseed >= 0.0 && rseed < 1.0)) return set_error(SPK_HIP_EINVAL, "init_seeds: rseed outside [0,1)");
  int rc = bits_leave(h);   // the spins may live on the planes
  if (rc) return rc;
  const double xbase = rseed * (double)kIM;
  const uint32_t a100 = park_warmup100();
  hipLaunchKernelGGL(k_init_seeds, dim3(grid_for(h->N, 256)), dim3(256), 0, h->stream, h->seed, h->g, xbase,
                     a100);
  rc = check_hip(hipGetLastError(), "init_seeds launch");
  if (rc) return rc;
  return check_hip(hipStreamSynchronize(h->stream), "init_seeds sync");
}

int spk_hip_sweeps(spk_hip_lattice *h, int nsweeps, int64_t *naccept_total, int64_t *naccept_per_sweep,
                   double *energy_per_sweep) {
  if (!h) return set_error(SPK_HIP_EINVAL, "null handle");
  if (nsweeps < 0) return set_error(SPK_HIP_EINVAL, "nsweeps < 0");
  if (nsweeps == 0) return SPK_HIP_OK;
  int rc;
  int64_t e0 = 0;
  if (energy_per_sweep) {
    if (h->energy_valid) e0 = h->energy;
    else if ((rc = compute_energy(h, &e0))) return rc;
  }
  const size_t row = 2 * sizeof(unsigned long long);   // a folded row
  int64_t nlaunch = 0;
  if (const size_t lds = resident_bytes(h)) {
    if ((rc = bits_leave(h)) || (rc = ensure_acc(h, std::min(nsweeps, kInitialAccRows))) ||
        (rc = ensure_pinned(h, nsweeps)) || (rc = run_resident(h, nsweeps, lds, &nlaunch)))
      return rc;
    return finish_sweeps(h, nsweeps, nlaunch, e0, naccept_total, naccept_per_sweep, energy_per_sweep);
  }
  // two-state sq/4n: the state moves to the bit planes and stays there
  if (bits_eligible(h) && !h->bits_active && (rc = bits_enter(h))) return rc;
  if ((rc = ensure_pinned(h, nsweeps))) return rc;
  if (h->chunk > 0) {
    // graph replays (graph_pieces); each piece's rows are folded into
    // accc[0, n) and copied to the pinned rows [done, done + n) in stream order
    const std::vector<int> pieces = graph_pieces(h->chunk, nsweeps);
    if ((rc = ensure_acc(h, pieces[0]))) return rc;
    std::vector<hipGraphExec_t> execs(pieces.size());
    for (size_t k = 0; k < pieces.size(); k++)
      if ((rc = piece_graph(h, pieces[k], &execs[k]))) return rc;
    if ((rc = check_hip(hipEventRecord(h->ev0, h->stream), "event"))) return rc;
    int done = 0;
    for (size_t k = 0; k < pieces.size(); k++) {
      if ((rc = check_hip(hipGraphLaunch(execs[k], h->stream), "graph launch")) ||
          (rc = check_hip(hipMemcpyAsync(h->pinned + (size_t)done * 2, h->accc, row * pieces[k],
                                         hipMemcpyDeviceToHost, h->stream),
                          "accumulator readback")))
        return rc;
      done += pieces[k];
    }
    for (int n : pieces) nlaunch += sweep_launches(h, n);
  } else {
    // every launch from the host: rows [0, nsweeps)
    if ((rc = ensure_acc(h, nsweeps))) return rc;
    const int64_t words = (int64_t)nsweeps * kAccRow;
    if ((rc = check_hip(hipEventRecord(h->ev0, h->stream), "event"))) return rc;
    hipLaunchKernelGGL(k_zero_u64, dim3(grid_for(words, 256)), dim3(256), 0, h->stream, h->acc, words);
    if ((rc = check_hip(hipGetLastError(), "zero launch")) || (rc = launch_sweeps(h, nsweeps, h->acc, &nlaunch)) ||
        (rc = launch_fold(h, h->acc, nsweeps, h->accc)) ||
        (rc = check_hip(hipMemcpyAsync(h->pinned, h->accc, row * nsweeps, hipMemcpyDeviceToHost, h->stream),
                        "accumulator readback")))
      return rc;
  }
  if ((rc = check_hip(hipEventRecord(h->ev1, h->stream), "event"))) return rc;
  return finish_sweeps(h, nsweeps, nlaunch, e0, naccept_total, naccept_per_sweep, energy_per_sweep);
}

int spk_hip_sweeps_noalloc(spk_hip_lattice *h, int nsweeps, unsigned long long *rows) {
  if (!h || !rows) return set_error(SPK_HIP_EINVAL, "null argument");
  if (nsweeps < 0 || nsweeps > kNoAllocSweeps)
    return set_error(SPK_HIP_EINVAL, "nsweeps %d outside [0, %d] for sweeps without allocation", nsweeps,
                     kNoAllocSweeps);
  if (nsweeps == 0) return SPK_HIP_OK;
  unsigned long long *acc = g_scratch[h->device];   // looked up when the handle was created
  if (!acc) return set_error(SPK_HIP_ESTATE, "no scratch rows on device %d", h->device);
  // the state in the byte layout (an unpack when the planes hold it: kernels
  // only, into buffers that exist), then per-colour launches on the stream
  int rc = bits_leave(h);
  if (rc) return rc;
  const int64_t words = (int64_t)nsweeps * kAccRow;
  hipLaunchKernelGGL(k_zero_u64, dim3(grid_for(words, 256)), dim3(256), 0, h->stream, acc, words);
  int64_t nlaunch = 0;
  if ((rc = check_hip(hipGetLastError(), "zero launch")) || (rc = launch_sweeps(h, nsweeps, acc, &nlaunch)) ||
      (rc = launch_fold(h, acc, nsweeps, rows)))
    return rc;
  h->energy_valid = false;   // the energy changed on the device only
  return SPK_HIP_OK;
}

int spk_hip_prepare(spk_hip_lattice *h, int nsweeps) {
  if (!h) return set_error(SPK_HIP_EINVAL, "null handle");
  if (nsweeps < 0) return set_error(SPK_HIP_EINVAL, "nsweeps < 0");
  if (nsweeps == 0 || resident_bytes(h)) return SPK_HIP_OK;
  int rc;
  if (bits_eligible(h) && !h->bits_active && (rc = bits_enter(h))) return rc;
  if ((rc = ensure_pinned(h, nsweeps))) return rc;
  if (h->chunk > 0) {
    const std::vector<int> pieces = graph_pieces(h->chunk, nsweeps);
    if ((rc = ensure_acc(h, pieces[0]))) return rc;
    for (int n : pieces) {
      hipGraphExec_t exec;
      if ((rc = piece_graph(h, n, &exec))) return rc;
    }
  } else if ((rc = ensure_acc(h, nsweeps))) {
    return rc;
  }
  return check_hip(hipStreamSynchronize(h->stream), "prepare sync");
}

int spk_hip_energy(spk_hip_lattice *h, double *energy) {
  if (!h || !energy) return set_error(SPK_HIP_EINVAL, "null argument");
  int64_t e;
  int rc = compute_energy(h, &e);
  if (rc) return rc;
  *energy = (double)e;
  return SPK_HIP_OK;
}

int spk_hip_checksum(spk_hip_lattice *h, uint64_t *sum) {
  if (!h || !sum) return set_error(SPK_HIP_EINVAL, "null argument");
  int rc;
  if ((rc = bits_leave(h)) || (rc = ensure_acc(h, 1))) return rc;
  if ((rc = check_hip(hipMemsetAsync(h->acc, 0, sizeof(unsigned long long), h->stream), "memset"))) return rc;
  with_type(h->stype, [&]<typename T>() {
    hipLaunchKernelGGL((k_checksum<T>), dim3(grid_for(h->N, 256)), dim3(256), 0, h->stream,
                       static_cast<const T *>(h->spin), h->seed, h->g, h->acc);
  });
  if ((rc = check_hip(hipGetLastError(), "checksum launch"))) return rc;
  unsigned long long v = 0;
  if ((rc = check_hip(hipMemcpyAsync(&v, h->acc, sizeof v, hipMemcpyDeviceToHost, h->stream), "readback")) ||
      (rc = check_hip(hipStreamSynchronize(h->stream), "sync")))
    return rc;
  *sum = v;
  return SPK_HIP_OK;
}

int spk_hip_sync(spk_hip_lattice *h) {
  if (!h) return set_error(SPK_HIP_EINVAL, "null handle");
  return check_hip(hipStreamSynchronize(h->stream), "sync");
}

int spk_hip_set_launch(spk_hip_lattice *h, int block) {
  if (!h) return set_error(SPK_HIP_EINVAL, "null handle");
  if (block < 64 || block > 1024 || (block & (block - 1)))
    return set_error(SPK_HIP_EINVAL, "block must be a power of two in [64, 1024]");
  h->block = block;
  drop_graphs(h);
  return SPK_HIP_OK;
}

int spk_hip_set_resident(spk_hip_lattice *h, int enable) {
  if (!h) return set_error(SPK_HIP_EINVAL, "null handle");
  h->resident = enable ? 1 : 0;
  return SPK_HIP_OK;
}

int spk_hip_set_bitplanes(spk_hip_lattice *h, int enable) {
  if (!h) return set_error(SPK_HIP_EINVAL, "null handle");
  if (enable != 0 && enable != 1 && enable != 8 && enable != 16 && enable != 32)
    return set_error(SPK_HIP_EINVAL, "bit planes: 0 (off), 1 (on), or 8, 16, 32 sites per thread");
  const int rc = bits_leave(h);   // the plane layout depends on the sites per thread
  if (rc) return rc;
  h->bitplanes = enable;
  drop_graphs(h);   // captured launches name one kernel or the other
  return SPK_HIP_OK;
}

int spk_hip_set_graph(spk_hip_lattice *h, int chunk_sweeps) {
  if (!h) return set_error(SPK_HIP_EINVAL, "null handle");
  if (chunk_sweeps < 0 || chunk_sweeps > kMaxChunk)
    return set_error(SPK_HIP_EINVAL, "chunk_sweeps must be in [0, %d]", kMaxChunk);
  drop_graphs(h);
  h->chunk = chunk_sweeps;
  return SPK_HIP_OK;
}

int spk_hip_last_timing(spk_hip_lattice *h, double *ms, int64_t *nlaunch) {
  if (!h) return set_error(SPK_HIP_EINVAL, "null handle");
  if (ms) *ms = h->last_ms;
  if (nlaunch) *nlaunch = h->last_nlaunch;
  return SPK_HIP_OK;
}

}  // extern "C"
