// Two-state sq/4n lattices (Ising, Potts Q=2) as bit planes: the sweep path
// for the 2d Ising workloads.
//
// A colour pass of sweep color/strict (src/app_lattice.cpp:1023-1061,
// sweep_nomask_strict) on a checkerboard updates every site of one colour
// from its four neighbours, all of the other colour.  With two states a site
// is one bit (spin - 1), so a 32-bit word holds 32 sites of one sub-grid row
// and the four neighbour counts of all 32 come out of a bit-sliced adder over
// four "differs" words; only the Park-Miller draws and the accept test stay
// per site.  The per-site seeds (the reference's ranstrict[] generators,
// src/app_lattice.cpp:346-351) keep their 4 bytes and dominate the traffic.
//
// The planes are a shadow of the u8 colour-blocked state: the first sweeps
// call packs them, and they hold the state across later sweeps calls until
// a call that reads or writes the u8 layout (energy, transfers, checksum,
// set-up) unpacks them, so every other kernel sees the usual layout.
//
// Layout.  Colour c, row b (0..hy-1), word w (0..pw-1, pw = ceil(hx/32)):
//   bits[(c*hy + b)*pw + w], bit i = site a = 32*w + i of that row
//   (bits of a >= hx are padding).
// A thread owns SPT (8, 16 or 32) consecutive sites of one word: thread t of
// a colour takes part q = t % (32/SPT) of word t / (32/SPT).  Its local site
// i (0..SPT-1) has its seed at
//   c*seedc + ((((t>>6)*(SPT/4) + (i>>2))*64 + (t&63))*4 + (i&3)
// so a wavefront's j-th 16-byte seed load (local sites 4j..4j+3 of every
// lane) covers 1 KiB of consecutive bytes.
#pragma once

#include <type_traits>

#include "spk_device.h"

namespace spk {

struct BitGeom {
  int pw;           // words per sub-grid row
  int spt;          // sites per thread (8, 16, 32)
  int tpw;          // threads per word (32 / spt)
  int64_t threads;  // threads per colour with a real word (hy*pw*tpw)
  int64_t nt;       // threads per colour rounded up to 64
  int64_t seedc;    // shadow seed slots per colour (nt*spt)
  int r;            // bit of site hx-1 in a row's last word
  uint32_t lastmask;   // real-site bits of a row's last word
  FastDiv fpw;      // division by pw
};

__host__ __device__ inline BitGeom bit_geom(const Geom &g, int spt) {
  BitGeom bg;
  bg.pw = (g.hx + 31) / 32;
  bg.spt = spt;
  bg.tpw = 32 / spt;
  bg.threads = (int64_t)g.hy * bg.pw * bg.tpw;
  bg.nt = (bg.threads + 63) / 64 * 64;
  bg.seedc = bg.nt * spt;
  bg.r = (g.hx - 1) & 31;
  bg.lastmask = bg.r == 31 ? 0xffffffffu : (2u << bg.r) - 1u;
  bg.fpw = fastdiv_make((uint32_t)bg.pw);
  return bg;
}

__host__ __device__ __forceinline__ int64_t bit_seed_index(const BitGeom &bg, int c, uint32_t t, int i) {
  return (int64_t)c * bg.seedc + ((((int64_t)(t >> 6) * (bg.spt / 4) + (i >> 2)) * 64 + (t & 63)) * 4 + (i & 3));
}

// --------------------------------------------------------------- conversions
// Word (c, b, w) of the planes from the u8 state: bit i = spin - 1 of site
// a = 32w+i (a < hx).  Spins are read a 4-byte group at a time (px is a
// multiple of 4, so a group never leaves its row).
__host__ __device__ __forceinline__ uint32_t bits_pack_word(const uint8_t *__restrict__ spin, const Geom &g, int c,
                                                            int b, int w) {
  const uint8_t *row = spin + (int64_t)c * g.Mp + (int64_t)b * g.px;
  uint32_t out = 0;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const int a = 32 * w + 4 * j;
    if (a >= g.hx) break;
    const uint32_t q = *reinterpret_cast<const uint32_t *>(row + a);
#pragma unroll
    for (int v = 0; v < 4; v++)
      if (a + v < g.hx) out |= (((q >> (8 * v)) - 1u) & 1u) << (4 * j + v);
  }
  return out;
}

// The inverse: spins 1 + bit for the real sites of word (c, b, w); the pad
// bytes of a row keep their values.
__host__ __device__ __forceinline__ void bits_unpack_word(uint8_t *__restrict__ spin, const Geom &g, int c, int b,
                                                          int w, uint32_t word) {
  uint8_t *row = spin + (int64_t)c * g.Mp + (int64_t)b * g.px;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const int a = 32 * w + 4 * j;
    if (a >= g.hx) break;
    uint32_t q = *reinterpret_cast<const uint32_t *>(row + a);
#pragma unroll
    for (int v = 0; v < 4; v++)
      if (a + v < g.hx) {
        const uint32_t s = 1u + ((word >> (4 * j + v)) & 1u);
        q = (q & ~(0xffu << (8 * v))) | (s << (8 * v));
      }
    *reinterpret_cast<uint32_t *>(row + a) = q;
  }
}

// Shadow slot s (0 .. 2*seedc-1) -> colour, thread, local site.
__host__ __device__ __forceinline__ void bit_seed_slot(const BitGeom &bg, int64_t s, int &c, uint32_t &t, int &i) {
  c = s >= bg.seedc;
  const int64_t r = s - (c ? bg.seedc : 0);
  const int64_t per = 64 * (int64_t)bg.spt;   // slots of 64 threads
  const int64_t grp = r / per;
  const int within = (int)(r - grp * per);
  const int j = within >> 8;                  // 16-byte load index
  const int l = (within >> 2) & 63;           // lane
  i = 4 * j + (within & 3);
  t = (uint32_t)(grp * 64 + l);
}

// Canonical seed offset of shadow slot (c, t, i), or -1 for a padding slot.
__host__ __device__ __forceinline__ int64_t bit_seed_home(const Geom &g, const BitGeom &bg, int c, uint32_t t, int i) {
  if ((int64_t)t >= bg.threads) return -1;
  const uint32_t wd = t / (uint32_t)bg.tpw, q = t % (uint32_t)bg.tpw;
  uint32_t b, w;
  fastdiv(bg.fpw, wd, b, w);
  const int a = 32 * (int)w + (int)q * bg.spt + i;
  if (a >= g.hx) return -1;
  return (int64_t)c * g.Mp + (int64_t)b * g.px + a;
}

// The conversions move seeds in 16-byte groups: group X (0 .. seedc/2 - 1)
// is local sites 4j..4j+3 of thread t of colour c, i.e. shadow slots
// 4*unit .. 4*unit+3.  Within each block of 64 threads, consecutive X take
// j fastest, then t: the canonical side of consecutive groups is consecutive
// 16-byte pieces of one row, and each run of spt/4 groups lands on 8
// consecutive shadow groups of one j.
__host__ __device__ __forceinline__ void bit_seed_group(const BitGeom &bg, int64_t X, int &c, uint32_t &t, int &j,
                                                        int64_t &unit) {
  const int64_t per = bg.seedc / 4;             // groups per colour
  c = X >= per;
  const int64_t r = X - (c ? per : 0);
  const int64_t units = 16 * (int64_t)bg.spt;   // groups of a block of 64 threads
  const int64_t blk = r / units;
  const int x = (int)(r - blk * units);
  const int lpt = bg.spt / 4;                   // groups per thread
  const int tl = x / lpt;
  j = x - tl * lpt;
  t = (uint32_t)(blk * 64 + tl);
  unit = c * per + blk * units + (int64_t)j * 64 + tl;
}

// Real sites (0..4) of the group starting at local site 4j of thread t, and
// the canonical offset of the first.  Sites of a thread lie on one row and
// a = 32w + q*spt + 4j is a multiple of 4, so a full group is one aligned
// 16-byte piece.
__host__ __device__ __forceinline__ int bit_seed_group_home(const Geom &g, const BitGeom &bg, int c, uint32_t t, int j,
                                                            int64_t &home) {
  home = bit_seed_home(g, bg, c, t, 4 * j);
  if (home < 0) return 0;
  int n = 1;
  while (n < 4 && bit_seed_home(g, bg, c, t, 4 * j + n) >= 0) n++;
  return n;
}

// canonical seeds -> shadow group X (padding slots hold 1)
__host__ __device__ __forceinline__ void bits_seed_group_in(const int32_t *__restrict__ seed, const Geom &g,
                                                            const BitGeom &bg, int64_t X, int32_t *__restrict__ sseed) {
  int c, j;
  uint32_t t;
  int64_t unit, home;
  bit_seed_group(bg, X, c, t, j, unit);
  const int n = bit_seed_group_home(g, bg, c, t, j, home);
  int4 v = make_int4(1, 1, 1, 1);
  if (n == 4) {
    v = *reinterpret_cast<const int4 *>(seed + home);
  } else if (n > 0) {
    v.x = seed[home];
    if (n > 1) v.y = seed[home + 1];
    if (n > 2) v.z = seed[home + 2];
  }
  reinterpret_cast<int4 *>(sseed)[unit] = v;
}

// shadow group X -> canonical seeds of its real sites
__host__ __device__ __forceinline__ void bits_seed_group_out(const int32_t *__restrict__ sseed, const Geom &g,
                                                             const BitGeom &bg, int64_t X, int32_t *__restrict__ seed) {
  int c, j;
  uint32_t t;
  int64_t unit, home;
  bit_seed_group(bg, X, c, t, j, unit);
  const int n = bit_seed_group_home(g, bg, c, t, j, home);
  if (n == 0) return;
  const int4 v = reinterpret_cast<const int4 *>(sseed)[unit];
  if (n == 4) {
    *reinterpret_cast<int4 *>(seed + home) = v;
    return;
  }
  seed[home] = v.x;
  if (n > 1) seed[home + 1] = v.y;
  if (n > 2) seed[home + 2] = v.z;
}

// ------------------------------------------------------ bitwise primitives
// Truth table of f(x, y, z) over bits in the order of v_bitop3_b32's
// immediate: bit (x<<2 | y<<1 | z).
template <class F>
constexpr uint32_t truth3(F f) {
  uint32_t m = 0;
  for (int k = 0; k < 8; k++)
    if (f((k >> 2) & 1, (k >> 1) & 1, k & 1)) m |= 1u << k;
  return m;
}

// f applied bitwise to (x, y, z), f given by its truth table
template <uint32_t TT>
__host__ __device__ __forceinline__ uint32_t bop3(uint32_t x, uint32_t y, uint32_t z) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_amdgcn_bitop3_b32(x, y, z, TT);
#else
  uint32_t r = 0;
  for (int k = 0; k < 8; k++)
    if ((TT >> k) & 1u) r |= ((k & 4) ? x : ~x) & ((k & 2) ? y : ~y) & ((k & 1) ? z : ~z);
  return r;
#endif
}

// ~0 when bit i of x is set, else 0
__host__ __device__ __forceinline__ uint32_t bit_mask(uint32_t x, int i) {
#ifdef __HIP_DEVICE_COMPILE__
  return (uint32_t)__builtin_amdgcn_sbfe((int)x, (unsigned)i, 1u);
#else
  return 0u - ((x >> i) & 1u);
#endif
}


// (acc << 1) | (x >> 31): the sign bit of x shifted in at bit 0
__host__ __device__ __forceinline__ uint32_t shift_in_sign(uint32_t acc, uint32_t x) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_amdgcn_alignbit(acc, x, 31u);
#else
  return (acc << 1) | (x >> 31);
#endif
}

constexpr uint32_t kSel = truth3([](int x, int y, int z) { return x ? y : z; });            // x ? y : z
constexpr uint32_t kAnd3 = truth3([](int x, int y, int z) { return x & y & z; });           // x & y & z
constexpr uint32_t kKeep = truth3([](int x, int y, int z) { return x & (!y | z); });        // x & (~y | z)
constexpr uint32_t kOrAnd = truth3([](int x, int y, int z) { return (x & y) | z; });        // (x & y) | z
constexpr uint32_t kAndNot2 = truth3([](int x, int y, int z) { return x & !y & !z; });      // x & ~y & ~z
constexpr uint32_t kBothDiff = truth3([](int x, int y, int z) { return (x ^ y) & (x ^ z); });   // y and z unlike x
constexpr uint32_t kXor3 = truth3([](int x, int y, int z) { return x ^ y ^ z; });            // x ^ y ^ z
constexpr uint32_t kNor3 = truth3([](int x, int y, int z) { return !(x | y | z); });         // ~(x | y | z)

// ~0 when bit 30 of x is set, else 0.  Opaque to the compiler, which would
// otherwise fold it with the xor that follows into two shifts, an xor and
// an arithmetic shift (four instructions for three).
__host__ __device__ __forceinline__ uint32_t bit30_mask(uint32_t x) {
#ifdef __HIP_DEVICE_COMPILE__
  uint32_t m;
  asm("v_bfe_i32 %0, %1, 30, 1" : "=v"(m) : "v"(x));
  return m;
#else
  return bit_mask(x, 30);
#endif
}

// -------------------------------------------------------------- the events
// The SPT site events of one thread: part q (sites 32q/TPW .. +SPT-1) of
// word w of a colour-c row whose x-neighbours sit at a and a+1 (p = 1) or
// a-1 and a (p = 0).  own is the row's word; up and down the other colour's
// words of the rows above and below; x the other colour's word of the same
// row and w; xe its neighbour word (w+1 for p = 1, w-1 for p = 0, wrapping
// at the row's ends).  The events are those of AppIsing::
// site_event_rejection (src/app_ising.cpp:104-136) and, for Q=2,
// AppPotts::site_event_rejection (src/app_potts.cpp:112-146), which draw
// the same way: the first draw proposes 1 (below half) or 2; a change (the
// proposal differs from the spin) with dE = ef - ei = 4 - 2*ei > 0 at T > 0
// takes a second draw and is kept when it is below thr[dE]; a change with
// dE <= 0 is kept without one; T = 0 rejects uphill changes without one.
// SINGLE: AppIsingSingle::site_event_rejection (src/app_ising_single.cpp:
// 39-77) instead: the proposal is the flip, without a draw; a flip with
// dE <= 0 is kept, and an uphill one at T > 0 takes the one draw and is
// kept when it is below thr[dE].

// The x-neighbour word shifted into place: site i's neighbour at a+1
// (p = 1) or a-1 (p = 0); across a row's ends the periodic image (a = 0 or
// hx-1).  FULLW: every word of the row holds 32 sites (bg.r == 31), so the
// shift is one funnel of x and xe.
template <bool FULLW>
__host__ __device__ __forceinline__ uint32_t bits_xshift(uint32_t x, uint32_t xe, int w, int p, const BitGeom &bg) {
  if constexpr (FULLW) {
    // p = 1: site 0 of xe enters at bit 31; p = 0: site 31 of xe at bit 0
#ifdef __HIP_DEVICE_COMPILE__
    return p ? __builtin_amdgcn_alignbit(xe, x, 1u) : __builtin_amdgcn_alignbit(x, xe, 31u);
#else
    return p ? (x >> 1) | (xe << 31) : (x << 1) | (xe >> 31);
#endif
  } else {
    const int last = bg.pw - 1;
    if (p) {
      const uint32_t e = xe & 1u;   // site 0 of word w+1, or of word 0 after the last word
      const uint32_t at = w == last ? 1u << bg.r : 0x80000000u;
      return ((x >> 1) & ~at) | (e ? at : 0u);
    }
    const uint32_t e = (xe >> (w == 0 ? bg.r : 31)) & 1u;   // site 31 of word w-1, or site hx-1
    return (x << 1) | e;
  }
}

// Unlike-neighbour counts ei = s0 + 2*s1 + 4*s2 of all 32 sites of a word,
// and upw: the sites with ei <= 1, where a change is uphill (dE = 4 or 2).
// s2 = ~upw & ~s1 (ei 4 alone has neither).
struct BitsEi {
  uint32_t s0, s1, upw;
};

__host__ __device__ __forceinline__ BitsEi bits_ei(uint32_t own, uint32_t up, uint32_t down, uint32_t x, uint32_t xs) {
  // per neighbour pair (up, down) and (x, xs): t = one of the two unlike
  // own, c = both; ei = t1 + t2 + 2*(c1 + c2), and c3 = t1 & t2 is the
  // carry of t1 + t2 (it excludes c1 and c2)
  const uint32_t t1 = up ^ down, c1 = bop3<kBothDiff>(own, up, down);
  const uint32_t t2 = x ^ xs, c2 = bop3<kBothDiff>(own, x, xs);
  const uint32_t c3 = t1 & t2;
  return BitsEi{t1 ^ t2, bop3<kXor3>(c1, c2, c3), bop3<kNor3>(c1, c2, c3)};
}

// The SPT events of sites 0..SPT-1 of own_q, upw_q and s0_q (the word's own
// bits, upw and s0 shifted down so the thread's first site is bit 0).  Per
// site: proposal draw, second draw, accept; masks are 0 or ~0.  At T = 0
// both thresholds are 0, so no second draw passes; the draw itself is not
// taken (nzt clears it).  The proposal is 2 when the first draw is at least
// half = 2^30 (bit 30 of a value below 2^31; the caller checks half).  A
// site's "kept" flag is the sign bit of keep, and enters the result at bit
// 0 as the earlier sites move up (sites run from the last down, so site i
// ends at bit i).  sq holds the sites' seeds; sink(j, words) receives the
// new seeds of local sites 4j..4j+3.  Returns the kept sites (bit i = site
// i; bits SPT and up are 0).
template <int SPT, bool SINGLE, class Sink>
__host__ __device__ __forceinline__ uint32_t bits_sites(uint32_t own_q, uint32_t upw_q, uint32_t s0_q, const Rule &r,
                                                        const Words<4> (&sq)[SPT / 4], Sink &&sink) {
  const uint32_t nzt = r.zero_t ? 0u : 0xffffffffu;
  const int32_t thr2 = r.zero_t ? 0 : r.thr[2], thr4 = r.zero_t ? 0 : r.thr[4];
  uint32_t ch_q = 0;
#pragma unroll
  for (int j = SPT / 4 - 1; j >= 0; j--) {
    Words<4> nq;
#pragma unroll
    for (int v = 3; v >= 0; v--) {
      const int i = 4 * j + v;
      const uint32_t upm = bit_mask(upw_q, i);
      const int32_t thr = (int32_t)bop3<kSel>(bit_mask(s0_q, i), (uint32_t)thr2, (uint32_t)thr4);   // ei 1: dE 2
      uint32_t keep;
      if constexpr (SINGLE) {
        const int32_t s = (int32_t)sq[j].w[v];
        const int32_t u = pm_step(s);
        const uint32_t pass = (uint32_t)u - (uint32_t)thr;   // sign bit: u < thr
        nq.w[v] = bop3<kSel>(upm & nzt, (uint32_t)u, (uint32_t)s);
        keep = bop3<kKeep>(0xffffffffu, upm, pass);
      } else {
        const int32_t u1 = pm_step((int32_t)sq[j].w[v]);
        const uint32_t flipm = bit30_mask((uint32_t)u1) ^ bit_mask(own_q, i);   // proposal 2 differs from the spin
        const int32_t u2 = pm_step(u1);
        const uint32_t pass = (uint32_t)u2 - (uint32_t)thr;   // sign bit: u2 < thr
        const uint32_t drawm = bop3<kAnd3>(flipm, upm, nzt);
        nq.w[v] = bop3<kSel>(drawm, (uint32_t)u2, (uint32_t)u1);
        keep = bop3<kKeep>(flipm, upm, pass);
      }
      ch_q = shift_in_sign(ch_q, keep);
    }
    sink(j, nq);
  }
  return ch_q;
}

// All of the above for part q of word w: returns the changed sites as a
// word mask (bit i = site 32w+i) in ch, and (changes, their dE sum).
template <int SPT, bool SINGLE, class Sink>
__host__ __device__ __forceinline__ Counts bits_events(uint32_t own, uint32_t up, uint32_t down, uint32_t x,
                                                       uint32_t xe, int w, int p, int q, const BitGeom &bg,
                                                       const Rule &r, const Words<4> (&sq)[SPT / 4], uint32_t &ch,
                                                       Sink &&sink) {
  static_assert(SPT == 8 || SPT == 16 || SPT == 32, "8, 16 or 32 sites per thread");
  const BitsEi ei = bits_ei(own, up, down, x, bits_xshift<false>(x, xe, w, p, bg));
  const int i0 = q * SPT;   // this thread's part of the word, shifted down so site i is bit i
  const uint32_t ch_q = bits_sites<SPT, SINGLE>(own >> i0, ei.upw >> i0, ei.s0 >> i0, r, sq, sink);
  ch = SPT == 32 ? ch_q : ch_q << i0;
  if (w == bg.pw - 1) ch &= bg.lastmask;
  Counts out;
  out.nacc = __builtin_popcount(ch);
  // sum over changes of dE = 4 - 2*ei
  const int sum_ei = __builtin_popcount(ch & ei.s0) + 2 * __builtin_popcount(ch & ei.s1) +
                     4 * __builtin_popcount(bop3<kAndNot2>(ch, ei.upw, ei.s1));
  out.de = 4 * out.nacc - 2 * sum_ei;
  return out;
}

// One colour pass's work of thread t of colour c on the planes in place:
// its SPT site events (bits_events), the new seeds and changed bits written
// back.
template <int SPT, bool SINGLE = false>
__host__ __device__ __forceinline__ Counts bits_thread(uint32_t *__restrict__ bits, int32_t *__restrict__ sseed,
                                                       const Geom &g, const BitGeom &bg, const Rule &r, int c,
                                                       uint32_t t) {
  constexpr int TPW = 32 / SPT;
  const uint32_t wd = t / TPW;
  const int q = (int)(t % TPW);
  uint32_t b, wu;
  fastdiv(bg.fpw, wd, b, wu);
  const int w = (int)wu;
  const int last = bg.pw - 1;
  const int o = 1 - c;
  const int p = ((int)b + c) & 1;   // x-neighbours at a and a+1 (p = 1) or a-1 (p = 0)
  const int64_t pw = bg.pw;
  const uint32_t *ob = bits + ((int64_t)o * g.hy) * pw;
  const int64_t self = ((int64_t)c * g.hy + b) * pw + w;
  const int wn = p ? (w == last ? 0 : w + 1) : (w == 0 ? last : w - 1);

  // ---- loads: five plane words and the thread's seeds, before any is used
  const uint32_t own = bits[self];
  const uint32_t up = ob[(int64_t)wrap_dec((int)b, g.hy) * pw + w];
  const uint32_t down = ob[(int64_t)wrap_inc((int)b, g.hy) * pw + w];
  const uint32_t x = ob[(int64_t)b * pw + w];
  const uint32_t xe = ob[(int64_t)b * pw + wn];
  Words<4> sq[SPT / 4];
#pragma unroll
  for (int j = 0; j < SPT / 4; j++) sq[j] = *reinterpret_cast<const Words<4> *>(sseed + bit_seed_index(bg, c, t, 4 * j));
  fence_schedule();

  uint32_t ch;
  const Counts out = bits_events<SPT, SINGLE>(own, up, down, x, xe, w, p, q, bg, r, sq, ch,
                                              [&](int j, const Words<4> &nq) {
                                                *reinterpret_cast<Words<4> *>(sseed + bit_seed_index(bg, c, t, 4 * j)) = nq;
                                              });
  // ---- this thread's part of the word
  if (ch) {
    const uint32_t nw = own ^ ch;
    if constexpr (SPT == 32) {
      bits[self] = nw;
    } else if constexpr (SPT == 16) {
      reinterpret_cast<uint16_t *>(bits + self)[q] = (uint16_t)(nw >> (16 * q));
    } else {
      reinterpret_cast<uint8_t *>(bits + self)[q] = (uint8_t)(nw >> (8 * q));
    }
  }
  return out;
}

// ------------------------------------------------- two sweeps per launch
// A lattice whose sub-grid rows hold 33..64 plane words (sq/4n, nx 2050 ..
// 4096) can run two whole sweeps (colour passes 0, 1, 0, 1) in one launch,
// reading each seed once and writing it once per two sweeps instead of per
// sweep.  A workgroup owns a band of kPairBand rows starting at y0 and holds
// the planes of the band and kPairHalo rows either side in LDS.  Pass p
// (colour p & 1) updates the rows y0-3+p <= rho < y0+kPairBand+3-p: each
// pass shrinks the span that is still exact by one row at each end, and the
// band's own rows are exact after all four.  Rows are taken periodically,
// so rho may lie outside 0..hy-1.  The workgroup's threads form kPairGroups
// groups of kPairStage (8 sites each: the spt-8 layout); group g takes the
// rows y0-3+c+g+kPairGroups*j (slot j) of colour c in every pass, so each
// thread keeps the seeds of its rows in registers from their one load (two
// slots ahead of first use) to the one store of the band's own rows after
// the colour's last pass; outside the ising/single rule one colour-0 slot
// waits in LDS (PairLds::parked) through pass 1.  A barrier separates the passes (a pass reads only the other
// colour's rows and writes only its own).  Planes and seeds go from one
// pair of buffers to another, so no workgroup reads what another writes
// during a launch.
constexpr int kPairBand = 16;                            // rows a workgroup owns
constexpr int kPairHalo = 4;                             // rows of input either side
constexpr int kPairRows = kPairBand + 2 * kPairHalo;     // plane rows in LDS
constexpr int kPairMaxPw = 64;                           // words per row, at most
constexpr int kPairMinPw = 33;                           // and at least
constexpr int kPairStage = 4 * kPairMaxPw;               // threads per group (8 sites each)
constexpr int kPairGroups = 4;
constexpr int kPairThreads = kPairGroups * kPairStage;   // the workgroup
constexpr int kPairSlots0 = (kPairBand + 6 + kPairGroups - 1) / kPairGroups;   // colour-0 rows per thread
constexpr int kPairSlots1 = (kPairBand + 4 + kPairGroups - 1) / kPairGroups;   // colour-1 rows per thread
constexpr int kPairMinRows = 64;                         // rows (hy) a lattice needs for pairs

struct PairLds {
  uint32_t planes[2][kPairRows][kPairMaxPw];             // colour, row - (y0 - kPairHalo), word
  Words<4> parked[2][kPairThreads];                      // one slot of seeds per thread, through pass 1
};

// sq/4n with 33..64 plane words per row, at least kPairMinRows rows, and
// shadow seeds (2 colours x hy*4*pw threads x 8 sites, 4 bytes each) below
// 2^31 bytes, so a seed's byte offset fits 31 bits
__host__ __device__ inline bool pair_fits(const Geom &g) {
  const int pw = (g.hx + 31) / 32;
  return g.lat == SPK_HIP_SQ_4N && pw >= kPairMinPw && pw <= kPairMaxPw && g.hy >= kPairMinRows &&
         (int64_t)g.hy * pw * 4 * 8 * 2 * 4 < (int64_t(1) << 31);
}

__host__ __device__ __forceinline__ int pair_row(int rho, int hy) {
  const int r = rho % hy;
  return r < 0 ? r + hy : r;
}

// Row of slot j of group g for colour c, and whether pass p works on it.
__host__ __device__ __forceinline__ int pair_slot_row(int y0, int c, int g, int j) {
  return y0 - 3 + c + g + kPairGroups * j;
}
__host__ __device__ __forceinline__ bool pair_pass_row(int y0, int p, int rho) {
  return rho >= y0 - 3 + p && rho < y0 + kPairBand + 3 - p;
}

// Global offset of local sites 4j..4j+3 of group thread tau on physical row
// prow of colour c: bit_seed_index of the spt-8 layout (thread t =
// prow*4*pw + tau) in 32-bit arithmetic, which pair_fits allows.
__host__ __device__ __forceinline__ uint32_t pair_seed_index(const BitGeom &bg, int c, int prow, uint32_t tau, int j) {
  const uint32_t t = (uint32_t)prow * (uint32_t)(4 * bg.pw) + tau;
  return (uint32_t)c * (uint32_t)bg.seedc + ((((t >> 6) * 2u + (uint32_t)j) * 64u + (t & 63u)) * 4u);
}

// Those 4 seeds in the shadow at base.  The byte offset is 32-bit (below
// 2^31 by pair_fits), so the shadow's address stays scalar.  FULL (pw =
// kPairMaxPw): a row's 4*pw threads are whole 64-thread blocks of the
// layout, so the row's part of the offset (c, prow: the same on every
// lane) goes into that scalar address and the thread's is the same on
// every row.
template <bool FULL, typename S>
__host__ __device__ __forceinline__ auto pair_seeds(S *base, const BitGeom &bg, int c, int prow, uint32_t tau, int j) {
  using B = std::conditional_t<std::is_const_v<S>, const char, char>;
  using W = std::conditional_t<std::is_const_v<S>, const Words<4>, Words<4>>;
  B *b = reinterpret_cast<B *>(base);
  if constexpr (FULL) {
    static_assert(4 * kPairMaxPw % 64 == 0, "rows of whole blocks");
    b += ((int64_t)c * bg.seedc + (int64_t)prow * (4 * kPairMaxPw * 8)) * 4;
    return reinterpret_cast<W *>(b + ((tau >> 6) * 2048u + (uint32_t)j * 1024u + (tau & 63u) * 16u));
  } else {
    return reinterpret_cast<W *>(b + pair_seed_index(bg, c, prow, tau, j) * 4u);
  }
}

// Plane word i (0 .. 2*kPairRows*pw - 1) of the band: its global offset, and
// its place in LDS.
__host__ __device__ __forceinline__ int64_t pair_plane_src(const Geom &g, const BitGeom &bg, int y0, int i, int &c,
                                                           int &lr, int &w) {
  const int per = kPairRows * bg.pw;
  c = i >= per;
  const int r = i - (c ? per : 0);
  lr = r / bg.pw;
  w = r - lr * bg.pw;
  return ((int64_t)c * g.hy + pair_row(y0 - kPairHalo + lr, g.hy)) * bg.pw + w;
}

// Plane word i (0 .. 2*kPairBand*pw - 1) of the band's own rows out of LDS.
__host__ __device__ __forceinline__ void pair_store_plane(const PairLds &L, uint32_t *__restrict__ bits_out,
                                                          const Geom &g, const BitGeom &bg, int y0, int i) {
  const int per = kPairBand * bg.pw;
  const int c = i >= per;
  const int r = i - (c ? per : 0);
  const int rr = r / bg.pw, w = r - rr * bg.pw;
  if (y0 + rr < g.hy) bits_out[((int64_t)c * g.hy + y0 + rr) * bg.pw + w] = L.planes[c][kPairHalo + rr][w];
}

// Row rho's x-neighbour side for colour c: 1 when a site's x-neighbours
// sit at a and a+1, 0 for a-1 and a.  FULL rows (the form needs hy even):
// rho's row has rho's parity and y0 is even, so for group grp's rows
// (y0-3+c+grp+4j) the side is (1 + grp) & 1 in every pass.
template <bool FULL>
__host__ __device__ __forceinline__ int pair_side(const Geom &g, int rho, int c, int grp) {
  if constexpr (FULL) return (1 + grp) & 1;
  return (pair_row(rho, g.hy) + c) & 1;
}

// LDS byte offsets of thread tau's words at a row of colour c of group
// grp: its word w (own, above, below, beside), the x-neighbour word, and
// its byte of its own word (w*4 + q = tau), all from the start of plane row
// grp, so that what remains of a row's offset (its LDS row less grp: 1 + c
// + 4j for slot j) is the same on every lane.  FULL: the same for every row
// of the group (pair_side), so computed once.
struct PairLane {
  uint32_t w4, wn4, b;
};

template <bool FULL>
__host__ __device__ __forceinline__ PairLane pair_lane(const Geom &g, const BitGeom &bg, int rho, int c, int grp,
                                                       uint32_t tau) {
  const int w = (int)(tau >> 2), last = bg.pw - 1;
  const int px = pair_side<FULL>(g, rho, c, grp);
  const int wn = FULL ? (w + (px ? 1 : kPairMaxPw - 1)) & (kPairMaxPw - 1)
                      : px ? (w == last ? 0 : w + 1) : (w == 0 ? last : w - 1);
  const uint32_t rowg = (uint32_t)grp * (kPairMaxPw * 4);
  return PairLane{rowg + (uint32_t)w * 4u, rowg + (uint32_t)wn * 4u, rowg + tau};
}

// Byte `at` of the plane rows of colour c from LDS row lr0.
__host__ __device__ __forceinline__ uint8_t *pair_lds(PairLds &L, int c, int lr0, uint32_t at) {
  return reinterpret_cast<uint8_t *>(&L.planes[c][lr0][0]) + at;
}
__host__ __device__ __forceinline__ uint32_t pair_lds_word(const PairLds &L, int c, int lr0, uint32_t at) {
  return *reinterpret_cast<const uint32_t *>(reinterpret_cast<const uint8_t *>(&L.planes[c][lr0][0]) + at);
}

// The five plane words pass p's thread reads at row rho of group grp: its
// own colour's word, and the other colour's words above, below, beside and
// next along (bits_events).  rho's LDS row must lie in 1..kPairRows-2.
struct PairWords {
  uint32_t own, up, down, x, xe;
};

__host__ __device__ __forceinline__ PairWords pair_words(const PairLds &L, int y0, int p, int rho, int grp,
                                                         const PairLane &ln) {
  const int c = p & 1, o = c ^ 1;
  const int lr0 = rho - (y0 - kPairHalo) - grp;
  return PairWords{pair_lds_word(L, c, lr0, ln.w4), pair_lds_word(L, o, lr0 - 1, ln.w4),
                   pair_lds_word(L, o, lr0 + 1, ln.w4), pair_lds_word(L, o, lr0, ln.w4),
                   pair_lds_word(L, o, lr0, ln.wn4)};
}

// A thread's changes over one sweep's two passes, on the band's own rows:
// n, and the sums over them of s0, s1 and s2 (ei = s0 + 2*s1 + 4*s2) in
// 10-bit fields of e.  A thread owns at most 4 rows a pass, 64 sites a
// sweep, so no field overflows.
struct PairCounts {
  uint32_t n = 0, e = 0;
};

__host__ __device__ __forceinline__ Counts pair_counts(const PairCounts &a) {
  const int e0 = (int)(a.e & 1023u), e1 = (int)((a.e >> 10) & 1023u), e2 = (int)(a.e >> 20);
  return Counts{(int)a.n, 4 * (int)a.n - 2 * (e0 + 2 * e1 + 4 * e2)};   // dE = 4 - 2*ei a change
}

// Pass p's work at row rho for thread tau (< 4*pw), given its plane words
// (pair_words) and its seeds in sq: the 8 site events (bits_sites), the
// changed bits into the LDS planes, the new seeds back into sq and, after
// the colour's last pass (p >= 2), for the band's own rows, to seed_out.
// The changes of the band's own rows go into acc.  FULL: pw = kPairMaxPw,
// with 32 sites in the last word as in every other, and hy even.
template <bool SINGLE, bool FULL>
__host__ __device__ __forceinline__ void pair_thread(PairLds &L, const Geom &g, const BitGeom &bg, const Rule &r,
                                                     int y0, int p, int rho, int grp, uint32_t tau, const PairLane &ln,
                                                     const PairWords &pwd, Words<4> (&sq)[2],
                                                     int32_t *__restrict__ seed_out, PairCounts &acc) {
  const int c = p & 1;
  const int lr0 = rho - (y0 - kPairHalo) - grp;
  const int prow = pair_row(rho, g.hy);
  const int w = (int)(tau >> 2), q = (int)(tau & 3);
  const int px = pair_side<FULL>(g, rho, c, grp);
  const bool owned = rho >= y0 && rho < y0 + kPairBand && rho < g.hy;
  const BitsEi ei = bits_ei(pwd.own, pwd.up, pwd.down, pwd.x, bits_xshift<FULL>(pwd.x, pwd.xe, w, px, bg));
  const int i0 = 8 * q;
  const uint32_t own_q = pwd.own >> i0, upw_q = ei.upw >> i0, s0_q = ei.s0 >> i0, s1_q = ei.s1 >> i0;
  uint32_t ch_q = bits_sites<8, SINGLE>(own_q, upw_q, s0_q, r, sq, [&](int j, const Words<4> &nq) {
    sq[j] = nq;   // sq[j] is not read after this
    if (p >= 2 && owned) *pair_seeds<FULL>(seed_out, bg, c, prow, tau, j) = nq;
  });
  if (!FULL && w == bg.pw - 1) ch_q &= bg.lastmask >> i0;
  // the thread's byte of its own word; the word's other threads write the
  // other bytes and use only their own bits of what they read
  *pair_lds(L, c, lr0, ln.b) = (uint8_t)(own_q ^ ch_q);
  if (owned) {   // s2 = ~upw & ~s1 (ei 4), so s2 need not stay live
    acc.n += (uint32_t)__builtin_popcount(ch_q);
    acc.e += (uint32_t)__builtin_popcount(ch_q & s0_q) + ((uint32_t)__builtin_popcount(ch_q & s1_q) << 10) +
             ((uint32_t)__builtin_popcount(bop3<kAndNot2>(ch_q, upw_q, s1_q)) << 20);
  }
}

}  // namespace spk
