// spk_emul.hip -- host-only emulation of libspk_hip.so's colour pass, for CPU
// tests: runs the very sweep_sites<> / gather<> / blocked_index code the
// kernels use, one thread's work at a time, under AddressSanitizer.  Test
// infrastructure; not part of the library.
//
// usage: spk_emul app lat nx ny nz nspins kT nsweeps stype vec in.bin out.bin
//   stype: 0 uint8, 1 uint16, 2 int32 spin storage; vec: sites per thread (4)
//   SPK_EMUL_BITS=8|16|32 in the environment runs two-state sq/4n lattices
//   on the bit-plane path (spk_bits.h), that many sites per thread;
//   SPK_EMUL_BITS=pair runs them two sweeps per launch (k_sweep_bits_pair's
//   band passes), the sweeps after the last pair on the 8-sites-per-thread
//   pass.
//   in.bin : int32 spin[N], int32 seed[N] (reference order)
//   out.bin: int32 spin[N], int32 seed[N], int64 naccept[nsweeps],
//            int64 energy[nsweeps]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <algorithm>
#include <type_traits>
#include <utility>
#include <vector>

#include "spk_bits.h"
#include "spk_device.h"

using namespace spk;

static long long S_slots(const Geom &g) { return (long long)g.K * g.Mp; }

static Geom geom(int lat, int nx, int ny, int nz) {
  Geom g{};
  geom_init(g, lat, nx, ny, nz);
  return g;
}

extern "C" int spk_hip_metropolis_table(double, int, double *, int32_t *);

template <int LAT, typename T>
static long long energy(const std::vector<T> &spin, const Geom &g) {
  constexpr int NN = NeighCount<LAT>::value;
  long long e = 0;
  for (int c = 0; c < g.K; c++)
    for (int t = 0; t < g.M; t++) {
      int a = t % g.hx, rest = t / g.hx;
      int32_t nb[NN];
      gather<LAT, T>(spin.data(), g, c, a, rest % g.hy, rest / g.hy, nb);
      for (int m = 0; m < NN; m++) e += nb[m] != (int32_t)spin.at((long long)c * g.Mp + a + (long long)g.px * rest);
    }
  return e;
}

template <int LAT, typename T, int V>
static int run(int app, int nx, int ny, int nz, int q, double kt, int nsw, const char *in, const char *out) {
  Geom g = geom(LAT, nx, ny, nz);
  if (g.px % V) { fprintf(stderr, "px %% V != 0\n"); return 5; }
  const long long N = (long long)nx * ny * nz;
  const long long S = (long long)g.K * g.Mp;   // stored slots, row pads included
  Rule r{};
  r.app = app; r.nspins = (app == SPK_HIP_POTTS || app == SPK_HIP_POTTS_NEIGH || app == SPK_HIP_POTTS_NEIGHONLY) ? q : 2;
  r.zero_t = kt == 0.0;
  int32_t thr[27];
  spk_hip_metropolis_table(kt > 0 ? kt : 1.0, maxneigh_of(LAT), nullptr, thr);
  r.half = thr[0];
  for (int d = 1; d <= maxneigh_of(LAT); d++) r.thr[d] = thr[d];
  std::vector<int32_t> sn(N), kn(N);
  // exact-size heap buffers so ASan sees any out-of-range pack or segment;
  // pad slots get poison values that would show if a site ever read one
  std::vector<T> s(S, (T)0);
  std::vector<int32_t> k(S, 0x7ffffff0);
  FILE *f = fopen(in, "rb");
  if (!f || fread(sn.data(), 4, N, f) != (size_t)N || fread(kn.data(), 4, N, f) != (size_t)N) return 2;
  fclose(f);
  std::vector<char> hit(S, 0);
  for (long long t = 0; t < N; t++) {
    int i = t % nx, j = (t / nx) % ny, kk = t / ((long long)nx * ny);
    long long p = blocked_index(g, i, j, kk);
    if (p < 0 || p >= S || hit[p]) { fprintf(stderr, "blocked_index not injective\n"); return 3; }
    hit[p] = 1;
    s[p] = (T)sn[t];
    k[p] = kn[t];
  }
  // the per-colour kernel's workgroup tiles cover every pack of a colour
  // exactly once, for every workgroup size
  for (int block = 64; block <= 1024; block *= 2) {
    const TileMap m = make_tile_map(g, block);
    std::vector<int> seen(g.Mp / 4, 0);
    for (int tile = 0; tile < m.nxc * m.nyg * m.nzg; tile++)
      for (int tid = 0; tid < block; tid++) {
        int a0, b, e;
        if (!tile_site(g, m, (uint32_t)tile, (uint32_t)tid, a0, b, e)) continue;
        if (a0 % 4 || a0 >= g.px || b >= g.hy || e >= g.hz) { fprintf(stderr, "tile site out of range\n"); return 11; }
        seen[(a0 + (long long)g.px * (b + (long long)g.hy * e)) / 4]++;
      }
    for (int v : seen)
      if (v != 1) { fprintf(stderr, "tiles do not cover each pack once (block %d)\n", block); return 12; }
  }
  std::vector<long long> nacc(nsw), eng(nsw);
  long long na = 0, e = energy<LAT, T>(s, g);
  int w = 0;
  const char *bz = getenv("SPK_EMUL_BITS");
  if (bz && LAT == SPK_HIP_SQ_4N && r.nspins == 2 && sizeof(T) == 1 &&
      (app < SPK_HIP_POTTS_NEIGH || app == SPK_HIP_ISING_SINGLE)) {
    // bit-plane path (spk_bits.h): pack, every pass thread by thread, unpack
    const bool pairs = std::string(bz) == "pair";
    const int spt = pairs ? 8 : atoi(bz);
    if (spt != 8 && spt != 16 && spt != 32) { fprintf(stderr, "SPK_EMUL_BITS must be 8, 16, 32 or pair\n"); return 16; }
    if (pairs && !pair_fits(g)) { fprintf(stderr, "lattice does not fit two sweeps per launch\n"); return 20; }
    if (r.half != (1 << 30)) { fprintf(stderr, "proposal threshold is not 2^30\n"); return 24; }
    const BitGeom bg = bit_geom(g, spt);
    const int64_t words = (int64_t)g.hy * bg.pw;
    // every real site has exactly one shadow seed slot
    std::vector<int> home_hits(S_slots(g), 0);
    for (int64_t sl = 0; sl < 2 * bg.seedc; sl++) {
      int c, i;
      uint32_t t;
      bit_seed_slot(bg, sl, c, t, i);
      if (bit_seed_index(bg, c, t, i) != sl) { fprintf(stderr, "bit seed slot does not invert\n"); return 13; }
      const int64_t h = bit_seed_home(g, bg, c, t, i);
      if (h >= (int64_t)S_slots(g)) { fprintf(stderr, "bit seed home out of range\n"); return 14; }
      if (h >= 0) home_hits[h]++;
    }
    for (int c = 0; c < 2; c++)
      for (int b = 0; b < g.hy; b++)
        for (int a = 0; a < g.px; a++)
          if (home_hits[(int64_t)c * g.Mp + (int64_t)b * g.px + a] != (a < g.hx ? 1 : 0)) {
            fprintf(stderr, "bit seed slots do not cover each site once\n");
            return 15;
          }
    // the conversion kernels' 16-byte groups: each shadow group once, at
    // the slot bit_seed_index gives its first site
    {
      std::vector<int> unit_hits(bg.seedc / 2, 0);
      for (int64_t X = 0; X < bg.seedc / 2; X++) {
        int c, j;
        uint32_t t;
        int64_t unit;
        bit_seed_group(bg, X, c, t, j, unit);
        if (unit < 0 || unit >= bg.seedc / 2 || bit_seed_index(bg, c, t, 4 * j) != 4 * unit) {
          fprintf(stderr, "bit seed group does not match the slot order\n");
          return 17;
        }
        unit_hits[unit]++;
      }
      for (int v : unit_hits)
        if (v != 1) { fprintf(stderr, "bit seed groups do not cover each shadow group once\n"); return 18; }
    }
    std::vector<uint32_t> bits(2 * words);
    std::vector<int32_t> sseed(2 * bg.seedc, 0x7ffffff0);
    for (int64_t q = 0; q < 2 * words; q++) {
      const int c = q >= words;
      const int64_t t = q - (c ? words : 0);
      bits[q] = bits_pack_word(reinterpret_cast<const uint8_t *>(s.data()), g, c, (int)(t / bg.pw), (int)(t % bg.pw));
    }
    for (int64_t X = 0; X < bg.seedc / 2; X++) bits_seed_group_in(k.data(), g, bg, X, sseed.data());
    for (int64_t sl = 0; sl < 2 * bg.seedc; sl++) {   // the groups agree with the slot-by-slot map
      int c, i;
      uint32_t t;
      bit_seed_slot(bg, sl, c, t, i);
      const int64_t h = bit_seed_home(g, bg, c, t, i);
      if (sseed[sl] != (h < 0 ? 1 : k[h])) { fprintf(stderr, "shadow seed group mismatch\n"); return 19; }
    }
    const bool single = app == SPK_HIP_ISING_SINGLE;
    // two sweeps per launch: every band, pass by pass (the barrier), each
    // group's threads in turn (a pass's threads touch disjoint words), the
    // seeds of each thread's slots held in its "registers" from the band's
    // start, from (bits, sseed) into (bits2, sseed2), as k_sweep_bits_pair
    std::vector<uint32_t> bits2(pairs ? 2 * words : 0, 0xdeadbeefu);
    std::vector<int32_t> sseed2(pairs ? 2 * bg.seedc : 0, 0x7ffffff1);
    PairLds *L = pairs ? new PairLds : nullptr;
    // as launch_pair picks k_sweep_bits_pair<., true>
    const bool full = bg.pw == kPairMaxPw && bg.r == 31 && g.hy % 2 == 0;
    std::vector<Words<4>> regs(pairs ? (size_t)kPairThreads * (kPairSlots0 + kPairSlots1) * 2 : 0);
    while (pairs && w + 2 <= nsw) {
      long long de[2] = {0, 0}, nc[2] = {0, 0};
      for (int y0 = 0; y0 < g.hy; y0 += kPairBand) {
        memset(L, 0xa5, sizeof(PairLds));   // stale LDS must not matter
        for (int i = 0; i < 2 * kPairRows * bg.pw; i++) {
          int c, lr, wd;
          const int64_t at = pair_plane_src(g, bg, y0, i, c, lr, wd);
          if (at < 0 || at >= 2 * words || lr < 0 || lr >= kPairRows || wd >= bg.pw) {
            fprintf(stderr, "pair plane index out of range\n");
            return 21;
          }
          L->planes[c][lr][wd] = bits[at];
        }
        // the loads every thread issues at the band's start
        auto reg = [&](int c, int grp, uint32_t tau, int j) -> Words<4> * {
          const size_t k = ((size_t)(grp * kPairStage + tau) * (kPairSlots0 + kPairSlots1) + (c ? kPairSlots0 : 0) + j) * 2;
          return &regs[k];
        };
        for (int grp = 0; grp < kPairGroups; grp++)
          for (uint32_t tau = 0; tau < (uint32_t)kPairStage; tau++)
            for (int c = 0; c < 2; c++)
              for (int j = 0; j < (c ? kPairSlots1 : kPairSlots0); j++) {
                const int rho = std::min(pair_slot_row(y0, c, grp, j), y0 + kPairBand + 2 - c);
                const uint32_t ta = tau < (uint32_t)(4 * bg.pw) ? tau : 0u;
                for (int h = 0; h < 2; h++) {
                  const int64_t at = pair_seed_index(bg, c, pair_row(rho, g.hy), ta, h);
                  if (at + 4 > 2 * bg.seedc) { fprintf(stderr, "pair seed index out of range\n"); return 22; }
                  // the kernel's address of those seeds (a scalar row part for full rows)
                  const int64_t kat = reinterpret_cast<const int32_t *>(
                                          full ? pair_seeds<true>(sseed.data(), bg, c, pair_row(rho, g.hy), ta, h)
                                               : pair_seeds<false>(sseed.data(), bg, c, pair_row(rho, g.hy), ta, h)) -
                                      sseed.data();
                  if (kat != at) { fprintf(stderr, "pair seed address differs from the layout's index\n"); return 24; }
                  memcpy(&reg(c, grp, tau, j)[h], &sseed[at], sizeof(Words<4>));
                }
              }
        for (int p = 0; p < 4; p++) {
          const int c = p & 1;
          const int slots = c ? kPairSlots1 : (p == 2 ? kPairSlots0 - 1 : kPairSlots0);
          for (int grp = 0; grp < kPairGroups; grp++)
            for (uint32_t tau = 0; tau < (uint32_t)(4 * bg.pw); tau++)
              for (int j = 0; j < slots; j++) {
                const int rho = pair_slot_row(y0, c, grp, j);
                if (!pair_pass_row(y0, p, rho)) continue;
                const int lr = rho - (y0 - kPairHalo);
                if (lr < 1 || lr > kPairRows - 2) { fprintf(stderr, "pair row outside its LDS rows\n"); return 23; }
                Words<4>(&sq)[2] = *reinterpret_cast<Words<4>(*)[2]>(reg(c, grp, tau, j));
                PairCounts a;
                auto run = [&](auto single_c, auto full_c) {
                  constexpr bool SG = decltype(single_c)::value, FL = decltype(full_c)::value;
                  if (FL && pair_side<true>(g, rho, c, grp) != pair_side<false>(g, rho, c, grp)) {
                    fprintf(stderr, "full-row x side differs from the row's\n");
                    exit(25);
                  }
                  const PairLane ln = pair_lane<FL>(g, bg, rho, c, grp, tau);
                  const PairWords pwd = pair_words(*L, y0, p, rho, grp, ln);
                  pair_thread<SG, FL>(*L, g, bg, r, y0, p, rho, grp, tau, ln, pwd, sq, sseed2.data(), a);
                };
                using T_ = std::true_type;
                using F_ = std::false_type;
                if (single) full ? run(T_{}, T_{}) : run(T_{}, F_{});
                else full ? run(F_{}, T_{}) : run(F_{}, F_{});
                const Counts cnt = pair_counts(a);
                nc[p >> 1] += cnt.nacc;
                de[p >> 1] += cnt.de;
              }
        }
        for (int i = 0; i < 2 * kPairBand * bg.pw; i++) pair_store_plane(*L, bits2.data(), g, bg, y0, i);
      }
      std::swap(bits, bits2);
      std::swap(sseed, sseed2);
      for (int h = 0; h < 2; h++, w++) {
        na += nc[h];
        e += 2 * de[h];
        nacc[w] = na;
        eng[w] = e;
      }
      for (int64_t q = 0; q < 2 * words; q++) {
        const int c = q >= words;
        const int64_t t = q - (c ? words : 0);
        bits_unpack_word(reinterpret_cast<uint8_t *>(s.data()), g, c, (int)(t / bg.pw), (int)(t % bg.pw), bits[q]);
      }
      if (e != energy<LAT, T>(s, g)) { fprintf(stderr, "incremental energy drifted (two sweeps per launch)\n"); return 4; }
    }
    delete L;
    if (pairs) {   // the seeds of the pairs' result, for the sweeps left and the output
      for (int64_t X = 0; X < bg.seedc / 2; X++) bits_seed_group_out(sseed.data(), g, bg, X, k.data());
    }
    for (; w < nsw; w++) {
      long long de = 0;
      for (int c = 0; c < 2; c++)
        for (uint32_t t = 0; t < (uint32_t)bg.threads; t++) {
          const Counts cnt = spt == 8    ? (single ? bits_thread<8, true>(bits.data(), sseed.data(), g, bg, r, c, t)
                                               : bits_thread<8>(bits.data(), sseed.data(), g, bg, r, c, t))
                             : spt == 16 ? (single ? bits_thread<16, true>(bits.data(), sseed.data(), g, bg, r, c, t)
                                               : bits_thread<16>(bits.data(), sseed.data(), g, bg, r, c, t))
                                         : (single ? bits_thread<32, true>(bits.data(), sseed.data(), g, bg, r, c, t)
                                               : bits_thread<32>(bits.data(), sseed.data(), g, bg, r, c, t));
          na += cnt.nacc;
          de += cnt.de;
        }
      for (int64_t q = 0; q < 2 * words; q++) {
        const int c = q >= words;
        const int64_t t = q - (c ? words : 0);
        bits_unpack_word(reinterpret_cast<uint8_t *>(s.data()), g, c, (int)(t / bg.pw), (int)(t % bg.pw), bits[q]);
      }
      for (int64_t X = 0; X < bg.seedc / 2; X++) bits_seed_group_out(sseed.data(), g, bg, X, k.data());
      e += 2 * de;
      if (e != energy<LAT, T>(s, g)) { fprintf(stderr, "incremental energy drifted (bit planes)\n"); return 4; }
      nacc[w] = na;
      eng[w] = e;
    }
  }
  if (app >= SPK_HIP_POTTS_NEIGH) {
    // proposal variants: the ordered one-site-per-slot path
    for (; w < nsw; w++) {
      long long de = 0;
      for (int c = 0; c < g.K; c++)
        for (int t = 0; t < g.Mp; t++) {
          Counts cnt{0, 0};
          if (app == SPK_HIP_POTTS_NEIGH)
            cnt = variant_site<LAT, T, SPK_HIP_POTTS_NEIGH>(s.data(), k.data(), g, r, r.thr, c, (uint32_t)t);
          else if (app == SPK_HIP_POTTS_NEIGHONLY)
            cnt = variant_site<LAT, T, SPK_HIP_POTTS_NEIGHONLY>(s.data(), k.data(), g, r, r.thr, c, (uint32_t)t);
          else
            cnt = variant_site<LAT, T, SPK_HIP_ISING_SINGLE>(s.data(), k.data(), g, r, r.thr, c, (uint32_t)t);
          na += cnt.nacc;
          de += cnt.de;
        }
      e += 2 * de;
      if (e != energy<LAT, T>(s, g)) { fprintf(stderr, "incremental energy drifted (variant)\n"); return 4; }
      nacc[w] = na;
      eng[w] = e;
    }
  }
  for (; w < nsw; w++) {
    long long de = 0;
    for (int c = 0; c < g.K; c++)
      for (int t = 0; t < g.Mp / V; t++) {
        int a0, b, e;
        pack_site(g, (uint32_t)t, a0, b, e);
        // the fast division must agree with the plain one
        if (a0 != (t % (g.px / V)) * V || b != (t / (g.px / V)) % g.hy || e != (t / (g.px / V)) / g.hy) {
          fprintf(stderr, "pack_site disagrees with / and %%\n");
          return 10;
        }
        // the kernels' dispatch: the two-state shortcut for u8 storage;
        // 32-bit offsets on even sweeps, 64-bit on odd ones (both forms run)
        const bool two = r.nspins == 2 && sizeof(T) == 1;
        Counts cnt;
        if (w % 2 == 0)
          cnt = two ? sweep_sites<LAT, T, V, true, uint32_t>(s.data(), k.data(), g, r, r.thr, c, a0, b, e)
                    : sweep_sites<LAT, T, V, false, uint32_t>(s.data(), k.data(), g, r, r.thr, c, a0, b, e);
        else
          cnt = two ? sweep_sites<LAT, T, V, true, int64_t>(s.data(), k.data(), g, r, r.thr, c, a0, b, e)
                    : sweep_sites<LAT, T, V, false, int64_t>(s.data(), k.data(), g, r, r.thr, c, a0, b, e);
        na += cnt.nacc;
        de += cnt.de;
      }
    e += 2 * de;
    if (e != energy<LAT, T>(s, g)) { fprintf(stderr, "incremental energy drifted\n"); return 4; }
    nacc[w] = na;
    eng[w] = e;
  }
  for (long long t = 0; t < N; t++) {
    int i = t % nx, j = (t / nx) % ny, kk = t / ((long long)nx * ny);
    long long p = blocked_index(g, i, j, kk);
    sn[t] = (int32_t)s[p];
    kn[t] = k[p];
  }
  f = fopen(out, "wb");
  fwrite(sn.data(), 4, N, f);
  fwrite(kn.data(), 4, N, f);
  fwrite(nacc.data(), 8, nsw, f);
  fwrite(eng.data(), 8, nsw, f);
  fclose(f);
  return 0;
}

template <int LAT>
static int by_type(int stype, int vec, int app, int nx, int ny, int nz, int q, double kt, int nsw, const char *in,
                   const char *out) {
  switch (stype * 10 + vec) {
    case 4: return run<LAT, uint8_t, 4>(app, nx, ny, nz, q, kt, nsw, in, out);
    case 14: return run<LAT, uint16_t, 4>(app, nx, ny, nz, q, kt, nsw, in, out);
    case 24: return run<LAT, int32_t, 4>(app, nx, ny, nz, q, kt, nsw, in, out);
  }
  return 6;
}

int main(int argc, char **argv) {
  if (argc != 13) { fprintf(stderr, "usage\n"); return 1; }
  int app = atoi(argv[1]), lat = atoi(argv[2]), nx = atoi(argv[3]), ny = atoi(argv[4]), nz = atoi(argv[5]);
  int q = atoi(argv[6]);
  double kt = atof(argv[7]);
  int nsw = atoi(argv[8]), stype = atoi(argv[9]), vec = atoi(argv[10]);
  const char *in = argv[11], *out = argv[12];
  switch (lat) {
    case SPK_HIP_SQ_4N: return by_type<SPK_HIP_SQ_4N>(stype, vec, app, nx, ny, nz, q, kt, nsw, in, out);
    case SPK_HIP_SQ_8N: return by_type<SPK_HIP_SQ_8N>(stype, vec, app, nx, ny, nz, q, kt, nsw, in, out);
    case SPK_HIP_SC_6N: return by_type<SPK_HIP_SC_6N>(stype, vec, app, nx, ny, nz, q, kt, nsw, in, out);
    default: return by_type<SPK_HIP_SC_26N>(stype, vec, app, nx, ny, nz, q, kt, nsw, in, out);
  }
}
