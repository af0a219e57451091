// spk_emul.hip -- host-only emulation of libspk_hip.so's colour pass, for CPU
// tests: runs the very sweep_sites<> / gather<> / blocked_index code the
// kernels use, one thread's work at a time, under AddressSanitizer.  Test
// infrastructure; not part of the library.
//
// usage: spk_emul app lat nx ny nz nspins kT nsweeps stype vec in.bin out.bin
//   stype: 0 uint8, 1 uint16, 2 int32 spin storage; vec: sites per thread (4)
//   in.bin : int32 spin[N], int32 seed[N] (reference order)
//   out.bin: int32 spin[N], int32 seed[N], int64 naccept[nsweeps],
//            int64 energy[nsweeps]
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "spk_device.h"

using namespace spk;

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
  r.app = app; r.nspins = app == SPK_HIP_POTTS ? q : 2; r.zero_t = kt == 0.0;
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
  std::vector<long long> nacc(nsw), eng(nsw);
  long long na = 0, e = energy<LAT, T>(s, g);
  const int per_row = g.px / V;
  for (int w = 0; w < nsw; w++) {
    long long de = 0;
    for (int c = 0; c < g.K; c++)
      for (int t = 0; t < g.Mp / V; t++) {
        int a0 = (t % per_row) * V, rest = t / per_row;
        Counts cnt = sweep_sites<LAT, T, V>(s.data(), k.data(), g, r, r.thr, c, a0, rest % g.hy, rest / g.hy);
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
