// spk_emul.hip -- host-only emulation of libspk_hip.so's colour pass, for CPU
// tests: runs the very gather<> / site_event<> / blocked_index code the
// kernels use, one site at a time, under AddressSanitizer.  Test
// infrastructure; not part of the library.
//
// usage: spk_emul app lat nx ny nz nspins kT nsweeps in.bin out.bin
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
  g.lat = lat; g.nx = nx; g.ny = ny; g.nz = nz;
  switch (lat) {
    case SPK_HIP_SQ_4N: g.K = 2; g.hx = nx / 2; g.hy = ny; g.hz = 1; break;
    case SPK_HIP_SC_6N: g.K = 2; g.hx = nx / 2; g.hy = ny; g.hz = nz; break;
    case SPK_HIP_SQ_8N: g.K = 4; g.hx = nx / 2; g.hy = ny / 2; g.hz = 1; break;
    default: g.K = 8; g.hx = nx / 2; g.hy = ny / 2; g.hz = nz / 2; break;
  }
  g.M = nx * ny * nz / g.K;
  return g;
}

extern "C" int spk_hip_metropolis_table(double, int, double *, int32_t *);

template <int LAT>
static void pass(std::vector<int32_t> &spin, std::vector<int32_t> &seed, const Geom &g,
                 const Rule &r, int color, long long &nacc, long long &de) {
  constexpr int NN = NeighCount<LAT>::value;
  for (int t = 0; t < g.M; t++) {
    int a = t % g.hx, rest = t / g.hx, b = rest % g.hy, e = rest / g.hy;
    int32_t nb[NN];
    gather<LAT>(spin.data(), g, color, a, b, e, nb);
    long long self = (long long)color * g.M + t;
    SiteOutcome o = site_event<NN>(r, r.thr, spin.at(self), seed.at(self), nb);
    seed.at(self) = o.seed;
    if (o.changed) spin.at(self) = o.spin;
    nacc += o.changed;
    de += o.dE;
  }
}

template <int LAT>
static long long energy(const std::vector<int32_t> &spin, const Geom &g) {
  constexpr int NN = NeighCount<LAT>::value;
  long long e = 0;
  for (int c = 0; c < g.K; c++)
    for (int t = 0; t < g.M; t++) {
      int a = t % g.hx, rest = t / g.hx, b = rest % g.hy, z = rest / g.hy;
      int32_t nb[NN];
      gather<LAT>(spin.data(), g, c, a, b, z, nb);
      for (int m = 0; m < NN; m++) e += nb[m] != spin.at((long long)c * g.M + t);
    }
  return e;
}

template <int LAT>
static int run(int app, int nx, int ny, int nz, int q, double kt, int nsw, const char *in, const char *out) {
  Geom g = geom(LAT, nx, ny, nz);
  const long long N = (long long)nx * ny * nz;
  Rule r{};
  r.app = app; r.nspins = app == SPK_HIP_POTTS ? q : 2; r.zero_t = kt == 0.0;
  int32_t thr[27];
  spk_hip_metropolis_table(kt > 0 ? kt : 1.0, maxneigh_of(LAT), nullptr, thr);
  r.half = thr[0];
  for (int d = 1; d <= maxneigh_of(LAT); d++) r.thr[d] = thr[d];
  std::vector<int32_t> sn(N), kn(N), s(N), k(N);
  FILE *f = fopen(in, "rb");
  if (!f || fread(sn.data(), 4, N, f) != (size_t)N || fread(kn.data(), 4, N, f) != (size_t)N) return 2;
  fclose(f);
  // natural -> blocked; check it is a bijection onto [0, N)
  std::vector<char> hit(N, 0);
  for (long long t = 0; t < N; t++) {
    int i = t % nx, j = (t / nx) % ny, kk = t / ((long long)nx * ny);
    long long p = blocked_index(g, i, j, kk);
    if (p < 0 || p >= N || hit[p]) { fprintf(stderr, "blocked_index not a bijection\n"); return 3; }
    hit[p] = 1;
    s[p] = sn[t];
    k[p] = kn[t];
  }
  std::vector<long long> nacc(nsw), eng(nsw);
  long long na = 0, e = energy<LAT>(s, g);
  for (int w = 0; w < nsw; w++) {
    long long de = 0;
    for (int c = 0; c < g.K; c++) pass<LAT>(s, k, g, r, c, na, de);
    e += 2 * de;
    if (e != energy<LAT>(s, g)) { fprintf(stderr, "incremental energy drifted\n"); return 4; }
    nacc[w] = na;
    eng[w] = e;
  }
  for (long long t = 0; t < N; t++) {
    int i = t % nx, j = (t / nx) % ny, kk = t / ((long long)nx * ny);
    long long p = blocked_index(g, i, j, kk);
    sn[t] = s[p];
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

int main(int argc, char **argv) {
  if (argc != 11) { fprintf(stderr, "usage\n"); return 1; }
  int app = atoi(argv[1]), lat = atoi(argv[2]), nx = atoi(argv[3]), ny = atoi(argv[4]), nz = atoi(argv[5]);
  int q = atoi(argv[6]);
  double kt = atof(argv[7]);
  int nsw = atoi(argv[8]);
  switch (lat) {
    case SPK_HIP_SQ_4N: return run<SPK_HIP_SQ_4N>(app, nx, ny, nz, q, kt, nsw, argv[9], argv[10]);
    case SPK_HIP_SQ_8N: return run<SPK_HIP_SQ_8N>(app, nx, ny, nz, q, kt, nsw, argv[9], argv[10]);
    case SPK_HIP_SC_6N: return run<SPK_HIP_SC_6N>(app, nx, ny, nz, q, kt, nsw, argv[9], argv[10]);
    default: return run<SPK_HIP_SC_26N>(app, nx, ny, nz, q, kt, nsw, argv[9], argv[10]);
  }
}
