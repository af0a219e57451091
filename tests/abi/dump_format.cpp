// dump_format.cpp -- checks the dump writer's formatting (front/spk_dump_writer.cpp)
// against plain fprintf, the way DumpText::write_text formats a line
// (src/dump_text.cpp:405-421): integers as %d, coordinates and energies as
// %g, each value followed by a space.  Host only; test infrastructure.
//
// usage: dump_format out_dir   (exit 0 when every case matches)
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "spk_dump_writer.h"

using namespace spkfront;

static std::string slurp(const std::string &p) {
  std::string s;
  FILE *f = std::fopen(p.c_str(), "rb");
  if (!f) return s;
  char buf[65536];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) s.append(buf, n);
  std::fclose(f);
  return s;
}

// the same snapshot written one fprintf per value
static void expected(const DumpJob &d, FILE *fp) {
  std::fputs(d.header.c_str(), fp);
  for (int64_t g = 0; g < d.nx * d.ny * d.nz; g++) {
    const int64_t i = g % d.nx, j = (g / d.nx) % d.ny, k = g / (d.nx * d.ny);
    for (DumpField f : d.fields) {
      switch (f) {
        case DumpField::Id: std::fprintf(fp, "%lld ", (long long)(g + 1)); break;
        case DumpField::Site: std::fprintf(fp, "%d ", d.spin[g]); break;
        case DumpField::Seed: std::fprintf(fp, "%d ", d.seed[g]); break;
        case DumpField::X: std::fprintf(fp, "%g ", i * d.latconst + d.boxlo[0]); break;
        case DumpField::Y: std::fprintf(fp, "%g ", j * d.latconst + d.boxlo[1]); break;
        case DumpField::Z:
          std::fprintf(fp, "%g ", d.dimension == 3 ? k * d.latconst + d.boxlo[2] : 0.5 * (d.boxlo[2] + d.boxhi[2]));
          break;
        case DumpField::Energy: {
          int e = 0;
          for (int dk = -1; dk <= 1; dk++) {
            if (d.dimension == 2 && dk) continue;
            for (int dj = -1; dj <= 1; dj++)
              for (int di = -1; di <= 1; di++) {
                const int nb = std::abs(di) + std::abs(dj) + std::abs(dk);
                if (nb == 0 || (d.nearest_only && nb != 1)) continue;
                const int64_t a = (i + di + d.nx) % d.nx, b = (j + dj + d.ny) % d.ny, c = (k + dk + d.nz) % d.nz;
                e += d.spin[a + d.nx * (b + d.ny * c)] != d.spin[g];
              }
          }
          std::fprintf(fp, "%g ", (double)e);
          break;
        }
      }
    }
    std::fprintf(fp, "\n");
  }
}

static DumpJob make(int nx, int ny, int nz, int dim, bool nearest, double latconst, double lo, unsigned seed) {
  DumpJob d;
  d.nx = nx;
  d.ny = ny;
  d.nz = nz;
  d.dimension = dim;
  d.nearest_only = nearest;
  d.latconst = latconst;
  for (int m = 0; m < 3; m++) {
    d.boxlo[m] = lo * (m + 1);
    d.boxhi[m] = d.boxlo[m] + latconst * (m == 0 ? nx : m == 1 ? ny : nz);
  }
  d.header = "ITEM: TIMESTEP\n7 0.25\n";
  d.fields = {DumpField::Id, DumpField::Site, DumpField::Seed, DumpField::X,
              DumpField::Y,  DumpField::Z,    DumpField::Energy};
  const int64_t n = (int64_t)nx * ny * nz;
  std::srand(seed);
  for (int64_t g = 0; g < n; g++) {
    d.spin.push_back(1 + std::rand() % 3);
    d.seed.push_back(g % 5 == 0 ? 2147483646 : 1 + std::rand());
  }
  return d;
}

int main(int argc, char **argv) {
  if (argc != 2) return 2;
  const std::string dir = argv[1];
  struct Case {
    int nx, ny, nz, dim;
    bool nearest;
    double latconst, lo;
  } cases[] = {{16, 12, 1, 2, false, 1.0, 0.0}, {8, 6, 4, 3, true, 1.5, -0.5}, {6, 4, 10, 3, false, 0.37, 2.25},
               {20000, 3, 1, 2, true, 1.0, 0.0}};   // the last fills the 1 MiB buffer several times
  int bad = 0, k = 0;
  for (const Case &c : cases) {
    DumpJob d = make(c.nx, c.ny, c.nz, c.dim, c.nearest, c.latconst, c.lo, 100 + k);
    const std::string got = dir + "/got" + std::to_string(k), want = dir + "/want" + std::to_string(k);
    FILE *fw = std::fopen(want.c_str(), "w");
    expected(d, fw);
    std::fclose(fw);
    d.fp = std::fopen(got.c_str(), "w");
    d.close = true;
    if (!write_snapshot(d)) bad++;
    if (slurp(got) != slurp(want)) {
      std::fprintf(stderr, "case %d differs\n", k);
      bad++;
    }
    k++;
  }
  // through the writer thread, two snapshots into one file, in order
  {
    DumpJob a = make(12, 10, 1, 2, true, 1.0, 0.0, 7), b = make(12, 10, 1, 2, true, 1.0, 0.0, 8);
    const std::string got = dir + "/gotq", want = dir + "/wantq";
    FILE *fw = std::fopen(want.c_str(), "w");
    expected(a, fw);
    expected(b, fw);
    std::fclose(fw);
    FILE *fp = std::fopen(got.c_str(), "w");
    a.fp = b.fp = fp;
    DumpWriter w;
    w.submit(std::move(a));
    w.submit(std::move(b));
    if (!w.drain()) bad++;
    std::fclose(fp);
    if (slurp(got) != slurp(want)) {
      std::fprintf(stderr, "queued snapshots differ\n");
      bad++;
    }
  }
  std::printf("%s\n", bad ? "FAIL" : "ok");
  return bad ? 1 : 0;
}
