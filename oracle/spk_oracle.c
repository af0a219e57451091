/*
 * spk_oracle.c -- CPU restatement of SPPARKS' strict-mode colour sweep.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path (see spk_oracle.h).
 * Nothing in the product links or loads this file.
 *
 * Written from the reference's behaviour, each routine citing the file:line
 * it restates.  Where the reference's arithmetic fixes the result bit for bit
 * (Park-Miller step, double conversions, truncations, libm exp on uphill
 * moves) the restatement performs the same operations in the same order; it
 * deliberately does NOT use the tricks the GPU path uses (integer
 * thresholds, jump-ahead seeding, colour-blocked layout), so that agreement
 * between the two is evidence rather than tautology.
 *
 * Build: gcc -O2 -ffp-contract=off -fopenmp (no -ffast-math, no -march: the
 * reference is built with plain -O on x86-64, which has no FMA contraction).
 *
 * Parity unpinned: orc_color's numbering of the sq/4n, sc/6n and sc/26n
 * colours.  It restates Lattice::id2color (src/lattice.cpp:234-283) as
 * read; the partition invariants the tests check hold equally for a
 * relabelled numbering, and no reference-held vector covers these lattices
 * (only the sq/8n numbering is fixed, by the recorded 128x128 Potts run).
 */
#include "spk_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Park-Miller minimal standard constants (src/random_park.cpp:24-28) */
#define ORC_IA 16807
#define ORC_IM 2147483647
#define ORC_IQ 127773
#define ORC_IR 2836
static const double ORC_AM = 1.0 / ORC_IM;

/* ----------------------------------------------------------------------
   Park-Miller step in Schrage's overflow-free form, then the double
   conversion (src/random_park.cpp:79-86).
   ---------------------------------------------------------------------- */
double orc_pm_uniform(int *seed)
{
  int hi = *seed / ORC_IQ;
  int lo = *seed - hi * ORC_IQ;
  int next = ORC_IA * lo - ORC_IR * hi;
  if (next < 0) next += ORC_IM;
  *seed = next;
  return ORC_AM * next;
}

/* integer RN in 1..n (src/random_park.cpp:92-97) */
int orc_pm_irandom(int *seed, int n)
{
  int v = (int)(orc_pm_uniform(seed) * n) + 1;
  return v > n ? n : v;
}

/* RandomPark(double) constructor (src/random_park.cpp:45-49) */
int orc_pm_from_double(double rseed)
{
  int s = (int)(rseed * ORC_IM);
  return s == 0 ? 1 : s;
}

/* RandomPark::tagreset (src/random_park.cpp:67-73): fmod-based seed from
   (rseed, offset) followed by `warmup` discarded steps. */
int orc_pm_tagreset(double rseed, int64_t offset, int warmup)
{
  double x = rseed * ORC_IM;      /* separate statements: no contraction */
  x = x + (double)offset;
  int s = (int)fmod(x, (double)ORC_IM);
  if (s < 0) s = -s;
  if (s == 0) s = 1;
  for (int w = 0; w < warmup; w++) orc_pm_uniform(&s);
  return s;
}

/* ----------------------------------------------------------------------
   Marsaglia / Zaman "RANMAR" as used by SPPARKS' master RNG
   (src/random_mars.cpp:43-83 init, :89-105 uniform).
   ---------------------------------------------------------------------- */
void orc_ranmars_init(orc_ranmars *r, int seed)
{
  while (seed > 900000000) seed -= 900000000;
  int ij = (seed - 1) / 30082;
  int kl = (seed - 1) - 30082 * ij;
  int i = (ij / 177) % 177 + 2;
  int j = ij % 177 + 2;
  int k = (kl / 169) % 178 + 1;
  int l = kl % 169;
  for (int n = 1; n <= 97; n++) {
    double s = 0.0, t = 0.5;
    for (int b = 0; b < 24; b++) {
      int m = ((i * j) % 179) * k % 179;
      i = j; j = k; k = m;
      l = (53 * l + 1) % 169;
      if ((l * m) % 64 >= 32) s = s + t;
      t = 0.5 * t;
    }
    r->u[n] = s;
  }
  r->c = 362436.0 / 16777216.0;
  r->cd = 7654321.0 / 16777216.0;
  r->cm = 16777213.0 / 16777216.0;
  r->i97 = 97;
  r->j97 = 33;
  orc_ranmars_uniform(r);   /* init() discards one draw (random_mars.cpp:82) */
}

double orc_ranmars_uniform(orc_ranmars *r)
{
  double uni = r->u[r->i97] - r->u[r->j97];
  if (uni < 0.0) uni += 1.0;
  r->u[r->i97] = uni;
  if (--r->i97 == 0) r->i97 = 97;
  if (--r->j97 == 0) r->j97 = 97;
  r->c -= r->cd;
  if (r->c < 0.0) r->c += r->cm;
  uni -= r->c;
  if (uni < 0.0) uni += 1.0;
  return uni;
}

/* ----------------------------------------------------------------------
   Lattice geometry.
   ---------------------------------------------------------------------- */

/* maxneigh per style (src/create_sites.cpp:530-540) */
int orc_maxneigh(int lattice)
{
  switch (lattice) {
    case ORC_SQ_4N: return 4;
    case ORC_SQ_8N: return 8;
    case ORC_SC_6N: return 6;
    case ORC_SC_26N: return 26;
  }
  return 0;
}

static int orc_is3d(int lattice)
{
  return lattice == ORC_SC_6N || lattice == ORC_SC_26N;
}

/* Lattice::ncolors with delcolor = delevent+delpropensity = 1 for Potts and
   Ising (src/lattice.cpp:187-228): 0 when an edge is odd. */
int orc_ncolors(const orc_desc *d)
{
  int odd = (d->nx % 2) || (d->ny % 2) || (orc_is3d(d->lattice) && d->nz % 2);
  if (odd) return 0;
  switch (d->lattice) {
    case ORC_SQ_4N: return 2;
    case ORC_SQ_8N: return 4;
    case ORC_SC_6N: return 2;
    case ORC_SC_26N: return 8;
  }
  return 0;
}

/* Lattice::id2color (src/lattice.cpp:234-283), ids 1-based, x fastest */
int orc_color(const orc_desc *d, int64_t gid)
{
  int64_t g = gid - 1;
  int64_t i = g % d->nx;
  int64_t j = (g % ((int64_t)d->nx * d->ny)) / d->nx;
  int64_t k = g / ((int64_t)d->nx * d->ny);
  switch (d->lattice) {
    case ORC_SQ_4N: return (int)((i + j) % 2) + 1;
    case ORC_SQ_8N: return (int)(2 * (j % 2) + i % 2) + 1;
    case ORC_SC_6N: return (int)((i + j + k) % 2) + 1;
    case ORC_SC_26N: return (int)(4 * (k % 2) + 2 * (j % 2) + i % 2) + 1;
  }
  return 0;
}

/* the same colour from coordinates (no division) */
static inline int orc_color_ijk(int lattice, int64_t i, int64_t j, int64_t k)
{
  switch (lattice) {
    case ORC_SQ_4N: return (int)((i + j) & 1) + 1;
    case ORC_SQ_8N: return (int)(2 * (j & 1) + (i & 1)) + 1;
    case ORC_SC_6N: return (int)((i + j + k) & 1) + 1;
    default: return (int)(4 * (k & 1) + 2 * (j & 1) + (i & 1)) + 1;
  }
}

/* Neighbour stencil by the reference's distance criterion
   (src/create_sites.cpp:1303-1371): offsets in {-1,0,1}^dim whose length r
   satisfies cutlo-eps < r < cuthi+eps, cutlo = 1, cuthi = 1 (4n/6n),
   sqrt(2) (8n) or sqrt(3) (26n). */
int orc_stencil(int lattice, int *offs)
{
  const double eps = 1.0e-6;
  double cuthi;
  int is3 = orc_is3d(lattice);
  switch (lattice) {
    case ORC_SQ_4N: case ORC_SC_6N: cuthi = 1.0; break;
    case ORC_SQ_8N: cuthi = sqrt(2.0); break;
    case ORC_SC_26N: cuthi = sqrt(3.0); break;
    default: return 0;
  }
  int n = 0;
  for (int di = -1; di <= 1; di++)
    for (int dj = -1; dj <= 1; dj++)
      for (int dk = (is3 ? -1 : 0); dk <= (is3 ? 1 : 0); dk++) {
        double r = sqrt((double)(di * di + dj * dj + dk * dk));
        if (r > 1.0 - eps && r < cuthi + eps) {
          offs[3 * n + 0] = di;
          offs[3 * n + 1] = dj;
          offs[3 * n + 2] = dk;
          n++;
        }
      }
  return n;
}

/* periodic connectivity (src/create_sites.cpp:589-640) */
void orc_neighbors(const orc_desc *d, int *neigh)
{
  int offs[27 * 3];
  int ns = orc_stencil(d->lattice, offs);
  int64_t nx = d->nx, ny = d->ny, nz = d->nz;
  int64_t N = nx * ny * nz;
  for (int64_t g = 0; g < N; g++) {
    int64_t i = g % nx, j = (g / nx) % ny, k = g / (nx * ny);
    for (int m = 0; m < ns; m++) {
      int64_t a = i + offs[3 * m], b = j + offs[3 * m + 1], c = k + offs[3 * m + 2];
      if (a < 0) a += nx;
      if (a >= nx) a -= nx;
      if (b < 0) b += ny;
      if (b >= ny) b -= ny;
      if (c < 0) c += nz;
      if (c >= nz) c -= nz;
      neigh[g * ns + m] = (int)(a + nx * (b + ny * c));
    }
  }
}

/* ----------------------------------------------------------------------
   Initial state.
   ---------------------------------------------------------------------- */

/* Set::set_range with loop all: one generator, one irandom(range) per
   global id 1..N in order (src/set.cpp:598-627). */
void orc_set_range(int seed0, int lo, int hi, int64_t n, int *spin)
{
  int s = seed0;
  int range = hi - lo + 1;
  for (int64_t g = 0; g < n; g++) spin[g] = orc_pm_irandom(&s, range) - 1 + lo;
}

/* siteseeds[i] = tagreset(rseed, id[i], 100) (src/app_lattice.cpp:287-295) */
void orc_site_seeds(double rseed, int64_t n, int *seeds)
{
#pragma omp parallel for schedule(static)
  for (int64_t g = 0; g < n; g++) seeds[g] = orc_pm_tagreset(rseed, g + 1, 100);
}

/* seed -> set site range -> first run's init draws (src/input.cpp:942-951,
   src/set.cpp:598-603, src/app_lattice.cpp:274-295). */
int orc_script_init(const orc_desc *d, int master_seed, int lo, int hi,
                    int *spin, int *seeds)
{
  if (master_seed <= 0) return -1;
  int64_t n = (int64_t)d->nx * d->ny * d->nz;
  orc_ranmars m;
  orc_ranmars_init(&m, master_seed);
  int s0 = orc_pm_from_double(orc_ranmars_uniform(&m));  /* set: 1 draw */
  orc_set_range(s0, lo, hi, n, spin);
  orc_ranmars_uniform(&m);           /* ranapp constructor              */
  orc_ranmars_uniform(&m);           /* ranapp reset seed               */
  orc_ranmars_uniform(&m);           /* ranstrict constructor           */
  double rseed = orc_ranmars_uniform(&m);   /* per-site seed base       */
  orc_site_seeds(rseed, n, seeds);
  return 0;
}

/* ----------------------------------------------------------------------
   Site energy and the rejection event.
   ---------------------------------------------------------------------- */

typedef struct {
  int ns;
  int offs[27 * 3];
  int64_t nx, ny, nz;
} orc_geom;

static void orc_geom_init(const orc_desc *d, orc_geom *G)
{
  G->ns = orc_stencil(d->lattice, G->offs);
  G->nx = d->nx; G->ny = d->ny; G->nz = d->nz;
}

/* neighbour ids of site (i,j,k) with periodic wrap, in stencil order:
   the wrapped coordinates i-1,i,i+1 (likewise j, k) are formed once and the
   stencil entries index them */
static inline void orc_site_nbrs(const orc_geom *G, int64_t i, int64_t j, int64_t k,
                                 int64_t *nb)
{
  int64_t xs[3] = {i == 0 ? G->nx - 1 : i - 1, i, i == G->nx - 1 ? 0 : i + 1};
  int64_t ys[3] = {j == 0 ? G->ny - 1 : j - 1, j, j == G->ny - 1 ? 0 : j + 1};
  int64_t zs[3] = {k == 0 ? G->nz - 1 : k - 1, k, k == G->nz - 1 ? 0 : k + 1};
  for (int m = 0; m < G->ns; m++)
    nb[m] = xs[G->offs[3 * m] + 1] +
            G->nx * (ys[G->offs[3 * m + 1] + 1] + G->ny * zs[G->offs[3 * m + 2] + 1]);
}

/* site_energy: number of unlike neighbours, duplicates counted with
   multiplicity (src/app_potts.cpp:97-104, src/app_ising.cpp:88-95) */
static inline double orc_site_energy_nb(int ns, const int64_t *nb, const int *spin, int me)
{
  int e = 0;
  for (int m = 0; m < ns; m++)
    if (spin[nb[m]] != me) e++;
  return (double)e;
}

/* The distinct neighbour values unlike `me`, in order of first appearance
   along the neighbour list (src/app_potts_neigh.cpp:78-88, the same loop in
   src/app_potts_neighonly.cpp:78-88).  Returns their number. */
static int orc_unique_unlike(int ns, const int64_t *nb, const int *spin, int me, int *unique)
{
  int nevent = 0;
  for (int j = 0; j < ns; j++) {
    int value = spin[nb[j]];
    if (value == me) continue;
    int m;
    for (m = 0; m < nevent; m++)
      if (value == unique[m]) break;
    if (m < nevent) continue;
    unique[nevent++] = value;
  }
  return nevent;
}

/* One strict site event; mask bookkeeping omitted (it never skips a site
   in color/strict mode).  nb holds the site's neighbour ids in the
   reference's order.  Returns 1 when the spin changed.
     potts          AppPotts::site_event_rejection (src/app_potts.cpp:112-146)
     ising          AppIsing::site_event_rejection (src/app_ising.cpp:104-136)
     potts/neigh    AppPottsNeigh (src/app_potts_neigh.cpp:71-116): draw
                    iran = (int)(maxneigh*u); no event when iran >= nevent
     potts/neighonly AppPottsNeighOnly (src/app_potts_neighonly.cpp:71-118):
                    no draw at all when nevent == 0; iran = (int)(nevent*u)
     ising/single   AppIsingSingle (src/app_ising_single.cpp:39-77): the flip
                    is proposed without a draw */
static int orc_event_nb(const orc_desc *d, int ns, const int64_t *nb, int *spin,
                        int64_t g, int *seed, double t_inverse)
{
  int old = spin[g];
  double einitial = orc_site_energy_nb(ns, nb, spin, old);
  if (d->app == ORC_APP_POTTS) {
    int q = d->nspins;
    int iran = (int)(q * orc_pm_uniform(seed)) + 1;
    if (iran > q) iran = q;
    spin[g] = iran;
  } else if (d->app == ORC_APP_ISING) {
    spin[g] = orc_pm_uniform(seed) < 0.5 ? 1 : 2;
  } else if (d->app == ORC_APP_POTTS_NEIGH) {
    int unique[26];
    int nevent = orc_unique_unlike(ns, nb, spin, old, unique);
    int iran = (int)(ns * orc_pm_uniform(seed));
    if (iran >= nevent) return 0;
    spin[g] = unique[iran];
  } else if (d->app == ORC_APP_POTTS_NEIGHONLY) {
    int unique[26];
    int nevent = orc_unique_unlike(ns, nb, spin, old, unique);
    if (nevent == 0) return 0;
    int iran = (int)(nevent * orc_pm_uniform(seed));
    if (iran >= nevent) iran = nevent - 1;
    spin[g] = unique[iran];
  } else {
    spin[g] = old == 1 ? 2 : 1;
  }
  double efinal = orc_site_energy_nb(ns, nb, spin, spin[g]);
  if (efinal <= einitial) {
    /* downhill or level: accept */
  } else if (d->temperature == 0.0) {
    spin[g] = old;
  } else if (orc_pm_uniform(seed) > exp((einitial - efinal) * t_inverse)) {
    spin[g] = old;
  }
  return spin[g] != old;
}

static int orc_event(const orc_desc *d, const orc_geom *G, int *spin,
                     int64_t g, int *seed, double t_inverse)
{
  int64_t nb[26];
  orc_site_nbrs(G, g % G->nx, (g / G->nx) % G->ny, g / (G->nx * G->ny), nb);
  return orc_event_nb(d, G->ns, nb, spin, g, seed, t_inverse);
}

int orc_site_event(const orc_desc *d, int *spin, int64_t g, int *seed)
{
  orc_geom G;
  orc_geom_init(d, &G);
  double t_inverse = d->temperature != 0.0 ? 1.0 / d->temperature : 0.0;
  return orc_event(d, &G, spin, g, seed, t_inverse);
}

double orc_energy(const orc_desc *d, const int *spin)
{
  orc_geom G;
  orc_geom_init(d, &G);
  int64_t n = G.nx * G.ny * G.nz;
  double e = 0.0;
  /* integer partial sums are exact; the double total equals the
     reference's left-to-right double sum because every partial sum is an
     integer below 2^53 */
  int64_t tot = 0;
  (void)n;
#pragma omp parallel for collapse(2) reduction(+:tot) schedule(static)
  for (int64_t k = 0; k < G.nz; k++)
    for (int64_t j = 0; j < G.ny; j++)
      for (int64_t i = 0; i < G.nx; i++) {
        int64_t nb[26];
        int64_t g = i + G.nx * (j + G.ny * k);
        orc_site_nbrs(&G, i, j, k, nb);
        tot += (int64_t)orc_site_energy_nb(G.ns, nb, spin, spin[g]);
      }
  e = (double)tot;
  return e;
}

/* AppLattice::iterate_rejection colour loop (src/app_lattice.cpp:632-711)
   with sweep_nomask_strict (:796-806): for colour 1..K, every site of that
   colour in ascending id, its own seed loaded into the shared generator and
   stored back after the event. */
int orc_sweeps(const orc_desc *d, int *spin, int *seeds, int nsweeps,
               int64_t *naccept_out, double *energy_out, int nthreads)
{
  int K = orc_ncolors(d);
  if (K == 0) return -1;
  orc_geom G;
  orc_geom_init(d, &G);
  int64_t n = G.nx * G.ny * G.nz;
  double t_inverse = d->temperature != 0.0 ? 1.0 / d->temperature : 0.0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  int64_t naccept = 0;
  for (int sw = 0; sw < nsweeps; sw++) {
    for (int c = 1; c <= K; c++) {
      int64_t acc = 0;
      /* sites of one colour share no neighbour, so their order is free */
#pragma omp parallel for collapse(2) reduction(+:acc) schedule(static)
      for (int64_t k = 0; k < G.nz; k++)
        for (int64_t j = 0; j < G.ny; j++)
          for (int64_t i = 0; i < G.nx; i++) {
            int64_t g = i + G.nx * (j + G.ny * k);
            if (orc_color_ijk(d->lattice, i, j, k) != c) continue;
            int64_t nb[26];
            orc_site_nbrs(&G, i, j, k, nb);
            int s = seeds[g];
            acc += orc_event_nb(d, G.ns, nb, spin, g, &s, t_inverse);
            seeds[g] = s;
          }
      naccept += acc;
    }
    if (naccept_out) naccept_out[sw] = naccept;
    if (energy_out) energy_out[sw] = orc_energy(d, spin);
  }
  (void)n;
  return 0;
}

/* Sweep count of `run T` (App::run src/app.cpp:93, AppLattice::setup
   src/app_lattice.cpp:401-420, iterate_rejection :704-706). */
int64_t orc_run_nsweeps(double t0, double runtime, double dt_sweep, double *t_end)
{
  double stoptime = t0 + runtime;
  double time = t0;
  double dt = dt_sweep;
  if (stoptime - time < dt) dt = stoptime - time;
  int64_t n = 0;
  if (stoptime > time) {
    for (;;) {
      n++;
      time += dt;
      if (time >= stoptime) break;
    }
  }
  if (t_end) *t_end = time;
  return n;
}
