/*
 * spk_oracle.h -- CPU restatement of SPPARKS' strict-mode Metropolis colour
 * sweep (sweep color/strict) for app_style potts / ising and the variants
 * potts/neigh, potts/neighonly and ising/single.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against; it is loaded only by tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py.  The product (libspk_hip.so and the spk_hip
 * front end) never links, loads or calls it.
 *
 * Each function cites the reference file:line whose behaviour it restates
 * (paths relative to the SPPARKS source tree).
 */
#ifndef SPK_ORACLE_H
#define SPK_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* app_style potts, ising, and the strict-capable variants that differ only
   in the proposed move: potts/neigh (src/app_potts_neigh.cpp), potts/neighonly
   (src/app_potts_neighonly.cpp), ising/single (src/app_ising_single.cpp) */
enum { ORC_APP_POTTS = 0, ORC_APP_ISING = 1, ORC_APP_POTTS_NEIGH = 2, ORC_APP_POTTS_NEIGHONLY = 3,
       ORC_APP_ISING_SINGLE = 4 };
enum { ORC_SQ_4N = 0, ORC_SQ_8N = 1, ORC_SC_6N = 2, ORC_SC_26N = 3 };

typedef struct {
  int app;            /* ORC_APP_* */
  int lattice;        /* ORC_* lattice style */
  int nx, ny, nz;     /* nz = 1 for 2d lattices */
  int nspins;         /* Potts Q; ignored (2) for the Ising styles */
  double temperature; /* kT; 0.0 selects the zero-temperature branch */
} orc_desc;

/* --- Park-Miller minimal standard (src/random_park.cpp) --------------- */
double orc_pm_uniform(int *seed);                     /* :79-86  */
int    orc_pm_irandom(int *seed, int n);              /* :92-97  */
int    orc_pm_from_double(double rseed);              /* :45-49  */
int    orc_pm_tagreset(double rseed, int64_t offset, int warmup); /* :67-73 */

/* --- Marsaglia master RNG (src/random_mars.cpp:43-105) ----------------- */
typedef struct { double u[98]; double c, cd, cm; int i97, j97; } orc_ranmars;
void   orc_ranmars_init(orc_ranmars *r, int seed);
double orc_ranmars_uniform(orc_ranmars *r);

/* --- lattice geometry (src/lattice.cpp:187-283, src/create_sites.cpp) -- */
int  orc_maxneigh(int lattice);
int  orc_ncolors(const orc_desc *d);         /* 0 if edges incommensurate */
int  orc_color(const orc_desc *d, int64_t gid);   /* 1-based gid -> 1..K */
/* neighbour stencil: ns*3 ints (di,dj,dk) in the reference's loop order */
int  orc_stencil(int lattice, int *offs);
/* explicit neighbour gids (0-based), N*maxneigh, periodic wrap */
void orc_neighbors(const orc_desc *d, int *neigh);

/* --- initial state ------------------------------------------------------ */
/* set site range lo hi, loop all (src/set.cpp:582-627) */
void orc_set_range(int seed0, int lo, int hi, int64_t n, int *spin);
/* per-site strict seeds (src/app_lattice.cpp:287-295) */
void orc_site_seeds(double rseed, int64_t n, int *seeds);
/* whole script chain: seed S -> set site range lo hi -> first run's
   four master draws -> site seeds.  Returns 0 on success. */
int  orc_script_init(const orc_desc *d, int master_seed, int lo, int hi,
                     int *spin, int *seeds);

/* --- the hot path ------------------------------------------------------- */
/* total energy, sum of site_energy (src/diag_energy.cpp:53-66) */
double orc_energy(const orc_desc *d, const int *spin);
/* nsweeps strict colour sweeps (src/app_lattice.cpp:632-711,796-819 with
   src/app_potts.cpp:97-146 / src/app_ising.cpp:88-136).  naccept_out and
   energy_out (nullable) receive cumulative Naccept and the energy after each
   sweep.  nthreads <= 0 uses the OpenMP default.  Returns 0 on success. */
int  orc_sweeps(const orc_desc *d, int *spin, int *seeds, int nsweeps,
                int64_t *naccept_out, double *energy_out, int nthreads);
/* one site update with the shared per-site RNG, exposed for unit tests */
int  orc_site_event(const orc_desc *d, int *spin, int64_t g, int *seed);

/* --- run-time bookkeeping ------------------------------------------------ */
/* number of sweeps `run T` performs from time t0 (src/app_lattice.cpp:
   382-421, 704-708): dt = min(dt_sweep, T), time += dt until >= t0+T */
int64_t orc_run_nsweeps(double t0, double runtime, double dt_sweep,
                        double *t_end);

#ifdef __cplusplus
}
#endif
#endif
