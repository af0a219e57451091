/*
 * spk_hip.h -- C ABI of libspk_hip.so: SPPARKS' strict-mode Metropolis colour
 * sweep (`sweep color/strict`, app_style potts / ising) on one MI355X.
 *
 * Plain C: pointers, sizes and status codes only.  Every call returns 0 on
 * success or a negative SPK_HIP_E* code; spk_hip_last_error() then holds a
 * thread-local message.  Device memory belongs to the handle; host arrays
 * belong to the caller and are copied.  Each handle owns one HIP stream and
 * is not thread-safe; distinct handles may be driven from distinct threads.
 *
 * Host arrays use the reference's site order: 0-based index
 * g = id-1 = i + nx*(j + ny*k), x fastest (src/create_sites.cpp:344-350).
 * On the device, spins and seeds are held colour-blocked (all sites of colour
 * 1, then colour 2, ...), which is invisible through this interface.
 *
 * Reference interfaces replaced (paths relative to the SPPARKS source tree):
 *   - the per-colour sweep `(this->*sweep)(n, list)` bound to
 *     AppLattice::sweep_nomask_strict / sweep_mask_strict
 *     (src/app_lattice.h:165-170, src/app_lattice.cpp:305-314, :674-677,
 *     :796-819), which calls AppPotts / AppIsing::site_event_rejection and
 *     site_energy (src/app_lattice.h:73-74, src/app_potts.cpp:97-146,
 *     src/app_ising.cpp:88-136);
 *   - the per-site RandomPark seeds `siteseeds` (src/app_lattice.h:119,
 *     src/app_lattice.cpp:287-295, src/random_park.cpp:67-86);
 *   - DiagEnergy::compute and AppLattice::stats reductions
 *     (src/diag_energy.cpp:53-66, src/app_lattice.cpp:917-934).
 */
#ifndef SPK_HIP_H
#define SPK_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPK_HIP_ABI_VERSION 1

/* status codes */
#define SPK_HIP_OK        0
#define SPK_HIP_EINVAL   -1   /* bad argument / incommensurate lattice    */
#define SPK_HIP_EDEVICE  -2   /* HIP runtime error                        */
#define SPK_HIP_ENOMEM   -3   /* device or host allocation failed         */
#define SPK_HIP_ESTATE   -4   /* call not valid in the handle's state     */
#define SPK_HIP_EVALUE   -5   /* site values outside the app's range      */

/* app styles (src/app_potts.h:18-19, src/app_ising.h:18-19) */
#define SPK_HIP_POTTS 0
#define SPK_HIP_ISING 1
/* proposal variants (app_style potts/neigh Q, potts/neighonly Q, ising/single) */
#define SPK_HIP_POTTS_NEIGH 2
#define SPK_HIP_POTTS_NEIGHONLY 3
#define SPK_HIP_ISING_SINGLE 4

/* lattice styles (src/lattice.cpp:44-56): colours 2 / 4 / 2 / 8 */
#define SPK_HIP_SQ_4N  0
#define SPK_HIP_SQ_8N  1
#define SPK_HIP_SC_6N  2
#define SPK_HIP_SC_26N 3

typedef struct spk_hip_desc {
  int app;             /* SPK_HIP_POTTS, SPK_HIP_ISING or a variant       */
  int lattice;         /* SPK_HIP_SQ_4N ... SPK_HIP_SC_26N                */
  int nx, ny, nz;      /* periodic box in lattice units; nz = 1 in 2d     */
  int nspins;          /* Potts Q (`app_style potts[/neigh[only]] Q`);
                          ignored (2) for ising and ising/single       */
  double temperature;  /* kT (`temperature` command); 0.0 = T=0 branch    */
} spk_hip_desc;

typedef struct spk_hip_lattice spk_hip_lattice;   /* opaque handle */

/* Thread-local message for the last failing call on this thread. */
const char *spk_hip_last_error(void);

/* Number of colours the strict sweep uses for `d` (Lattice::ncolors,
   src/lattice.cpp:187-228); SPK_HIP_EINVAL when an edge is odd. */
int spk_hip_ncolors(const spk_hip_desc *d);

/* Metropolis table of a temperature (host only, no device needed):
   tab[d] = exp((ei-ef)*t_inverse) for ef-ei = d = 1..maxd, t_inverse =
   1.0/temperature, exactly as src/app_potts.cpp:130 / src/app_ising.cpp:120
   evaluate it with this process's libm; thr[d] = the smallest Park-Miller
   state s with (1.0/2147483647)*s > tab[d] (2147483647 if none), so that the
   reference's rejection `uniform() > tab[d]` is `s >= thr[d]`.  thr[0] gets
   the smallest s with (1.0/2147483647)*s >= 0.5 (the Ising proposal cut).
   Either output may be NULL; arrays hold maxd+1 entries. */
int spk_hip_metropolis_table(double temperature, int maxd, double *tab, int32_t *thr);

/* The `seed S` command's master generator (RanMars::init / uniform,
   src/random_mars.cpp:43-105; src/input.cpp:942-951), host only: writes the
   first n uniform() draws that follow init(S).  A strict run that starts
   with `set site range` uses draw 0 for the set generator
   (RandomPark(double), src/random_park.cpp:45-49; src/set.cpp:598) and draw
   4 as the site-seed base rseed (src/app_lattice.cpp:274-292). */
int spk_hip_master_draws(int seed, int n, double *out);

/* Allocate device state for one lattice (AppLattice::init's colour sets and
   siteseeds, src/app_lattice.cpp:190-295).  Spins start at 1, seeds at 1. */
int spk_hip_create(const spk_hip_desc *d, spk_hip_lattice **out);
void spk_hip_destroy(spk_hip_lattice *h);

/* Change kT between runs (AppLattice::set_temperature,
   src/app_lattice.cpp:893-898); rebuilds the Metropolis table. */
int spk_hip_set_temperature(spk_hip_lattice *h, double temperature);

/* Copy spins and/or seeds (either pointer may be NULL), length nx*ny*nz,
   reference site order.  Upload checks 1 <= spin <= Q (AppPotts::init_app,
   src/app_potts.cpp:78-91) and 1 <= seed <= 2^31-2. */
int spk_hip_upload(spk_hip_lattice *h, const int32_t *spin, const int32_t *seed);
int spk_hip_download(spk_hip_lattice *h, int32_t *spin, int32_t *seed);

/* Device-side initial state.
   set_range: `set site range lo hi` with loop all (src/set.cpp:582-627):
     site id g gets irandom(hi-lo+1)-1+lo from a RandomPark started at
     seed0 and advanced g times (Park-Miller jump-ahead).
   set_value: `set site value v`.
   init_seeds: siteseeds[i] = tagreset(rseed, id[i], 100)
     (src/app_lattice.cpp:287-295, src/random_park.cpp:67-73). */
int spk_hip_set_range(spk_hip_lattice *h, int seed0, int lo, int hi);
int spk_hip_set_value(spk_hip_lattice *h, int value);
int spk_hip_init_seeds(spk_hip_lattice *h, double rseed);

/* Run `nsweeps` strict sweeps (colours 1..K each), as nsweeps iterations of
   AppLattice::iterate_rejection (src/app_lattice.cpp:632-711).
   naccept_total (nullable): accepted flips over the call (added, not set).
   naccept_per_sweep / energy_per_sweep (nullable, length nsweeps):
   cumulative accepted flips since the call began and the total energy
   (DiagEnergy, src/diag_energy.cpp:53-66) after each sweep. */
int spk_hip_sweeps(spk_hip_lattice *h, int nsweeps, int64_t *naccept_total,
                   int64_t *naccept_per_sweep, double *energy_per_sweep);

/* spk_hip_sweeps for callers that cannot allocate memory when they run
   sweeps: code under HIP graph capture, serving loops, harnesses that
   forbid hipMalloc.  No counterpart in the reference.  It allocates
   nothing and does not wait: it queues nsweeps sweeps (0..256) on the
   handle's stream, through the per-colour kernels on the byte layout
   (bit planes are unpacked first), with the accumulator rows in a
   library-owned device array of 2 MiB per GPU, and writes 2*nsweeps
   uint64 to `rows`, device memory the caller owns: for sweep s,
   rows[2s] = accepted flips and rows[2s+1] = the sum of ef-ei over them as
   a two's-complement int64, so the energy after sweep s is the energy
   before the call plus twice the running sum.  Results are those of
   spk_hip_sweeps once the stream reaches them (spk_hip_sync).  All calls on
   one GPU share the array: calls that overlap on one GPU, from different
   handles or threads, race. */
int spk_hip_sweeps_noalloc(spk_hip_lattice *h, int nsweeps, unsigned long long *rows);

/* Total energy: sum over sites of #unlike neighbours
   (src/diag_energy.cpp:53-66, exact in double). */
int spk_hip_energy(spk_hip_lattice *h, double *energy);

/* State checksum, independent of the device layout: the wrapping sum over
   sites g (0-based, reference order) of mix64(g*0xD1B54A32D192ED03 +
   spin*0xABC98388FB8FAC03 + seed), with mix64 the splitmix64 finaliser
   (spin and seed taken as unsigned 32-bit).  Lets a caller compare
   states without copying them back. */
int spk_hip_checksum(spk_hip_lattice *h, uint64_t *sum);

/* Block until all work queued on the handle's stream has finished. */
int spk_hip_sync(spk_hip_lattice *h);

/* Launch configuration of the per-colour sweep kernel (results do not
   depend on it: colour/strict is decomposition independent).
   block = threads per workgroup (a power of two in 64..1024); a workgroup
   covers a tile of sub-grid rows (x, y and z), rows in y and z. */
int spk_hip_set_launch(spk_hip_lattice *h, int block);

/* Two-state sq/4n lattices (Ising, ising/single, or Potts with Q = 2) sweep on bit planes
   by default: the spins are held one bit per site, 32 sites of a sub-grid
   row per word, and the seeds in a wavefront-interleaved order.  The first
   spk_hip_sweeps call packs the planes; they then hold the state across
   sweeps calls, and a call that needs the byte layout (upload, download,
   energy, checksum, set_range, set_value, init_seeds, set_bitplanes, the
   one-workgroup kernel) unpacks them first.  By default (enable = 1) a
   lattice whose sub-grid rows hold 33..64 plane words (nx 2050..4096) and
   at least 64 rows runs two whole sweeps per launch: a workgroup owns a
   band of 16 rows, keeps its planes with 4 halo rows either side in LDS
   and its threads' seeds in registers through the four colour passes, so
   each seed is read and written once per two sweeps; the last n % 4 sweeps
   of a call take the per-colour kernel with 8 sites per thread (the same
   layout), and every sweep of other lattices the per-colour kernel with 32
   sites per thread.  enable = 0 keeps such lattices on the byte-per-site kernel;
   8, 16 or 32 selects the per-colour kernel with that many sites per thread
   for every sweep.  Results are identical either way.  No counterpart in
   the reference (a storage choice of this implementation). */
int spk_hip_set_bitplanes(spk_hip_lattice *h, int enable);

/* Lattices whose spins and seeds fit one workgroup's LDS (up to about 30 k
   sites at one byte per spin on MI355X: the reference's CPU-sized cases)
   run every sweep of a spk_hip_sweeps call in one launch of a single
   workgroup, the state read and written once.  On by default; enable = 0
   sends them through the per-colour kernels.  Results are identical either
   way.  No counterpart in the reference (a launch choice of this
   implementation). */
int spk_hip_set_resident(spk_hip_lattice *h, int enable);

/* Sweeps run as replays of captured HIP graphs: a call of n sweeps replays
   n / chunk_sweeps graphs of chunk_sweeps sweeps (default 32, at most 1024),
   then one graph of the remainder (n % chunk_sweeps sweeps).  A
   graph is captured on first use and kept until a setting that changes the
   launches (set_launch, set_bitplanes, set_temperature, set_graph) drops
   the cache.  chunk_sweeps = 0 launches every kernel from the host.
   Results are identical either way. */
int spk_hip_set_graph(spk_hip_lattice *h, int chunk_sweeps);

/* Set-up for a later spk_hip_sweeps(h, nsweeps, ...): captures the graphs
   that call will replay and, for a lattice on bit planes, packs the planes,
   so the call itself captures and converts nothing.  Runs no sweep and
   changes no result; calling it is optional. */
int spk_hip_prepare(spk_hip_lattice *h, int nsweeps);

/* Device-side timing of the last spk_hip_sweeps call: milliseconds between
   HIP events recorded on the handle's stream around the kernel sequence,
   and the number of sweep-kernel launches it made. */
int spk_hip_last_timing(spk_hip_lattice *h, double *ms, int64_t *nlaunch);

/* ======================================================================
   Batched replicas: many independent lattices (each its own script: app,
   lattice, size, Q, kT, seed, sweep count) resident in one device pool and
   advanced together.  Each colour step of a block is one persistent kernel
   over a tile work-list spanning every replica that is due a colour pass
   at that step; a replica's K*n passes of an n-sweep block are spread evenly
   over the block's steps, always in colour order 1..K.  Every replica's
   trajectory is identical to running it alone (the strict sweep is
   decomposition independent), which the tests check.  Blocks after the
   first are replayed as HIP graphs.  No counterpart in the reference, which
   runs one lattice per process (its closest relative is `-partition`,
   src/spparks.cpp:58-65).
   ====================================================================== */

typedef struct spk_hip_batch spk_hip_batch;   /* opaque handle */

typedef struct spk_hip_replica_desc {
  spk_hip_desc lattice;   /* app, lattice style, edges, Q, kT               */
  int seed;               /* the replica script's `seed S`                  */
  int lo, hi;             /* `set site range lo hi`; lo == hi: `set site value lo` */
  int nsweeps;            /* strict sweeps to run (`run` length)            */
} spk_hip_replica_desc;

/* Device pool of pool_bytes for up to max_replicas replicas. */
int spk_hip_batch_create(int64_t pool_bytes, int max_replicas, spk_hip_batch **out);
void spk_hip_batch_destroy(spk_hip_batch *b);

/* Bytes a replica of this description occupies in a pool (admission). */
int spk_hip_replica_bytes(const spk_hip_desc *d, int64_t *bytes);

/* Admit a replica: allocate it in the pool and build its initial state on
   the device (set range / value, strict seeds) exactly as its script would
   (seed, set, first run's draws).  SPK_HIP_ENOMEM when the pool has no room
   or all replica slots are taken; SPK_HIP_EINVAL for a replica of 2^30 or
   more sites (run such a lattice alone).  The handle is unchanged then. */
int spk_hip_batch_add(spk_hip_batch *b, const spk_hip_replica_desc *r, int *id);
/* Retire a replica and return its pool space. */
int spk_hip_batch_remove(spk_hip_batch *b, int id);

/* Spins / seeds of one replica in the reference's site order (NULL skips). */
int spk_hip_batch_upload(spk_hip_batch *b, int id, const int32_t *spin, const int32_t *seed);
int spk_hip_batch_download(spk_hip_batch *b, int id, int32_t *spin, int32_t *seed);

/* Advance every replica with sweeps left by min(left, max_sweeps) sweeps.
   use_graph != 0 captures the block's launches into a HIP graph (cached by
   block shape) and replays it. */
int spk_hip_batch_run(spk_hip_batch *b, int max_sweeps, int use_graph);

/* Per replica: sweeps done and left, accepted flips and energy now. */
int spk_hip_batch_status(spk_hip_batch *b, int id, int *done, int *left, int64_t *naccept,
                         double *energy);
/* Cumulative Naccept and energy after each of the first n sweeps. */
int spk_hip_batch_trace(spk_hip_batch *b, int id, int n, int64_t *naccept, double *energy);

/* Launch shape of the step kernel: block = threads per tile (64..1024,
   multiple of 64), grid = persistent workgroups.  0, 0 = autotune: the next
   run times candidate shapes on its own steps (results do not depend on the
   shape) and keeps the fastest. */
int spk_hip_batch_set_launch(spk_hip_batch *b, int block, int grid);

/* Last run: device milliseconds between events around its steps, site
   updates it made, step kernels it launched, and the launch shape used. */
int spk_hip_batch_last_timing(spk_hip_batch *b, double *ms, int64_t *updates, int64_t *nlaunch,
                              int *block, int *grid);

/* Pool occupancy: bytes in use, largest free extent, live replicas. */
int spk_hip_batch_pool(spk_hip_batch *b, int64_t *used, int64_t *largest_free, int *live);

/* spk_hip_checksum of one replica. */
int spk_hip_batch_checksum(spk_hip_batch *b, int id, uint64_t *sum);

/* ======================================================================
   Serving: continuous batching of a stream of solves.  Requests are
   admitted in arrival order as soon as they have arrived and the pool
   (the HBM budget) and a replica slot have room; every block advances all
   admitted requests together through spk_hip_batch_run; a request whose
   sweeps are done is checksummed, reported and retired at once, freeing
   its memory and slot for the next arrivals.  Times are seconds on a
   monotonic clock from the start of the call; arrivals in the future are
   waited for.  A request larger than the whole pool is refused
   (SPK_HIP_ENOMEM), and so is a request that is not a valid solve, such as
   an odd edge or 2^30 sites or more (SPK_HIP_EINVAL) or a set range
   outside 1..Q (SPK_HIP_EVALUE); the other requests are served as usual.  No counterpart in the reference.
   ====================================================================== */

typedef struct spk_hip_request {
  spk_hip_replica_desc replica;   /* the solve: lattice, seed, set, sweeps */
  double arrive;                  /* seconds after the start of the call   */
} spk_hip_request;

typedef struct spk_hip_result {
  double admit;                   /* admitted into the pool               */
  double finish;                  /* sweeps done and state checksummed    */
  int64_t naccept;
  double energy;
  uint64_t checksum;              /* spk_hip_checksum of the final state  */
  int status;                     /* 0 done; SPK_HIP_ENOMEM, _EINVAL or
                                     _EVALUE: refused (see spk_hip_serve) */
} spk_hip_result;

typedef struct spk_hip_serve_stats {
  double wall;                    /* start to last completion             */
  int64_t site_updates;           /* over completed requests              */
  int blocks;                     /* batch runs                           */
  int max_live;                   /* most requests resident at once       */
  int64_t max_pool_used;          /* bytes                                */
  double busy;                    /* seconds inside batch runs            */
} spk_hip_serve_stats;

/* req must be sorted by arrival time.  block_sweeps: sweeps per block;
   block, grid: launch shape of the step kernels (0, 0 = autotune). */
int spk_hip_serve(const spk_hip_request *req, int n, int64_t pool_bytes, int max_live, int block_sweeps,
                  int block, int grid, spk_hip_result *out, spk_hip_serve_stats *stats);

#ifdef __cplusplus
}
#endif
#endif
