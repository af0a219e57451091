/* consumer.c -- a plain C99 program that drives libspk_hip.so and
 * libspparks_hip.so through include/spk_hip.h and include/spparks_hip.h
 * only, as a C caller of the reference's library would.
 * Test infrastructure: built with gcc -std=c99 -pedantic -Werror, so both
 * headers are checked to be plain C.
 *
 * usage: consumer host   host-only entry points (no GPU needed)
 *        consumer gpu    a two-replica batch against the same scripts run
 *                        alone, and a script through spparks_command
 * Prints one "key values" line per result; exits non-zero on a failed call
 * or a mismatch.
 */
#include <inttypes.h>
#include <stdio.h>
#include <string.h>

#include "spk_hip.h"
#include "spparks_hip.h"

static int fail(const char *what) {
  fprintf(stderr, "%s failed: %s\n", what, spk_hip_last_error());
  return 1;
}

static int host_part(void) {
  spk_hip_desc d = {SPK_HIP_POTTS, SPK_HIP_SC_26N, 10, 8, 6, 10000, 0.5};
  printf("ncolors %d\n", spk_hip_ncolors(&d));
  d.nx = 7;
  const int rc = spk_hip_ncolors(&d);
  printf("odd_edge %d %s\n", rc, spk_hip_last_error());
  double tab[5];
  int32_t thr[5];
  if (spk_hip_metropolis_table(2.269, 4, tab, thr)) return fail("spk_hip_metropolis_table");
  for (int i = 0; i <= 4; i++) printf("thr %d %" PRId32 " %a\n", i, thr[i], tab[i]);
  double draws[5];
  if (spk_hip_master_draws(12345, 5, draws)) return fail("spk_hip_master_draws");
  for (int i = 0; i < 5; i++) printf("draw %d %a\n", i, draws[i]);
  return 0;
}

/* The replica's script run alone through a lattice handle: `seed S`, `set
   site range lo hi`, then `run` of nsweeps (master draw 0 seeds the set
   generator, draw 4 is the site-seed base). */
static int alone(const spk_hip_replica_desc *r, uint64_t *sum, int64_t *nacc, double *energy) {
  double draws[5];
  if (spk_hip_master_draws(r->seed, 5, draws)) return fail("spk_hip_master_draws");
  int s0 = (int)(draws[0] * 2147483647.0);
  if (s0 == 0) s0 = 1;
  spk_hip_lattice *h;
  if (spk_hip_create(&r->lattice, &h)) return fail("spk_hip_create");
  *nacc = 0;
  int rc = spk_hip_set_range(h, s0, r->lo, r->hi) || spk_hip_init_seeds(h, draws[4]) ||
           spk_hip_sweeps(h, r->nsweeps, nacc, NULL, NULL) || spk_hip_energy(h, energy) ||
           spk_hip_checksum(h, sum);
  if (rc) fail("lattice calls");
  spk_hip_destroy(h);
  return rc;
}

static int gpu_part(void) {
  spk_hip_replica_desc rep[2] = {
      {{SPK_HIP_ISING, SPK_HIP_SQ_4N, 512, 512, 1, 2, 2.269}, 1000, 1, 2, 100},
      {{SPK_HIP_POTTS, SPK_HIP_SC_26N, 32, 32, 32, 1000, 0.5}, 1256, 1, 1000, 60},
  };
  spk_hip_batch *b;
  if (spk_hip_batch_create(1LL << 30, 16, &b)) return fail("spk_hip_batch_create");
  int id[2];
  for (int k = 0; k < 2; k++)
    if (spk_hip_batch_add(b, &rep[k], &id[k])) return fail("spk_hip_batch_add");
  int left = 1;
  while (left > 0) {
    if (spk_hip_batch_run(b, 32, 1)) return fail("spk_hip_batch_run");
    left = 0;
    for (int k = 0; k < 2; k++) {
      int done, l;
      if (spk_hip_batch_status(b, id[k], &done, &l, NULL, NULL)) return fail("spk_hip_batch_status");
      left += l;
    }
  }
  int bad = 0;
  for (int k = 0; k < 2; k++) {
    uint64_t hb, ha;
    int64_t nb, na;
    double eb, ea;
    int done, l;
    if (spk_hip_batch_checksum(b, id[k], &hb) || spk_hip_batch_status(b, id[k], &done, &l, &nb, &eb))
      return fail("batch state");
    if (alone(&rep[k], &ha, &na, &ea)) return 1;
    printf("replica %d batch %" PRIu64 " %" PRId64 " %.0f alone %" PRIu64 " %" PRId64 " %.0f\n", k, hb, nb, eb,
           ha, na, ea);
    bad += hb != ha || nb != na || eb != ea;
  }
  spk_hip_batch_destroy(b);

  /* the reference's library calls: Potts Q=100 sq/8n 128x128, 100 sweeps */
  char *args[] = {"consumer", "-screen", "none", "-log", "none"};
  void *spk = NULL;
  spparks_open_no_mpi(5, args, &spk);
  if (!spk) return fail("spparks_open_no_mpi");
  static const char *script[] = {
      "seed 56789", "app_style potts 100", "dimension 2", "lattice sq/8n 1.0",
      "region box block 0 128 0 128 -0.5 0.5", "create_box box", "create_sites box",
      "set site range 1 100", "sweep color/strict", "temperature 0.5", "diag_style energy",
      "stats 0.5", "run 1.0"};
  for (size_t i = 0; i < sizeof script / sizeof script[0]; i++) {
    char line[128];
    snprintf(line, sizeof line, "%s", script[i]);
    if (!spparks_command(spk, line)) {
      fprintf(stderr, "spparks_command(%s) failed: %s\n", script[i], spparks_error_message(spk));
      return 1;
    }
  }
  const int *nglobal = (const int *)spparks_extract(spk, "nglobal");
  printf("library nglobal %d energy %.0f\n", nglobal ? *nglobal : -1, spparks_energy(spk));
  spparks_close(spk);
  return bad;
}

int main(int argc, char **argv) {
  if (argc == 2 && !strcmp(argv[1], "host")) return host_part();
  if (argc == 2 && !strcmp(argv[1], "gpu")) return gpu_part();
  fprintf(stderr, "usage: consumer host|gpu\n");
  return 2;
}
