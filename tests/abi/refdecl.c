/*
 * refdecl.c -- a caller written against the reference's library interface:
 * it declares the seven entry points exactly as src/library.h:31-38 does,
 * with MPI_Comm as MPICH's int handle (no MPI is installed here), includes
 * no header of this project, and links libspparks_hip.so.  Test
 * infrastructure, not part of the library.
 *
 * usage: refdecl SCRIPT    (a 2d sq/8n Potts script, e.g. the 128x128 one)
 * prints: nglobal, energy after the script, two site coordinates from
 * "xyz", which iarray/darray names exist, the energy after writing stripes
 * through "site", and the energy after one more sweep next to the energy
 * recomputed on the host from the refreshed "site" array.
 */
#include <stdio.h>
#include <string.h>

typedef int MPI_Comm;                           /* MPICH's handle type */
#define MPI_COMM_WORLD ((MPI_Comm)0x44000000)   /* MPICH's value */

void spparks_open(int, char **, MPI_Comm, void **);
void spparks_open_no_mpi(int, char **, void **);
void spparks_close(void *);
void spparks_file(void *, char *);
char *spparks_command(void *, char *);
void *spparks_extract(void *, char *);
double spparks_energy(void *);

/* sum over sites of unlike neighbours on a periodic nx x ny sq/8n lattice */
static double host_energy(const int *s, int nx, int ny) {
  double e = 0.0;
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++)
      for (int dj = -1; dj <= 1; dj++)
        for (int di = -1; di <= 1; di++) {
          if (!di && !dj) continue;
          const int ii = (i + di + nx) % nx, jj = (j + dj + ny) % ny;
          e += s[ii + nx * jj] != s[i + nx * j];
        }
  return e;
}

int main(int argc, char **argv) {
  if (argc != 2) {
    fprintf(stderr, "usage: refdecl SCRIPT\n");
    return 2;
  }
  char a0[] = "refdecl", a1[] = "-screen", a2[] = "none", a3[] = "-log", a4[] = "none";
  char *args[] = {a0, a1, a2, a3, a4};
  void *spk = NULL;
  spparks_open(5, args, MPI_COMM_WORLD, &spk);
  if (!spk) {
    fprintf(stderr, "spparks_open returned no instance\n");
    return 1;
  }
  spparks_file(spk, argv[1]);
  char n_nglobal[] = "nglobal", n_xyz[] = "xyz", n_site[] = "site", n_i1[] = "iarray1", n_i2[] = "iarray2",
       n_d1[] = "darray1";
  const int *nglobal = (const int *)spparks_extract(spk, n_nglobal);
  if (!nglobal) {
    fprintf(stderr, "no nglobal\n");
    return 1;
  }
  const int n = *nglobal, nx = 128, ny = n / 128;
  printf("nglobal %d energy %.0f\n", n, spparks_energy(spk));
  double **xyz = (double **)spparks_extract(spk, n_xyz);
  if (!xyz) {
    fprintf(stderr, "no xyz\n");
    return 1;
  }
  printf("xyz 129 %g %g %g\n", xyz[129][0], xyz[129][1], xyz[129][2]);
  printf("xyz %d %g %g %g\n", n - 1, xyz[n - 1][0], xyz[n - 1][1], xyz[n - 1][2]);
  int *site = (int *)spparks_extract(spk, n_site);
  int *i1 = (int *)spparks_extract(spk, n_i1);
  printf("iarray1_is_site %d iarray2_null %d darray1_null %d\n", site && i1 == site,
         spparks_extract(spk, n_i2) == NULL, spparks_extract(spk, n_d1) == NULL);
  /* rows of 1 and 2 alternate: every site has 6 unlike neighbours */
  for (int g = 0; g < n; g++) site[g] = 1 + (g / nx) % 2;
  printf("striped energy %.0f\n", spparks_energy(spk));
  char run[] = "run 0.01";
  if (!spparks_command(spk, run)) {
    fprintf(stderr, "run failed\n");
    return 1;
  }
  printf("after_run energy %.0f host %.0f\n", spparks_energy(spk), host_energy(site, nx, ny));
  spparks_close(spk);
  return 0;
}
