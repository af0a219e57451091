/*
 * spparks_hip.h -- library interface of libspparks_hip.so: the spk_hip front
 * end (SPPARKS input-script commands for on-lattice Potts / Ising models,
 * strict colour sweep on the GPU) driven from C, C++ or Python ctypes.
 *
 * Same entry points, argument meaning and return conventions as the
 * reference's library interface (src/library.h:31-38, src/library.cpp:37-147)
 * and its ctypes wrapper (python/spparks.py:19-95), so that code written
 * against libspparks.so can load this library instead.  Differences:
 *   - spparks_open takes no MPI communicator argument; one GPU per instance;
 *   - an input error does not exit the process: the failing call returns
 *     (spparks_command returns NULL) and spparks_error_message() holds the
 *     reference's error text.
 */
#ifndef SPPARKS_HIP_H
#define SPPARKS_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* create an instance (src/library.cpp:37-64); argv as on the command line */
void spparks_open(int argc, char **argv, void **ptr);
void spparks_open_no_mpi(int argc, char **argv, void **ptr);
/* destroy (src/library.cpp:70-74) */
void spparks_close(void *ptr);
/* run an input script (src/library.cpp:80-84) */
void spparks_file(void *ptr, char *file);
/* run one command; returns the command word or NULL on error
   (src/library.cpp:90-94) */
char *spparks_command(void *ptr, char *str);
/* pointer to internal data (src/library.cpp:100-106, App::extract
   src/app.cpp:163-196): "site" / "iarray1" -> int* spins in id order,
   "id" -> int* ids, "nglobal" / "nlocal" / "dimension" -> int*,
   "boxxlo" ... "boxzhi" -> double*; additionally "seeds" -> int* strict
   seeds and "time" -> double*.  Lattice data live on the GPU, so the
   pointer is a host snapshot taken by the call; it stays valid until the
   next call on the instance.  NULL for unknown names. */
void *spparks_extract(void *ptr, char *name);
/* total energy of the current spins (src/library.cpp:118-147) */
double spparks_energy(void *ptr);
/* text of the last error on this instance, "" if none */
const char *spparks_error_message(void *ptr);

#ifdef __cplusplus
}
#endif
#endif
