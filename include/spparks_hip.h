/*
 * spparks_hip.h -- library interface of libspparks_hip.so: the spk_hip front
 * end (SPPARKS input-script commands for on-lattice Potts / Ising models,
 * strict colour sweep on the GPU) driven from C, C++ or Python ctypes.
 *
 * Same entry points, argument meaning and return conventions as the
 * reference's library interface (src/library.h:31-38, src/library.cpp:37-147)
 * and its ctypes wrapper (python/spparks.py:19-95), so that code written
 * against libspparks.so can load this library instead.  Differences:
 *   - spparks_open's communicator argument is accepted and ignored: one GPU
 *     per instance, no MPI;
 *   - an input error does not exit the process: the failing call returns
 *     (spparks_command returns NULL) and spparks_error_message() holds the
 *     reference's error text.
 */
#ifndef SPPARKS_HIP_H
#define SPPARKS_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* create an instance (src/library.cpp:37-64); argv as on the command line.
   spparks_open's third argument stands where the reference takes an
   MPI_Comm (src/library.h:31): a pointer-sized slot that also carries an
   integer handle (MPICH's int MPI_Comm) in the same argument register under
   the x86-64 and AArch64 C calling conventions, so a caller compiled
   against the reference's prototype links and runs unchanged.  The value
   is never read. */
void spparks_open(int argc, char **argv, void *comm, void **ptr);
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
   "id" -> int* ids, "xyz" -> double** (xyz[i][0..2], the coordinates
   create_sites gives site i), "nglobal" / "nlocal" / "dimension" -> int*,
   "boxxlo" ... "boxzhi" -> double*; "iarrayN" for N > 1 and "darrayN" are
   NULL (the potts and ising styles hold one integer and no double per
   site); additionally "seeds" -> int* strict seeds and "time" ->
   double*.
   Lattice data live on the GPU, so "site", "iarray1" and "seeds" return a
   host array that the library keeps in step with the device, as the
   reference's live arrays are: every later call on the instance first
   copies values written through the array to the device (checked against
   1..Q, as AppPotts::init_app does) and, after a command or file, refreshes
   the array from the device.  The pointer stays valid until the lattice is
   recreated (create_sites, clear) or the instance is closed.  NULL for
   unknown names. */
void *spparks_extract(void *ptr, char *name);
/* total energy of the current spins (src/library.cpp:118-147) */
double spparks_energy(void *ptr);
/* text of the last error on this instance, "" if none */
const char *spparks_error_message(void *ptr);

#ifdef __cplusplus
}
#endif
#endif
