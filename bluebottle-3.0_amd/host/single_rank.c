/*
 * The MPI layer of the reference (src/mpi_comm.c) on one GPU that holds the
 * whole domain: one rank, no neighbours, no windows.
 *
 * A program built from the reference's host objects (src/bluebottle.c,
 * domain.c, particle.c, recorder.c, scalar.c, rng.c, vtk.c) calls these
 * functions and reads these globals; with src/mpi_comm.c and the CUDA
 * objects left out of the link they come from this library:
 *   mpi_startup          selects device 0, initialises MPI when the program
 *                        links an MPI library (the reference's host code calls
 *                        MPI_Allreduce and MPI_Wtime itself), rank 0 of 1;
 *   mpi_dom_comm_init    no halo windows: there is no neighbour;
 *   mpi_parts_init       no MPI particle type: particles never leave the GPU;
 *   mpi_cuda_exchange_G* the periodic ghost refresh of the field on the GPU
 *                        (cuda_exchange), the work the reference's exchange
 *                        does on one rank with periodic faces (its self
 *                        exchange, src/mpi_comm.c:257-405);
 *   mpi_free, mpi_end    release nothing / finalise MPI if it was started.
 * The send/receive buffers and window handles of the particle partial sums
 * (src/cuda_physalis.cu:41-58, src/mpi_comm.c:59-136) exist so that
 * src/scalar.c links; on one rank nothing uses them.
 *
 * MPI is not linked into the library: MPI_Init, MPI_Initialized and
 * MPI_Finalize are looked up in the running program, so the library works
 * the same with or without an MPI library next to it.
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

#include "bluebottle.h"

/* The reference's recorder.c owns this clock (src/recorder.c:41); it is
 * set at start-up (src/mpi_comm.c:78).  Weak: a program without that file
 * has none. */
extern time_t startwalltime __attribute__((weak));

/* Partial-sum exchange buffers of the particle scalar heat path
 * (src/cuda_physalis.cu:41-58), read by src/scalar.c:115-224. */
real *_sum_send_e = NULL, *_sum_send_w = NULL, *_sum_send_n = NULL,
     *_sum_send_s = NULL, *_sum_send_t = NULL, *_sum_send_b = NULL;
real *_sum_recv_e = NULL, *_sum_recv_w = NULL, *_sum_recv_n = NULL,
     *_sum_recv_s = NULL, *_sum_recv_t = NULL, *_sum_recv_b = NULL;
/* MPI window handles (src/mpi_comm.c:71-136): storage of one OpenMPI
 * handle (a pointer) each. */
void *parts_recv_win_e = NULL, *parts_recv_win_w = NULL,
     *parts_recv_win_n = NULL, *parts_recv_win_s = NULL,
     *parts_recv_win_t = NULL, *parts_recv_win_b = NULL;

static int mpi_started = 0;

typedef int (*mpi_init_fn)(int *, char ***);
typedef int (*mpi_flag_fn)(int *);
typedef int (*mpi_void_fn)(void);

void mpi_startup(int argc, char *argv[])
{
  /* src/mpi_comm.c:39-79 */
  cuda_device_init(0);
  mpi_init_fn init = (mpi_init_fn) dlsym(RTLD_DEFAULT, "MPI_Init");
  mpi_flag_fn initialized = (mpi_flag_fn) dlsym(RTLD_DEFAULT,
                                                "MPI_Initialized");
  if (init != NULL) {
    int flag = 0;
    if (initialized != NULL) initialized(&flag);
    if (!flag) {
      init(&argc, &argv);
      mpi_started = 1;
    }
  }
  rank = 0;
  nprocs = 1;
  if (&startwalltime != NULL) startwalltime = time(NULL);
  printf("N%d >> One rank on GPU 0 holds the whole domain\n", rank);
}

void mpi_dom_comm_init(void)
{
  /* src/mpi_comm.c:81-96: no neighbour, no window */
}

void mpi_parts_init(void)
{
  /* src/mpi_comm.c:471-...: no particle leaves the GPU */
}

void mpi_cuda_exchange_Gcc(real *array) { cuda_exchange(array, GCC); }
void mpi_cuda_exchange_Gfx(real *array) { cuda_exchange(array, GFX); }
void mpi_cuda_exchange_Gfy(real *array) { cuda_exchange(array, GFY); }
void mpi_cuda_exchange_Gfz(real *array) { cuda_exchange(array, GFZ); }

void mpi_free(void)
{
  /* src/mpi_comm.c:1262-1309: no window was created */
}

void mpi_end(void)
{
  /* src/mpi_comm.c:1311-1315 */
  if (!mpi_started) return;
  mpi_void_fn fin = (mpi_void_fn) dlsym(RTLD_DEFAULT, "MPI_Finalize");
  if (rank == 0) printf("N%d >> Finalizing MPI... \n", rank);
  if (fin != NULL) fin();
  mpi_started = 0;
}
