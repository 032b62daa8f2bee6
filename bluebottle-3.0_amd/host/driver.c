/*
 * The timestep driver: src/bluebottle.c:28-436 restated for one GPU holding
 * the whole domain.  The MPI halo exchanges become cuda_exchange (periodic
 * planes on the device); the call order of the reference is kept.
 *
 * Lamb iterations: with particles the loop runs until the coefficient change
 * falls below lamb_residual (src/bluebottle.c:200-276).  Without particles
 * the reference's loop count is undefined (cuda_lamb_err leaves its error
 * uninitialised, src/cuda_physalis.cu:864,899-902) and every extra pass
 * recomputes the same fields, so exactly one pass is made.
 */
#include <float.h>
#include <sys/time.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "bluebottle.h"

extern int use_restart;
extern int recorder_enabled;
int bb_verbose = 1;

/* Particle hooks, provided by the particle/Physalis module when linked. */
void parts_read_input(void) __attribute__((weak));
void bb_parts_setup(void) __attribute__((weak));
void bb_parts_teardown(void) __attribute__((weak));
void cuda_part_BC(void) __attribute__((weak));
void cuda_part_BC_star(void) __attribute__((weak));
void cuda_part_p_fill(void) __attribute__((weak));
void cuda_move_parts_sub(void) __attribute__((weak));
void cuda_lamb(void) __attribute__((weak));
void cuda_lamb_err(real *error, int *num) __attribute__((weak));
void cuda_update_part_velocity(void) __attribute__((weak));
void cuda_update_part_position(void) __attribute__((weak));
void cuda_parts_internal(void) __attribute__((weak));

/* Learned warm start hooks (bb_warm.hip); they return at once unless
 * pp_warmstart is LEARNED. */
void bb_warm_setup(void);
void bb_warm_teardown(void);

static double wall_seconds(void)
{
  struct timeval tv;
  gettimeofday(&tv, NULL);
  return tv.tv_sec + 1e-6 * tv.tv_usec;
}

/* cuda_exchange of u, v, w (GFX, GFY, GFZ), in one launch */
void cuda_exchange_vel(real *a, real *b, real *c);
static void exchange_vel(real *a, real *b, real *c)
{
  cuda_exchange_vel(a, b, c);
}

static void pressure_solve(void)
{
  if (pp_precision == PP_MIXED) cuda_PP_cg_mixed();
  else if (NPARTS > 0) cuda_PP_cg();
  else cuda_PP_cg_noparts();
}

int bb_lamb_iters = 0;
real bb_lamb_err = 0.;
int bb_scalar_iters = 0;               /* scalar iterations of the last step */
real bb_scalar_err = 0.;

int bluebottle_setup(void)
{
  /* src/bluebottle.c:31-142 */
  domain_read_input();
  recorder_read_config();
  domain_fill();
  cuda_dom_malloc_host();
  cuda_dom_malloc_dev();
  if (SCALAR >= 1) {
    cuda_scalar_malloc_host();
    cuda_scalar_malloc_dev();
  }
  cuda_blocks_init();
  domain_init_fields();
  if (SCALAR >= 1) scalar_init_fields();
  NPARTS = 0;
  nparts = 0;
  if (parts_read_input) parts_read_input();
  cuda_part_malloc_host();
  if (!use_restart) {
    vtk_recorder_init();
    recorder_PP_init("solver_expd.rec");
    recorder_scalar_init("iter_etime.rec");
  }
  if (use_restart) {
    in_restart();
    if (ttime >= duration) {
      printf("\n...simulation completed");
      restart_stop = 1;
      exit(EXIT_SUCCESS);
    }
  }
  if (bb_parts_setup) bb_parts_setup();
  cuda_dom_push();
  cuda_exchange(_p, GCC);
  exchange_vel(_u, _v, _w);
  if (SCALAR >= 1) {
    cuda_scalar_push();
    cuda_exchange(_s, GCC);
  }
  cuda_build_cages();
  cuda_find_dt();
  cuda_dom_BC();
  cuda_exchange(_p, GCC);
  exchange_vel(_u, _v, _w);
  if (SCALAR >= 1) {
    cuda_scalar_BC(_s);
    cuda_scalar_BC(_s0);
    cuda_exchange(_s, GCC);
    cuda_exchange(_s0, GCC);
  }
  cuda_PP_init_jacobi_preconditioner();
  bb_warm_setup();
  /* initial output (src/bluebottle.c:159-162) */
  if (!use_restart) recorder_out();
  return 0;
}

int bluebottle_step(void)
{
  /* One pass of src/bluebottle.c:168-407. */
  ttime += dt;
  rec_cgns_flow_ttime_out += dt;
  rec_cgns_part_ttime_out += dt;
  rec_vtk_ttime_out += dt;
  stepnum++;
  int iter = 0;
  real iter_err = FLT_MAX;
  int iter_num = -1;
  if (rank == 0 && bb_verbose) {
    printf("\nEXPD: Time = %e of %e (dt = %e) (stepnum = %d).\n", ttime,
      duration, dt, stepnum);
    fflush(stdout);
  }
  cuda_compute_forcing();
  if (SCALAR >= 1) cuda_compute_boussinesq();
  compute_vel_BC();

  double mitertimestart = wall_seconds();
  while (iter_err > lamb_residual) {
    cuda_U_star();
    exchange_vel(_u_star, _v_star, _w_star);
    if (nparts > 0) cuda_part_BC_star();
    cuda_dom_BC_star();
    exchange_vel(_u_star, _v_star, _w_star);

    cuda_solvability();
    if (nparts > 0) cuda_part_BC_star();
    cuda_dom_BC_star();
    exchange_vel(_u_star, _v_star, _w_star);

    pressure_solve();
    cuda_exchange(_phi, GCC);
    cuda_dom_BC_p(_phi);

    cuda_project();
    exchange_vel(_u, _v, _w);
    if (nparts > 0) cuda_part_BC();
    cuda_dom_BC();
    exchange_vel(_u, _v, _w);

    cuda_update_p();
    if (nparts > 0) {
      cuda_part_BC();
      cuda_part_p_fill();
    }
    cuda_dom_BC_p(_p);
    cuda_exchange(_p, GCC);

    if (NPARTS > 0) {
      cuda_move_parts_sub();
      cuda_lamb();
      cuda_lamb_err(&iter_err, &iter_num);
    } else {
      iter_err = 0.;
    }
    iter++;
    if (iter == lamb_max_iter) break;
  }
  double mitertime = wall_seconds() - mitertimestart;
  bb_lamb_iters = iter;
  bb_lamb_err = iter_err;
  if (rank == 0 && bb_verbose && NPARTS > 0) {
    if (iter < lamb_max_iter)
      printf("  The Lamb's coefficients converged in %d iterations\n", iter);
    else
      printf("  Reached the max number of Lamb's iterations. Continuing...\n");
  }

  /* scalar inner iteration (src/bluebottle.c:293-341) */
  int scalar_iter = 0;
  real scalar_iter_err = FLT_MAX;
  double sitertime = 0.;
  if (SCALAR >= 1) {
    double sitertimestart = wall_seconds();
    while (scalar_iter_err > lamb_residual) {
      cuda_scalar_BC(_s0);
      cuda_scalar_part_BC(_s0);
      cuda_exchange(_s0, GCC);
      cuda_scalar_solve();
      cuda_exchange(_s, GCC);
      cuda_scalar_BC(_s);
      cuda_scalar_part_fill();
      cuda_exchange(_s, GCC);
      cuda_scalar_lamb();
      scalar_iter_err = cuda_scalar_lamb_err();
      scalar_iter++;
      if (scalar_iter == lamb_max_iter) break;
    }
    sitertime = wall_seconds() - sitertimestart;
    bb_scalar_iters = scalar_iter;
    bb_scalar_err = scalar_iter_err;
    cuda_store_s();
    cuda_scalar_update_part();
  }

  if (NPARTS > 0) {
    cuda_move_parts_sub();
    cuda_update_part_velocity();
    cuda_parts_internal();
    cuda_dom_BC();
    exchange_vel(_u, _v, _w);
  }

  cuda_store_u();
  dt0 = dt;
  cuda_find_dt();
  recorder_scalar("iter_etime.rec", ttime, iter, mitertime, iter_err,
    scalar_iter, sitertime, scalar_iter_err);
  recorder_out();

  if (NPARTS > 0) {
    cuda_update_part_position();
    cuda_build_cages();
    cuda_PP_init_jacobi_preconditioner();
  }

  /* restart file (src/bluebottle.c:397-398); 1 asks the caller to stop */
  if (restart_recorder_write()) return 1;

  if (dt < 1.e-10 || dt > 1.e10) {
    printf("N%d >> The solution time step has diverged. Ending simulation.\n",
      rank);
    exit(EXIT_FAILURE);
  }
  fflush(stdout);
  return 0;
}

void bluebottle_teardown(void)
{
  cuda_block();
  bb_warm_teardown();
  if (bb_parts_teardown) bb_parts_teardown();
  if (SCALAR >= 1) cuda_scalar_free();
  cuda_dom_free();
  domain_free();
}

int bluebottle_main(int argc, char *argv[])
{
  parse_cmdline_args(argc, argv);
  if (rank == 0) printf("N%d >> Running Bluebottle (MI355X)...\n", rank);
  bluebottle_setup();
  if (rank == 0) {
    printf("\n=====BLUEBOTTLE============================");
    printf("======================================\n");
    fflush(stdout);
  }
  while (ttime <= duration)
    if (bluebottle_step()) break;
  if (rank == 0) {
    printf("\nEXPD: The simulation has reached its specified duration\n");
    printf("========================================");
    printf("========================================\n\n");
  }
  bluebottle_teardown();
  return 0;
}
