/*
 * FP64 host oracle of the Bluebottle per-timestep flow path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the GPU library is compared
 * against; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it, and never as the thing measured or shipped.
 *
 * Arrays are x-fastest with one ghost layer: element (i, j, k) of a grid
 * with (ni, nj, nk) interior points sits at i + j*(ni+2) + k*(ni+2)*(nj+2).
 */
#ifndef BB_ORACLE_H
#define BB_ORACLE_H

/* Set up from <dir>/input/flow.config (+ decomp.config); 0 on success. */
int orc_setup(const char *dir);
void orc_teardown(void);
/* One full timestep; returns the PCG iteration count of its (last) solve. */
int orc_step(void);

/* Named arrays: p p0 phi u u0 v v0 w w0 u_star v_star w_star conv_u conv_v
 * conv_w conv0_u conv0_v conv0_w diff_u diff_v diff_w diff0_u diff0_v
 * diff0_w rhs invM.  Returns NULL for an unknown name. */
double *orc_field(const char *name);
/* Integer arrays: flag_u flag_v flag_w phase phase_shell. */
int *orc_ifield(const char *name);
/* Interior extents of a grid (0 = Gcc, 1 = Gfx, 2 = Gfy, 3 = Gfz). */
void orc_dims(int grid, int *ni, int *nj, int *nk);

/* Scalars: dt dt0 ttime stepnum niter resid true_resid sp_rhs lamb_iters
 * lamb_err nparts */
double orc_scalar(const char *name);
void orc_set_scalar(const char *name, double v);

/* Individual phases of the step (for per-function parity tests). */
void orc_forcing(void);
void orc_vel_bc_ramp(void);
void orc_u_star(void);
void orc_bc_star(void);
void orc_bc_all(void);
void orc_bc_p(double *a);
void orc_exchange(double *a, int grid);
void orc_solvability(void);
int orc_pcg(void);
void orc_project(void);
void orc_update_p(void);
void orc_store_u(void);
void orc_find_dt(void);
void orc_jacobi(void);
void orc_build_flags(void);
/* (-A) x over interior cells with the face flags (ghosts of x as given). */
void orc_apply_A(const double *x, double *Ax);

/* Particles (input/part.config).  orc_build_cages rebuilds phase,
 * phase_shell and the face flags. */
void orc_build_cages(void);
int orc_nparts(void);
/* 126 values: x y z u v w ox oy oz udot vdot wdot oxdot oydot ozdot
 * Fx Fy Fz Lx Ly Lz iFx iFy iFz iLx iLy iLz axx..azz, then 15 each of
 * pnm_re pnm_im phinm_re phinm_im chinm_re chinm_im. */
int orc_part_state(int n, double *out);
int orc_part_set_coeffs(int n, const double *c90);
int orc_part_set_motion(int n, const double *m15);
int orc_part_contacts(int n, int *iSt, double *St, int *ncoll2);
void orc_collide_once(void);
void orc_move_parts_sub(void);
int orc_part_scalar(int n, double *s_q_anm52);
int orc_part_set_scalar(int n, const double *s_q_anm52);
void orc_scalar_lamb_once(void);
void orc_scalar_part_bc_s(void);
/* Lamb velocity (U[0..2]) and n = 0 pressure term (U[3]) of particle n at
 * (x, y, z) relative to its centre. */
void orc_lamb_eval(int n, double x, double y, double z, double *U);
/* One Physalis update (quadrature, coefficients, forces) on the current
 * fields; returns the Lamb error. */
double orc_lamb_once(void);
void orc_lebedev(double *theta, double *phi, double *w);

/* Canonical reductions, exposed so tests can pin them. */
double orc_canon_sum(const double *terms, int nx, int ny, int nz);
double orc_canon_linear(const double *terms, long n);

#endif
