/*
 * Definitions of the global state that the C ABI passes between the host
 * driver and the device entry points (src/bluebottle.c:440-576,
 * src/particle.c:27-50, src/scalar.c globals, src/mpi_comm.c rank/nprocs).
 *
 * A host program that defines these symbols itself (as the reference's
 * bluebottle.c does) interposes its own copies over these ones at load time.
 */
#include "bluebottle.h"

int rank = 0;
int nprocs = 1;

dom_struct DOM;
dom_struct *dom = NULL;
BC bc;
gradP_struct gradP;
g_struct g;
BC_s bc_s;

real dt, dt0, CFL, duration, ttime;
int stepnum;
real rho_f, mu, nu;
real pp_residual;
int pp_max_iter;
int lamb_max_iter;
real lamb_residual, lamb_relax, lamb_cut;
real osci_f;
int init_cond;
int out_plane;
real v_bc_tdelay, p_bc_tdelay, g_bc_tdelay;
real pid_int, pid_back, Kp, Ki, Kd;
real turbA, turbl, turb_k0;
long int cpumem, gpumem;

int SCALAR;
real s_D, s_k, s_beta, s_ref, lamb_cut_scalar, s_init, s_init_rand;

int NPARTS = 0;
int nparts = 0;

real fixed_dt = 0.;
unsigned long long perturb_seed = 0ULL;
real perturb_amp = 0.;
int pp_precision = PP_PLAIN;
int pp_warmstart = PP_WARM_OFF;
real pp_inner_tol = 1e-3;

real *p, *u, *v, *w;
real *p0, *u0, *v0, *w0;
real *phi;
real *u_star, *v_star, *w_star;
real *f_x, *f_y, *f_z;
real *conv_u, *conv_v, *conv_w;
real *conv0_u, *conv0_v, *conv0_w;
real *diff_u, *diff_v, *diff_w;
real *diff0_u, *diff0_v, *diff0_w;
int *flag_u, *flag_v, *flag_w;
int *phase, *phase_shell;

real *_p, *_u, *_v, *_w;
real *_p0, *_u0, *_v0, *_w0;
real *_phi;
real *_invM;
real *_u_star, *_v_star, *_w_star;
real *_conv_u, *_conv_v, *_conv_w;
real *_conv0_u, *_conv0_v, *_conv0_w;
real *_diff_u, *_diff_v, *_diff_w;
real *_diff0_u, *_diff0_v, *_diff0_w;
real *_rhs_p, *_r_q, *_z_q, *_p_q, *_pb_q, *_Apb_q;
int *_flag_u, *_flag_v, *_flag_w;
int *_phase, *_phase_shell;

pp_stats pp_last;
