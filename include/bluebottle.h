/*
 * bluebottle.h -- C ABI of the MI355X per-timestep flow path.
 *
 * This header is the drop-in boundary.  It keeps the reference's
 * global-variable calling convention: the host driver fills the structs and
 * scalars below, then calls the `void cuda_*(void)` entry points, which read
 * and write those globals.  Every entry point is `extern "C"` and is
 * implemented with hand-written HIP kernels for gfx950 in libbluebottle.so.
 *
 * Struct layouts are field-for-field those of the reference so that a host
 * compiled against the reference headers sees the same bytes:
 *   grid_info     src/domain.h:52-95
 *   dom_struct    src/domain.h:168-210
 *   BC            src/bluebottle.h:662-747
 *   gradP_struct  src/bluebottle.h:875-885
 *   g_struct      src/bluebottle.h:910-920
 *   BC_s          src/scalar.h:68-87
 *
 * Host arrays (u, v, w, p, ...) use the reference memory layouts
 * (GFX_LOC is y-fastest, GFY_LOC is z-fastest, GCC_LOC/GFZ_LOC x-fastest,
 * src/bluebottle.h:70-73).  Device arrays are x-fastest for every grid; the
 * push/pull entry points convert between the two.
 */
#ifndef BLUEBOTTLE_MI355X_H
#define BLUEBOTTLE_MI355X_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef double real;                 /* src/bluebottle.h:45-47 (-DDOUBLE) */

/* Location macros, reference host layouts (src/bluebottle.h:70-73). */
#define GCC_LOC(II, JJ, KK, S1, S2) ((II) + (JJ)*(S1) + (KK)*(S2))
#define GFX_LOC(II, JJ, KK, S1, S2) ((JJ) + (KK)*(S1) + (II)*(S2))
#define GFY_LOC(II, JJ, KK, S1, S2) ((KK) + (II)*(S1) + (JJ)*(S2))
#define GFZ_LOC(II, JJ, KK, S1, S2) ((II) + (JJ)*(S1) + (KK)*(S2))

#define PI 3.14159265358979323846
#define DOM_BUF 1                    /* src/bluebottle.h:141 */
#define CHAR_BUF_SIZE 256
#define FILE_NAME_SIZE 256
#define ROOT_DIR "."
#define INPUT_DIR "input"
#define OUTPUT_DIR "output"

/* Boundary condition types (src/bluebottle.h:218-254). */
#define PERIODIC 0
#define DIRICHLET 1
#define NEUMANN 2
#define PRECURSOR 3

/* Initial conditions (src/bluebottle.h:268-340). */
#define QUIESCENT 0
#define SHEAR 1
#define CHANNEL 2
#define TAYLOR_GREEN 3
#define TAYLOR_GREEN_3 4
#define TURB_CHANNEL 5
#define TURBULENT 6

/* Solvability plane (src/bluebottle.h:353-425). */
#define HOMOGENEOUS 10
#define WEST 0
#define EAST 1
#define SOUTH 2
#define NORTH 3
#define BOTTOM 4
#define TOP 5

/* Grid identifiers (src/bluebottle.h:437-473). */
#define GCC 0
#define GFX 1
#define GFY 2
#define GFZ 3

/* A neighbour that does not exist (the reference uses MPI_PROC_NULL). */
#define BB_PROC_NULL (-1)

/* Pressure-solver forms selected by the optional `pp_precision` key. */
#define PP_PLAIN 0                   /* FP64 Jacobi-PCG, the reference form */
#define PP_MIXED 1                   /* FP32 PCG inside FP64 defect correction */

/* Warm-start selection for the pressure solve (`pp_warmstart` key). */
#define PP_WARM_OFF 0
#define PP_WARM_LEARNED 1

typedef struct grid_info {           /* src/domain.h:52-95 */
  int is, ie, in, isb, ieb, inb;
  int js, je, jn, jsb, jeb, jnb;
  int ks, ke, kn, ksb, keb, knb;
  int _is, _ie, _isb, _ieb;
  int _js, _je, _jsb, _jeb;
  int _ks, _ke, _ksb, _keb;
  int s1, s1b, s2, s2b, s3, s3b;
  int s2_i, s2_j, s2_k, s2b_i, s2b_j, s2b_k;
} grid_info;

typedef struct dom_struct {          /* src/domain.h:168-210 */
  grid_info Gcc;
  grid_info Gfx;
  grid_info Gfy;
  grid_info Gfz;
  real xs, xe, xl; int xn; real dx;
  real ys, ye, yl; int yn; real dy;
  real zs, ze, zl; int zn; real dz;
  int rank;
  int e, w, n, s, t, b;
  int I, Is, Ie, In;
  int J, Js, Je, Jn;
  int K, Ks, Ke, Kn;
  int S1, S2, S3;
} dom_struct;

typedef struct BC {                  /* src/bluebottle.h:662-747 */
  int pW, pE, pS, pN, pB, pT;
  int uW; real uWDm, uWD, uWDa;
  int uE; real uEDm, uED, uEDa;
  int uS; real uSDm, uSD, uSDa;
  int uN; real uNDm, uND, uNDa;
  int uB; real uBDm, uBD, uBDa;
  int uT; real uTDm, uTD, uTDa;
  int vW; real vWDm, vWD, vWDa;
  int vE; real vEDm, vED, vEDa;
  int vS; real vSDm, vSD, vSDa;
  int vN; real vNDm, vND, vNDa;
  int vB; real vBDm, vBD, vBDa;
  int vT; real vTDm, vTD, vTDa;
  int wW; real wWDm, wWD, wWDa;
  int wE; real wEDm, wED, wEDa;
  int wS; real wSDm, wSD, wSDa;
  int wN; real wNDm, wND, wNDa;
  int wB; real wBDm, wBD, wBDa;
  int wT; real wTDm, wTD, wTDa;
  real dsW, dsE, dsS, dsN, dsB, dsT;
} BC;

typedef struct gradP_struct {        /* src/bluebottle.h:875-885 */
  real x, xm, xa;
  real y, ym, ya;
  real z, zm, za;
} gradP_struct;

typedef struct g_struct {            /* src/bluebottle.h:910-920 */
  real x, xm, xa;
  real y, ym, ya;
  real z, zm, za;
} g_struct;

typedef struct BC_s {                /* src/scalar.h:68-87 */
  int sW; real sWD, sWN;
  int sE; real sED, sEN;
  int sN; real sND, sNN;
  int sS; real sSD, sSN;
  int sT; real sTD, sTN;
  int sB; real sBD, sBN;
} BC_s;

/* ---- Global state (defined in the library; src/bluebottle.c:440-576) ---- */
extern int rank;                     /* always 0: one GPU holds the domain */
extern int nprocs;                   /* always 1 */
extern dom_struct DOM;
extern dom_struct *dom;
extern BC bc;
extern gradP_struct gradP;
extern g_struct g;
extern BC_s bc_s;

extern real dt, dt0, CFL, duration, ttime;
extern int stepnum;
extern real rho_f, mu, nu;
extern real pp_residual;
extern int pp_max_iter;
extern int lamb_max_iter;
extern real lamb_residual, lamb_relax, lamb_cut;
extern real osci_f;
extern int init_cond;
extern int out_plane;
extern real v_bc_tdelay, p_bc_tdelay, g_bc_tdelay;
extern real pid_int, pid_back, Kp, Ki, Kd;
extern real turbA, turbl, turb_k0;
extern long int cpumem, gpumem;

/* Scalar transport parameters (src/domain.c:166-206). */
extern int SCALAR;
extern real s_D, s_k, s_beta, s_ref, lamb_cut_scalar, s_init, s_init_rand;

/* ----------------- Particles (src/particle.h, src/physalis.h) ------------- */
#define NNODES 26                 /* Lebedev nodes, src/particle.h:42 */
#define MAX_COEFFS 15             /* src/particle.h:55 */
#define S_MAX_COEFFS 25           /* src/particle.h:70 */
#define MAX_NEIGHBORS 12          /* src/particle.h:83 */
#define NCOEFFS 15                /* src/physalis.h:52 */
#define DIV_ST -1e-10             /* src/physalis.h:40 */
/* Node codes of quadrature nodes cut by a wall (src/physalis.h:64-196). */
#define WEST_WALL_D -100
#define WEST_WALL_N -101
#define EAST_WALL_D -110
#define EAST_WALL_N -111
#define SOUTH_WALL_D -120
#define SOUTH_WALL_N -121
#define NORTH_WALL_D -130
#define NORTH_WALL_N -131
#define BOTTOM_WALL_D -140
#define BOTTOM_WALL_N -141
#define TOP_WALL_D -150
#define TOP_WALL_N -151

/* One particle, member for member as src/particle.h:343-446 (the host and
 * device arrays hold this layout; 3128 bytes on x86-64). */
typedef struct part_struct {
  int N;
  real r;
  real x, y, z;
  real u, v, w;
  real u0, v0, w0;
  real udot, vdot, wdot;
  real udot0, vdot0, wdot0;
  real axx, axy, axz, ayx, ayy, ayz, azx, azy, azz;
  real ox, oy, oz;
  real ox0, oy0, oz0;
  real oxdot, oydot, ozdot;
  real oxdot0, oydot0, ozdot0;
  real Fx, Fy, Fz;
  real Lx, Ly, Lz;
  real aFx, aFy, aFz;
  real aLx, aLy, aLz;
  real kFx, kFy, kFz;
  real iFx, iFy, iFz;
  real iLx, iLy, iLz;
  int nodes[NNODES];
  real rho;
  real E;
  real sigma;
  int order;
  real rs;
  int ncoeff;
  real spring_k;
  real spring_x, spring_y, spring_z;
  real spring_l;
  int translating;
  int rotating;
  real St[MAX_NEIGHBORS];
  int iSt[MAX_NEIGHBORS];
  real e_dry;
  real coeff_fric;
  real pnm_re[MAX_COEFFS];
  real pnm_im[MAX_COEFFS];
  real phinm_re[MAX_COEFFS];
  real phinm_im[MAX_COEFFS];
  real chinm_re[MAX_COEFFS];
  real chinm_im[MAX_COEFFS];
  real pnm_re0[MAX_COEFFS];
  real pnm_im0[MAX_COEFFS];
  real phinm_re0[MAX_COEFFS];
  real phinm_im0[MAX_COEFFS];
  real chinm_re0[MAX_COEFFS];
  real chinm_im0[MAX_COEFFS];
  int ncoll_part;
  int ncoll_wall;
  real s;
  int update;
  real srs;
  real q;
  real cp;
  int sorder;
  int sncoeff;
  real anm_re[S_MAX_COEFFS];
  real anm_im[S_MAX_COEFFS];
  real anm_re0[S_MAX_COEFFS];
  real anm_im0[S_MAX_COEFFS];
} part_struct;

/* Bin grid of the collision search (src/particle.h:545-650). */
typedef struct bin_grid_info {
  int _is, _ie, _isb, _ieb, in, inb;
  int _js, _je, _jsb, _jeb, jn, jnb;
  int _ks, _ke, _ksb, _keb, kn, knb;
  int s1, s2, s3, s1b, s2b, s3b;
  int s2_i, s2_j, s2_k, s2b_i, s2b_j, s2b_k;
} bin_grid_info;

typedef struct bin_struct {
  bin_grid_info Gcc;
  real xs, xe, xl;
  int xn;
  real dx;
  real ys, ye, yl;
  int yn;
  real dy;
  real zs, ze, zl;
  int zn;
  real dz;
} bin_struct;

/* Particle globals (src/particle.c:27-50). */
extern int NPARTS;                   /* particles in the domain */
extern int nparts;                   /* particles held (all of them here) */
extern part_struct *parts;           /* host copy */
extern part_struct *_parts;          /* device copy */
extern bin_struct bins;
extern real interaction_length_ratio;
extern real bin_size;
extern int ncoeffs_max;
extern int s_ncoeffs_max;
extern real volume_fraction, total_parts_mass, total_parts_volume, rho_avg;

/* Optional trailing config keys (absent from the reference grammar; a file
 * without them parses exactly as before). */
extern real fixed_dt;                /* > 0: use this dt every step */
extern unsigned long long perturb_seed;
extern real perturb_amp;             /* > 0: seeded uniform(-a, a) on u, v, w */
extern int pp_precision;             /* PP_PLAIN or PP_MIXED */
extern int pp_warmstart;             /* PP_WARM_OFF or PP_WARM_LEARNED */
extern real pp_inner_tol;            /* relative tolerance of a mixed inner solve */

/* Host arrays in reference layouts (src/bluebottle.c:482-490). */
extern real *p, *u, *v, *w;
extern real *p0, *u0, *v0, *w0;
extern real *phi;
extern real *u_star, *v_star, *w_star;
extern real *f_x, *f_y, *f_z;
extern real *conv_u, *conv_v, *conv_w;
extern real *conv0_u, *conv0_v, *conv0_w;
extern real *diff_u, *diff_v, *diff_w;
extern real *diff0_u, *diff0_v, *diff0_w;
extern int *flag_u, *flag_v, *flag_w;
extern int *phase, *phase_shell;

/* Device arrays (x-fastest, ghosted; src/bluebottle.c:492-515). */
extern real *_p, *_u, *_v, *_w;
extern real *_p0, *_u0, *_v0, *_w0;
extern real *_phi;
extern real *_invM;
extern real *_u_star, *_v_star, *_w_star;
extern real *_conv_u, *_conv_v, *_conv_w;
extern real *_conv0_u, *_conv0_v, *_conv0_w;
extern real *_diff_u, *_diff_v, *_diff_w;
extern real *_diff0_u, *_diff0_v, *_diff0_w;
extern real *_rhs_p, *_r_q, *_z_q, *_p_q, *_pb_q, *_Apb_q;
extern int *_flag_u, *_flag_v, *_flag_w;
extern int *_phase, *_phase_shell;

/* Per-solve statistics of the last pressure solve (the values the reference
 * writes to record/solver_expd.rec, src/cuda_solver.cu:236-240). */
typedef struct pp_stats {
  int niter;          /* PCG iterations (inner iterations summed for mixed) */
  int nouter;         /* defect-correction sweeps (1 for the plain form) */
  real resid;         /* sqrt((r, M^-1 r) / (b, b)) at exit, FP64 */
  real true_resid;    /* ||b - A phi||_2 / ||b||_2 at exit, FP64 */
  real sp_rhs;        /* (b, b) */
  double etime;       /* seconds */
} pp_stats;
extern pp_stats pp_last;

/* ---------------- Host C functions (src/domain.h:269-462) ---------------- */
void parse_cmdline_args(int argc, char *argv[]);   /* src/domain.c:33-70 */
void domain_read_input(void);                      /* src/domain.c:72-916 */
void domain_fill(void);                            /* src/domain.c:918-1487 */
void compute_vel_BC(void);                         /* src/domain.c:1931-2082 */
void domain_init_fields(void);                     /* src/domain.c:2083-2142 */
void domain_free(void);
void rng_init(unsigned long long seed);            /* src/rng.c:30-41 */
unsigned long long rng_uint64(void);               /* src/rng.c:44-51 */
double rng_dbl(void);                              /* src/rng.c:78-81 */
void recorder_PP_init(const char *name);           /* src/recorder.c:157-188 */
/* Restart (src/domain.c:3005-3261, src/recorder.c:43-98, 338-391): the
 * reference's binary restart.config-<rank> in the reference layouts. */
extern int use_restart, restart_stop, rec_restart_complete;
extern real rec_restart_dt, rec_cgns_flow_dt, rec_cgns_part_dt, rec_vtk_dt;
extern int rec_vtk_stepnum_out;
extern real rec_cgns_flow_ttime_out, rec_cgns_part_ttime_out,
  rec_vtk_ttime_out;
void recorder_read_config(void);                   /* src/recorder.c:43-98 */
void out_restart(void);                            /* src/domain.c:3005-3092 */
void in_restart(void);                             /* src/domain.c:3094-3261 */
int restart_recorder_write(void);                  /* src/recorder.c:338-391 */
void recorder_PP(const char *name, int niter, real resid, real etime); /* :190-221 */
void recorder_scalar_init(const char *name);       /* src/scalar.c:226-262 */
void recorder_scalar(const char *name, real ttime, int mniter,
  real mitertime, real merr, int sniter, real sitertime,
  real serr);                                      /* src/scalar.c:264-295 */
/* ParaView output under <case>/output (src/vtk.c:29-405): out.pvd, and per
 * output out_<n>.pvtr/.vtr (p, phase, phase_shell, vel) and out_<n>.pvtp/.vtp
 * (particle centres, n, r, pvel).  recorder_out pulls and writes when
 * VTK_FLOW_FIELD in record.config says an output is due. */
void recorder_out(void);                           /* src/recorder.c:91-155 */
void vtk_recorder_init(void);                      /* src/vtk.c:29-38 */
void vtk_recorder_write(void);                     /* src/vtk.c:40-60 */
void init_VTK(void);                               /* src/vtk.c:62-92 */
void out_VTK(void);                                /* src/vtk.c:94-129 */
void dom_out_VTK(void);                            /* src/vtk.c:131-309 */
void part_out_VTK(void);                           /* src/vtk.c:311-404 */

/* Driver: src/bluebottle.c:28-436 split into its phases so that a caller can
 * step the solver one timestep at a time.  bluebottle_main() is the whole
 * program (setup, `while (ttime <= duration)` loop, teardown). */
int bluebottle_setup(void);          /* :31-142, returns 0 on success */
int bluebottle_step(void);           /* one pass of :168-407; 1 when a
                                        restart file was written and
                                        RESTART_STOP asks to end the run */
void bluebottle_teardown(void);      /* :416-430 */
int bluebottle_main(int argc, char *argv[]);

/* --------------- Device entry points (libbluebottle.so) ------------------ */
int cuda_device_count(void);                 /* src/bluebottle.h:2604 */
void cuda_device_init(int device);           /* :2616 */
void cuda_block(void);                       /* :2642, device synchronise */
void cuda_dom_malloc_host(void);             /* src/cuda_bluebottle.cu:76-141 */
void cuda_dom_malloc_dev(void);              /* src/cuda_bluebottle.cu:144-432 */
void cuda_dom_push(void);                    /* :443-490, host layouts -> device */
void cuda_dom_pull(void);                    /* :492-509, device -> host layouts */
void cuda_dom_pull_phase(void);              /* :1455, phase and phase_shell */
void cuda_dom_pull_debug(void);              /* :1464-1505, phi, u*, conv, diff,
                                                flags */
void cuda_dom_pull_restart(void);            /* :1507-1530, p0, u0, conv0, diff0 */
void cuda_compute_turb_forcing(void);        /* :1985-2093; exits with a message
                                                when init_cond is TURBULENT
                                                (not supported), else nothing */
void cuda_wall_shear_stress(void);           /* :2764-2835, output/wss.dat */
void printMemInfo(void);                     /* src/cuda_scalar.cu:972-986 */
void cuda_blocks_init(void);                 /* :511-1006 (launch shapes) */
void cuda_update_bc(void);                   /* :435-440 */
void cuda_dom_BC(void);                      /* :1218-1438 + periodic planes */
void cuda_dom_BC_star(void);                 /* :2111-2311 + periodic planes */
void cuda_dom_BC_p(real *array);             /* :2538-2589 + periodic planes */
void cuda_build_cages(void);                 /* src/cuda_particle.cu:1516-1646 */
void cuda_find_dt(void);                     /* :1753-1808 */
void cuda_compute_forcing(void);             /* :1811-1982 (constant gradP) */
void cuda_U_star(void);                      /* :2095-2108 */
void cuda_solvability(void);                 /* :2313-2492 */
void cuda_project(void);                     /* :2495-2503 */
void cuda_update_p(void);                    /* :2506-2535 */
void cuda_store_u(void);                     /* :2592-2616 (pointer rotation) */
void cuda_dom_free(void);                    /* :2619 */
void cuda_PP_init_jacobi_preconditioner(void);  /* src/cuda_solver.cu:32-36 */
void cuda_PP_cg(void);                       /* src/cuda_solver.cu:39-300 */
void cuda_PP_cg_noparts(void);               /* src/cuda_solver.cu:574-761 */
/* No counterpart in the reference: the FP32-inner / FP64-outer form. */
void cuda_PP_cg_mixed(void);
/* No counterpart in the reference: cuda_PP_cg (plain FP64 Jacobi-PCG, the
 * same kernels and bits) for callers that cannot allocate device memory
 * (code under graph capture, serving loops).  Its scratch (CG state, the
 * second search direction, tile partials) lives in a static 64 MiB device
 * array, which limits the grid to about 7.9 million cells including ghosts;
 * calls that overlap on one GPU race on it.  Returns 0, or 1 (no scratch
 * address on this device), 2 (grid too large) or 3 (learned initial guess
 * selected: its network allocates) before doing any work. */
int cuda_PP_cg_noalloc(void);
/* Particles and Physalis.  On one GPU every particle is held once; a
 * particle whose cage or quadrature nodes cross a periodic boundary is
 * handled through its periodic image positions x +- L (the same values the
 * reference's ghost copies carry, src/particle_kernel.cu:1166-1210). */
void parts_read_input(void);                 /* src/particle.c:66-391 */
void parts_init(void);                       /* src/particle.c:393-554 */
void init_bins(void);                        /* src/particle.c:744-834 */
void parts_free(void);                       /* src/particle.c:876-879 */
void cuda_part_malloc_host(void);            /* src/cuda_particle.cu:47-55, host
                                                phase and phase_shell */
void cuda_part_malloc_dev(void);             /* src/cuda_particle.cu */
void cuda_part_push(void);                   /* host parts -> device */
void cuda_part_pull(void);                   /* device parts -> host */
void cuda_part_pull_debug(void);             /* src/cuda_particle.cu:254-265 */
void cuda_part_free(void);                   /* :1268-1284 */
/* src/cuda_particle.cu:287-1186 move particles between ranks and copy
 * ghosts near periodic faces; on one GPU the kernels evaluate the periodic
 * images themselves and cuda_update_part_position wraps centres, so these
 * have nothing to do. */
void cuda_transfer_parts_i(void);
void cuda_transfer_parts_j(void);
void cuda_transfer_parts_k(void);
void cuda_init_physalis(void);               /* src/cuda_physalis.cu:61-135 */
void cuda_part_BC(void);                     /* src/cuda_particle.cu:1649-1663 */
void cuda_part_BC_star(void);                /* :1665-1679 */
void cuda_part_BC_p(void);                   /* :1681-1686 */
void cuda_part_p_fill(void);                 /* :1688-1693 */
void cuda_parts_internal(void);              /* :1695-1706 */
void cuda_lamb(void);                        /* src/cuda_physalis.cu:137-220 */
void cuda_lamb_err(real *error, int *number);   /* :859-903 */
void cuda_move_parts_sub(void);              /* src/cuda_particle.cu:1335-1473 */
void cuda_update_part_velocity(void);        /* :1477-1496 */
void cuda_update_part_position(void);        /* :1498-1513 + periodic wrap */
/* Learned initial guess of the pressure solve (pp_warmstart LEARNED); no
 * counterpart in the reference, which starts every solve from phi = 0
 * (src/solver_kernel.cu:258-283).  infer places alpha * U-Net(b, p0) in
 * _phi before a solve (returns 1 if it did); train feeds the finished
 * solve to the network while it is learning. */
int cuda_PP_warmstart_infer(void);
void cuda_PP_warmstart_train(void);
/* Scalar transport (src/scalar.h, src/cuda_scalar.cu): host arrays in the
 * reference layout and device arrays, both Gcc-sized (src/scalar.c:41-53). */
extern real *s, *s0, *s_conv, *s_conv0, *s_diff, *s_diff0;
extern real *_s, *_s0, *_s_conv, *_s_conv0, *_s_diff, *_s_diff0;
void scalar_init_fields(void);               /* src/scalar.c:55-88 */
void scalar_part_init(void);                 /* src/scalar.c:90-115 */
void cuda_scalar_malloc_host(void);          /* src/cuda_scalar.cu:18-32 */
void cuda_scalar_malloc_dev(void);           /* :35-61 */
void cuda_scalar_push(void);                 /* :64-99 */
void cuda_scalar_pull(void);                 /* :102-106 */
void cuda_scalar_pull_debug(void);           /* :108-115, s_conv, s_diff */
void cuda_scalar_pull_restart(void);         /* :117-126, s0, s_conv0, s_diff0 */
void cuda_scalar_free(void);                 /* :129-146 */
void cuda_compute_boussinesq(void);          /* :149-154 */
void cuda_scalar_BC(real *array);            /* :157-237 */
void cuda_scalar_part_BC(real *array);       /* :240-244 */
void cuda_scalar_part_fill(void);            /* :247-250 */
void cuda_scalar_solve(void);                /* :253-257 */
void cuda_scalar_lamb(void);                 /* :866-911 */
real cuda_scalar_lamb_err(void);             /* :914-946 */
void cuda_store_s(void);                     /* :949-957 */
void cuda_scalar_update_part(void);          /* :959-969 */
/* Periodic ghost refresh of one field; replaces mpi_cuda_exchange_G*
 * (src/mpi_comm.c:257-405) on a single GPU.  `grid` is GCC/GFX/GFY/GFZ. */
void cuda_exchange(real *array, int grid);
/* The MPI layer (src/mpi_comm.h:267-631) on one rank: no neighbour and no
 * window; the exchanges refresh the periodic ghosts on the GPU
 * (cuda_exchange).  MPI_Init/MPI_Finalize run when the program links an
 * MPI library (found at run time; the library does not link MPI). */
void mpi_startup(int argc, char *argv[]);    /* src/mpi_comm.c:39-79 */
void mpi_dom_comm_init(void);                /* :81-96 */
void mpi_parts_init(void);                   /* :471 */
void mpi_cuda_exchange_Gcc(real *array);     /* :257-405 */
void mpi_cuda_exchange_Gfx(real *array);
void mpi_cuda_exchange_Gfy(real *array);
void mpi_cuda_exchange_Gfz(real *array);
void mpi_free(void);                         /* :1262-1309 */
void mpi_end(void);                          /* :1311-1315 */
/* Exchange buffers and window handles that src/scalar.c:115-224 names
 * (src/cuda_physalis.cu:41-58, src/mpi_comm.c:71-136); unused on one rank. */
extern real *_sum_send_e, *_sum_send_w, *_sum_send_n, *_sum_send_s,
  *_sum_send_t, *_sum_send_b;
extern real *_sum_recv_e, *_sum_recv_w, *_sum_recv_n, *_sum_recv_s,
  *_sum_recv_t, *_sum_recv_b;
extern void *parts_recv_win_e, *parts_recv_win_w, *parts_recv_win_n,
  *parts_recv_win_s, *parts_recv_win_t, *parts_recv_win_b;
void cuda_check_errors(int line);            /* src/cuda_bluebottle.cu:37-42 */

#ifdef __cplusplus
}
#endif

#endif
