/*
 * Scalar transport in the FP64 host oracle (included by oracle.c).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows:
 *   initial field       src/scalar.c:55-88
 *   boundary cells      src/scalar_kernel.cu:4-158, src/cuda_scalar.cu:157-237
 *   Boussinesq forcing  src/scalar_kernel.cu:160-206, src/cuda_scalar.cu:149-154
 *   transport step      src/scalar_kernel.cu:208-301
 *   store               src/cuda_scalar.cu:949-957
 *   step order          src/bluebottle.c:189-192, 293-341
 * Particle heat transfer is not restated (the library refuses scalar runs
 * with particles).
 */

static int O_scalar;
static double O_sD, O_sk, O_sinit, O_sinit_rand;
static int O_sbc_type[6];
static double O_sbc_val[6];
static double *s_, *s0_, *sc_, *sc0_, *sd_, *sd0_;

static void o_scalar_parse(void)
{
  char rest[MAXL], word[64];
  O_scalar = find_line("SCALAR", rest) ? atoi(rest) : 0;
  O_sD = find_line("s_D", rest) ? atof(rest) : 0.;
  O_sk = find_line("s_k", rest) ? atof(rest) : 0.;
  O_sinit = O_sinit_rand = 0.;
  if (find_line("init_s", rest)) sscanf(rest, "%lf %lf", &O_sinit, &O_sinit_rand);
  const char *F = "WESNBT";
  for (int f6 = 0; f6 < 6; f6++) {
    char key[32];
    snprintf(key, sizeof(key), "bc_s.s%c", F[f6]);
    O_sbc_type[f6] = PERIODIC;
    O_sbc_val[f6] = 0.;
    if (!find_line(key, rest)) continue;
    double v = 0.;
    int n = sscanf(rest, "%63s %lf", word, &v);
    if (strcmp(word, "DIRICHLET") == 0) O_sbc_type[f6] = DIRICHLET;
    else if (strcmp(word, "NEUMANN") == 0) O_sbc_type[f6] = NEUMANN;
    O_sbc_val[f6] = n >= 2 ? v : 0.;
  }
}

static void o_scalar_alloc(void)
{
  s_ = dal(O.cc.n); s0_ = dal(O.cc.n); sc_ = dal(O.cc.n);
  sc0_ = dal(O.cc.n); sd_ = dal(O.cc.n); sd0_ = dal(O.cc.n);
}

static void o_scalar_free(void)
{
  double *d[] = {s_, s0_, sc_, sc0_, sd_, sd0_};
  for (size_t i = 0; i < sizeof(d) / sizeof(d[0]); i++) free(d[i]);
  s_ = s0_ = sc_ = sc0_ = sd_ = sd0_ = NULL;
}

/* every cell, ghosts included: the C library generator seeded with the
 * cell's index */
static void o_scalar_init(void)
{
  for (int k = 0; k <= O.nz + 1; k++)
    for (int j = 0; j <= O.ny + 1; j++)
      for (int i = 0; i <= O.nx + 1; i++) {
        long c = IDX(O.cc, i, j, k);
        srand((unsigned) c);
        s_[c] = O_sinit + O_sinit_rand * (2. * rand() / (double) RAND_MAX - 1.);
        s0_[c] = s_[c];
      }
}

/* ghost cells of the Dirichlet and Neumann faces, interior rows only */
static void o_scalar_bc(double *a)
{
  const double h[3] = {O.dx, O.dy, O.dz};
  for (int f6 = 0; f6 < 6; f6++) {
    int type = O_sbc_type[f6];
    if (type != DIRICHLET && type != NEUMANN) continue;
    int dir = f6 / 2, hi = f6 % 2;
    double val = O_sbc_val[f6];
    int n1 = dir == 0 ? O.ny : O.nx, n2 = dir == 2 ? O.ny : O.nz;
    int e = dir == 0 ? O.nx : (dir == 1 ? O.ny : O.nz);
    for (int t2 = 1; t2 <= n2; t2++)
      for (int t1 = 1; t1 <= n1; t1++) {
        int gi[3], ii[3];
        int g = hi ? e + 1 : 0, in = hi ? e : 1;
        if (dir == 0) { gi[0] = g; gi[1] = t1; gi[2] = t2; }
        else if (dir == 1) { gi[0] = t1; gi[1] = g; gi[2] = t2; }
        else { gi[0] = t1; gi[1] = t2; gi[2] = g; }
        ii[0] = gi[0]; ii[1] = gi[1]; ii[2] = gi[2];
        ii[dir] = in;
        long G = IDX(O.cc, gi[0], gi[1], gi[2]);
        long I = IDX(O.cc, ii[0], ii[1], ii[2]);
        if (type == DIRICHLET) a[G] = 2. * val - a[I];
        else if (hi) a[G] = a[I] + h[dir] * val;
        else a[G] = a[I] - val * h[dir];
      }
  }
}

/* f += -g s_beta (mean of the two cells beside the face - s_ref) on the
 * faces 1..n+1 along the component and every row across it */
static void o_boussinesq(void)
{
  double bg[3];
  for (int c = 0; c < 3; c++) bg[c] = -O_g[c] * O_sbeta;
  for (int k = 0; k <= O.nz + 1; k++) for (int j = 0; j <= O.ny + 1; j++)
    for (int i = 1; i <= O.nx + 1; i++) {
      long c1 = IDX(O.cc, i, j, k), c0 = c1 - 1;
      fx_[IDX(O.fx, i, j, k)] += bg[0] * (0.5*(s_[c0] + s_[c1]) - O_sref);
    }
  for (int k = 0; k <= O.nz + 1; k++) for (int j = 1; j <= O.ny + 1; j++)
    for (int i = 0; i <= O.nx + 1; i++) {
      long c1 = IDX(O.cc, i, j, k), c0 = c1 - O.cc.sy;
      fy_[IDX(O.fy, i, j, k)] += bg[1] * (0.5*(s_[c0] + s_[c1]) - O_sref);
    }
  for (int k = 1; k <= O.nz + 1; k++) for (int j = 0; j <= O.ny + 1; j++)
    for (int i = 0; i <= O.nx + 1; i++) {
      long c1 = IDX(O.cc, i, j, k), c0 = c1 - O.cc.sz;
      fz_[IDX(O.fz, i, j, k)] += bg[2] * (0.5*(s_[c0] + s_[c1]) - O_sref);
    }
}

/* one Adams-Bashforth step of ds/dt + div(u s) = D lap s with the start
 * of step velocities u0, v0, w0; s changes in fluid cells only */
static void o_scalar_solve(void)
{
  double ab0 = 0.5 * O.dt / O.dt0, ab = 1. + ab0;
  double ddx = 1. / O.dx, ddy = 1. / O.dy, ddz = 1. / O.dz;
  #pragma omp parallel for schedule(static)
  for (int k = 1; k <= O.nz; k++)
    for (int j = 1; j <= O.ny; j++)
      for (int i = 1; i <= O.nx; i++) {
        long c = IDX(O.cc, i, j, k);
        double s111 = s0_[c];
        double s011 = s0_[c - 1], s211 = s0_[c + 1];
        double s101 = s0_[c - O.cc.sy], s121 = s0_[c + O.cc.sy];
        double s110 = s0_[c - O.cc.sz], s112 = s0_[c + O.cc.sz];
        double u0w = u0_[IDX(O.fx, i, j, k)], u0e = u0_[IDX(O.fx, i + 1, j, k)];
        double v0s = v0_[IDX(O.fy, i, j, k)], v0n = v0_[IDX(O.fy, i, j + 1, k)];
        double w0b = w0_[IDX(O.fz, i, j, k)], w0t = w0_[IDX(O.fz, i, j, k + 1)];
        double fxs = (u0e * 0.5 * (s211 + s111) - u0w * 0.5 * (s111 + s011)) * ddx;
        double fys = (v0n * 0.5 * (s121 + s111) - v0s * 0.5 * (s111 + s101)) * ddy;
        double fzs = (w0t * 0.5 * (s112 + s111) - w0b * 0.5 * (s111 + s110)) * ddz;
        double conv = fxs + fys + fzs;
        sc_[c] = conv;
        double lx = O_sD * (s211 - 2.*s111 + s011) * ddx * ddx;
        double ly = O_sD * (s121 - 2.*s111 + s101) * ddy * ddy;
        double lz = O_sD * (s112 - 2.*s111 + s110) * ddz * ddz;
        double diff = lx + ly + lz;
        sd_[c] = diff;
        double rate = O.dt0 > 0 ? (ab*diff - ab0*sd0_[c]) - (ab*conv - ab0*sc0_[c])
                                : diff - conv;
        double next = s111 + O.dt * rate;
        s_[c] += (next - s_[c]) * (phase_[c] == -1);
      }
}

static void o_store_s(void)
{
  memcpy(s0_, s_, O.cc.n * sizeof(double));
  memcpy(sc0_, sc_, O.cc.n * sizeof(double));
  memcpy(sd0_, sd_, O.cc.n * sizeof(double));
}

/* the scalar iteration after the Lamb loop; without particles its error is
 * DBL_MIN after the first pass */
static void o_scalar_step(void)
{
  int it = 0;
  double err = FLT_MAX;
  while (err > O_lamb_residual) {
    o_scalar_bc(s0_);
    orc_exchange(s0_, 0);
    o_scalar_solve();
    orc_exchange(s_, 0);
    o_scalar_bc(s_);
    orc_exchange(s_, 0);
    err = DBL_MIN;
    it++;
    if (it == O_lamb_max_iter) break;
  }
  o_store_s();
}
