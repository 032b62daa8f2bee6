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
 *   particle heat       src/scalar.c:90-115, src/scalar_kernel.cu:303-577,
 *                       1143-1360, src/cuda_scalar.cu:866-969
 *
 * Pinning (tests/test_oracle_scalar.py): Fourier-mode decay, conservation
 * of the flux form, steady wall profiles, the Boussinesq force of a
 * uniform field, the diffusivity limit of the time step; particle heat by
 * exact quadrature of a linear field, the conducting-sphere flux 4 pi a
 * dT, the series in particle cells and the particle temperature update.
 */

static double O_sD, O_sk, O_sinit, O_sinit_rand, O_slamb_cut;
static int O_sbc_type[6];
static double O_sbc_val[6];
static double *s_, *s0_, *sc_, *sc0_, *sd_, *sd0_;

static void o_scalar_parse(void)
{
  char rest[MAXL], word[64];
  O_scalar = find_line("SCALAR", rest) ? atoi(rest) : 0;
  O_sD = find_line("s_D", rest) ? atof(rest) : 0.;
  O_sk = find_line("s_k", rest) ? atof(rest) : 0.;
  O_slamb_cut = find_line("s_lamb_cut", rest) ? atof(rest) : 0.;
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

/* --------------------------- particle heat ------------------------------ */
static int o_sncmax_;                 /* s_ncoeffs_max */
static const int o_snn[O_SNCOEFF] = {0, 1, 1, 1, 2, 2, 2, 2, 2, 3, 3, 3, 3, 3,
  3, 3, 4, 4, 4, 4, 4, 4, 4, 4, 4};
static const int o_smm[O_SNCOEFF] = {0, -1, 0, 1, -2, -1, 0, 1, 2, -3, -2, -1,
  0, 1, 2, 3, -4, -3, -2, -1, 0, 1, 2, 3, 4};

/* The reference's associated Legendre table has its m < 0 cases commented
 * out, and its switch falls through every n to `return 0`
 * (src/particle_kernel.cu:1928-1990, src/physalis_kernel.cu:57-116): the
 * m < 0 terms of the scalar series and quadrature are zero. */
static double o_sPnm(int n, int m, double th)
{
  return m < 0 ? 0. : o_Pnm(n, m, th);
}

static void o_scalar_part_init(void)
{
  o_sncmax_ = 0;
  for (int n = 0; n < NP_; n++) {
    int snc = 0;
    for (int j = 0; j <= P_[n].sorder; j++) snc += 2*j + 1;
    if (snc > o_sncmax_) o_sncmax_ = snc;
  }
  for (int n = 0; n < NP_; n++) {
    OP *Q = &P_[n];
    Q->srs = Q->srs * Q->r;
    Q->q = 0.0;
    Q->sncoeff = 0;
    for (int j = 0; j <= Q->sorder; j++) Q->sncoeff += 2*j + 1;
    for (int j = 0; j < O_SNCOEFF; j++)
      Q->anm_re[j] = Q->anm_im[j] = Q->anm_re0[j] = Q->anm_im0[j] = 0.;
  }
}

/* scalar at one node on the srs sphere, minus the particle's scalar
 * (src/scalar_kernel.cu:303-527); the node's code from the scalar BCs */
static double o_snode(OP *Q, int node)
{
  double th = o_nt[node], ph = o_np[node];
  double xp = Q->srs * sin(th) * cos(ph);
  double yp = Q->srs * sin(th) * sin(ph);
  double zp = Q->srs * cos(th);
  double x = xp + Q->x, y = yp + Q->y, z = zp + Q->z;
  int ic = 0, jc = 0, kc = 0;
  for (int im = 0; im < 27; im++) {
    int sx = im % 3 - 1, sy = (im / 3) % 3 - 1, sz = im / 9 - 1;
    if ((sx && !O.per[0]) || (sy && !O.per[1]) || (sz && !O.per[2])) continue;
    double xi = xp + oimg(Q->x, sx, O.xl), yi = yp + oimg(Q->y, sy, O.yl),
           zi = zp + oimg(Q->z, sz, O.zl);
    int a = o_cell((xi - (O.xs - O.dx)) * (1. / O.dx));
    int b = o_cell((yi - (O.ys - O.dy)) * (1. / O.dy));
    int c = o_cell((zi - (O.zs - O.dz)) * (1. / O.dz));
    if (im == 13 || (a >= 1 && a <= O.nx && b >= 1 && b <= O.ny &&
                     c >= 1 && c <= O.nz)) {
      x = xi; y = yi; z = zi; ic = a; jc = b; kc = c;
      if (a >= 1 && a <= O.nx && b >= 1 && b <= O.ny && c >= 1 && c <= O.nz)
        break;
    }
  }
  /* node code: faces W E S N B T, Dirichlet before Neumann */
  int code = -1;
  const double gap[6] = {x - O.xs, x - O.xe, y - O.ys, y - O.ye, z - O.zs,
                         z - O.ze};
  for (int f6 = 0; f6 < 6 && code == -1; f6++) {
    int out = (f6 & 1) ? gap[f6] >= 0 : gap[f6] < 0;
    if (!out) continue;
    if (O_sbc_type[f6] == DIRICHLET) code = -100 - 10 * f6;
    else if (O_sbc_type[f6] == NEUMANN) code = -101 - 10 * f6;
  }
  Q->nodes[node] = code;
  int oobi = ic < 1 || ic > O.nx, oobj = jc < 1 || jc > O.ny,
      oobk = kc < 1 || kc > O.nz;
  int i = o_clamp(ic, 1, O.nx), j = o_clamp(jc, 1, O.ny), k = o_clamp(kc, 1, O.nz);
  long C = IDX(O.cc, i, j, k);
  double ss = o_interp(s_, &O.cc, i, j, k, x, y, z, (i - 0.5) * O.dx + O.xs,
    (j - 0.5) * O.dy + O.ys, (k - 0.5) * O.dz + O.zs) - Q->s;
  /* a node beyond a face: the face's Dirichlet value, or the ghost cell
   * beside the clamped cell for a Neumann face */
  double wall[3];
  long nb[6] = {C - 1, C + 1, C - O.cc.sy, C + O.cc.sy, C - O.cc.sz, C + O.cc.sz};
  for (int d = 0; d < 3; d++) {
    wall[d] = (code == -100 - 20 * d) * O_sbc_val[2 * d]
            + (code == -110 - 20 * d) * O_sbc_val[2 * d + 1]
            + (code == -101 - 20 * d) * s_[nb[2 * d]]
            + (code == -111 - 20 * d) * s_[nb[2 * d + 1]]
            - Q->s;
  }
  return (1-oobi) * (1-oobj) * (1-oobk) * (code == -1) * ss
       + oobi * (1-oobj) * (1-oobk) * (code < -1) * wall[0]
       + (1-oobi) * oobj * (1-oobk) * (code < -1) * wall[1]
       + (1-oobi) * (1-oobj) * oobk * (code < -1) * wall[2];
}

/* quadrature and coefficients anm = <Y_nm, s - s_p> / (rsa^n - ars^(n+1)),
 * heat flux q = 2 sqrt(pi) a a_00 (src/scalar_kernel.cu:529-577,
 * 1247-1272) */
static void o_scalar_lamb(void)
{
  for (int n = 0; n < NP_; n++) {
    OP *Q = &P_[n];
    double ss[O_NNODES];
    for (int node = 0; node < O_NNODES; node++) ss[node] = o_snode(Q, node);
    double ars = Q->r / Q->srs, rsa = Q->srs / Q->r;
    for (int coeff = 0; coeff < Q->sncoeff; coeff++) {
      int nn = o_snn[coeff], m = o_smm[coeff];
      double re = 0., im = 0.;
      for (int node = 0; node < O_NNODES; node++) {
        double th = o_nt[node], ph = o_np[node];
        double N_nm = o_Nnm(nn, m), P_nm = o_sPnm(nn, m, th);
        double Ire = N_nm*P_nm*ss[node]*cos(m * ph);
        double Iim = -N_nm*P_nm*ss[node]*sin(m * ph);
        double wgt = node < 6 ? o_A1 : (node < 18 ? o_A2 : o_A3);
        if (node == 0) { re = Ire * o_A1; im = Iim * o_A1; }
        else { re += wgt * Ire; im += wgt * Iim; }
      }
      double A = pow(rsa, nn) - pow(ars, nn + 1.);
      Q->anm_re[coeff] = re / A;
      Q->anm_im[coeff] = im / A;
    }
    if (Q->sncoeff > 0) Q->q = 2. * sqrt(PI) * Q->r * Q->anm_re[0];
  }
}

/* largest relative coefficient change (src/scalar_kernel.cu:1274-1339,
 * src/cuda_scalar.cu:914-946), then anm0 = anm */
static double o_scalar_err(void)
{
  double err = DBL_MIN;
  if (NP_ > 0) {
    double best = 0.;
    for (int n = 0; n < NP_; n++) {
      OP *Q = &P_[n];
      const double *cur[2] = {Q->anm_re, Q->anm_im};
      const double *old[2] = {Q->anm_re0, Q->anm_im0};
      double mx = DBL_MIN;
      for (int coeff = 0; coeff < o_sncmax_; coeff++) {
        double m = DBL_MIN;
        for (int i = 0; i < 2; i++) {
          double cv = cur[i][coeff], zv = cur[i][0];
          int flag = (fabs(cv) > fabs(O_slamb_cut*zv)) * (fabs(cv) > 1.e-16) *
                     (fabs(zv) > DBL_MIN);
          double div = fabs(cv);
          div += (1.e-16 - div) * (div < 1.e-16);
          double e = (double) flag * fabs(cv - old[i][coeff]) / div;
          m += (e - m) * (e > m);
        }
        mx += (m - mx) * (m > mx);
      }
      if (mx > best) best = mx;
    }
    err = best;
    for (int n = 0; n < NP_; n++)
      for (int coeff = 0; coeff < o_sncmax_; coeff++) {
        P_[n].anm_re0[coeff] = P_[n].anm_re[coeff];
        P_[n].anm_im0[coeff] = P_[n].anm_im[coeff];
      }
  }
  return err;
}

/* the analytic scalar in particle cells: the series in shell cells, the
 * particle's scalar inside (src/scalar_kernel.cu:1143-1227) */
static void o_scalar_part_bc(double *a)
{
  if (NP_ == 0) return;
  for (int k = 1; k <= O.nz; k++)
    for (int j = 1; j <= O.ny; j++)
      for (int i = 1; i <= O.nx; i++) {
        long C = IDX(O.cc, i, j, k);
        int P = phase_[C];
        if (P < 0) continue;
        const OP *Q = &P_[P];
        double x = (i - 0.5) * O.dx + O.xs, y = (j - 0.5) * O.dy + O.ys,
               z = (k - 0.5) * O.dz + O.zs;
        x -= onear(x, Q->x, O.xl, O.per[0]);
        y -= onear(y, Q->y, O.yl, O.per[1]);
        z -= onear(z, Q->z, O.zl, O.per[2]);
        double r, th, ph;
        o_xyz2rtp(x, y, z, &r, &th, &ph);
        double ar = Q->r / r, ra = r / Q->r;
        double v = Q->s;
        for (int n = 0; n <= Q->sorder; n++) {
          int coeff = 0;
          for (int jj = 0; jj < n; jj++) coeff += 2*jj + 1;
          double X = 0.;
          for (int m = -n; m <= n; m++) {
            X += o_Nnm(n, m) * o_sPnm(n, m, th)
              * (Q->anm_re[coeff]*cos(m*ph) - Q->anm_im[coeff]*sin(m*ph));
            coeff++;
          }
          v += (pow(ra, n) - pow(ar, n + 1)) * X;
        }
        a[C] = shell_[C] < 1 ? v : Q->s;
      }
}

static void o_scalar_part_fill(double *a)
{
  if (NP_ == 0) return;
  for (int k = 1; k <= O.nz; k++)
    for (int j = 1; j <= O.ny; j++)
      for (int i = 1; i <= O.nx; i++) {
        long C = IDX(O.cc, i, j, k);
        if (phase_[C] > -1) a[C] = P_[phase_[C]].s;
      }
}

/* particle scalar from the heat flux (src/scalar_kernel.cu:1351-1360) */
static void o_update_part_scalar(void)
{
  for (int n = 0; n < NP_; n++) {
    OP *Q = &P_[n];
    double vol = 4./3. * PI * Q->r*Q->r*Q->r;
    double m = vol * Q->rho;
    Q->s += (float) Q->update * Q->q * O_sk * O.dt / m / Q->cp;
  }
}

/* the scalar iteration after the Lamb loop (src/bluebottle.c:293-341) */
static int O_scalar_iters;
static double O_scalar_err;

static void o_scalar_step(void)
{
  int it = 0;
  double err = FLT_MAX;
  while (err > O_lamb_residual) {
    o_scalar_bc(s0_);
    o_scalar_part_bc(s0_);
    orc_exchange(s0_, 0);
    o_scalar_solve();
    orc_exchange(s_, 0);
    o_scalar_bc(s_);
    o_scalar_part_fill(s_);
    orc_exchange(s_, 0);
    if (NP_ > 0) o_scalar_lamb();
    err = o_scalar_err();
    it++;
    if (it == O_lamb_max_iter) break;
  }
  O_scalar_iters = it;
  O_scalar_err = err;
  o_store_s();
  o_update_part_scalar();
}

/* Scalar state of particle n: s, q, then anm_re and anm_im (25 each). */
int orc_part_scalar(int n, double *out)
{
  if (n < 0 || n >= NP_) return -1;
  const OP *Q = &P_[n];
  out[0] = Q->s;
  out[1] = Q->q;
  memcpy(out + 2, Q->anm_re, sizeof(Q->anm_re));
  memcpy(out + 2 + O_SNCOEFF, Q->anm_im, sizeof(Q->anm_im));
  return 0;
}

/* Set the scalar state of particle n (s, q, anm_re[25], anm_im[25]). */
int orc_part_set_scalar(int n, const double *in)
{
  if (n < 0 || n >= NP_) return -1;
  OP *Q = &P_[n];
  Q->s = in[0];
  Q->q = in[1];
  memcpy(Q->anm_re, in + 2, sizeof(Q->anm_re));
  memcpy(Q->anm_im, in + 2 + O_SNCOEFF, sizeof(Q->anm_im));
  return 0;
}

/* One pass of the particle scalar quadrature on the current s (the
 * coefficients and heat fluxes), and the analytic particle values written
 * into s, for closed-form tests. */
void orc_scalar_lamb_once(void)
{
  o_scalar_lamb();
}

void orc_scalar_part_bc_s(void)
{
  o_scalar_part_bc(s_);
}
