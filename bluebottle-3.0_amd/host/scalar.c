/*
 * Scalar transport: initial field and particle scalar state
 * (src/scalar.c:55-115).
 */
#include <stdio.h>
#include <stdlib.h>

#include "bluebottle.h"

void scalar_init_fields(void)
{
  /* src/scalar.c:55-88: every cell, ghosts included, draws from the C
   * library generator seeded with the cell's global index */
  const grid_info *G = &dom[0].Gcc;
  for (int i = 0; i < G->s3b; i++) {
    s_conv[i] = 0.0;
    s_conv0[i] = 0.0;
    s_diff[i] = 0.0;
    s_diff0[i] = 0.0;
  }
  for (int i = G->_isb; i <= G->_ieb; i++)
    for (int j = G->_jsb; j <= G->_jeb; j++)
      for (int k = G->_ksb; k <= G->_keb; k++) {
        int ii = i + G->isb, jj = j + G->jsb, kk = k + G->ksb;
        int lc = GCC_LOC(i, j, k, G->s1b, G->s2b);
        int gc = GCC_LOC(ii, jj, kk, DOM.Gcc.s1b, DOM.Gcc.s2b);
        srand(gc);
        s[lc] = s_init + s_init_rand * (2. * rand() / (real) RAND_MAX - 1.);
        s0[lc] = s[lc];
      }
}

void scalar_part_init(void)
{
  /* src/scalar.c:90-115 */
  for (int i = 0; i < nparts; i++) {
    part_struct *P = &parts[i];
    P->srs = P->srs * P->r;
    P->q = 0.0;
    P->sncoeff = 0;
    for (int j = 0; j <= P->sorder; j++) P->sncoeff += 2*j + 1;
    if (P->sncoeff > S_MAX_COEFFS) {
      printf("Maximum order is 4.");
      exit(EXIT_FAILURE);
    }
    for (int j = 0; j < S_MAX_COEFFS; j++) {
      P->anm_re[j] = 0.;
      P->anm_im[j] = 0.;
      P->anm_re0[j] = 0.;
      P->anm_im0[j] = 0.;
    }
  }
}
