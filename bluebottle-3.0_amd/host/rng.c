/*
 * Seeded 64-bit generator: the Ranq1 xorshift + multiplier method of
 * Press et al., Numerical Recipes 3rd ed., section 7.1 (the generator the
 * reference uses, src/rng.c:24-81).  The state and its update are those of
 * the published method; rng_dbl maps to [0, 1) by multiplying with 2^-64.
 */
#include "bluebottle.h"

static unsigned long long rng_state = 4101842887655102017ULL;

void rng_init(unsigned long long seed)
{
  rng_state = 4101842887655102017ULL ^ seed;
  rng_state = rng_uint64();
}

unsigned long long rng_uint64(void)
{
  rng_state ^= rng_state >> 21;
  rng_state ^= rng_state << 35;
  rng_state ^= rng_state >> 4;
  return rng_state * 2685821657736338717ULL;
}

double rng_dbl(void)
{
  return 5.42101086242752217e-20 * (double) rng_uint64();
}
