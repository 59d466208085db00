/*
 * The reference's example programs (example1.cu:19-54, example2.cu:42-60,
 * example3.cu:43-60) as one plain-C program against the drop-in header
 * ParaLanczos.h and libparalanczos.so, with their results checked against
 * FP64 LAPACK answers for the example matrix rounded to FP32 (the device
 * path's input precision): relative error <= 1e-5 per eigenvalue, 1e-4 per
 * determinant. Exit status 0 when every check passes.
 *
 *   gcc -O2 -I include legacy_examples.c -L paralanczos_amd -lparalanczos \
 *       -Wl,-rpath,<dir of libparalanczos.so> -lm -o legacy_examples
 */
#include <math.h>
#include <stdio.h>

#include "ParaLanczos.h"

/* the 8 x 8 symmetric example matrix of example1.cu:19-26 (row-major) */
static double A[64] = {
    11.615, 12.6022, 2.7449, 0.6018, 5.7419, -0.8989, 6.2368, 1.488,
    12.6022, 32.5907, 6.8222, 3.8457, 3.8868, -1.3879, 7.2553, 1.3472,
    2.7449, 6.8222, 11.3095, 1.99, -3.8608, 0.9067, 3.4324, 3.3344,
    0.6018, 3.8457, 1.99, 4.7412, -0.0663, -1.2597, -0.2133, -1.6123,
    5.7419, 3.8868, -3.8608, -0.0663, 18.0182, -1.7439, -1.5991, 1.2881,
    -0.8989, -1.3879, 0.9067, -1.2597, -1.7439, 1.4386, 1.4156, 1.8073,
    6.2368, 7.2553, 3.4324, -0.2133, -1.5991, 1.4156, 8.2406, 2.8208,
    1.488, 1.3472, 3.3344, -1.6123, 1.2881, 1.8073, 2.8208, 4.7334};

/* ascending eigenvalues and determinants of the principal submatrices */
static const double EIG_1357[4] = {0.5130670074085307, 2.732051392112843, 7.034317355852152, 33.2244622691855};
static const double DET_1357 = 327.5992900908204;
static const double EIG_0123[4] = {3.614286799788453, 6.086225025211212, 9.731245978986411, 40.82463992016554};
static const double DET_0123 = 8738.994819784279;
static const double EIG_4567[4] = {0.3851825248030271, 3.346166795501465, 10.176441052970794, 18.523008991768357};
static const double DET_4567 = 242.9526873637158;
static const double EIG_ALL[8] = {0.1670855789103259, 0.6041754031344736, 2.4214609195113304, 4.885823629808334, 7.604127696806458, 11.184628177287234, 21.31756945285902, 44.50232623087808};
static const double DET_ALL = 96362.70358759053;

static int failures = 0;

static void check(const char *what, const double *got, const double *want, int n, double rtol) {
  for (int i = 0; i < n; ++i) {
    const double err = fabs(got[i] - want[i]) / fabs(want[i]);
    if (!(err <= rtol)) {
      printf("FAIL %s[%d]: %.12g vs %.12g (rel %.3g)\n", what, i, got[i], want[i], err);
      ++failures;
    }
  }
}

int main(void) {
  double eig[32], det[8];

  /* example1.cu:39-47: one subset {1, 3, 5, 7} through calculate_subeigs */
  int s1357[4] = {1, 3, 5, 7};
  calculate_subeigs(A, 8, eig, det, s1357, 4);
  check("subeigs eig", eig, EIG_1357, 4, 1e-5);
  check("subeigs det", det, &DET_1357, 1, 1e-4);

  /* example2.cu / example3.cu: calculate_subeigs3 over 32 indices, eight
     subsets of four: {0..3}, {4..7} repeated */
  int idx[32];
  for (int i = 0; i < 32; ++i) idx[i] = i % 8;
  calculate_subeigs3(A, 32, 8, eig, det, idx, 4);
  for (int s = 0; s < 8; ++s) {
    const int lo = (s % 2) == 0;
    check("subeigs3 eig", eig + 4 * s, lo ? EIG_0123 : EIG_4567, 4, 1e-5);
    check("subeigs3 det", det + s, lo ? &DET_0123 : &DET_4567, 1, 1e-4);
  }

  /* calculate_eigs (ParaLanczos.cu:1685-1722): the whole 8 x 8 matrix */
  calculate_eigs(A, 8, eig, det);
  check("eigs eig", eig, EIG_ALL, 8, 1e-5);
  check("eigs det", det, &DET_ALL, 1, 1e-4);

  printf("%s (%d failures)\n", failures ? "FAILED" : "all example checks passed", failures);
  return failures ? 1 : 0;
}
