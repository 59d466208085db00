/*
 * ParaLanczos.h - drop-in host entry points of btuomanen/ParaLanczos,
 * served by libparalanczos.so on MI355X (gfx950).
 *
 * Plain C (extern "C"), compilable by any C or C++ compiler: no CUDA or HIP
 * qualifiers. Same names, argument meaning, host-array ownership, stderr
 * messages and stdout printing as the reference; see each declaration for
 * the reference line it replaces and the documented differences.
 *
 * Shared behaviour:
 *   - mat is a caller-owned host array of FP64 values, row-major n x n;
 *     it is converted to FP32 on upload (the device path is FP32, with the
 *     Sturm bisection in FP64), so eigenvalues carry FP32 input rounding.
 *   - Eigenvalues are written ascending per subset, Det = product of them.
 *   - No cudaDeviceReset-equivalent: other users of the GPU are unaffected.
 *   - Every eigenvalue is printed to stdout as the reference does, unless
 *     the environment variable PARALANCZOS_QUIET is set to a non-empty value
 *     other than "0".
 *   - Calls are synchronous and serialised by an internal mutex.
 *   - Where the reference printed an argument error and carried on (writing
 *     out of bounds), these return after the message instead.
 */
#ifndef PARALANCZOS
#define PARALANCZOS

#ifdef __cplusplus
extern "C" {
#endif

/* ParaLanczos.h:39, ParaLanczos.cu:1642-1682: prints the n x n tridiagonal
   matrix with diagonal alpha[i] and off-diagonals T(i,i+1) = beta[i+1]. */
void print_tridiagonal(double *alpha, double *beta, int n);

/* ParaLanczos.h:41, ParaLanczos.cu:1685-1722: all n eigenvalues of the
   n x n matrix (n % 4 == 0, n <= 128) into Eigenvalues[0..n-1], Det[0]. */
void calculate_eigs(double *mat, int n, double *Eigenvalues, double *Det);

/* ParaLanczos.h:43, ParaLanczos.cu:1725-1765: one subset of size subn
   (subn % 4 == 0, n <= 128 as the reference checks) of the n x n matrix;
   Eigenvalues[0..subn-1], Det[0]. */
void calculate_subeigs(double *mat, int n, double *Eigenvalues, double *Det, int *subIndex,
                       int subn);

/* ParaLanczos.h:45, ParaLanczos.cu:1767-1808 (deprecated there): n indices
   forming n/subn subsets of the matsize x matsize matrix; Eigenvalues[0..n-1],
   Det[0] only (the reference copies back one determinant). */
void calculate_subeigs2(double *mat, int n, int matsize, double *Eigenvalues, double *Det,
                        int *subIndex, int subn);

/* ParaLanczos.h:47, ParaLanczos.cu:1811-1863: subIndexLen indices forming
   subIndexLen/subn subsets (subn % 4 == 0, subn <= 128, subIndexLen % 32 == 0)
   of the matsize x matsize matrix; Eigenvalues[0..subIndexLen-1],
   Det[0..subIndexLen/subn-1]. */
void calculate_subeigs3(double *mat, int subIndexLen, int matsize, double *Eigenvalues,
                        double *Det, int *subIndex, int subn);

/* Host-array entry point behind the Matlab front-ends
   Matlab/gpu_subsvd.m:14-59, gpu_multi_subsvd.m:13-68, gpu_multi_subsvd2.m:13-73
   and gpu_svd.m:13-57, which launch POS_SYM_SUBMATRIX_KER(2)/POS_SYM_MAT_KER via
   parallel.gpu.CUDAKernel + feval. g is M x M FP64 (column- or row-major: it
   is symmetric), subIndex holds L 0-based indices forming L/minor subsets of
   size minor. Writes eigs[0..L-1] (ascending per subset), det[0..L/minor-1],
   err[0..L/minor-1] (0 = ok). Performs no argument validation beyond memory
   safety (the .m wrappers validate and return -1,-1,-1 as the reference's
   do); returns 0 on success, -1 on failure. */
int pl_multi_subsvd(const double *g, int M, const int *subIndex, int L, int minor,
                    double *eigs, double *det, int *err);

#ifdef __cplusplus
}
#endif
#endif
