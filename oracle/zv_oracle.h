/*
 * CPU oracle for the slater_stoch ZVMC two-electron-integral hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (scemama/slater_stoch, files cited per function in
 * zv_oracle.c).  It is the checker for the HIP library: only tests, the smoke
 * check and the CPU-baseline leg of the benchmark may load it.  The product
 * (libzvmc.so) never links or calls it.
 *
 * Array conventions are the reference's Fortran ones, so that the same
 * buffers can be handed to the oracle and to the HIP C-ABI:
 *   center[3*(i-1)+d], npower[3*(i-1)+d]          (Fortran center(3,n))
 *   g_gauss[ldg*(i-1)+m], c_gauss[ldg*n*t+ldg*(i-1)+m]  (c_gauss(ldg,n,2))
 *   samples r1[3*m+d] (sample m, Cartesian d), pi0[m]
 * Orbital indices are 1-based as in the reference.
 */
#ifndef ZV_ORACLE_H
#define ZV_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

int zvo_set_basis(int n, const double *center, const int *npower, const int *n_sto,
                  const int *i_type, const double *zeta, const int *n_gauss, int ldg,
                  const double *c_gauss, const double *g_gauss, double a_zv);
int zvo_nbasis(void);

double zvo_exp(double x);
double zvo_u_orb(int i, double r);
double zvo_u_gauss(int i, double r);

void zvo_pair_samples(int i, int k, int ns, const double *r1, const double *r2, double *ut1,
                      double *ut2, double *rho1, double *rho2, double *rhog1, double *rhog2,
                      double *jac1, double *jac2);
void zvo_pi0(int ns, const double *r1, const double *r2, double *pi0);

int64_t zvo_nint(int n);
int64_t zvo_build_mapping(int n, int *is, int *js, int *ks, int *ls);
int zvo_compare_nuclei(int ni, int nj, int nk, int nl);

int zvo_mc_fixed(int nbl, int npts, const double *r1, const double *r2, const double *pi0,
                 int64_t nq, const int *is, const int *js, const int *ks, const int *ls,
                 double *mean, double *err, double *es_block, double *eg_block);

double zvo_gauss_ijkl(int i, int k, int j, int l);
double zvo_bielec_integral(double alpha, double beta, double delta, double gama,
                           const double *A, const double *B, const double *C, const double *D,
                           const int *pA, const int *pB, const int *pC, const int *pD);
void zvo_precond(double *precond);
void zvo_gauss_eri(int64_t nq, const int *is, const int *js, const int *ks, const int *ls,
                   const double *precond, double *out, int64_t *n_screened);

double zvo_rint_sum(int n_pt_out, double rho, const double *d1);
double zvo_rint(int n, double rho);

void zvo_philox4x32_10(const uint32_t *ctr, const uint32_t *key, uint32_t *out);

/* exact one-centre Slater integral (ik|jl) of x^nx y^ny z^nz r^nsto e^{-zeta r}
 * factors ordered i, k, j, l (src/compute_int_slater.F90); NAN above d */
double zvo_mono_exact(const int *nx, const int *ny, const int *nz, const int *nsto,
                      const double *zeta);
double zvo_gaunt(int l1, int l2, int l3, int m1, int m2, int m3);
void zvo_ylm_coef(int a, int b, int c, int l, int m, double *re, double *im);

/* overlap, nuclear attraction and kinetic integrals over the Gaussian fits
 * of the loaded basis, n x n column-major (src/one_elect.F90); nuclei:
 * charge[nnuc], nuc[3*c+d] */
void zvo_one_e(int nnuc, const double *charge, const double *nuc, double *S, double *V,
               double *K);

#ifdef __cplusplus
}
#endif
#endif
