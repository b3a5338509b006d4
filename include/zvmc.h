/*
 * zvmc.h -- C-ABI of the MI355X ZVMC two-electron-integral engine.
 *
 * Drop-in for the hot path of scemama/slater_stoch's two_e_slater.  The
 * reference has no library API: its hot path is inline in the main program
 * and in Fortran external procedures with implicit interfaces.  Each entry
 * point below names the reference code it replaces (file:line, paths
 * relative to the reference tree).  The Fortran host binds these with
 * iso_c_binding (slater_stoch_amd/fortran/zvmc_mod.F90).
 *
 * Conventions
 *  - Plain pointers to host memory, Fortran (column-major) layouts:
 *      center(3,n), npower(3,n), c_gauss(ldg,n,2), g_gauss(ldg,n),
 *      samples r(3,m) (x,y,z of sample 1, then sample 2, ...).
 *  - Orbital indices are 1-based (as in the reference); quartet counts int64.
 *  - Quartet arrays (ERIs, means, errors) are in the reference order of
 *    src/build_mapping_ijkl.F90:16-43 (kcp = 1..nint).
 *  - Every call returns 0 on success or a ZV_ERR_* code; zv_last_error()
 *    returns the message.  The Fortran host turns nonzero into `stop`.
 *  - Contexts on a device share one HIP stream and four side streams,
 *    made by the first context and kept until the process ends; side
 *    streams are joined back inside a call, and calls are synchronous for
 *    the caller.  A context is not
 *    thread-safe: drive it from one host thread.  Calls on different
 *    contexts from several threads run one at a time (a process-wide lock
 *    in every call that creates, drives or destroys a context): three
 *    threads driving three contexts at once hung on MI355X.
 */
#ifndef ZVMC_H
#define ZVMC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ZV_OK 0
#define ZV_ERR_ARG 1     /* invalid argument or unsupported input        */
#define ZV_ERR_HIP 2     /* HIP runtime / library error                   */
#define ZV_ERR_STATE 3   /* call out of order (e.g. no basis set)         */
#define ZV_ERR_NOMEM 4   /* device allocation failed                      */
#define ZV_ERR_NODEV 5   /* no usable gfx950 device                       */

/* sampling modes for zv_mc_blocks */
#define ZV_SAMPLES_FIXED 0   /* caller-supplied (r1, r2, pi0) list        */
#define ZV_SAMPLES_PHILOX 1  /* on-device Philox4x32-10 + inverse erfc    */

/* estimator precisions for zv_mc_blocks */
#define ZV_PREC_FP64_PARITY 0 /* reference operation order and FP32 casts,
                                 bit-identical to the CPU oracle          */
#define ZV_PREC_FP64 1        /* tiled FP64 estimator, split S-G form      */
#define ZV_PREC_FP32 2        /* tiled FP32 estimator, FP64 block sums     */
#define ZV_PREC_BF16 3        /* FP32 fields and kernel values; the two
                                 products K A1 A2 + K B1 B2 from factors
                                 rounded to BF16, summed by dot2 in FP32;
                                 FP64 block sums                            */

typedef struct zv_context zv_context;

/* Context on HIP device `device` (one stream).  Replaces the module state of
 * src/common_data.F90:1-58 that the reference keeps in globals. */
int zv_create(int device, zv_context **ctx);
int zv_destroy(zv_context *ctx);
int zv_last_error(const zv_context *ctx, char *buf, int len);

/* Basis as filled by src/read_basis.F90:27-98 and
 * src/build_gaussian_expansion_of_orbitals.F90:9-471 (module common_data):
 * center(3,n) bohr, npower(3,n), n_STO(n), i_type(n) (1 or 2), g_slater(n),
 * n_gauss(n), c_gauss(ldg,n,2), g_gauss(ldg,n), a_ZV (src/two_e_slater.F90:153).
 * n_gauss(i) <= 32.  Also builds the pair table (G_center of
 * src/two_e_slater.F90:339-347) and the quartet index of
 * src/build_mapping_ijkl.F90. */
int zv_set_basis(zv_context *ctx, int n, const double *center, const int *npower,
                 const int *n_sto, const int *i_type, const double *zeta, const int *n_gauss,
                 int ldg, const double *c_gauss, const double *g_gauss, double a_zv);

/* nint = np(np+1)/2 with np = n(n+1)/2 (src/build_mapping_ijkl.F90:12-13). */
int64_t zv_nint(int n);

/* is/js/ks/ls of src/build_mapping_ijkl.F90:9-44 (host computation, 1-based). */
int zv_build_mapping(int n, int *is, int *js, int *ks, int *ls);

/* Analytic Gaussian-fit ERIs with the reference's Schwarz driver,
 * src/two_e_slater.F90:219-284: precond(i,k) = (ik|ik) (out, n x n, may be
 * NULL), ijkl_gaus(kcp) for all nint quartets (out; NULL keeps it on the
 * device only, see zv_get_result), n_screened = quartets dropped by the
 * 1e-15 Cauchy-Schwarz test (out, may be NULL).
 * Per quartet this is gauss_ijkl (src/gauss_ijkl.F90:51-154).  The work
 * plan of long quartet lists is built on up to 8 host threads inside the
 * call (joined before it returns). */
int zv_gauss_eri(zv_context *ctx, double *ijkl_gaus, double *precond, int64_t *n_screened);

/* Exact Slater integrals of the one-centre quartets (all four AOs on one
 * atom), the do_exact_monocenter branch of src/two_e_slater.F90:318-365
 * (per quartet src/compute_int_slater.F90 -> src/ijkl_slater_xnynzn.F90):
 * ijkl[kcp] (in/out, nint) is overwritten with the exact value for every
 * one-centre kcp and err[kcp] (in/out, may be NULL) with 0, as the
 * reference's final assembly does (src/two_e_slater.F90:489-493, 558-563);
 * other entries are left as they are.  n_mono (out, may be NULL) = number
 * of one-centre quartets.  AOs above d (4F, 5G) on any atom give
 * ZV_ERR_ARG, where the reference stops. */
int zv_exact_monocenter(zv_context *ctx, double *ijkl, double *err, int64_t *n_mono);

/* One-electron integrals over the Gaussian fits of the loaded basis, the
 * loop of src/one_e_slater.F90:87-95 over src/one_elect.F90: overlap S,
 * nuclear attraction V (src/one_electron_Vij.F90: -sum over nuclei of
 * charge x NAI) and kinetic K (src/one_electron_Kij.F90), each n x n
 * column-major (out; S(i,k) at S[(i-1) + n (k-1)]).  Nuclei: charge[nnuc]
 * and nuc[3 c + d] in bohr (src/read_geometry.F90: charge = atomic number).
 * AOs above degree 6 (with the r^2 of i_type 2) give ZV_ERR_ARG. */
int zv_one_e(zv_context *ctx, int nnuc, const double *charge, const double *nuc, double *S,
             double *V, double *K);

/* gauss_ijkl(i,k,j,l) for an explicit list of nq quartets (1-based orbital
 * indices; no screening).  src/gauss_ijkl.F90:51 batched. */
int zv_gauss_ijkl_list(zv_context *ctx, int64_t nq, const int *qi, const int *qk,
                       const int *qj, const int *ql, double *out);

/* zv_gauss_ijkl_list without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, kernels and values (bit for bit) as zv_gauss_ijkl_list;
 * the device work runs in a fixed static device array of the library
 * (2 MiB per device that loads it, held until the process ends), so
 * nq <= ZV_NOALLOC_MAXQ (ZV_ERR_ARG above).  The primitive-pair tables of
 * the basis must exist: the first zv_gauss_eri or zv_gauss_ijkl_list after
 * zv_set_basis builds them (ZV_ERR_STATE otherwise).  Every call on a
 * device uses that device's one array: calls run one at a time under the
 * library's call lock, and callers that bypass it (several processes on one
 * GPU have separate arrays) must not overlap calls on one device. */
#define ZV_NOALLOC_MAXQ 65536
int zv_gauss_ijkl_list_noalloc(zv_context *ctx, int64_t nq, const int *qi, const int *qk,
                               const int *qj, const int *ql, double *out);

/* ZVMC correction blocks, src/two_e_slater.F90:386-496: nbl blocks of npts
 * samples (npts a multiple of 40, src/two_e_slater.F90:148).  mode
 * ZV_SAMPLES_FIXED reads r1(3,nbl*npts), r2(3,nbl*npts), pi0(nbl*npts);
 * ZV_SAMPLES_PHILOX draws them on the device from (seed, stream)
 * (src/draw_configuration.F90:1-29) and ignores r1/r2/pi0.  Outputs:
 * mean(kcp) = block average of e_S - e_G, err(kcp) = the reference's
 * standard error sqrt|<x^2>-<x>^2|/sqrt(nbl), optional block_means(kcp,nbl)
 * (may be NULL).  A NULL mean or err keeps that array on the device only
 * (zv_get_result). */
int zv_mc_blocks(zv_context *ctx, int nbl, int64_t npts, int mode, uint64_t seed,
                 uint64_t stream, int precision, const double *r1, const double *r2,
                 const double *pi0, double *mean, double *err, double *block_means);

/* Device-resident results.  No counterpart in the reference: for callers
 * that keep the per-quartet arrays in HBM between calls (a sample-size
 * policy that only needs the largest standard error, a benchmark that times
 * the job without the PCIe copies).  The last zv_gauss_eri (ZV_RESULT_ERI)
 * and zv_mc_blocks (ZV_RESULT_MEAN, ZV_RESULT_STDERR) results stay on the
 * device in kcp order until the next such call or zv_set_basis; a NULL host
 * output skips only the copy.  zv_get_result copies one to host (nint
 * doubles); zv_result_max reduces it on the device.  ZV_ERR_STATE when the
 * result has not been computed for the current basis.  Neither call changes
 * what zv_last_timings reports. */
#define ZV_RESULT_ERI 0
#define ZV_RESULT_MEAN 1
#define ZV_RESULT_STDERR 2
int zv_get_result(zv_context *ctx, int which, double *host);
int zv_result_max(zv_context *ctx, int which, double *value);

/* Targeted refinement of device-resident Monte Carlo results (no
 * counterpart in the reference, which gives every quartet the same npts,
 * src/two_e_slater.F90:148): the quartets whose kept standard error exceeds
 * `target` get a fresh run of nbl (>= 2) blocks of npts Philox samples from
 * (seed, stream) -- the group estimator restricted to the tiles that hold
 * them, with the densities and block statistics of
 * src/two_e_slater.F90:386-496 -- and their kept mean and standard error are
 * replaced by that run's; every other quartet keeps its values.  Needs the
 * ZV_RESULT_MEAN / ZV_RESULT_STDERR results of a zv_mc_blocks call on the
 * current basis.  precision: ZV_PREC_FP64, ZV_PREC_FP32 or ZV_PREC_BF16.
 * n_refined (out, may be NULL): quartets replaced; max_err (out, may be
 * NULL): the largest kept standard error afterwards.  A time-to-target
 * policy: one pilot zv_mc_blocks, then zv_mc_refine until max_err <= target,
 * each round with more samples and another stream. */
int zv_mc_refine(zv_context *ctx, int nbl, int64_t npts, uint64_t seed, uint64_t stream,
                 int precision, double target, int64_t *n_refined, double *max_err);

/* Samples drawn by the device sampler for global sample indices
 * m0 .. m0+m-1 (ZV_SAMPLES_PHILOX), copied to the host: r1(3,m), r2(3,m),
 * pi0(m).  src/draw_configuration.F90:1-29 + src/compute_jacobian_and_pi0.F90:8-17. */
int zv_draw_samples(zv_context *ctx, uint64_t seed, uint64_t stream, int64_t m0, int64_t m,
                    double *r1, double *r2, double *pi0);

/* Pair transform, densities and Jacobians for pair (i,k), i >= k, at ns
 * caller samples r(3,ns): u_ik(r1) (3,ns), u_ik(r2), rho, rho_G at both
 * points and both Jacobians.  src/two_e_slater.F90:401-419,
 * src/compute_densities.F90:1-93, src/compute_jacobian_and_pi0.F90:1-33. */
int zv_pair_samples(zv_context *ctx, int i, int k, int ns, const double *r1, const double *r2,
                    double *ut1, double *ut2, double *rho1, double *rho2, double *rhog1,
                    double *rhog2, double *jac1, double *jac2);

/* PSD cleanup of the ERI list, src/svd_clean.F90:1-426 (rank, q power
 * iterations, Gaussian test matrices from `seed`): eri(kcp) in/out.
 * Removes negative eigencomponents of W((i,j),(k,l)) = <ij|kl>. */
int zv_svd_clean(zv_context *ctx, int rank, int q, uint64_t seed, double *eri,
                 int *n_removed);

/* Negative-eigenvalue removal through a dense symmetric eigensolve of the
 * same matrix, src/davidson_clean.F90:30-120 (dense branch). */
int zv_davidson_clean(zv_context *ctx, double *eri, int *n_removed);

/* Cleanups of the pair-packed ERI supermatrix V[P][Q] = (ik|jl), P = (i,k),
 * Q = (j,l) the AO pairs i >= k, j >= l (N = n(n+1)/2, pair index
 * i(i-1)/2 + k), built on the device from eri(kcp).  V is positive
 * semidefinite in exact arithmetic; Monte Carlo noise leaves small negative
 * eigenvalues, which both calls remove (eri in/out, written back as
 * (V[P][Q] + V[Q][P]) / 2).
 *
 * zv_rsvd_clean: randomized range finder of rank min(rank, N) with q power
 * iterations (src/randomized_svd.F90:22-65: Gaussian test matrix from
 * `seed`, Z = V P, then Z <- V (V^T Z) with re-orthonormalization each
 * iteration), followed by an FP64 Rayleigh-Ritz step in place of the SVD
 * of src/svd_clean.F90:159-281 (B = Q^T V Q, symmetric eigensolver); the
 * negative Ritz pairs are removed, V -= lambda (Q y)(Q y)^T.  `sketch`
 * selects the precision of the range finder's N x N x r products:
 * ZV_SKETCH_FP32 (SGEMM on an FP32 copy of V) or ZV_SKETCH_FP64; the
 * orthonormalizations and the Rayleigh-Ritz step are FP64.  ritz (out, may
 * be NULL): the min(rank, N) Ritz values, ascending; n_removed (out, may be
 * NULL): the negative pairs removed, those converged in the sense of
 * src/svd_clean.F90:286-296 (|u.v| >= 0.99, for a symmetric matrix a
 * residual ||V x - lambda x|| <= sqrt(1 - 0.99^2) |lambda|). */
#define ZV_SKETCH_FP64 0
#define ZV_SKETCH_FP32 1
int zv_rsvd_clean(zv_context *ctx, int rank, int q, uint64_t seed, int sketch, double *eri,
                  double *ritz, int *n_removed);

/* Block Davidson for the nev (1..64) lowest eigenpairs of the same V (the
 * role of src/davidson.F90:1-159 in the iterative branch of
 * src/davidson_clean.F90:120-323): the residuals of the unconverged Ritz
 * vectors as corrections (no diagonal preconditioner: diag V spans orders of
 * magnitude while the wanted eigenvalues lie near zero), an orthonormal
 * search space of up to max(8 nev, 48) vectors with thick restart, FP64
 * throughout.  precond ZV_DAVIDSON_SHIFT_INVERT replaces the residual by
 * (V - sigma I)^-1 r, with the Cholesky factor of V - sigma I for a shift
 * sigma below the lowest eigenvalue (the factorization succeeding certifies
 * it; one extra N x N matrix): after 8 plain iterations sigma = theta_1 -
 * (2 ||r_1|| + 1e-3 |theta_1|), lowered fourfold on each failed
 * factorization and moved up to theta_1 - 2 ||r_1|| whenever the residuals
 * have dropped a hundredfold (at most 4 factorizations).  Stops when every
 * residual norm ||V x - theta x|| <= tol (Hartree) or after max_iter
 * iterations; the negative eigenpairs whose residual is within tol are
 * removed from V (eri in/out).  evals (out, nev, may be NULL): the Ritz
 * values ascending; resid (out, nev, may be NULL): their residual norms;
 * n_removed, iters (out, may be NULL). */
#define ZV_DAVIDSON_PLAIN 0
#define ZV_DAVIDSON_SHIFT_INVERT 1
int zv_davidson_lowest(zv_context *ctx, int nev, double tol, int max_iter, int precond,
                       double *eri, double *evals, double *resid, int *n_removed, int *iters);

/* Progress lines of the long-running calls, written to stdout: level 1
 * prints what the reference prints unconditionally in svd_clean ("Vectors
 * a -- b / N" per randomized pass, src/svd_clean.F90:153-154, and the
 * largest/smallest singular value of the pass, :285); level 2 adds one line
 * per finished zv_mc_blocks block (the call then waits for each block, so
 * the line reports work done); level 0 (the default) prints nothing.  From
 * level 1 on, zv_gauss_eri also writes the host time of its Schwarz
 * phases' planning to stderr (lines starting "[zvmc]").  No counterpart
 * in the reference, whose prints cannot be switched off. */
int zv_set_verbose(zv_context *ctx, int level);

/* Tuning knobs (0 = the library's automatic choice, the default).  No
 * counterpart in the reference (whose only knobs are the OpenMP schedule and
 * the input deck); for autotuners and for tests that pin each code path.
 *  ZV_TUNE_EST_SPLIT  sample splits per estimator tile (1 = unsplit; capped
 *                     at 64, at 4 LDS stages per split and at 1 GiB of
 *                     partial sums)
 *  ZV_TUNE_MC_CHUNK   largest number of samples per estimator chunk (rounded
 *                     down to a multiple of 160, at least 160)
 *  ZV_TUNE_ERI_ITEMS  the analytic shell planner cuts a Schwarz phase into
 *                     work items of at most 1/value of its cost
 *  ZV_TUNE_ERI_STREAMS streams that run the analytic kernel buckets side by
 *                     side (1..5; 1 runs them one after another, which
 *                     times every bucket on its own)
 *  ZV_TUNE_EST_TILES  estimator tile shapes: 0 the default set, 1 half the
 *                     units per thread along the longer side (more, smaller
 *                     tiles for small molecules)
 * Values < 0 give ZV_ERR_ARG; the setting holds for the context's later
 * calls and changes only speed, except ZV_TUNE_EST_SPLIT and
 * ZV_TUNE_MC_CHUNK, which also change the estimator's summation order
 * (rounding only). */
#define ZV_TUNE_EST_SPLIT 0
#define ZV_TUNE_MC_CHUNK 1
#define ZV_TUNE_ERI_ITEMS 2
#define ZV_TUNE_ERI_STREAMS 3
#define ZV_TUNE_EST_TILES 4
#define ZV_TUNE_N 8
int zv_set_tuning(zv_context *ctx, int key, int64_t value);

/* Milliseconds spent in the last call's kernels, by phase:
 * t[0] sampler, t[1] densities, t[2] estimator, t[3] block statistics,
 * t[4] primitive pairs, t[5] analytic ERIs, t[6] copies, t[7] cleanup;
 * n[k] = number of launches of that phase.  Arrays of 8. */
int zv_last_timings(const zv_context *ctx, double *ms, int64_t *launches);

/* Per ERI kernel of the last zv_gauss_eri / zv_gauss_ijkl_list call, arrays
 * of 236.  Entries 0..35: component kernels, index 6*LB + LK (LB >= LK the
 * Hermite orders of the two AO pairs, 5 = the runtime-order kernel);
 * entries 36..235: shell kernels, index 36 + 25*v + 5*LU + LL (LU, LL the
 * orders of the uniform-side and lane-side shell pairs; v = 2*(lane batch-1)
 * + mode, lane batch 1..4, mode 0 direct, 1 wave-sum reduction).  ms: kernel milliseconds;
 * launches; units: primitive-quartet evaluations (times the lane-side
 * components for shell kernels); flops: algorithmic FP64 flops (Boys values
 * and R table once per primitive quartet, contractions per component).
 * Diagnostic entry point with no counterpart in the reference (which prints
 * only a CPU time, src/two_e_slater.F90:286-287). */
int zv_last_eri_stats(const zv_context *ctx, double *ms, int64_t *launches,
                      double *units, double *flops);

#ifdef __cplusplus
}
#endif
#endif /* ZVMC_H */
