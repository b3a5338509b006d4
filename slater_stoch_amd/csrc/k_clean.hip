// PSD cleanup of the ERI matrix on the GPU.
//
// Both cleanups act on the reference's product-basis matrix
// W((i,j),(k,l)) = <ij|kl>, N = n^2, column-major W(ii,jj) with
// ii = i + (j-1) n, jj = k + (l-1) n, filled from the unique quartets through
// the eight permutations of src/svd_clean.F90:97-130.
//
// zv_svd_clean follows src/svd_clean.F90:143-367: passes of a randomized
// range finder (Gaussian test matrix, Z = W_work P, q power iterations with QR
// re-orthonormalization), SVD of Y = Z^T W_work, U = Z U_Y; a singular triple
// with u.v < 0 is a negative eigenvalue, |u.v| < 0.99 an unconverged one;
// converged triples are deflated from W_work and negative ones removed from
// W; passes stop when N vectors are accepted or max|D| < 1e-12.  Finally
// moy = (1/8) sum of the eight permutations (src/svd_clean.F90:372-405).
// GEMMs are rocBLAS DGEMM.  Orthonormal bases come from shifted CholeskyQR3
// (Fukaya, Kannan, Nakatsukasa, Yamamoto, Yanagisawa, SIAM J. Sci. Comput.
// 42 (2020) A477: one Cholesky QR of the shifted Gram matrix, then two plain
// ones; DGEMM + a 512 x 512 Cholesky + TRSM per round) with rocSOLVER
// geqrf/orgqr as the fallback when a Cholesky factorization fails (rank-
// deficient panels).  The SVD of the wide r x N matrix Y = Z^T W_work goes
// through an orthonormal basis Q of Y^T: Y = (Y Q) Q^T, and the r x r factor
// Y Q is decomposed by rocSOLVER gesvdj (one-sided Jacobi).  Both give the same subspace and the
// same singular triples as the reference's Householder QR and full SVD; they
// only avoid the tall-skinny factorizations that dominate on the GPU.
//
// zv_davidson_clean is the dense branch of src/davidson_clean.F90:30-120:
// symmetric eigendecomposition (rocSOLVER syevd) and W -= sum_{e<0} e u u^T.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>

#include "zvmc_internal.h"

namespace {

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

__device__ inline void tri_decode(int64_t t, int64_t &P, int64_t &Q)
{
    int64_t p = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while (p * (p + 1) / 2 > t) p--;
    while ((p + 1) * (p + 2) / 2 <= t) p++;
    P = p;
    Q = t - p * (p + 1) / 2;
}

// (i,j,k,l) of the eight permutations, src/svd_clean.F90:99-128
__device__ inline void perm8(int s, int i, int j, int k, int l, int &a, int &b, int &c, int &d)
{
    switch (s) {
    case 0: a = i; b = j; c = k; d = l; break;
    case 1: a = k; b = j; c = i; d = l; break;
    case 2: a = k; b = l; c = i; d = j; break;
    case 3: a = i; b = l; c = k; d = j; break;
    case 4: a = j; b = i; c = l; d = k; break;
    case 5: a = j; b = k; c = l; d = i; break;
    case 6: a = l; b = k; c = j; d = i; break;
    default: a = l; b = i; c = j; d = k; break;
    }
}

// quartet t of the pair triangle -> reference (i,j,k,l) and its kcp
__device__ inline int64_t quartet_ref(int64_t t, const int *pi, const int *pk,
                                      const int64_t *kbase, int n, int &i, int &j, int &k, int &l)
{
    int64_t P, Q;
    tri_decode(t, P, Q);
    i = pi[P];
    k = pk[P];
    j = pi[Q];
    l = pk[Q];
    return zv_kcp(kbase, n, i, j, k, l);
}

__global__ void k_build_W(int64_t ntri, const double *eri, double *W, int n, const int *pi,
                          const int *pk, const int64_t *kbase)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntri) return;
    int i, j, k, l;
    int64_t kcp = quartet_ref(t, pi, pk, kbase, n, i, j, k, l);
    double v = eri[kcp];
    int64_t N = (int64_t)n * n;
    for (int s = 0; s < 8; s++) {
        int a, b, c, d;
        perm8(s, i, j, k, l, a, b, c, d);
        int64_t ii = (a - 1) + (int64_t)(b - 1) * n, jj = (c - 1) + (int64_t)(d - 1) * n;
        W[ii + jj * N] = v;
    }
}

// moy(kcp): average of the eight permutations (svd_clean) or the first one
// (davidson_clean, src/davidson_clean.F90:110-114)
__global__ void k_gather(int64_t ntri, const double *W, double *eri, int n, const int *pi,
                         const int *pk, const int64_t *kbase, int avg8)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntri) return;
    int i, j, k, l;
    int64_t kcp = quartet_ref(t, pi, pk, kbase, n, i, j, k, l);
    int64_t N = (int64_t)n * n;
    double m = 0.0;
    int ns = avg8 ? 8 : 1;
    for (int s = 0; s < ns; s++) {
        int a, b, c, d;
        perm8(s, i, j, k, l, a, b, c, d);
        int64_t ii = (a - 1) + (int64_t)(b - 1) * n, jj = (c - 1) + (int64_t)(d - 1) * n;
        m = avg8 ? m + W[ii + jj * N] * 0.125 : W[ii + jj * N];
    }
    eri[kcp] = m;
}

__global__ void k_fill0(double *p, int64_t count)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < count) p[t] = 0.0;
}

// Philox4x32-10 normal deviates (Box-Muller, cosine branch as the reference,
// src/svd_clean.F90:161-167), column k of an N x r matrix
__device__ inline void philox(uint32_t &c0, uint32_t &c1, uint32_t &c2, uint32_t &c3, uint32_t k0,
                              uint32_t k1)
{
    for (int r = 0; r < 10; r++) {
        uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
        uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

__global__ void k_randn(int64_t total, uint64_t seed, uint32_t pass, double *P)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    uint32_t c0 = (uint32_t)t, c1 = (uint32_t)(t >> 32), c2 = pass, c3 = 0x5c1ea2u;
    philox(c0, c1, c2, c3, (uint32_t)seed, (uint32_t)(seed >> 32));
    double u1 = ((double)(((uint64_t)(c0 >> 5) << 26) | (c1 >> 6)) + 0.5) * 0x1.0p-53;
    double u2 = ((double)(((uint64_t)(c2 >> 5) << 26) | (c3 >> 6)) + 0.5) * 0x1.0p-53;
    P[t] = sqrt(-2.0 * log(u1)) * cos(6.283185307179586477 * u2);
}

// one workgroup per column: x(:,k) /= ||x(:,k)||
__global__ void k_colnormalize(int64_t N, double *X)
{
    __shared__ double red[256];
    double *x = X + (int64_t)blockIdx.x * N;
    double s = 0.0;
    for (int64_t t = threadIdx.x; t < N; t += 256) s += x[t] * x[t];
    red[threadIdx.x] = s;
    __syncthreads();
    for (unsigned o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    double inv = 1.0 / sqrt(red[0]);
    for (int64_t t = threadIdx.x; t < N; t += 256) x[t] *= inv;
}

// dots[k] = U(:,k) . Vt(k,:)   (U N x r ld N, Vt r x N ld r)
__global__ void k_coldot(int64_t N, int r, const double *U, const double *Vt, double *dots)
{
    __shared__ double red[256];
    int k = blockIdx.x;
    double s = 0.0;
    for (int64_t t = threadIdx.x; t < N; t += 256) s += U[t + (int64_t)k * N] * Vt[k + t * r];
    red[threadIdx.x] = s;
    __syncthreads();
    for (unsigned o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) dots[k] = red[0];
}

// out(:,k) = U(:,k) * f(D_k): mode 0 |D|, mode 1 min(D, 0)
__global__ void k_colscale(int64_t N, int r, const double *U, const double *D, double *out,
                           int mode)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N * r) return;
    int k = (int)(t / N);
    double d = D[k];
    double f = mode == 0 ? fabs(d) : (d < 0.0 ? d : 0.0);
    out[t] = U[t] * f;
}

// G += (f * trace G) I, one workgroup: the shift of the first CholeskyQR
// round (trace G = ||Z||_F^2 bounds ||Z||_2^2 from above)
__global__ void k_shift_diag(double *G, int r, double f)
{
    __shared__ double red[256];
    double s = 0.0;
    for (int k = threadIdx.x; k < r; k += 256) s += G[(int64_t)k * r + k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (unsigned o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    const double shift = f * red[0];
    for (int k = threadIdx.x; k < r; k += 256) G[(int64_t)k * r + k] += shift;
}

// flag = 1 when the Cholesky factor R (upper, in G) has min R_kk < thr max
// R_kk (or a non-positive / non-finite pivot), else 0
__global__ void k_pivot_check(const double *G, int r, double thr, rocblas_int *flag)
{
    __shared__ double mn[256], mx[256];
    double lo = INFINITY, hi = 0.0;
    for (int k = threadIdx.x; k < r; k += 256) {
        const double d = G[(int64_t)k * r + k];
        lo = d < lo ? d : lo;
        hi = d > hi ? d : hi;
        if (!(d > 0.0) || !isfinite(d)) lo = -1.0;
    }
    mn[threadIdx.x] = lo;
    mx[threadIdx.x] = hi;
    __syncthreads();
    for (unsigned o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            mn[threadIdx.x] = fmin(mn[threadIdx.x], mn[threadIdx.x + o]);
            mx[threadIdx.x] = fmax(mx[threadIdx.x], mx[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *flag = (mn[0] > thr * mx[0]) ? 0 : 1;
}

// G[t] = sum_b part[b * rr + t], b < nb
__global__ void k_sum_parts(int64_t rr, int nb, const double *part, double *G)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= rr) return;
    double s = 0.0;
    for (int b = 0; b < nb; b++) s += part[(int64_t)b * rr + t];
    G[t] = s;
}

} // namespace

// one rocBLAS handle per context, bound to the context's stream
static int blas_handle(zv_context *c, rocblas_handle *h)
{
    if (!c->blas) {
        rocblas_handle nh;
        if (rocblas_create_handle(&nh) != rocblas_status_success)
            return zv_fail(c, ZV_ERR_HIP, "rocblas_create_handle failed");
        rocblas_set_stream(nh, c->stream);
        c->blas = nh;
    }
    *h = (rocblas_handle)c->blas;
    return ZV_OK;
}

void zv_blas_destroy(zv_context *c)
{
    if (c->blas) rocblas_destroy_handle((rocblas_handle)c->blas);
    c->blas = nullptr;
}

#define RB(c, call)                                                                               \
    do {                                                                                          \
        rocblas_status s_ = (call);                                                               \
        if (s_ != rocblas_status_success)                                                         \
            return zv_fail((c), ZV_ERR_HIP, "%s failed: status %d", #call, (int)s_);              \
    } while (0)

// W (N x N) from the eri list in kcp order (host), 8 permutations
static int build_W(zv_context *c, const double *eri_host, double *d_eri, double *W)
{
    int64_t ntri = c->nint, N = (int64_t)c->n * c->n;
    ZV_HIPCHECK(c, hipMemcpyAsync(d_eri, eri_host, sizeof(double) * ntri, hipMemcpyHostToDevice,
                                  c->stream));
    hipLaunchKernelGGL(k_fill0, dim3(nblk(N * N, 256)), dim3(256), 0, c->stream, W, N * N);
    hipLaunchKernelGGL(k_build_W, dim3(nblk(ntri, 256)), dim3(256), 0, c->stream, ntri,
                       (const double *)d_eri, W, c->n, c->d_pair_i, c->d_pair_k, c->d_kbase);
    ZV_HIPCHECK(c, hipGetLastError());
    return ZV_OK;
}

static int qr_inplace(zv_context *c, rocblas_handle h, double *Z, int64_t N, int r, double *tau)
{
    RB(c, rocsolver_dgeqrf(h, (rocblas_int)N, r, Z, (rocblas_int)N, tau));
    RB(c, rocsolver_dorgqr(h, (rocblas_int)N, r, r, Z, (rocblas_int)N, tau));
    return ZV_OK;
}

// Scratch of the orthonormalization: G (r x r), a copy of the panel
// (N x r), Householder tau (r), three Cholesky info words and two pivot
// flags.
struct OrthoWork {
    double *G, *backup, *tau;
    rocblas_int *info;
    double *part; // kGramSplit + 1 partial r x r products
};

// C (r x r) = A^T B for N x r panels A, B (leading dimension N): rocBLAS runs
// the single r x r x N product at 3.3 TFLOP/s (N = 40000, r = 512: too few
// output tiles), so the rows are split into kGramSplit chunks computed as one
// strided batch (plus the remainder rows), then summed.
constexpr int kGramSplit = 16;

static int gram_tn(zv_context *c, rocblas_handle h, const double *A, const double *B, int64_t N,
                   int r, double *C, double *part)
{
    const double one = 1.0, zero = 0.0;
    const int64_t chunk = N / kGramSplit, rr = (int64_t)r * r;
    if (chunk < 256) {
        RB(c, rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, r, r,
                            (rocblas_int)N, &one, A, (rocblas_int)N, B, (rocblas_int)N, &zero, C,
                            r));
        return ZV_OK;
    }
    RB(c, rocblas_dgemm_strided_batched(h, rocblas_operation_transpose, rocblas_operation_none, r,
                                        r, (rocblas_int)chunk, &one, A, (rocblas_int)N, chunk, B,
                                        (rocblas_int)N, chunk, &zero, part, r, rr, kGramSplit));
    const int64_t rest = N - chunk * kGramSplit;
    int nb = kGramSplit;
    if (rest > 0) {
        RB(c, rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, r, r,
                            (rocblas_int)rest, &one, A + chunk * kGramSplit, (rocblas_int)N,
                            B + chunk * kGramSplit, (rocblas_int)N, &zero,
                            part + (int64_t)kGramSplit * rr, r));
        nb++;
    }
    hipLaunchKernelGGL(k_sum_parts, dim3((unsigned)((rr + 255) / 256)), dim3(256), 0, c->stream,
                       rr, nb, (const double *)part, C);
    ZV_HIPCHECK(c, hipGetLastError());
    return ZV_OK;
}

// Z (N x r, N >= r) <- an orthonormal basis of its columns, in place:
// shifted CholeskyQR3; Householder QR from the saved panel when one of the
// three Cholesky factorizations fails, or when a factor of the unshifted
// rounds has a pivot ratio below 1e-6.  For a panel with condition number
// below 1/u the shifted round leaves cond(Q1) <= ~1/sqrt(11 N r u) (pivot
// ratio >= 6.7e-6 at N r = 3.5e4, >= 1.6e-4 at 2e7); a rank-deficient panel
// gives ratios near sqrt(u) = 1e-8, where a "successful" Cholesky would
// return columns outside the panel's range.
static int orthonormalize(zv_context *c, rocblas_handle h, double *Z, int64_t N, int r,
                          const OrthoWork &w, int *n_fallback)
{
    const double one = 1.0;
    const double f = 11.0 * ((double)N * r + (double)r * (r + 1)) * 0x1.0p-53;
    ZV_HIPCHECK(c, hipMemcpyAsync(w.backup, Z, sizeof(double) * N * r, hipMemcpyDeviceToDevice,
                                  c->stream));
    for (int round = 0; round < 3; round++) {
        ZV_CHECK(gram_tn(c, h, Z, Z, N, r, w.G, w.part));
        if (round == 0) {
            hipLaunchKernelGGL(k_shift_diag, dim3(1), dim3(256), 0, c->stream, w.G, r, f);
            ZV_HIPCHECK(c, hipGetLastError());
        }
        RB(c, rocsolver_dpotrf(h, rocblas_fill_upper, r, w.G, r, w.info + round));
        if (round > 0) {
            hipLaunchKernelGGL(k_pivot_check, dim3(1), dim3(256), 0, c->stream, (const double *)w.G,
                               r, 1.0e-6, w.info + 2 + round);
            ZV_HIPCHECK(c, hipGetLastError());
        }
        RB(c, rocblas_dtrsm(h, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                            rocblas_diagonal_non_unit, (rocblas_int)N, r, &one, w.G, r, Z,
                            (rocblas_int)N));
    }
    rocblas_int hinfo[5] = {0, 0, 0, 0, 0};
    ZV_HIPCHECK(c, hipMemcpyAsync(hinfo, w.info, sizeof(hinfo), hipMemcpyDeviceToHost, c->stream));
    ZV_HIPCHECK(c, hipStreamSynchronize(c->stream));
    if (!(hinfo[0] | hinfo[1] | hinfo[2] | hinfo[3] | hinfo[4])) return ZV_OK;
    ZV_HIPCHECK(c, hipMemcpyAsync(Z, w.backup, sizeof(double) * N * r, hipMemcpyDeviceToDevice,
                                  c->stream));
    if (n_fallback) (*n_fallback)++;
    return qr_inplace(c, h, Z, N, r, w.tau);
}

int zv_run_svd_clean(zv_context *c, int rank, int q, uint64_t seed, double *eri, int *n_removed)
{
    const int64_t N = (int64_t)c->n * c->n;
    if (N > 65536) return zv_fail(c, ZV_ERR_ARG, "svd_clean: n^2 = %lld too large", (long long)N);
    const int r = (int)std::min<int64_t>(rank, N);
    rocblas_handle h;
    ZV_CHECK(blas_handle(c, &h));
    PhaseTimer tm(c, PH_CLEAN);
    // workspace: W, Wwork (N^2), P, Z, U, S (N r), Yt/Vt (N r), UY, Rt, Vs, G (r^2),
    // small vectors
    size_t nn = (size_t)N * N, nr = (size_t)N * r, rr = (size_t)r * r;
    size_t words = 2 * nn + 6 * nr + 4 * rr + 8 * (size_t)r + c->nint + 16;
    ZV_CHECK(zv_ensure(c, c->b_misc, sizeof(double) * words));
    double *W = (double *)c->b_misc.p, *Ww = W + nn, *P = Ww + nn, *Z = P + nr, *U = Z + nr,
           *S = U + nr, *Yt = S + nr, *Vt = Yt + nr, *UY = Vt + nr, *Rt = UY + rr, *Vs = Rt + rr,
           *G = Vs + rr, *D = G + rr, *E = D + r, *tau = E + r, *dots = tau + r,
           *d_eri = dots + r + 4;
    // info words: [0] SVD, [1, 6) orthonormalization, [6] Jacobi sweeps
    rocblas_int *info = (rocblas_int *)(d_eri + c->nint);
    // power iterations keep their panel copy in S, the SVD's in U (both are
    // written only after the SVD)
    ZV_CHECK(zv_ensure(c, c->b_part, sizeof(double) * (kGramSplit + 1) * rr));
    double *part = (double *)c->b_part.p;
    const OrthoWork ow_iter{G, S, tau, info + 1, part}, ow_svd{G, U, tau, info + 1, part};
    int n_fallback = 0;
    ZV_CHECK(build_W(c, eri, d_eri, Ww));
    ZV_HIPCHECK(c, hipMemcpyAsync(W, Ww, sizeof(double) * nn, hipMemcpyDeviceToDevice, c->stream));
    const double one = 1.0, zero = 0.0, mone = -1.0;
    int64_t npass = 0;
    int removed = 0;
    std::vector<double> hD(r), hdots(r);
    // the reference loops until N vectors are accepted; every pass either
    // deflates a converged triple or ends the loop through max|D| < 1e-12, the
    // bound below only guards against a pathological spectrum
    const uint32_t max_pass = (uint32_t)(4 * (N / r) + 16);
    for (uint32_t pass = 0; npass < N; pass++) {
        if (pass >= max_pass)
            return zv_fail(c, ZV_ERR_STATE, "svd_clean: no convergence after %u passes", pass);
        if (c->verbose) {
            printf(" Vectors %lld -- %lld / %lld\n", (long long)npass + 1,
                   (long long)std::min<int64_t>(npass + r, N), (long long)N);
            fflush(stdout);
        }
        hipLaunchKernelGGL(k_randn, dim3(nblk(nr, 256)), dim3(256), 0, c->stream, (int64_t)nr,
                           seed, pass, P);
        hipLaunchKernelGGL(k_colnormalize, dim3(r), dim3(256), 0, c->stream, N, P);
        ZV_HIPCHECK(c, hipGetLastError());
        RB(c, rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, N, r, N, &one, Ww, N,
                            P, N, &zero, Z, N));
        ZV_CHECK(orthonormalize(c, h, Z, N, r, ow_iter, &n_fallback));
        for (int it = 0; it < q; it++) {
            RB(c, rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, N, r, N,
                                &one, Ww, N, Z, N, &zero, P, N));
            RB(c, rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, N, r, N, &one,
                                Ww, N, P, N, &zero, Z, N));
            ZV_CHECK(orthonormalize(c, h, Z, N, r, ow_iter, &n_fallback));
        }
        // SVD of Y = Z^T W_work (r x N): Yt = Y^T, Q = orthonormal basis of
        // Yt (in S), Rt = Y Q (r x r) = UY diag(D) Vs, so Y = UY diag(D) (Vs Q^T)
        RB(c, rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, N, r, N, &one,
                            Ww, N, Z, N, &zero, Yt, N));
        ZV_HIPCHECK(c, hipMemcpyAsync(S, Yt, sizeof(double) * nr, hipMemcpyDeviceToDevice,
                                      c->stream));
        ZV_CHECK(orthonormalize(c, h, S, N, r, ow_svd, &n_fallback));
        ZV_CHECK(gram_tn(c, h, Yt, S, N, r, Rt, part));
        // one-sided Jacobi: accurate singular vectors inside the clusters of
        // the late passes (rocSOLVER's divide-and-conquer gesdd lost the
        // small negative eigenvalues of a 110-AO ERI matrix there; gesvd is
        // accurate but 660 ms at r = 512)
        RB(c, rocsolver_dgesvdj(h, rocblas_svect_singular, rocblas_svect_singular, r, r, Rt, r,
                                0.0, E, 100, info + 6, D, UY, r, Vs, r, info));
        RB(c, rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_transpose, r, N, r, &one,
                            Vs, r, S, N, &zero, Vt, r));
        RB(c, rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, N, r, r, &one, Z, N,
                            UY, r, &zero, U, N));
        hipLaunchKernelGGL(k_coldot, dim3(r), dim3(256), 0, c->stream, N, r, U, Vt, dots);
        ZV_HIPCHECK(c, hipGetLastError());
        ZV_HIPCHECK(c, hipMemcpyAsync(hD.data(), D, sizeof(double) * r, hipMemcpyDeviceToHost,
                                      c->stream));
        ZV_HIPCHECK(c, hipMemcpyAsync(hdots.data(), dots, sizeof(double) * r,
                                      hipMemcpyDeviceToHost, c->stream));
        rocblas_int hinfo = 0;
        ZV_HIPCHECK(c, hipMemcpyAsync(&hinfo, info, sizeof(rocblas_int), hipMemcpyDeviceToHost,
                                      c->stream));
        ZV_HIPCHECK(c, hipStreamSynchronize(c->stream));
        if (hinfo != 0) return zv_fail(c, ZV_ERR_HIP, "%d: SVD failed", (int)hinfo);
        if (c->verbose) {
            printf(" Largest/Smallest singular value : %g %g\n", (double)(float)hD[0],
                   (double)(float)hD[r - 1]);
            if (n_fallback) printf(" Householder QR fallbacks so far: %d\n", n_fallback);
            fflush(stdout);
        }
        // src/svd_clean.F90:286-296
        npass += r;
        double dmax = 0.0;
        int negative = 0;
        for (int k = 0; k < r; k++) {
            if (hdots[k] < 0.0) hD[k] = -hD[k];
            if (fabs(hdots[k]) < 0.99) {
                hD[k] = 0.0;
                npass -= 1;
            }
            dmax = std::max(dmax, fabs(hD[k]));
            if (hD[k] < 0.0) negative++;
        }
        removed += negative;
        ZV_HIPCHECK(c, hipMemcpyAsync(D, hD.data(), sizeof(double) * r, hipMemcpyHostToDevice,
                                      c->stream));
        // W_work -= U |D| Vt ; W += U min(D,0) Vt  (src/svd_clean.F90:334-353)
        hipLaunchKernelGGL(k_colscale, dim3(nblk(nr, 256)), dim3(256), 0, c->stream, N, r, U, D,
                           S, 0);
        RB(c, rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, N, N, r, &mone, S,
                            N, Vt, r, &one, Ww, N));
        if (negative) { // min(D, 0) = 0 for every triple otherwise
            hipLaunchKernelGGL(k_colscale, dim3(nblk(nr, 256)), dim3(256), 0, c->stream, N, r, U,
                               D, S, 1);
            RB(c, rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, N, N, r, &one,
                                S, N, Vt, r, &one, W, N));
        }
        ZV_HIPCHECK(c, hipGetLastError());
        if (dmax < 1.0e-12) break; // src/svd_clean.F90:359
    }
    hipLaunchKernelGGL(k_gather, dim3(nblk(c->nint, 256)), dim3(256), 0, c->stream, c->nint,
                       (const double *)W, d_eri, c->n, c->d_pair_i, c->d_pair_k, c->d_kbase, 1);
    ZV_HIPCHECK(c, hipGetLastError());
    ZV_HIPCHECK(c, hipMemcpyAsync(eri, d_eri, sizeof(double) * c->nint, hipMemcpyDeviceToHost,
                                  c->stream));
    ZV_HIPCHECK(c, hipStreamSynchronize(c->stream));
    if (n_removed) *n_removed = removed;
    return ZV_OK;
}

int zv_run_davidson_clean(zv_context *c, double *eri, int *n_removed)
{
    const int64_t N = (int64_t)c->n * c->n;
    if (N > 65536) return zv_fail(c, ZV_ERR_ARG, "davidson_clean: n^2 = %lld too large",
                                  (long long)N);
    rocblas_handle h;
    ZV_CHECK(blas_handle(c, &h));
    PhaseTimer tm(c, PH_CLEAN);
    size_t nn = (size_t)N * N;
    size_t words = 2 * nn + 2 * (size_t)N + c->nint + 16;
    ZV_CHECK(zv_ensure(c, c->b_misc, sizeof(double) * words));
    double *W = (double *)c->b_misc.p, *U = W + nn, *ev = U + nn, *E = ev + N, *d_eri = E + N;
    rocblas_int *info = (rocblas_int *)(d_eri + c->nint);
    ZV_CHECK(build_W(c, eri, d_eri, W));
    ZV_HIPCHECK(c, hipMemcpyAsync(U, W, sizeof(double) * nn, hipMemcpyDeviceToDevice, c->stream));
    RB(c, rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, (rocblas_int)N, U,
                           (rocblas_int)N, ev, E, info));
    std::vector<double> hev(N);
    rocblas_int hinfo = 0;
    ZV_HIPCHECK(c, hipMemcpyAsync(hev.data(), ev, sizeof(double) * N, hipMemcpyDeviceToHost,
                                  c->stream));
    ZV_HIPCHECK(c, hipMemcpyAsync(&hinfo, info, sizeof(rocblas_int), hipMemcpyDeviceToHost,
                                  c->stream));
    ZV_HIPCHECK(c, hipStreamSynchronize(c->stream));
    if (hinfo != 0) return zv_fail(c, ZV_ERR_HIP, "info=%d: DSYGV Diagonalization failed", hinfo);
    int l = 0; // eigenvalues ascend: the first l are negative (src/davidson_clean.F90:92-96)
    while (l < N && hev[l] < 0.0) l++;
    if (l > 0) {
        // W -= U_l diag(e_l) U_l^T, written as W += (U_l diag(-e_l)) U_l^T
        std::vector<double> hd(l);
        for (int k = 0; k < l; k++) hd[k] = -hev[k];
        ZV_HIPCHECK(c, hipMemcpyAsync(E, hd.data(), sizeof(double) * l, hipMemcpyHostToDevice,
                                      c->stream));
        const double one = 1.0;
        ZV_CHECK(zv_ensure(c, c->b_part, sizeof(double) * (size_t)N * l));
        double *Sl = (double *)c->b_part.p;
        hipLaunchKernelGGL(k_colscale, dim3(nblk((int64_t)N * l, 256)), dim3(256), 0, c->stream,
                           N, l, U, E, Sl, 0);
        ZV_HIPCHECK(c, hipGetLastError());
        RB(c, rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_transpose, N, N, l, &one,
                            Sl, N, U, N, &one, W, N));
    }
    hipLaunchKernelGGL(k_gather, dim3(nblk(c->nint, 256)), dim3(256), 0, c->stream, c->nint,
                       (const double *)W, d_eri, c->n, c->d_pair_i, c->d_pair_k, c->d_kbase, 0);
    ZV_HIPCHECK(c, hipGetLastError());
    ZV_HIPCHECK(c, hipMemcpyAsync(eri, d_eri, sizeof(double) * c->nint, hipMemcpyDeviceToHost,
                                  c->stream));
    ZV_HIPCHECK(c, hipStreamSynchronize(c->stream));
    if (n_removed) *n_removed = l;
    return ZV_OK;
}

// ====================================================================== pair supermatrix
// The pair-packed ERI supermatrix V[P + Q N] = (ik|jl), P = (i,k), Q = (j,l)
// AO pairs (i >= k, j >= l, lexicographic, N = n(n+1)/2), symmetric and
// positive semidefinite in exact arithmetic (1/r12 is a positive definite
// kernel); Monte Carlo noise leaves small negative eigenvalues.  Two GPU
// cleanups act on it:
//
// zv_rsvd_clean: randomized range finder of rank r with q power iterations
// (src/randomized_svd.F90:22-65: Gaussian test matrix, Z = V P, Z <- V (V^T Z)
// with re-orthonormalization), then Rayleigh-Ritz in FP64 (B = Q^T V Q,
// symmetric eigendecomposition) instead of the SVD of Q^T V, so that the sign
// of every Ritz value is explicit; negative Ritz pairs are removed,
// V -= sum lambda (Q y)(Q y)^T.  The sketch products (the N x N x r GEMMs of
// the range finder) run in FP32 when asked (SGEMM on an FP32 copy of V), the
// orthonormalizations and the Rayleigh-Ritz step in FP64.
//
// zv_davidson_lowest: block Davidson for the nev lowest eigenpairs of V
// (the role of src/davidson.F90:1-159 in src/davidson_clean.F90:120-323):
// Ritz vectors of the search space, residuals R = V X - X Theta; the
// residuals of the unconverged Ritz vectors, orthogonalized against the
// space and orthonormalized (CholeskyQR3, Householder fallback), extend it;
// thick restart to the current Ritz vectors when the space is full.  Every
// projection keeps the basis orthonormal, so the small problem is a standard
// symmetric one (the reference's generalized one lost definiteness).  No
// diagonal (Jacobi) preconditioner: the wanted eigenvalues lie near zero
// while diag V = (P|P) spans orders of magnitude, so R / (diag V - theta)
// damps exactly the components the wanted vectors carry (on a noisy 30-AO
// matrix it had not converged after 2000 iterations; without it, 167).  The
// optional shift-invert preconditioner (V - sigma I)^-1, with sigma below
// the spectrum, serves instead (shift_factor).  The converged negative pairs
// are removed from V.
namespace {

// V from the quartet list: the kcp entry of (P, Q) at V[P,Q] and V[Q,P]
__global__ void k_build_V(int64_t ntri, const double *eri, double *V, int64_t N, int n,
                          const int *pi, const int *pk, const int64_t *kbase)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntri) return;
    int i, j, k, l;
    const int64_t kcp = quartet_ref(t, pi, pk, kbase, n, i, j, k, l);
    int64_t P, Q;
    tri_decode(t, P, Q);
    const double v = eri[kcp];
    V[P + Q * N] = v;
    V[Q + P * N] = v;
}

// eri[kcp] = (V[P,Q] + V[Q,P]) / 2
__global__ void k_gather_V(int64_t ntri, const double *V, double *eri, int64_t N, int n,
                           const int *pi, const int *pk, const int64_t *kbase)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntri) return;
    int i, j, k, l;
    const int64_t kcp = quartet_ref(t, pi, pk, kbase, n, i, j, k, l);
    int64_t P, Q;
    tri_decode(t, P, Q);
    eri[kcp] = 0.5 * (V[P + Q * N] + V[Q + P * N]);
}

__global__ void k_to_f32(int64_t n, const double *x, float *y)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) y[t] = (float)x[t];
}

__global__ void k_to_f64(int64_t n, const float *x, double *y)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) y[t] = (double)x[t];
}

// B = (B + B^T) / 2 for an r x r matrix (leading dimension r)
__global__ void k_symmetrize(int r, double *B)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)r * r) return;
    const int i = (int)(t % r), j = (int)(t / r);
    if (i >= j) return;
    const double m = 0.5 * (B[i + (int64_t)j * r] + B[j + (int64_t)i * r]);
    B[i + (int64_t)j * r] = m;
    B[j + (int64_t)i * r] = m;
}

// R = AX - X diag(theta), one thread per entry
__global__ void k_residual(int64_t N, int k, const double *AX, const double *X,
                           const double *theta, double *R)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N * k) return;
    const int j = (int)(t / N);
    R[t] = AX[t] - X[t] * theta[j];
}

// out[j] = ||X(:,j)||, one workgroup per column
__global__ void k_colnorms(int64_t N, const double *X, double *out)
{
    __shared__ double red[256];
    const double *x = X + (int64_t)blockIdx.x * N;
    double s = 0.0;
    for (int64_t t = threadIdx.x; t < N; t += 256) s += x[t] * x[t];
    red[threadIdx.x] = s;
    __syncthreads();
    for (unsigned o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sqrt(red[0]);
}

// fill the lower triangle of the m x m leading block of H (ld) from its upper one
__global__ void k_mirror(int m, int ld, double *H)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)m * m) return;
    const int i = (int)(t % m), j = (int)(t / m);
    if (i > j) H[i + (int64_t)j * ld] = H[j + (int64_t)i * ld];
}

// F = V - sigma I
__global__ void k_shifted_copy(int64_t N, const double *V, double sigma, double *F)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N * N) return;
    F[t] = V[t] - ((t % N) == (t / N) ? sigma : 0.0);
}

} // namespace

// V (N x N, device) from the host quartet list (kcp order)
static int build_V(zv_context *c, const double *eri_host, double *d_eri, double *V, int64_t N)
{
    ZV_HIPCHECK(c, hipMemcpyAsync(d_eri, eri_host, sizeof(double) * c->nint,
                                  hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_build_V, dim3(nblk(c->nint, 256)), dim3(256), 0, c->stream, c->nint,
                       (const double *)d_eri, V, N, c->n, c->d_pair_i, c->d_pair_k, c->d_kbase);
    ZV_HIPCHECK(c, hipGetLastError());
    return ZV_OK;
}

static int gather_V(zv_context *c, const double *V, double *d_eri, double *eri_host, int64_t N)
{
    hipLaunchKernelGGL(k_gather_V, dim3(nblk(c->nint, 256)), dim3(256), 0, c->stream, c->nint, V,
                       d_eri, N, c->n, c->d_pair_i, c->d_pair_k, c->d_kbase);
    ZV_HIPCHECK(c, hipGetLastError());
    ZV_HIPCHECK(c, hipMemcpyAsync(eri_host, d_eri, sizeof(double) * c->nint,
                                  hipMemcpyDeviceToHost, c->stream));
    ZV_HIPCHECK(c, hipStreamSynchronize(c->stream));
    return ZV_OK;
}

// V -= X diag(lam) X^T for the first m columns of X (N x m) with lam[j] < 0
// (scaled copy in Xs)
static int remove_pairs(zv_context *c, rocblas_handle h, double *V, int64_t N, const double *X,
                        const double *d_lam, int m, double *Xs)
{
    if (m <= 0) return ZV_OK;
    const double one = 1.0, mone = -1.0;
    hipLaunchKernelGGL(k_colscale, dim3(nblk(N * m, 256)), dim3(256), 0, c->stream, N, m, X,
                       d_lam, Xs, 1);
    ZV_HIPCHECK(c, hipGetLastError());
    RB(c, rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_transpose, (rocblas_int)N,
                        (rocblas_int)N, m, &mone, Xs, (rocblas_int)N, X, (rocblas_int)N, &one, V,
                        (rocblas_int)N));
    return ZV_OK;
}

int zv_run_rsvd_clean(zv_context *c, int rank, int q, uint64_t seed, int sketch32, double *eri,
                      double *ritz, int *n_removed)
{
    const int64_t N = c->npair;
    const int r = (int)std::min<int64_t>(rank, N);
    rocblas_handle h;
    ZV_CHECK(blas_handle(c, &h));
    PhaseTimer tm(c, PH_CLEAN);
    const size_t nn = (size_t)N * N, nr = (size_t)N * r, rr = (size_t)r * r;
    // V, its FP32 copy (sketch32), Z, AQ, backup (N r), FP32 panels (2 N r
    // floats = N r doubles), B, (r^2), lam, E, tau (r), the quartet list, info
    const size_t words = nn + (sketch32 ? nn / 2 + 1 : 0) + 4 * nr + rr + 3 * (size_t)r +
                         (size_t)c->nint + 16;
    ZV_CHECK(zv_ensure(c, c->b_misc, sizeof(double) * words));
    ZV_CHECK(zv_ensure(c, c->b_part, sizeof(double) * (kGramSplit + 1) * rr));
    double *V = (double *)c->b_misc.p;
    float *V32 = (float *)(V + nn);
    double *Z = V + nn + (sketch32 ? nn / 2 + 1 : 0), *AQ = Z + nr, *bk = AQ + nr;
    float *F1 = (float *)(bk + nr), *F2 = F1 + nr;
    double *B = bk + 2 * nr, *lam = B + rr, *E = lam + r, *tau = E + r, *d_eri = tau + r;
    rocblas_int *info = (rocblas_int *)(d_eri + c->nint);
    double *part = (double *)c->b_part.p;
    const OrthoWork ow{B, bk, tau, info + 1, part};
    int n_fallback = 0;
    ZV_CHECK(build_V(c, eri, d_eri, V, N));
    const double one = 1.0, zero = 0.0;
    const float onef = 1.0f, zerof = 0.0f;
    hipLaunchKernelGGL(k_randn, dim3(nblk(nr, 256)), dim3(256), 0, c->stream, (int64_t)nr, seed,
                       0u, AQ);
    hipLaunchKernelGGL(k_colnormalize, dim3(r), dim3(256), 0, c->stream, N, AQ);
    ZV_HIPCHECK(c, hipGetLastError());
    // Z = V P, then q times Z <- V (V Z) (V symmetric: V^T = V), each product
    // pair followed by an FP64 orthonormalization
    if (sketch32) {
        hipLaunchKernelGGL(k_to_f32, dim3(nblk(nn, 256)), dim3(256), 0, c->stream, (int64_t)nn,
                           (const double *)V, V32);
        hipLaunchKernelGGL(k_to_f32, dim3(nblk(nr, 256)), dim3(256), 0, c->stream, (int64_t)nr,
                           (const double *)AQ, F1);
        RB(c, rocblas_sgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)N, r,
                            (rocblas_int)N, &onef, V32, (rocblas_int)N, F1, (rocblas_int)N,
                            &zerof, F2, (rocblas_int)N));
        hipLaunchKernelGGL(k_to_f64, dim3(nblk(nr, 256)), dim3(256), 0, c->stream, (int64_t)nr,
                           (const float *)F2, Z);
    } else {
        RB(c, rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)N, r,
                            (rocblas_int)N, &one, V, (rocblas_int)N, AQ, (rocblas_int)N, &zero, Z,
                            (rocblas_int)N));
    }
    ZV_HIPCHECK(c, hipGetLastError());
    ZV_CHECK(orthonormalize(c, h, Z, N, r, ow, &n_fallback));
    for (int it = 0; it < q; it++) {
        if (sketch32) {
            hipLaunchKernelGGL(k_to_f32, dim3(nblk(nr, 256)), dim3(256), 0, c->stream,
                               (int64_t)nr, (const double *)Z, F1);
            RB(c, rocblas_sgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)N,
                                r, (rocblas_int)N, &onef, V32, (rocblas_int)N, F1, (rocblas_int)N,
                                &zerof, F2, (rocblas_int)N));
            RB(c, rocblas_sgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)N,
                                r, (rocblas_int)N, &onef, V32, (rocblas_int)N, F2, (rocblas_int)N,
                                &zerof, F1, (rocblas_int)N));
            hipLaunchKernelGGL(k_to_f64, dim3(nblk(nr, 256)), dim3(256), 0, c->stream,
                               (int64_t)nr, (const float *)F1, Z);
        } else {
            RB(c, rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)N,
                                r, (rocblas_int)N, &one, V, (rocblas_int)N, Z, (rocblas_int)N,
                                &zero, AQ, (rocblas_int)N));
            RB(c, rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)N,
                                r, (rocblas_int)N, &one, V, (rocblas_int)N, AQ, (rocblas_int)N,
                                &zero, Z, (rocblas_int)N));
        }
        ZV_HIPCHECK(c, hipGetLastError());
        ZV_CHECK(orthonormalize(c, h, Z, N, r, ow, &n_fallback));
    }
    // Rayleigh-Ritz in FP64: B = Z^T V Z, B = Y diag(lam) Y^T (ascending)
    RB(c, rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)N, r,
                        (rocblas_int)N, &one, V, (rocblas_int)N, Z, (rocblas_int)N, &zero, AQ,
                        (rocblas_int)N));
    ZV_CHECK(gram_tn(c, h, Z, AQ, N, r, B, part));
    hipLaunchKernelGGL(k_symmetrize, dim3(nblk((int64_t)rr, 256)), dim3(256), 0, c->stream, r, B);
    ZV_HIPCHECK(c, hipGetLastError());
    RB(c, rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, r, B, r, lam, E, info));
    std::vector<double> hlam(r);
    rocblas_int hinfo = 0;
    ZV_HIPCHECK(c, hipMemcpyAsync(hlam.data(), lam, sizeof(double) * r, hipMemcpyDeviceToHost,
                                  c->stream));
    ZV_HIPCHECK(c, hipMemcpyAsync(&hinfo, info, sizeof(rocblas_int), hipMemcpyDeviceToHost,
                                  c->stream));
    ZV_HIPCHECK(c, hipStreamSynchronize(c->stream));
    if (hinfo != 0) return zv_fail(c, ZV_ERR_HIP, "rsvd_clean: eigensolver info %d", (int)hinfo);
    int neg = 0;
    while (neg < r && hlam[neg] < 0.0) neg++;
    int removed = 0;
    if (neg > 0) {
        // Ritz vectors of the negative values, X = Z Y(:, 0:neg) in AQ, and
        // their residuals ||V x - lambda x||: a pair is removed only when
        // converged in the sense of src/svd_clean.F90:286-296, |u.v| >= 0.99
        // for its singular vectors, which for a symmetric matrix is a
        // residual of at most sqrt(1 - 0.99^2) |lambda|
        RB(c, rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)N, neg,
                            r, &one, Z, (rocblas_int)N, B, r, &zero, AQ, (rocblas_int)N));
        double *VX = Z; // Z is no longer needed
        RB(c, rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)N, neg,
                            (rocblas_int)N, &one, V, (rocblas_int)N, AQ, (rocblas_int)N, &zero, VX,
                            (rocblas_int)N));
        hipLaunchKernelGGL(k_residual, dim3(nblk((int64_t)N * neg, 256)), dim3(256), 0,
                           c->stream, N, neg, (const double *)VX, (const double *)AQ,
                           (const double *)lam, bk);
        hipLaunchKernelGGL(k_colnorms, dim3(neg), dim3(256), 0, c->stream, N, (const double *)bk,
                           E);
        ZV_HIPCHECK(c, hipGetLastError());
        std::vector<double> hres(neg);
        ZV_HIPCHECK(c, hipMemcpyAsync(hres.data(), E, sizeof(double) * neg, hipMemcpyDeviceToHost,
                                      c->stream));
        ZV_HIPCHECK(c, hipStreamSynchronize(c->stream));
        // accepted pairs first (columns moved in place, values in lam)
        const double accept = std::sqrt(1.0 - 0.99 * 0.99);
        std::vector<double> hacc;
        for (int j = 0; j < neg; j++) {
            if (!(hres[j] <= accept * std::fabs(hlam[j]))) continue;
            if (removed != j)
                ZV_HIPCHECK(c, hipMemcpyAsync(AQ + (size_t)removed * N, AQ + (size_t)j * N,
                                              sizeof(double) * N, hipMemcpyDeviceToDevice,
                                              c->stream));
            hacc.push_back(hlam[j]);
            removed++;
        }
        if (removed > 0) {
            ZV_HIPCHECK(c, hipMemcpyAsync(E, hacc.data(), sizeof(double) * removed,
                                          hipMemcpyHostToDevice, c->stream));
            ZV_CHECK(remove_pairs(c, h, V, N, AQ, E, removed, bk));
        }
    }
    if (c->verbose) {
        printf(" rsvd_clean: rank %d, %d power iterations (%s sketch), Ritz values %g .. %g, "
               "%d negative, %d converged and removed\n",
               r, q, sketch32 ? "FP32" : "FP64", hlam[0], hlam[r - 1], neg, removed);
        fflush(stdout);
    }
    if (ritz) std::copy(hlam.begin(), hlam.end(), ritz);
    if (n_removed) *n_removed = removed;
    return gather_V(c, V, d_eri, eri, N);
}

// Shift-invert preconditioner for the lowest eigenpairs: the Cholesky
// factor of V - sigma I with sigma below the lowest eigenvalue (the
// factorization succeeds exactly when V - sigma I is positive definite,
// which certifies sigma < lambda_min), so that a correction
// (V - sigma I)^-1 r = x + (sigma - theta)(V - sigma I)^-1 x is an
// inverse-iteration step toward the eigenvalues just above sigma.  sigma
// starts at theta_1 - (2 ||r_1|| + 1e-3 |theta_1|) and steps down four
// times further on each failed factorization (at most 12); once the
// residuals have dropped a hundredfold it is moved up to theta_1 -
// 2 ||r_1|| (at most 4 factorizations).  Returns false when no shift
// factorized (the iteration then goes on unpreconditioned).
static int shift_factor(zv_context *c, rocblas_handle h, const double *V, int64_t N,
                        double theta1, double gap, double *F, rocblas_int *info, double *sigma,
                        bool *ok)
{
    *ok = false;
    for (int tries = 0; tries < 12; tries++, gap *= 4.0) {
        if (tries > 0 && gap == 0.0) break;
        const double sg = theta1 - gap;
        hipLaunchKernelGGL(k_shifted_copy, dim3(nblk(N * N, 256)), dim3(256), 0, c->stream, N, V,
                           sg, F);
        ZV_HIPCHECK(c, hipGetLastError());
        RB(c, rocsolver_dpotrf(h, rocblas_fill_upper, (rocblas_int)N, F, (rocblas_int)N, info));
        rocblas_int hinfo = 0;
        ZV_HIPCHECK(c, hipMemcpyAsync(&hinfo, info, sizeof(rocblas_int), hipMemcpyDeviceToHost,
                                      c->stream));
        ZV_HIPCHECK(c, hipStreamSynchronize(c->stream));
        if (c->verbose > 1) {
            printf(" davidson: shift %.6e %s\n", sg, hinfo == 0 ? "factorized" : "not definite");
            fflush(stdout);
        }
        if (hinfo == 0) {
            *sigma = sg;
            *ok = true;
            return ZV_OK;
        }
    }
    return ZV_OK;
}

int zv_run_davidson_lowest(zv_context *c, int nev, double tol, int max_iter, int precond,
                           double *eri, double *evals, double *resid, int *n_removed, int *iters)
{
    const int64_t N = c->npair;
    const int k = nev;
    // search space of up to 8 blocks (at least 48 vectors), never beyond N
    const int mmax = (int)std::min<int64_t>(N, std::max(8 * k, 48));
    if (k < 1 || 2 * k > mmax)
        return zv_fail(c, ZV_ERR_ARG, "davidson: nev %d out of range for %lld pairs", nev,
                       (long long)N);
    rocblas_handle h;
    ZV_CHECK(blas_handle(c, &h));
    PhaseTimer tm(c, PH_CLEAN);
    const size_t nn = (size_t)N * N, nm = (size_t)N * mmax, nk = (size_t)N * k;
    const size_t mm = (size_t)mmax * mmax;
    // V, S and AS (N mmax), X, AX, R, T, AT, bk (N k), H, Hw (mmax^2), C
    // (mmax k), theta, E, tau, norms (mmax), G (k^2), the quartet list, info
    const size_t words = nn + 2 * nm + 6 * nk + 2 * mm + (size_t)mmax * k + 4 * (size_t)mmax +
                         (size_t)k * k + (size_t)c->nint + 16 + (precond ? nn : 0);
    ZV_CHECK(zv_ensure(c, c->b_misc, sizeof(double) * words));
    ZV_CHECK(zv_ensure(c, c->b_part, sizeof(double) * (kGramSplit + 1) * (size_t)k * k));
    double *V = (double *)c->b_misc.p, *S = V + nn, *AS = S + nm, *X = AS + nm, *AX = X + nk,
           *R = AX + nk, *T = R + nk, *AT = T + nk, *bk = AT + nk, *H = bk + nk,
           *Hw = H + mm, *C = Hw + mm, *theta = C + (size_t)mmax * k, *E = theta + mmax,
           *tau = E + mmax, *nrm = tau + mmax, *G = nrm + mmax, *d_eri = G + (size_t)k * k;
    rocblas_int *info = (rocblas_int *)(d_eri + c->nint);
    double *F = precond ? d_eri + c->nint + 16 : nullptr; // Cholesky factor of V - sigma I
    double *part = (double *)c->b_part.p;
    const OrthoWork ow{G, bk, tau, info + 1, part};
    int n_fallback = 0;
    const double one = 1.0, zero = 0.0, mone = -1.0;
    ZV_CHECK(build_V(c, eri, d_eri, V, N));
    // start: k Gaussian vectors, orthonormalized
    hipLaunchKernelGGL(k_randn, dim3(nblk(nk, 256)), dim3(256), 0, c->stream, (int64_t)nk,
                       (uint64_t)0x0DA71D50u, 0u, S);
    ZV_HIPCHECK(c, hipGetLastError());
    ZV_CHECK(orthonormalize(c, h, S, N, k, ow, &n_fallback));
    RB(c, rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)N, k,
                        (rocblas_int)N, &one, V, (rocblas_int)N, S, (rocblas_int)N, &zero, AS,
                        (rocblas_int)N));
    RB(c, rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, k, k,
                        (rocblas_int)N, &one, S, (rocblas_int)N, AS, (rocblas_int)N, &zero, H,
                        mmax));
    int m = k, it = 0;
    std::vector<double> htheta(mmax), hnrm(k);
    bool converged = false;
    bool have_pc = false;
    int nfact = 0;
    double sigma = 0.0, r_at_fact = 0.0;
    for (;;) {
        // Ritz pairs of the space: Hw = H(0:m, 0:m) (upper triangle valid)
        hipLaunchKernelGGL(k_mirror, dim3(nblk((int64_t)m * m, 256)), dim3(256), 0, c->stream, m,
                           mmax, H);
        ZV_HIPCHECK(c, hipGetLastError());
        ZV_HIPCHECK(c, hipMemcpy2DAsync(Hw, sizeof(double) * m, H, sizeof(double) * mmax,
                                        sizeof(double) * m, m, hipMemcpyDeviceToDevice,
                                        c->stream));
        RB(c, rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, m, Hw, m, theta, E,
                               info));
        RB(c, rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)N, k,
                            m, &one, S, (rocblas_int)N, Hw, m, &zero, X, (rocblas_int)N));
        RB(c, rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)N, k,
                            m, &one, AS, (rocblas_int)N, Hw, m, &zero, AX, (rocblas_int)N));
        hipLaunchKernelGGL(k_residual, dim3(nblk(nk, 256)), dim3(256), 0, c->stream, N, k,
                           (const double *)AX, (const double *)X, (const double *)theta, R);
        hipLaunchKernelGGL(k_colnorms, dim3(k), dim3(256), 0, c->stream, N, (const double *)R,
                           nrm);
        ZV_HIPCHECK(c, hipGetLastError());
        rocblas_int hinfo = 0;
        ZV_HIPCHECK(c, hipMemcpyAsync(htheta.data(), theta, sizeof(double) * k,
                                      hipMemcpyDeviceToHost, c->stream));
        ZV_HIPCHECK(c, hipMemcpyAsync(hnrm.data(), nrm, sizeof(double) * k, hipMemcpyDeviceToHost,
                                      c->stream));
        ZV_HIPCHECK(c, hipMemcpyAsync(&hinfo, info, sizeof(rocblas_int), hipMemcpyDeviceToHost,
                                      c->stream));
        ZV_HIPCHECK(c, hipStreamSynchronize(c->stream));
        if (hinfo != 0) return zv_fail(c, ZV_ERR_HIP, "davidson: eigensolver info %d", (int)hinfo);
        double rmax = 0.0;
        for (int j = 0; j < k; j++) rmax = std::max(rmax, hnrm[j]);
        if (c->verbose > 1) {
            printf(" davidson iteration %d: space %d, lowest %.6e, max residual %.3e\n", it, m,
                   htheta[0], rmax);
            fflush(stdout);
        }
        if (rmax <= tol) {
            converged = true;
            break;
        }
        if (it >= max_iter) break;
        it++;
        // shift-invert preconditioner: first after 8 plain iterations (the
        // Ritz values have left the bulk of the spectrum), then moved up
        // toward theta_1 each time the residuals have dropped a hundredfold
        if (precond && nfact < 4 && it >= 8 && (!have_pc || rmax < 0.01 * r_at_fact)) {
            const double gap = have_pc ? 2.0 * hnrm[0] + 1e-300
                                       : 2.0 * hnrm[0] + 1e-3 * std::fabs(htheta[0]) + 1e-300;
            bool ok = false;
            double sg = 0.0;
            ZV_CHECK(shift_factor(c, h, V, N, htheta[0], gap, F, info, &sg, &ok));
            nfact++;
            if (ok) {
                have_pc = true;
                sigma = sg;
                r_at_fact = rmax;
            } else if (have_pc) {
                // the factor of the previous shift was overwritten: refactor it
                bool ok2 = false;
                ZV_CHECK(shift_factor(c, h, V, N, sigma, 0.0, F, info, &sg, &ok2));
                have_pc = ok2;
                r_at_fact = rmax;
            }
        }
        // the residuals of the unconverged Ritz vectors extend the space
        int nt = 0;
        for (int j = 0; j < k; j++)
            if (hnrm[j] > tol) {
                ZV_HIPCHECK(c, hipMemcpyAsync(T + (size_t)nt * N, R + (size_t)j * N,
                                              sizeof(double) * N, hipMemcpyDeviceToDevice,
                                              c->stream));
                nt++;
            }
        if (have_pc) // T <- (V - sigma I)^-1 T
            RB(c, rocsolver_dpotrs(h, rocblas_fill_upper, (rocblas_int)N, nt, F, (rocblas_int)N, T,
                                   (rocblas_int)N));
        if (m + nt > mmax) {
            // thick restart: the Ritz vectors span the new space (orthonormal)
            ZV_HIPCHECK(c, hipMemcpyAsync(S, X, sizeof(double) * nk, hipMemcpyDeviceToDevice,
                                          c->stream));
            ZV_HIPCHECK(c, hipMemcpyAsync(AS, AX, sizeof(double) * nk, hipMemcpyDeviceToDevice,
                                          c->stream));
            std::vector<double> hH((size_t)k * mmax, 0.0);
            for (int j = 0; j < k; j++) hH[(size_t)j * mmax + j] = htheta[j];
            ZV_HIPCHECK(c, hipMemcpyAsync(H, hH.data(), sizeof(double) * k * mmax,
                                          hipMemcpyHostToDevice, c->stream));
            m = k;
        }
        // T <- T - S (S^T T), twice, an orthonormal basis of T, once more
        for (int pass = 0; pass < 3; pass++) {
            RB(c, rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, m, nt,
                                (rocblas_int)N, &one, S, (rocblas_int)N, T, (rocblas_int)N, &zero,
                                C, m));
            RB(c, rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)N,
                                nt, m, &mone, S, (rocblas_int)N, C, m, &one, T, (rocblas_int)N));
            if (pass > 0) ZV_CHECK(orthonormalize(c, h, T, N, nt, ow, &n_fallback));
        }
        RB(c, rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)N, nt,
                            (rocblas_int)N, &one, V, (rocblas_int)N, T, (rocblas_int)N, &zero, AT,
                            (rocblas_int)N));
        // new columns of H: S^T AT above, T^T AT on the diagonal block
        RB(c, rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, m, nt,
                            (rocblas_int)N, &one, S, (rocblas_int)N, AT, (rocblas_int)N, &zero,
                            H + (size_t)m * mmax, mmax));
        RB(c, rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, nt, nt,
                            (rocblas_int)N, &one, T, (rocblas_int)N, AT, (rocblas_int)N, &zero,
                            H + (size_t)m * mmax + m, mmax));
        ZV_HIPCHECK(c, hipMemcpyAsync(S + (size_t)m * N, T, sizeof(double) * N * nt,
                                      hipMemcpyDeviceToDevice, c->stream));
        ZV_HIPCHECK(c, hipMemcpyAsync(AS + (size_t)m * N, AT, sizeof(double) * N * nt,
                                      hipMemcpyDeviceToDevice, c->stream));
        m += nt;
    }
    // the converged negative pairs are removed (columns moved to the front)
    int neg = 0;
    std::vector<double> hneg;
    for (int j = 0; j < k; j++) {
        if (!(htheta[j] < 0.0 && hnrm[j] <= tol)) continue;
        if (neg != j)
            ZV_HIPCHECK(c, hipMemcpyAsync(X + (size_t)neg * N, X + (size_t)j * N,
                                          sizeof(double) * N, hipMemcpyDeviceToDevice, c->stream));
        hneg.push_back(htheta[j]);
        neg++;
    }
    if (neg > 0) {
        ZV_HIPCHECK(c, hipMemcpyAsync(E, hneg.data(), sizeof(double) * neg, hipMemcpyHostToDevice,
                                      c->stream));
        ZV_CHECK(remove_pairs(c, h, V, N, X, E, neg, bk));
    }
    if (c->verbose) {
        printf(" davidson: %d lowest eigenvalues of the pair supermatrix, %s after %d "
               "iterations, lowest %g, %d negative removed%s\n",
               k, converged ? "converged" : "not all converged", it, htheta[0], neg,
               have_pc ? ", shift-invert preconditioned" : "");
        fflush(stdout);
    }
    if (evals) std::copy(htheta.begin(), htheta.begin() + k, evals);
    if (resid) std::copy(hnrm.begin(), hnrm.end(), resid);
    if (n_removed) *n_removed = neg;
    if (iters) *iters = it;
    return gather_V(c, V, d_eri, eri, N);
}
