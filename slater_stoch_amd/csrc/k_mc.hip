// ZVMC correction on the GPU: sampler, pair transform + densities, estimator
// and block statistics.
//
// Reference flow (src/two_e_slater.F90:386-496): per block of npts samples,
// draw (r1, r2) in walker blocks of 40 (src/draw_configuration.F90), map them
// per pair P = (i,k) to u_P(r) = f_P |r| r + G_P, evaluate rho_P = phi_i phi_k
// and its Gaussian-fit counterpart at both points (src/compute_densities.F90)
// plus the Jacobians and pi_0 (src/compute_jacobian_and_pi0.F90), then for
// every quartet (P,Q) accumulate J_P(r1) J_Q(r2)/pi_0 rho/|u_P(r1)-u_Q(r2)|.
//
// Layout in HBM: per chunk of Mc samples, every per-(pair, sample) field is a
// plane [Mc][npair_pad] with the pair index fastest, so that a wavefront that
// walks consecutive pairs at one sample reads one contiguous segment.
// Quartet accumulators are kept in pair-triangle order t = P(P+1)/2 + Q,
// P >= Q (P = electron-1 pair), which is the reference's set of canonical
// quartets; results are scattered to the reference kcp order at the end.
#include <cmath>

#include "zv_math.h"
#include "zvmc_internal.h"

namespace {

constexpr double kTwoPi = 2.0 * 3.14159265358979323846;
constexpr double kPi0Factor = 1.0 / ((kTwoPi * kTwoPi) * kTwoPi);
constexpr int kParityFields = 12;
constexpr int kSplitFields = 10;

// ------------------------------------------------------------------ Philox
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
__device__ inline void philox(uint32_t &c0, uint32_t &c1, uint32_t &c2, uint32_t &c3,
                              uint32_t k0, uint32_t k1)
{
#pragma unroll
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

// 53 random bits -> (0,1)
__device__ inline double u53(uint32_t a, uint32_t b)
{
    uint64_t v = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);
    return ((double)v + 0.5) * 0x1.0p-53;
}

// Sample g of stream `stream`: six uniforms from three Philox blocks,
// x = 2 - 2u, r = sqrt(2) erfc^-1(x) (src/draw_configuration.F90:8-25).
__device__ inline void draw_one(uint64_t seed, uint64_t stream, uint64_t g, double r[6])
{
    double u[6];
#pragma unroll
    for (uint32_t t = 0; t < 3; t++) {
        uint32_t c0 = (uint32_t)g, c1 = (uint32_t)(g >> 32), c2 = (uint32_t)stream,
                 c3 = ((uint32_t)(stream >> 32) << 2) | t;
        philox(c0, c1, c2, c3, (uint32_t)seed, (uint32_t)(seed >> 32));
        u[2 * t] = u53(c0, c1);
        u[2 * t + 1] = u53(c2, c3);
    }
    const double s2 = 1.41421356237309504880;
#pragma unroll
    for (int d = 0; d < 6; d++) r[d] = s2 * erfcinv(2.0 - 2.0 * u[d]);
}

__global__ void k_draw(uint64_t seed, uint64_t stream, uint64_t g0, int M, int Mpad, double *r1,
                       double *r2, double *pi0)
{
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= Mpad) return;
    if (m >= M) { // padding samples: finite, never used with weight
        for (int d = 0; d < 3; d++) { r1[3 * m + d] = 0.0; r2[3 * m + d] = 0.0; }
        pi0[m] = 1.0;
        return;
    }
    double r[6];
    draw_one(seed, stream, g0 + (uint64_t)m, r);
    for (int d = 0; d < 3; d++) {
        r1[3 * m + d] = r[d];
        r2[3 * m + d] = r[3 + d];
    }
    double m1 = (r[0] * r[0] + r[1] * r[1]) + r[2] * r[2];
    double m2 = (r[3] * r[3] + r[4] * r[4]) + r[5] * r[5];
    pi0[m] = zv_exp(-0.5 * (m1 + m2)) * kPi0Factor;
}

// ------------------------------------------------------------------ densities
struct Basis {
    const double *center, *zeta, *c1, *c2, *g;
    const int *npower, *nsto, *ngauss;
};

// src/u_orb.F90:15-33 (c_contract = 1)
__device__ inline double u_orb(int nsto, double zeta, double r)
{
    double e = zv_exp(-zeta * r);
    double r2;
    switch (nsto) {
    case 0: return 1.0 * e;
    case 1: return (1.0 * r) * e;
    case 2: r2 = r * r; return (1.0 * r2) * e;
    case 3: r2 = r * r; return ((1.0 * r2) * r) * e;
    default: r2 = r * r; return ((1.0 * r2) * r2) * e;
    }
}

// src/u_gauss.F90:8-12
__device__ inline double u_gauss(const Basis &B, int o, double r)
{
    const double *c1 = B.c1 + (size_t)o * ZV_GMAX, *c2 = B.c2 + (size_t)o * ZV_GMAX,
                 *g = B.g + (size_t)o * ZV_GMAX;
    double r2 = r * r;
    double u = 0.0;
    int ng = B.ngauss[o];
    for (int m = 0; m < ng; m++) u = u + (c1[m] + c2[m] * r2) * zv_exp(-g[m] * r2);
    return u;
}

__device__ inline double poly_at(const Basis &B, int o, const double *u, double *rr)
{
    double dx = u[0] - B.center[3 * o], dy = u[1] - B.center[3 * o + 1],
           dz = u[2] - B.center[3 * o + 2];
    double p = 1.0;
    for (int t = 0; t < B.npower[3 * o]; t++) p = p * dx;
    for (int t = 0; t < B.npower[3 * o + 1]; t++) p = p * dy;
    for (int t = 0; t < B.npower[3 * o + 2]; t++) p = p * dz;
    *rr = sqrt((dx * dx + dy * dy) + dz * dz);
    return p;
}

// rho_ik and rho^G_ik at the point u: src/compute_densities.F90:16-49
__device__ inline void pair_density(const Basis &B, int oi, int ok, const double *u, double &rho,
                                    double &rhog)
{
    double ri, rk;
    double pi_ = poly_at(B, oi, u, &ri);
    double pk_ = poly_at(B, ok, u, &rk);
    double pik = pi_ * pk_;
    rho = (pik * u_orb(B.nsto[oi], B.zeta[oi], ri)) * u_orb(B.nsto[ok], B.zeta[ok], rk);
    rhog = (pik * u_gauss(B, oi, ri)) * u_gauss(B, ok, rk);
}

struct PairSample {
    double u1[3], u2[3], rho1, rho2, rhog1, rhog2, jac1, jac2;
};

// Everything the estimator needs for pair (oi,ok) and sample (a, b):
// src/two_e_slater.F90:406-414 and src/compute_jacobian_and_pi0.F90:22-29
__device__ inline void pair_sample(const Basis &B, int oi, int ok, double f, double jf,
                                   const double G[3], const double *a, const double *b,
                                   PairSample &s)
{
    double m1 = (a[0] * a[0] + a[1] * a[1]) + a[2] * a[2];
    double m2 = (b[0] * b[0] + b[1] * b[1]) + b[2] * b[2];
    double r1m = sqrt(m1), r2m = sqrt(m2);
    for (int d = 0; d < 3; d++) {
        s.u1[d] = (f * a[d]) * r1m + G[d];
        s.u2[d] = (f * b[d]) * r2m + G[d];
    }
    pair_density(B, oi, ok, s.u1, s.rho1, s.rhog1);
    pair_density(B, oi, ok, s.u2, s.rho2, s.rhog2);
    double f1 = jf * r1m, f2 = jf * r2m;
    s.jac1 = 0.25 * fabs((f1 * f1) * f1);
    s.jac2 = 0.25 * fabs((f2 * f2) * f2);
}

struct PairTab {
    const int *pi, *pk;
    const double *f, *jf, *G; // G: [3][npp]
};

// One thread per (pair, sample); FP64 fields for the parity estimator:
// u1x u1y u1z u2x u2y u2z rho1 rho2 rhog1 rhog2 jac1 jac2, each [Mc][npp].
__global__ void k_dens_parity(Basis B, PairTab T, int np, int npp, int Mv, int Mc,
                              const double *__restrict__ r1, const double *__restrict__ r2,
                              double *__restrict__ out)
{
    int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (int64_t)Mv * npp) return;
    int p = (int)(gid % npp), m = (int)(gid / npp);
    if (p >= np) return;
    double G[3] = {T.G[p], T.G[npp + p], T.G[2 * npp + p]};
    PairSample s;
    pair_sample(B, T.pi[p] - 1, T.pk[p] - 1, T.f[p], T.jf[p], G, r1 + 3 * m, r2 + 3 * m, s);
    size_t plane = (size_t)Mc * npp, at = (size_t)m * npp + p;
    out[0 * plane + at] = s.u1[0];
    out[1 * plane + at] = s.u1[1];
    out[2 * plane + at] = s.u1[2];
    out[3 * plane + at] = s.u2[0];
    out[4 * plane + at] = s.u2[1];
    out[5 * plane + at] = s.u2[2];
    out[6 * plane + at] = s.rho1;
    out[7 * plane + at] = s.rho2;
    out[8 * plane + at] = s.rhog1;
    out[9 * plane + at] = s.rhog2;
    out[10 * plane + at] = s.jac1;
    out[11 * plane + at] = s.jac2;
}

// Split form for the tiled estimator: S - G = K (Delta_1 rho_2 + rho^G_1 Delta_2)
// with K = J_1 J_2 / (pi_0 r12) and Delta = rho - rho^G.  Fields [Mc][npp]:
//   electron 1: u1x u1y u1z A1 = (J1/pi0) Delta1, B1 = (J1/pi0) rhoG1
//   electron 2: u2x u2y u2z A2 = J2 rho2,        B2 = J2 Delta2
// Padding pairs/samples get zero weights and points 2000 bohr apart.
template <typename T>
__global__ void k_dens_split(Basis B, PairTab Tb, int np, int npp, int Mv, int Mc,
                             const double *__restrict__ r1, const double *__restrict__ r2,
                             const double *__restrict__ pi0, T *__restrict__ out)
{
    int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (int64_t)Mc * npp) return;
    int p = (int)(gid % npp), m = (int)(gid / npp);
    size_t plane = (size_t)Mc * npp, at = (size_t)m * npp + p;
    T v[kSplitFields];
    if (p >= np || m >= Mv) {
        v[0] = (T)1000.0; v[1] = v[2] = (T)0.0; v[3] = v[4] = (T)0.0;
        v[5] = (T)-1000.0; v[6] = v[7] = (T)0.0; v[8] = v[9] = (T)0.0;
    } else {
        double G[3] = {Tb.G[p], Tb.G[npp + p], Tb.G[2 * npp + p]};
        PairSample s;
        pair_sample(B, Tb.pi[p] - 1, Tb.pk[p] - 1, Tb.f[p], Tb.jf[p], G, r1 + 3 * m,
                    r2 + 3 * m, s);
        double c1 = s.jac1 / pi0[m];
        v[0] = (T)s.u1[0]; v[1] = (T)s.u1[1]; v[2] = (T)s.u1[2];
        v[3] = (T)(c1 * (s.rho1 - s.rhog1));
        v[4] = (T)(c1 * s.rhog1);
        v[5] = (T)s.u2[0]; v[6] = (T)s.u2[1]; v[7] = (T)s.u2[2];
        v[8] = (T)(s.jac2 * s.rho2);
        v[9] = (T)(s.jac2 * (s.rho2 - s.rhog2));
    }
#pragma unroll
    for (int f = 0; f < kSplitFields; f++) out[f * plane + at] = v[f];
}

// ------------------------------------------------------------------ estimator
__device__ inline void tri_decode(int64_t t, int64_t &P, int64_t &Q)
{
    int64_t p = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while (p * (p + 1) / 2 > t) p--;
    while ((p + 1) * (p + 2) / 2 <= t) p++;
    P = p;
    Q = t - p * (p + 1) / 2;
}

// Reference estimator, one thread per quartet (src/two_e_slater.F90:435-450):
// FP32 casts of J/pi0 and of 1/sqrt(r12^2), sequential 40-sample sums.
__global__ void k_est_parity(const double *__restrict__ pd, const double *__restrict__ pi0,
                             int Mv, int Mc, int npp, int64_t ntri, double *__restrict__ eS,
                             double *__restrict__ eG)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntri) return;
    int64_t P, Q;
    tri_decode(t, P, Q);
    size_t plane = (size_t)Mc * npp;
    const double *u1x = pd, *u1y = pd + plane, *u1z = pd + 2 * plane;
    const double *u2x = pd + 3 * plane, *u2y = pd + 4 * plane, *u2z = pd + 5 * plane;
    const double *rh1 = pd + 6 * plane, *rh2 = pd + 7 * plane;
    const double *rg1 = pd + 8 * plane, *rg2 = pd + 9 * plane;
    const double *jc1 = pd + 10 * plane, *jc2 = pd + 11 * plane;
    double s_S = eS[t], s_G = eG[t];
    for (int w = 0; w < Mv; w += ZV_NW) {
        double sS = 0.0, sG = 0.0;
        for (int kw = 0; kw < ZV_NW; kw++) {
            size_t a = (size_t)(w + kw) * npp + P, b = (size_t)(w + kw) * npp + Q;
            double dx = u1x[a] - u2x[b];
            double dy = u1y[a] - u2y[b];
            double dz = u1z[a] - u2z[b];
            double r12 = (dx * dx + dy * dy) + dz * dz;
            double jac = jc1[a] * jc2[b];
            double fac = (double)((float)jac / (float)pi0[w + kw]);
            double wS = (fac * rh1[a]) * rh2[b];
            double wG = (fac * rg1[a]) * rg2[b];
            double rinv = (double)(1.0f / sqrtf((float)r12));
            sS = sS + wS * rinv;
            sG = sG + wG * rinv;
        }
        s_S = s_S + sS;
        s_G = s_G + sG;
    }
    eS[t] = s_S;
    eG[t] = s_G;
}

template <typename T> struct TileCfg;
template <> struct TileCfg<float> { static constexpr int SUB = 16; };
template <> struct TileCfg<double> { static constexpr int SUB = 8; };

// FP32: the hardware reciprocal square root (about 1 ulp).  FP64: the
// hardware estimate (about 5e-8 relative) and one Newton step,
// y = y0 + y0 (1 - x y0^2) / 2, about 5e-15 relative (measured over 1e-9..1e9
// on MI355X); r^2 here is a squared distance of two distinct sample points,
// never 0, infinity or NaN, so the library's special-case handling is not
// needed.
__device__ inline float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ inline double rsq(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    const double e = fma(-(x * y), y, 1.0);
    return fma(y * 0.5, e, y);
}

// Tiled estimator: a workgroup owns a 64x64 tile of (P, Q) pairs with
// P-tile >= Q-tile; each thread a 4x4 register tile.  Samples stream through
// LDS in stages of SUB; per quartet-sample: 3 sub, r^2 (mul + 2 fma), one
// rsqrt, t = A1 A2 + B1 B2 (mul + fma), acc += t rinv (fma).  Stage sums in
// T, flushed to FP64 accumulators after each stage.
template <typename T>
__global__ __launch_bounds__(256) void k_est_tile(const T *__restrict__ e, int Mround, int Mc,
                                                  int npp, int np, double *__restrict__ acc)
{
    constexpr int SUB = TileCfg<T>::SUB;
    __shared__ T sP[5][SUB][ZV_TILE];
    __shared__ T sQ[5][SUB][ZV_TILE];
    int64_t tt = blockIdx.x;
    int64_t bp64, bq64;
    tri_decode(tt, bp64, bq64);
    int bp = (int)bp64, bq = (int)bq64;
    int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    size_t plane = (size_t)Mc * npp;
    double acc64[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc64[i][j] = 0.0;

    for (int m0 = 0; m0 < Mround; m0 += SUB) {
        for (int idx = threadIdx.x; idx < 5 * SUB * ZV_TILE; idx += 256) {
            int f = idx / (SUB * ZV_TILE), rem = idx % (SUB * ZV_TILE);
            int m = rem / ZV_TILE, q = rem % ZV_TILE;
            size_t row = (size_t)(m0 + m) * npp;
            sP[f][m][q] = e[f * plane + row + (size_t)bp * ZV_TILE + q];
            sQ[f][m][q] = e[(5 + f) * plane + row + (size_t)bq * ZV_TILE + q];
        }
        __syncthreads();
        T a[4][4];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) a[i][j] = (T)0;
        for (int m = 0; m < SUB; m++) {
            T px[4], py[4], pz[4], pa[4], pb[4], qx[4], qy[4], qz[4], qa[4], qb[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                px[i] = sP[0][m][ty * 4 + i]; py[i] = sP[1][m][ty * 4 + i];
                pz[i] = sP[2][m][ty * 4 + i]; pa[i] = sP[3][m][ty * 4 + i];
                pb[i] = sP[4][m][ty * 4 + i];
                qx[i] = sQ[0][m][tx * 4 + i]; qy[i] = sQ[1][m][tx * 4 + i];
                qz[i] = sQ[2][m][tx * 4 + i]; qa[i] = sQ[3][m][tx * 4 + i];
                qb[i] = sQ[4][m][tx * 4 + i];
            }
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    T dx = px[i] - qx[j], dy = py[i] - qy[j], dz = pz[i] - qz[j];
                    T r2 = fma(dz, dz, fma(dy, dy, dx * dx));
                    T w = fma(pb[i], qb[j], pa[i] * qa[j]);
                    a[i][j] = fma(w, rsq(r2), a[i][j]);
                }
        }
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) acc64[i][j] += (double)a[i][j];
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        int P = bp * ZV_TILE + ty * 4 + i;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            int Q = bq * ZV_TILE + tx * 4 + j;
            if (P < np && Q <= P) {
                size_t t = (size_t)zv_tri(P, Q);
                acc[t] = acc[t] + acc64[i][j];
            }
        }
    }
}

// ------------------------------------------------------------------ statistics
// End of a block (src/two_e_slater.F90:463-481): x = e_S/npts - e_G/npts,
// sums of x and x^2, accumulators cleared.
__global__ void k_block_end(int64_t ntri, double f, double *accS, double *accG, double *s1,
                            double *s2, double *blk)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntri) return;
    double x = accG ? accS[t] * f - accG[t] * f : accS[t] * f;
    s1[t] = s1[t] + x;
    s2[t] = s2[t] + x * x;
    accS[t] = 0.0;
    if (accG) accG[t] = 0.0;
    if (blk) blk[t] = x;
}

// src/two_e_slater.F90:488-496
__global__ void k_final_stats(int64_t ntri, int nbl, const double *s1, const double *s2,
                              double *mean, double *err)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntri) return;
    double m1 = s1[t] / nbl, m2 = s2[t] / nbl;
    mean[t] = m1;
    err[t] = sqrt(fabs(m2 - m1 * m1)) / sqrt((double)nbl);
}

__global__ void k_fill(double *p, int64_t count, double v)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < count) p[t] = v;
}

// triangle order -> reference kcp order
__global__ void k_scatter_ref(int64_t ntri, const double *__restrict__ in, double *__restrict__ out,
                              const int *pi, const int *pk, const int64_t *kbase, int n)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntri) return;
    int64_t P, Q;
    tri_decode(t, P, Q);
    int i = pi[P], k = pk[P], j = pi[Q], l = pk[Q];
    out[zv_kcp(kbase, n, i, j, k, l)] = in[t];
}

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

} // namespace

// ====================================================================== host
static Basis dev_basis(zv_context *c)
{
    return Basis{c->d_center, c->d_zeta, c->d_c1, c->d_c2, c->d_g,
                 c->d_npower, c->d_nsto, c->d_ngauss};
}
static PairTab dev_pairs(zv_context *c)
{
    return PairTab{c->d_pair_i, c->d_pair_k, c->d_pair_f, c->d_pair_jf, c->d_pair_G};
}

int zv_fill_dev(zv_context *c, double *p, int64_t count, double v)
{
    if (count <= 0) return ZV_OK;
    hipLaunchKernelGGL(k_fill, dim3(nblk(count, 256)), dim3(256), 0, c->stream, p, count, v);
    ZV_HIPCHECK(c, hipGetLastError());
    return ZV_OK;
}

int zv_run_scatter_ref(zv_context *c, const double *d_tri, double *host_out)
{
    int64_t ntri = c->nint;
    ZV_CHECK(zv_ensure(c, c->b_out, sizeof(double) * ntri));
    {
        PhaseTimer tm(c, PH_COPY);
        hipLaunchKernelGGL(k_scatter_ref, dim3(nblk(ntri, 256)), dim3(256), 0, c->stream, ntri,
                           d_tri, (double *)c->b_out.p, c->d_pair_i, c->d_pair_k, c->d_kbase,
                           c->n);
        ZV_HIPCHECK(c, hipGetLastError());
        ZV_HIPCHECK(c, hipMemcpyAsync(host_out, c->b_out.p, sizeof(double) * ntri,
                                      hipMemcpyDeviceToHost, c->stream));
    }
    ZV_HIPCHECK(c, hipStreamSynchronize(c->stream));
    return ZV_OK;
}

int zv_run_draw(zv_context *c, uint64_t seed, uint64_t stream, int64_t m0, int64_t m, double *r1,
                double *r2, double *pi0)
{
    const int64_t chunk = 1 << 20;
    ZV_CHECK(zv_ensure(c, c->b_r1, sizeof(double) * 3 * chunk));
    ZV_CHECK(zv_ensure(c, c->b_r2, sizeof(double) * 3 * chunk));
    ZV_CHECK(zv_ensure(c, c->b_pi0, sizeof(double) * chunk));
    for (int64_t s = 0; s < m; s += chunk) {
        int M = (int)((m - s) < chunk ? (m - s) : chunk);
        {
            PhaseTimer tm(c, PH_SAMPLE);
            hipLaunchKernelGGL(k_draw, dim3(nblk(M, 256)), dim3(256), 0, c->stream, seed, stream,
                               (uint64_t)(m0 + s), M, M, (double *)c->b_r1.p, (double *)c->b_r2.p,
                               (double *)c->b_pi0.p);
            ZV_HIPCHECK(c, hipGetLastError());
        }
        ZV_HIPCHECK(c, hipMemcpyAsync(r1 + 3 * s, c->b_r1.p, sizeof(double) * 3 * M,
                                      hipMemcpyDeviceToHost, c->stream));
        ZV_HIPCHECK(c, hipMemcpyAsync(r2 + 3 * s, c->b_r2.p, sizeof(double) * 3 * M,
                                      hipMemcpyDeviceToHost, c->stream));
        ZV_HIPCHECK(c, hipMemcpyAsync(pi0 + s, c->b_pi0.p, sizeof(double) * M,
                                      hipMemcpyDeviceToHost, c->stream));
        ZV_HIPCHECK(c, hipStreamSynchronize(c->stream));
    }
    return ZV_OK;
}

int zv_run_pair_samples(zv_context *c, int i, int k, int ns, const double *r1, const double *r2,
                        double *ut1, double *ut2, double *rho1, double *rho2, double *rhog1,
                        double *rhog2, double *jac1, double *jac2)
{
    // evaluate through the parity density kernel on a one-pair view
    int npp = c->npair_pad;
    int p = zv_pidx(i, k);
    ZV_CHECK(zv_ensure(c, c->b_r1, sizeof(double) * 3 * ns));
    ZV_CHECK(zv_ensure(c, c->b_r2, sizeof(double) * 3 * ns));
    ZV_CHECK(zv_ensure(c, c->b_pd, sizeof(double) * kParityFields * (size_t)ns * npp));
    ZV_HIPCHECK(c, hipMemcpyAsync(c->b_r1.p, r1, sizeof(double) * 3 * ns, hipMemcpyHostToDevice,
                                  c->stream));
    ZV_HIPCHECK(c, hipMemcpyAsync(c->b_r2.p, r2, sizeof(double) * 3 * ns, hipMemcpyHostToDevice,
                                  c->stream));
    {
        PhaseTimer tm(c, PH_DENS);
        hipLaunchKernelGGL(k_dens_parity, dim3(nblk((int64_t)ns * npp, 256)), dim3(256), 0,
                           c->stream, dev_basis(c), dev_pairs(c), c->npair, npp, ns, ns,
                           (const double *)c->b_r1.p, (const double *)c->b_r2.p,
                           (double *)c->b_pd.p);
        ZV_HIPCHECK(c, hipGetLastError());
    }
    std::vector<double> h((size_t)kParityFields * ns);
    size_t plane = (size_t)ns * npp;
    for (int f = 0; f < kParityFields; f++)
        ZV_HIPCHECK(c, hipMemcpy2DAsync(h.data() + (size_t)f * ns, sizeof(double),
                                        (double *)c->b_pd.p + f * plane + p,
                                        sizeof(double) * npp, sizeof(double), ns,
                                        hipMemcpyDeviceToHost, c->stream));
    ZV_HIPCHECK(c, hipStreamSynchronize(c->stream));
    for (int m = 0; m < ns; m++) {
        for (int d = 0; d < 3; d++) {
            ut1[3 * m + d] = h[(size_t)d * ns + m];
            ut2[3 * m + d] = h[(size_t)(3 + d) * ns + m];
        }
        rho1[m] = h[(size_t)6 * ns + m];
        rho2[m] = h[(size_t)7 * ns + m];
        rhog1[m] = h[(size_t)8 * ns + m];
        rhog2[m] = h[(size_t)9 * ns + m];
        jac1[m] = h[(size_t)10 * ns + m];
        jac2[m] = h[(size_t)11 * ns + m];
    }
    return ZV_OK;
}

// Chunk size (samples) for a per-chunk buffer budget.
static int pick_chunk(zv_context *c, int64_t npts, int precision)
{
    size_t per = (precision == ZV_PREC_FP64_PARITY) ? kParityFields * 8
                 : (precision == ZV_PREC_FP64)      ? kSplitFields * 8
                                                    : kSplitFields * 4;
    size_t budget = (size_t)2 << 30; // 2 GiB of per-(pair, sample) data
    int64_t mc = (int64_t)(budget / (per * (size_t)c->npair_pad));
    mc = mc / 80 * 80;
    if (mc < 80) mc = 80;
    if (mc > 65520) mc = 65520;
    int64_t need = (npts + 79) / 80 * 80;
    if (mc > need) mc = need;
    return (int)mc;
}

int zv_run_mc(zv_context *c, int nbl, int64_t npts, int mode, uint64_t seed, uint64_t stream,
              int precision, const double *r1, const double *r2, const double *pi0, double *mean,
              double *err, double *block_means)
{
    const int64_t ntri = c->nint;
    const int np = c->npair, npp = c->npair_pad;
    const int Mc = pick_chunk(c, npts, precision);
    size_t fbytes = precision == ZV_PREC_FP32 ? 4 : 8;
    size_t nfield = precision == ZV_PREC_FP64_PARITY ? kParityFields : kSplitFields;
    ZV_CHECK(zv_ensure(c, c->b_r1, sizeof(double) * 3 * Mc));
    ZV_CHECK(zv_ensure(c, c->b_r2, sizeof(double) * 3 * Mc));
    ZV_CHECK(zv_ensure(c, c->b_pi0, sizeof(double) * Mc));
    ZV_CHECK(zv_ensure(c, c->b_pd, fbytes * nfield * (size_t)Mc * npp));
    ZV_CHECK(zv_ensure(c, c->b_accS, sizeof(double) * ntri));
    if (precision == ZV_PREC_FP64_PARITY) ZV_CHECK(zv_ensure(c, c->b_accG, sizeof(double) * ntri));
    ZV_CHECK(zv_ensure(c, c->b_s1, sizeof(double) * ntri));
    ZV_CHECK(zv_ensure(c, c->b_s2, sizeof(double) * ntri));
    if (block_means) ZV_CHECK(zv_ensure(c, c->b_blk, sizeof(double) * ntri));
    double *accS = (double *)c->b_accS.p;
    double *accG = precision == ZV_PREC_FP64_PARITY ? (double *)c->b_accG.p : nullptr;
    double *s1 = (double *)c->b_s1.p, *s2 = (double *)c->b_s2.p;
    {
        PhaseTimer tm(c, PH_STATS);
        ZV_CHECK(zv_fill_dev(c, accS, ntri, 0.0));
        if (accG) ZV_CHECK(zv_fill_dev(c, accG, ntri, 0.0));
        ZV_CHECK(zv_fill_dev(c, s1, ntri, 0.0));
        ZV_CHECK(zv_fill_dev(c, s2, ntri, 0.0));
    }
    Basis B = dev_basis(c);
    PairTab PT = dev_pairs(c);
    double *d_r1 = (double *)c->b_r1.p, *d_r2 = (double *)c->b_r2.p, *d_pi0 = (double *)c->b_pi0.p;
    for (int b = 0; b < nbl; b++) {
        for (int64_t m0 = 0; m0 < npts; m0 += Mc) {
            int Mv = (int)((npts - m0) < Mc ? (npts - m0) : Mc);
            uint64_t g0 = (uint64_t)b * (uint64_t)npts + (uint64_t)m0;
            int sub = precision == ZV_PREC_FP32 ? TileCfg<float>::SUB : TileCfg<double>::SUB;
            int Mround = (Mv + sub - 1) / sub * sub;
            if (mode == ZV_SAMPLES_FIXED) {
                PhaseTimer tm(c, PH_COPY);
                ZV_HIPCHECK(c, hipMemcpyAsync(d_r1, r1 + 3 * g0, sizeof(double) * 3 * Mv,
                                              hipMemcpyHostToDevice, c->stream));
                ZV_HIPCHECK(c, hipMemcpyAsync(d_r2, r2 + 3 * g0, sizeof(double) * 3 * Mv,
                                              hipMemcpyHostToDevice, c->stream));
                ZV_HIPCHECK(c, hipMemcpyAsync(d_pi0, pi0 + g0, sizeof(double) * Mv,
                                              hipMemcpyHostToDevice, c->stream));
            } else {
                PhaseTimer tm(c, PH_SAMPLE);
                hipLaunchKernelGGL(k_draw, dim3(nblk(Mv, 256)), dim3(256), 0, c->stream, seed,
                                   stream, g0, Mv, Mv, d_r1, d_r2, d_pi0);
                ZV_HIPCHECK(c, hipGetLastError());
            }
            if (precision == ZV_PREC_FP64_PARITY) {
                {
                    PhaseTimer tm(c, PH_DENS);
                    hipLaunchKernelGGL(k_dens_parity, dim3(nblk((int64_t)Mv * npp, 256)),
                                       dim3(256), 0, c->stream, B, PT, np, npp, Mv, Mc, d_r1,
                                       d_r2, (double *)c->b_pd.p);
                    ZV_HIPCHECK(c, hipGetLastError());
                }
                PhaseTimer tm(c, PH_EST);
                hipLaunchKernelGGL(k_est_parity, dim3(nblk(ntri, 256)), dim3(256), 0, c->stream,
                                   (const double *)c->b_pd.p, d_pi0, Mv, Mc, npp, ntri, accS,
                                   accG);
                ZV_HIPCHECK(c, hipGetLastError());
            } else {
                int nbt = npp / ZV_TILE;
                unsigned ntiles = (unsigned)((int64_t)nbt * (nbt + 1) / 2);
                if (precision == ZV_PREC_FP32) {
                    {
                        PhaseTimer tm(c, PH_DENS);
                        hipLaunchKernelGGL(k_dens_split<float>, dim3(nblk((int64_t)Mc * npp, 256)),
                                           dim3(256), 0, c->stream, B, PT, np, npp, Mv, Mc, d_r1,
                                           d_r2, d_pi0, (float *)c->b_pd.p);
                        ZV_HIPCHECK(c, hipGetLastError());
                    }
                    PhaseTimer tm(c, PH_EST);
                    hipLaunchKernelGGL(k_est_tile<float>, dim3(ntiles), dim3(256), 0, c->stream,
                                       (const float *)c->b_pd.p, Mround, Mc, npp, np, accS);
                    ZV_HIPCHECK(c, hipGetLastError());
                } else {
                    {
                        PhaseTimer tm(c, PH_DENS);
                        hipLaunchKernelGGL(k_dens_split<double>,
                                           dim3(nblk((int64_t)Mc * npp, 256)), dim3(256), 0,
                                           c->stream, B, PT, np, npp, Mv, Mc, d_r1, d_r2, d_pi0,
                                           (double *)c->b_pd.p);
                        ZV_HIPCHECK(c, hipGetLastError());
                    }
                    PhaseTimer tm(c, PH_EST);
                    hipLaunchKernelGGL(k_est_tile<double>, dim3(ntiles), dim3(256), 0, c->stream,
                                       (const double *)c->b_pd.p, Mround, Mc, npp, np, accS);
                    ZV_HIPCHECK(c, hipGetLastError());
                }
            }
        }
        {
            PhaseTimer tm(c, PH_STATS);
            hipLaunchKernelGGL(k_block_end, dim3(nblk(ntri, 256)), dim3(256), 0, c->stream, ntri,
                               1.0 / (double)npts, accS, accG, s1, s2,
                               block_means ? (double *)c->b_blk.p : nullptr);
            ZV_HIPCHECK(c, hipGetLastError());
        }
        if (block_means)
            ZV_CHECK(zv_run_scatter_ref(c, (const double *)c->b_blk.p,
                                        block_means + (size_t)b * ntri));
    }
    ZV_CHECK(zv_ensure(c, c->b_out2, sizeof(double) * ntri));
    {
        PhaseTimer tm(c, PH_STATS);
        // mean -> accS, err -> out2 (triangle order), then scatter
        hipLaunchKernelGGL(k_final_stats, dim3(nblk(ntri, 256)), dim3(256), 0, c->stream, ntri,
                           nbl, s1, s2, accS, (double *)c->b_out2.p);
        ZV_HIPCHECK(c, hipGetLastError());
    }
    ZV_CHECK(zv_run_scatter_ref(c, accS, mean));
    ZV_CHECK(zv_run_scatter_ref(c, (const double *)c->b_out2.p, err));
    return ZV_OK;
}
