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
#include <algorithm>
#include <array>
#include <cmath>

#include "zv_math.h"
#include "zvmc_internal.h"

namespace {

constexpr double kTwoPi = 2.0 * 3.14159265358979323846;
constexpr double kPi0Factor = 1.0 / ((kTwoPi * kTwoPi) * kTwoPi);
constexpr int kParityFields = 12;

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

// ------------------------------------------------------------------ group estimator
// All AO pairs of one shell pair (the Cartesian components of two STO
// shells) share f_P, G_P and the Jacobian, so u_P(r) and the 1/r12 kernel
// of a quartet depend on the two shell pairs only (src/two_e_slater.F90:
// 406-411,437-446).  In the split form
//   S - G = sum_m K(m) [A1_P A2_Q + B1_P B2_Q],  K = 1/|u1_P - u2_Q|,
// the weights fold J/pi0 and J into the component fields, so K is needed
// once per (shell pair, shell pair, sample) and each component quartet costs
// a multiply and two FMAs.
//
// Component fields live in shell-pair order (position pos = the shell
// pair's first position + component, pair index comp[pos]), planes
// [4][npos][Mc] with the sample fastest: A1 = (J1/pi0)(rho1 - rhoG1),
// B1 = (J1/pi0) rhoG1, A2 = J2 rho2, B2 = J2 (rho2 - rhoG2); the points per
// shell pair in [6][nsp][Mc]: u1 xyz, u2 xyz.  Samples m >= Mv get zero
// weights and points 2000 bohr apart.
//
// One thread per (shell pair, sample), the lanes of a wave on consecutive
// samples of one shell pair: the transform, the Jacobians and the radial
// parts (STO and 20-Gaussian fit, 84 exponentials) are computed once for
// the shell pair and every component only multiplies its Cartesian
// polynomials in, with the operations of pair_density in the same order.
template <typename T>
__global__ void k_dens_group(Basis B, PairTab Tb, int npp, const int *__restrict__ pos_pair,
                             const int *__restrict__ sp_first, int npos, int nsp, int Mv, int Mc,
                             const double *__restrict__ r1, const double *__restrict__ r2,
                             const double *__restrict__ pi0, T *__restrict__ comp,
                             T *__restrict__ grp, const int *__restrict__ sp_list)
{
    // sp_list (may be null): the shell pairs to evaluate, one per grid row
    const int mblocks = (Mc + blockDim.x - 1) / blockDim.x;
    const int sp = sp_list ? sp_list[blockIdx.x / mblocks] : blockIdx.x / mblocks;
    const int m = (blockIdx.x % mblocks) * blockDim.x + threadIdx.x;
    if (m >= Mc) return;
    const int pos0 = sp_first[sp], nc = sp_first[sp + 1] - pos0;
    const size_t cplane = (size_t)npos * Mc, gplane = (size_t)nsp * Mc;
    T *g = grp + (size_t)sp * Mc + m;
    if (m >= Mv) {
        const T u[6] = {(T)1000.0, (T)0, (T)0, (T)-1000.0, (T)0, (T)0};
#pragma unroll
        for (int f = 0; f < 6; f++) g[f * gplane] = u[f];
        for (int t = 0; t < nc; t++)
#pragma unroll
            for (int f = 0; f < 4; f++) comp[f * cplane + (size_t)(pos0 + t) * Mc + m] = (T)0;
        return;
    }
    // shell-pair quantities from the first component (src/two_e_slater.F90:406-414)
    const int p0 = pos_pair[pos0];
    const int oi0 = Tb.pi[p0] - 1, ok0 = Tb.pk[p0] - 1;
    const double f = Tb.f[p0], jf = Tb.jf[p0];
    const double *a = r1 + 3 * m, *b = r2 + 3 * m;
    const double m1 = (a[0] * a[0] + a[1] * a[1]) + a[2] * a[2];
    const double m2 = (b[0] * b[0] + b[1] * b[1]) + b[2] * b[2];
    const double r1m = sqrt(m1), r2m = sqrt(m2);
    double u1[3], u2[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double G = Tb.G[d * npp + p0];
        u1[d] = (f * a[d]) * r1m + G;
        u2[d] = (f * b[d]) * r2m + G;
    }
    const double f1 = jf * r1m, f2 = jf * r2m;
    const double jac1 = 0.25 * fabs((f1 * f1) * f1), jac2 = 0.25 * fabs((f2 * f2) * f2);
    const double c1 = jac1 / pi0[m];
    // radial parts of the two shells at both points (src/u_orb.F90, src/u_gauss.F90)
    double ri1, rk1, ri2, rk2;
    (void)poly_at(B, oi0, u1, &ri1);
    (void)poly_at(B, ok0, u1, &rk1);
    (void)poly_at(B, oi0, u2, &ri2);
    (void)poly_at(B, ok0, u2, &rk2);
    const double oi1 = u_orb(B.nsto[oi0], B.zeta[oi0], ri1), ok1 = u_orb(B.nsto[ok0], B.zeta[ok0], rk1);
    const double oi2 = u_orb(B.nsto[oi0], B.zeta[oi0], ri2), ok2 = u_orb(B.nsto[ok0], B.zeta[ok0], rk2);
    const double gi1 = u_gauss(B, oi0, ri1), gk1 = u_gauss(B, ok0, rk1);
    const double gi2 = u_gauss(B, oi0, ri2), gk2 = u_gauss(B, ok0, rk2);
#pragma unroll
    for (int d = 0; d < 3; d++) {
        g[d * gplane] = (T)u1[d];
        g[(3 + d) * gplane] = (T)u2[d];
    }
    for (int t = 0; t < nc; t++) {
        const int p = pos_pair[pos0 + t];
        const int oi = Tb.pi[p] - 1, ok = Tb.pk[p] - 1;
        double rr;
        const double pik1 = poly_at(B, oi, u1, &rr) * poly_at(B, ok, u1, &rr);
        const double pik2 = poly_at(B, oi, u2, &rr) * poly_at(B, ok, u2, &rr);
        const double rho1 = (pik1 * oi1) * ok1, rhog1 = (pik1 * gi1) * gk1;
        const double rho2 = (pik2 * oi2) * ok2, rhog2 = (pik2 * gi2) * gk2;
        T *o = comp + (size_t)(pos0 + t) * Mc + m;
        o[0] = (T)(c1 * (rho1 - rhog1));
        o[cplane] = (T)(c1 * rhog1);
        o[2 * cplane] = (T)(jac2 * rho2);
        o[3 * cplane] = (T)(jac2 * (rho2 - rhog2));
    }
}

// samples per LDS stage of the group estimator, measured on MI355X: deeper
// stages cost more LDS than they gain in latency hiding, and a 16-sample FP32
// stage holds more prefetch registers than two waves per SIMD allow (FP32
// estimator 13.5 ms per block at 16, 11.8 ms at 8 on the 110-AO molecule)
// two BF16 values in one register, for the dot2 instruction
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
__device__ inline bf16x2 bf16_pair(float x, float y)
{
    bf16x2 v = {(__bf16)x, (__bf16)y};
    return v;
}

template <typename T> struct GroupSub;
template <> struct GroupSub<float> { static constexpr int SUB = 8; };
template <> struct GroupSub<double> { static constexpr int SUB = 8; };

// A workgroup owns a tile of 16 TR row units x 16 TC column units (units of
// NR resp. NC components of one shell pair); thread (tx, ty) owns TR row
// units and TC column units: TR*TC kernel values and (TR NR) x (TC NC)
// component quartets per sample.  Rows carry electron 1, columns electron
// 2; an element is written only when its row pair P >= its column pair Q,
// the reference's assignment (the electron-1 pair is the larger one), so
// every quartet is written once.  Samples stream through LDS in stages of
// SUB; stage sums in T are flushed to FP64.
template <typename T, int NR, int NC, int TR, int TC, bool SPLIT, bool BF>
__global__ __launch_bounds__(256) void k_est_group(
    const T *__restrict__ comp, const T *__restrict__ grp, const int *__restrict__ urow_pos,
    const int *__restrict__ urow_sp, int nurow, const int *__restrict__ ucol_pos,
    const int *__restrict__ ucol_sp, int nucol, const int2 *__restrict__ tiles,
    const int *__restrict__ pos_pair, int Mround, int mspan, int Mc, int npos, int nsp_pad,
    double *__restrict__ acc, double *__restrict__ mpart, int64_t ntri)
{
    constexpr int SUB = GroupSub<T>::SUB;
    constexpr int RU = 16 * TR, CU = 16 * TC, ROWS = RU * NR, COLS = CU * NC;
    constexpr int ER = TR * NR, EC = TC * NC; // elements per thread
    // Rows of one sample are padded so that the stage store, whose lanes run
    // over SUB samples of one row, spreads over the LDS banks.
    constexpr int PAD = sizeof(T) == 8 ? 8 : 4;
    constexpr int ROWP = ROWS + PAD, COLP = COLS + PAD, RUP = RU + PAD, CUP = CU + PAD;
    __shared__ T sA1[SUB][ROWP], sB1[SUB][ROWP], sA2[SUB][COLP], sB2[SUB][COLP];
    __shared__ T sU1[3][SUB][RUP], sU2[3][SUB][CUP];
    const int2 tl = tiles[blockIdx.x];
    const int ru0 = tl.x * RU, cu0 = tl.y * CU;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const size_t cplane = (size_t)npos * Mc, gplane = (size_t)nsp_pad * Mc;
    double acc64[ER][EC];
#pragma unroll
    for (int i = 0; i < ER; i++)
#pragma unroll
        for (int j = 0; j < EC; j++) acc64[i][j] = 0.0;

    // Load slots of this thread (fixed for every stage): source offset inside
    // a stage of the component / point planes, -1 for padding.  The next
    // stage's values are loaded into registers while the current one is
    // computed, then stored to LDS behind the barrier.
    constexpr int LA = (SUB * ROWS + 255) / 256, LB = (SUB * COLS + 255) / 256;
    constexpr int LU = (3 * SUB * RU + 255) / 256, LV = (3 * SUB * CU + 255) / 256;
    const int tid = threadIdx.x;
    // Slot idx = (row, sample) with the sample fastest, matching the
    // sample-fastest planes: lo = source offset inside a stage (pos * Mc + m,
    // or sp * Mc + m for a point), -1 for padding; li = flat LDS index.
    int sa[LA], sb[LB], su[LU], sv[LV], la[LA], lb[LB], lu[LU], lv[LV];
#pragma unroll
    for (int k = 0; k < LA; k++) {
        const int idx = tid + 256 * k, m = idx % SUB, r = idx / SUB, gu = ru0 + r / NR;
        sa[k] = (idx < SUB * ROWS && gu < nurow) ? (urow_pos[gu] + r % NR) * Mc + m : -1;
        la[k] = idx < SUB * ROWS ? m * ROWP + r : -1;
    }
#pragma unroll
    for (int k = 0; k < LB; k++) {
        const int idx = tid + 256 * k, m = idx % SUB, r = idx / SUB, gu = cu0 + r / NC;
        sb[k] = (idx < SUB * COLS && gu < nucol) ? (ucol_pos[gu] + r % NC) * Mc + m : -1;
        lb[k] = idx < SUB * COLS ? m * COLP + r : -1;
    }
#pragma unroll
    for (int k = 0; k < LU; k++) {
        const int idx = tid + 256 * k, m = idx % SUB, q = idx / SUB, f = q / RU, gu = ru0 + q % RU;
        su[k] = (idx < 3 * SUB * RU && gu < nurow) ? urow_sp[gu] * Mc + m : -1;
        lu[k] = idx < 3 * SUB * RU ? (f * SUB + m) * RUP + q % RU : -1;
    }
#pragma unroll
    for (int k = 0; k < LV; k++) {
        const int idx = tid + 256 * k, m = idx % SUB, q = idx / SUB, f = q / CU, gu = cu0 + q % CU;
        sv[k] = (idx < 3 * SUB * CU && gu < nucol) ? ucol_sp[gu] * Mc + m : -1;
        lv[k] = idx < 3 * SUB * CU ? (f * SUB + m) * CUP + q % CU : -1;
    }
    T va[LA], vb[LA], vc[LB], vd[LB], vu[LU], vv[LV];
    auto gload = [&](int m0) {
        const T *c0 = comp + m0, *g0 = grp + m0;
#pragma unroll
        for (int k = 0; k < LA; k++) {
            va[k] = sa[k] >= 0 ? c0[sa[k]] : (T)0;
            vb[k] = sa[k] >= 0 ? c0[cplane + sa[k]] : (T)0;
        }
#pragma unroll
        for (int k = 0; k < LB; k++) {
            vc[k] = sb[k] >= 0 ? c0[2 * cplane + sb[k]] : (T)0;
            vd[k] = sb[k] >= 0 ? c0[3 * cplane + sb[k]] : (T)0;
        }
#pragma unroll
        for (int k = 0; k < LU; k++) {
            const int f = (tid + 256 * k) / SUB / RU;
            vu[k] = su[k] >= 0 ? g0[f * gplane + su[k]] : (f == 0 ? (T)1000.0 : (T)0);
        }
#pragma unroll
        for (int k = 0; k < LV; k++) {
            const int f = (tid + 256 * k) / SUB / CU;
            vv[k] = sv[k] >= 0 ? g0[(3 + f) * gplane + sv[k]] : (f == 0 ? (T)-1000.0 : (T)0);
        }
    };
    auto lstore = [&]() {
        T *fA1 = &sA1[0][0], *fB1 = &sB1[0][0], *fA2 = &sA2[0][0], *fB2 = &sB2[0][0];
        T *fU1 = &sU1[0][0][0], *fU2 = &sU2[0][0][0];
#pragma unroll
        for (int k = 0; k < LA; k++)
            if (la[k] >= 0) {
                fA1[la[k]] = va[k];
                fB1[la[k]] = vb[k];
            }
#pragma unroll
        for (int k = 0; k < LB; k++)
            if (lb[k] >= 0) {
                fA2[lb[k]] = vc[k];
                fB2[lb[k]] = vd[k];
            }
#pragma unroll
        for (int k = 0; k < LU; k++)
            if (lu[k] >= 0) fU1[lu[k]] = vu[k];
#pragma unroll
        for (int k = 0; k < LV; k++)
            if (lv[k] >= 0) fU2[lv[k]] = vv[k];
    };
    // SPLIT: sample split blockIdx.y covers [ms0, ms1); split 0 adds into
    // acc, split s > 0 into row s - 1 of mpart (folded in at the end of the
    // block).  Without splits the chunk is walked whole (a separate
    // instantiation: the general bounds cost the FP32 kernels 15 %).
    int ms0 = 0, ms1 = Mround;
    if constexpr (SPLIT) {
        ms0 = (int)blockIdx.y * mspan;
        ms1 = ms0 + mspan < Mround ? ms0 + mspan : Mround;
        if (blockIdx.y > 0) acc = mpart + (size_t)(blockIdx.y - 1) * ntri;
    }
    if (ms0 < ms1) gload(ms0);
    for (int m0 = ms0; m0 < ms1; m0 += SUB) {
        lstore();
        __syncthreads();
        if (m0 + SUB < ms1) gload(m0 + SUB);
        // FP32 sums a stage in T and flushes it to FP64; FP64 accumulates
        // directly.
        T a[ER][EC];
#pragma unroll
        for (int i = 0; i < ER; i++)
#pragma unroll
            for (int j = 0; j < EC; j++) a[i][j] = (T)0;
        T u1[TR][3], u2[TC][3], a1[ER], b1[ER], a2[EC], b2[EC];
        auto fetch = [&](int m) {
#pragma unroll
            for (int i = 0; i < TR; i++)
#pragma unroll
                for (int d = 0; d < 3; d++) u1[i][d] = sU1[d][m][ty * TR + i];
#pragma unroll
            for (int j = 0; j < TC; j++)
#pragma unroll
                for (int d = 0; d < 3; d++) u2[j][d] = sU2[d][m][tx * TC + j];
#pragma unroll
            for (int i = 0; i < ER; i++) {
                a1[i] = sA1[m][ty * ER + i];
                b1[i] = sB1[m][ty * ER + i];
            }
#pragma unroll
            for (int j = 0; j < EC; j++) {
                a2[j] = sA2[m][tx * EC + j];
                b2[j] = sB2[m][tx * EC + j];
            }
        };
#pragma unroll
        for (int m = 0; m < SUB; m++) {
            fetch(m);
            T K[TR][TC];
#pragma unroll
            for (int i = 0; i < TR; i++)
#pragma unroll
                for (int j = 0; j < TC; j++) {
                    const T dx = u1[i][0] - u2[j][0], dy = u1[i][1] - u2[j][1],
                            dz = u1[i][2] - u2[j][2];
                    K[i][j] = rsq(fma(dz, dz, fma(dy, dy, dx * dx)));
                }
            // K scales the side with fewer products: the row components
            // (ER x TC per field) or the column components (TR x EC)
            constexpr bool SCALE_COLS = TR * EC < ER * TC;
            if constexpr (BF) {
                // BF16 products: both factors of K A1 A2 + K B1 B2 rounded to
                // BF16 (round to nearest even) and the pair summed by one
                // dot2 into the FP32 stage sum
                static_assert(sizeof(T) == 4, "BF16 products only in the FP32 kernels");
                if constexpr (!SCALE_COLS) {
                    bf16x2 pr[ER][TC], pc[EC];
#pragma unroll
                    for (int i = 0; i < ER; i++)
#pragma unroll
                        for (int j = 0; j < TC; j++)
                            pr[i][j] = bf16_pair(K[i / NR][j] * a1[i], K[i / NR][j] * b1[i]);
#pragma unroll
                    for (int j = 0; j < EC; j++) pc[j] = bf16_pair(a2[j], b2[j]);
#pragma unroll
                    for (int i = 0; i < ER; i++)
#pragma unroll
                        for (int j = 0; j < TC; j++)
#pragma unroll
                            for (int c = 0; c < NC; c++)
                                a[i][j * NC + c] = __builtin_amdgcn_fdot2_f32_bf16(
                                    pr[i][j], pc[j * NC + c], a[i][j * NC + c], false);
                } else {
                    bf16x2 pr[ER], pc[TR][EC];
#pragma unroll
                    for (int i = 0; i < ER; i++) pr[i] = bf16_pair(a1[i], b1[i]);
#pragma unroll
                    for (int u = 0; u < TR; u++)
#pragma unroll
                        for (int j = 0; j < EC; j++)
                            pc[u][j] = bf16_pair(K[u][j / NC] * a2[j], K[u][j / NC] * b2[j]);
#pragma unroll
                    for (int i = 0; i < ER; i++)
#pragma unroll
                        for (int j = 0; j < EC; j++)
                            a[i][j] = __builtin_amdgcn_fdot2_f32_bf16(pr[i], pc[i / NR][j],
                                                                      a[i][j], false);
                }
            } else if constexpr (!SCALE_COLS) {
                T ka[ER][TC], kb[ER][TC];
#pragma unroll
                for (int i = 0; i < ER; i++)
#pragma unroll
                    for (int j = 0; j < TC; j++) {
                        ka[i][j] = K[i / NR][j] * a1[i];
                        kb[i][j] = K[i / NR][j] * b1[i];
                    }
#pragma unroll
                for (int i = 0; i < ER; i++)
#pragma unroll
                    for (int j = 0; j < TC; j++)
#pragma unroll
                        for (int c = 0; c < NC; c++) {
                            if constexpr (sizeof(T) == 8)
                                acc64[i][j * NC + c] =
                                    fma(ka[i][j], a2[j * NC + c],
                                        fma(kb[i][j], b2[j * NC + c], acc64[i][j * NC + c]));
                            else
                                a[i][j * NC + c] =
                                    fma(ka[i][j], a2[j * NC + c],
                                        fma(kb[i][j], b2[j * NC + c], a[i][j * NC + c]));
                        }
            } else {
                T kc[TR][EC], kd[TR][EC];
#pragma unroll
                for (int u = 0; u < TR; u++)
#pragma unroll
                    for (int j = 0; j < EC; j++) {
                        kc[u][j] = K[u][j / NC] * a2[j];
                        kd[u][j] = K[u][j / NC] * b2[j];
                    }
#pragma unroll
                for (int i = 0; i < ER; i++)
#pragma unroll
                    for (int j = 0; j < EC; j++) {
                        if constexpr (sizeof(T) == 8)
                            acc64[i][j] = fma(a1[i], kc[i / NR][j],
                                              fma(b1[i], kd[i / NR][j], acc64[i][j]));
                        else
                            a[i][j] = fma(a1[i], kc[i / NR][j], fma(b1[i], kd[i / NR][j], a[i][j]));
                    }
            }
        }
        if constexpr (sizeof(T) != 8) {
#pragma unroll
            for (int i = 0; i < ER; i++)
#pragma unroll
                for (int j = 0; j < EC; j++) acc64[i][j] += (double)a[i][j];
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < TR; i++) {
        const int gu = ru0 + ty * TR + i;
        if (gu >= nurow) continue;
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const int P = pos_pair[urow_pos[gu] + r];
#pragma unroll
            for (int j = 0; j < TC; j++) {
                const int gv = cu0 + tx * TC + j;
                if (gv >= nucol) continue;
#pragma unroll
                for (int c = 0; c < NC; c++) {
                    const int Q = pos_pair[ucol_pos[gv] + c];
                    if (P >= Q) {
                        const size_t t = (size_t)zv_tri(P, Q);
                        acc[t] = acc[t] + acc64[i * NR + r][j * NC + c];
                    }
                }
            }
        }
    }
}

// ------------------------------------------------------------------ unit-pair estimator
// The estimator for a short list of unit pairs (targeted refinement): one
// workgroup per (row unit, column unit) pair, its threads over the samples of
// the chunk, with the kernel value and products of k_est_group (K in T,
// K A1 A2 + K B1 B2), summed per thread in FP64 and over the workgroup at
// the end; an element is written only when its row pair P >= its column pair
// Q, as in k_est_group.  up[6 u ..]: row position, row shell pair, row
// components, column position, column shell pair, column components.
template <typename T>
__global__ __launch_bounds__(256) void k_est_pairs(const T *__restrict__ comp,
                                                   const T *__restrict__ grp,
                                                   const int *__restrict__ up,
                                                   const int *__restrict__ pos_pair, int Mv,
                                                   int Mc, int npos, int nsp_pad,
                                                   double *__restrict__ acc)
{
    const int *u = up + 6 * blockIdx.x;
    const int rpos = u[0], rsp = u[1], nr = u[2], cpos = u[3], csp = u[4], nc = u[5];
    const size_t cplane = (size_t)npos * Mc, gplane = (size_t)nsp_pad * Mc;
    double s[6][6];
#pragma unroll
    for (int i = 0; i < 6; i++)
#pragma unroll
        for (int j = 0; j < 6; j++) s[i][j] = 0.0;
    for (int m = threadIdx.x; m < Mv; m += blockDim.x) {
        const T dx = grp[(size_t)rsp * Mc + m] - grp[3 * gplane + (size_t)csp * Mc + m];
        const T dy = grp[gplane + (size_t)rsp * Mc + m] - grp[4 * gplane + (size_t)csp * Mc + m];
        const T dz =
            grp[2 * gplane + (size_t)rsp * Mc + m] - grp[5 * gplane + (size_t)csp * Mc + m];
        const T K = rsq(fma(dz, dz, fma(dy, dy, dx * dx)));
        T a2[6], b2[6];
#pragma unroll
        for (int j = 0; j < 6; j++) {
            a2[j] = j < nc ? comp[2 * cplane + (size_t)(cpos + j) * Mc + m] : (T)0;
            b2[j] = j < nc ? comp[3 * cplane + (size_t)(cpos + j) * Mc + m] : (T)0;
        }
#pragma unroll
        for (int i = 0; i < 6; i++) {
            if (i >= nr) break;
            const T ka = K * comp[(size_t)(rpos + i) * Mc + m];
            const T kb = K * comp[cplane + (size_t)(rpos + i) * Mc + m];
#pragma unroll
            for (int j = 0; j < 6; j++) s[i][j] += (double)fma(kb, b2[j], ka * a2[j]);
        }
    }
    __shared__ double red[4][36];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < 6; i++)
#pragma unroll
        for (int j = 0; j < 6; j++) {
            double v = s[i][j];
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
            if (lane == 0) red[w][6 * i + j] = v;
        }
    __syncthreads();
    const int t = threadIdx.x;
    if (t < 36) {
        const int i = t / 6, j = t % 6;
        if (i < nr && j < nc) {
            const int P = pos_pair[rpos + i], Q = pos_pair[cpos + j];
            if (P >= Q) acc[zv_tri(P, Q)] += (red[0][t] + red[1][t]) + (red[2][t] + red[3][t]);
        }
    }
}

// block end for the selected quartets only (their triangle slots tri[x])
__global__ void k_sel_block_end(int64_t nsel, const int64_t *__restrict__ tri, double f,
                                double *accS, double *s1, double *s2)
{
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= nsel) return;
    const double v = accS[tri[x]] * f;
    accS[tri[x]] = 0.0;
    s1[x] += v;
    s2[x] += v * v;
}

// mean and standard error of the selected quartets (k_final_stats) into the
// kept results at their reference positions
__global__ void k_sel_final(int64_t nsel, int nbl, const double *__restrict__ s1,
                            const double *__restrict__ s2, const int64_t *__restrict__ list,
                            double *res_mean, double *res_err)
{
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= nsel) return;
    const double m1 = s1[x] / nbl, m2 = s2[x] / nbl;
    res_mean[list[x]] = m1;
    res_err[list[x]] = sqrt(fabs(m2 - m1 * m1)) / sqrt((double)nbl);
}

__global__ void k_sel_zero(int64_t nsel, const int64_t *__restrict__ tri, double *accS,
                           double *s1, double *s2)
{
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= nsel) return;
    accS[tri[x]] = 0.0;
    s1[x] = 0.0;
    s2[x] = 0.0;
}

// ------------------------------------------------------------------ statistics
// End of a block (src/two_e_slater.F90:463-481): x = e_S/npts - e_G/npts,
// sums of x and x^2, accumulators cleared.
// With sample splits, their nextra partial rows are added to e_S (in split
// order) and cleared.
__global__ void k_block_end(int64_t ntri, double f, double *accS, double *accG, double *s1,
                            double *s2, double *blk, int nextra, double *mpart)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntri) return;
    double eS = accS[t];
    for (int s = 0; s < nextra; s++) {
        eS += mpart[(size_t)s * ntri + t];
        mpart[(size_t)s * ntri + t] = 0.0;
    }
    double x = accG ? eS * f - accG[t] * f : eS * f;
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

int zv_keep_result(zv_context *c, const double *d_tri, int which, double *host_out)
{
    int64_t ntri = c->nint;
    DevBuf &res = c->b_res[which];
    c->res_valid[which] = false;
    ZV_CHECK(zv_ensure(c, res, sizeof(double) * ntri));
    {
        PhaseTimer tm(c, PH_COPY);
        hipLaunchKernelGGL(k_scatter_ref, dim3(nblk(ntri, 256)), dim3(256), 0, c->stream, ntri,
                           d_tri, (double *)res.p, c->d_pair_i, c->d_pair_k, c->d_kbase, c->n);
        ZV_HIPCHECK(c, hipGetLastError());
        if (host_out)
            ZV_HIPCHECK(c, hipMemcpyAsync(host_out, res.p, sizeof(double) * ntri,
                                          hipMemcpyDeviceToHost, c->stream));
    }
    ZV_HIPCHECK(c, hipStreamSynchronize(c->stream));
    c->res_valid[which] = true;
    return ZV_OK;
}

namespace {
// per-workgroup maxima of x[0 .. n), grid-stride
__global__ void k_block_max(int64_t n, const double *x, double *part)
{
    __shared__ double red[256];
    double m = -INFINITY;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * blockDim.x)
        m = fmax(m, x[t]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (unsigned o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
} // namespace

int zv_run_result_max(zv_context *c, int which, double *value)
{
    constexpr int nb = 1024;
    ZV_CHECK(zv_ensure(c, c->b_out2, sizeof(double) * nb));
    double *part = (double *)c->b_out2.p;
    hipLaunchKernelGGL(k_block_max, dim3(nb), dim3(256), 0, c->stream, c->nint,
                       (const double *)c->b_res[which].p, part);
    ZV_HIPCHECK(c, hipGetLastError());
    double h[nb];
    ZV_HIPCHECK(c, hipMemcpyAsync(h, part, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    ZV_HIPCHECK(c, hipStreamSynchronize(c->stream));
    double m = -INFINITY;
    for (int b = 0; b < nb; b++) m = fmax(m, h[b]);
    *value = m;
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
    // bytes per sample of per-(pair, sample) data
    size_t per = (precision == ZV_PREC_FP64_PARITY)
                     ? (size_t)kParityFields * 8 * c->npair_pad
                     : (size_t)(precision == ZV_PREC_FP64 ? 8 : 4) *
                           (4 * (size_t)c->npair + 6 * (size_t)c->nsp);
    size_t budget = (size_t)2 << 30; // 2 GiB of per-(pair, sample) data
    // a multiple of 160: of the 40-walker blocks of the parity estimator and
    // of every LDS stage depth (8, 16, 32), so a rounded-up stage never
    // reads past the chunk
    int64_t mc = (int64_t)(budget / per);
    if (c->tune[ZV_TUNE_MC_CHUNK] > 0) mc = std::min<int64_t>(mc, c->tune[ZV_TUNE_MC_CHUNK]);
    mc = mc / 160 * 160;
    if (mc < 160) mc = 160;
    if (mc > 65440) mc = 65440;
    int64_t need = (npts + 159) / 160 * 160;
    if (mc > need) mc = need;
    return (int)mc;
}

// Sample splits of the group estimator: a molecule with few tiles would
// leave most CUs idle while each tile walks the whole chunk, so the chunk's
// stages are split over up to 64 workgroups per tile, aiming at about 1024
// workgroups in all (about one resident round: 4 per CU by LDS), with at
// least 4 stages per split and at most 1 GiB of partial rows.
static int pick_split(zv_context *c, int Mc, int precision, int64_t tiles)
{
    if (tiles <= 0) return 1;
    const bool single = precision == ZV_PREC_FP32 || precision == ZV_PREC_BF16;
    const int sub = single ? GroupSub<float>::SUB : GroupSub<double>::SUB;
    int64_t n = (1024 + tiles - 1) / tiles;
    // a caller's choice (zv_set_tuning) replaces the count, within the
    // same bounds of stages and partial-row memory
    if (c->tune[ZV_TUNE_EST_SPLIT] > 0) n = c->tune[ZV_TUNE_EST_SPLIT];
    n = std::min<int64_t>(n, 64);
    n = std::min<int64_t>(n, std::max<int64_t>(1, Mc / sub / 4));
    n = std::min<int64_t>(n, 1 + ((int64_t)1 << 27) / std::max<int64_t>(1, c->nint));
    return (int)std::max<int64_t>(n, 1);
}

// ---- group estimator tables -------------------------------------------------
// unit classes: 1, 3 or 6 components of one shell pair
constexpr int kUnitRows[3] = {1, 3, 6};
// units per thread (TR row units, TC column units) for each pair of classes
// Two tile-shape sets (ZV_TUNE_EST_TILES): set 1 halves the units per thread
// along the longer side, four times fewer units per tile for the narrow
// classes: more tiles (and less sample splitting) for small molecules.
constexpr int kUnitsPerThread[2][3][3][2] = {{{{4, 4}, {4, 2}, {4, 1}},
                                             {{2, 4}, {2, 2}, {2, 1}},
                                             {{1, 4}, {1, 2}, {1, 1}}},
                                            {{{2, 2}, {2, 1}, {2, 1}},
                                             {{1, 2}, {1, 1}, {1, 1}},
                                             {{1, 2}, {1, 1}, {1, 1}}}};

// Units (greedy split of each shell pair's components into 6, 3 and 1),
// ordered by the first shell of their shell pair, and the tiles of every
// pair of classes that hold at least one element with P >= Q: a row unit
// of shell pair (I, K) and a column unit of (J, L) can hold P >= Q only
// when I >= J.
static int group_ready(zv_context *c)
{
    const int set = c->tune[ZV_TUNE_EST_TILES] == 1 ? 1 : 0;
    if (c->grp_ready && c->grp_tileset == set) return ZV_OK;
    c->grp_tileset = set;
    const int nsp = c->nsp, npair = c->npair;
    std::vector<int> sp_I(nsp);
    for (int I = 0; I < c->nshell; I++)
        for (int K = 0; K <= I; K++) sp_I[I * (I + 1) / 2 + K] = I;
    std::vector<int> upos[3], usp[3], uI[3];
    for (int sp = 0; sp < nsp; sp++) {
        int at = c->h_sp_coff[sp];
        int rem = c->h_sp_coff[sp + 1] - at;
        for (int k = 2; k >= 0; k--)
            while (rem >= kUnitRows[k]) {
                upos[k].push_back(at);
                usp[k].push_back(sp);
                uI[k].push_back(sp_I[sp]);
                at += kUnitRows[k];
                rem -= kUnitRows[k];
            }
    }
    std::vector<int> &g = c->h_gmeta;
    g.clear();
    g.insert(g.end(), c->h_sp_comp.begin(), c->h_sp_comp.end()); // pos -> pair
    c->grp_possp_off = (int64_t)g.size();
    std::vector<int> pos_sp(npair);
    for (int sp = 0; sp < nsp; sp++)
        for (int t = c->h_sp_coff[sp]; t < c->h_sp_coff[sp + 1]; t++) pos_sp[t] = sp;
    g.insert(g.end(), pos_sp.begin(), pos_sp.end());
    g.insert(g.end(), c->h_sp_coff.begin(), c->h_sp_coff.end()); // sp -> first position
    for (int k = 0; k < 3; k++) {
        c->grp_nunit[k] = (int)upos[k].size();
        c->grp_unit_off[k] = (int64_t)g.size();
        g.insert(g.end(), upos[k].begin(), upos[k].end());
        g.insert(g.end(), usp[k].begin(), usp[k].end());
    }
    c->h_pair_ucls.assign(npair, 0);
    c->h_pair_uidx.assign(npair, 0);
    for (int k = 0; k < 3; k++)
        for (int u = 0; u < (int)upos[k].size(); u++)
            for (int r = 0; r < kUnitRows[k]; r++) {
                const int P = c->h_sp_comp[upos[k][u] + r];
                c->h_pair_ucls[P] = k;
                c->h_pair_uidx[P] = u;
            }
    if (g.size() & 1) g.push_back(0); // int2 alignment of the tiles
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            const int RU = 16 * kUnitsPerThread[set][a][b][0];
            const int CU = 16 * kUnitsPerThread[set][a][b][1];
            const int nr = c->grp_nunit[a], ncl = c->grp_nunit[b];
            const int nrt = (nr + RU - 1) / RU, nct = (ncl + CU - 1) / CU;
            c->grp_tile_off[3 * a + b] = (int64_t)g.size() / 2;
            int cnt = 0;
            for (int rt = 0; rt < nrt; rt++) {
                const int maxI = uI[a][std::min(nr, (rt + 1) * RU) - 1];
                for (int ct = 0; ct < nct; ct++) {
                    const int minI = uI[b][ct * CU];
                    if (maxI < minI) continue;
                    g.push_back(rt);
                    g.push_back(ct);
                    cnt++;
                }
            }
            c->grp_ntile[3 * a + b] = cnt;
        }
    c->grp_nsp_pad = nsp;
    ZV_CHECK(zv_ensure(c, c->b_gmeta, sizeof(int) * g.size()));
    ZV_HIPCHECK(c, hipMemcpyAsync(c->b_gmeta.p, g.data(), sizeof(int) * g.size(),
                                  hipMemcpyHostToDevice, c->stream));
    ZV_HIPCHECK(c, hipStreamSynchronize(c->stream));
    c->grp_ready = true;
    return ZV_OK;
}

template <typename T, bool BF, int NR, int NC, int TR, int TC>
static void launch_group_t(zv_context *c, hipStream_t st, unsigned ntile, const T *comp,
                           const T *grp, const int *urp, const int *urs, int nur, const int *ucp,
                           const int *ucs, int nuc, const int2 *tiles, const int *pos_pair,
                           int Mround, int nsplit, int mspan, int Mc, double *acc,
                           double *mpart)
{
    if (nsplit > 1)
        hipLaunchKernelGGL((k_est_group<T, NR, NC, TR, TC, true, BF>), dim3(ntile, (unsigned)nsplit),
                           dim3(256), 0, st, comp, grp, urp, urs, nur, ucp, ucs, nuc, tiles,
                           pos_pair, Mround, mspan, Mc, c->npair, c->grp_nsp_pad, acc, mpart,
                           c->nint);
    else
        hipLaunchKernelGGL((k_est_group<T, NR, NC, TR, TC, false, BF>), dim3(ntile), dim3(256), 0, st,
                           comp, grp, urp, urs, nur, ucp, ucs, nuc, tiles, pos_pair, Mround,
                           mspan, Mc, c->npair, c->grp_nsp_pad, acc, mpart, c->nint);
}

// One launch per pair of unit classes; each is too small to fill the device
// alone, so they run side by side on the main and the side streams (they
// write disjoint quartets) and are joined back into the main stream.  With
// nsplit > 1 every tile is also split over the samples of the chunk.
template <typename T, bool BF>
static int run_group_estimator(zv_context *c, const T *comp, const T *grp, int Mround, int Mc,
                               int nsplit, double *acc, double *mpart)
{
    constexpr int SUB = GroupSub<T>::SUB;
    const int nstage = Mround / SUB;
    const int per = (nstage + nsplit - 1) / nsplit;
    const int mspan = per * SUB;
    const int *g = (const int *)c->b_gmeta.p;
    const int *pos_pair = g;
    hipStream_t sts[1 + zv_context::NSIDE] = {c->stream};
    for (int k = 0; k < zv_context::NSIDE; k++) sts[1 + k] = c->side[k];
    ZV_HIPCHECK(c, hipEventRecord(c->ev_fork, c->stream));
    for (int k = 0; k < zv_context::NSIDE; k++)
        ZV_HIPCHECK(c, hipStreamWaitEvent(c->side[k], c->ev_fork, 0));
    // largest launches first
    int order[9];
    for (int k = 0; k < 9; k++) order[k] = k;
    std::sort(order, order + 9, [&](int x, int y) { return c->grp_ntile[x] > c->grp_ntile[y]; });
    int nl = 0;
    for (int oi = 0; oi < 9; oi++) {
        const int a = order[oi] / 3, b = order[oi] % 3;
        {
            const unsigned nt = (unsigned)c->grp_ntile[3 * a + b];
            if (!nt) continue;
            hipStream_t st = sts[nl++ % (1 + zv_context::NSIDE)];
            const int *urp = g + c->grp_unit_off[a], *urs = urp + c->grp_nunit[a];
            const int *ucp = g + c->grp_unit_off[b], *ucs = ucp + c->grp_nunit[b];
            const int nur = c->grp_nunit[a], nuc = c->grp_nunit[b];
            const int2 *tiles = (const int2 *)g + c->grp_tile_off[3 * a + b];
#define ZV_GROUP(NR, NC, TR, TC)                                                                  \
    launch_group_t<T, BF, NR, NC, TR, TC>(c, st, nt, comp, grp, urp, urs, nur, ucp, ucs, nuc, tiles, \
                                      pos_pair, Mround, nsplit, mspan, Mc, acc, mpart)
            switch (9 * c->grp_tileset + 3 * a + b) {
            case 0: ZV_GROUP(1, 1, 4, 4); break;
            case 1: ZV_GROUP(1, 3, 4, 2); break;
            case 2: ZV_GROUP(1, 6, 4, 1); break;
            case 3: ZV_GROUP(3, 1, 2, 4); break;
            case 4: ZV_GROUP(3, 3, 2, 2); break;
            case 5: ZV_GROUP(3, 6, 2, 1); break;
            case 6: ZV_GROUP(6, 1, 1, 4); break;
            case 7: ZV_GROUP(6, 3, 1, 2); break;
            case 8: ZV_GROUP(6, 6, 1, 1); break;
            case 9: ZV_GROUP(1, 1, 2, 2); break;
            case 10: ZV_GROUP(1, 3, 2, 1); break;
            case 11: ZV_GROUP(1, 6, 2, 1); break;
            case 12: ZV_GROUP(3, 1, 1, 2); break;
            case 13: ZV_GROUP(3, 3, 1, 1); break;
            case 14: ZV_GROUP(3, 6, 1, 1); break;
            case 15: ZV_GROUP(6, 1, 1, 2); break;
            case 16: ZV_GROUP(6, 3, 1, 1); break;
            case 17: ZV_GROUP(6, 6, 1, 1); break;
            }
#undef ZV_GROUP
            ZV_HIPCHECK(c, hipGetLastError());
        }
    }
    for (int k = 0; k < zv_context::NSIDE; k++) {
        ZV_HIPCHECK(c, hipEventRecord(c->ev_join[k], c->side[k]));
        ZV_HIPCHECK(c, hipStreamWaitEvent(c->stream, c->ev_join[k], 0));
    }
    return ZV_OK;
}

// nbl blocks of npts samples over every quartet; the mean is left in accS
// and the standard error in b_out2, both in triangle order.
static int run_mc_core(zv_context *c, int nbl, int64_t npts, int mode, uint64_t seed,
                       uint64_t stream, int precision, const double *r1, const double *r2,
                       const double *pi0, double *block_means)
{
    const int64_t ntri = c->nint;
    const int np = c->npair, npp = c->npair_pad;
    const int Mc = pick_chunk(c, npts, precision);
    const bool single = precision == ZV_PREC_FP32 || precision == ZV_PREC_BF16;
    size_t fbytes = single ? 4 : 8;
    if (precision != ZV_PREC_FP64_PARITY) ZV_CHECK(group_ready(c));
    const size_t ncomp = (size_t)4 * Mc * np, ngrp = (size_t)6 * Mc * c->nsp;
    ZV_CHECK(zv_ensure(c, c->b_r1, sizeof(double) * 3 * Mc));
    ZV_CHECK(zv_ensure(c, c->b_r2, sizeof(double) * 3 * Mc));
    ZV_CHECK(zv_ensure(c, c->b_pi0, sizeof(double) * Mc));
    ZV_CHECK(zv_ensure(c, c->b_pd, precision == ZV_PREC_FP64_PARITY
                                       ? fbytes * kParityFields * (size_t)Mc * npp
                                       : fbytes * (ncomp + ngrp)));
    ZV_CHECK(zv_ensure(c, c->b_accS, sizeof(double) * ntri));
    if (precision == ZV_PREC_FP64_PARITY) ZV_CHECK(zv_ensure(c, c->b_accG, sizeof(double) * ntri));
    ZV_CHECK(zv_ensure(c, c->b_s1, sizeof(double) * ntri));
    ZV_CHECK(zv_ensure(c, c->b_s2, sizeof(double) * ntri));
    if (block_means) ZV_CHECK(zv_ensure(c, c->b_blk, sizeof(double) * ntri));
    int64_t ntiles = 0;
    for (int k = 0; k < 9; k++) ntiles += c->grp_ntile[k];
    const int nsplit = precision == ZV_PREC_FP64_PARITY ? 1 : pick_split(c, Mc, precision, ntiles);
    if (nsplit > 1) ZV_CHECK(zv_ensure(c, c->b_mpart, sizeof(double) * (nsplit - 1) * ntri));
    double *mpart = nsplit > 1 ? (double *)c->b_mpart.p : nullptr;
    double *accS = (double *)c->b_accS.p;
    double *accG = precision == ZV_PREC_FP64_PARITY ? (double *)c->b_accG.p : nullptr;
    double *s1 = (double *)c->b_s1.p, *s2 = (double *)c->b_s2.p;
    {
        PhaseTimer tm(c, PH_STATS);
        ZV_CHECK(zv_fill_dev(c, accS, ntri, 0.0));
        if (accG) ZV_CHECK(zv_fill_dev(c, accG, ntri, 0.0));
        ZV_CHECK(zv_fill_dev(c, s1, ntri, 0.0));
        ZV_CHECK(zv_fill_dev(c, s2, ntri, 0.0));
        if (mpart) ZV_CHECK(zv_fill_dev(c, mpart, (int64_t)(nsplit - 1) * ntri, 0.0));
    }
    Basis B = dev_basis(c);
    PairTab PT = dev_pairs(c);
    double *d_r1 = (double *)c->b_r1.p, *d_r2 = (double *)c->b_r2.p, *d_pi0 = (double *)c->b_pi0.p;
    for (int b = 0; b < nbl; b++) {
        for (int64_t m0 = 0; m0 < npts; m0 += Mc) {
            int Mv = (int)((npts - m0) < Mc ? (npts - m0) : Mc);
            uint64_t g0 = (uint64_t)b * (uint64_t)npts + (uint64_t)m0;
            int sub = single ? GroupSub<float>::SUB : GroupSub<double>::SUB;
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
                const int *g = (const int *)c->b_gmeta.p;
                const int *pos_pair = g, *sp_first = g + c->grp_possp_off + np;
                const dim3 grid(nblk(Mc, 256) * (unsigned)c->nsp);
                if (single) {
                    float *comp = (float *)c->b_pd.p, *grp = comp + ncomp;
                    {
                        PhaseTimer tm(c, PH_DENS);
                        hipLaunchKernelGGL(k_dens_group<float>, grid, dim3(256), 0, c->stream, B,
                                           PT, npp, pos_pair, sp_first, np, c->nsp, Mv, Mc, d_r1,
                                           d_r2, d_pi0, comp, grp, (const int *)nullptr);
                        ZV_HIPCHECK(c, hipGetLastError());
                    }
                    PhaseTimer tm(c, PH_EST);
                    if (precision == ZV_PREC_BF16)
                        ZV_CHECK((run_group_estimator<float, true>(c, comp, grp, Mround, Mc,
                                                                   nsplit, accS, mpart)));
                    else
                        ZV_CHECK((run_group_estimator<float, false>(c, comp, grp, Mround, Mc,
                                                                    nsplit, accS, mpart)));
                } else {
                    double *comp = (double *)c->b_pd.p, *grp = comp + ncomp;
                    {
                        PhaseTimer tm(c, PH_DENS);
                        hipLaunchKernelGGL(k_dens_group<double>, grid, dim3(256), 0, c->stream, B,
                                           PT, npp, pos_pair, sp_first, np, c->nsp, Mv, Mc, d_r1,
                                           d_r2, d_pi0, comp, grp, (const int *)nullptr);
                        ZV_HIPCHECK(c, hipGetLastError());
                    }
                    PhaseTimer tm(c, PH_EST);
                    ZV_CHECK((run_group_estimator<double, false>(c, comp, grp, Mround, Mc,
                                                                 nsplit, accS, mpart)));
                }
            }
        }
        {
            PhaseTimer tm(c, PH_STATS);
            hipLaunchKernelGGL(k_block_end, dim3(nblk(ntri, 256)), dim3(256), 0, c->stream, ntri,
                               1.0 / (double)npts, accS, accG, s1, s2,
                               block_means ? (double *)c->b_blk.p : nullptr, nsplit - 1, mpart);
            ZV_HIPCHECK(c, hipGetLastError());
        }
        if (block_means)
            ZV_CHECK(zv_run_scatter_ref(c, (const double *)c->b_blk.p,
                                        block_means + (size_t)b * ntri));
        if (c->verbose >= 2) {
            ZV_HIPCHECK(c, hipStreamSynchronize(c->stream));
            printf(" MC block %d / %d done (%lld samples, %lld quartets)\n", b + 1, nbl,
                   (long long)npts, (long long)ntri);
            fflush(stdout);
        }
    }
    ZV_CHECK(zv_ensure(c, c->b_out2, sizeof(double) * ntri));
    {
        PhaseTimer tm(c, PH_STATS);
        // mean -> accS, err -> out2 (triangle order), then scatter
        hipLaunchKernelGGL(k_final_stats, dim3(nblk(ntri, 256)), dim3(256), 0, c->stream, ntri,
                           nbl, s1, s2, accS, (double *)c->b_out2.p);
        ZV_HIPCHECK(c, hipGetLastError());
    }
    return ZV_OK;
}

int zv_run_mc(zv_context *c, int nbl, int64_t npts, int mode, uint64_t seed, uint64_t stream,
              int precision, const double *r1, const double *r2, const double *pi0, double *mean,
              double *err, double *block_means)
{
    ZV_CHECK(run_mc_core(c, nbl, npts, mode, seed, stream, precision, r1, r2, pi0, block_means));
    ZV_CHECK(zv_keep_result(c, (const double *)c->b_accS.p, ZV_RESULT_MEAN, mean));
    ZV_CHECK(zv_keep_result(c, (const double *)c->b_out2.p, ZV_RESULT_STDERR, err));
    return ZV_OK;
}

// ---- targeted refinement ------------------------------------------------------
namespace {
// kcp positions with err > target: count, then (after the list is sized) fill
__global__ void k_select(int64_t n, const double *__restrict__ err, double target,
                         unsigned long long *count, int64_t *list, int64_t cap)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n || !(err[t] > target)) return;
    const unsigned long long at = atomicAdd(count, 1ull);
    if (list && (int64_t)at < cap) list[at] = t;
}

// reference position kcp -> electron-1 pair (is, ks), electron-2 pair (js, ls)
// and the triangle slot; the (l, k, j) run holding kcp is the last base-table
// entry not above it (src/build_mapping_ijkl.F90:16-43 order)
__global__ void k_kcp_pairs(int64_t nsel, const int64_t *__restrict__ list,
                            const int64_t *__restrict__ kbase, int n, int *pr, int *pc,
                            int64_t *tri)
{
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= nsel) return;
    const int64_t kcp = list[x];
    int64_t lo = 0, hi = (int64_t)n * n * n; // first entry above kcp
    while (lo < hi) {
        const int64_t mid = (lo + hi) / 2;
        if (kbase[mid] <= kcp) lo = mid + 1;
        else hi = mid;
    }
    const int64_t e = lo - 1;
    const int j = (int)(e % n) + 1, k = (int)((e / n) % n) + 1, l = (int)(e / ((int64_t)n * n)) + 1;
    const int start = (k > j ? k : j) + ((l > k && j >= k) ? 1 : 0);
    const int i = start + (int)(kcp - kbase[e]);
    const int P = zv_pidx(i, k), Q = zv_pidx(j, l);
    pr[x] = P > Q ? P : Q;
    pc[x] = P > Q ? Q : P;
    tri[x] = P > Q ? zv_tri(P, Q) : zv_tri(Q, P);
}

// kept results of the selected quartets <- this run's (triangle order)
__global__ void k_merge_sel(int64_t nsel, const int64_t *__restrict__ list,
                            const int64_t *__restrict__ tri, const double *__restrict__ mean_t,
                            const double *__restrict__ err_t, double *res_mean, double *res_err)
{
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= nsel) return;
    res_mean[list[x]] = mean_t[tri[x]];
    res_err[list[x]] = err_t[tri[x]];
}
} // namespace

int zv_run_mc_refine(zv_context *c, int nbl, int64_t npts, uint64_t seed, uint64_t stream,
                     int precision, double target, int64_t *n_refined, double *max_err)
{
    if (!c->res_valid[ZV_RESULT_MEAN] || !c->res_valid[ZV_RESULT_STDERR])
        return zv_fail(c, ZV_ERR_STATE,
                       "zv_mc_refine needs the device results of a zv_mc_blocks run");
    ZV_CHECK(group_ready(c));
    const int64_t ntri = c->nint;
    double *res_mean = (double *)c->b_res[ZV_RESULT_MEAN].p;
    double *res_err = (double *)c->b_res[ZV_RESULT_STDERR].p;
    // 1. the quartets above the target
    ZV_CHECK(zv_ensure(c, c->b_misc, 64));
    unsigned long long *d_count = (unsigned long long *)c->b_misc.p;
    ZV_HIPCHECK(c, hipMemsetAsync(d_count, 0, sizeof(unsigned long long), c->stream));
    hipLaunchKernelGGL(k_select, dim3(nblk(ntri, 256)), dim3(256), 0, c->stream, ntri,
                       (const double *)res_err, target, d_count, (int64_t *)nullptr, (int64_t)0);
    ZV_HIPCHECK(c, hipGetLastError());
    unsigned long long hcount = 0;
    ZV_HIPCHECK(c, hipMemcpyAsync(&hcount, d_count, sizeof(hcount), hipMemcpyDeviceToHost,
                                  c->stream));
    ZV_HIPCHECK(c, hipStreamSynchronize(c->stream));
    const int64_t nsel = (int64_t)hcount;
    if (n_refined) *n_refined = nsel;
    if (nsel == 0) return zv_run_result_max(c, ZV_RESULT_STDERR, max_err);
    // list (int64), pairs (2 int), triangle slots (int64) in b_list
    const size_t lbytes = (size_t)nsel * (8 + 8 + 8);
    ZV_CHECK(zv_ensure(c, c->b_list, lbytes));
    int64_t *d_list = (int64_t *)c->b_list.p, *d_tri = d_list + nsel;
    int *d_pr = (int *)(d_tri + nsel), *d_pc = d_pr + nsel;
    ZV_HIPCHECK(c, hipMemsetAsync(d_count, 0, sizeof(unsigned long long), c->stream));
    hipLaunchKernelGGL(k_select, dim3(nblk(ntri, 256)), dim3(256), 0, c->stream, ntri,
                       (const double *)res_err, target, d_count, d_list, nsel);
    hipLaunchKernelGGL(k_kcp_pairs, dim3(nblk(nsel, 256)), dim3(256), 0, c->stream, nsel,
                       (const int64_t *)d_list, (const int64_t *)c->d_kbase, c->n, d_pr, d_pc,
                       d_tri);
    ZV_HIPCHECK(c, hipGetLastError());
    // 2. the unit pairs holding them (row unit of P, column unit of Q) and
    // the shell pairs those units belong to
    std::vector<int> hpr(nsel), hpc(nsel);
    ZV_HIPCHECK(c, hipMemcpyAsync(hpr.data(), d_pr, sizeof(int) * nsel, hipMemcpyDeviceToHost,
                                  c->stream));
    ZV_HIPCHECK(c, hipMemcpyAsync(hpc.data(), d_pc, sizeof(int) * nsel, hipMemcpyDeviceToHost,
                                  c->stream));
    ZV_HIPCHECK(c, hipStreamSynchronize(c->stream));
    std::vector<std::array<int, 4>> pairs(nsel);
    for (int64_t x = 0; x < nsel; x++)
        pairs[x] = {c->h_pair_ucls[hpr[x]], c->h_pair_uidx[hpr[x]], c->h_pair_ucls[hpc[x]],
                    c->h_pair_uidx[hpc[x]]};
    std::sort(pairs.begin(), pairs.end());
    pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
    const int nup = (int)pairs.size();
    const std::vector<int> &g = c->h_gmeta;
    std::vector<int> hup(6 * (size_t)nup), hsp;
    for (int x = 0; x < nup; x++) {
        const int cr = pairs[x][0], ur = pairs[x][1], cc = pairs[x][2], uc = pairs[x][3];
        const int *rtab = g.data() + c->grp_unit_off[cr], *ctab = g.data() + c->grp_unit_off[cc];
        int *o = hup.data() + 6 * (size_t)x;
        o[0] = rtab[ur];
        o[1] = rtab[c->grp_nunit[cr] + ur];
        o[2] = kUnitRows[cr];
        o[3] = ctab[uc];
        o[4] = ctab[c->grp_nunit[cc] + uc];
        o[5] = kUnitRows[cc];
        hsp.push_back(o[1]);
        hsp.push_back(o[4]);
    }
    std::sort(hsp.begin(), hsp.end());
    hsp.erase(std::unique(hsp.begin(), hsp.end()), hsp.end());
    const int nspl = (int)hsp.size();
    ZV_CHECK(zv_ensure(c, c->b_part, sizeof(int) * (hup.size() + hsp.size())));
    int *d_up = (int *)c->b_part.p, *d_spl = d_up + hup.size();
    ZV_HIPCHECK(c, hipMemcpyAsync(d_up, hup.data(), sizeof(int) * hup.size(),
                                  hipMemcpyHostToDevice, c->stream));
    ZV_HIPCHECK(c, hipMemcpyAsync(d_spl, hsp.data(), sizeof(int) * hsp.size(),
                                  hipMemcpyHostToDevice, c->stream));
    // 3. nbl blocks of npts samples (the samples of a full zv_mc_blocks run
    // with this seed and stream): densities of those shell pairs, the
    // unit-pair estimator, statistics of the selected quartets
    const bool single = precision != ZV_PREC_FP64;
    const int Mc = pick_chunk(c, npts, precision);
    const int np = c->npair, npp = c->npair_pad;
    const size_t fbytes = single ? 4 : 8;
    const size_t ncomp = (size_t)4 * Mc * np, ngrp = (size_t)6 * Mc * c->nsp;
    ZV_CHECK(zv_ensure(c, c->b_r1, sizeof(double) * 3 * Mc));
    ZV_CHECK(zv_ensure(c, c->b_r2, sizeof(double) * 3 * Mc));
    ZV_CHECK(zv_ensure(c, c->b_pi0, sizeof(double) * Mc));
    ZV_CHECK(zv_ensure(c, c->b_pd, fbytes * (ncomp + ngrp)));
    ZV_CHECK(zv_ensure(c, c->b_accS, sizeof(double) * ntri));
    ZV_CHECK(zv_ensure(c, c->b_s1, sizeof(double) * 2 * nsel));
    double *accS = (double *)c->b_accS.p, *s1 = (double *)c->b_s1.p, *s2 = s1 + nsel;
    double *d_r1 = (double *)c->b_r1.p, *d_r2 = (double *)c->b_r2.p, *d_pi0 = (double *)c->b_pi0.p;
    const int *pos_pair = (const int *)c->b_gmeta.p;
    const int *sp_first = pos_pair + c->grp_possp_off + np;
    const Basis B = dev_basis(c);
    const PairTab PT = dev_pairs(c);
    {
        PhaseTimer tm(c, PH_STATS);
        hipLaunchKernelGGL(k_sel_zero, dim3(nblk(nsel, 256)), dim3(256), 0, c->stream, nsel,
                           (const int64_t *)d_tri, accS, s1, s2);
        ZV_HIPCHECK(c, hipGetLastError());
    }
    for (int b = 0; b < nbl; b++) {
        for (int64_t m0 = 0; m0 < npts; m0 += Mc) {
            const int Mv = (int)((npts - m0) < Mc ? (npts - m0) : Mc);
            const uint64_t g0 = (uint64_t)b * (uint64_t)npts + (uint64_t)m0;
            {
                PhaseTimer tm(c, PH_SAMPLE);
                hipLaunchKernelGGL(k_draw, dim3(nblk(Mv, 256)), dim3(256), 0, c->stream, seed,
                                   stream, g0, Mv, Mv, d_r1, d_r2, d_pi0);
                ZV_HIPCHECK(c, hipGetLastError());
            }
            const dim3 dgrid(nblk(Mc, 256) * (unsigned)nspl);
            if (single) {
                float *comp = (float *)c->b_pd.p, *grp = comp + ncomp;
                {
                    PhaseTimer tm(c, PH_DENS);
                    hipLaunchKernelGGL(k_dens_group<float>, dgrid, dim3(256), 0, c->stream, B, PT,
                                       npp, pos_pair, sp_first, np, c->nsp, Mv, Mc, d_r1, d_r2,
                                       d_pi0, comp, grp, (const int *)d_spl);
                    ZV_HIPCHECK(c, hipGetLastError());
                }
                PhaseTimer tm(c, PH_EST);
                hipLaunchKernelGGL(k_est_pairs<float>, dim3(nup), dim3(256), 0, c->stream,
                                   (const float *)comp, (const float *)grp, (const int *)d_up,
                                   pos_pair, Mv, Mc, np, c->grp_nsp_pad, accS);
                ZV_HIPCHECK(c, hipGetLastError());
            } else {
                double *comp = (double *)c->b_pd.p, *grp = comp + ncomp;
                {
                    PhaseTimer tm(c, PH_DENS);
                    hipLaunchKernelGGL(k_dens_group<double>, dgrid, dim3(256), 0, c->stream, B,
                                       PT, npp, pos_pair, sp_first, np, c->nsp, Mv, Mc, d_r1,
                                       d_r2, d_pi0, comp, grp, (const int *)d_spl);
                    ZV_HIPCHECK(c, hipGetLastError());
                }
                PhaseTimer tm(c, PH_EST);
                hipLaunchKernelGGL(k_est_pairs<double>, dim3(nup), dim3(256), 0, c->stream,
                                   (const double *)comp, (const double *)grp, (const int *)d_up,
                                   pos_pair, Mv, Mc, np, c->grp_nsp_pad, accS);
                ZV_HIPCHECK(c, hipGetLastError());
            }
        }
        PhaseTimer tm(c, PH_STATS);
        hipLaunchKernelGGL(k_sel_block_end, dim3(nblk(nsel, 256)), dim3(256), 0, c->stream, nsel,
                           (const int64_t *)d_tri, 1.0 / (double)npts, accS, s1, s2);
        ZV_HIPCHECK(c, hipGetLastError());
    }
    {
        PhaseTimer tm(c, PH_STATS);
        hipLaunchKernelGGL(k_sel_final, dim3(nblk(nsel, 256)), dim3(256), 0, c->stream, nsel, nbl,
                           (const double *)s1, (const double *)s2, (const int64_t *)d_list,
                           res_mean, res_err);
        ZV_HIPCHECK(c, hipGetLastError());
    }
    return zv_run_result_max(c, ZV_RESULT_STDERR, max_err);
}
