// Analytic Gaussian-fit ERIs on the GPU (the reference's gauss_ijkl,
// src/gauss_ijkl.F90:51-154, and its Schwarz driver src/two_e_slater.F90:219-284).
//
// Each AO is phi(r) = poly(r - A) * sum_m d_m exp(-g_m |r - A|^2), where the
// reference's i_type 2 orbitals carry an extra r^2 = x^2 + y^2 + z^2 that it
// expands into three Cartesian terms (src/gauss_ijkl.F90:86-102).  Here a
// primitive pair (a, b) of AO pair (i, k) is written once as a Hermite
// expansion (McMurchie-Davidson):
//
//   phi_i^a phi_k^b = c_ab sum_{tuv} E_tuv Lambda_tuv(r; p, P)
//
// with c_ab = d_a d_b exp(-alpha beta/p |AB|^2) and the three r^2 terms
// summed into E_tuv.  The primitive quartet is then
//
//   (ab|cd) = C / (p q sqrt(p+q)) sum_tuv E^ab_tuv sum_t'u'v' (-1)^(t'+u'+v')
//             E^cd_t'u'v' R_{t+t',u+u',v+v'}(alpha = pq/(p+q), P - Q)
//
// with C = 34.9868366552, the reference's truncated 2 pi^(5/2)
// (src/bielec_integral.F90:29), and the contracted integral is the sum over
// the kept primitive pairs.  Primitive pairs whose Gaussian-product factor is
// cut by the reference (exponent > 40, src/gaussian_product.F90:25, or factor
// < 1e-15, src/give_explicit_poly_and_gaussian.F90:46) are dropped the same
// way.  The Boys values follow src/rint_sum.F90: series for T < 1, erf plus
// upward recursion for T >= 1.
//
// Work layout: quartets are grouped by the Hermite orders (LB, LK) of their
// two pairs (LB >= LK) and each class runs a kernel templated on (LB, LK) so
// the R table and the contraction are fully unrolled in registers.  One
// wavefront owns one quartet (or one slice of its bra primitive pairs): the
// bra primitive pair is wave-uniform, the 64 lanes stride over the ket
// primitive pairs.
#include <algorithm>
#include <cmath>
#include <cstdlib>

#include "zvmc_internal.h"

namespace {

constexpr double kPi52 = 34.9868366552;
constexpr double kSqrtPi = 1.77245385090551602729816748334;

constexpr int hcount(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }
// Hermite index of (t,u,v): by total s = t+u+v, then t descending, u descending
constexpr int hidx(int t, int u, int v)
{
    int s = t + u + v;
    return s * (s + 1) * (s + 2) / 6 + (s - t) * (s - t + 1) / 2 + (s - t - u);
}
struct HDecode {
    int t[969], u[969], v[969];
    constexpr HDecode() : t(), u(), v()
    {
        for (int s = 0; s <= 16; s++)
            for (int tt = s; tt >= 0; tt--)
                for (int uu = s - tt; uu >= 0; uu--) {
                    int h = hidx(tt, uu, s - tt - uu);
                    t[h] = tt;
                    u[h] = uu;
                    v[h] = s - tt - uu;
                }
    }
};
constexpr HDecode kH{};

// ------------------------------------------------------------------ pair data
struct PP {
    double *p, *ip, *Px, *Py, *Pz, *c; // [pp_total], compacted per AO pair
    double *E;                          // [e_total]: pair P at e_off[P] + h*cap + e
    int *cnt;                           // kept entries per AO pair
    const int64_t *off, *eoff;          // per pair
    const int *cap;                     // n_gauss(i) n_gauss(k)
    const int *L;                       // Hermite order per pair
};

struct BasisE {
    const double *center, *c1, *c2, *g;
    const int *npower, *itype, *ngauss;
};

// 1D Hermite coefficients E[a][b][t] (a <= A, b <= B), E[0][0][0] = 1.
constexpr int kEA = 5, kET = 9; // per-axis power <= 4 (validated in zv_set_basis)
__device__ void hermite_1d(double ip2, double xpa, double xpb, int A, int Bm,
                           double (&E)[kEA][kEA][kET])
{
    for (int a = 0; a <= A; a++)
        for (int b = 0; b <= Bm; b++)
            for (int t = 0; t < kET; t++) E[a][b][t] = 0.0;
    E[0][0][0] = 1.0;
    for (int a = 0; a <= A; a++) {
        if (a > 0)
            for (int t = 0; t <= a; t++) {
                double v = xpa * E[a - 1][0][t];
                if (t > 0) v += ip2 * E[a - 1][0][t - 1];
                if (t + 1 <= a - 1) v += (t + 1) * E[a - 1][0][t + 1];
                E[a][0][t] = v;
            }
        for (int b = 1; b <= Bm; b++)
            for (int t = 0; t <= a + b; t++) {
                double v = xpb * E[a][b - 1][t];
                if (t > 0) v += ip2 * E[a][b - 1][t - 1];
                if (t + 1 <= a + b - 1) v += (t + 1) * E[a][b - 1][t + 1];
                E[a][b][t] = v;
            }
    }
}

// One wavefront per AO pair: all n_gauss(i) x n_gauss(k) primitive pairs,
// cut as in the reference, compacted in order (ballot prefix), with their
// Hermite coefficients summed over the r^2 variants of i_type 2 orbitals.
__global__ __launch_bounds__(64) void k_primpairs(BasisE B, const int *pair_i, const int *pair_k,
                                                  PP pp)
{
    int P = blockIdx.x;
    int lane = threadIdx.x;
    int i = pair_i[P] - 1, k = pair_k[P] - 1;
    int ngi = B.ngauss[i], ngk = B.ngauss[k];
    int cap = ngi * ngk;
    int L = pp.L[P];
    int H = hcount(L);
    const double *A = B.center + 3 * i, *Bc = B.center + 3 * k;
    int ti = B.itype[i], tk = B.itype[k];
    const int *pa = B.npower + 3 * i, *pb = B.npower + 3 * k;
    const double *di = (ti == 1 ? B.c1 : B.c2) + (size_t)i * ZV_GMAX;
    const double *dk = (tk == 1 ? B.c1 : B.c2) + (size_t)k * ZV_GMAX;
    int64_t off = pp.off[P], eoff = pp.eoff[P];
    int base = 0;
    for (int e0 = 0; e0 < cap; e0 += 64) {
        int e = e0 + lane;
        bool keep = false;
        double p = 0, Pc[3] = {0, 0, 0}, coef = 0;
        int a = 0, b = 0;
        if (e < cap) {
            a = e / ngk;
            b = e % ngk;
            double al = B.g[(size_t)i * ZV_GMAX + a], be = B.g[(size_t)k * ZV_GMAX + b];
            // src/gaussian_product.F90:17-32
            p = al + be;
            double pinv = 1.0 / (al + be);
            double ab = (al * be) * pinv;
            double x0 = A[0] - Bc[0], x1 = A[1] - Bc[1], x2 = A[2] - Bc[2];
            double kk = ab * ((x0 * x0 + x1 * x1) + x2 * x2);
            if (kk <= 40.0) {
                double fk = exp(-kk);
                if (!(fk < 1.0e-15)) {
                    keep = true;
                    for (int d = 0; d < 3; d++) Pc[d] = (al * A[d] + be * Bc[d]) * pinv;
                    coef = (di[a] * dk[b]) * fk;
                }
            }
        }
        uint64_t mask = __ballot(keep);
        int slot = base + __popcll(mask & ((1ull << lane) - 1ull));
        base += __popcll(mask);
        if (!keep) continue;
        int64_t at = off + slot;
        pp.p[at] = p;
        pp.ip[at] = 1.0 / p;
        pp.Px[at] = Pc[0];
        pp.Py[at] = Pc[1];
        pp.Pz[at] = Pc[2];
        pp.c[at] = coef;
        // Hermite expansion, summed over the r^2 variants
        double ip2 = 0.5 / p;
        double E1[3][kEA][kEA][kET];
        int Amax[3], Bmax[3];
        for (int d = 0; d < 3; d++) {
            Amax[d] = pa[d] + (ti == 2 ? 2 : 0);
            Bmax[d] = pb[d] + (tk == 2 ? 2 : 0);
            hermite_1d(ip2, Pc[d] - A[d], Pc[d] - Bc[d], Amax[d], Bmax[d], E1[d]);
        }
        double E3[165];
        for (int h = 0; h < H; h++) E3[h] = 0.0;
        int nvi = ti == 2 ? 3 : 1, nvk = tk == 2 ? 3 : 1;
        for (int vi = 0; vi < nvi; vi++)
            for (int vk = 0; vk < nvk; vk++) {
                int ai[3], bk[3];
                for (int d = 0; d < 3; d++) {
                    ai[d] = pa[d] + ((ti == 2 && d == vi) ? 2 : 0);
                    bk[d] = pb[d] + ((tk == 2 && d == vk) ? 2 : 0);
                }
                for (int t = 0; t <= ai[0] + bk[0]; t++)
                    for (int u = 0; u <= ai[1] + bk[1]; u++)
                        for (int v = 0; v <= ai[2] + bk[2]; v++)
                            E3[hidx(t, u, v)] += E1[0][ai[0]][bk[0]][t] *
                                                 E1[1][ai[1]][bk[1]][u] *
                                                 E1[2][ai[2]][bk[2]][v];
            }
        for (int h = 0; h < H; h++) pp.E[eoff + (int64_t)h * cap + slot] = E3[h];
    }
    if (lane == 0) pp.cnt[P] = base;
}

// ------------------------------------------------------------------ Boys
// F_n(T) = int_0^1 t^2n exp(-T t^2) dt for n = 0..L, as src/rint_sum.F90:
// T < 1: Taylor series of F_L (20 terms) and downward recursion;
// T >= 1: F_0 from erf and upward recursion, exp(-T) = 0 beyond T = 80.
template <int L> struct BoysCoef {
    double c[20];
    constexpr BoysCoef() : c()
    {
        double f = 1.0;
        for (int k = 0; k < 20; k++) {
            if (k > 0) f = f * k;
            c[k] = ((k & 1) ? -1.0 : 1.0) / (f * (2 * L + 2 * k + 1));
        }
    }
};

template <int L> __device__ inline void boys(double T, double *F)
{
    if (T < 1.0) {
        constexpr BoysCoef<L> bc{};
        double s = bc.c[19];
#pragma unroll
        for (int k = 18; k >= 0; k--) s = fma(s, T, bc.c[k]);
        F[L] = s;
        if (L > 0) {
            double e = exp(-T);
#pragma unroll
            for (int m = L - 1; m >= 0; m--) F[m] = (2.0 * T * F[m + 1] + e) * (1.0 / (2 * m + 1));
        }
    } else {
        double u_inv = rsqrt(T);
        double u = T * u_inv;
        F[0] = ((0.5 * u_inv) * kSqrtPi) * erf(u);
        if (L > 0) {
            double e = T > 80.0 ? 0.0 : exp(-T);
            double tri = (0.5 * u_inv) * u_inv;
#pragma unroll
            for (int m = 0; m < L; m++) F[m + 1] = (F[m] * (2 * m + 1) - e) * tri;
        }
    }
}

// generic-L version (runtime loop)
__device__ inline void boys_rt(int L, double T, double *F)
{
    if (T < 1.0) {
        double term = 1.0, s = 1.0 / (2 * L + 1);
        for (int k = 1; k < 40; k++) {
            term = -term * T / k;
            double d = term / (2 * L + 2 * k + 1);
            s += d;
            if (fabs(d) < 1e-18) break;
        }
        F[L] = s;
        double e = exp(-T);
        for (int m = L - 1; m >= 0; m--) F[m] = (2.0 * T * F[m + 1] + e) / (2 * m + 1);
    } else {
        double u_inv = rsqrt(T);
        double u = T * u_inv;
        F[0] = ((0.5 * u_inv) * kSqrtPi) * erf(u);
        double e = T > 80.0 ? 0.0 : exp(-T);
        double tri = (0.5 * u_inv) * u_inv;
        for (int m = 0; m < L; m++) F[m + 1] = (F[m] * (2 * m + 1) - e) * tri;
    }
}

// ------------------------------------------------------------------ R table
// In place, levels n = L..0, totals s descending inside a level so that the
// level-(n+1) entries a value needs are still intact:
//   R^n_{t+1,u,v} = t R^{n+1}_{t-1,u,v} + X R^{n+1}_{t,u,v}   (same for u, v)
//   R^n_{000} = (-2 alpha)^n F_n
template <int L>
__device__ inline void hermite_R(double alpha, double X, double Y, double Z, const double *F,
                                 double *R)
{
    double m2a = -2.0 * alpha;
    double pw = 1.0;
    double pws[L + 1];
#pragma unroll
    for (int n = 0; n <= L; n++) {
        pws[n] = pw;
        pw *= m2a;
    }
#pragma unroll
    for (int n = L; n >= 0; n--) {
#pragma unroll
        for (int s = L - n; s >= 1; s--) {
#pragma unroll
            for (int t = s; t >= 0; t--) {
#pragma unroll
                for (int u = s - t; u >= 0; u--) {
                    int v = s - t - u;
                    double val;
                    if (t > 0) {
                        val = X * R[hidx(t - 1, u, v)];
                        if (t > 1) val = fma((double)(t - 1), R[hidx(t - 2, u, v)], val);
                    } else if (u > 0) {
                        val = Y * R[hidx(t, u - 1, v)];
                        if (u > 1) val = fma((double)(u - 1), R[hidx(t, u - 2, v)], val);
                    } else {
                        val = Z * R[hidx(t, u, v - 1)];
                        if (v > 1) val = fma((double)(v - 1), R[hidx(t, u, v - 2)], val);
                    }
                    R[hidx(t, u, v)] = val;
                }
            }
        }
        R[0] = pws[n] * F[n];
    }
}

__device__ inline void hermite_R_rt(int L, double alpha, double X, double Y, double Z,
                                    const double *F, double *R)
{
    double m2a = -2.0 * alpha;
    for (int n = L; n >= 0; n--) {
        for (int s = L - n; s >= 1; s--)
            for (int t = s; t >= 0; t--)
                for (int u = s - t; u >= 0; u--) {
                    int v = s - t - u;
                    double val;
                    if (t > 0) {
                        val = X * R[hidx(t - 1, u, v)];
                        if (t > 1) val += (t - 1) * R[hidx(t - 2, u, v)];
                    } else if (u > 0) {
                        val = Y * R[hidx(t, u - 1, v)];
                        if (u > 1) val += (u - 1) * R[hidx(t, u - 2, v)];
                    } else {
                        val = Z * R[hidx(t, u, v - 1)];
                        if (v > 1) val += (v - 1) * R[hidx(t, u, v - 2)];
                    }
                    R[hidx(t, u, v)] = val;
                }
        double pw = 1.0;
        for (int t = 0; t < n; t++) pw *= m2a;
        R[0] = pw * F[n];
    }
}

// ------------------------------------------------------------------ classes
enum QMode { Q_DIAG = 0, Q_TRI = 1, Q_RECT = 2, Q_LIST = 3 };

struct QSrc {
    int mode;
    const int *listB, *listK; // pair lists of the two classes
    int nB, nK;
    const int *qb, *qk;       // explicit bra / ket pairs (Q_LIST)
    int64_t nq;
    const double *pre;        // Schwarz diagonal per pair (Q_TRI, Q_RECT), may be null
};

__device__ inline void quartet_of(const QSrc &S, int64_t x, int &Pb, int &Pk, int64_t &oidx)
{
    if (S.mode == Q_DIAG) {
        Pb = Pk = S.listB[x];
        oidx = zv_tri(Pb, Pb);
    } else if (S.mode == Q_TRI) {
        // strict lower triangle of the list: x -> (xi > yi)
        int64_t xi = (int64_t)((sqrt(8.0 * (double)x + 1.0) + 1.0) * 0.5);
        while (xi * (xi - 1) / 2 > x) xi--;
        while ((xi + 1) * xi / 2 <= x) xi++;
        int64_t yi = x - xi * (xi - 1) / 2;
        Pb = S.listB[xi];
        Pk = S.listB[yi];
        oidx = Pb > Pk ? zv_tri(Pb, Pk) : zv_tri(Pk, Pb);
    } else if (S.mode == Q_RECT) {
        Pb = S.listB[x / S.nK];
        Pk = S.listK[x % S.nK];
        oidx = Pb > Pk ? zv_tri(Pb, Pk) : zv_tri(Pk, Pb);
    } else {
        Pb = S.qb[x];
        Pk = S.qk[x];
        oidx = x;
    }
}

__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Loop order: each lane owns one ket primitive pair (its position, exponent,
// coefficient and Hermite coefficients stay in registers); the bra primitive
// pairs are walked in the inner loop with wave-uniform (scalar) loads, so the
// hot loop issues no vector memory traffic.
template <int LB, int LK>
__global__ __launch_bounds__(256) void k_eri_class(PP pp, QSrc S, int nsplit, double *out,
                                                   double *part)
{
    constexpr int HB = hcount(LB), HK = hcount(LK), L = LB + LK, HR = hcount(L);
    int lane = threadIdx.x & 63;
    int64_t item = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (item >= S.nq * nsplit) return;
    int64_t x = item / nsplit;
    int sp = (int)(item % nsplit);
    int Pb, Pk;
    int64_t oidx;
    quartet_of(S, x, Pb, Pk, oidx);
    double acc = 0.0;
    bool skip = S.pre && (S.mode == Q_TRI || S.mode == Q_RECT) &&
                fabs(S.pre[Pb] * S.pre[Pk]) < 1.0e-30;
    if (!skip) {
        int nb = pp.cnt[Pb], nk = pp.cnt[Pk];
        int chunk = (nb + nsplit - 1) / nsplit;
        int a0 = sp * chunk, a1 = min(nb, a0 + chunk);
        int64_t ob = pp.off[Pb], ok = pp.off[Pk];
        int64_t eb = pp.eoff[Pb], ek = pp.eoff[Pk];
        int capb = pp.cap[Pb], capk = pp.cap[Pk];
        for (int c0 = 0; c0 < nk; c0 += 64) {
            int c = c0 + lane;
            bool live = c < nk;
            int cc = live ? c : c0; // in-range index for idle lanes
            double q = pp.p[ok + cc], iq = pp.ip[ok + cc];
            double Qx = pp.Px[ok + cc], Qy = pp.Py[ok + cc], Qz = pp.Pz[ok + cc];
            double ck = live ? pp.c[ok + cc] : 0.0;
            double Ek[HK];
#pragma unroll
            for (int h = 0; h < HK; h++)
                Ek[h] = ((kH.t[h] + kH.u[h] + kH.v[h]) & 1) ? -pp.E[ek + (int64_t)h * capk + cc]
                                                             : pp.E[ek + (int64_t)h * capk + cc];
            for (int a = a0; a < a1; a++) {
                double p = pp.p[ob + a], ip = pp.ip[ob + a];
                double X = pp.Px[ob + a] - Qx, Y = pp.Py[ob + a] - Qy, Z = pp.Pz[ob + a] - Qz;
                double cb = pp.c[ob + a];
                double s = p + q;
                double is = 1.0 / s;
                double alpha = (p * q) * is;
                double T = alpha * ((X * X + Y * Y) + Z * Z);
                double F[L + 1];
                boys<L>(T, F);
                double R[HR];
                hermite_R<L>(alpha, X, Y, Z, F, R);
                double sum = 0.0;
#pragma unroll
                for (int hb = 0; hb < HB; hb++) {
                    double w = 0.0;
#pragma unroll
                    for (int hk = 0; hk < HK; hk++)
                        w = fma(Ek[hk], R[hidx(kH.t[hb] + kH.t[hk], kH.u[hb] + kH.u[hk],
                                                kH.v[hb] + kH.v[hk])], w);
                    sum = fma(pp.E[eb + (int64_t)hb * capb + a], w, sum);
                }
                double pref = ((kPi52 * ip) * iq) * sqrt(is);
                acc = fma((cb * ck) * pref, sum, acc);
            }
        }
    }
    acc = wave_sum(acc);
    if (lane == 0) {
        if (nsplit == 1) out[oidx] = acc;
        else part[item] = acc;
    }
}

// Same contraction with runtime orders (pairs beyond l = 4 per side).
__global__ __launch_bounds__(64) void k_eri_generic(PP pp, QSrc S, int nsplit, double *out,
                                                    double *part)
{
    int lane = threadIdx.x & 63;
    int64_t item = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (item >= S.nq * nsplit) return;
    int64_t x = item / nsplit;
    int sp = (int)(item % nsplit);
    int Pb, Pk;
    int64_t oidx;
    quartet_of(S, x, Pb, Pk, oidx);
    double acc = 0.0;
    bool skip = S.pre && (S.mode == Q_TRI || S.mode == Q_RECT) &&
                fabs(S.pre[Pb] * S.pre[Pk]) < 1.0e-30;
    if (!skip) {
        int LB = pp.L[Pb], LK = pp.L[Pk], L = LB + LK;
        int HB = hcount(LB), HK = hcount(LK);
        int nb = pp.cnt[Pb], nk = pp.cnt[Pk];
        int chunk = (nb + nsplit - 1) / nsplit;
        int a0 = sp * chunk, a1 = min(nb, a0 + chunk);
        int64_t ob = pp.off[Pb], ok = pp.off[Pk];
        int64_t eb = pp.eoff[Pb], ek = pp.eoff[Pk];
        int capb = pp.cap[Pb], capk = pp.cap[Pk];
        double F[17], R[969];
        for (int a = a0; a < a1; a++) {
            double p = pp.p[ob + a], ip = pp.ip[ob + a];
            double Px = pp.Px[ob + a], Py = pp.Py[ob + a], Pz = pp.Pz[ob + a];
            double cb = pp.c[ob + a];
            for (int c = lane; c < nk; c += 64) {
                double q = pp.p[ok + c], iq = pp.ip[ok + c];
                double X = Px - pp.Px[ok + c], Y = Py - pp.Py[ok + c], Z = Pz - pp.Pz[ok + c];
                double s = p + q;
                double is = 1.0 / s;
                double alpha = (p * q) * is;
                double T = alpha * ((X * X + Y * Y) + Z * Z);
                boys_rt(L, T, F);
                hermite_R_rt(L, alpha, X, Y, Z, F, R);
                double sum = 0.0;
                for (int hb = 0; hb < HB; hb++) {
                    double w = 0.0;
                    for (int hk = 0; hk < HK; hk++) {
                        double ekv = pp.E[ek + (int64_t)hk * capk + c];
                        double r = R[hidx(kH.t[hb] + kH.t[hk], kH.u[hb] + kH.u[hk],
                                          kH.v[hb] + kH.v[hk])];
                        w += ((kH.t[hk] + kH.u[hk] + kH.v[hk]) & 1) ? -ekv * r : ekv * r;
                    }
                    sum += pp.E[eb + (int64_t)hb * capb + a] * w;
                }
                double pref = ((kPi52 * ip) * iq) * sqrt(is);
                acc += ((cb * pp.c[ok + c]) * pref) * sum;
            }
        }
    }
    acc = wave_sum(acc);
    if (lane == 0) {
        if (nsplit == 1) out[oidx] = acc;
        else part[item] = acc;
    }
}

// ordered sum of the split partials of each quartet
__global__ void k_reduce_parts(QSrc S, int nsplit, const double *part, double *out)
{
    int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= S.nq) return;
    int Pb, Pk;
    int64_t oidx;
    quartet_of(S, x, Pb, Pk, oidx);
    double s = 0.0;
    for (int t = 0; t < nsplit; t++) s += part[x * nsplit + t];
    out[oidx] = s;
}

__global__ void k_diag_pre(int np, const double *tri, double *pre)
{
    int P = blockIdx.x * blockDim.x + threadIdx.x;
    if (P < np) pre[P] = tri[zv_tri(P, P)];
}

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

template <int LB, int LK>
void launch_t(unsigned grid, hipStream_t st, PP pp, QSrc S, int ns, double *out, double *part)
{
    hipLaunchKernelGGL((k_eri_class<LB, LK>), dim3(grid), dim3(256), 0, st, pp, S, ns, out, part);
}

// LB >= LK, both <= 4
bool launch_class_t(int LB, int LK, unsigned grid, hipStream_t st, PP pp, QSrc S, int ns,
                    double *out, double *part)
{
    switch (LB * 5 + LK) {
    case 0: launch_t<0, 0>(grid, st, pp, S, ns, out, part); return true;
    case 5: launch_t<1, 0>(grid, st, pp, S, ns, out, part); return true;
    case 6: launch_t<1, 1>(grid, st, pp, S, ns, out, part); return true;
    case 10: launch_t<2, 0>(grid, st, pp, S, ns, out, part); return true;
    case 11: launch_t<2, 1>(grid, st, pp, S, ns, out, part); return true;
    case 12: launch_t<2, 2>(grid, st, pp, S, ns, out, part); return true;
    case 15: launch_t<3, 0>(grid, st, pp, S, ns, out, part); return true;
    case 16: launch_t<3, 1>(grid, st, pp, S, ns, out, part); return true;
    case 17: launch_t<3, 2>(grid, st, pp, S, ns, out, part); return true;
    case 18: launch_t<3, 3>(grid, st, pp, S, ns, out, part); return true;
    case 20: launch_t<4, 0>(grid, st, pp, S, ns, out, part); return true;
    case 21: launch_t<4, 1>(grid, st, pp, S, ns, out, part); return true;
    case 22: launch_t<4, 2>(grid, st, pp, S, ns, out, part); return true;
    case 23: launch_t<4, 3>(grid, st, pp, S, ns, out, part); return true;
    case 24: launch_t<4, 4>(grid, st, pp, S, ns, out, part); return true;
    }
    return false;
}

} // namespace

// ====================================================================== host
static int ensure_primpairs(zv_context *c, PP &pp)
{
    int np = c->npair;
    // layout inside b_pp: 6 planes of pp_total doubles; b_ppE: e_total doubles;
    // b_ppcnt: cnt[np] ints, off[np], eoff[np] int64, cap[np], L[np] ints
    ZV_CHECK(zv_ensure(c, c->b_pp, sizeof(double) * 6 * (size_t)c->pp_total));
    ZV_CHECK(zv_ensure(c, c->b_ppE, sizeof(double) * (size_t)c->e_total));
    size_t meta = sizeof(int) * np * 3 + sizeof(int64_t) * np * 2 + 64;
    ZV_CHECK(zv_ensure(c, c->b_ppcnt, meta));
    double *base = (double *)c->b_pp.p;
    size_t T = (size_t)c->pp_total;
    pp.p = base;
    pp.ip = base + T;
    pp.Px = base + 2 * T;
    pp.Py = base + 3 * T;
    pp.Pz = base + 4 * T;
    pp.c = base + 5 * T;
    pp.E = (double *)c->b_ppE.p;
    char *m = (char *)c->b_ppcnt.p;
    int64_t *d_off = (int64_t *)m;
    int64_t *d_eoff = d_off + np;
    int *d_cnt = (int *)(d_eoff + np);
    int *d_cap = d_cnt + np;
    int *d_L = d_cap + np;
    pp.off = d_off;
    pp.eoff = d_eoff;
    pp.cnt = d_cnt;
    pp.cap = d_cap;
    pp.L = d_L;
    if (c->pp_ready) return ZV_OK;
    std::vector<int> cap(np);
    for (int P = 0; P < np; P++) cap[P] = (int)(c->h_pp_off[P + 1] - c->h_pp_off[P]);
    ZV_HIPCHECK(c, hipMemcpyAsync(d_off, c->h_pp_off.data(), sizeof(int64_t) * np,
                                  hipMemcpyHostToDevice, c->stream));
    ZV_HIPCHECK(c, hipMemcpyAsync(d_eoff, c->h_e_off.data(), sizeof(int64_t) * np,
                                  hipMemcpyHostToDevice, c->stream));
    ZV_HIPCHECK(c, hipMemcpyAsync(d_cap, cap.data(), sizeof(int) * np, hipMemcpyHostToDevice,
                                  c->stream));
    ZV_HIPCHECK(c, hipMemcpyAsync(d_L, c->h_pair_L.data(), sizeof(int) * np,
                                  hipMemcpyHostToDevice, c->stream));
    BasisE B{c->d_center, c->d_c1, c->d_c2, c->d_g, c->d_npower, c->d_itype, c->d_ngauss};
    {
        PhaseTimer tm(c, PH_PRIM);
        hipLaunchKernelGGL(k_primpairs, dim3(np), dim3(64), 0, c->stream, B, c->d_pair_i,
                           c->d_pair_k, pp);
        ZV_HIPCHECK(c, hipGetLastError());
    }
    c->h_cnt.assign(np, 0);
    ZV_HIPCHECK(c, hipMemcpyAsync(c->h_cnt.data(), d_cnt, sizeof(int) * np, hipMemcpyDeviceToHost,
                                  c->stream));
    ZV_HIPCHECK(c, hipStreamSynchronize(c->stream));
    c->pp_ready = true;
    return ZV_OK;
}

// primitive quartets a class launch evaluates (screened quartets excluded)
double class_pq(const zv_context *c, int mode, const std::vector<int> &lb,
                       const std::vector<int> &lk, const std::vector<double> *pre)
{
    const std::vector<int> &cnt = c->h_cnt;
    auto keep = [&](int P, int Q) {
        return !pre || fabs((*pre)[P] * (*pre)[Q]) >= 1.0e-30;
    };
    double s = 0.0;
    if (mode == Q_DIAG) {
        for (int P : lb) s += (double)cnt[P] * cnt[P];
    } else if (mode == Q_TRI) {
        for (size_t x = 0; x < lb.size(); x++)
            for (size_t y = 0; y < x; y++)
                if (keep(lb[x], lb[y])) s += (double)cnt[lb[x]] * cnt[lb[y]];
    } else {
        for (int P : lb)
            for (int Q : lk)
                if (keep(P, Q)) s += (double)cnt[P] * cnt[Q];
    }
    return s;
}

// Launch one class over the quartet source S, splitting bra primitive pairs
// when there are too few quartets to fill the device.
static int launch_class(zv_context *c, const PP &pp, QSrc S, int LB, int LK, double *out)
{
    if (S.nq <= 0) return ZV_OK;
    const int64_t target_waves = 16384;
    int nsplit = 1;
    if (S.nq < target_waves) nsplit = (int)std::min<int64_t>(32, (target_waves + S.nq - 1) / S.nq);
    double *part = nullptr;
    if (nsplit > 1) {
        ZV_CHECK(zv_ensure(c, c->b_part, sizeof(double) * S.nq * nsplit));
        part = (double *)c->b_part.p;
    }
    int64_t items = S.nq * nsplit;
    static const bool trace = getenv("ZV_TRACE") != nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (trace) {
        hipEventCreate(&t0);
        hipEventCreate(&t1);
        hipEventRecord(t0, c->stream);
    }
    PhaseTimer tm(c, PH_ERI, LB * 6 + LK);
    if (LB <= ZV_LMAX_T && LK <= ZV_LMAX_T) {
        if (!launch_class_t(LB, LK, nblk(items * 64, 256), c->stream, pp, S, nsplit, out, part))
            return zv_fail(c, ZV_ERR_STATE, "no ERI kernel for class (%d,%d)", LB, LK);
    } else {
        hipLaunchKernelGGL(k_eri_generic, dim3(nblk(items * 64, 64)), dim3(64), 0, c->stream, pp,
                           S, nsplit, out, part);
    }
    ZV_HIPCHECK(c, hipGetLastError());
    if (nsplit > 1) {
        hipLaunchKernelGGL(k_reduce_parts, dim3(nblk(S.nq, 256)), dim3(256), 0, c->stream, S,
                           nsplit, (const double *)part, out);
        ZV_HIPCHECK(c, hipGetLastError());
    }
    if (trace) {
        hipEventRecord(t1, c->stream);
        hipEventSynchronize(t1);
        float ms = 0.f;
        hipEventElapsedTime(&ms, t0, t1);
        fprintf(stderr, "[zvmc] eri class (%d,%d) mode %d quartets %lld split %d: %.3f ms\n", LB,
                LK, S.mode, (long long)S.nq, nsplit, ms);
        hipEventDestroy(t0);
        hipEventDestroy(t1);
    }
    return ZV_OK;
}

// pairs grouped by Hermite order (classes of the generic kernel are merged:
// every pair with L > ZV_LMAX_T goes through k_eri_generic, which reads L
// per pair, so they form one list with a pseudo-order of ZV_LMAX_T + 1)
static void class_lists(zv_context *c, std::vector<std::vector<int>> &lists)
{
    lists.assign(ZV_LMAX_T + 2, {});
    for (int P = 0; P < c->npair; P++) {
        int L = c->h_pair_L[P];
        lists[L > ZV_LMAX_T ? ZV_LMAX_T + 1 : L].push_back(P);
    }
}

int zv_run_gauss_eri(zv_context *c, double *ijkl, double *precond, int64_t *n_screened)
{
    PP pp;
    ZV_CHECK(ensure_primpairs(c, pp));
    int np = c->npair;
    std::vector<std::vector<int>> lists;
    class_lists(c, lists);
    int ncls = (int)lists.size();
    std::vector<int> flat;
    std::vector<int64_t> lo(ncls + 1, 0);
    for (int t = 0; t < ncls; t++) {
        lo[t] = (int64_t)flat.size();
        flat.insert(flat.end(), lists[t].begin(), lists[t].end());
    }
    lo[ncls] = (int64_t)flat.size();
    ZV_CHECK(zv_ensure(c, c->b_list, sizeof(int) * flat.size()));
    ZV_HIPCHECK(c, hipMemcpyAsync(c->b_list.p, flat.data(), sizeof(int) * flat.size(),
                                  hipMemcpyHostToDevice, c->stream));
    const int *dl = (const int *)c->b_list.p;
    ZV_CHECK(zv_ensure(c, c->b_tri, sizeof(double) * c->nint));
    ZV_CHECK(zv_ensure(c, c->b_misc, sizeof(double) * np));
    double *tri = (double *)c->b_tri.p, *pre = (double *)c->b_misc.p;
    auto order_of = [&](int t) { return t > ZV_LMAX_T ? ZV_LMAX_T + 1 : t; };
    // phase 1: diagonal quartets (ik|ik) = precond
    for (int t = 0; t < ncls; t++) {
        QSrc S{Q_DIAG, dl + lo[t], dl + lo[t], (int)lists[t].size(), (int)lists[t].size(),
               nullptr, nullptr, (int64_t)lists[t].size(), nullptr};
        ZV_CHECK(launch_class(c, pp, S, order_of(t), order_of(t), tri));
    }
    hipLaunchKernelGGL(k_diag_pre, dim3(nblk(np, 256)), dim3(256), 0, c->stream, np, tri, pre);
    ZV_HIPCHECK(c, hipGetLastError());
    std::vector<double> hpre(np);
    ZV_HIPCHECK(c, hipMemcpyAsync(hpre.data(), pre, sizeof(double) * np, hipMemcpyDeviceToHost,
                                  c->stream));
    // phase 2: off-diagonal quartets with the Cauchy-Schwarz test
    for (int tb = 0; tb < ncls; tb++)
        for (int tk = 0; tk <= tb; tk++) {
            int nb = (int)lists[tb].size(), nk = (int)lists[tk].size();
            QSrc S;
            if (tb == tk)
                S = QSrc{Q_TRI, dl + lo[tb], dl + lo[tb], nb, nb, nullptr, nullptr,
                         (int64_t)nb * (nb - 1) / 2, pre};
            else
                S = QSrc{Q_RECT, dl + lo[tb], dl + lo[tk], nb, nk, nullptr, nullptr,
                         (int64_t)nb * nk, pre};
            ZV_CHECK(launch_class(c, pp, S, order_of(tb), order_of(tk), tri));
        }
    ZV_CHECK(zv_run_scatter_ref(c, tri, ijkl));
    // work accounting is done on request (zv_last_eri_stats)
    c->last_lists = lists;
    c->last_pre = hpre;
    c->pq_pending = true;
    if (precond) {
        int n = c->n;
        for (int i = 1; i <= n; i++)
            for (int k = 1; k <= i; k++) {
                double v = hpre[zv_pidx(i, k)];
                precond[(size_t)(k - 1) * n + (i - 1)] = v;
                precond[(size_t)(i - 1) * n + (k - 1)] = v;
            }
    }
    if (n_screened) {
        // quartets (P >= Q) with |pre_P pre_Q| < 1e-30 (src/two_e_slater.F90:253-262)
        std::vector<double> v(np);
        for (int P = 0; P < np; P++) v[P] = fabs(hpre[P]);
        std::sort(v.begin(), v.end());
        int64_t cnt = 0;
        int hi = np - 1;
        for (int a = 0; a < np; a++) {
            while (hi >= 0 && v[a] * v[hi] >= 1.0e-30) hi--;
            // partners b <= hi satisfy the test; count unordered pairs with b <= a
            int64_t m = std::min(hi, a) + 1;
            cnt += m;
        }
        *n_screened = cnt;
    }
    return ZV_OK;
}

int zv_run_gauss_list(zv_context *c, int64_t nq, const int *qi, const int *qk, const int *qj,
                      const int *ql, double *out)
{
    PP pp;
    ZV_CHECK(ensure_primpairs(c, pp));
    int ncls = ZV_LMAX_T + 2;
    auto order = [&](int P) {
        int L = c->h_pair_L[P];
        return L > ZV_LMAX_T ? ZV_LMAX_T + 1 : L;
    };
    // bucket by (bra order >= ket order)
    std::vector<std::vector<int64_t>> bucket(ncls * ncls);
    std::vector<int> pb(nq), pk(nq);
    for (int64_t t = 0; t < nq; t++) {
        int i = std::max(qi[t], qk[t]), k = std::min(qi[t], qk[t]);
        int j = std::max(qj[t], ql[t]), l = std::min(qj[t], ql[t]);
        int P = zv_pidx(i, k), Q = zv_pidx(j, l);
        if (order(P) < order(Q)) std::swap(P, Q);
        pb[t] = P;
        pk[t] = Q;
        bucket[order(P) * ncls + order(Q)].push_back(t);
    }
    std::vector<int> qb, qk2;
    std::vector<int64_t> perm;
    std::vector<int64_t> lo(ncls * ncls + 1, 0);
    for (int b = 0; b < ncls * ncls; b++) {
        lo[b] = (int64_t)perm.size();
        for (int64_t t : bucket[b]) {
            perm.push_back(t);
            qb.push_back(pb[t]);
            qk2.push_back(pk[t]);
        }
    }
    lo[ncls * ncls] = (int64_t)perm.size();
    ZV_CHECK(zv_ensure(c, c->b_list, sizeof(int) * 2 * nq));
    int *dqb = (int *)c->b_list.p, *dqk = dqb + nq;
    ZV_HIPCHECK(c, hipMemcpyAsync(dqb, qb.data(), sizeof(int) * nq, hipMemcpyHostToDevice,
                                  c->stream));
    ZV_HIPCHECK(c, hipMemcpyAsync(dqk, qk2.data(), sizeof(int) * nq, hipMemcpyHostToDevice,
                                  c->stream));
    ZV_CHECK(zv_ensure(c, c->b_out, sizeof(double) * nq));
    double *dout = (double *)c->b_out.p;
    for (int b = 0; b < ncls * ncls; b++) {
        int64_t cnt = lo[b + 1] - lo[b];
        if (!cnt) continue;
        QSrc S{Q_LIST, nullptr, nullptr, 0, 0, dqb + lo[b], dqk + lo[b], cnt, nullptr};
        ZV_CHECK(launch_class(c, pp, S, b / ncls, b % ncls, dout + lo[b]));
    }
    std::vector<double> h(nq);
    ZV_HIPCHECK(c, hipMemcpyAsync(h.data(), dout, sizeof(double) * nq, hipMemcpyDeviceToHost,
                                  c->stream));
    ZV_HIPCHECK(c, hipStreamSynchronize(c->stream));
    for (int64_t t = 0; t < nq; t++) out[perm[t]] = h[t];
    return ZV_OK;
}

// per-class primitive-quartet counts of the last zv_gauss_eri (host work)
void zv_eri_account(const zv_context *cc)
{
    zv_context *c = const_cast<zv_context *>(cc);
    if (!c->pq_pending) return;
    const auto &lists = c->last_lists;
    int ncls = (int)lists.size();
    auto order_of = [&](int t) { return t > ZV_LMAX_T ? ZV_LMAX_T + 1 : t; };
    for (int t = 0; t < ncls; t++)
        c->cls_pq[order_of(t) * 6 + order_of(t)] += class_pq(c, Q_DIAG, lists[t], lists[t], nullptr);
    for (int tb = 0; tb < ncls; tb++)
        for (int tk = 0; tk <= tb; tk++)
            c->cls_pq[order_of(tb) * 6 + order_of(tk)] +=
                class_pq(c, tb == tk ? Q_TRI : Q_RECT, lists[tb], lists[tk], &c->last_pre);
    c->pq_pending = false;
}
