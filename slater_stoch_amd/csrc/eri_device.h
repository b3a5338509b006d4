// Device-side pieces shared by the ERI kernels (component and shell
// kernels): Hermite index tables, the primitive-pair layout, the Boys
// function of src/rint_sum.F90 and the McMurchie-Davidson R table.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <utility>
#include <vector>

#include "zvmc_internal.h"

namespace zveri {

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

// ------------------------------------------------------------------ E sparsity
// Structural zeros of a pair's Hermite coefficients.  For AOs with Cartesian
// powers a and b, E_tuv = E^x_t E^y_u E^z_v is nonzero only for t <= a_x +
// b_x, u <= a_y + b_y, v <= a_z + b_z (a box over the per-axis sums S = a +
// b); an r^2-type AO (i_type 2) sums its three variants with 2 added to one
// power, so a pair with nr such AOs has the union of the boxes of S plus 2
// on nr axes.  HermPatterns lists, for every pair order L <= ZV_LMAX_T, each
// nonzero set that can occur (nr = 0, 1, 2 and |S| + 2 nr = L), with its
// Hermite indices in ascending order.  k_primpairs writes exact zeros
// outside the set, so sums over it equal sums over all indices.
constexpr int kPatMax = 22, kNzMax = 25;
constexpr uint64_t herm_box(int bx, int by, int bz)
{
    uint64_t m = 0;
    for (int t = 0; t <= bx; t++)
        for (int u = 0; u <= by; u++)
            for (int v = 0; v <= bz; v++) m |= 1ull << hidx(t, u, v);
    return m;
}
// nonzero set of per-axis sums S with nr r^2-type AOs
constexpr uint64_t herm_union(int sx, int sy, int sz, int nr)
{
    if (nr == 0) return herm_box(sx, sy, sz);
    uint64_t m = 0;
    for (int v = 0; v < 3; v++)
        m |= herm_union(sx + (v == 0 ? 2 : 0), sy + (v == 1 ? 2 : 0), sz + (v == 2 ? 2 : 0),
                        nr - 1);
    return m;
}
struct HermPatterns {
    int n[ZV_LMAX_T + 1] = {};
    uint64_t mask[ZV_LMAX_T + 1][kPatMax] = {};
    int nz[ZV_LMAX_T + 1][kPatMax] = {};
    int h[ZV_LMAX_T + 1][kPatMax][kNzMax] = {};
    int nzmax[ZV_LMAX_T + 1] = {};
    constexpr HermPatterns()
    {
        for (int L = 0; L <= ZV_LMAX_T; L++)
            for (int nr = 0; 2 * nr <= L; nr++) {
                const int s = L - 2 * nr;
                for (int sx = s; sx >= 0; sx--)
                    for (int sy = s - sx; sy >= 0; sy--) {
                        const uint64_t m = herm_union(sx, sy, s - sx - sy, nr);
                        bool seen = false;
                        for (int p = 0; p < n[L]; p++) seen = seen || mask[L][p] == m;
                        if (seen) continue;
                        const int p = n[L]++;
                        mask[L][p] = m;
                        for (int hh = 0; hh < 64; hh++)
                            if ((m >> hh) & 1) h[L][p][nz[L][p]++] = hh;
                        if (nz[L][p] > nzmax[L]) nzmax[L] = nz[L][p];
                    }
            }
    }
    // pattern index of a nonzero set of order L, -1 if it is not listed
    constexpr int find(int L, uint64_t m) const
    {
        for (int p = 0; p < n[L]; p++)
            if (mask[L][p] == m) return p;
        return -1;
    }
};
constexpr HermPatterns kHP{};
static_assert(kHP.n[ZV_LMAX_T] <= kPatMax && kHP.nzmax[ZV_LMAX_T] <= kNzMax, "pattern table");

// ------------------------------------------------------------------ pair data
struct PP {
    double *p, *ip, *Px, *Py, *Pz, *c; // [pp_total], compacted per AO pair
    double *E;                          // [e_total]: pair P at e_off[P] + h*cap + e
    int *cnt;                           // kept entries per AO pair
    const int64_t *off, *eoff;          // per pair
    const int *cap;                     // n_gauss(i) n_gauss(k)
    const int *L;                       // Hermite order per pair
    const double *btab;                 // Boys Taylor table [L][k][i] (boys_table)
};

// ------------------------------------------------------------------ Boys
// F_n(T) = int_0^1 t^2n exp(-T t^2) dt for n = 0..L, as src/rint_sum.F90:
// T < 1: Taylor series of F_L (20 terms) and downward recursion;
// T >= 1: F_0 from erf and upward recursion, exp(-T) = 0 beyond T = 80.
// From T = 36 on erf(sqrt T) = 1 - erfc(6) rounds to 1.0 (erfc(6) = 2.2e-17
// is below half an ulp of 1), so the call is skipped there: same value,
// and a wave whose lanes all have T >= 36 (distant pairs) skips the erf.
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
        double ef = 1.0;
        if (T < 36.0) ef = erf(u);
        F[0] = ((0.5 * u_inv) * kSqrtPi) * ef;
        if (L > 0) {
            double e = T > 80.0 ? 0.0 : exp(-T);
            double tri = (0.5 * u_inv) * u_inv;
#pragma unroll
            for (int m = 0; m < L; m++) F[m + 1] = (F[m] * (2 * m + 1) - e) * tri;
        }
    }
}

// Tabulated Boys function (the shell kernels): for T < 36 a Taylor
// expansion of F_L about the nearest grid point T_i = i / 2,
//   F_L(T_i + d) = sum_k (-1)^k F_{L+k}(T_i) d^k / k!,  |d| <= 1/4, k <= 14
// (remainder below 1e-21 of F_L), then the downward recursion
//   F_m = (2 T F_{m+1} + exp(-T)) / (2m + 1),
// which is stable; the table holds (-1)^k F_{L+k}(T_i) / k! for L <= 8,
// computed on the host in extended precision from the series
// F_m(T) = exp(-T) sum_j (2T)^j / ((2m+1)(2m+3)...(2m+2j+1)) (zv_boys_table).
// From T = 36 on the branch of boys<L> (F_0 = sqrt(pi/T)/2, erf rounding to
// 1, upward recursion).  Relative error below 1e-15 against the series in
// extended precision; one code path for all T < 36, where boys<L> splits at
// T = 1 and evaluates erf above it.  lt: the kernel's LDS copy of the
// L-slice, [k][i].
constexpr int kBoysK = 14, kBoysN = 73; // orders 0..kBoysK, grid 0, 0.5, .., 36
constexpr int kBoysLmax = 8;
constexpr int kBoysSlice = (kBoysK + 1) * kBoysN;

template <int L> __device__ inline void boys_table(double T, const double *lt, double *F)
{
    if (T < 36.0) {
        const int i = (int)(T * 2.0 + 0.5);
        const double d = T - 0.5 * (double)i;
        double s = lt[kBoysK * kBoysN + i];
#pragma unroll
        for (int k = kBoysK - 1; k >= 0; k--) s = fma(s, d, lt[k * kBoysN + i]);
        F[L] = s;
        if (L > 0) {
            const double e = exp(-T);
#pragma unroll
            for (int m = L - 1; m >= 0; m--)
                F[m] = fma(2.0 * T, F[m + 1], e) * (1.0 / (2 * m + 1));
        }
    } else {
        double u_inv = rsqrt(T);
        F[0] = (0.5 * u_inv) * kSqrtPi;
        if (L > 0) {
            double e = T > 80.0 ? 0.0 : exp(-T);
            double tri = (0.5 * u_inv) * u_inv;
#pragma unroll
            for (int m = 0; m < L; m++) F[m + 1] = (F[m] * (2 * m + 1) - e) * tri;
        }
    }
}

// the L-slice of the table into LDS (one workgroup of 64 threads)
template <int L> __device__ inline void boys_table_load(const double *btab, double *lt)
{
    const double *src = btab + (size_t)L * kBoysSlice;
    for (int t = threadIdx.x; t < kBoysSlice; t += 64) lt[t] = src[t];
    __syncthreads();
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

// 1/sqrt(x) for a positive normal x (an exponent sum p + q in the shell
// kernels): the hardware estimate (about 5e-8 relative) and two Newton
// steps y' = y + y (1 - x y^2) / 2, a few ulp; it replaces a correctly
// rounded division and square root (1/s and sqrt(1/s)) in the prefactor.
__device__ inline double rsqrt_pos(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    double e = fma(-(x * y), y, 1.0);
    y = fma(y * 0.5, e, y);
    e = fma(-(x * y), y, 1.0);
    return fma(y * 0.5, e, y);
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

__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

} // namespace zveri

// ---- host side of the shell kernels (k_eri_shell.hip) ------------------
int zv_shell_ready(zv_context *c, const zveri::PP &pp);
int zv_shell_run(zv_context *c, const zveri::PP &pp, const std::vector<std::pair<int, int>> &sq,
                 DevBuf &bitem, DevBuf &bgroup, DevBuf &bpart, std::vector<char> &hplan,
                 int64_t *ngroups);
int zv_shell_scatter(zv_context *c, const DevBuf &bgroup, const DevBuf &bpart, int64_t ngroups,
                     const double *pre, double *tri);
int zv_zero(zv_context *c, double *x, int64_t n);
