// Exact one-centre Slater integrals on the GPU: the do_exact_monocenter
// branch of src/two_e_slater.F90:318-365 (src/compute_int_slater.F90,
// src/ijkl_slater_xnynzn.F90, src/ijkl_slater.F90, src/bigF.F90,
// src/smallg.F90, src/compute_lmin_lmax.F90, src/gaunt.F90).
//
// For four AOs x^a y^b z^c r^nsto e^{-zeta r} on one centre (a+b+c <= 2,
// as the reference's formulas require), each angular factor
// (x/r)^a (y/r)^b (z/r)^c is a short sum of complex spherical harmonics
// (Condon-Shortley phase; the table below), the radial part is r^(n-1) with
// n = 1 + a + b + c + nsto, and with the multipole expansion of 1/r12
//
//   (ik|jl) = Re sum conj(c_i) c_k conj(c_j) c_l 4 pi sum_L F_L / (2L + 1)
//             (-1)^M G(l_i L l_k; m_i -M m_k) G(l_j L l_l; m_j M m_l),
//   M = m_k - m_i = m_j - m_l,
//
// G(l1 l2 l3; m1 m2 m3) = int Y*_{l1 m1} Y_{l2 m2} Y_{l3 m3} (Gaunt, from
// Wigner 3j symbols by Racah's formula) and F_L the radial Coulomb integral
// in closed form.  One thread per one-centre quartet.
#include <algorithm>
#include <cstring>
#include <map>
#include <vector>

#include "zvmc_internal.h"

namespace {

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// (x/r)^a (y/r)^b (z/r)^c = sum_t coef_t Y_{l_t m_t}, monomials of degree
// <= 2 in the order 1, x, y, z, xx, yy, zz, xy, xz, yz
struct YTerm {
    int l, m;
    double re, im;
};
struct YExp {
    int n;
    YTerm t[4];
};
constexpr double kS4pi = 3.5449077018110318;    // sqrt(4 pi)
constexpr double kS4pi3 = 2.046653415892977;    // sqrt(4 pi / 3)
constexpr double kS2pi3 = 1.4472025091165353;   // sqrt(2 pi / 3)
constexpr double kS2pi15 = 0.6472086375185664;  // sqrt(2 pi / 15)
constexpr double kD0 = 0.5284436396808014;      // (2/3) sqrt(pi / 5)
__constant__ YExp kYExp[10] = {
    {1, {{0, 0, kS4pi, 0.0}}},
    {2, {{1, -1, kS2pi3, 0.0}, {1, 1, -kS2pi3, 0.0}}},
    {2, {{1, -1, 0.0, kS2pi3}, {1, 1, 0.0, kS2pi3}}},
    {1, {{1, 0, kS4pi3, 0.0}}},
    {4, {{0, 0, kS4pi / 3.0, 0.0}, {2, -2, kS2pi15, 0.0}, {2, 0, -kD0, 0.0}, {2, 2, kS2pi15, 0.0}}},
    {4, {{0, 0, kS4pi / 3.0, 0.0}, {2, -2, -kS2pi15, 0.0}, {2, 0, -kD0, 0.0}, {2, 2, -kS2pi15, 0.0}}},
    {2, {{0, 0, kS4pi / 3.0, 0.0}, {2, 0, 2.0 * kD0, 0.0}}},
    {2, {{2, -2, 0.0, kS2pi15}, {2, 2, 0.0, -kS2pi15}}},
    {2, {{2, -1, kS2pi15, 0.0}, {2, 1, -kS2pi15, 0.0}}},
    {2, {{2, -1, 0.0, kS2pi15}, {2, 1, 0.0, kS2pi15}}},
};

__host__ __device__ inline int monomial(int a, int b, int c)
{
    const int d = a + b + c;
    if (d == 0) return 0;
    if (d == 1) return a ? 1 : (b ? 2 : 3);
    if (a == 2) return 4;
    if (b == 2) return 5;
    if (c == 2) return 6;
    return a && b ? 7 : (a ? 8 : 9);
}

__device__ inline double fct(int n)
{
    double f = 1.0;
    for (int i = 2; i <= n; i++) f *= (double)i;
    return f;
}

// Wigner 3j symbol, Racah's formula
__device__ double w3j(int j1, int j2, int j3, int m1, int m2, int m3)
{
    if (m1 + m2 + m3 != 0 || j3 < abs(j1 - j2) || j3 > j1 + j2) return 0.0;
    if (abs(m1) > j1 || abs(m2) > j2 || abs(m3) > j3) return 0.0;
    const double tri =
        fct(j1 + j2 - j3) * fct(j1 - j2 + j3) * fct(-j1 + j2 + j3) / fct(j1 + j2 + j3 + 1);
    const double pre = sqrt(tri * fct(j1 + m1) * fct(j1 - m1) * fct(j2 + m2) * fct(j2 - m2) *
                            fct(j3 + m3) * fct(j3 - m3));
    const int kmin = max(0, max(j2 - j3 - m1, j1 - j3 + m2));
    const int kmax = min(j1 + j2 - j3, min(j1 - m1, j2 + m2));
    double s = 0.0;
    for (int k = kmin; k <= kmax; k++)
        s += ((k & 1) ? -1.0 : 1.0) /
             (fct(k) * fct(j3 - j2 + k + m1) * fct(j3 - j1 + k - m2) * fct(j1 + j2 - j3 - k) *
              fct(j1 - k - m1) * fct(j2 - k + m2));
    return ((abs(j1 - j2 - m3) & 1) ? -1.0 : 1.0) * pre * s;
}

// int dOmega Y*_{l1 m1} Y_{l2 m2} Y_{l3 m3}
__device__ double gaunt(int l1, int l2, int l3, int m1, int m2, int m3)
{
    constexpr double kInv4pi = 0.07957747154594767; // 1 / (4 pi)
    const double f = sqrt((2.0 * l1 + 1.0) * (2.0 * l2 + 1.0) * (2.0 * l3 + 1.0) * kInv4pi);
    return ((abs(m1) & 1) ? -1.0 : 1.0) * f * w3j(l1, l2, l3, 0, 0, 0) *
           w3j(l1, l2, l3, -m1, m2, m3);
}

__device__ inline double ipow(double x, int n)
{
    double r = 1.0;
    for (int i = 0; i < n; i++) r *= x;
    return r;
}

// F_L = int r1^k1 e^{-g1 r1} [r1^-(L+1) int_0^r1 r2^(k2+L) e^{-g2 r2}
//                             + r1^L int_r1^inf r2^(k2-L-1) e^{-g2 r2}] dr1
__device__ double radial_F(int k1, double g1, int k2, double g2, int L)
{
    const double g = g1 + g2;
    double v = fct(k2 + L) / ipow(g2, k2 + L + 1) * fct(k1 - L - 1) / ipow(g1, k1 - L);
    for (int part = 0; part < 2; part++) {
        const int p = part == 0 ? -L - 1 : L;
        double s = 0.0;
        for (int i = 0; i <= k2 + p; i++)
            s += fct(k2 + p) * fct(k1 + k2 - 1 - i) / fct(k2 + p - i) / ipow(g2, i + 1) /
                 ipow(g, k1 + k2 - i);
        v += part == 0 ? s : -s;
    }
    return v;
}

// quartets q[4 t .. 4 t + 3] = (i, k, j, l), 1-based; value and kcp out
__global__ void k_mono_exact(int64_t nq, const int *__restrict__ q, const int *__restrict__ npower,
                             const int *__restrict__ nsto, const double *__restrict__ zeta,
                             const int64_t *__restrict__ kbase, int n, double *__restrict__ val,
                             int64_t *__restrict__ kcp)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nq) return;
    int ao[4], mono[4], nr[4];
    double z[4];
    for (int s = 0; s < 4; s++) {
        ao[s] = q[4 * t + s] - 1;
        const int a = npower[3 * ao[s]], b = npower[3 * ao[s] + 1], c = npower[3 * ao[s] + 2];
        mono[s] = monomial(a, b, c);
        nr[s] = 1 + a + b + c + nsto[ao[s]];
        z[s] = zeta[ao[s]];
    }
    const int k1 = nr[0] + nr[1], k2 = nr[2] + nr[3];
    const double g1 = z[0] + z[1], g2 = z[2] + z[3];
    // F_L for L = 0..4 (l <= 2 per factor)
    double F[5];
    for (int L = 0; L < 5; L++) F[L] = L <= min(k1 - 1, k2 - 1) ? radial_F(k1, g1, k2, g2, L) : 0.0;
    const YExp &e0 = kYExp[mono[0]], &e1 = kYExp[mono[1]], &e2 = kYExp[mono[2]],
               &e3 = kYExp[mono[3]];
    double sum = 0.0;
    for (int a = 0; a < e0.n; a++)
        for (int b = 0; b < e1.n; b++) {
            const YTerm &ti = e0.t[a], &tk = e1.t[b];
            const int M = tk.m - ti.m;
            // conj(c_i) c_k
            const double xr = ti.re * tk.re + ti.im * tk.im, xi = ti.re * tk.im - ti.im * tk.re;
            for (int c = 0; c < e2.n; c++)
                for (int d = 0; d < e3.n; d++) {
                    const YTerm &tj = e2.t[c], &tl = e3.t[d];
                    if (tj.m - tl.m != M) continue;
                    // conj(c_j) c_l
                    const double yr = tj.re * tl.re + tj.im * tl.im,
                                 yi = tj.re * tl.im - tj.im * tl.re;
                    const double cr = xr * yr - xi * yi; // real part of the product
                    if (fabs(cr) <= 1e-20 && fabs(xr * yi + xi * yr) <= 1e-20) continue;
                    const int Lmax = min(ti.l + tk.l, tj.l + tl.l);
                    double s = 0.0;
                    for (int L = abs(M); L <= Lmax; L++)
                        s += gaunt(ti.l, L, tk.l, ti.m, -M, tk.m) *
                             gaunt(tj.l, L, tl.l, tj.m, M, tl.m) * F[L] / (2.0 * L + 1.0);
                    // the angular sum is real for real AOs: only Re(coefficient) survives
                    sum += ((abs(M) & 1) ? -1.0 : 1.0) * cr * s;
                }
        }
    constexpr double k4pi = 12.566370614359172;
    val[t] = k4pi * sum;
    kcp[t] = zv_kcp(kbase, n, ao[0] + 1, ao[2] + 1, ao[1] + 1, ao[3] + 1);
}

} // namespace

// One-centre quartets (all four AOs on one centre) in the reference's
// canonical orientation, replaced by their exact values; their standard
// errors set to 0 (src/two_e_slater.F90:489-493, 558-563).
int zv_run_exact_monocenter(zv_context *c, double *ijkl, double *err, int64_t *n_mono)
{
    const int n = c->n;
    // AOs grouped by centre (bitwise equal coordinates: the AOs of one atom)
    std::map<std::vector<double>, std::vector<int>> groups;
    for (int i = 1; i <= n; i++) {
        std::vector<double> key(c->h_center.begin() + 3 * (i - 1),
                                c->h_center.begin() + 3 * i);
        groups[key].push_back(i);
    }
    std::vector<int> q;
    for (const auto &g : groups) {
        const std::vector<int> &ao = g.second;
        for (int i : ao) {
            const int deg = c->h_npower[3 * (i - 1)] + c->h_npower[3 * (i - 1) + 1] +
                            c->h_npower[3 * (i - 1) + 2];
            if (deg > 2)
                return zv_fail(c, ZV_ERR_ARG,
                               "exact one-centre integrals: AO %d has angular degree %d; the "
                               "reference's formulas (src/ijkl_slater_xnynzn.F90) stop above d",
                               i, deg);
        }
        std::vector<std::pair<int, int>> pairs; // (i, k), i >= k, in pair-index order
        for (int i : ao)
            for (int k : ao)
                if (k <= i) pairs.push_back({i, k});
        std::sort(pairs.begin(), pairs.end(), [](const auto &x, const auto &y) {
            return zv_pidx(x.first, x.second) < zv_pidx(y.first, y.second);
        });
        for (size_t P = 0; P < pairs.size(); P++)
            for (size_t Q = 0; Q <= P; Q++) {
                q.push_back(pairs[P].first);
                q.push_back(pairs[P].second);
                q.push_back(pairs[Q].first);
                q.push_back(pairs[Q].second);
            }
    }
    const int64_t nq = (int64_t)q.size() / 4;
    if (n_mono) *n_mono = nq;
    if (nq == 0) return ZV_OK;
    ZV_CHECK(zv_ensure(c, c->b_misc, sizeof(int) * q.size() + sizeof(double) * nq +
                                         sizeof(int64_t) * nq + 64));
    int *dq = (int *)c->b_misc.p;
    double *dval = (double *)(((uintptr_t)(dq + q.size()) + 15) & ~(uintptr_t)15);
    int64_t *dkcp = (int64_t *)(dval + nq);
    std::vector<double> val(nq);
    std::vector<int64_t> kcp(nq);
    ZV_HIPCHECK(c, hipMemcpyAsync(dq, q.data(), sizeof(int) * q.size(), hipMemcpyHostToDevice,
                                  c->stream));
    {
        PhaseTimer tm(c, PH_ERI);
        hipLaunchKernelGGL(k_mono_exact, dim3(nblk(nq, 128)), dim3(128), 0, c->stream, nq, dq,
                           c->d_npower, c->d_nsto, c->d_zeta, c->d_kbase, n, dval, dkcp);
        ZV_HIPCHECK(c, hipGetLastError());
    }
    ZV_HIPCHECK(c, hipMemcpyAsync(val.data(), dval, sizeof(double) * nq, hipMemcpyDeviceToHost,
                                  c->stream));
    ZV_HIPCHECK(c, hipMemcpyAsync(kcp.data(), dkcp, sizeof(int64_t) * nq, hipMemcpyDeviceToHost,
                                  c->stream));
    ZV_HIPCHECK(c, hipStreamSynchronize(c->stream));
    for (int64_t t = 0; t < nq; t++) {
        if (kcp[t] < 0 || kcp[t] >= c->nint)
            return zv_fail(c, ZV_ERR_STATE, "exact one-centre integrals: bad quartet index");
        ijkl[kcp[t]] = val[t];
        if (err) err[kcp[t]] = 0.0;
    }
    return ZV_OK;
}
