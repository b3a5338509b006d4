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
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <map>
#include <thread>

#include "eri_device.h"
#include "zvmc_internal.h"

using namespace zveri;

namespace {

// ------------------------------------------------------------------ pair data
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

// ------------------------------------------------------------------ classes
// Explicit quartet lists of AO pairs (bra, ket).  Q_LIST writes quartet x to
// out[x]; Q_PAIRS writes it to its triangle slot out[tri(max, min)].
enum QMode { Q_LIST = 0, Q_PAIRS = 1 };

struct QSrc {
    int mode;
    const int *qb, *qk;
    int64_t nq;
};

__device__ inline void quartet_of(const QSrc &S, int64_t x, int &Pb, int &Pk, int64_t &oidx)
{
    Pb = S.qb[x];
    Pk = S.qk[x];
    oidx = S.mode == Q_LIST ? x : (Pb > Pk ? zv_tri(Pb, Pk) : zv_tri(Pk, Pb));
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
    // wave index made provably uniform so the per-quartet data come in by scalar loads
    int64_t item = (int64_t)blockIdx.x * (blockDim.x >> 6) +
                   __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (item >= S.nq * nsplit) return;
    int64_t x = item / nsplit;
    int sp = (int)(item % nsplit);
    int Pb, Pk;
    int64_t oidx;
    quartet_of(S, x, Pb, Pk, oidx);
    double acc = 0.0;
    {
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
    // wave index made provably uniform so the per-quartet data come in by scalar loads
    int64_t item = (int64_t)blockIdx.x * (blockDim.x >> 6) +
                   __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (item >= S.nq * nsplit) return;
    int64_t x = item / nsplit;
    int sp = (int)(item % nsplit);
    int Pb, Pk;
    int64_t oidx;
    quartet_of(S, x, Pb, Pk, oidx);
    double acc = 0.0;
    {
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
// Device pointers of the primitive-pair tables inside b_pp, b_ppE, b_ppcnt.
// The Boys Taylor table of boys_table (eri_device.h), computed once per
// process in extended precision (x87 long double on the host):
// [L][k][i] = (-1)^k F_{L+k}(i / 2) / k!, L <= kBoysLmax, k <= kBoysK,
// F_m(T) = exp(-T) sum_j (2T)^j / ((2m+1)(2m+3)...(2m+2j+1)), all terms
// positive, summed until they fall below 1e-22 of the sum.
static const std::vector<double> &boys_table_host()
{
    static std::vector<double> tab;
    if (!tab.empty()) return tab;
    constexpr int NM = kBoysLmax + kBoysK + 1;
    std::vector<long double> F((size_t)NM * kBoysN);
    for (int m = 0; m < NM; m++)
        for (int i = 0; i < kBoysN; i++) {
            const long double T = 0.5L * i;
            long double term = 1.0L / (2 * m + 1), sum = term;
            for (int j = 1; j < 1000; j++) {
                term *= 2.0L * T / (2 * m + 2 * j + 1);
                sum += term;
                if (term < 1e-22L * sum) break;
            }
            F[(size_t)m * kBoysN + i] = sum * expl(-T);
        }
    std::vector<double> t((size_t)(kBoysLmax + 1) * kBoysSlice);
    for (int L = 0; L <= kBoysLmax; L++) {
        long double fact = 1.0L;
        for (int k = 0; k <= kBoysK; k++) {
            if (k > 0) fact *= k;
            for (int i = 0; i < kBoysN; i++)
                t[(size_t)L * kBoysSlice + (size_t)k * kBoysN + i] =
                    (double)(((k & 1) ? -1.0L : 1.0L) * F[(size_t)(L + k) * kBoysN + i] / fact);
        }
    }
    tab.swap(t);
    return tab;
}

// the table on the context's device (once per context)
static int ensure_boys_table(zv_context *c)
{
    const std::vector<double> &t = boys_table_host();
    ZV_CHECK(zv_ensure(c, c->b_btab, sizeof(double) * t.size()));
    if (c->btab_ready) return ZV_OK;
    ZV_HIPCHECK(c, hipMemcpyAsync(c->b_btab.p, t.data(), sizeof(double) * t.size(),
                                  hipMemcpyHostToDevice, c->stream));
    c->btab_ready = true;
    return ZV_OK;
}

static PP pp_view(zv_context *c)
{
    PP pp;
    pp.btab = (const double *)c->b_btab.p;
    const int np = c->npair;
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
    pp.off = d_off;
    pp.eoff = d_eoff;
    pp.cnt = d_cnt;
    pp.cap = d_cnt + np;
    pp.L = d_cnt + 2 * np;
    return pp;
}

static int ensure_primpairs(zv_context *c, PP &pp)
{
    int np = c->npair;
    ZV_CHECK(ensure_boys_table(c));
    pp.btab = (const double *)c->b_btab.p;
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

// FP64 flops per primitive quartet of the (LB, LK) component kernel: Boys
// values (3 per order plus ~25 for erf/exp/sqrt/division), the Hermite R
// recursion (2 FMA per entry over all levels), the ket-Hermite contraction
// (one FMA per pair of Hermite indices), the bra sum and the prefactor.
static double component_flops(int LB, int LK)
{
    int L = LB + LK;
    double HB = hcount(LB), HK = hcount(LK);
    double HR4 = (L + 1) * (L + 2) * (L + 3) * (L + 4) / 24.0; // R entries over all levels
    return 25.0 + 3.0 * (L + 1) + 4.0 * (HR4 - 1.0) + 2.0 * HB * HK + 2.0 * HB + 12.0;
}

// Launch one class over an explicit quartet list S, splitting bra primitive
// pairs when there are too few quartets to fill the device.  pq and flops
// (primitive quartets and algorithmic flops of the list) feed the statistics.
// Split of a class launch over the bra primitive pairs: short lists are
// spread over up to 32 waves per quartet, about 16384 waves in all.  The
// partial sums then take nq * nsplit < 16384 + nq doubles.
static int class_split(int64_t nq)
{
    const int64_t target_waves = 16384;
    if (nq >= target_waves) return 1;
    return (int)std::min<int64_t>(32, (target_waves + nq - 1) / nq);
}

// The kernels of one class: no allocation, no timing.
static int launch_class_raw(zv_context *c, const PP &pp, QSrc S, int LB, int LK, double *out,
                            double *part, int nsplit)
{
    int64_t items = S.nq * nsplit;
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
    return ZV_OK;
}

static int launch_class(zv_context *c, const PP &pp, QSrc S, int LB, int LK, double *out,
                        double pq, double flops)
{
    if (S.nq <= 0) return ZV_OK;
    const int nsplit = class_split(S.nq);
    double *part = nullptr;
    if (nsplit > 1) {
        ZV_CHECK(zv_ensure(c, c->b_part, sizeof(double) * S.nq * nsplit));
        part = (double *)c->b_part.p;
    }
    const int cls = LB * 6 + LK;
    c->cls_pq[cls] += pq;
    c->cls_flops[cls] += flops;
    PhaseTimer tm(c, PH_ERI, cls);
    return launch_class_raw(c, pp, S, LB, LK, out, part, nsplit);
}

// Component quartets (bra, ket) of pairs beyond the templated orders, written
// to their triangle slots by the runtime-order kernel.
static int run_generic_pairs(zv_context *c, const PP &pp, const std::vector<int> &qb,
                             const std::vector<int> &qk, double *tri)
{
    const int64_t nq = (int64_t)qb.size();
    if (!nq) return ZV_OK;
    double pq = 0.0, flops = 0.0;
    for (int64_t t = 0; t < nq; t++) {
        double v = (double)c->h_cnt[qb[t]] * c->h_cnt[qk[t]];
        pq += v;
        flops += v * component_flops(c->h_pair_L[qb[t]], c->h_pair_L[qk[t]]);
    }
    c->h_plist.resize(2 * nq);
    std::copy(qb.begin(), qb.end(), c->h_plist.begin());
    std::copy(qk.begin(), qk.end(), c->h_plist.begin() + nq);
    ZV_CHECK(zv_ensure(c, c->b_plist, sizeof(int) * 2 * nq));
    int *d = (int *)c->b_plist.p;
    ZV_HIPCHECK(c, hipMemcpyAsync(d, c->h_plist.data(), sizeof(int) * 2 * nq,
                                  hipMemcpyHostToDevice, c->stream));
    QSrc S{Q_PAIRS, d, d + nq, nq};
    return launch_class(c, pp, S, ZV_LMAX_T + 1, ZV_LMAX_T + 1, tri, pq, flops);
}

// The reference's Schwarz driver (src/two_e_slater.F90:230-273) over shell
// quartets.  Phase 1: the diagonal shell quartets (x|x), which hold every
// (P|P) = precond(i,k).  Phase 2: the shell quartets (x|y), x > y, that hold
// at least one component quartet passing |precond_P precond_Q| >= 1e-30;
// the test itself is applied per component quartet when the results are
// written.  Shell pairs of order above ZV_LMAX_T go through explicit
// component lists.
int zv_run_gauss_eri(zv_context *c, double *ijkl, double *precond, int64_t *n_screened)
{
    PP pp;
    ZV_CHECK(ensure_primpairs(c, pp));
    ZV_CHECK(zv_shell_ready(c, pp));
    const int np = c->npair, nsp = c->nsp;
    ZV_CHECK(zv_ensure(c, c->b_tri, sizeof(double) * c->nint));
    ZV_CHECK(zv_ensure(c, c->b_misc, sizeof(double) * np));
    double *tri = (double *)c->b_tri.p, *pre = (double *)c->b_misc.p;
    ZV_CHECK(zv_zero(c, tri, c->nint));
    auto templ = [&](int s) { return c->h_sp_L[s] <= ZV_LMAX_T; };
    auto comps = [&](int s) {
        return std::make_pair(c->h_sp_comp.data() + c->h_sp_coff[s],
                              c->h_sp_comp.data() + c->h_sp_coff[s + 1]);
    };
    // phase 1
    std::vector<std::pair<int, int>> sq;
    std::vector<int> gb, gk;
    for (int s = 0; s < nsp; s++) {
        if (!c->h_sp_np[s]) continue;
        if (templ(s)) {
            sq.push_back({s, s});
        } else {
            auto r = comps(s);
            for (const int *P = r.first; P != r.second; P++) {
                gb.push_back(*P);
                gk.push_back(*P);
            }
        }
    }
    int64_t ng1 = 0, ng2 = 0;
    ZV_CHECK(zv_shell_run(c, pp, sq, c->b_sitem1, c->b_sgroup1, c->b_part1, c->h_plan1, &ng1));
    ZV_CHECK(zv_shell_scatter(c, c->b_sgroup1, c->b_part1, ng1, nullptr, tri));
    ZV_CHECK(run_generic_pairs(c, pp, gb, gk, tri));
    hipLaunchKernelGGL(k_diag_pre, dim3(nblk(np, 256)), dim3(256), 0, c->stream, np, tri, pre);
    ZV_HIPCHECK(c, hipGetLastError());
    std::vector<double> hpre(np);
    ZV_HIPCHECK(c, hipMemcpyAsync(hpre.data(), pre, sizeof(double) * np, hipMemcpyDeviceToHost,
                                  c->stream));
    ZV_HIPCHECK(c, hipStreamSynchronize(c->stream));
    // phase 2
    const auto tp2 = std::chrono::steady_clock::now();
    std::vector<double> spmax(nsp, 0.0);
    for (int s = 0; s < nsp; s++) {
        auto r = comps(s);
        for (const int *P = r.first; P != r.second; P++)
            spmax[s] = std::max(spmax[s], fabs(hpre[*P]));
    }
    // the lists of shell pairs x in [x0, x1), on host threads over ranges of
    // x of equal work (~x each), joined in order: the same lists as one pass
    struct Lists {
        std::vector<std::pair<int, int>> sq;
        std::vector<int> gb, gk;
    };
    auto build = [&](int x0, int x1, Lists &o) {
        for (int x = x0; x < x1; x++) {
            if (!c->h_sp_np[x]) continue;
            auto rx = comps(x);
            if (!templ(x))
                for (const int *P = rx.first; P != rx.second; P++)
                    for (const int *Q = rx.first; Q != rx.second; Q++)
                        if (*P > *Q && fabs(hpre[*P] * hpre[*Q]) >= 1.0e-30) {
                            o.gb.push_back(*P);
                            o.gk.push_back(*Q);
                        }
            for (int y = 0; y < x; y++) {
                if (!c->h_sp_np[y] || spmax[x] * spmax[y] < 1.0e-30) continue;
                if (templ(x) && templ(y)) {
                    o.sq.push_back({x, y});
                    continue;
                }
                auto ry = comps(y);
                for (const int *P = rx.first; P != rx.second; P++)
                    for (const int *Q = ry.first; Q != ry.second; Q++)
                        if (fabs(hpre[*P] * hpre[*Q]) >= 1.0e-30) {
                            o.gb.push_back(*P);
                            o.gk.push_back(*Q);
                        }
            }
        }
    };
    const unsigned hw = std::thread::hardware_concurrency();
    const int nt = nsp < 512 ? 1 : (int)std::max(1u, std::min(16u, hw ? hw : 1u));
    std::vector<Lists> lists(nt);
    std::vector<int> xb(nt + 1);
    for (int t = 0; t <= nt; t++) xb[t] = (int)std::lround(nsp * std::sqrt((double)t / nt));
    xb[nt] = nsp;
    {
        std::vector<std::thread> th;
        int started = 1;
        for (int t = 1; t < nt; t++, started++) {
            try {
                th.emplace_back(build, xb[t], xb[t + 1], std::ref(lists[t]));
            } catch (...) {
                break;
            }
        }
        for (int t = started; t < nt; t++) build(xb[t], xb[t + 1], lists[t]);
        build(xb[0], xb[1], lists[0]);
        for (auto &x : th) x.join();
    }
    sq.clear();
    gb.clear();
    gk.clear();
    for (const Lists &o : lists) {
        sq.insert(sq.end(), o.sq.begin(), o.sq.end());
        gb.insert(gb.end(), o.gb.begin(), o.gb.end());
        gk.insert(gk.end(), o.gk.begin(), o.gk.end());
    }
    if (c->verbose)
        fprintf(stderr, "[zvmc] Schwarz phase 2 list: %zu shell quartets, %zu component quartets, "
                "%.1f ms\n", sq.size(), gb.size(),
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tp2)
                    .count());
    ZV_CHECK(zv_shell_run(c, pp, sq, c->b_sitem2, c->b_sgroup2, c->b_part2, c->h_plan2, &ng2));
    ZV_CHECK(zv_shell_scatter(c, c->b_sgroup2, c->b_part2, ng2, pre, tri));
    // the off-diagonal components of (x|x) get the test now that precond is known
    ZV_CHECK(zv_shell_scatter(c, c->b_sgroup1, c->b_part1, ng1, pre, tri));
    ZV_CHECK(run_generic_pairs(c, pp, gb, gk, tri));
    ZV_CHECK(zv_keep_result(c, tri, ZV_RESULT_ERI, ijkl));
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
    std::vector<double> bpq(ncls * ncls, 0.0), bfl(ncls * ncls, 0.0);
    for (int b = 0; b < ncls * ncls; b++) {
        lo[b] = (int64_t)perm.size();
        for (int64_t t : bucket[b]) {
            perm.push_back(t);
            qb.push_back(pb[t]);
            qk2.push_back(pk[t]);
            double v = (double)c->h_cnt[pb[t]] * c->h_cnt[pk[t]];
            bpq[b] += v;
            bfl[b] += v * component_flops(c->h_pair_L[pb[t]], c->h_pair_L[pk[t]]);
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
        QSrc S{Q_LIST, dqb + lo[b], dqk + lo[b], cnt};
        ZV_CHECK(launch_class(c, pp, S, b / ncls, b % ncls, dout + lo[b], bpq[b], bfl[b]));
    }
    std::vector<double> h(nq);
    ZV_HIPCHECK(c, hipMemcpyAsync(h.data(), dout, sizeof(double) * nq, hipMemcpyDeviceToHost,
                                  c->stream));
    ZV_HIPCHECK(c, hipStreamSynchronize(c->stream));
    for (int64_t t = 0; t < nq; t++) out[perm[t]] = h[t];
    return ZV_OK;
}

// ---- allocation-free list mode (zv_gauss_ijkl_list_noalloc) ---------------
// The device memory of the allocation-free entry point: a file-scope array,
// one copy per device that loads the library, held until the process ends.
// Layout for up to kNaQ quartets: bra pair, ket pair and list position of
// each quartet in class order (int), the values in class order and in list
// order (double), then the split partial sums of one class launch (fewer
// than kNaPart doubles, see class_split).  Every entry that is read is
// written earlier in the same call, so nothing needs clearing.
constexpr size_t kNaQ = ZV_NOALLOC_MAXQ;
constexpr size_t kNaPart = 32768;
constexpr size_t kNaBytes = 3 * kNaQ * sizeof(int) + (2 * kNaQ + kNaPart) * sizeof(double);
static_assert((3 * kNaQ * sizeof(int)) % sizeof(double) == 0, "double alignment");
static __device__ __align__(256) char scratch[kNaBytes];

// Device address of `scratch` per device, from hipGetSymbolAddress at the
// first context on that device (guarded by the library's call lock).  On the
// host the array's name is not a device address; only these are used.
static std::map<int, char *> g_scratch_ptr;

int zv_noalloc_setup(zv_context *c)
{
    if (g_scratch_ptr.count(c->device)) return ZV_OK;
    char *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return zv_hip(c, e, "hipGetSymbolAddress(scratch)");
    if (!ptr) return zv_fail(c, ZV_ERR_HIP, "hipGetSymbolAddress(scratch) returned null");
    g_scratch_ptr[c->device] = ptr;
    return ZV_OK;
}

// list position -> class order, written back in list order
__global__ void k_unpermute(int64_t nq, const int *__restrict__ pos,
                            const double *__restrict__ in, double *__restrict__ out)
{
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x < nq) out[pos[x]] = in[x];
}

// zv_run_gauss_list without allocation: the same buckets, kernels and
// summation order (so the same values, bit for bit), with the device work
// in `scratch` and the host bookkeeping in static arrays (calls run one at
// a time under the library's call lock).  Needs the primitive-pair tables
// of the current basis, which the first zv_gauss_eri or zv_gauss_ijkl_list
// after zv_set_basis builds.
int zv_run_gauss_list_noalloc(zv_context *c, int64_t nq, const int *qi, const int *qk,
                              const int *qj, const int *ql, double *out)
{
    auto sp = g_scratch_ptr.find(c->device);
    if (sp == g_scratch_ptr.end())
        return zv_fail(c, ZV_ERR_STATE, "allocation-free scratch is not set up on device %d",
                       c->device);
    if (nq < 0 || nq > (int64_t)kNaQ)
        return zv_fail(c, ZV_ERR_ARG, "%lld quartets: the allocation-free list holds at most %lld",
                       (long long)nq, (long long)kNaQ);
    if (!c->pp_ready)
        return zv_fail(c, ZV_ERR_STATE,
                       "the primitive-pair tables of this basis are not built: call zv_gauss_eri "
                       "or zv_gauss_ijkl_list once after zv_set_basis");
    const PP pp = pp_view(c);
    constexpr int ncls = ZV_LMAX_T + 2;
    static int h_qb[kNaQ], h_qk[kNaQ], h_pos[kNaQ];
    static unsigned char h_bk[kNaQ];
    int64_t lo[ncls * ncls + 1] = {};
    auto order = [&](int P) {
        const int L = c->h_pair_L[P];
        return L > ZV_LMAX_T ? ZV_LMAX_T + 1 : L;
    };
    for (int64_t t = 0; t < nq; t++) {
        const int i = std::max(qi[t], qk[t]), k = std::min(qi[t], qk[t]);
        const int j = std::max(qj[t], ql[t]), l = std::min(qj[t], ql[t]);
        int P = zv_pidx(i, k), Q = zv_pidx(j, l);
        if (order(P) < order(Q)) std::swap(P, Q);
        h_bk[t] = (unsigned char)(order(P) * ncls + order(Q));
        lo[h_bk[t] + 1]++;
    }
    for (int b = 0; b < ncls * ncls; b++) lo[b + 1] += lo[b];
    {
        int64_t at[ncls * ncls];
        for (int b = 0; b < ncls * ncls; b++) at[b] = lo[b];
        // stable within a bucket: the order of zv_run_gauss_list
        for (int64_t t = 0; t < nq; t++) {
            const int i = std::max(qi[t], qk[t]), k = std::min(qi[t], qk[t]);
            const int j = std::max(qj[t], ql[t]), l = std::min(qj[t], ql[t]);
            int P = zv_pidx(i, k), Q = zv_pidx(j, l);
            if (order(P) < order(Q)) std::swap(P, Q);
            const int64_t x = at[h_bk[t]]++;
            h_qb[x] = P;
            h_qk[x] = Q;
            h_pos[x] = (int)t;
        }
    }
    char *base = sp->second;
    int *dqb = (int *)base, *dqk = dqb + kNaQ, *dpos = dqk + kNaQ;
    double *dsorted = (double *)(dpos + kNaQ), *dout = dsorted + kNaQ, *dpart = dout + kNaQ;
    ZV_HIPCHECK(c, hipMemcpyAsync(dqb, h_qb, sizeof(int) * nq, hipMemcpyHostToDevice, c->stream));
    ZV_HIPCHECK(c, hipMemcpyAsync(dqk, h_qk, sizeof(int) * nq, hipMemcpyHostToDevice, c->stream));
    ZV_HIPCHECK(c, hipMemcpyAsync(dpos, h_pos, sizeof(int) * nq, hipMemcpyHostToDevice,
                                  c->stream));
    for (int b = 0; b < ncls * ncls; b++) {
        const int64_t cnt = lo[b + 1] - lo[b];
        if (!cnt) continue;
        const int nsplit = class_split(cnt);
        if (nsplit > 1 && (size_t)(cnt * nsplit) > kNaPart)
            return zv_fail(c, ZV_ERR_STATE, "split partials exceed the scratch");
        QSrc S{Q_LIST, dqb + lo[b], dqk + lo[b], cnt};
        ZV_CHECK(launch_class_raw(c, pp, S, b / ncls, b % ncls, dsorted + lo[b],
                                  nsplit > 1 ? dpart : nullptr, nsplit));
    }
    hipLaunchKernelGGL(k_unpermute, dim3(nblk(nq, 256)), dim3(256), 0, c->stream, nq,
                       (const int *)dpos, (const double *)dsorted, dout);
    ZV_HIPCHECK(c, hipGetLastError());
    ZV_HIPCHECK(c, hipMemcpyAsync(out, dout, sizeof(double) * nq, hipMemcpyDeviceToHost,
                                  c->stream));
    ZV_HIPCHECK(c, hipStreamSynchronize(c->stream));
    return ZV_OK;
}
