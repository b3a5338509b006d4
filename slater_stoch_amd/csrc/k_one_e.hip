// One-electron integrals over the Gaussian fits: overlap S, nuclear
// attraction V and kinetic K of every ordered AO pair (i, k), the loop of
// src/one_e_slater.F90:87-95 over src/one_elect.F90.
//
// Each AO is the sum over its fit terms (src/one_elect.F90:20-72: i_type 1
// one term c_gauss(:,1) with the AO's powers; i_type 2 three terms
// c_gauss(:,2) with +2 on x, y or z; i_type 3 both) of its Gaussians.  Per
// primitive pair, McMurchie-Davidson with 1D Hermite coefficients E:
//   S = e^{-mu AB^2} prod_d E^{a_d b_d}_0 sqrt(pi/p), zero when
//       mu AB^2 > -ln 1e-15 (src/one_electron_sij.F90:44-46);
//   K = -1/2 of the bra Laplacian, per dimension
//       a(a-1) S(a-2) - 2 alpha (2a+1) S(a) + 4 alpha^2 S(a+2)
//       (src/one_electron_Kij.F90, with the same cut through S);
//   V = -sum_C Z_C 2pi/p e^{-mu AB^2} sum_tuv E_t E_u E_v R_tuv(p, P-C),
//       unscreened as src/NAI_pol_mult.F90 (whose own polynomial method
//       drops coefficients below 1e-9; this evaluation keeps every term).
// One workgroup per ordered AO pair; threads stride over primitive pairs
// and fit terms.
#include <cmath>
#include <vector>

#include "eri_device.h"
#include "zvmc_internal.h"

using namespace zveri;

namespace {

constexpr int kOeA = 8;   // largest bra power + 2 (5G: 4 + 2; 2S, 3S: 2 + 2)
constexpr int kOeT = 13;  // largest t + 1 (bra power + 2 + ket power)
constexpr double kPi = 3.14159265358979323846;

// fit term ii of an AO of type t: coefficient set (1 or 2) and the
// coordinate that carries +2 (-1: none)
__device__ inline void fit_term(int type, int ii, int &set, int &shift)
{
    if (type == 1) {
        set = 1;
        shift = -1;
    } else if (type == 2) {
        set = 2;
        shift = ii;
    } else {
        set = ii == 0 ? 1 : 2;
        shift = ii - 1;
    }
}

// e[i][t] = E^{i,b}_t for i = 0..imax, without the e^{-mu X^2} factor
__device__ void hermite_1d(int imax, int b, double p, double xpa, double xpb,
                           double (&e)[kOeA + 1][kOeT])
{
    double cur[kOeT];
    for (int t = 0; t < kOeT; t++) cur[t] = 0.0;
    cur[0] = 1.0;
    const double h = 0.5 / p;
    for (int j = 0; j < b; j++) {
        double nxt[kOeT];
        for (int t = 0; t <= j + 1; t++)
            nxt[t] = (t > 0 ? h * cur[t - 1] : 0.0) + xpb * cur[t] + (t + 1) * cur[t + 1];
        for (int t = 0; t <= j + 1; t++) cur[t] = nxt[t];
    }
    for (int t = 0; t < kOeT; t++) e[0][t] = cur[t];
    for (int i = 0; i < imax; i++) {
        for (int t = 0; t < kOeT; t++) e[i + 1][t] = 0.0;
        for (int t = 0; t <= i + 1 + b && t < kOeT; t++)
            e[i + 1][t] = (t > 0 ? h * e[i][t - 1] : 0.0) + xpa * e[i][t] +
                          (t + 1 < kOeT ? (t + 1) * e[i][t + 1] : 0.0);
    }
}

struct OneEArgs {
    const double *center, *zeta, *c1, *c2, *g;
    const int *npower, *itype, *ngauss;
    int n, ldg, nnuc;
    const double *charge, *nuc; // nuc[3 c + d]
    double *S, *V, *K;          // column-major n x n
};

__global__ __launch_bounds__(256) void k_one_e(OneEArgs A)
{
    const int i = blockIdx.x % A.n, k = blockIdx.x / A.n;
    const int ti = A.itype[i], tk = A.itype[k];
    const int nci = ti == 1 ? 1 : (ti == 2 ? 3 : 4), nck = tk == 1 ? 1 : (tk == 2 ? 3 : 4);
    const int ngi = A.ngauss[i], ngk = A.ngauss[k];
    const int nitem = ngi * ngk * nci * nck;
    const double *Ai = A.center + 3 * i, *Bk = A.center + 3 * k;
    const double thr = 34.538776394910684; // -ln(1e-15)
    double s = 0.0, v = 0.0, kin = 0.0;
    for (int w = threadIdx.x; w < nitem; w += blockDim.x) {
        const int a = w % ngi, b = (w / ngi) % ngk, ii = (w / (ngi * ngk)) % nci,
                  jj = w / (ngi * ngk * nci);
        int seti, shi, setk, shk;
        fit_term(ti, ii, seti, shi);
        fit_term(tk, jj, setk, shk);
        int pa[3], pb[3];
        for (int d = 0; d < 3; d++) {
            pa[d] = A.npower[3 * i + d] + (d == shi ? 2 : 0);
            pb[d] = A.npower[3 * k + d] + (d == shk ? 2 : 0);
        }
        const double ca = (seti == 1 ? A.c1 : A.c2)[(size_t)i * A.ldg + a];
        const double cb = (setk == 1 ? A.c1 : A.c2)[(size_t)k * A.ldg + b];
        const double dd = ca * cb;
        const double al = A.g[(size_t)i * A.ldg + a], be = A.g[(size_t)k * A.ldg + b];
        const double p = al + be, mu = al * be / p;
        double P[3], ab2 = 0.0;
        for (int d = 0; d < 3; d++) {
            P[d] = (al * Ai[d] + be * Bk[d]) / p;
            ab2 += (Ai[d] - Bk[d]) * (Ai[d] - Bk[d]);
        }
        const double arg = mu * ab2, kab = exp(-arg);
        double e[3][kOeA + 1][kOeT];
        for (int d = 0; d < 3; d++) hermite_1d(pa[d] + 2, pb[d], p, P[d] - Ai[d], P[d] - Bk[d], e[d]);
        if (!(arg > thr)) {
            const double sq = sqrt(kPi / p);
            double s1[3];
            for (int d = 0; d < 3; d++) s1[d] = e[d][pa[d]][0] * sq;
            double lap = 0.0;
            for (int d = 0; d < 3; d++) {
                const int na = pa[d];
                double d2 = -2.0 * al * (2 * na + 1) * e[d][na][0] * sq +
                            4.0 * al * al * e[d][na + 2][0] * sq;
                if (na >= 2) d2 += na * (na - 1.0) * e[d][na - 2][0] * sq;
                lap += d2 * s1[(d + 1) % 3] * s1[(d + 2) % 3];
            }
            s += dd * kab * (s1[0] * s1[1] * s1[2]);
            kin += dd * kab * (-0.5 * lap);
        }
        const int tx = pa[0] + pb[0], ty = pa[1] + pb[1], tz = pa[2] + pb[2];
        const int L = tx + ty + tz;
        double F[kOeT], R[hcount(kOeT - 1)];
        for (int c = 0; c < A.nnuc; c++) {
            const double X = P[0] - A.nuc[3 * c], Y = P[1] - A.nuc[3 * c + 1],
                         Z = P[2] - A.nuc[3 * c + 2];
            boys_rt(L, p * (X * X + Y * Y + Z * Z), F);
            hermite_R_rt(L, p, X, Y, Z, F, R);
            double acc = 0.0;
            for (int t = 0; t <= tx; t++)
                for (int u = 0; u <= ty; u++)
                    for (int q = 0; q <= tz; q++)
                        acc += e[0][pa[0]][t] * e[1][pa[1]][u] * e[2][pa[2]][q] * R[hidx(t, u, q)];
            v -= dd * A.charge[c] * (2.0 * kPi / p) * kab * acc;
        }
    }
    __shared__ double red[3][256];
    red[0][threadIdx.x] = s;
    red[1][threadIdx.x] = v;
    red[2][threadIdx.x] = kin;
    __syncthreads();
    for (unsigned o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int r = 0; r < 3; r++) red[r][threadIdx.x] += red[r][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const size_t at = (size_t)i + (size_t)k * A.n;
        A.S[at] = red[0][0];
        A.V[at] = red[1][0];
        A.K[at] = red[2][0];
    }
}

} // namespace

int zv_run_one_e(zv_context *c, int nnuc, const double *charge, const double *nuc, double *S,
                 double *V, double *K)
{
    const int n = c->n;
    for (int i = 0; i < n; i++) {
        const int deg = c->h_npower[3 * i] + c->h_npower[3 * i + 1] + c->h_npower[3 * i + 2];
        if (deg + (c->h_itype[i] != 1 ? 2 : 0) > kOeA - 2)
            return zv_fail(c, ZV_ERR_ARG, "one-electron integrals: AO %d above degree 6", i + 1);
    }
    const size_t nn = (size_t)n * n;
    ZV_CHECK(zv_ensure(c, c->b_misc, sizeof(double) * (3 * nn + 4 * (size_t)nnuc + 8)));
    double *dS = (double *)c->b_misc.p, *dV = dS + nn, *dK = dV + nn, *dq = dK + nn,
           *dnuc = dq + nnuc;
    ZV_HIPCHECK(c, hipMemcpyAsync(dq, charge, sizeof(double) * nnuc, hipMemcpyHostToDevice,
                                  c->stream));
    ZV_HIPCHECK(c, hipMemcpyAsync(dnuc, nuc, sizeof(double) * 3 * nnuc, hipMemcpyHostToDevice,
                                  c->stream));
    // device fit arrays are padded to ZV_GMAX entries per AO
    OneEArgs a{c->d_center, c->d_zeta, c->d_c1, c->d_c2, c->d_g, c->d_npower, c->d_itype,
               c->d_ngauss, n, ZV_GMAX, nnuc, dq, dnuc, dS, dV, dK};
    {
        PhaseTimer tm(c, PH_ERI);
        hipLaunchKernelGGL(k_one_e, dim3((unsigned)nn), dim3(256), 0, c->stream, a);
        ZV_HIPCHECK(c, hipGetLastError());
    }
    ZV_HIPCHECK(c, hipMemcpyAsync(S, dS, sizeof(double) * nn, hipMemcpyDeviceToHost, c->stream));
    ZV_HIPCHECK(c, hipMemcpyAsync(V, dV, sizeof(double) * nn, hipMemcpyDeviceToHost, c->stream));
    ZV_HIPCHECK(c, hipMemcpyAsync(K, dK, sizeof(double) * nn, hipMemcpyDeviceToHost, c->stream));
    ZV_HIPCHECK(c, hipStreamSynchronize(c->stream));
    return ZV_OK;
}
