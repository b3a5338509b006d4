// Shell-blocked analytic ERIs: the reference's gauss_ijkl
// (src/gauss_ijkl.F90:51-154) for all Cartesian components of a shell
// quartet at once, inside its Schwarz driver (src/two_e_slater.F90:230-273).
//
// The AOs of one shell (e.g. the six 3D functions of an atom) share center,
// Gaussian exponents and contraction coefficients, so all AO pairs of a
// shell pair share the primitive-pair data p, P, c_ab and the reference's
// cuts; only the Hermite coefficients E differ.  For a shell quartet (U|L)
// one primitive quartet (a, c) needs one Boys evaluation and one R table:
//
//   (u|l) += c_a c_c pref(a,c) sum_hu E^u_hu(a) X^l_hu(a,c),
//   X^l_hu(a,c) = sum_hl (-1)^(t+u+v of hl) E^l_hl(c) R_{hu+hl}(a,c)
//
// for every component u of U and l of L.
//
// Work item of one wavefront: a shell quartet, a range of lane-side
// components, a range of uniform-side components (up to 64 in the reduction
// mode, at most kDirectMaxU in the direct mode) and a slice of the
// uniform-side primitive pairs.  Lanes own lane-side primitive pairs (their
// E in registers, reloaded per lane component); the uniform primitive pairs
// are walked in the inner loop through an AoS copy ([a][component][h]) read
// with scalar loads, so each uniform E value is one SGPR operand shared by
// the 64 lanes.  Per-lane sums over the uniform components live in LDS;
// each item writes its partial sums, and one scatter kernel adds the slices,
// applies the Schwarz test per component quartet and writes triangle order.
//
// E^l of one lane component is nonzero only on a pattern of Hermite indices
// fixed by the AOs' Cartesian powers (HermPatterns); the classes listed in
// kShellSparse read the pattern's rows of a compacted copy of E and run the
// X contraction over them only, the same terms in the same order minus
// exact zeros.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <thread>
#include <type_traits>
#include <vector>

#include "eri_device.h"
#include "zvmc_internal.h"

using namespace zveri;

namespace {

struct SPD {
    const int *comp;      // AO pair index of each component
    const int *coff;      // first component of each shell pair (nsp + 1)
    const int64_t *puoff; // AoS primitive data offset (doubles) per shell pair
    const int64_t *euoff; // AoS E offset (doubles) per shell pair
    const double *pu;     // [a][8]: p, 1/p, Px, Py, Pz, c, 0, 0
    const double *eu;     // [a][component][h]  (scalar reads, direct mode)
    const double *eu2;    // [a][h][component]  (lane-per-component reads, reduction mode)
    const int *cpat;      // per component: Hermite pattern index | nonzero count << 8
    const double *ec;     // lane-side E: the pattern's nonzero rows, offsets and stride of E
};

struct ShellItem {
    int su, sl;   // shell pairs on the uniform and the lane side
    int a0, a1;   // uniform primitive-pair slice
    int cl0, cl1; // lane-side components
    int cu0, nu;  // uniform-side components cu0 .. cu0 + nu - 1
    int64_t part; // partial sums of this item: [nu][cl1 - cl0]
};

struct ShellGroup {
    int su, sl, cu0, nu, cl0, nl, ns, pad; // ns slices of [nu][nl] partials from part on
    int64_t part;
};

// Uniform components of a direct-mode item: the planner sends ranges of
// kDirectMaxU + 1 or more uniform components to the reduction mode, so a
// direct item holds a whole shell pair of at most this many components or
// the short tail of a longer one (checked on the host before each launch).
// The direct kernel's LDS accumulators are sized for it.
constexpr int kDirectMaxU = 3;

// AoS copies of a shell pair's primitive data and E (one block per shell pair)
__global__ __launch_bounds__(256) void k_shellpairs(PP pp, SPD sd, double *pu, double *eu,
                                                    double *eu2)
{
    const int s = blockIdx.x;
    const int c0 = sd.coff[s], nc = sd.coff[s + 1] - c0;
    const int rep = sd.comp[c0];
    const int np = pp.cnt[rep];
    const int H = hcount(pp.L[rep]);
    const int64_t off = pp.off[rep];
    double *dp = pu + sd.puoff[s];
    for (int t = threadIdx.x; t < np * 8; t += blockDim.x) {
        const int64_t at = off + (t >> 3);
        double v = 0.0;
        switch (t & 7) {
        case 0: v = pp.p[at]; break;
        case 1: v = pp.ip[at]; break;
        case 2: v = pp.Px[at]; break;
        case 3: v = pp.Py[at]; break;
        case 4: v = pp.Pz[at]; break;
        case 5: v = pp.c[at]; break;
        default: break;
        }
        dp[t] = v;
    }
    double *de = eu + sd.euoff[s], *de2 = eu2 + sd.euoff[s];
    const int64_t tot = (int64_t)np * nc * H;
    for (int64_t t = threadIdx.x; t < tot; t += blockDim.x) {
        const int h = (int)(t % H);
        const int64_t r = t / H;
        const int cu = (int)(r % nc);
        const int a = (int)(r / nc);
        const int P = sd.comp[c0 + cu];
        const double v = pp.E[pp.eoff[P] + (int64_t)h * pp.cap[P] + a];
        de[t] = v;
        de2[((int64_t)a * H + h) * nc + cu] = v;
    }
}

// Lane-side E without its structural zeros: per component (AO pair), the
// rows of the nonzero Hermite indices of its pattern, in ascending order, at
// the pair's E offset with the same row stride (one block per component).
__global__ __launch_bounds__(256) void k_compact_e(PP pp, SPD sd, double *ec)
{
    const int q = blockIdx.x;
    const int P = sd.comp[q];
    const int L = pp.L[P];
    if (L > ZV_LMAX_T) return;
    const int pid = sd.cpat[q] & 0xff, nz = sd.cpat[q] >> 8;
    const int n = pp.cnt[P], cap = pp.cap[P];
    const int64_t eo = pp.eoff[P];
    for (int t = threadIdx.x; t < nz * n; t += blockDim.x) {
        const int k = t / n, a = t - k * n;
        ec[eo + (int64_t)k * cap + a] = pp.E[eo + (int64_t)kHP.h[L][pid][k] * cap + a];
    }
}

// Calls f(std::integral_constant<int, p>) for the runtime pattern index p of
// order LL (a chain of uniform compares), so the body sees the pattern's
// nonzero Hermite indices as constants.
template <int LL, int I = 0, class F> __device__ __forceinline__ void pat_dispatch(int p, F &&f)
{
    if constexpr (I < kHP.n[LL]) {
        if (p == I)
            f(std::integral_constant<int, I>{});
        else
            pat_dispatch<LL, I + 1>(p, f);
    }
}

// X[hu] over one lane component: the sum over the nonzero Hermite indices
// hl of its pattern pat, in ascending order, of (-1)^(t+u+v of hl) sc(E[k])
// R_{hu+hl}, E[k] the k-th nonzero coefficient; with INIT the first term
// sets X (hl = 0, even), else every term is added.  The skipped terms are
// exact zeros, so X equals the sum over all indices.
template <int LU, int LL, bool INIT, int NZ, class Sc>
__device__ __forceinline__ void lane_contract(int pat, const double (&E)[NZ], const double *R,
                                              double (&X)[hcount(LU)], Sc sc)
{
    constexpr int HU = hcount(LU);
    auto term = [&](int k, int hl) {
        const double ek = sc(E[k]);
        const double e = ((kH.t[hl] + kH.u[hl] + kH.v[hl]) & 1) ? -ek : ek;
#pragma unroll
        for (int hu = 0; hu < HU; hu++) {
            const double r =
                R[hidx(kH.t[hu] + kH.t[hl], kH.u[hu] + kH.u[hl], kH.v[hu] + kH.v[hl])];
            X[hu] = INIT && k == 0 ? e * r : fma(e, r, X[hu]);
        }
    };
    pat_dispatch<LL>(pat, [&](auto PI) {
        constexpr int P = decltype(PI)::value;
        static_assert(kHP.h[LL][P][0] == 0, "pattern order");
#pragma unroll
        for (int k = 0; k < kHP.nz[LL][P]; k++) term(k, kHP.h[LL][P][k]);
    });
}

// The uniform-side arrays and the output are separate restrict arguments so
// the compiler may read the former through the scalar cache (no store of
// the kernel can reach them).  CLB lane-side components are processed per
// pass: they share the Boys values and the R table of a primitive quartet.

// Direct mode (few uniform components): every lane contracts its X with the
// E of every uniform component, scalar E operands, nu * HU FMAs per lane and
// primitive quartet; per-lane sums over the uniform components live in LDS.
template <int LU, int LL, int CLB, bool PFU, bool TB, bool SP>
__global__ __launch_bounds__(64) void k_eri_shell(PP pp, SPD sd,
                                                  const ShellItem *__restrict__ items,
                                                  const double *__restrict__ pu_all,
                                                  const double *__restrict__ eu_all,
                                                  double *__restrict__ part)
{
    constexpr int HU = hcount(LU), L = LU + LL, HR = hcount(L);
    constexpr int NZ = SP ? kHP.nzmax[LL] : hcount(LL); // lane E per component
    __shared__ double accs[kDirectMaxU * CLB * 64]; // it.nu <= kDirectMaxU (host-checked)
    __shared__ double lbt[TB ? kBoysSlice : 1];     // Boys Taylor table, this L
    if constexpr (TB) boys_table_load<L>(pp.btab, lbt);
    const int lane = threadIdx.x;
    int vzero = 0;
    if constexpr (PFU) asm volatile("v_mov_b32 %0, 0" : "=v"(vzero));
    const ShellItem it = items[blockIdx.x];
    const int ub = sd.coff[it.su], lb = sd.coff[it.sl];
    const int ncu = sd.coff[it.su + 1] - ub;
    const int repl = sd.comp[lb];
    const int nl = pp.cnt[repl];
    const int64_t ol = pp.off[repl];
    const double *pu = pu_all + sd.puoff[it.su];
    const double *eu = eu_all + sd.euoff[it.su] + (int64_t)it.cu0 * HU;
    const int64_t erow = (int64_t)ncu * HU;
    const int nlc = it.cl1 - it.cl0;
    // the item's lane components are whole passes of CLB (host-checked)
    for (int cl = it.cl0; cl < it.cl1; cl += CLB) {
        int64_t el[CLB];
        int capl[CLB], pat[CLB], nzl[CLB];
#pragma unroll
        for (int j = 0; j < CLB; j++) {
            const int Pl = sd.comp[lb + cl + j];
            el[j] = pp.eoff[Pl];
            capl[j] = pp.cap[Pl];
            const int cp = sd.cpat[lb + cl + j];
            pat[j] = cp & 0xff;
            nzl[j] = cp >> 8;
        }
        for (int u = 0; u < it.nu * CLB; u++) accs[u * 64 + lane] = 0.0;
        for (int c0 = 0; c0 < nl; c0 += 64) {
            const int c = c0 + lane;
            const bool live = c < nl;
            const int cc = live ? c : c0; // in-range index for idle lanes
            const double q = pp.p[ol + cc], iq = pp.ip[ol + cc];
            const double Qx = pp.Px[ol + cc], Qy = pp.Py[ol + cc], Qz = pp.Pz[ol + cc];
            const double ck = live ? pp.c[ol + cc] : 0.0;
            // the lane E of each component (SP: its nonzero ones, pattern order)
            double El[CLB][NZ];
#pragma unroll
            for (int j = 0; j < CLB; j++)
#pragma unroll
                for (int k = 0; k < NZ; k++) {
                    if constexpr (SP) {
                        El[j][k] = k < nzl[j] ? sd.ec[el[j] + (int64_t)k * capl[j] + cc] : 0.0;
                    } else {
                        const double e = pp.E[el[j] + (int64_t)k * capl[j] + cc];
                        El[j][k] = ((kH.t[k] + kH.u[k] + kH.v[k]) & 1) ? -e : e;
                    }
                }
            // PFU: the uniform primitive pair's data (p, 1/p, P, c and the E of
            // the item's uniform components) are loaded into vector registers
            // one iteration ahead instead of through the scalar cache, whose
            // latency each iteration otherwise waits for.  The opaque zero
            // makes their addresses lane-dependent for the compiler.
            constexpr int NUE = PFU ? kDirectMaxU * HU : 1;
            double nU[6], nEu[NUE];
            auto load_u = [&](int a) {
                const double *pa = pu + (int64_t)a * 8 + vzero;
#pragma unroll
                for (int t = 0; t < 6; t++) nU[t] = pa[t];
                const double *ea = eu + (int64_t)a * erow + vzero;
#pragma unroll
                for (int t = 0; t < NUE; t++) nEu[t] = t < it.nu * HU ? ea[t] : 0.0;
            };
            if constexpr (PFU) load_u(it.a0);
            for (int a = it.a0; a < it.a1; a++) {
                double U[6], Eu[NUE];
                if constexpr (PFU) {
#pragma unroll
                    for (int t = 0; t < 6; t++) U[t] = nU[t];
#pragma unroll
                    for (int t = 0; t < NUE; t++) Eu[t] = nEu[t];
                    if (a + 1 < it.a1) load_u(a + 1);
                } else {
                    const double *pa = pu + (int64_t)a * 8;
#pragma unroll
                    for (int t = 0; t < 6; t++) U[t] = pa[t];
                }
                const double p = U[0], ip = U[1];
                const double X = U[2] - Qx, Y = U[3] - Qy, Z = U[4] - Qz;
                const double s = p + q;
                const double rs = rsqrt_pos(s), is = rs * rs;
                const double alpha = (p * q) * is;
                const double T = alpha * ((X * X + Y * Y) + Z * Z);
                double F[L + 1];
                if constexpr (TB)
                    boys_table<L>(T, lbt, F);
                else
                    boys<L>(T, F);
                double R[HR];
                hermite_R<L>(alpha, X, Y, Z, F, R);
                const double f = (U[5] * ck) * (((kPi52 * ip) * iq) * rs);
                double Xv[CLB][HU];
                if constexpr (SP) {
                    // HU independent FMA chains per lane component
#pragma unroll
                    for (int j = 0; j < CLB; j++)
                        lane_contract<LU, LL, true>(pat[j], El[j], R, Xv[j],
                                                    [](double e) { return e; });
                } else {
                    // hl outer, hu inner: CLB * HU independent FMA chains
#pragma unroll
                    for (int j = 0; j < CLB; j++)
#pragma unroll
                        for (int hu = 0; hu < HU; hu++) Xv[j][hu] = El[j][0] * R[hu];
#pragma unroll
                    for (int hl = 1; hl < NZ; hl++)
#pragma unroll
                        for (int hu = 0; hu < HU; hu++)
#pragma unroll
                            for (int j = 0; j < CLB; j++)
                                Xv[j][hu] = fma(El[j][hl],
                                                R[hidx(kH.t[hu] + kH.t[hl], kH.u[hu] + kH.u[hl],
                                                       kH.v[hu] + kH.v[hl])],
                                                Xv[j][hu]);
                }
#pragma unroll
                for (int j = 0; j < CLB; j++)
#pragma unroll
                    for (int hu = 0; hu < HU; hu++) Xv[j][hu] *= f;
                const double *ea = eu + (int64_t)a * erow;
#pragma unroll
                for (int u = 0; u < kDirectMaxU; u++) {
                    if (u >= it.nu) break;
                    // independent partial sums keep the FMA chains short
                    constexpr int NP = CLB == 1 ? 8 : CLB == 2 ? 4 : 2;
                    double y[CLB][NP];
#pragma unroll
                    for (int j = 0; j < CLB; j++)
#pragma unroll
                        for (int t = 0; t < NP; t++) y[j][t] = 0.0;
#pragma unroll
                    for (int hu = 0; hu < HU; hu++) {
                        const double e = PFU ? Eu[(u * HU + hu) % NUE] : ea[u * HU + hu];
#pragma unroll
                        for (int j = 0; j < CLB; j++)
                            y[j][hu % NP] = fma(e, Xv[j][hu], y[j][hu % NP]);
                    }
#pragma unroll
                    for (int j = 0; j < CLB; j++) {
                        double t = y[j][0];
#pragma unroll
                        for (int r = 1; r < NP; r++) t += y[j][r];
                        accs[(u * CLB + j) * 64 + lane] += t;
                    }
                }
            }
        }
        for (int u = 0; u < it.nu; u++)
#pragma unroll
            for (int j = 0; j < CLB; j++) {
                const double v = wave_sum(accs[(u * CLB + j) * 64 + lane]);
                if (lane == 0) part[it.part + (int64_t)u * nlc + (cl + j - it.cl0)] = v;
            }
    }
}

constexpr int pow2ceil(int x)
{
    int p = 1;
    while (p < x) p <<= 1;
    return p;
}

// Reduction-mode batch size per class compiled for at least two waves per
// SIMD (the register allocator then fits 256 registers, spilling if it
// must); 0: none.  (2, 1) with sparse passes of 3 needed it while its Boys
// function was the erf form (260 registers otherwise; 554 against 908 ms);
// with the table (two reduction rounds, shell_red_rounds) it takes 202
// registers by itself, and no class uses it now.
constexpr int kShellTwoWaves[5][5] = {{0, 0, 0, 0, 0},
                                      {0, 0, 0, 0, 0},
                                      {0, 0, 0, 0, 0},
                                      {0, 0, 0, 0, 0},
                                      {0, 0, 0, 0, 0}};
constexpr int shell_waves(int LU, int LL, int B) { return kShellTwoWaves[LU][LL] == B ? 2 : 1; }

// Reduction mode (nu >= 4): the uniform E does not depend on the lane
// primitive pair, so for each uniform primitive pair a the wave first forms
// X_hu = sum over all lane primitive pairs (every lane chunk, the prefactor
// folded into the lane E), sums it over the 64 lanes through LDS, and lane u
// then contracts the sum with the E of its own uniform component u.  One LDS
// reduction per (a, pass).
// PF: lane data of the next chunk loaded one chunk ahead (kShellPrefetch)
template <int LU, int LL, int CLB, bool PF, bool TB, bool SP, int NCH>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(shell_waves(LU, LL, CLB))))
void k_eri_shell_red(PP pp, SPD sd,
                                                      const ShellItem *__restrict__ items,
                                                      const double *__restrict__ pu_all,
                                                      const double *__restrict__ eu2_all,
                                                      double *__restrict__ part)
{
    constexpr int HU = hcount(LU), L = LU + LL, HR = hcount(L);
    constexpr int NZ = SP ? kHP.nzmax[LL] : hcount(LL); // lane E per component
    constexpr int NR = CLB * HU; // rows summed over the wave
    constexpr int RC = NR / NCH; // rows per transpose round
    static_assert(NR % NCH == 0, "whole transpose rounds");
    constexpr int S = RC | 1;    // odd lane stride of the transpose buffer
    __shared__ double lds[64 * S + NR];
    __shared__ double lbt[TB ? kBoysSlice : 1]; // Boys Taylor table, this L
    if constexpr (TB) boys_table_load<L>(pp.btab, lbt);
    const int lane = threadIdx.x;
    const ShellItem it = items[blockIdx.x];
    const int ub = sd.coff[it.su], lb = sd.coff[it.sl];
    const int ncu = sd.coff[it.su + 1] - ub;
    const int repl = sd.comp[lb];
    const int nl = pp.cnt[repl];
    const int64_t ol = pp.off[repl];
    const double *pu = pu_all + sd.puoff[it.su];
    const int nlc = it.cl1 - it.cl0;
    const bool own = lane < it.nu; // lane owns uniform component cu0 + lane
    // lanes past the uniform range read component cu0 (in range) and drop it
    const double *eu2 = eu2_all + sd.euoff[it.su] + it.cu0 + (own ? lane : 0);
    // the item's lane components are whole passes of CLB (host-checked)
    for (int cl = it.cl0; cl < it.cl1; cl += CLB) {
        int64_t el[CLB];
        int capl[CLB], pat[CLB], nzl[CLB];
#pragma unroll
        for (int j = 0; j < CLB; j++) {
            const int Pl = sd.comp[lb + cl + j];
            el[j] = pp.eoff[Pl];
            capl[j] = pp.cap[Pl];
            const int cp = sd.cpat[lb + cl + j];
            pat[j] = cp & 0xff;
            nzl[j] = cp >> 8;
        }
        double acc[CLB];
#pragma unroll
        for (int j = 0; j < CLB; j++) acc[j] = 0.0;
        // lane data of the chunk in flight (PF)
        double nq = 0.0, niq = 0.0, nQx = 0.0, nQy = 0.0, nQz = 0.0, nck = 0.0, nE[CLB][NZ];
        auto load_lane = [&](int c0) {
            const int c = c0 + lane;
            const bool live = c < nl;
            const int cc = live ? c : c0;
            nq = pp.p[ol + cc];
            niq = pp.ip[ol + cc];
            nQx = pp.Px[ol + cc];
            nQy = pp.Py[ol + cc];
            nQz = pp.Pz[ol + cc];
            nck = live ? pp.c[ol + cc] : 0.0;
#pragma unroll
            for (int j = 0; j < CLB; j++)
#pragma unroll
                for (int k = 0; k < NZ; k++) {
                    if constexpr (SP)
                        nE[j][k] = k < nzl[j] ? sd.ec[el[j] + (int64_t)k * capl[j] + cc] : 0.0;
                    else
                        nE[j][k] = cl + j < it.cl1 ? pp.E[el[j] + (int64_t)k * capl[j] + cc] : 0.0;
                }
        };
        if constexpr (PF) load_lane(0);
        for (int a = it.a0; a < it.a1; a++) {
            const double *pa = pu + (int64_t)a * 8;
            const double p = pa[0], ip = pa[1], Px = pa[2], Py = pa[3], Pz = pa[4], cb = pa[5];
            const double *ea = eu2 + (int64_t)a * HU * ncu;
            double Xa[CLB][HU];
#pragma unroll
            for (int j = 0; j < CLB; j++)
#pragma unroll
                for (int hu = 0; hu < HU; hu++) Xa[j][hu] = 0.0;
            for (int c0 = 0; c0 < nl; c0 += 64) {
                const int c = c0 + lane;
                const bool live = c < nl;
                const int cc = live ? c : c0; // in-range index for idle lanes
                double q, iq, Qx, Qy, Qz, ck, Er[CLB][NZ];
                if constexpr (PF) {
                    // this chunk's lane data were loaded one chunk ahead; the
                    // next chunk's (or, after the last, the first chunk of the
                    // next uniform primitive pair) go in flight now, behind
                    // this chunk's Boys, R and X work
                    q = nq;
                    iq = niq;
                    Qx = nQx;
                    Qy = nQy;
                    Qz = nQz;
                    ck = nck;
#pragma unroll
                    for (int j = 0; j < CLB; j++)
#pragma unroll
                        for (int k = 0; k < NZ; k++) Er[j][k] = nE[j][k];
                    // (one chunk in all: its data stay in the prefetch set)
                    const int n0 = c0 + 64 < nl ? c0 + 64 : 0;
                    if (c0 + 64 < nl || (a + 1 < it.a1 && nl > 64)) load_lane(n0);
                } else {
                    q = pp.p[ol + cc];
                    iq = pp.ip[ol + cc];
                    Qx = pp.Px[ol + cc];
                    Qy = pp.Py[ol + cc];
                    Qz = pp.Pz[ol + cc];
                    ck = live ? pp.c[ol + cc] : 0.0;
                }
                const double X = Px - Qx, Y = Py - Qy, Z = Pz - Qz;
                const double s = p + q;
                const double rs = rsqrt_pos(s), is = rs * rs;
                const double alpha = (p * q) * is;
                const double T = alpha * ((X * X + Y * Y) + Z * Z);
                double F[L + 1];
                if constexpr (TB)
                    boys_table<L>(T, lbt, F);
                else
                    boys<L>(T, F);
                double R[HR];
                hermite_R<L>(alpha, X, Y, Z, F, R);
                const double f = (cb * ck) * (((kPi52 * ip) * iq) * rs);
                // the lane E (dense: with the prefactor and the Hermite sign
                // folded in here; SP: lane_contract applies both)
                double El[CLB][NZ];
#pragma unroll
                for (int j = 0; j < CLB; j++)
#pragma unroll
                    for (int k = 0; k < NZ; k++) {
                        // (guarded loads: as a condition the guard keeps the
                        // compiler from issuing them one at a time next to
                        // their uses, which serialized their latency; cl + j <
                        // it.cl1 always holds)
                        if constexpr (SP) {
                            if constexpr (PF)
                                El[j][k] = Er[j][k];
                            else
                                El[j][k] =
                                    k < nzl[j] ? sd.ec[el[j] + (int64_t)k * capl[j] + cc] : 0.0;
                        } else {
                            double e;
                            if constexpr (PF)
                                e = Er[j][k];
                            else
                                e = cl + j < it.cl1 ? pp.E[el[j] + (int64_t)k * capl[j] + cc] : 0.0;
                            El[j][k] = ((kH.t[k] + kH.u[k] + kH.v[k]) & 1) ? -(e * f) : e * f;
                        }
                    }
                if constexpr (SP) {
#pragma unroll
                    for (int j = 0; j < CLB; j++)
                        lane_contract<LU, LL, false>(pat[j], El[j], R, Xa[j],
                                                     [f](double e) { return e * f; });
                } else {
#pragma unroll
                    for (int hl = 0; hl < NZ; hl++)
#pragma unroll
                        for (int hu = 0; hu < HU; hu++)
#pragma unroll
                            for (int j = 0; j < CLB; j++)
                                Xa[j][hu] = fma(El[j][hl],
                                                R[hidx(kH.t[hu] + kH.t[hl], kH.u[hu] + kH.u[hl],
                                                       kH.v[hu] + kH.v[hl])],
                                                Xa[j][hu]);
                }
            }
            // the lane's uniform E, loaded here so that its latency hides
            // behind the LDS reduction below
            double ev[HU];
#pragma unroll
            for (int hu = 0; hu < HU; hu++) ev[hu] = ea[(int64_t)hu * ncu];
            // row r = j HU + hu summed over the 64 lanes into lds[64 S + r],
            // RC rows per round through the transpose buffer (NCH rounds
            // keep the buffer small enough for the Boys table beside it)
#pragma unroll
            for (int ch = 0; ch < NCH; ch++) {
                if (ch > 0) __syncthreads(); // the previous round's reads are done
#pragma unroll
                for (int rr = 0; rr < RC; rr++) {
                    const int r = ch * RC + rr;
                    lds[lane * S + rr] = Xa[r / HU][r % HU];
                }
                __syncthreads();
                if constexpr (RC <= 64) {
                    // P lanes per segment, 64 / P segments of each row
                    constexpr int P = pow2ceil(RC), G = 64 / P, SEG = 64 / G;
                    const int r = lane & (P - 1), k = lane / P;
                    double t4[4] = {0.0, 0.0, 0.0, 0.0};
                    if (r < RC) {
#pragma unroll
                        for (int i = 0; i < SEG; i++) t4[i & 3] += lds[(k * SEG + i) * S + r];
                    }
                    double v = (t4[0] + t4[1]) + (t4[2] + t4[3]);
#pragma unroll
                    for (int o = P; o < 64; o <<= 1) v += __shfl_xor(v, o, 64);
                    if (k == 0 && r < RC) lds[64 * S + ch * RC + r] = v;
                } else {
                    for (int r = lane; r < RC; r += 64) {
                        double t4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                        for (int i = 0; i < 64; i++) t4[i & 3] += lds[i * S + r];
                        lds[64 * S + ch * RC + r] = (t4[0] + t4[1]) + (t4[2] + t4[3]);
                    }
                }
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < CLB; j++) {
                double y[2] = {0.0, 0.0};
#pragma unroll
                for (int hu = 0; hu < HU; hu++)
                    y[hu & 1] = fma(ev[hu], lds[64 * S + j * HU + hu], y[hu & 1]);
                acc[j] += y[0] + y[1];
            }
        }
#pragma unroll
        for (int j = 0; j < CLB; j++)
            if (own) part[it.part + (int64_t)lane * nlc + (cl + j - it.cl0)] = acc[j];
    }
}

// Sum the slices of each group and write triangle order.  A diagonal shell
// quartet (su == sl) holds every component quartet twice; the (P >= Q) copy
// is written.  With pre, off-diagonal component quartets failing
// |pre_P pre_Q| >= 1e-30 are written as zero (src/two_e_slater.F90:268-272).
__global__ __launch_bounds__(256) void k_shell_scatter(const ShellGroup *groups, SPD sd,
                                                       const double *part, const double *pre,
                                                       double *tri)
{
    const ShellGroup g = groups[blockIdx.x];
    const int n = g.nu * g.nl;
    const int bu = sd.coff[g.su] + g.cu0, bl = sd.coff[g.sl] + g.cl0;
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        double v = 0.0;
        for (int s = 0; s < g.ns; s++) v += part[g.part + (int64_t)s * n + t];
        const int u = t / g.nl, l = t - u * g.nl;
        const int P = sd.comp[bu + u], Q = sd.comp[bl + l];
        if (g.su == g.sl && P < Q) continue;
        if (pre && P != Q && fabs(pre[P] * pre[Q]) < 1.0e-30) v = 0.0;
        tri[P > Q ? zv_tri(P, Q) : zv_tri(Q, P)] = v;
    }
}

__global__ void k_zero(double *x, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0.0;
}

// Largest lane-component batch of each (LU, LL) class, direct and reduction
// mode: the lane components of a shell pair run in passes of this many (and
// one pass of the remainder); the components of one pass share the Boys
// values and the R table of every primitive quartet, so wider passes redo
// them less often.  Chosen so that the R table, the pass's lane E and its X
// sums stay in the register file without spilling, and by shell-pair sizes:
// an order-0 pair has 1 component, order 1 has 3 (p x s), order 2 has 9
// (p x p) or 6 (d x s), so 3 covers them in as few passes as 4 does with
// fewer registers.  (2, 2) in reduction mode runs passes of 2 at two waves
// per SIMD, which beats passes of 3 at one wave (949 against 1088 ms on the
// 200-AO molecule, dense lane E).  The wider passes of (2, 3), (2, 4),
// (3, 3) and (4, 2) in reduction mode and of (3, 3) and (1, 4) in direct
// mode pay off with the sparse lane E (kShellSparse) only.
constexpr int kClbMax = 4;
constexpr int kShellClb[5][5] = {{1, 3, 3, 4, 4},
                                 {1, 3, 3, 4, 4},
                                 {1, 3, 3, 3, 1},
                                 {1, 3, 3, 3, 1},
                                 {1, 3, 2, 1, 1}};
constexpr int kShellClbRed[5][5] = {{1, 3, 3, 4, 2},
                                    {1, 3, 3, 3, 1},
                                    {1, 3, 2, 3, 2},
                                    {1, 3, 3, 2, 1},
                                    {1, 2, 2, 1, 1}};

// Reduction-mode classes whose lane data are loaded one chunk ahead: those
// at one wave per SIMD anyway, with room for a second set of lane registers
// without spilling (single-component passes of order-4 quartets stay at
// two waves per SIMD without it and keep that), and (2, 1), whose passes
// of 3 keep two waves with it (241 registers; 507 → 478 ms).
constexpr bool kShellPrefetch[5][5] = {{false, false, false, false, false},
                                       {false, false, false, true, true},
                                       {false, true, false, true, false},
                                       {false, true, false, true, false},
                                       {false, true, true, false, false}};
constexpr bool shell_prefetch(int LU, int LL, int B)
{
    return kShellPrefetch[LU][LL] && !(B == 1 && LU + LL <= 4);
}
// Instances that evaluate the Boys function from the Taylor table in LDS
// (boys_table, 8.8 KB per workgroup) instead of erf and the series (boys):
// all whose LDS then still allows two workgroups of one wave per SIMD (20 KB
// each), and the reduction kernels of order 4 and up, which run one wave
// per SIMD anyway, up to 40 KB (four per CU).
constexpr int shell_lds_bytes(int LU, int B, bool red, int nch = 1)
{
    const int HU = hcount(LU), NR = B * HU, S = (NR / nch) | 1;
    return red ? (64 * S + NR) * 8 : kDirectMaxU * B * 64 * 8;
}
constexpr bool shell_table_fits(int LU, int LL, int B, bool red, int nch)
{
    const int bytes = shell_lds_bytes(LU, B, red, nch) + kBoysSlice * 8;
    return LU + LL <= kBoysLmax &&
           (bytes <= 20480 || (red && LU + LL >= 4 && bytes <= 40960));
}
// Reduction instances whose transpose buffer would leave no room for the
// Boys table sum their rows in two rounds of half the rows instead, where
// that makes it fit ((2, 1) with passes of 3, (4, x) with passes of 2).
constexpr int shell_red_rounds(int LU, int LL, int B)
{
    const int NR = B * hcount(LU);
    return !shell_table_fits(LU, LL, B, true, 1) && NR % 2 == 0 &&
                   shell_table_fits(LU, LL, B, true, 2)
               ? 2
               : 1;
}
constexpr bool shell_boys_table(int LU, int LL, int B, bool red)
{
    return shell_table_fits(LU, LL, B, red, red ? shell_red_rounds(LU, LL, B) : 1);
}
// Direct-mode classes whose uniform data are loaded one iteration ahead into
// vector registers (uniform side of order 0 or 1: at most 18 doubles): the
// ones already at one wave per SIMD without it (lane order 3 and 4) and
// still free of spills with it.
constexpr bool shell_prefetch_u(int LU, int LL, int B)
{
    return LU <= 1 && LL >= 3 && !(LL == 4 && LU + B >= 4);
}

// Classes that skip the structural zeros of the lane E (lane_contract), by
// (uniform order, lane order): direct mode, reduction mode.  Chosen by
// per-class kernel time on the 200-AO molecule: the pattern dispatch costs
// scalar compares and registers, so it pays where the X contraction is
// large against the Boys function and R table (uniform order >= 2 in
// reduction mode) and not in the direct mode's short per-uniform-pair
// iterations of uniform order <= 1 (lane order 3: 278 against 103 ms),
// except (1, 4) with passes of 4 (about even); nor for (4, 4), whose 22
// patterns of up to 25 coefficients spill (315 against 145 ms).
constexpr bool kShellSparse[5][5] = {{false, false, false, false, false},
                                     {false, false, false, false, true},
                                     {false, false, false, false, false},
                                     {false, false, false, true, false},
                                     {false, false, false, false, false}};
constexpr bool kShellSparseRed[5][5] = {{false, false, false, false, false},
                                        {false, false, false, false, false},
                                        {false, true, true, true, true},
                                        {false, false, true, true, true},
                                        {false, false, true, true, false}};
constexpr bool shell_sparse(int LU, int LL, bool red)
{
    return red ? kShellSparseRed[LU][LL] : kShellSparse[LU][LL];
}

template <int LU, int LL, int B>
void launch_shell_b(bool red, unsigned grid, hipStream_t st, const PP &pp, const SPD &sd,
                    const ShellItem *items, double *part)
{
    if (red)
        hipLaunchKernelGGL((k_eri_shell_red<LU, LL, B, shell_prefetch(LU, LL, B),
                                            shell_boys_table(LU, LL, B, true),
                                            shell_sparse(LU, LL, true),
                                            shell_red_rounds(LU, LL, B)>),
                           dim3(grid), dim3(64), 0, st, pp, sd, items, sd.pu, sd.eu2, part);
    else
        hipLaunchKernelGGL((k_eri_shell<LU, LL, B, shell_prefetch_u(LU, LL, B),
                                        shell_boys_table(LU, LL, B, false),
                                        shell_sparse(LU, LL, false)>),
                           dim3(grid), dim3(64), 0, st, pp, sd, items, sd.pu, sd.eu, part);
}

// instances up to the class's batch limit only (the larger ones are never
// planned); false for a batch outside it
template <int LU, int LL>
bool launch_shell_t(int clb, bool red, unsigned grid, hipStream_t st, const PP &pp,
                    const SPD &sd, const ShellItem *items, double *part)
{
    constexpr int MR = kShellClbRed[LU][LL], MD = kShellClb[LU][LL];
    if (clb < 1 || clb > (red ? MR : MD)) return false;
    switch (clb) {
    case 1: launch_shell_b<LU, LL, 1>(red, grid, st, pp, sd, items, part); return true;
    case 2:
        if constexpr (MR >= 2 || MD >= 2)
            launch_shell_b<LU, LL, 2>(red, grid, st, pp, sd, items, part);
        return true;
    case 3:
        if constexpr (MR >= 3 || MD >= 3)
            launch_shell_b<LU, LL, 3>(red, grid, st, pp, sd, items, part);
        return true;
    case 4:
        if constexpr (MR >= 4 || MD >= 4)
            launch_shell_b<LU, LL, 4>(red, grid, st, pp, sd, items, part);
        return true;
    }
    return false;
}

template <int LU>
bool launch_shell_u(int LL, int clb, bool red, unsigned grid, hipStream_t st, const PP &pp,
                    const SPD &sd, const ShellItem *items, double *part)
{
    switch (LL) {
    case 0: return launch_shell_t<LU, 0>(clb, red, grid, st, pp, sd, items, part);
    case 1: return launch_shell_t<LU, 1>(clb, red, grid, st, pp, sd, items, part);
    case 2: return launch_shell_t<LU, 2>(clb, red, grid, st, pp, sd, items, part);
    case 3: return launch_shell_t<LU, 3>(clb, red, grid, st, pp, sd, items, part);
    case 4: return launch_shell_t<LU, 4>(clb, red, grid, st, pp, sd, items, part);
    }
    return false;
}

bool launch_shell(int LU, int LL, int clb, bool red, unsigned grid, hipStream_t st,
                  const PP &pp, const SPD &sd, const ShellItem *items, double *part)
{
    switch (LU) {
    case 0: return launch_shell_u<0>(LL, clb, red, grid, st, pp, sd, items, part);
    case 1: return launch_shell_u<1>(LL, clb, red, grid, st, pp, sd, items, part);
    case 2: return launch_shell_u<2>(LL, clb, red, grid, st, pp, sd, items, part);
    case 3: return launch_shell_u<3>(LL, clb, red, grid, st, pp, sd, items, part);
    case 4: return launch_shell_u<4>(LL, clb, red, grid, st, pp, sd, items, part);
    }
    return false;
}

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }
inline int hc(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }

// Boys values, R recursion (2 FMA per entry over all levels) and prefactor
double boys_r_flops(int L)
{
    double HR4 = (L + 1) * (L + 2) * (L + 3) * (L + 4) / 24.0;
    return 25.0 + 3.0 * (L + 1) + 4.0 * (HR4 - 1.0) + 12.0;
}

// X contraction and scaling per primitive quartet and lane component with
// nz nonzero Hermite coefficients
double x_flops(int LU, double nz) { return 2.0 * hc(LU) * nz + hc(LU); }

// nonzero Hermite coefficients of the lane components [l0, l1) of shell pair s
int64_t lane_nz(const zv_context *c, int s, int l0, int l1)
{
    return c->h_sp_nzcum[c->h_sp_coff[s] + l1] - c->h_sp_nzcum[c->h_sp_coff[s] + l0];
}

// uniform components per item and the smallest count that uses the
// reduction mode
constexpr int kMaxUDirect = ZV_MAXU, kMaxURed = 64, kRedMin = kDirectMaxU + 1;

// Estimated kernel cost of finishing, in lane flop-equivalents: the direct
// mode per primitive quartet and pass (scalar-operand FMAs run about three
// times slower than register-only work); the reduction mode per uniform
// primitive pair and pass (LDS transpose, row sums, two barriers and the
// load of the uniform E, mostly latency at one wave per SIMD).
double finish_direct(int LU, int clb, int nu) { return 3.0 * 2.0 * nu * clb * hc(LU); }
double finish_red(int LU, int clb)
{
    const double NR = clb * hc(LU);
    return 64.0 * (4.0 * NR + 128.0 * std::ceil(NR / 64.0)) + 40000.0;
}

// bucket b = 25 (2 (clb - 1) + red) + 5 LU + LL
constexpr int kNB = 25 * 2 * kClbMax;

// A plan's items (bucket b at [lo[b], lo[b + 1])) and groups live in the
// caller's host buffer (items, then groups), from which they are copied to
// the device as they are.
struct Plan {
    ShellItem *items = nullptr;
    size_t nitems = 0;
    int lo[kNB + 1] = {};
    ShellGroup *groups = nullptr;
    size_t ngroups = 0;
    int64_t npart = 0;
    double units[kNB] = {}, flops[kNB] = {}; // per bucket
    double host_ms[4] = {};                  // orientation, items, join, bucket sort
    bool bad = false;                        // an item the kernels cannot take
};

// lane components [0, full) run in passes of the class's batch limit, the
// remainder [full, ncl) in one pass
inline int lane_batch(int LU, int LL, bool red)
{
    return red ? kShellClbRed[LU][LL] : kShellClb[LU][LL];
}
inline int full_passes_end(int LU, int LL, int ncl, bool red)
{
    const int b = lane_batch(LU, LL, red);
    return ncl / b * b;
}

// uniform-component ranges of one orientation: ranges of kMaxURed (reduction
// mode, from kRedMin components on) or kMaxUDirect components; a short tail
// of a reduction-mode pair below kRedMin runs in direct mode
inline int uniform_step(int ncu) { return ncu >= kRedMin ? kMaxURed : kMaxUDirect; }
inline bool uniform_red(int ncu, int nu) { return ncu >= kRedMin && nu >= kRedMin; }

// Kernel-time cost model of (U|L): lanes padded to 64, Boys and R redone per
// pass of lane components and per uniform range; nz: mean nonzero Hermite
// coefficients of a lane component.
double pass_cost(int LU, int LL, int clb, int nu, bool red, double npu, double lanes, double nz)
{
    const double work = lanes * npu * (boys_r_flops(LU + LL) + clb * x_flops(LU, nz));
    return work + (red ? npu * finish_red(LU, clb) : lanes * npu * finish_direct(LU, clb, nu));
}

double item_cost_of(int LU, int LL, int ncu, int ncl, double npu, double lanes)
{
    const double nz = hc(LL);
    const int step = uniform_step(ncu);
    double cost = 0.0;
    for (int cu0 = 0; cu0 < ncu; cu0 += step) {
        const int nu = std::min(step, ncu - cu0);
        const bool red = uniform_red(ncu, nu);
        const int bm = lane_batch(LU, LL, red), n2 = full_passes_end(LU, LL, ncl, red);
        cost += n2 / bm * pass_cost(LU, LL, bm, nu, red, npu, lanes, nz);
        if (ncl > n2) cost += pass_cost(LU, LL, ncl - n2, nu, red, npu, lanes, nz);
    }
    return cost;
}

// The planned cost of an orientation is npu (lanes A + B) with A and B fixed
// by the (order, components) types of the uniform and the lane shell pair
// (pass_cost is linear in npu and lanes), tabulated once per plan.
struct CostTable {
    std::vector<int> type;        // per shell pair
    std::vector<double> npu, lanes; // per shell pair
    std::vector<double> A, B;     // [type u][type l]
    int nt = 0;
    explicit CostTable(const zv_context *c)
    {
        const int nsp = c->nsp;
        std::vector<std::pair<int, int>> keys;
        type.resize(nsp);
        npu.resize(nsp);
        lanes.resize(nsp);
        for (int s = 0; s < nsp; s++) {
            const std::pair<int, int> k{c->h_sp_L[s], c->h_sp_coff[s + 1] - c->h_sp_coff[s]};
            auto it = std::find(keys.begin(), keys.end(), k);
            type[s] = (int)(it - keys.begin());
            if (it == keys.end()) keys.push_back(k);
            npu[s] = c->h_sp_np[s];
            lanes[s] = std::ceil(c->h_sp_np[s] / 64.0) * 64.0;
        }
        nt = (int)keys.size();
        A.resize((size_t)nt * nt);
        B.resize((size_t)nt * nt);
        for (int u = 0; u < nt; u++)
            for (int l = 0; l < nt; l++) {
                const auto [LU, ncu] = keys[u];
                const auto [LL, ncl] = keys[l];
                // shell quartets of these orders only (the others go through
                // component lists and have no shell kernel)
                if (LU > ZV_LMAX_T || LL > ZV_LMAX_T) continue;
                const double b = item_cost_of(LU, LL, ncu, ncl, 1.0, 0.0);
                A[(size_t)u * nt + l] = item_cost_of(LU, LL, ncu, ncl, 1.0, 1.0) - b;
                B[(size_t)u * nt + l] = b;
            }
    }
    double cost(int su, int sl) const
    {
        const size_t k = (size_t)type[su] * nt + type[sl];
        return npu[su] * (lanes[sl] * A[k] + B[k]);
    }
};

// Host threads for the planner: up to 16 (the host share of one GPU), one
// for short lists.
int plan_threads(size_t n)
{
    if (n < 20000) return 1;
    const unsigned hw = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(16u, hw ? hw : 1u));
}

// run f(t) for t = 0 .. nt-1, on nt - 1 extra threads and this one; a
// range whose thread cannot be started runs here (f(t) are independent)
template <typename F> void run_threads(int nt, F &&f)
{
    std::vector<std::thread> th;
    int started = 1;
    for (int t = 1; t < nt; t++, started++) {
        try {
            th.emplace_back(f, t);
        } catch (...) {
            break;
        }
    }
    for (int t = started; t < nt; t++) f(t);
    f(0);
    for (auto &x : th) x.join();
}

// Items of one run of oriented shell quartets, as the sequential planner
// would make them but with partial-sum offsets from 0.
struct PlanPart {
    std::vector<ShellGroup> groups;
    std::vector<ShellItem> tmp[kNB];
    std::vector<float> tcost[kNB];
    double units[kNB] = {}, flops[kNB] = {};
    int64_t npart = 0;
};

struct Orient {
    int su, sl;
    double cost;
};

void plan_items(const zv_context *c, const Orient *ors, size_t n, double budget, PlanPart &pp)
{
    for (size_t q = 0; q < n; q++) {
        const Orient &o = ors[q];
        const int LU = c->h_sp_L[o.su], LL = c->h_sp_L[o.sl], L = LU + LL;
        const int ncu = c->h_sp_coff[o.su + 1] - c->h_sp_coff[o.su];
        const int ncl = c->h_sp_coff[o.sl + 1] - c->h_sp_coff[o.sl];
        const int npu = c->h_sp_np[o.su], npl = c->h_sp_np[o.sl];
        const double lanes = std::ceil(npl / 64.0) * 64.0;
        const int k = 5 * LU + LL;
        const double HU = hc(LU);
        const double nz = hc(LL);
        const int step = uniform_step(ncu);
        bool br_counted = false; // Boys and R counted once per shell quartet
        for (int cu0 = 0; cu0 < ncu; cu0 += step) {
            const int nu = std::min(step, ncu - cu0);
            const bool red = uniform_red(ncu, nu);
            const int bm = lane_batch(LU, LL, red), n2 = full_passes_end(LU, LL, ncl, red);
            // segment [0, n2) in passes of bm components, [n2, ncl) in one pass
            for (int seg = 0; seg < 2; seg++) {
                const int clb = seg == 0 ? bm : ncl - n2;
                const int s0 = seg == 0 ? 0 : n2, s1 = seg == 0 ? n2 : ncl;
                if (clb <= 0) continue;
                const int npass = (s1 - s0) / clb;
                if (npass <= 0) continue;
                const double wp = pass_cost(LU, LL, clb, nu, red, npu, lanes, nz); // per pass
                const int k1 = (int)std::min<double>(
                    npass, std::max(1.0, std::ceil(npass * wp / budget)));
                const int b = 25 * (2 * (clb - 1) + (red ? 1 : 0)) + k;
                for (int r = 0; r < k1; r++) {
                    const int p0 = (int)((int64_t)npass * r / k1);
                    const int p1 = (int)((int64_t)npass * (r + 1) / k1);
                    const int l0 = s0 + p0 * clb, l1 = s0 + p1 * clb;
                    const int nlr = l1 - l0;
                    const double cr = (p1 - p0) * wp;
                    const int ns =
                        (int)std::min<double>(npu, std::max(1.0, std::ceil(cr / budget)));
                    pp.groups.push_back(ShellGroup{o.su, o.sl, cu0, nu, l0, nlr, ns, 0, pp.npart});
                    // algorithmic flops: Boys and R once per primitive quartet
                    // (however many lane passes redo them); per lane component
                    // the X contraction over its nonzero Hermite coefficients,
                    // then either nu direct dot products per primitive quartet
                    // or the sum over lane primitives per primitive quartet
                    // plus nu dot products per uniform primitive
                    const double pq = (double)npu * npl;
                    pp.units[b] += pq * nlr;
                    const double xf =
                        2.0 * HU * (double)lane_nz(c, o.sl, l0, l1) + nlr * HU; // sum of x_flops
                    double fl = pq * ((br_counted ? 0.0 : boys_r_flops(L)) + xf);
                    br_counted = true;
                    fl += red ? nlr * (pq * HU + (double)npu * 2.0 * nu * HU)
                              : nlr * pq * 2.0 * nu * HU;
                    pp.flops[b] += fl;
                    for (int sl = 0; sl < ns; sl++) {
                        const int a0 = (int)((int64_t)npu * sl / ns);
                        const int a1 = (int)((int64_t)npu * (sl + 1) / ns);
                        pp.tmp[b].push_back(ShellItem{o.su, o.sl, a0, a1, l0, l1, cu0, nu,
                                                      pp.npart + (int64_t)sl * nu * nlr});
                        pp.tcost[b].push_back((float)(cr * (a1 - a0) / npu));
                    }
                    pp.npart += (int64_t)ns * nu * nlr;
                }
            }
        }
    }
}

// Orientation, splitting and ordering of a list of shell quartets (x, y),
// both shell pairs of order <= ZV_LMAX_T with kept primitive pairs.  Items
// are cut along lane-component passes first (no extra work), then along
// uniform primitive-pair slices (partial sums added by the scatter).  Long
// lists are planned on host threads over contiguous runs of quartets whose
// items are joined in order, so the plan is the same for any thread count.
// the kernels assume whole lane passes (no guard per lane component) and,
// in direct mode, at most kDirectMaxU uniform components
inline bool item_ok(const ShellItem &it, int clb, bool red)
{
    return it.cl1 > it.cl0 && (it.cl1 - it.cl0) % clb == 0 &&
           (red || (it.nu >= 1 && it.nu <= kDirectMaxU));
}

void plan_shell(const zv_context *c, const std::vector<std::pair<int, int>> &sq, Plan &pl,
                std::vector<char> &hbuf)
{
    using clk = std::chrono::steady_clock;
    auto ms_since = [](clk::time_point t) {
        return std::chrono::duration<double, std::milli>(clk::now() - t).count();
    };
    auto tt = clk::now();
    const size_t nq = sq.size();
    const int nt = plan_threads(nq);
    const CostTable ct(c);
    std::vector<Orient> ors(nq);
    run_threads(nt, [&](int t) {
        for (size_t q = nq * t / nt; q < nq * (t + 1) / nt; q++) {
            const int x = sq[q].first, y = sq[q].second;
            double cxy = ct.cost(x, y), cyx = x == y ? cxy : ct.cost(y, x);
            ors[q] = cxy <= cyx ? Orient{x, y, cxy} : Orient{y, x, cyx};
        }
    });
    double total = 0.0;
    for (const Orient &o : ors) total += o.cost;
    // items of at most 1/40000 of the total cost: larger quartets are split,
    // so the sorted queue ends on short items
    const double nitems =
        c->tune[ZV_TUNE_ERI_ITEMS] > 0 ? (double)c->tune[ZV_TUNE_ERI_ITEMS] : 40000.0;
    const double budget = std::max(total / nitems, 2.0e6);
    pl.host_ms[0] = ms_since(tt);
    tt = clk::now();
    std::vector<PlanPart> parts(nt);
    run_threads(nt, [&](int t) {
        const size_t q0 = nq * t / nt, q1 = nq * (t + 1) / nt;
        plan_items(c, ors.data() + q0, q1 - q0, budget, parts[t]);
    });
    pl.host_ms[1] = ms_since(tt);
    tt = clk::now();
    // join the runs in order: partial-sum offsets shifted by the runs before
    pl.npart = 0;
    size_t ng = 0, nb_items[kNB] = {};
    for (const PlanPart &pp : parts) {
        ng += pp.groups.size();
        for (int b = 0; b < kNB; b++) nb_items[b] += pp.tmp[b].size();
    }
    size_t at = 0;
    for (int b = 0; b < kNB; b++) {
        pl.lo[b] = (int)at;
        at += nb_items[b];
    }
    pl.lo[kNB] = (int)at;
    hbuf.resize(sizeof(ShellItem) * at + sizeof(ShellGroup) * ng);
    pl.items = (ShellItem *)hbuf.data();
    pl.nitems = at;
    pl.groups = (ShellGroup *)(hbuf.data() + sizeof(ShellItem) * at);
    pl.ngroups = ng;
    std::vector<int64_t> off(nt, 0);
    std::vector<size_t> goff(nt, 0);
    for (int t = 0, g0 = 0; t < nt; t++) {
        PlanPart &pp = parts[t];
        off[t] = pl.npart;
        goff[t] = (size_t)g0;
        g0 += (int)pp.groups.size();
        for (int b = 0; b < kNB; b++) {
            pl.units[b] += pp.units[b];
            pl.flops[b] += pp.flops[b];
        }
        pl.npart += pp.npart;
    }
    run_threads(nt, [&](int t) {
        ShellGroup *dst = pl.groups + goff[t];
        for (ShellGroup g : parts[t].groups) {
            g.part += off[t];
            *dst++ = g;
        }
    });
    pl.host_ms[2] = ms_since(tt);
    tt = clk::now();
    // each bucket in descending cost (quarter-octave bins, counting sort), so the
    // long items start first and the queue ends on short ones; buckets are
    // independent and spread over the threads
    constexpr int NBIN = 256;
    static_assert(NBIN <= 256, "bins are stored as bytes");
    std::vector<char> bad(nt, 0);
    run_threads(nt, [&](int t) {
        std::vector<uint8_t> bin;
        std::vector<const ShellItem *> src;
        std::vector<float> cst;
        for (int b = t; b < kNB; b += nt) {
            const size_t n = nb_items[b];
            if (!n) continue;
            const int clb = b / 50 + 1;
            const bool red = ((b / 25) & 1) != 0;
            src.clear();
            cst.clear();
            std::vector<int64_t> shift;
            for (int u = 0; u < nt; u++)
                for (size_t i = 0; i < parts[u].tmp[b].size(); i++) {
                    src.push_back(&parts[u].tmp[b][i]);
                    cst.push_back(parts[u].tcost[b][i]);
                    shift.push_back(off[u]);
                }
            bin.resize(n);
            int cnt[NBIN + 1] = {};
            for (size_t i = 0; i < n; i++) {
                const int e = (int)std::floor(4.0 * std::log2(std::max(cst[i], 1.0f)));
                bin[i] = (uint8_t)(NBIN - 1 - std::min(e, NBIN - 1)); // large cost -> small bin
                cnt[bin[i] + 1]++;
            }
            for (int k = 0; k < NBIN; k++) cnt[k + 1] += cnt[k];
            for (size_t i = 0; i < n; i++) {
                ShellItem it = *src[i];
                it.part += shift[i];
                if (!item_ok(it, clb, red)) bad[t] = 1;
                pl.items[pl.lo[b] + cnt[bin[i]]++] = it;
            }
        }
    });
    for (char x : bad) pl.bad = pl.bad || x;
    pl.host_ms[3] = ms_since(tt);
}

} // namespace

// ====================================================================== host
static SPD spd_of(zv_context *c);

// Shell-pair tables on the device and the AoS copies of the uniform side.
int zv_shell_ready(zv_context *c, const PP &pp)
{
    const int nsp = c->nsp;
    size_t meta =
        sizeof(int) * (2 * (size_t)c->npair + nsp + 1) + sizeof(int64_t) * 2 * nsp + 64;
    ZV_CHECK(zv_ensure(c, c->b_spmeta, meta));
    if (c->sp_ready) return ZV_OK;
    // Hermite pattern of each component with a shell-kernel order
    c->h_sp_cpat.assign(c->npair, 0);
    for (int q = 0; q < c->npair; q++) {
        const int P = c->h_sp_comp[q];
        const int L = c->h_pair_L[P];
        if (L > ZV_LMAX_T) continue;
        const int i = c->h_pair_i[P] - 1, k = c->h_pair_k[P] - 1;
        int S[3], deg = 0;
        for (int d = 0; d < 3; d++) {
            S[d] = c->h_npower[3 * i + d] + c->h_npower[3 * k + d];
            deg += S[d];
        }
        const int nr = (c->h_itype[i] == 2) + (c->h_itype[k] == 2);
        const int pid = deg + 2 * nr == L ? kHP.find(L, herm_union(S[0], S[1], S[2], nr)) : -1;
        if (pid < 0)
            return zv_fail(c, ZV_ERR_STATE, "AO pair %d: no Hermite pattern of order %d", P, L);
        c->h_sp_cpat[q] = pid | kHP.nz[L][pid] << 8;
    }
    c->h_sp_nzcum.assign(c->npair + 1, 0);
    for (int q = 0; q < c->npair; q++)
        c->h_sp_nzcum[q + 1] = c->h_sp_nzcum[q] + (c->h_sp_cpat[q] >> 8);
    std::vector<int64_t> puoff(nsp), euoff(nsp);
    c->h_sp_np.assign(nsp, 0);
    int64_t pu_tot = 0, eu_tot = 0;
    for (int s = 0; s < nsp; s++) {
        const int c0 = c->h_sp_coff[s], nc = c->h_sp_coff[s + 1] - c0;
        const int np = c->h_cnt[c->h_sp_comp[c0]];
        c->h_sp_np[s] = np;
        puoff[s] = pu_tot;
        euoff[s] = eu_tot;
        pu_tot += 8 * (int64_t)np;
        eu_tot += (int64_t)np * nc * hc(c->h_sp_L[s]);
    }
    ZV_CHECK(zv_ensure(c, c->b_spu, sizeof(double) * std::max<int64_t>(pu_tot, 1)));
    // two layouts of the uniform-side E: [a][comp][h] then [a][h][comp]
    ZV_CHECK(zv_ensure(c, c->b_spe, sizeof(double) * std::max<int64_t>(2 * eu_tot, 1)));
    ZV_CHECK(zv_ensure(c, c->b_spec, sizeof(double) * std::max<int64_t>(c->e_total, 1)));
    c->sp_eu_total = eu_tot;
    int64_t *d_puoff = (int64_t *)c->b_spmeta.p;
    int64_t *d_euoff = d_puoff + nsp;
    int *d_comp = (int *)(d_euoff + nsp);
    int *d_coff = d_comp + c->npair;
    int *d_cpat = d_coff + nsp + 1;
    ZV_HIPCHECK(c, hipMemcpyAsync(d_puoff, puoff.data(), sizeof(int64_t) * nsp,
                                  hipMemcpyHostToDevice, c->stream));
    ZV_HIPCHECK(c, hipMemcpyAsync(d_euoff, euoff.data(), sizeof(int64_t) * nsp,
                                  hipMemcpyHostToDevice, c->stream));
    ZV_HIPCHECK(c, hipMemcpyAsync(d_comp, c->h_sp_comp.data(), sizeof(int) * c->npair,
                                  hipMemcpyHostToDevice, c->stream));
    ZV_HIPCHECK(c, hipMemcpyAsync(d_coff, c->h_sp_coff.data(), sizeof(int) * (nsp + 1),
                                  hipMemcpyHostToDevice, c->stream));
    ZV_HIPCHECK(c, hipMemcpyAsync(d_cpat, c->h_sp_cpat.data(), sizeof(int) * c->npair,
                                  hipMemcpyHostToDevice, c->stream));
    SPD sd = spd_of(c);
    {
        PhaseTimer tm(c, PH_PRIM);
        hipLaunchKernelGGL(k_shellpairs, dim3(nsp), dim3(256), 0, c->stream, pp, sd,
                           (double *)c->b_spu.p, (double *)c->b_spe.p,
                           (double *)c->b_spe.p + eu_tot);
        ZV_HIPCHECK(c, hipGetLastError());
        hipLaunchKernelGGL(k_compact_e, dim3(c->npair), dim3(256), 0, c->stream, pp, sd,
                           (double *)c->b_spec.p);
        ZV_HIPCHECK(c, hipGetLastError());
    }
    ZV_HIPCHECK(c, hipStreamSynchronize(c->stream));
    c->sp_ready = true;
    return ZV_OK;
}

static SPD spd_of(zv_context *c)
{
    int64_t *d_puoff = (int64_t *)c->b_spmeta.p;
    int64_t *d_euoff = d_puoff + c->nsp;
    int *d_comp = (int *)(d_euoff + c->nsp);
    int *d_coff = d_comp + c->npair;
    const int *d_cpat = d_coff + c->nsp + 1;
    const double *eu = (const double *)c->b_spe.p;
    return SPD{d_comp, d_coff, d_puoff, d_euoff, (const double *)c->b_spu.p, eu,
               eu + c->sp_eu_total, d_cpat, (const double *)c->b_spec.p};
}

// Plan and launch the shell kernels of one Schwarz phase over the shell
// quartets sq; the partial sums stay in bpart for zv_shell_scatter.
int zv_shell_run(zv_context *c, const PP &pp, const std::vector<std::pair<int, int>> &sq,
                 DevBuf &bitem, DevBuf &bgroup, DevBuf &bpart, std::vector<char> &hplan,
                 int64_t *ngroups)
{
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    Plan pl;
    plan_shell(c, sq, pl, hplan);
    const auto t1 = clk::now();
    *ngroups = (int64_t)pl.ngroups;
    if (!pl.nitems) return ZV_OK;
    if (pl.bad) {
        // the first offending item, for the message
        for (int b = 0; b < kNB; b++)
            for (int t = pl.lo[b]; t < pl.lo[b + 1]; t++) {
                const ShellItem &it = pl.items[t];
                if (!item_ok(it, b / 50 + 1, ((b / 25) & 1) != 0))
                    return zv_fail(c, ZV_ERR_STATE, "shell item with lane components [%d, %d) "
                                   "and %d uniform components in bucket %d", it.cl0, it.cl1,
                                   it.nu, b);
            }
    }
    const size_t bi = sizeof(ShellItem) * pl.nitems;
    const size_t bg = sizeof(ShellGroup) * pl.ngroups;
    const auto t2 = clk::now();
    ZV_CHECK(zv_ensure(c, bitem, bi));
    ZV_CHECK(zv_ensure(c, bgroup, bg));
    ZV_CHECK(zv_ensure(c, bpart, sizeof(double) * pl.npart));
    ZV_HIPCHECK(c, hipMemcpyAsync(bitem.p, hplan.data(), bi, hipMemcpyHostToDevice, c->stream));
    ZV_HIPCHECK(c, hipMemcpyAsync(bgroup.p, hplan.data() + bi, bg, hipMemcpyHostToDevice,
                                  c->stream));
    if (c->verbose) {
        const auto t3 = clk::now();
        auto ms = [](clk::duration d) { return std::chrono::duration<double, std::milli>(d).count(); };
        fprintf(stderr, "[zvmc] shell plan: %zu quartets -> %zu items, %zu groups; plan %.1f ms "
                "(orientation %.1f, items %.1f, join %.1f, bucket sort %.1f), device alloc + "
                "copy %.1f ms\n", sq.size(), pl.nitems, pl.ngroups, ms(t1 - t0), pl.host_ms[0],
                pl.host_ms[1], pl.host_ms[2], pl.host_ms[3], ms(t3 - t2));
    }
    const SPD sd = spd_of(c);
    const ShellItem *di = (const ShellItem *)bitem.p;
    // The buckets write disjoint partial sums, so they run side by side on
    // the context's stream and its side streams (a bucket's last items leave
    // CUs idle that the next bucket fills), joined back before the scatter.
    // Largest planned cost first, each to the stream with the least planned
    // cost so far.
    // bucket b = 25 v + 5 LU + LL, v = 2 (clb - 1) + red: statistics slot 36 + b
    static_assert(36 + kNB <= ZV_NSTAT, "statistics slots");
    constexpr int NST = 1 + zv_context::NSIDE;
    hipStream_t sts[NST] = {c->stream};
    for (int k = 0; k < zv_context::NSIDE; k++) sts[1 + k] = c->side[k];
    int order[kNB], nb = 0;
    for (int b = 0; b < kNB; b++)
        if (pl.lo[b + 1] > pl.lo[b]) order[nb++] = b;
    std::sort(order, order + nb, [&](int x, int y) { return pl.flops[x] > pl.flops[y]; });
    double load[NST] = {};
    // phase time: one span around the fork and join; the buckets' own spans
    // overlap and only go to their statistics slots
    PhaseTimer wall(c, PH_ERI);
    ZV_HIPCHECK(c, hipEventRecord(c->ev_fork, c->stream));
    for (int k = 0; k < zv_context::NSIDE; k++)
        ZV_HIPCHECK(c, hipStreamWaitEvent(c->side[k], c->ev_fork, 0));
    for (int t = 0; t < nb; t++) {
        const int b = order[t];
        const int n = pl.lo[b + 1] - pl.lo[b];
        const int v = b / 25, k = b % 25, LU = k / 5, LL = k % 5;
        const int cls = 36 + b;
        c->cls_pq[cls] += pl.units[b];
        c->cls_flops[cls] += pl.flops[b];
        // ZV_TUNE_ERI_STREAMS limits the streams (1: buckets one after another)
        const int nst = c->tune[ZV_TUNE_ERI_STREAMS] > 0
                            ? (int)std::min<int64_t>(c->tune[ZV_TUNE_ERI_STREAMS], NST)
                            : NST;
        int s = 0;
        for (int q = 1; q < nst; q++)
            if (load[q] < load[s]) s = q;
        load[s] += pl.flops[b];
        PhaseTimer tm(c, PH_N, cls, sts[s]);
        if (!launch_shell(LU, LL, v / 2 + 1, (v & 1) != 0, (unsigned)n, sts[s], pp, sd,
                          di + pl.lo[b], (double *)bpart.p))
            return zv_fail(c, ZV_ERR_STATE, "no shell kernel for (%d,%d)", LU, LL);
        ZV_HIPCHECK(c, hipGetLastError());
    }
    for (int k = 0; k < zv_context::NSIDE; k++) {
        ZV_HIPCHECK(c, hipEventRecord(c->ev_join[k], c->side[k]));
        ZV_HIPCHECK(c, hipStreamWaitEvent(c->stream, c->ev_join[k], 0));
    }
    return ZV_OK;
}

int zv_shell_scatter(zv_context *c, const DevBuf &bgroup, const DevBuf &bpart, int64_t ngroups,
                     const double *pre, double *tri)
{
    if (ngroups <= 0) return ZV_OK;
    hipLaunchKernelGGL(k_shell_scatter, dim3((unsigned)ngroups), dim3(256), 0, c->stream,
                       (const ShellGroup *)bgroup.p, spd_of(c), (const double *)bpart.p, pre,
                       tri);
    ZV_HIPCHECK(c, hipGetLastError());
    return ZV_OK;
}

int zv_zero(zv_context *c, double *x, int64_t n)
{
    if (n <= 0) return ZV_OK;
    hipLaunchKernelGGL(k_zero, dim3(nblk(n, 256)), dim3(256), 0, c->stream, x, n);
    ZV_HIPCHECK(c, hipGetLastError());
    return ZV_OK;
}
