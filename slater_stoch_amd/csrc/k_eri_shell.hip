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
// components, up to ZV_MAXU uniform-side components and a slice of the
// uniform-side primitive pairs.  Lanes own lane-side primitive pairs (their
// E in registers, reloaded per lane component); the uniform primitive pairs
// are walked in the inner loop through an AoS copy ([a][component][h]) read
// with scalar loads, so each uniform E value is one SGPR operand shared by
// the 64 lanes.  Per-lane sums over the uniform components live in LDS;
// each item writes its partial sums, and one scatter kernel adds the slices,
// applies the Schwarz test per component quartet and writes triangle order.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
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
    const double *eu;     // [a][component][h]
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

// AoS copies of a shell pair's primitive data and E (one block per shell pair)
__global__ __launch_bounds__(256) void k_shellpairs(PP pp, SPD sd, double *pu, double *eu)
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
    double *de = eu + sd.euoff[s];
    const int64_t tot = (int64_t)np * nc * H;
    for (int64_t t = threadIdx.x; t < tot; t += blockDim.x) {
        const int h = (int)(t % H);
        const int64_t r = t / H;
        const int cu = (int)(r % nc);
        const int a = (int)(r / nc);
        const int P = sd.comp[c0 + cu];
        de[t] = pp.E[pp.eoff[P] + (int64_t)h * pp.cap[P] + a];
    }
}

template <int LU, int LL>
__global__ __launch_bounds__(64) void k_eri_shell(PP pp, SPD sd, const ShellItem *items,
                                                  double *part)
{
    constexpr int HU = hcount(LU), HL = hcount(LL), L = LU + LL, HR = hcount(L);
    __shared__ double accs[ZV_MAXU * 64];
    const int lane = threadIdx.x;
    const ShellItem it = items[blockIdx.x];
    const int ub = sd.coff[it.su], lb = sd.coff[it.sl];
    const int ncu = sd.coff[it.su + 1] - ub;
    const int repl = sd.comp[lb];
    const int nl = pp.cnt[repl];
    const int64_t ol = pp.off[repl];
    const double *pu = sd.pu + sd.puoff[it.su];
    const double *eu = sd.eu + sd.euoff[it.su] + (int64_t)it.cu0 * HU;
    const int64_t erow = (int64_t)ncu * HU;
    const int nlc = it.cl1 - it.cl0;
    for (int cl = it.cl0; cl < it.cl1; cl++) {
        const int Pl = sd.comp[lb + cl];
        const int64_t el = pp.eoff[Pl];
        const int capl = pp.cap[Pl];
        for (int u = 0; u < it.nu; u++) accs[u * 64 + lane] = 0.0;
        for (int c0 = 0; c0 < nl; c0 += 64) {
            const int c = c0 + lane;
            const bool live = c < nl;
            const int cc = live ? c : c0; // in-range index for idle lanes
            const double q = pp.p[ol + cc], iq = pp.ip[ol + cc];
            const double Qx = pp.Px[ol + cc], Qy = pp.Py[ol + cc], Qz = pp.Pz[ol + cc];
            const double ck = live ? pp.c[ol + cc] : 0.0;
            double El[HL];
#pragma unroll
            for (int h = 0; h < HL; h++) {
                const double e = pp.E[el + (int64_t)h * capl + cc];
                El[h] = ((kH.t[h] + kH.u[h] + kH.v[h]) & 1) ? -e : e;
            }
            for (int a = it.a0; a < it.a1; a++) {
                const double *pa = pu + (int64_t)a * 8;
                const double p = pa[0], ip = pa[1];
                const double X = pa[2] - Qx, Y = pa[3] - Qy, Z = pa[4] - Qz;
                const double s = p + q;
                const double is = 1.0 / s;
                const double alpha = (p * q) * is;
                const double T = alpha * ((X * X + Y * Y) + Z * Z);
                double F[L + 1];
                boys<L>(T, F);
                double R[HR];
                hermite_R<L>(alpha, X, Y, Z, F, R);
                const double f = (pa[5] * ck) * (((kPi52 * ip) * iq) * sqrt(is));
                double Xv[HU];
#pragma unroll
                for (int hu = 0; hu < HU; hu++) {
                    double w = 0.0;
#pragma unroll
                    for (int hl = 0; hl < HL; hl++)
                        w = fma(El[hl],
                                R[hidx(kH.t[hu] + kH.t[hl], kH.u[hu] + kH.u[hl],
                                       kH.v[hu] + kH.v[hl])],
                                w);
                    Xv[hu] = w * f;
                }
                const double *ea = eu + (int64_t)a * erow;
                for (int u = 0; u < it.nu; u++) {
                    double y = 0.0;
#pragma unroll
                    for (int hu = 0; hu < HU; hu++) y = fma(ea[u * HU + hu], Xv[hu], y);
                    accs[u * 64 + lane] += y;
                }
            }
        }
        for (int u = 0; u < it.nu; u++) {
            const double v = wave_sum(accs[u * 64 + lane]);
            if (lane == 0) part[it.part + (int64_t)u * nlc + (cl - it.cl0)] = v;
        }
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

template <int LU, int LL>
void launch_shell_t(unsigned grid, hipStream_t st, const PP &pp, const SPD &sd,
                    const ShellItem *items, double *part)
{
    hipLaunchKernelGGL((k_eri_shell<LU, LL>), dim3(grid), dim3(64), 0, st, pp, sd, items, part);
}

template <int LU>
bool launch_shell_u(int LL, unsigned grid, hipStream_t st, const PP &pp, const SPD &sd,
                    const ShellItem *items, double *part)
{
    switch (LL) {
    case 0: launch_shell_t<LU, 0>(grid, st, pp, sd, items, part); return true;
    case 1: launch_shell_t<LU, 1>(grid, st, pp, sd, items, part); return true;
    case 2: launch_shell_t<LU, 2>(grid, st, pp, sd, items, part); return true;
    case 3: launch_shell_t<LU, 3>(grid, st, pp, sd, items, part); return true;
    case 4: launch_shell_t<LU, 4>(grid, st, pp, sd, items, part); return true;
    }
    return false;
}

bool launch_shell(int LU, int LL, unsigned grid, hipStream_t st, const PP &pp, const SPD &sd,
                  const ShellItem *items, double *part)
{
    switch (LU) {
    case 0: return launch_shell_u<0>(LL, grid, st, pp, sd, items, part);
    case 1: return launch_shell_u<1>(LL, grid, st, pp, sd, items, part);
    case 2: return launch_shell_u<2>(LL, grid, st, pp, sd, items, part);
    case 3: return launch_shell_u<3>(LL, grid, st, pp, sd, items, part);
    case 4: return launch_shell_u<4>(LL, grid, st, pp, sd, items, part);
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

// Contraction flops per primitive quartet, lane component and uniform range of nu
double contract_flops(int LU, int LL, int nu)
{
    double HU = hc(LU), HL = hc(LL);
    return 2.0 * HU * HL + HU + 2.0 * nu * HU;
}

struct Plan {
    std::vector<ShellItem> items; // class k = 5 LU + LL at [lo[k], lo[k + 1])
    int lo[26] = {};
    std::vector<ShellGroup> groups;
    int64_t npart = 0;
    double units[25] = {}, flops[25] = {};
};

// Kernel-time cost model of (U|L): lanes padded to 64, Boys and R redone per
// lane component and uniform range.
double item_cost(const zv_context *c, int su, int sl)
{
    const int LU = c->h_sp_L[su], LL = c->h_sp_L[sl];
    const int ncu = c->h_sp_coff[su + 1] - c->h_sp_coff[su];
    const int ncl = c->h_sp_coff[sl + 1] - c->h_sp_coff[sl];
    const double npu = c->h_sp_np[su], lanes = std::ceil(c->h_sp_np[sl] / 64.0) * 64.0;
    double cost = 0.0;
    for (int cu0 = 0; cu0 < ncu; cu0 += ZV_MAXU) {
        const int nu = std::min(ZV_MAXU, ncu - cu0);
        cost += lanes * npu * ncl * (boys_r_flops(LU + LL) + contract_flops(LU, LL, nu));
    }
    return cost;
}

// Orientation, splitting and ordering of a list of shell quartets (x, y),
// both shell pairs of order <= ZV_LMAX_T with kept primitive pairs.
void plan_shell(const zv_context *c, const std::vector<std::pair<int, int>> &sq, Plan &pl)
{
    struct Or {
        int su, sl;
        double cost;
    };
    std::vector<Or> ors;
    ors.reserve(sq.size());
    double total = 0.0;
    for (const auto &xy : sq) {
        const int x = xy.first, y = xy.second;
        double cxy = item_cost(c, x, y), cyx = x == y ? cxy : item_cost(c, y, x);
        Or o = cxy <= cyx ? Or{x, y, cxy} : Or{y, x, cyx};
        ors.push_back(o);
        total += o.cost;
    }
    const double budget = std::max(total / 40000.0, 2.0e6);
    std::vector<std::pair<double, ShellItem>> tmp[25];
    for (const Or &o : ors) {
        const int LU = c->h_sp_L[o.su], LL = c->h_sp_L[o.sl], L = LU + LL;
        const int ncu = c->h_sp_coff[o.su + 1] - c->h_sp_coff[o.su];
        const int ncl = c->h_sp_coff[o.sl + 1] - c->h_sp_coff[o.sl];
        const int npu = c->h_sp_np[o.su], npl = c->h_sp_np[o.sl];
        const double lanes = std::ceil(npl / 64.0) * 64.0;
        const int k = 5 * LU + LL;
        for (int cu0 = 0; cu0 < ncu; cu0 += ZV_MAXU) {
            const int nu = std::min(ZV_MAXU, ncu - cu0);
            const double w = lanes * npu * (boys_r_flops(L) + contract_flops(LU, LL, nu));
            const int k1 = (int)std::min<double>(ncl, std::max(1.0, std::ceil(w * ncl / budget)));
            for (int r = 0; r < k1; r++) {
                const int l0 = (int)((int64_t)ncl * r / k1), l1 = (int)((int64_t)ncl * (r + 1) / k1);
                const int nlr = l1 - l0;
                const double cr = w * nlr;
                const int ns = (int)std::min<double>(npu, std::max(1.0, std::ceil(cr / budget)));
                ShellGroup g{o.su, o.sl, cu0, nu, l0, nlr, ns, 0, pl.npart};
                pl.groups.push_back(g);
                for (int s = 0; s < ns; s++) {
                    const int a0 = (int)((int64_t)npu * s / ns), a1 = (int)((int64_t)npu * (s + 1) / ns);
                    ShellItem it{o.su, o.sl, a0, a1, l0, l1, cu0, nu,
                                 pl.npart + (int64_t)s * nu * nlr};
                    tmp[k].push_back({cr * (a1 - a0) / npu, it});
                    const double pq = (double)(a1 - a0) * npl;
                    pl.units[k] += pq * nlr;
                    pl.flops[k] += pq * (boys_r_flops(L) + nlr * contract_flops(LU, LL, nu));
                }
                pl.npart += (int64_t)ns * nu * nlr;
            }
        }
    }
    pl.items.clear();
    for (int k = 0; k < 25; k++) {
        pl.lo[k] = (int)pl.items.size();
        std::stable_sort(tmp[k].begin(), tmp[k].end(),
                         [](const std::pair<double, ShellItem> &a,
                            const std::pair<double, ShellItem> &b) { return a.first > b.first; });
        for (const auto &e : tmp[k]) pl.items.push_back(e.second);
    }
    pl.lo[25] = (int)pl.items.size();
}

} // namespace

// ====================================================================== host
// Shell-pair tables on the device and the AoS copies of the uniform side.
int zv_shell_ready(zv_context *c, const PP &pp)
{
    const int nsp = c->nsp;
    size_t meta = sizeof(int) * ((size_t)c->npair + nsp + 1) + sizeof(int64_t) * 2 * nsp + 64;
    ZV_CHECK(zv_ensure(c, c->b_spmeta, meta));
    if (c->sp_ready) return ZV_OK;
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
    ZV_CHECK(zv_ensure(c, c->b_spe, sizeof(double) * std::max<int64_t>(eu_tot, 1)));
    int64_t *d_puoff = (int64_t *)c->b_spmeta.p;
    int64_t *d_euoff = d_puoff + nsp;
    int *d_comp = (int *)(d_euoff + nsp);
    int *d_coff = d_comp + c->npair;
    ZV_HIPCHECK(c, hipMemcpyAsync(d_puoff, puoff.data(), sizeof(int64_t) * nsp,
                                  hipMemcpyHostToDevice, c->stream));
    ZV_HIPCHECK(c, hipMemcpyAsync(d_euoff, euoff.data(), sizeof(int64_t) * nsp,
                                  hipMemcpyHostToDevice, c->stream));
    ZV_HIPCHECK(c, hipMemcpyAsync(d_comp, c->h_sp_comp.data(), sizeof(int) * c->npair,
                                  hipMemcpyHostToDevice, c->stream));
    ZV_HIPCHECK(c, hipMemcpyAsync(d_coff, c->h_sp_coff.data(), sizeof(int) * (nsp + 1),
                                  hipMemcpyHostToDevice, c->stream));
    SPD sd{d_comp, d_coff, d_puoff, d_euoff, (const double *)c->b_spu.p,
           (const double *)c->b_spe.p};
    {
        PhaseTimer tm(c, PH_PRIM);
        hipLaunchKernelGGL(k_shellpairs, dim3(nsp), dim3(256), 0, c->stream, pp, sd,
                           (double *)c->b_spu.p, (double *)c->b_spe.p);
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
    return SPD{d_comp, d_coff, d_puoff, d_euoff, (const double *)c->b_spu.p,
               (const double *)c->b_spe.p};
}

// Plan and launch the shell kernels of one Schwarz phase over the shell
// quartets sq; the partial sums stay in bpart for zv_shell_scatter.
int zv_shell_run(zv_context *c, const PP &pp, const std::vector<std::pair<int, int>> &sq,
                 DevBuf &bitem, DevBuf &bgroup, DevBuf &bpart, std::vector<char> &hplan,
                 int64_t *ngroups)
{
    Plan pl;
    plan_shell(c, sq, pl);
    *ngroups = (int64_t)pl.groups.size();
    if (pl.items.empty()) return ZV_OK;
    const size_t bi = sizeof(ShellItem) * pl.items.size();
    const size_t bg = sizeof(ShellGroup) * pl.groups.size();
    hplan.resize(bi + bg);
    std::memcpy(hplan.data(), pl.items.data(), bi);
    std::memcpy(hplan.data() + bi, pl.groups.data(), bg);
    ZV_CHECK(zv_ensure(c, bitem, bi));
    ZV_CHECK(zv_ensure(c, bgroup, bg));
    ZV_CHECK(zv_ensure(c, bpart, sizeof(double) * pl.npart));
    ZV_HIPCHECK(c, hipMemcpyAsync(bitem.p, hplan.data(), bi, hipMemcpyHostToDevice, c->stream));
    ZV_HIPCHECK(c, hipMemcpyAsync(bgroup.p, hplan.data() + bi, bg, hipMemcpyHostToDevice,
                                  c->stream));
    const SPD sd = spd_of(c);
    const ShellItem *di = (const ShellItem *)bitem.p;
    for (int k = 0; k < 25; k++) {
        const int n = pl.lo[k + 1] - pl.lo[k];
        if (!n) continue;
        const int LU = k / 5, LL = k % 5;
        const int cls = 36 + 6 * LU + LL;
        c->cls_pq[cls] += pl.units[k];
        c->cls_flops[cls] += pl.flops[k];
        PhaseTimer tm(c, PH_ERI, cls);
        if (!launch_shell(LU, LL, (unsigned)n, c->stream, pp, sd, di + pl.lo[k],
                          (double *)bpart.p))
            return zv_fail(c, ZV_ERR_STATE, "no shell kernel for (%d,%d)", LU, LL);
        ZV_HIPCHECK(c, hipGetLastError());
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
