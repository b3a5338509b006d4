// C-ABI of the ZVMC engine: context, basis upload, quartet indexing, error
// and timing plumbing.  Kernels live in k_mc.hip (sampling, densities,
// estimator, block statistics), k_eri.hip (analytic Gaussian-fit ERIs) and
// k_clean.hip (PSD cleanup).
#include <cstdarg>
#include <cstring>
#include <map>
#include <mutex>

#include "zvmc_internal.h"

// Every entry point that creates, drives or destroys a context holds one
// process-wide lock for the whole call (calls are synchronous), so contexts
// driven from several host threads run one call at a time: a run with three
// host threads, each driving its own context at the same time, hung on
// MI355X, and concurrent contexts are not supported.  Recursive: zv_create
// calls zv_destroy on failure.
static std::recursive_mutex &call_lock()
{
    static std::recursive_mutex m;
    return m;
}
#define ZV_SERIALIZE std::lock_guard<std::recursive_mutex> zv_serial_(call_lock())

int zv_fail(zv_context *c, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

int zv_hip(zv_context *c, hipError_t e, const char *what)
{
    return zv_fail(c, e == hipErrorOutOfMemory ? ZV_ERR_NOMEM : ZV_ERR_HIP, "%s: %s", what,
                   hipGetErrorString(e));
}

int zv_ensure(zv_context *c, DevBuf &b, size_t bytes)
{
    if (bytes <= b.bytes && b.p) return ZV_OK;
    if (b.p) {
        ZV_HIPCHECK(c, hipStreamSynchronize(c->stream));
        ZV_HIPCHECK(c, hipFree(b.p));
        b.p = nullptr;
        b.bytes = 0;
    }
    size_t want = bytes < 256 ? 256 : bytes;
    hipError_t e = hipMalloc(&b.p, want);
    if (e != hipSuccess) {
        b.p = nullptr;
        return zv_fail(c, ZV_ERR_NOMEM, "device allocation of %zu bytes failed: %s", want,
                       hipGetErrorString(e));
    }
    b.bytes = want;
    return ZV_OK;
}

// ---------------------------------------------------------------- timing
PhaseTimer::PhaseTimer(zv_context *c_, Phase p, int cls_, hipStream_t st_)
    : c(c_), ph(p), cls(cls_), st(st_ ? st_ : c_->stream)
{
    auto take = [&]() -> hipEvent_t {
        if (c->ev_used == c->ev_pool.size()) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            c->ev_pool.push_back(e);
        }
        return c->ev_pool[c->ev_used++];
    };
    a = take();
    b = take();
    // timing is best effort: a failed record only drops this measurement
    if (a && hipEventRecord(a, st) != hipSuccess) a = nullptr;
}

PhaseTimer::~PhaseTimer()
{
    if (!a || !b) return;
    if (hipEventRecord(b, st) != hipSuccess) return;
    c->pend_phase.push_back(ph);
    c->pend_cls.push_back(cls);
    c->pend_ev.emplace_back(a, b);
}

void zv_reset_timings(zv_context *c)
{
    for (int p = 0; p < PH_N; p++) {
        c->t_ms[p] = 0.0;
        c->t_n[p] = 0;
    }
    for (int k = 0; k < ZV_NSTAT; k++) {
        c->cls_ms[k] = 0.0;
        c->cls_n[k] = 0;
        c->cls_pq[k] = 0.0;
        c->cls_flops[k] = 0.0;
    }
    c->pend_phase.clear();
    c->pend_cls.clear();
    c->pend_ev.clear();
    c->ev_used = 0;
}

int zv_collect_timings(zv_context *c)
{
    ZV_HIPCHECK(c, hipStreamSynchronize(c->stream));
    for (size_t t = 0; t < c->pend_ev.size(); t++) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c->pend_ev[t].first, c->pend_ev[t].second) == hipSuccess) {
            if (c->pend_phase[t] < PH_N) {
                c->t_ms[c->pend_phase[t]] += ms;
                c->t_n[c->pend_phase[t]] += 1;
            }
            int k = c->pend_cls[t];
            if (k >= 0 && k < ZV_NSTAT) {
                c->cls_ms[k] += ms;
                c->cls_n[k] += 1;
            }
        }
    }
    c->pend_phase.clear();
    c->pend_cls.clear();
    c->pend_ev.clear();
    c->ev_used = 0;
    return ZV_OK;
}

// ---------------------------------------------------------------- quartets
extern "C" int64_t zv_nint(int n)
{
    if (n <= 0) return 0;
    int64_t np = (int64_t)n * (n + 1) / 2;
    return np * (np + 1) / 2;
}

// Base table of the reference quartet order (src/build_mapping_ijkl.F90:16-43):
// for each (l, k, j) the index of the first valid i, which runs over
// [max(k,j), n] minus i == j when l > k.
static std::vector<int64_t> build_kbase(int n)
{
    std::vector<int64_t> kb((size_t)n * n * n, 0);
    int64_t off = 0;
    for (int l = 1; l <= n; l++)
        for (int k = 1; k <= n; k++)
            for (int j = 1; j <= n; j++) {
                kb[((size_t)(l - 1) * n + (k - 1)) * n + (j - 1)] = off;
                if (j < l) continue;
                int start = (k > j ? k : j) + ((l > k && j >= k) ? 1 : 0);
                if (start <= n) off += n - start + 1;
            }
    return kb;
}

// Writes the quartets in the reference loop order and places each through the
// same base table the device scatter uses, so this also checks zv_kcp.
extern "C" int zv_build_mapping(int n, int *is, int *js, int *ks, int *ls)
{
    if (n <= 0 || n > 1024 || !is || !js || !ks || !ls) return ZV_ERR_ARG;
    std::vector<int64_t> kb = build_kbase(n);
    const int64_t nint = zv_nint(n);
    int64_t q = 0;
    for (int l = 1; l <= n; l++)
        for (int k = 1; k <= n; k++)
            for (int j = l; j <= n; j++) {
                int start = (k > j ? k : j) + ((l > k && j >= k) ? 1 : 0);
                for (int i = start; i <= n; i++) {
                    int64_t at = zv_kcp(kb.data(), n, i, j, k, l);
                    if (at != q || at >= nint) return ZV_ERR_STATE;
                    is[at] = i;
                    js[at] = j;
                    ks[at] = k;
                    ls[at] = l;
                    q++;
                }
            }
    return q == nint ? ZV_OK : ZV_ERR_STATE;
}

// ---------------------------------------------------------------- context
// The streams of every context on a device: made by the first context and
// kept until the process ends.  With streams of their own, contexts created
// after the first in a process ran the estimator 9-15 % slower on MI355X
// (their streams land on other hardware queues than the first context's);
// calls run one at a time (call_lock) and each ends with its streams
// drained, so contexts can share them.  Guarded by call_lock.
struct DevStreams {
    hipStream_t main = nullptr;
    hipStream_t side[zv_context::NSIDE] = {};
};
static std::map<int, DevStreams> g_streams;

static bool device_streams(int device, DevStreams &out)
{
    auto it = g_streams.find(device);
    if (it != g_streams.end()) {
        out = it->second;
        return true;
    }
    DevStreams s;
    bool ok = hipStreamCreate(&s.main) == hipSuccess;
    for (int k = 0; ok && k < zv_context::NSIDE; k++) ok = hipStreamCreate(&s.side[k]) == hipSuccess;
    if (!ok) {
        if (s.main) (void)hipStreamDestroy(s.main);
        for (hipStream_t x : s.side)
            if (x) (void)hipStreamDestroy(x);
        return false;
    }
    g_streams[device] = s;
    out = s;
    return true;
}

extern "C" int zv_create(int device, zv_context **out)
{
    ZV_SERIALIZE;
    if (!out) return ZV_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return ZV_ERR_NODEV;
    if (device < 0 || device >= ndev) return ZV_ERR_NODEV;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return ZV_ERR_NODEV;
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) return ZV_ERR_NODEV;
    zv_context *c = new zv_context();
    c->device = device;
    DevStreams ds;
    bool ok = hipSetDevice(device) == hipSuccess && device_streams(device, ds) &&
              hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) == hipSuccess;
    if (ok) {
        c->stream = ds.main;
        for (int k = 0; k < zv_context::NSIDE; k++) c->side[k] = ds.side[k];
    }
    for (int k = 0; ok && k < zv_context::NSIDE; k++)
        ok = hipEventCreateWithFlags(&c->ev_join[k], hipEventDisableTiming) == hipSuccess;
    if (ok) ok = zv_noalloc_setup(c) == ZV_OK;
    if (!ok) {
        zv_destroy(c);
        return ZV_ERR_HIP;
    }
    *out = c;
    return ZV_OK;
}

static void free_basis(zv_context *c)
{
    void *ptrs[] = {c->d_center, c->d_zeta, c->d_c1, c->d_c2, c->d_g, c->d_npower, c->d_nsto,
                    c->d_itype, c->d_ngauss, c->d_pair_i, c->d_pair_k, c->d_pair_f,
                    c->d_pair_jf, c->d_pair_G, c->d_kbase};
    // teardown is best effort: a failing free has nothing left to report to
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    c->d_center = c->d_zeta = c->d_c1 = c->d_c2 = c->d_g = nullptr;
    c->d_npower = c->d_nsto = c->d_itype = c->d_ngauss = nullptr;
    c->d_pair_i = c->d_pair_k = nullptr;
    c->d_pair_f = c->d_pair_jf = c->d_pair_G = nullptr;
    c->d_kbase = nullptr;
    c->pp_ready = false;
    c->sp_ready = false;
    c->grp_ready = false;
    for (bool &v : c->res_valid) v = false;
}

extern "C" int zv_destroy(zv_context *c)
{
    ZV_SERIALIZE;
    if (!c) return ZV_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    free_basis(c);
    zv_blas_destroy(c);
    DevBuf *bufs[] = {&c->b_r1, &c->b_r2, &c->b_pi0, &c->b_pd, &c->b_accS, &c->b_accG,
                      &c->b_s1, &c->b_s2, &c->b_out, &c->b_out2, &c->b_blk, &c->b_pp, &c->b_btab,
                      &c->b_ppE, &c->b_ppcnt, &c->b_tri, &c->b_list, &c->b_part, &c->b_misc,
                      &c->b_spmeta, &c->b_spu, &c->b_spe, &c->b_spec, &c->b_sitem1, &c->b_sgroup1,
                      &c->b_part1, &c->b_sitem2, &c->b_sgroup2, &c->b_part2, &c->b_plist,
                      &c->b_gmeta, &c->b_res[0], &c->b_res[1], &c->b_res[2], &c->b_mpart};
    for (DevBuf *b : bufs)
        if (b->p) (void)hipFree(b->p);
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    // the streams belong to the device (device_streams), not to the context
    for (int k = 0; k < zv_context::NSIDE; k++)
        if (c->ev_join[k]) (void)hipEventDestroy(c->ev_join[k]);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    delete c;
    return ZV_OK;
}

extern "C" int zv_last_error(const zv_context *c, char *buf, int len)
{
    if (!buf || len <= 0) return ZV_ERR_ARG;
    const char *m = c ? c->err.c_str() : "no context";
    snprintf(buf, len, "%s", m);
    return ZV_OK;
}

extern "C" int zv_set_verbose(zv_context *c, int level)
{
    if (!c) return ZV_ERR_ARG;
    c->verbose = level;
    return ZV_OK;
}

extern "C" int zv_set_tuning(zv_context *c, int key, int64_t value)
{
    ZV_SERIALIZE;
    if (!c) return ZV_ERR_ARG;
    if (key < 0 || key > ZV_TUNE_EST_TILES)
        return zv_fail(c, ZV_ERR_ARG, "unknown tuning key %d", key);
    if (value < 0) return zv_fail(c, ZV_ERR_ARG, "tuning value %lld < 0", (long long)value);
    if (key == ZV_TUNE_EST_TILES && value > 1)
        return zv_fail(c, ZV_ERR_ARG, "tile-shape set %lld: 0 or 1", (long long)value);
    c->tune[key] = value;
    return ZV_OK;
}

extern "C" int zv_last_timings(const zv_context *c, double *ms, int64_t *launches)
{
    if (!c) return ZV_ERR_ARG;
    for (int p = 0; p < PH_N; p++) {
        if (ms) ms[p] = c->t_ms[p];
        if (launches) launches[p] = c->t_n[p];
    }
    return ZV_OK;
}

extern "C" int zv_last_eri_stats(const zv_context *c, double *ms, int64_t *launches,
                                 double *units, double *flops)
{
    if (!c) return ZV_ERR_ARG;
    for (int k = 0; k < ZV_NSTAT; k++) {
        if (ms) ms[k] = c->cls_ms[k];
        if (launches) launches[k] = c->cls_n[k];
        if (units) units[k] = c->cls_pq[k];
        if (flops) flops[k] = c->cls_flops[k];
    }
    return ZV_OK;
}

template <typename T>
static int upload(zv_context *c, T **dst, const T *src, size_t count)
{
    ZV_HIPCHECK(c, hipMalloc((void **)dst, sizeof(T) * (count ? count : 1)));
    if (count) ZV_HIPCHECK(c, hipMemcpy(*dst, src, sizeof(T) * count, hipMemcpyHostToDevice));
    return ZV_OK;
}

static int hermite_count(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }

static int set_basis_tables(zv_context *c, int n, const double *center, const int *npower,
                            const int *n_sto, const int *i_type, const double *zeta,
                            const int *n_gauss, int ldg, const double *c_gauss,
                            const double *g_gauss, double a_zv);

extern "C" int zv_set_basis(zv_context *c, int n, const double *center, const int *npower,
                            const int *n_sto, const int *i_type, const double *zeta,
                            const int *n_gauss, int ldg, const double *c_gauss,
                            const double *g_gauss, double a_zv)
{
    ZV_SERIALIZE;
    if (!c) return ZV_ERR_ARG;
    if (n <= 0 || n > 1024) return zv_fail(c, ZV_ERR_ARG, "nbasis %d out of range 1..1024", n);
    if (!center || !npower || !n_sto || !i_type || !zeta || !n_gauss || !c_gauss || !g_gauss)
        return zv_fail(c, ZV_ERR_ARG, "null basis array");
    if (!(a_zv > 0.0)) return zv_fail(c, ZV_ERR_ARG, "a_ZV must be positive");
    for (int i = 0; i < n; i++) {
        if (n_gauss[i] < 1 || n_gauss[i] > ZV_GMAX || n_gauss[i] > ldg)
            return zv_fail(c, ZV_ERR_ARG, "n_gauss(%d)=%d outside 1..min(%d,ldg=%d)", i + 1,
                           n_gauss[i], ZV_GMAX, ldg);
        if (i_type[i] == 3)
            return zv_fail(c, ZV_ERR_ARG, "i_type=3 not yet coded (orbital %d)", i + 1);
        if (i_type[i] != 1 && i_type[i] != 2)
            return zv_fail(c, ZV_ERR_ARG, "i_type(%d)=%d not supported", i + 1, i_type[i]);
        if (!(zeta[i] > 0.0)) return zv_fail(c, ZV_ERR_ARG, "zeta(%d) must be positive", i + 1);
        for (int d = 0; d < 3; d++)
            if (npower[3 * i + d] < 0 || npower[3 * i + d] + (i_type[i] == 2 ? 2 : 0) > 4)
                return zv_fail(c, ZV_ERR_ARG,
                               "orbital %d: power per axis (incl. r^2 of i_type 2) above 4", i + 1);
        if (n_sto[i] < 0 || n_sto[i] > 4)
            return zv_fail(c, ZV_ERR_ARG, "n_STO(%d)=%d outside 0..4", i + 1, n_sto[i]);
    }
    // the largest pair order is that of the AO with the largest order paired
    // with itself; checked before the old basis is released
    int lmax = 0, imax = 0;
    for (int i = 0; i < n; i++) {
        const int l = npower[3 * i] + npower[3 * i + 1] + npower[3 * i + 2] +
                      (i_type[i] == 2 ? 2 : 0);
        if (l > lmax) lmax = l, imax = i;
    }
    if (2 * lmax > ZV_LPAIR_MAX)
        return zv_fail(c, ZV_ERR_ARG, "pair (%d,%d) angular momentum %d > %d", imax + 1, imax + 1,
                       2 * lmax, ZV_LPAIR_MAX);
    ZV_HIPCHECK(c, hipSetDevice(c->device));
    ZV_HIPCHECK(c, hipStreamSynchronize(c->stream));
    free_basis(c);
    // from here on a failure leaves no basis: later calls get ZV_ERR_STATE
    // instead of running on a half-built one
    const int rc = set_basis_tables(c, n, center, npower, n_sto, i_type, zeta, n_gauss, ldg,
                                    c_gauss, g_gauss, a_zv);
    if (rc != ZV_OK) {
        free_basis(c);
        c->n = 0;
        c->npair = c->npair_pad = 0;
        c->nint = 0;
    }
    return rc;
}

static int set_basis_tables(zv_context *c, int n, const double *center, const int *npower,
                            const int *n_sto, const int *i_type, const double *zeta,
                            const int *n_gauss, int ldg, const double *c_gauss,
                            const double *g_gauss, double a_zv)
{
    c->n = n;
    c->a_zv = a_zv;
    c->npair = n * (n + 1) / 2;
    c->npair_pad = (c->npair + ZV_TILE - 1) / ZV_TILE * ZV_TILE;
    c->nint = zv_nint(n);
    c->h_center.assign(center, center + 3 * n);
    c->h_npower.assign(npower, npower + 3 * n);
    c->h_nsto.assign(n_sto, n_sto + n);
    c->h_itype.assign(i_type, i_type + n);
    c->h_zeta.assign(zeta, zeta + n);
    c->h_ngauss.assign(n_gauss, n_gauss + n);
    c->h_c1.assign((size_t)n * ZV_GMAX, 0.0);
    c->h_c2.assign((size_t)n * ZV_GMAX, 0.0);
    c->h_g.assign((size_t)n * ZV_GMAX, 0.0);
    c->h_lao.assign(n, 0);
    for (int i = 0; i < n; i++) {
        for (int m = 0; m < n_gauss[i]; m++) {
            c->h_c1[(size_t)i * ZV_GMAX + m] = c_gauss[(size_t)ldg * i + m];
            c->h_c2[(size_t)i * ZV_GMAX + m] = c_gauss[(size_t)ldg * n + (size_t)ldg * i + m];
            c->h_g[(size_t)i * ZV_GMAX + m] = g_gauss[(size_t)ldg * i + m];
        }
        c->h_lao[i] = npower[3 * i] + npower[3 * i + 1] + npower[3 * i + 2] +
                      (i_type[i] == 2 ? 2 : 0);
    }
    // shells: consecutive AOs with identical center, fit and order
    auto same_fit = [&](int a, int b) {
        if (i_type[a] != i_type[b] || n_gauss[a] != n_gauss[b] || c->h_lao[a] != c->h_lao[b])
            return false;
        for (int d = 0; d < 3; d++)
            if (center[3 * a + d] != center[3 * b + d]) return false;
        for (int m = 0; m < n_gauss[a]; m++) {
            size_t ia = (size_t)a * ZV_GMAX + m, ib = (size_t)b * ZV_GMAX + m;
            if (c->h_g[ia] != c->h_g[ib] || c->h_c1[ia] != c->h_c1[ib] || c->h_c2[ia] != c->h_c2[ib])
                return false;
        }
        return true;
    };
    c->h_sh_first.clear();
    for (int i = 0; i < n; i++)
        if (i == 0 || i - c->h_sh_first.back() >= ZV_SHMAX || !same_fit(i - 1, i))
            c->h_sh_first.push_back(i);
    c->nshell = (int)c->h_sh_first.size();
    c->h_sh_first.push_back(n);
    c->nsp = c->nshell * (c->nshell + 1) / 2;
    c->h_sp_coff.assign(c->nsp + 1, 0);
    c->h_sp_L.assign(c->nsp, 0);
    c->h_sp_comp.clear();
    for (int I = 0; I < c->nshell; I++)
        for (int K = 0; K <= I; K++) {
            int sp = I * (I + 1) / 2 + K;
            c->h_sp_coff[sp] = (int)c->h_sp_comp.size();
            c->h_sp_L[sp] = c->h_lao[c->h_sh_first[I]] + c->h_lao[c->h_sh_first[K]];
            for (int i = c->h_sh_first[I]; i < c->h_sh_first[I + 1]; i++)
                for (int k = c->h_sh_first[K]; k < c->h_sh_first[K + 1] && k <= i; k++)
                    c->h_sp_comp.push_back(zv_pidx(i + 1, k + 1));
        }
    c->h_sp_coff[c->nsp] = (int)c->h_sp_comp.size();
    // pair table, lexicographic (i, k), i >= k
    int np = c->npair, npp = c->npair_pad;
    c->h_pair_i.assign(np, 0);
    c->h_pair_k.assign(np, 0);
    c->h_pair_L.assign(np, 0);
    c->h_pp_off.assign(np + 1, 0);
    c->h_e_off.assign(np + 1, 0);
    std::vector<int> pi(npp, 1), pk(npp, 1);
    std::vector<double> pf(npp, 0.0), pjf(npp, 0.0), pG(3 * (size_t)npp, 0.0);
    for (int i = 1; i <= n; i++)
        for (int k = 1; k <= i; k++) {
            int p = zv_pidx(i, k);
            c->h_pair_i[p] = i;
            c->h_pair_k[p] = k;
            int L = c->h_lao[i - 1] + c->h_lao[k - 1];
            if (L > ZV_LPAIR_MAX)
                return zv_fail(c, ZV_ERR_ARG, "pair (%d,%d) angular momentum %d > %d", i, k, L,
                               ZV_LPAIR_MAX);
            c->h_pair_L[p] = L;
            int64_t cap = (int64_t)n_gauss[i - 1] * n_gauss[k - 1];
            c->h_pp_off[p + 1] = cap;
            c->h_e_off[p + 1] = cap * hermite_count(L);
            double zi = zeta[i - 1], zk = zeta[k - 1];
            pi[p] = i;
            pk[p] = k;
            // src/two_e_slater.F90:344,406 and src/compute_jacobian_and_pi0.F90:21
            pf[p] = 0.5 * a_zv / (zi + zk);
            pjf[p] = a_zv / (zi + zk);
            for (int d = 0; d < 3; d++)
                pG[(size_t)d * npp + p] =
                    (zi * center[3 * (i - 1) + d] + zk * center[3 * (k - 1) + d]) / (zi + zk);
        }
    for (int p = 0; p < np; p++) {
        c->h_pp_off[p + 1] += c->h_pp_off[p];
        c->h_e_off[p + 1] += c->h_e_off[p];
    }
    c->pp_total = c->h_pp_off[np];
    c->e_total = c->h_e_off[np];
    // padded pairs: far away, zero weight
    for (int p = np; p < npp; p++)
        for (int d = 0; d < 3; d++) pG[(size_t)d * npp + p] = 1.0e3;

    std::vector<int64_t> kb = build_kbase(n);
    ZV_CHECK(upload(c, &c->d_center, c->h_center.data(), c->h_center.size()));
    ZV_CHECK(upload(c, &c->d_zeta, c->h_zeta.data(), c->h_zeta.size()));
    ZV_CHECK(upload(c, &c->d_c1, c->h_c1.data(), c->h_c1.size()));
    ZV_CHECK(upload(c, &c->d_c2, c->h_c2.data(), c->h_c2.size()));
    ZV_CHECK(upload(c, &c->d_g, c->h_g.data(), c->h_g.size()));
    ZV_CHECK(upload(c, &c->d_npower, c->h_npower.data(), c->h_npower.size()));
    ZV_CHECK(upload(c, &c->d_nsto, c->h_nsto.data(), c->h_nsto.size()));
    ZV_CHECK(upload(c, &c->d_itype, c->h_itype.data(), c->h_itype.size()));
    ZV_CHECK(upload(c, &c->d_ngauss, c->h_ngauss.data(), c->h_ngauss.size()));
    ZV_CHECK(upload(c, &c->d_pair_i, pi.data(), pi.size()));
    ZV_CHECK(upload(c, &c->d_pair_k, pk.data(), pk.size()));
    ZV_CHECK(upload(c, &c->d_pair_f, pf.data(), pf.size()));
    ZV_CHECK(upload(c, &c->d_pair_jf, pjf.data(), pjf.size()));
    ZV_CHECK(upload(c, &c->d_pair_G, pG.data(), pG.size()));
    ZV_CHECK(upload(c, &c->d_kbase, kb.data(), kb.size()));
    return ZV_OK;
}

// ---------------------------------------------------------------- wrappers
#define ZV_NEED_BASIS(c)                                                                          \
    do {                                                                                          \
        if (!(c)) return ZV_ERR_ARG;                                                              \
        if ((c)->n <= 0) return zv_fail((c), ZV_ERR_STATE, "zv_set_basis has not been called");  \
        ZV_HIPCHECK((c), hipSetDevice((c)->device));                                             \
        zv_reset_timings(c);                                                                      \
    } while (0)

extern "C" int zv_gauss_eri(zv_context *c, double *ijkl, double *precond, int64_t *n_screened)
{
    ZV_SERIALIZE;
    ZV_NEED_BASIS(c);
    int rc = zv_run_gauss_eri(c, ijkl, precond, n_screened);
    if (rc == ZV_OK) rc = zv_collect_timings(c);
    return rc;
}

extern "C" int zv_exact_monocenter(zv_context *c, double *ijkl, double *err, int64_t *n_mono)
{
    ZV_SERIALIZE;
    ZV_NEED_BASIS(c);
    if (!ijkl) return zv_fail(c, ZV_ERR_ARG, "null ijkl");
    int rc = zv_run_exact_monocenter(c, ijkl, err, n_mono);
    if (rc == ZV_OK) rc = zv_collect_timings(c);
    return rc;
}

extern "C" int zv_one_e(zv_context *c, int nnuc, const double *charge, const double *nuc,
                        double *S, double *V, double *K)
{
    ZV_SERIALIZE;
    ZV_NEED_BASIS(c);
    if (nnuc < 1 || !charge || !nuc) return zv_fail(c, ZV_ERR_ARG, "no nuclei");
    if (!S || !V || !K) return zv_fail(c, ZV_ERR_ARG, "null output");
    int rc = zv_run_one_e(c, nnuc, charge, nuc, S, V, K);
    if (rc == ZV_OK) rc = zv_collect_timings(c);
    return rc;
}

extern "C" int zv_gauss_ijkl_list(zv_context *c, int64_t nq, const int *qi, const int *qk,
                                  const int *qj, const int *ql, double *out)
{
    ZV_SERIALIZE;
    ZV_NEED_BASIS(c);
    if (nq < 0 || (nq > 0 && (!qi || !qk || !qj || !ql || !out)))
        return zv_fail(c, ZV_ERR_ARG, "bad quartet list");
    for (int64_t t = 0; t < nq; t++) {
        if (qi[t] < 1 || qi[t] > c->n || qk[t] < 1 || qk[t] > c->n || qj[t] < 1 ||
            qj[t] > c->n || ql[t] < 1 || ql[t] > c->n)
            return zv_fail(c, ZV_ERR_ARG, "quartet %lld has an orbital index outside 1..%d",
                           (long long)t + 1, c->n);
    }
    if (nq == 0) return ZV_OK;
    int rc = zv_run_gauss_list(c, nq, qi, qk, qj, ql, out);
    if (rc == ZV_OK) rc = zv_collect_timings(c);
    return rc;
}

extern "C" int zv_gauss_ijkl_list_noalloc(zv_context *c, int64_t nq, const int *qi,
                                          const int *qk, const int *qj, const int *ql,
                                          double *out)
{
    ZV_SERIALIZE;
    ZV_NEED_BASIS(c);
    if (nq < 0 || (nq > 0 && (!qi || !qk || !qj || !ql || !out)))
        return zv_fail(c, ZV_ERR_ARG, "bad quartet list");
    if (nq > ZV_NOALLOC_MAXQ)
        return zv_fail(c, ZV_ERR_ARG, "%lld quartets: the allocation-free list holds at most %d",
                       (long long)nq, ZV_NOALLOC_MAXQ);
    for (int64_t t = 0; t < nq; t++) {
        if (qi[t] < 1 || qi[t] > c->n || qk[t] < 1 || qk[t] > c->n || qj[t] < 1 ||
            qj[t] > c->n || ql[t] < 1 || ql[t] > c->n)
            return zv_fail(c, ZV_ERR_ARG, "quartet %lld has an orbital index outside 1..%d",
                           (long long)t + 1, c->n);
    }
    if (nq == 0) return ZV_OK;
    return zv_run_gauss_list_noalloc(c, nq, qi, qk, qj, ql, out);
}

extern "C" int zv_mc_blocks(zv_context *c, int nbl, int64_t npts, int mode, uint64_t seed,
                            uint64_t stream, int precision, const double *r1, const double *r2,
                            const double *pi0, double *mean, double *err, double *block_means)
{
    ZV_SERIALIZE;
    ZV_NEED_BASIS(c);
    if (nbl < 1) return zv_fail(c, ZV_ERR_ARG, "nbl must be >= 1");
    if (npts < ZV_NW || npts % ZV_NW != 0)
        return zv_fail(c, ZV_ERR_ARG, "npts_two_elec must be a multiple of nw");
    if (mode != ZV_SAMPLES_FIXED && mode != ZV_SAMPLES_PHILOX)
        return zv_fail(c, ZV_ERR_ARG, "unknown sampling mode %d", mode);
    if (mode == ZV_SAMPLES_FIXED && (!r1 || !r2 || !pi0))
        return zv_fail(c, ZV_ERR_ARG, "fixed-sample mode needs r1, r2, pi0");
    if (precision < ZV_PREC_FP64_PARITY || precision > ZV_PREC_BF16)
        return zv_fail(c, ZV_ERR_ARG, "unknown precision %d", precision);
    int rc = zv_run_mc(c, nbl, npts, mode, seed, stream, precision, r1, r2, pi0, mean, err,
                       block_means);
    if (rc == ZV_OK) rc = zv_collect_timings(c);
    return rc;
}

extern "C" int zv_get_result(zv_context *c, int which, double *host)
{
    // a query: leaves the timings of the last computing call as they are
    ZV_SERIALIZE;
    if (!c) return ZV_ERR_ARG;
    if (c->n <= 0) return zv_fail(c, ZV_ERR_STATE, "zv_set_basis has not been called");
    ZV_HIPCHECK(c, hipSetDevice(c->device));
    if (which < ZV_RESULT_ERI || which > ZV_RESULT_STDERR || !host)
        return zv_fail(c, ZV_ERR_ARG, "bad result index %d or null output", which);
    if (!c->res_valid[which])
        return zv_fail(c, ZV_ERR_STATE, "result %d has not been computed for this basis", which);
    ZV_HIPCHECK(c, hipMemcpy(host, c->b_res[which].p, sizeof(double) * c->nint,
                             hipMemcpyDeviceToHost));
    return ZV_OK;
}

extern "C" int zv_result_max(zv_context *c, int which, double *value)
{
    // a query: leaves the timings of the last computing call as they are
    ZV_SERIALIZE;
    if (!c) return ZV_ERR_ARG;
    if (c->n <= 0) return zv_fail(c, ZV_ERR_STATE, "zv_set_basis has not been called");
    ZV_HIPCHECK(c, hipSetDevice(c->device));
    if (which < ZV_RESULT_ERI || which > ZV_RESULT_STDERR || !value)
        return zv_fail(c, ZV_ERR_ARG, "bad result index %d or null output", which);
    if (!c->res_valid[which])
        return zv_fail(c, ZV_ERR_STATE, "result %d has not been computed for this basis", which);
    return zv_run_result_max(c, which, value);
}

extern "C" int zv_draw_samples(zv_context *c, uint64_t seed, uint64_t stream, int64_t m0,
                               int64_t m, double *r1, double *r2, double *pi0)
{
    ZV_SERIALIZE;
    if (!c) return ZV_ERR_ARG;
    ZV_HIPCHECK(c, hipSetDevice(c->device));
    zv_reset_timings(c);
    if (m < 0 || m0 < 0 || (m > 0 && (!r1 || !r2 || !pi0)))
        return zv_fail(c, ZV_ERR_ARG, "bad sample range or null output");
    if (m == 0) return ZV_OK;
    int rc = zv_run_draw(c, seed, stream, m0, m, r1, r2, pi0);
    if (rc == ZV_OK) rc = zv_collect_timings(c);
    return rc;
}

extern "C" int zv_pair_samples(zv_context *c, int i, int k, int ns, const double *r1,
                               const double *r2, double *ut1, double *ut2, double *rho1,
                               double *rho2, double *rhog1, double *rhog2, double *jac1,
                               double *jac2)
{
    ZV_SERIALIZE;
    ZV_NEED_BASIS(c);
    if (i < 1 || i > c->n || k < 1 || k > i)
        return zv_fail(c, ZV_ERR_ARG, "pair (%d,%d) must satisfy 1 <= k <= i <= %d", i, k, c->n);
    if (ns < 0 || (ns > 0 && (!r1 || !r2 || !ut1 || !ut2 || !rho1 || !rho2 || !rhog1 ||
                              !rhog2 || !jac1 || !jac2)))
        return zv_fail(c, ZV_ERR_ARG, "bad sample count or null array");
    if (ns == 0) return ZV_OK;
    int rc = zv_run_pair_samples(c, i, k, ns, r1, r2, ut1, ut2, rho1, rho2, rhog1, rhog2, jac1,
                                 jac2);
    if (rc == ZV_OK) rc = zv_collect_timings(c);
    return rc;
}

extern "C" int zv_svd_clean(zv_context *c, int rank, int q, uint64_t seed, double *eri,
                            int *n_removed)
{
    ZV_SERIALIZE;
    ZV_NEED_BASIS(c);
    if (!eri) return zv_fail(c, ZV_ERR_ARG, "null eri");
    if (rank < 1 || q < 0) return zv_fail(c, ZV_ERR_ARG, "rank >= 1 and q >= 0 required");
    int rc = zv_run_svd_clean(c, rank, q, seed, eri, n_removed);
    if (rc == ZV_OK) rc = zv_collect_timings(c);
    return rc;
}

extern "C" int zv_davidson_clean(zv_context *c, double *eri, int *n_removed)
{
    ZV_SERIALIZE;
    ZV_NEED_BASIS(c);
    if (!eri) return zv_fail(c, ZV_ERR_ARG, "null eri");
    int rc = zv_run_davidson_clean(c, eri, n_removed);
    if (rc == ZV_OK) rc = zv_collect_timings(c);
    return rc;
}

extern "C" int zv_mc_refine(zv_context *c, int nbl, int64_t npts, uint64_t seed,
                            uint64_t stream, int precision, double target, int64_t *n_refined,
                            double *max_err)
{
    ZV_SERIALIZE;
    ZV_NEED_BASIS(c);
    if (nbl < 2) return zv_fail(c, ZV_ERR_ARG, "nbl must be >= 2");
    if (npts < ZV_NW || npts % ZV_NW != 0)
        return zv_fail(c, ZV_ERR_ARG, "npts_two_elec must be a multiple of nw");
    if (precision != ZV_PREC_FP64 && precision != ZV_PREC_FP32 && precision != ZV_PREC_BF16)
        return zv_fail(c, ZV_ERR_ARG, "zv_mc_refine runs the group estimator (FP64, FP32, BF16)");
    if (!(target > 0.0)) return zv_fail(c, ZV_ERR_ARG, "target must be positive");
    int rc = zv_run_mc_refine(c, nbl, npts, seed, stream, precision, target, n_refined, max_err);
    if (rc == ZV_OK) rc = zv_collect_timings(c);
    return rc;
}

extern "C" int zv_rsvd_clean(zv_context *c, int rank, int q, uint64_t seed, int sketch,
                             double *eri, double *ritz, int *n_removed)
{
    ZV_SERIALIZE;
    ZV_NEED_BASIS(c);
    if (!eri) return zv_fail(c, ZV_ERR_ARG, "null eri");
    if (rank < 1 || q < 0) return zv_fail(c, ZV_ERR_ARG, "rank >= 1 and q >= 0 required");
    if (sketch != ZV_SKETCH_FP64 && sketch != ZV_SKETCH_FP32)
        return zv_fail(c, ZV_ERR_ARG, "unknown sketch precision %d", sketch);
    int rc = zv_run_rsvd_clean(c, rank, q, seed, sketch == ZV_SKETCH_FP32, eri, ritz, n_removed);
    if (rc == ZV_OK) rc = zv_collect_timings(c);
    return rc;
}

extern "C" int zv_davidson_lowest(zv_context *c, int nev, double tol, int max_iter, int precond,
                                  double *eri, double *evals, double *resid, int *n_removed,
                                  int *iters)
{
    ZV_SERIALIZE;
    ZV_NEED_BASIS(c);
    if (!eri) return zv_fail(c, ZV_ERR_ARG, "null eri");
    if (nev < 1 || nev > 64) return zv_fail(c, ZV_ERR_ARG, "nev %d outside 1..64", nev);
    if (!(tol > 0.0) || max_iter < 1)
        return zv_fail(c, ZV_ERR_ARG, "tol > 0 and max_iter >= 1 required");
    if (precond != ZV_DAVIDSON_PLAIN && precond != ZV_DAVIDSON_SHIFT_INVERT)
        return zv_fail(c, ZV_ERR_ARG, "unknown Davidson preconditioner %d", precond);
    int rc = zv_run_davidson_lowest(c, nev, tol, max_iter, precond, eri, evals, resid, n_removed,
                                    iters);
    if (rc == ZV_OK) rc = zv_collect_timings(c);
    return rc;
}
