// Internal declarations of the ZVMC HIP library (context, device buffers,
// kernel launchers).  Host code and kernels are compiled for gfx950 only.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/zvmc.h"

#define ZV_GMAX 32     // max Gaussians per STO fit handled on the device
#define ZV_NW 40       // walker block of the reference estimator (src/common_data.F90:4)
#define ZV_TILE 64     // pair-table padding (the parity estimator's pair planes)
#define ZV_LMAX_T 4    // largest pair angular momentum with a templated ERI kernel
#define ZV_LPAIR_MAX 8 // largest pair angular momentum supported at all
#define ZV_SHMAX 15    // largest shell (Cartesian components sharing one Gaussian fit)
#define ZV_MAXU 36     // components of the uniform side handled by one shell work item
#define ZV_NSTAT 236   // ERI statistics slots (component and shell kernel variants)

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

// PH_N as a PhaseTimer's phase: a kernel-class span only (launches that
// overlap other timed launches; the phase takes one span around them)
enum Phase { PH_SAMPLE = 0, PH_DENS, PH_EST, PH_STATS, PH_PRIM, PH_ERI, PH_COPY, PH_CLEAN, PH_N };

struct zv_context {
    int device = 0;
    int verbose = 0;
    int64_t tune[ZV_TUNE_N] = {};      // zv_set_tuning values (0 = automatic)
    hipStream_t stream = nullptr;
    // extra streams (default priority) for independent launches that are
    // each too small to fill the device; forked from and joined back into
    // `stream` inside one call
    static constexpr int NSIDE = 4;
    hipStream_t side[NSIDE] = {};
    hipEvent_t ev_fork = nullptr, ev_join[NSIDE] = {};
    std::string err;

    // basis (host)
    int n = 0;
    int npair = 0;
    int npair_pad = 0;
    int64_t nint = 0;
    double a_zv = 0.0;
    std::vector<int> h_npower, h_nsto, h_itype, h_ngauss, h_lao;
    std::vector<double> h_center, h_zeta, h_c1, h_c2, h_g;
    std::vector<int> h_pair_i, h_pair_k;   // 1-based AO indices of pair p
    std::vector<int> h_pair_L;             // Hermite order of pair p
    std::vector<int64_t> h_pp_off;         // primitive-pair slot offset of pair p
    std::vector<int64_t> h_e_off;          // Hermite-coefficient offset of pair p
    int64_t pp_total = 0, e_total = 0;
    // shells: runs of consecutive AOs with the same center, Gaussian fit and
    // order; their pairs share every primitive-pair quantity but E
    int nshell = 0, nsp = 0;
    std::vector<int> h_sh_first;           // first AO (0-based) of each shell, nshell + 1
    std::vector<int> h_sp_coff;            // component offset of shell pair (I >= K), nsp + 1
    std::vector<int> h_sp_comp;            // AO pair index of each component
    std::vector<int> h_sp_L;               // Hermite order of the shell pair

    // basis (device)
    double *d_center = nullptr, *d_zeta = nullptr, *d_c1 = nullptr, *d_c2 = nullptr,
           *d_g = nullptr;
    int *d_npower = nullptr, *d_nsto = nullptr, *d_itype = nullptr, *d_ngauss = nullptr;
    int *d_pair_i = nullptr, *d_pair_k = nullptr;
    double *d_pair_f = nullptr, *d_pair_jf = nullptr, *d_pair_G = nullptr; // G: [3][npair_pad]
    int64_t *d_kbase = nullptr;

    // work buffers, grown on demand
    DevBuf b_r1, b_r2, b_pi0;          // samples of one chunk
    DevBuf b_pd;                       // per (pair, sample) data of one chunk
    DevBuf b_accS, b_accG;             // per quartet block accumulators (triangle order)
    DevBuf b_s1, b_s2;                 // per quartet sums over blocks
    DevBuf b_out, b_out2;              // reference-order outputs
    DevBuf b_blk;                      // optional block means
    // device-resident results in reference (kcp) order: ZV_RESULT_ERI,
    // ZV_RESULT_MEAN, ZV_RESULT_STDERR of the last zv_gauss_eri / zv_mc_blocks
    DevBuf b_res[3];
    bool res_valid[3] = {false, false, false};
    DevBuf b_mpart;                    // estimator partial sums of sample splits 1..n-1
    DevBuf b_pp;                       // primitive-pair data
    DevBuf b_btab;                     // Boys Taylor table (zv_boys_table)
    bool btab_ready = false;
    DevBuf b_ppE;                      // primitive-pair Hermite coefficients
    DevBuf b_ppcnt;                    // kept primitive pairs per AO pair
    DevBuf b_tri;                      // triangle-order ERIs
    DevBuf b_list;                     // class lists
    DevBuf b_part;                     // split partial sums
    DevBuf b_misc;
    bool pp_ready = false;
    // shell-pair data: component tables and the uniform-side copies
    DevBuf b_spmeta;                   // puoff[nsp], euoff[nsp], comp[npair], coff[nsp+1], cpat[npair]
    DevBuf b_spu, b_spe;               // per primitive pair: p,1/p,P,c; E [a][comp][h]
    DevBuf b_spec;                     // lane-side E without structural zeros (k_compact_e)
    std::vector<int> h_sp_np;          // kept primitive pairs per shell pair
    std::vector<int> h_sp_cpat;        // per component: Hermite pattern | nonzero count << 8
    std::vector<int64_t> h_sp_nzcum;   // running sum of the components' nonzero counts
    int64_t sp_eu_total = 0;           // doubles of one uniform-side E layout
    bool sp_ready = false;
    // group-blocked estimator: units of 1, 3 or 6 components of one shell
    // pair per class, tiles per pair of classes (built on first use per basis)
    bool grp_ready = false;
    DevBuf b_gmeta;
    int grp_nunit[3] = {};
    int64_t grp_unit_off[3] = {};     // int offsets in b_gmeta: pos, then sp, per class
    int grp_ntile[9] = {};
    int64_t grp_tile_off[9] = {};     // int2 offsets in b_gmeta
    int64_t grp_possp_off = 0;
    int grp_nsp_pad = 0;
    int grp_tileset = 0;               // tile-shape set the tables were built for
    std::vector<int> h_gmeta;
    std::vector<int> h_pair_ucls, h_pair_uidx; // unit class and index of each AO pair
    // planned shell work of the two Schwarz phases (host copies kept alive for the copies)
    DevBuf b_sitem1, b_sgroup1, b_part1, b_sitem2, b_sgroup2, b_part2, b_plist;
    std::vector<char> h_plan1, h_plan2;
    std::vector<int> h_plist;

    // timings of the last call: events recorded around each launch group
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<int> pend_phase;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pend_ev;
    double t_ms[PH_N] = {};
    int64_t t_n[PH_N] = {};
    std::vector<int> pend_cls;        // ERI class of each pending timer (-1: none)
    // ERI statistics: [0, 36) component kernels 6 LB + LK; [36, 236) shell
    // kernels 36 + 25 v + 5 LU + LL, v = 2 (lane batch - 1) + reduction mode
    double cls_ms[ZV_NSTAT] = {};
    int64_t cls_n[ZV_NSTAT] = {};
    double cls_pq[ZV_NSTAT] = {};     // primitive-quartet evaluations per class
    double cls_flops[ZV_NSTAT] = {};  // algorithmic FP64 flops per class
    std::vector<int> h_cnt;           // kept primitive pairs per AO pair (host copy)
    void *blas = nullptr;             // rocBLAS handle (cleanup GEMMs)
};
void zv_blas_destroy(zv_context *c);
// allocation-free list mode (k_eri.hip): device address of its scratch array
// on the context's device (once per device, at context creation), and the run
int zv_noalloc_setup(zv_context *c);
int zv_run_gauss_list_noalloc(zv_context *c, int64_t nq, const int *qi, const int *qk,
                              const int *qj, const int *ql, double *out);

// ---- helpers (zvmc_api.hip) -------------------------------------------
int zv_fail(zv_context *c, int code, const char *fmt, ...);
int zv_hip(zv_context *c, hipError_t e, const char *what);
int zv_ensure(zv_context *c, DevBuf &b, size_t bytes);

// triangle index of (P, Q), P >= Q (0-based pair indices)
__host__ __device__ inline int64_t zv_tri(int64_t P, int64_t Q) { return P * (P + 1) / 2 + Q; }

// 0-based pair index of (i, k), i >= k, 1-based AO indices
__host__ __device__ inline int zv_pidx(int i, int k) { return (i * (i - 1)) / 2 + (k - 1); }

// kcp (0-based) of canonical quartet (ik|jl) from the base table
__host__ __device__ inline int64_t zv_kcp(const int64_t *kbase, int n, int i, int j, int k, int l)
{
    int start = (k > j ? k : j) + ((l > k && j >= k) ? 1 : 0);
    return kbase[((int64_t)(l - 1) * n + (k - 1)) * n + (j - 1)] + (i - start);
}

// ---- timing ----------------------------------------------------------------
// Records a start event now and a stop event when it goes out of scope; the
// elapsed times are summed per phase by zv_collect_timings after a sync.
struct PhaseTimer {
    zv_context *c;
    Phase ph;
    int cls;
    hipEvent_t a = nullptr, b = nullptr;
    hipStream_t st = nullptr; // the context's stream unless given
    PhaseTimer(zv_context *c_, Phase p, int cls_ = -1, hipStream_t st_ = nullptr);
    ~PhaseTimer();
};
void zv_reset_timings(zv_context *c);
int zv_collect_timings(zv_context *c);

// ---- launchers (k_mc.hip, k_eri.hip, k_clean.hip) ---------------------
int zv_run_mc(zv_context *c, int nbl, int64_t npts, int mode, uint64_t seed, uint64_t stream,
              int precision, const double *r1, const double *r2, const double *pi0,
              double *mean, double *err, double *block_means);
int zv_run_draw(zv_context *c, uint64_t seed, uint64_t stream, int64_t m0, int64_t m,
                double *r1, double *r2, double *pi0);
int zv_run_pair_samples(zv_context *c, int i, int k, int ns, const double *r1, const double *r2,
                        double *ut1, double *ut2, double *rho1, double *rho2, double *rhog1,
                        double *rhog2, double *jac1, double *jac2);
int zv_run_gauss_eri(zv_context *c, double *ijkl, double *precond, int64_t *n_screened);
int zv_run_exact_monocenter(zv_context *c, double *ijkl, double *err, int64_t *n_mono);
int zv_run_one_e(zv_context *c, int nnuc, const double *charge, const double *nuc, double *S,
                 double *V, double *K);
int zv_run_gauss_list(zv_context *c, int64_t nq, const int *qi, const int *qk, const int *qj,
                      const int *ql, double *out);
int zv_run_scatter_ref(zv_context *c, const double *d_tri, double *host_out);
// d_tri (triangle order) -> result slot `which` in kcp order on the device,
// then to host_out unless it is null
int zv_keep_result(zv_context *c, const double *d_tri, int which, double *host_out);
int zv_run_result_max(zv_context *c, int which, double *value);
int zv_run_svd_clean(zv_context *c, int rank, int q, uint64_t seed, double *eri, int *n_removed);
int zv_run_davidson_clean(zv_context *c, double *eri, int *n_removed);
int zv_run_mc_refine(zv_context *c, int nbl, int64_t npts, uint64_t seed, uint64_t stream,
                     int precision, double target, int64_t *n_refined, double *max_err);
int zv_run_rsvd_clean(zv_context *c, int rank, int q, uint64_t seed, int sketch32, double *eri,
                      double *ritz, int *n_removed);
int zv_run_davidson_lowest(zv_context *c, int nev, double tol, int max_iter, int precond,
                           double *eri, double *evals, double *resid, int *n_removed, int *iters);

#define ZV_CHECK(call)                                                                            \
    do {                                                                                          \
        int rc_ = (call);                                                                         \
        if (rc_ != ZV_OK) return rc_;                                                             \
    } while (0)
#define ZV_HIPCHECK(c, call)                                                                      \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) return zv_hip((c), e_, #call);                                      \
    } while (0)
