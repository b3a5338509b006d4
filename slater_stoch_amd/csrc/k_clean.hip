// PSD cleanup of the ERI matrix (src/svd_clean.F90, src/davidson_clean.F90).
#include "zvmc_internal.h"

int zv_run_svd_clean(zv_context *c, int rank, int q, uint64_t seed, double *eri, int *n_removed)
{
    (void)rank; (void)q; (void)seed; (void)eri; (void)n_removed;
    return zv_fail(c, ZV_ERR_STATE, "zv_svd_clean: not built in this version");
}

int zv_run_davidson_clean(zv_context *c, double *eri, int *n_removed)
{
    (void)eri; (void)n_removed;
    return zv_fail(c, ZV_ERR_STATE, "zv_davidson_clean: not built in this version");
}
