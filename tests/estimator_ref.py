"""Cast-free estimator written with numpy over the oracle's pair data.

For quartet (P, Q) (P = (i,k) electron 1, Q = (j,l) electron 2) and sample m:
    t_m = J_P(r1) J_Q(r2) / pi0 * (rho_P(1) rho_Q(2) - rhoG_P(1) rhoG_Q(2)) / |u_P(r1) - u_Q(r2)|
(src/two_e_slater.F90:435-450 without the FP32 casts).  Returns the block
average over blocks (reference order) and, per quartet, the scale
sum_m |t_m| / M that bounds the rounding of any summation order.

t_m is evaluated in extended precision (numpy longdouble: the x87 format,
64-bit significand) in the algebraically identical form
    w ((rho1 - rhoG1) rho2 + rhoG1 (rho2 - rhoG2)) / r12.
In plain FP64 the reference formula loses digits to the cancellation of
rho1 rho2 against rhoG1 rhoG2 (the fit is close to the STO): on the 200-AO
workload one quartet in ten then differs from the exact value by more than
1e-12 sum|t|/M, which is the rounding of this reference, not of the code
under test.  Here the reference's own rounding is ~1e-19 relative.
"""
import numpy as np

from slater_stoch_amd import basis as B


def fp64_estimate(oracle, bas, r1, r2, pi0, nbl, npts, split=False, pick=None):
    """With split, also returns sum_m (|w A1 A2| + |w B1 B2|) / (r12 M) per
    quartet, the scale of the two products of the split form
    t = w ((rho1 - rhoG1) rho2 + rhoG1 (rho2 - rhoG2)) / r12.
    pick: reference (kcp, 0-based) positions to evaluate; default all."""
    n = bas.n
    is_, js_, ks_, ls_ = B.quartet_mapping(n)
    P = (is_ * (is_ - 1)) // 2 + ks_ - 1
    Q = (js_ * (js_ - 1)) // 2 + ls_ - 1
    if pick is not None:
        P, Q = P[pick], Q[pick]
    # pair data only for the pairs the quartets use
    used = np.unique(np.concatenate([P, Q]))
    slot = np.full(n * (n + 1) // 2, -1, np.int64)
    slot[used] = np.arange(used.size)
    M = r1.shape[0]
    U1 = np.zeros((used.size, M, 3))
    U2 = np.zeros((used.size, M, 3))
    A = np.zeros((used.size, 4, M))
    J = np.zeros((used.size, 2, M))
    for s, p in enumerate(used):
        i = int(np.floor((np.sqrt(8 * p + 1) - 1) / 2)) + 1
        while i * (i - 1) // 2 > p:
            i -= 1
        while i * (i + 1) // 2 <= p:
            i += 1
        k = p - i * (i - 1) // 2 + 1
        d = oracle.pair_samples(i, k, r1, r2)
        U1[s], U2[s] = d["ut1"], d["ut2"]
        A[s] = [d["rho1"], d["rho2"], d["rhog1"], d["rhog2"]]
        J[s] = [d["jac1"], d["jac2"]]
    P, Q = slot[P], slot[Q]
    mean = np.zeros(P.size)
    scale = np.zeros(P.size)
    scale2 = np.zeros(P.size)
    X = np.longdouble
    U1, U2, A, J, pi0x = U1.astype(X), U2.astype(X), A.astype(X), J.astype(X), pi0.astype(X)
    step = max(1, 100000 // M)
    for s in range(0, P.size, step):
        p, q = P[s:s + step], Q[s:s + step]
        d = U1[p] - U2[q]
        r12 = np.sqrt((d * d).sum(axis=2))
        w = J[p, 0] * J[q, 1] / pi0x[None, :]
        t = w * ((A[p, 0] - A[p, 2]) * A[q, 1] + A[p, 2] * (A[q, 1] - A[q, 3])) / r12
        mean[s:s + step] = t.reshape(-1, nbl, npts).mean(axis=2).mean(axis=1)
        scale[s:s + step] = np.abs(t).mean(axis=1)
        if split:
            s2 = np.abs(w * (A[p, 0] - A[p, 2]) * A[q, 1]) + np.abs(w * A[p, 2] * (A[q, 1] - A[q, 3]))
            scale2[s:s + step] = (s2 / r12).mean(axis=1)
    return (mean, scale, scale2) if split else (mean, scale)
