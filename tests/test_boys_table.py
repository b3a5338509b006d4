"""The tabulated Boys scheme of the shell kernels (boys_table in
csrc/eri_device.h) checked on the CPU against the Boys function summed in
extended precision: a 15-term Taylor expansion of F_L about the nearest
grid point T_i = i/2 (|d| <= 1/4), then the downward recursion
F_m = (2T F_{m+1} + exp(-T)) / (2m+1), for T < 36 and L <= 8.  The kernel
itself is pinned on the GPU (shell against component kernels at 1e-11, the
reference goldens at 1e-10); this pins the method: relative error below
2e-15 for every order, and the table's closed forms F_m(0) = 1/(2m+1)."""
import numpy as np

LD = np.longdouble
K, N, LMAX = 14, 73, 8


def series(m, T):
    """F_m(T) = exp(-T) sum_j (2T)^j / ((2m+1)(2m+3)...(2m+2j+1)), all
    terms positive, in extended precision."""
    T = LD(T)
    term = LD(1) / (2 * m + 1)
    s = term
    j = 0
    while True:
        j += 1
        term = term * 2 * T / (2 * m + 2 * j + 1)
        s += term
        if term < s * LD(1e-22):
            return s * np.exp(-T)


def table():
    F = np.array([[series(m, 0.5 * i) for i in range(N)] for m in range(LMAX + K + 1)],
                 dtype=LD)
    fact = np.cumprod([LD(1)] + [LD(k) for k in range(1, K + 1)])
    t = np.zeros((LMAX + 1, K + 1, N))
    for L in range(LMAX + 1):
        for k in range(K + 1):
            t[L, k] = ((-1) ** k * F[L + k] / fact[k]).astype(np.float64)
    return t


def boys_table(t, L, T):
    i = int(T * 2.0 + 0.5)
    d = T - 0.5 * i
    s = t[L, K, i]
    for k in range(K - 1, -1, -1):
        s = s * d + t[L, k, i]
    F = [0.0] * (L + 1)
    F[L] = s
    e = np.exp(-T)
    for m in range(L - 1, -1, -1):
        F[m] = (2.0 * T * F[m + 1] + e) / (2 * m + 1)
    return F


def test_table_scheme_against_extended_precision_series():
    t = table()
    # closed form at T = 0: F_m(0) = 1 / (2m + 1)
    for L in range(LMAX + 1):
        assert abs(t[L, 0, 0] - 1.0 / (2 * L + 1)) <= 1e-16
    rng = np.random.default_rng(4)
    Ts = np.concatenate([rng.uniform(0.0, 36.0, 150), [0.0, 0.2499, 0.25, 1.0, 35.99]])
    worst = 0.0
    for T in Ts:
        for L in (0, 1, 4, 8):
            F = boys_table(t, L, float(T))
            for m in range(L + 1):
                ref = series(m, T)
                worst = max(worst, float(abs(LD(F[m]) - ref) / ref))
    assert worst < 2e-15, worst
