"""Algorithmic work model of one energy+gradient evaluation.

Counting rules: FMA = 2 FLOP; add, mul, div, sqrt, exp = 1.  Per ordered
(pair, term) in the n-space form, the elements cost F_E(n) and one bra-side
gradient F_G(n); per unordered pair: nst (F_E + 2 F_G) off the diagonal and
nst (F_E + F_G) on it.  The table is the fixed model the throughput is
reported against (LiH: 413,856 FLOP per off-diagonal pair); the kernels may
execute fewer operations than the model counts."""

F_E = {2: 346, 3: 733, 5: 2100}
F_G = {2: 975, 3: 2272, 5: 7572}


def flops_per_eval(n: int, nst: int, K: int) -> float:
    if n not in F_E:
        raise KeyError(f"no FLOP model for n = {n}")
    off = K * (K - 1) // 2
    return float(nst * (off * (F_E[n] + 2 * F_G[n]) + K * (F_E[n] + F_G[n])))


def pairs_per_eval(K: int) -> int:
    return K * (K + 1) // 2
