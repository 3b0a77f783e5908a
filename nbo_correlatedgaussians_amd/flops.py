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


def dipole_flops_per_term_pair(n: int) -> int:
    """One (bra term, ket term) pair of the dipole pass in the L D L^T form,
    counted with the rules above: M and e (n(n+1)/2 + 3n adds), L D L^T with
    reciprocal pivots, forward and backward solves for three right-hand
    sides, e.st, the weight (2 subtractions, exp, sqrt, 5 multiplications) and
    the 1 + 3n accumulations.  LiH (n = 5): 295."""
    ns = n * (n + 1) // 2
    ldl = sum(2 * j + 2 + (n - 1 - j) * (2 * j + 1) for j in range(n))
    fwd = 3 * n * (n - 1)
    bwd = 3 * (n + n * (n - 1))
    return ns + 3 * n + ldl + fwd + bwd + 6 * n + 9 + 1 + 6 * n


def dipole_flops_per_eval(n: int, nst: int, K: int) -> float:
    return float(dipole_flops_per_term_pair(n) * nst * nst * (K * (K + 1) // 2))
