"""Algorithmic work of one energy+gradient evaluation.

Counting rules: FMA = 2 FLOP; add, subtract, multiply, divide, sqrt, exp = 1;
a reciprocal square root is a sqrt and a divide (2).  Newton refinements of
hardware seeds, selects and moves are not counted.

Two models:

* ``flops_per_eval_v0``: the reference's formulas written in n-space, per
  ordered (pair, term) F_E(n) for the elements and F_G(n) for one bra-side
  gradient; per unordered pair nst (F_E + 2 F_G) off the diagonal and
  nst (F_E + F_G) on it (LiH: 413,856 FLOP per off-diagonal pair).
* ``flops_per_eval``: the two-sided algorithm the pair kernel runs: per
  unordered (pair, term) one factorisation, the elements, the Coulomb and
  Stark derivatives, and both gradient halves (``two_sided_term``; LiH,
  whose mass matrix is diagonal plus a constant: 4,031 FLOP per pair-term,
  96,744 per off-diagonal pair).  The identities
  it uses (A' dif2 = -A dif1, B^-1 A' = I - B^-1 A, the collapsed kinetic and
  Coulomb gradient products) make it 4.3x smaller than the v0 model for LiH,
  so the v0 model overstates the work of the run and is not the roofline
  basis.
"""

F_E = {2: 346, 3: 733, 5: 2100}
F_G = {2: 975, 3: 2272, 5: 7572}


def flops_per_eval_v0(n: int, nst: int, K: int) -> float:
    if n not in F_E:
        raise KeyError(f"no v0 FLOP model for n = {n}")
    off = K * (K - 1) // 2
    return float(nst * (off * (F_E[n] + 2 * F_G[n]) + K * (F_E[n] + F_G[n])))


def mass_is_rank1(mass) -> bool:
    """True when every off-diagonal entry of the mass matrix is the same (the
    matrix is diagonal plus a constant; the kernels then form M a1 from one
    column sum)."""
    n = len(mass)
    if n < 2:
        return True
    c = mass[0][1]
    return all(mass[i][j] == c for i in range(n) for j in range(n) if i != j)


def two_sided_term(n: int, rank1: bool = False) -> dict:
    """FLOP of one unordered (bra, ket, term) evaluation of the two-sided
    pair term, by phase, for n pseudo-particles (ns = n(n+1)/2 packed
    entries, n3 = 3n shift entries, n(n+1)/2 Coulomb pairs of which n involve
    the reference particle).  'ket' is the ket-side gradient, which the
    diagonal pairs (k = l) do not need.  rank1: the mass matrix is diagonal
    plus a constant, and M a1 costs a column sum, 3 products and n3 FMAs."""
    ns, n3 = n * (n + 1) // 2, 3 * n
    chol = ns                                            # B = A + A'
    for j in range(n):                                   # Cholesky, 1/R_jj by rsq
        chol += 2 * j + 2 + 1 + (n - 1 - j) * (2 * j + 1)
    inv = sum(2 * (i - j) + 1 for j in range(n) for i in range(j + 1, n))   # R^-1
    inv += sum((i + 1) * 2 * (n - i) for i in range(n))                     # B^-1 = R^-T R^-1
    det = (n - 1) + 2                                    # det(B)^-3/2: the pivots' product, cubed
    elem = n3 + 2 * n3 * n + 2 * n3 + 6                  # E, st = B^-1 E, i3, ov
    ma1 = 3 * (n - 1) + 3 + 2 * n3 if rank1 else 2 * n3 * n
    kin = n3 + 2 * n3 * n + ma1 + 2 * n3                 # dif1, a1 = A dif1, M a1, i3k
    kin += 2 * ns + 2 + 4                                # <AMA, B^-1> (off-diagonal part doubled), tfac
    base = 2 + 1 + 6 + 51 + 8                            # rsq, 1/i2, z, Boys F0/2F0', ca cb
    ref_pair = base + 3                                  # (0, q): row q of Lb st set to cb st_q
    pp_pair = base + 3 + 3 + 2 + 12                      # (p, q): i2, dif, 2 La diagonal updates,
    #                                                      cb dif added to rows p, q of Lb st
    coul = n * ref_pair + (ns - n) * pp_pair
    stark = 2 * n + 4 + 3 + 2                            # zsum, alpha, H and S sums
    shared = 2 * n + 2 * n3 * n + 2 * n3 + 2 * n3 * n   # Stark column, A M a1, F - 8 A M a1,
    #                                                     B^-1 F (the ket side's Phi; Lb st is formed
    #                                                     in the Coulomb loop)
    side = (n3 + 5 + n3                                  # dif, the weights, e = 2 wo dif
            + 3 * n3                                     # Y_H = alpha e + wo Phi (Y_S = e)
            + 2 * ns + n * n * (2 * n - 1)               # U = 12 A'MA' + La, V = B^-1 U
            + ns * ((2 * n - 1) + 12 + 7 + 3)            # G = V B^-1 less dp, X_S, X_H
            + (2 * ns + 2 * n3 + 2))                     # accumulation of the values
    return {"factorisation": chol + inv + det, "elements": elem, "kinetic": kin,
            "coulomb": coul, "stark": stark, "shared_gradient": shared,
            "bra": side + 2 * n3,                        # Phi = Phi_ket + 8 M a1
            "ket": side}


def two_sided_term_flops(n: int, diagonal: bool = False, rank1: bool = False) -> int:
    ph = two_sided_term(n, rank1)
    return sum(ph.values()) - (ph["ket"] if diagonal else 0)


def flops_per_eval(n: int, nst: int, K: int, rank1: bool = False) -> float:
    """FLOP of one energy+gradient evaluation by the two-sided algorithm:
    nst terms for each of the K(K-1)/2 off-diagonal pairs (both gradient
    halves) and the K diagonal pairs (bra half only)."""
    off = K * (K - 1) // 2
    return float(nst * (off * two_sided_term_flops(n, rank1=rank1)
                        + K * two_sided_term_flops(n, True, rank1)))


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
