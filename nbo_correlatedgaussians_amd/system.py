"""System descriptions, the NECG input-file format, and parameter generation.

A *system* is everything the energy path needs besides the basis parameters:
the number of pseudo-particles ``n`` (the reference's ``ne``), the inverse-mass
tensor, the charges, the electric-field strength and the symmetry projector
(``nst`` terms, each a weight ``scvv`` with an n x n ket transform ``sml`` and
a 3n x 3n shift transform ``sms``).  The fields are the module globals that
``python/read_input.py:16-157`` fills (``python/shared.py:15-66``).

Parameter vector layout (``python/outprod.py:40-41``, ``python/matel.py:56-87``):
``x = [iv | C]`` with ``iv`` made of K blocks of ``blk = svh + skp`` numbers,
``[vech(L_k) column-major lower incl. diagonal (svh) | s_k particle-major xyz
(skp)]`` followed by the K linear coefficients ``C``.
"""

from __future__ import annotations

import json
import math
import os
from dataclasses import dataclass, field

import numpy as np

_DATA = os.path.join(os.path.dirname(__file__), "data")

# Particle masses in electron masses.
M_PROTON = 1836.15267
M_LI7_NUCLEUS = 12786.39
M_HE4_NUCLEUS = 7294.2995
M_BE9_NUCLEUS = 16424.2037    # 9.0121831 u less 4 electrons
M_C12_NUCLEUS = 21868.6617    # 12 u less 6 electrons


@dataclass
class System:
    """One non-Born-Oppenheimer system (physics block of an input file)."""

    n: int                      # pseudo-particles (reference: ne)
    nst: int                    # symmetry terms
    efs: float                  # electric-field strength
    mass: np.ndarray            # (n, n) inverse-mass tensor
    ncv: np.ndarray             # (n+1,) charges, ncv[0] = reference particle
    scvv_raw: np.ndarray        # (nst,) weights as written in the file
    sml: np.ndarray             # (nst, n, n); sml[p] = reference sml[:, :, p]
    sms: np.ndarray             # (nst, 3n, 3n); sms[p] = reference sms[:, :, p]
    name: str = "system"
    meta: dict = field(default_factory=dict)

    # -- derived sizes (python/shared.py:19-21) --------------------------
    @property
    def svh(self) -> int:
        return self.n * (self.n + 1) // 2

    @property
    def skp(self) -> int:
        return 3 * self.n

    @property
    def blk(self) -> int:
        return self.svh + self.skp

    @property
    def npairs(self) -> int:
        """Coulomb pairs (reference particle + n pseudo-particles)."""
        return self.n * (self.n + 1) // 2

    @property
    def scvv(self) -> np.ndarray:
        """Weights as the energy path uses them: divided by nst
        (``python/read_input.py:141``)."""
        return self.scvv_raw / float(self.nst)

    def siv(self, K: int) -> int:
        return K * self.blk

    def nx(self, K: int) -> int:
        return K * (self.blk + 1)

    def validate(self) -> None:
        n, nst = self.n, self.nst
        if n < 1 or nst < 1:
            raise ValueError("n and nst must be positive")
        for name, arr, shape in (
            ("mass", self.mass, (n, n)),
            ("ncv", self.ncv, (n + 1,)),
            ("scvv_raw", self.scvv_raw, (nst,)),
            ("sml", self.sml, (nst, n, n)),
            ("sms", self.sms, (nst, 3 * n, 3 * n)),
        ):
            if np.shape(arr) != shape:
                raise ValueError(f"{name} has shape {np.shape(arr)}, expected {shape}")
        if not np.allclose(self.mass, self.mass.T, rtol=0, atol=0):
            raise ValueError("mass tensor must be symmetric")

    def symmetry_report(self) -> dict:
        """Structural facts about the projector that the fused kernels use.

        * ``sms_is_kron``: every sms[p] equals kron(sml[p].T, I3), so the shift
          transform is the n x n ket transform acting on the n x 3 shift matrix.
        * ``closed``/``inverse_weights``: the sml matrices form a group and
          c_P = c_{P^-1}.
        """
        I3 = np.eye(3)
        kron = all(np.array_equal(self.sms[p], np.kron(self.sml[p].T, I3))
                   for p in range(self.nst))
        keys = {self.sml[p].tobytes(): p for p in range(self.nst)}
        closed = all((self.sml[a] @ self.sml[b]).tobytes() in keys
                     for a in range(self.nst) for b in range(self.nst))
        inv_ok = True
        for p in range(self.nst):
            inv = np.linalg.inv(self.sml[p])
            q = keys.get(np.round(inv, 12).tobytes())
            if q is None or self.scvv_raw[q] != self.scvv_raw[p]:
                inv_ok = False
        return {"sms_is_kron": kron, "closed": closed, "inverse_weights": inv_ok}


# ---------------------------------------------------------------------------
# Input-file format (python/read_input.py:16-157; Appendix layout:
# M ne nst efs | mass n*n row-major | ncv n+1 | iv siv | C M | scvv nst (raw)
# | sml nst*(n*n row-major) | sms nst*(3n*3n row-major)).
# ---------------------------------------------------------------------------

def _tok_float(t: str) -> float:
    return float(t.lower().replace("d", "e"))


def parse_tokens(tokens):
    """Parse whitespace tokens of an input file -> (System, K, iv, C)."""
    pos = 0

    def take(count, conv=_tok_float):
        nonlocal pos
        if pos + count > len(tokens):
            raise ValueError("input file ends early")
        vals = [conv(t) for t in tokens[pos:pos + count]]
        pos += count
        return vals

    K, n, nst = (int(v) for v in take(3, int))
    (efs,) = take(1)
    svh, skp = n * (n + 1) // 2, 3 * n
    mass = np.array(take(n * n)).reshape(n, n)
    ncv = np.array(take(n + 1))
    iv = np.array(take(K * (svh + skp)))
    C = np.array(take(K))
    scvv_raw = np.array(take(nst))
    sml = np.array(take(nst * n * n)).reshape(nst, n, n)
    sms = np.array(take(nst * skp * skp)).reshape(nst, skp, skp)
    sysd = System(n=n, nst=nst, efs=efs, mass=mass, ncv=ncv,
                  scvv_raw=scvv_raw, sml=sml, sms=sms)
    sysd.validate()
    return sysd, K, iv, C


def read_input(path: str):
    """Read an input file in the reference layout -> (System, K, iv, C)."""
    with open(path, "r", encoding="utf-8") as fh:
        return parse_tokens(fh.read().split())


def _fmt(v: float) -> str:
    return repr(float(v))


def format_input(sysd: System, iv: np.ndarray, C: np.ndarray) -> str:
    """Render an input file that the reference reader accepts."""
    K = len(C)
    n = sysd.n
    if len(iv) != K * sysd.blk:
        raise ValueError("iv length does not match K*blk")
    lines = [str(K), str(n), str(sysd.nst), _fmt(sysd.efs), ""]
    for i in range(n):
        lines.append(" ".join(_fmt(v) for v in sysd.mass[i]))
    lines.append("")
    lines += [_fmt(v) for v in sysd.ncv]
    lines.append("")
    lines += [_fmt(v) for v in iv]
    lines.append("")
    lines += [_fmt(v) for v in C]
    lines.append("")
    lines += [_fmt(v) for v in sysd.scvv_raw]
    lines.append("")
    for p in range(sysd.nst):
        for i in range(n):
            lines.append(" ".join(_fmt(v) for v in sysd.sml[p, i]))
        lines.append("")
    for p in range(sysd.nst):
        for i in range(3 * n):
            lines.append(" ".join(_fmt(v) for v in sysd.sms[p, i]))
        lines.append("")
    return "\n".join(lines) + "\n"


def write_input(path: str, sysd: System, iv: np.ndarray, C: np.ndarray) -> None:
    with open(path, "w", encoding="utf-8") as fh:
        fh.write(format_input(sysd, iv, C))


# ---------------------------------------------------------------------------
# Systems
# ---------------------------------------------------------------------------

def inverse_mass_tensor(m_ref: float, masses) -> np.ndarray:
    """Kinetic tensor of pseudo-particles measured from a reference particle:
    M_ii = (1/m_i + 1/m_0)/2, M_ij = 1/(2 m_0)."""
    masses = np.asarray(masses, dtype=np.float64)
    n = len(masses)
    M = np.full((n, n), 1.0 / (2.0 * m_ref))
    M[np.diag_indices(n)] = 0.5 * (1.0 / masses + 1.0 / m_ref)
    return M


def _perm_matrix(perm) -> np.ndarray:
    n = len(perm)
    P = np.zeros((n, n))
    for i, j in enumerate(perm):
        P[i, j] = 1.0
    return P


def _from_perms(name, n, efs, mass, ncv, perms, weights) -> System:
    sml = np.stack([_perm_matrix(p) for p in perms])
    sms = np.stack([np.kron(s.T, np.eye(3)) for s in sml])
    sysd = System(n=n, nst=len(perms), efs=efs, mass=mass,
                  ncv=np.asarray(ncv, dtype=np.float64),
                  scvv_raw=np.asarray(weights, dtype=np.float64),
                  sml=sml, sms=sms, name=name)
    sysd.validate()
    return sysd


def _load_oldquit():
    with open(os.path.join(_DATA, "lih_oldquit.json"), "r", encoding="utf-8") as fh:
        return json.load(fh)


def lih_system() -> System:
    """Non-BO LiH, Li nucleus as reference: proton + 4 electrons (n = 5),
    24 Young-operator terms and the Stark field of ``oldquit.dat``."""
    d = _load_oldquit()
    sysd = System(n=d["n"], nst=d["nst"], efs=d["efs"],
                  mass=np.array(d["mass"]), ncv=np.array(d["ncv"]),
                  scvv_raw=np.array(d["scvv_raw"]), sml=np.array(d["sml"]),
                  sms=np.array(d["sms"]), name="LiH")
    sysd.validate()
    return sysd


def lih_oldquit_basis():
    """The 12-function basis (iv, C) stored in ``oldquit.dat``."""
    d = _load_oldquit()
    return np.array(d["iv"]), np.array(d["C"])


def he_system() -> System:
    """He atom, nucleus as reference: two electrons (n = 2), singlet spatial
    projector {E, P12} with weights [1, 1], no field."""
    mass = inverse_mass_tensor(M_HE4_NUCLEUS, [1.0, 1.0])
    return _from_perms("He", 2, 0.0, mass, [2.0, -1.0, -1.0],
                       [[0, 1], [1, 0]], [1.0, 1.0])


def h2_system() -> System:
    """H2, one proton as reference: proton + 2 electrons (n = 3), electron
    exchange {E, P(e1 e2)} with weights [1, 1], no field."""
    mass = inverse_mass_tensor(M_PROTON, [M_PROTON, 1.0, 1.0])
    return _from_perms("H2", 3, 0.0, mass, [1.0, 1.0, -1.0, -1.0],
                       [[0, 1, 2], [0, 2, 1]], [1.0, 1.0])


def hydrogen_atom() -> System:
    """One electron around an infinitely heavy proton (n = 1, one term).
    Used by the closed-form checks: E(a) = 1.5 a - 2 sqrt(2 a / pi)."""
    mass = np.array([[0.5]])
    return _from_perms("H", 1, 0.0, mass, [1.0, -1.0], [[0]], [1.0])


def be_system() -> System:
    """Be-like atom, nucleus as reference: four electrons (n = 4), the group
    {E, P12, P34, P12 P34} with weights [1, 1, 1, 1] and a Stark field.
    A test system for n = 4 (not a spin-adapted projector)."""
    mass = inverse_mass_tensor(M_BE9_NUCLEUS, [1.0] * 4)
    return _from_perms("Be", 4, 0.001, mass, [4.0, -1.0, -1.0, -1.0, -1.0],
                       [[0, 1, 2, 3], [1, 0, 2, 3], [0, 1, 3, 2], [1, 0, 3, 2]],
                       [1.0, 1.0, 1.0, 1.0])


def c_system() -> System:
    """C-like atom, nucleus as reference: six electrons (n = 6), the group
    {E, a, b, ab} with a = P12, b = P34 P56 and characters [1, -1, 1, -1], and
    a Stark field.  A test system for n = 6 (not a spin-adapted projector)."""
    mass = inverse_mass_tensor(M_C12_NUCLEUS, [1.0] * 6)
    e = [0, 1, 2, 3, 4, 5]
    a = [1, 0, 2, 3, 4, 5]
    b = [0, 1, 3, 2, 5, 4]
    ab = [1, 0, 3, 2, 5, 4]
    return _from_perms("C", 6, -0.0005, mass, [6.0] + [-1.0] * 6, [e, a, b, ab],
                       [1.0, -1.0, 1.0, -1.0])


SYSTEMS = {"LiH": lih_system, "He": he_system, "H2": h2_system, "Be": be_system,
           "C": c_system}


# ---------------------------------------------------------------------------
# Parameter generation
# ---------------------------------------------------------------------------

def generate_parameters(sysd: System, K: int, seed: int,
                        diag=(0.1, 10.0), offdiag_sd=0.3, shift_sd=0.1,
                        coef_sd=1.0):
    """Random basis of K functions.

    Distributions: L_k lower-triangular with diagonal log-uniform in
    [diag[0], diag[1]] and off-diagonals N(0, offdiag_sd); s_k ~ N(0, shift_sd);
    C_k ~ N(0, coef_sd).  Draw order from ``numpy.random.default_rng(seed)``:
    all diagonals (K, n), all off-diagonals (K, svh - n), all shifts (K, 3n),
    then C (K).  Returns (iv, C).
    """
    n = sysd.n
    rng = np.random.default_rng(seed)
    d = np.exp(rng.uniform(math.log(diag[0]), math.log(diag[1]), size=(K, n)))
    off = rng.normal(0.0, offdiag_sd, size=(K, sysd.svh - n))
    s = rng.normal(0.0, shift_sd, size=(K, 3 * n))
    C = rng.normal(0.0, coef_sd, size=K)
    iv = np.empty((K, sysd.blk))
    for k in range(K):
        col = []
        oi = 0
        for j in range(n):            # vech: column j, rows j..n-1
            col.append(d[k, j])
            for _ in range(j + 1, n):
                col.append(off[k, oi])
                oi += 1
        iv[k, :sysd.svh] = col
        iv[k, sysd.svh:] = s[k]
    return iv.reshape(-1), C


def _vech_diagonal_slots(n: int):
    """Positions of L_00, L_11, ... in the column-major vech order."""
    out, k = [], 0
    for j in range(n):
        out.append(k)
        k += n - j
    return out


def template_statistics(sysd: System, iv_template: np.ndarray):
    """Per-column mean and standard deviation (population, ddof 0) of a
    basis's parameter blocks, with standard deviations below 1e-10 raised to
    1e-6 (input_file_generation/generate_lih_input.py:230-234)."""
    blocks = np.asarray(iv_template, dtype=np.float64).reshape(-1, sysd.blk)
    mean = blocks.mean(axis=0)
    sd = blocks.std(axis=0)
    return mean, np.where(sd < 1e-10, 1e-6, sd)


def _lognormal_fit(mean: float, sd: float):
    """(mu, sigma) of the log-normal variate with the given mean and standard
    deviation (input_file_generation/generate_lih_input.py:139-154): a
    non-positive mean becomes |mean| + 1e-3, a standard deviation below 1e-10
    becomes 5 % of the mean."""
    mean, sd = np.float64(mean), np.float64(sd)
    if mean <= 0:
        mean = abs(mean) + 1e-3
    if sd < 1e-10:
        sd = mean * 0.05
    s2 = np.log1p((sd / mean) ** 2)        # numpy's log/log1p/sqrt, so the
    return np.log(mean) - 0.5 * s2, np.sqrt(s2)   # draws match bit for bit


def generate_like_template(sysd: System, K: int, seed, iv_template: np.ndarray = None):
    """K functions drawn the way the reference's input generator draws them
    (input_file_generation/generate_lih_input.py:157-189, :228-245): every
    block is one normal draw of all blk columns with the template's column
    means and standard deviations, after which each diagonal entry of L is
    replaced, in column order, by a log-normal draw fitted to that column's
    mean and standard deviation; C_k = 1/K.  The stream is
    ``numpy.random.default_rng(seed)``.  The template defaults to the
    12-function basis of ``oldquit.dat`` (LiH).  Returns (iv, C) at full
    precision; the generator's file rounds them to 16 significant digits
    (``format_generated_input``)."""
    if K < 1:
        raise ValueError(f"K must be >= 1, got {K}")
    if iv_template is None:
        iv_template, _ = lih_oldquit_basis()
    mean, sd = template_statistics(sysd, iv_template)
    diag = _vech_diagonal_slots(sysd.n)
    fits = [_lognormal_fit(mean[i], sd[i]) for i in diag]
    rng = np.random.default_rng(seed)
    iv = np.empty((K, sysd.blk))
    for k in range(K):
        row = rng.normal(mean, sd)
        for i, (mu, sg) in zip(diag, fits):
            row[i] = rng.lognormal(mu, sg)
        iv[k] = row
    return iv.reshape(-1), np.full(K, 1.0 / K)


def _gen_float(v: float) -> str:
    return "0.0" if v == 0.0 else f"{v:25.15E}"


def format_generated_input(sysd: System, iv: np.ndarray, C: np.ndarray) -> str:
    """The input file in the generator's text layout, character for character
    (input_file_generation/generate_lih_input.py:212-307): header lines,
    mass rows, charges with one decimal, one parameter per line, C, raw
    weights (integers where whole), then sml and sms rows with one decimal.
    The reader (``read_input``) accepts it like any other input file."""
    K, n = len(C), sysd.n
    if len(iv) != K * sysd.blk:
        raise ValueError("iv length does not match K*blk")
    out = [str(K), str(n), str(sysd.nst), f"{sysd.efs:.10E}".replace("E", "d"), ""]
    out += ["  ".join(_gen_float(v) for v in row) for row in sysd.mass]
    out.append("")
    out += [f"{v:.1f}" for v in sysd.ncv]
    out += ["", ""]
    for blkv in np.asarray(iv, dtype=np.float64).reshape(K, sysd.blk):
        out += [f"   {_gen_float(v)}" for v in blkv]
        out.append("")
    out.append("")
    out += [f"  {_gen_float(v)}" for v in C]
    out.append("")
    out += [str(int(v)) if v == int(v) else f"{v:.6g}" for v in sysd.scvv_raw]
    out.append("")
    for mats in (sysd.sml, sysd.sms):
        for m in mats:
            out += ["  ".join(f"{v:.1f}" for v in row) for row in m]
            out.append("")
    return "\n".join(out) + "\n"


def generated_lih_input(K: int, seed):
    """Config-0-style LiH input: the oldquit.dat physics block with K
    functions drawn by ``generate_like_template`` and written by
    ``format_generated_input``.  Returns (text, System, iv, C) with iv and C
    as read back from the text (what a run on that file sees)."""
    sysd = lih_system()
    iv, C = generate_like_template(sysd, K, seed)
    text = format_generated_input(sysd, iv, C)
    s2, K2, iv2, C2 = parse_tokens(text.split())
    s2.name = "LiH"
    return text, s2, iv2, C2


def pack_x(iv: np.ndarray, C: np.ndarray) -> np.ndarray:
    return np.concatenate([np.asarray(iv, dtype=np.float64),
                           np.asarray(C, dtype=np.float64)])


def unpack_x(sysd: System, x: np.ndarray):
    x = np.asarray(x, dtype=np.float64)
    if x.ndim != 1 or x.size % (sysd.blk + 1):
        raise ValueError(f"x of length {x.size} is not K*(blk+1) with blk={sysd.blk}")
    K = x.size // (sysd.blk + 1)
    return K, x[:K * sysd.blk], x[K * sysd.blk:]


def vech_to_lower(v: np.ndarray, n: int) -> np.ndarray:
    """Inverse of the column-major lower-triangle packing
    (``python/utils.py:42-55``)."""
    L = np.zeros((n, n))
    k = 0
    for j in range(n):
        for i in range(j, n):
            L[i, j] = v[k]
            k += 1
    return L


def lower_to_vech(L: np.ndarray) -> np.ndarray:
    n = L.shape[0]
    return np.array([L[i, j] for j in range(n) for i in range(j, n)])
