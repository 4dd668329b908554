"""Cubic lattices, coupling shells and the benchmark model configurations.

The reference takes its Hamiltonian as coupling rows ``f1=..:f2=..:jij=..`` in an input
file (parameters.F90:2558-2846) and expands them over the supercell (jij.F90:123-446).
This module produces the same rows for simple cubic, BCC and FCC lattices from a list of
per-shell couplings, and the ELL tables from them through the library's host-side
table builder (``esp_build_table``).

Units: values are given in the reference's internal unit (energy / k_B).  "J = 1" in the
model descriptions below means jij = k_B eV in an input file, so temperatures are in J/k_B.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Optional

import numpy as np

K_B = 8.6173303e-5          # eV/K (constants.F90:48)

BASIS = {
    "sc": [(0.0, 0.0, 0.0)],
    "bcc": [(0.0, 0.0, 0.0), (0.5, 0.5, 0.5)],
    "fcc": [(0.0, 0.0, 0.0), (0.0, 0.5, 0.5), (0.5, 0.0, 0.5), (0.5, 0.5, 0.0)],
}
COORDINATION = {"sc": (6, 12, 8), "bcc": (8, 6, 12), "fcc": (12, 6, 24)}


@dataclasses.dataclass
class Rows:
    f1: np.ndarray            # [n,3]
    f2: np.ndarray            # [n,3]
    value: np.ndarray         # [n]
    dvec: Optional[np.ndarray] = None   # [n,3] unit vectors (DM only)

    @property
    def n(self):
        return len(self.value)


def shell_vectors(kind: str, nshell: int):
    """Neighbour vectors by distance shell: list over basis atoms a of list over shells of
    [(b, cell_offset, vector)] in a fixed, sorted order."""
    basis = np.array(BASIS[kind])
    rng = range(-3, 4)
    out = []
    for a, pa in enumerate(basis):
        cand = []
        for n0 in rng:
            for n1 in rng:
                for n2 in rng:
                    for b, pb in enumerate(basis):
                        v = np.array([n0, n1, n2], float) + pb - pa
                        d = float(np.linalg.norm(v))
                        if d < 1e-9:
                            continue
                        cand.append((round(d, 9), b, (n0, n1, n2), tuple(v)))
        dists = sorted({c[0] for c in cand})[:nshell]
        shells = []
        for d in dists:
            sel = sorted([c for c in cand if c[0] == d], key=lambda c: (c[3], c[1]))
            shells.append([(c[1], c[2], np.array(c[3])) for c in sel])
        out.append(shells)
    return out


def shell_rows(kind: str, couplings, dm: float = 0.0, dm_shell: int = 0) -> Rows:
    """Rows for per-shell couplings (couplings[k] for shell k+1).  With dm_shell > 0 the
    rows carry DM vectors along the bond direction (value dm) instead."""
    nshell = max(len(couplings), dm_shell)
    sv = shell_vectors(kind, nshell)
    basis = np.array(BASIS[kind])
    f1, f2, val, dv = [], [], [], []
    for a, shells in enumerate(sv):
        for k, shell in enumerate(shells):
            if dm_shell:
                if k + 1 != dm_shell:
                    continue
                c = dm
            else:
                c = couplings[k] if k < len(couplings) else 0.0
                if c == 0.0:
                    continue
            for b, cell, v in shell:
                f1.append(basis[a])
                f2.append(basis[b] + np.array(cell, float))
                val.append(c)
                dv.append(v / np.linalg.norm(v))
    r = Rows(np.array(f1, float).reshape(-1, 3), np.array(f2, float).reshape(-1, 3),
             np.array(val, float))
    if dm_shell:
        r.dvec = np.array(dv, float).reshape(-1, 3)
    return r


@dataclasses.dataclass
class Table:
    z: int
    cnt: np.ndarray    # [N] int32
    nbr: np.ndarray    # [z,N] int32, 1-based
    val: np.ndarray    # [z,N] or [z,N,3]


@dataclasses.dataclass
class Model:
    """A Hamiltonian in the reference's table layout (what esp_set_model takes)."""
    kind: str
    L: tuple
    nb: int
    basis_frac: np.ndarray
    mag: np.ndarray            # [N]
    J: Table
    B: Optional[Table] = None
    D: Optional[Table] = None
    si_par: Optional[np.ndarray] = None     # [nb]
    si_axis: Optional[np.ndarray] = None    # [nb,3]
    h_par: Optional[np.ndarray] = None
    h_axis: Optional[np.ndarray] = None
    rows: dict = dataclasses.field(default_factory=dict)

    @property
    def N(self):
        return len(self.mag)


def build_table(nb, basis_frac, basis_mag, L, rows: Rows, ncomp=1, open_boundary=False,
                builder=None) -> Table:
    """Expand rows with ``builder`` (default: the library's esp_build_table)."""
    from . import _lib as E
    ncell = int(L[0] * L[1] * L[2])
    N = ncell * nb
    zcap = max(rows.n, 1)
    cnt = np.zeros(N, np.int32)
    nbr = np.zeros((zcap, N), np.int32)
    val = np.zeros((zcap, N, ncomp) if ncomp == 3 else (zcap, N), np.float64)
    zmax = np.zeros(1, np.int32)
    Larr = np.array(L, np.int32)
    bf = E.f64(basis_frac)
    bm = E.f64(basis_mag)
    f1, f2, v = E.f64(rows.f1), E.f64(rows.f2), E.f64(rows.value)
    dv = E.f64(rows.dvec) if (ncomp == 3) else None
    if builder is None:
        rc = E.lib.esp_build_table(nb, E.ptr(bf, E.C.c_double), E.ptr(bm, E.C.c_double),
                                   E.ptr(Larr, E.C.c_int), int(open_boundary), rows.n,
                                   E.ptr(f1, E.C.c_double), E.ptr(f2, E.C.c_double),
                                   E.ptr(v, E.C.c_double), E.ptr(dv, E.C.c_double), ncomp, zcap,
                                   E.ptr(cnt, E.C.c_int), E.ptr(nbr, E.C.c_int),
                                   E.ptr(val, E.C.c_double), E.ptr(zmax, E.C.c_int))
        E.check(rc)
        z = int(zmax[0])
    else:
        z = builder(nb, bf, bm, Larr, open_boundary, f1, f2, v, dv, ncomp, zcap, cnt, nbr, val)
    return Table(z, cnt, np.ascontiguousarray(nbr[:z]), np.ascontiguousarray(val[:z]))


def cubic_model(kind: str, L: int, J=(1.0,), B: float = 0.0, D: float = 0.0,
                K: float = 0.0, K_axis=(0.0, 0.0, 1.0), field: float = 0.0,
                field_axis=(0.0, 0.0, 1.0), builder=None) -> Model:
    """Cubic-lattice model: per-shell J (internal units), nearest-neighbour B and D (D along
    the bond direction), single-ion Delta = K along K_axis (reference sign: Delta < 0 is easy
    axis), Zeeman field along field_axis."""
    nb = len(BASIS[kind])
    bf = np.array(BASIS[kind], float)
    bm = np.ones(nb)
    Ls = (L, L, L)
    rows = {"J": shell_rows(kind, list(J))}
    m = Model(kind, Ls, nb, bf, np.ones(L ** 3 * nb),
              build_table(nb, bf, bm, Ls, rows["J"], 1, builder=builder), rows=rows)
    if B:
        rows["B"] = shell_rows(kind, [B])
        m.B = build_table(nb, bf, bm, Ls, rows["B"], 1, builder=builder)
    if D:
        rows["D"] = shell_rows(kind, [], dm=D, dm_shell=1)
        m.D = build_table(nb, bf, bm, Ls, rows["D"], 3, builder=builder)
    if K:
        ax = np.array(K_axis, float)
        ax /= np.linalg.norm(ax)
        m.si_par = np.full(nb, K)
        m.si_axis = np.tile(ax, (nb, 1))
    if field:
        ax = np.array(field_axis, float)
        ax /= np.linalg.norm(ax)
        m.h_par = np.full(nb, field)
        m.h_axis = np.tile(ax, (nb, 1))
    return m


def shell_tables(kind: str, L: int, nshell: int, builder=None):
    """Neighbour tables of the first nshell distance shells of a cubic lattice (the shells of
    the J rows that spin correlations are reported for, jij.F90:449-503): cnt[nshell][N] and
    nbr[nshell][zs][N] (1-based site ids, zs the widest shell)."""
    nb = len(BASIS[kind])
    bf = np.array(BASIS[kind], float)
    Ls = (L, L, L)
    tabs = []
    for k in range(nshell):
        c = [0.0] * nshell
        c[k] = 1.0
        tabs.append(build_table(nb, bf, np.ones(nb), Ls, shell_rows(kind, c), 1, builder=builder))
    N = L ** 3 * nb
    zs = max(t.z for t in tabs)
    cnt = np.zeros((nshell, N), np.int32)
    nbr = np.zeros((nshell, zs, N), np.int32)
    for k, t in enumerate(tabs):
        cnt[k] = t.cnt
        nbr[k, :t.z] = t.nbr
    return cnt, nbr


def site_positions(kind: str, L: int) -> np.ndarray:
    """Cartesian positions [N][3] (cubic cell edge 1) in the reference site order
    ((x2 L + x1) L + x0) nb + a."""
    bf = np.array(BASIS[kind], float)
    nb = len(bf)
    x = np.arange(L)
    x2, x1, x0 = np.meshgrid(x, x, x, indexing="ij")
    cells = np.stack([x0.ravel(), x1.ravel(), x2.ravel()], axis=1).astype(float)
    return (cells[:, None, :] + bf[None, :, :]).reshape(-1, 3)


def ladder(mode: str, t0: float, t1: float, n: int) -> np.ndarray:
    """Temperature ladders of parameters.F90:293-317 ('lin', 'inv', 'log')."""
    t = np.empty(n)
    t[0] = t0
    if n > 1:
        if mode == "lin":
            step = (t1 - t0) / (n - 1)
            for k in range(n):
                t[k] = t0 + k * step
        elif mode == "inv":
            step = (1.0 / t1 - 1.0 / t0) / (n - 1)
            for k in range(n):
                t[k] = 1.0 / (1.0 / t0 + k * step)
        elif mode == "log":
            for k in range(n):
                t[k] = t0 * math.exp(math.log(t1 / t0) * (1.0 * k / (n - 1)))
        else:
            raise ValueError(mode)
    return t


def random_spins(N: int, R: int, seed: int, mag=None) -> np.ndarray:
    """Uniform-on-sphere initial spins [R][N][3] (the benchmark configurations ask for
    uniform on S^2; the reference's own 'rand' init, mcarlo.F90:275-286, is not uniform)."""
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(R, N, 3))
    v /= np.linalg.norm(v, axis=2, keepdims=True)
    if mag is not None:
        v *= np.asarray(mag)[None, :, None]
    return np.ascontiguousarray(v)


def ferro_spins(N: int, R: int, mag=None) -> np.ndarray:
    """'ferro' init (mcarlo.F90:270-273): S = (0, 0, m)."""
    s = np.zeros((R, N, 3))
    s[:, :, 2] = 1.0 if mag is None else np.asarray(mag)[None, :]
    return s


def sublattice_spins(m, dirs, R: int = 1) -> np.ndarray:
    """Spins [R][N][3] with basis atom a of every cell along dirs[a] (site index
    cell * nb + a, jij.F90:219-229), scaled by the moments."""
    d = np.asarray(dirs, float)
    d = d / np.linalg.norm(d, axis=1, keepdims=True)
    s = d[np.arange(m.N) % m.nb] * np.asarray(m.mag)[:, None]
    return np.ascontiguousarray(np.repeat(s[None], R, axis=0))


# The four-sublattice tetrahedral order of the FCC antiferromagnet: the four basis atoms of
# the conventional cell point to the corners of a tetrahedron (all nearest-neighbour angles
# arccos(-1/3), next-nearest neighbours parallel).
FCC_TETRAHEDRAL = ((1, 1, 1), (1, -1, -1), (-1, 1, -1), (-1, -1, 1))


# ---- the benchmark model configurations ------------------------------------------------
def config_c1(L=8):
    """Classical Heisenberg ferromagnet, SC, J1 = 1; 16 temperatures 0.5-2.0 (lin)."""
    return cubic_model("sc", L, J=(1.0,)), ladder("lin", 0.5, 2.0, 16)


def config_c2(L=16, R=64):
    """FCC, J1 = -1 (AF), J2 = +0.2, J3 = +0.05, easy-axis single-ion 0.1 along z
    (Delta = -0.1); geometric ladder 0.05-1.5 J, R replicas."""
    return (cubic_model("fcc", L, J=(-1.0, 0.2, 0.05), K=-0.1), ladder("log", 0.05, 1.5, R))


def config_c4(L=64, R=32):
    """Chiral magnet, SC: J1 = 1, DM 0.3 along the bond, biquadratic 0.05, easy-plane
    single-ion 0.02 (Delta = +0.02); geometric ladder 0.1-1.2 J."""
    return (cubic_model("sc", L, J=(1.0,), B=0.05, D=0.3, K=0.02), ladder("log", 0.1, 1.2, R))


def config_c5(L=32, R=64):
    """Frustrated FCC: J1 = -1, J2 = +0.1, NN biquadratic 0.1, single-ion Delta = -0.05;
    geometric ladder 0.02-0.8 J."""
    return (cubic_model("fcc", L, J=(-1.0, 0.1), B=0.1, K=-0.05), ladder("log", 0.02, 0.8, R))


# ---- the batched workload: many uneven PT chains ------------------------------------------
_STRUCT_CACHE: dict = {}


def cubic_model_fast(kind: str, L: int, J=(1.0,)):
    """cubic_model for exchange-only models, reusing the bond structure of an earlier call with
    the same (kind, L, number of shells): slots are built once with marker values k+1 for shell k
    and relabelled with the couplings (the table builder's output does not depend on the values
    otherwise, except that zero couplings are dropped -- callers pass non-zero J)."""
    key = (kind, L, len(J))
    if key not in _STRUCT_CACHE:
        base = cubic_model(kind, L, J=tuple(float(k + 1) for k in range(len(J))))
        _STRUCT_CACHE[key] = base
    base = _STRUCT_CACHE[key]
    shell = np.rint(base.J.val).astype(int) - 1
    val = np.asarray(J, float)[shell]
    t = Table(base.J.z, base.J.cnt, base.J.nbr, np.ascontiguousarray(val))
    return Model(kind, base.L, base.nb, base.basis_frac, base.mag, t,
                 rows={"J": shell_rows(kind, list(J))})


def config_c3(n_chains=256, seed=7, Ls=(8, 12, 16, 24, 32, 48), kinds=("sc", "bcc", "fcc"),
              r_range=(16, 32), t_range=(0.5, 2.5)):
    """Batched uneven workload: per chain a lattice kind and L drawn uniformly, J1 = 1 ferro and
    J2 ~ U[-0.3, 0.3], R ~ U{16..32} replicas on a geometric ladder over t_range.
    Returns a list of (kind, L, J2, R, tems) drawn from one master seed."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_chains):
        kind = kinds[int(rng.integers(len(kinds)))]
        L = Ls[int(rng.integers(len(Ls)))]
        J2 = float(rng.uniform(-0.3, 0.3))
        R = int(rng.integers(r_range[0], r_range[1] + 1))
        out.append((kind, L, J2, R, ladder("log", t_range[0], t_range[1], R)))
    return out
