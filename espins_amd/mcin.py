"""Reader for the subset of the ESpinS ``.mcin`` input format that the hot path needs.

Follows param_in_file / param_get_keyword / param_get_atoms_moments / param_get_parameters
(parameters.F90:1724-1805, 1808-1873, 2344-2555, 2558-2846): case-insensitive ``key = v``,
``key : v`` or ``key v`` lines with ``!``/``#`` comments, ``begin X ... end X`` blocks with an
optional unit line, coupling rows ``f1=x,y,z:f2=x,y,z:jij=v[:sh=n][:sig=s]`` (or ``c1=``/``c2=``
Cartesian), atoms sorted by species in order of first appearance.  Energies are converted
to the internal unit (Kelvin) exactly as jij.F90:163, bij.F90:126, dij.F90:126 and
parameters.F90:958,994 do; temperature ladders follow parameters.F90:271-326.
"""
from __future__ import annotations

import dataclasses
import math
import re
from typing import Optional

import numpy as np

from .lattice import Model, Rows, Table, build_table, ladder

K_B = 8.6173303e-5            # eV/K            constants.F90:48
HART = 27.21138602            # eV              constants.F90:49
BOHR_MAGN = 5.7883818012e-5   # eV/T            constants.F90:50
BOHR = 0.52917721092          # Angstrom


class McinError(ValueError):
    pass


@dataclasses.dataclass
class McinInput:
    lattice: np.ndarray                 # rows a1, a2, a3 (Angstrom)
    labels: list                        # species-sorted atom labels
    basis_frac: np.ndarray              # [nb,3]
    moments: np.ndarray                 # [nb]
    keywords: dict
    rows: dict                          # "J"/"B"/"D" -> Rows (values in eV)
    si: Optional[tuple] = None          # (par[nb] eV, axis[nb,3])
    field: Optional[tuple] = None       # (par[nb] Tesla, axis[nb,3])

    # ---- keyword helpers ----------------------------------------------------------
    def kw(self, name, default=None, kind=str):
        v = self.keywords.get(name)
        if v is None:
            return default
        if kind is bool:
            return v.strip().lower().lstrip(".").startswith("t")
        if kind is int:
            return int(float(v.split()[0]))
        if kind is float:
            return float(v.split()[0].replace("d", "e"))
        return v.strip()

    @property
    def supercell(self):
        v = self.keywords.get("supercell_size")
        return tuple(int(x) for x in v.replace(",", " ").split()) if v else (4, 4, 4)

    def temperatures(self):
        """parameters.F90:271-326 (+ ascending sort for PT, parameters.F90:601)."""
        mode = self.kw("tems_mode", "lin").lower()
        n = self.kw("tems_num", 1, int)
        if mode.startswith("man"):
            t = np.array([float(x) for x in self.keywords["tems"].replace(",", " ").split()][:n])
        else:
            t = ladder(mode[:3], self.kw("tem_start", 5.0, float), self.kw("tem_end", 5.0, float), n)
        if self.kw("pt", False, bool):
            t = np.sort(t)
        return t


def _strip(line: str) -> str:
    for c in "!#":
        k = line.find(c)
        if k >= 0:
            line = line[:k]
    return line.strip()


def parse(text: str) -> McinInput:
    lines = [_strip(l).lower() for l in text.splitlines()]
    blocks: dict = {}
    keywords: dict = {}
    i = 0
    while i < len(lines):
        l = lines[i]
        if not l:
            i += 1
            continue
        if l.startswith("begin"):
            name = l.split(None, 1)[1].strip() if len(l.split()) > 1 else ""
            body = []
            i += 1
            while i < len(lines) and not lines[i].startswith("end"):
                if lines[i]:
                    body.append(lines[i])
                i += 1
            if i >= len(lines):
                raise McinError(f"begin {name} without end")
            if name in blocks:
                raise McinError(f"block {name} found more than once")
            blocks[name] = body
            i += 1
            continue
        m = re.match(r"^([a-z_0-9]+)\s*[=:]?\s*(.*)$", l)
        if not m:
            raise McinError(f"cannot parse line: {l}")
        k, v = m.group(1), m.group(2).strip()
        if k in keywords:
            raise McinError(f"keyword {k} found more than once")
        keywords[k] = v
        i += 1

    if "unit_cell_cart" not in blocks:
        raise McinError("Did not find the cell information in the input file")
    cell = [b for b in blocks["unit_cell_cart"] if not b.startswith(("ang", "bohr"))]
    lat = np.array([[float(x) for x in r.split()[:3]] for r in cell])
    if blocks["unit_cell_cart"][0].startswith("bohr"):
        lat *= BOHR
    if lat.shape != (3, 3):
        raise McinError("Not correct the cell information in the input file")

    # atoms, species-sorted (parameters.F90:2470-2540)
    frac = "atoms_frac" in blocks
    if frac == ("atoms_cart" in blocks):
        raise McinError("need exactly one of atoms_frac / atoms_cart")
    ab = blocks["atoms_frac" if frac else "atoms_cart"]
    scale = 1.0
    if ab and ab[0].split()[0] in ("ang", "bohr"):
        scale = BOHR if ab[0].startswith("bohr") else 1.0
        ab = ab[1:]
    raw = []
    for r in ab:
        p = r.split()
        pos = np.array([float(x) for x in p[1:4]])
        if not frac:
            pos = np.linalg.solve(lat.T, pos * scale)
        for k in range(3):
            if abs(pos[k] - round(pos[k])) < 1e-5:
                pos[k] = round(pos[k])
        pos = pos - np.floor(pos)
        raw.append((p[0], pos, float(p[4])))
    order = []
    for lab, _, _ in raw:
        if lab not in order:
            order.append(lab)
    atoms = [a for lab in order for a in raw if a[0] == lab]
    labels = [a[0] for a in atoms]
    basis = np.array([a[1] for a in atoms])
    moments = np.array([a[2] for a in atoms])

    def rows_of(name, key, with_vec=False):
        if name not in blocks:
            return None
        body = blocks[name]
        econv = 1.0
        lconv = 1.0
        if body and not ("f1=" in body[0] or "c1=" in body[0]):
            units = body[0].replace(",", " ").split()
            for u in units:
                if u == "ryd":
                    econv = HART * 0.5
                elif u == "bohr":
                    lconv = BOHR
                elif u not in ("ev", "ang"):
                    raise McinError(f"Units in block {name} not recognised")
            body = body[1:]
        f1, f2, val, sh, sig = [], [], [], [], []
        for r in body:
            parts = [p.strip() for p in r.replace(" ", "").split(":")]
            def site(p, tag):
                if p.startswith(f"f{tag}="):
                    return np.array([float(x) for x in p[3:].split(",")])
                if p.startswith(f"c{tag}="):
                    c = np.array([float(x) for x in p[3:].split(",")]) * lconv
                    return np.linalg.solve(lat.T, c)
                raise McinError(f"malformed positions definition: {r}")
            f1.append(site(parts[0], 1))
            f2.append(site(parts[1], 2))
            if not parts[2].startswith(key + "="):
                raise McinError(f"malformed parameters definition: {r}")
            val.append(float(parts[2][len(key) + 1:]))
            s_, g_ = 0, 0.0                 # optional sh= and sig= (parameters.F90:2751-2766)
            for p in parts[3:]:
                if p.startswith("sh="):
                    s_ = int(p[3:])
                elif p.startswith("sig="):
                    g_ = float(p[4:])
            sh.append(s_)
            sig.append(g_)
        out = Rows(np.array(f1).reshape(-1, 3), np.array(f2).reshape(-1, 3), np.array(val) * econv)
        out.shell = np.array(sh, np.int32)
        out.sigma = np.array(sig) * econv
        return out

    rows = {}
    rj = rows_of("jij_parameters", "jij")
    if rj is not None:
        rows["J"] = rj
    if "bij_parameters" in blocks:
        rows["B"] = rows_of("bij_parameters", "bij")
    if "dij_parameters" in blocks:
        rd = rows_of("dij_parameters", "dij")
        vb = blocks.get("dij_vectors_cart") or blocks.get("dij_vectors_frac")
        if vb is None:
            raise McinError("no dij_vectors_frac or dij_vectors_cart block found")
        vecs = np.array([[float(x) for x in v.split()[:3]] for v in vb
                         if not v.startswith(("ang", "bohr"))])
        if "dij_vectors_frac" in blocks:
            vecs = vecs @ lat
        if len(vecs) != rd.n:
            raise McinError("Wrong number of lines in dij_vectors block")
        vecs = vecs / np.linalg.norm(vecs, axis=1, keepdims=True)
        rd.dvec = vecs
        rows["D"] = rd

    def axes(name):
        b = blocks.get(name + "_cart") or blocks.get(name + "_frac")
        if b is None:
            return None
        b = [x for x in b if not x.startswith(("ang", "bohr"))]
        lab = [x.split()[0] for x in b]
        ax = np.array([[float(v) for v in x.split()[1:4]] for x in b])
        par = np.array([float(x.split()[4]) for x in b])
        if (name + "_frac") in blocks:
            ax = ax @ lat
        ax = ax / np.linalg.norm(ax, axis=1, keepdims=True)
        # sorted to the species order (lsort=.true.): k-th line of a label -> k-th atom of it
        out_ax = np.zeros((len(labels), 3))
        out_par = np.zeros(len(labels))
        used = {}
        for k, l in enumerate(lab):
            idx = [i for i, x in enumerate(labels) if x == l]
            n = used.get(l, 0)
            if n >= len(idx):
                raise McinError(f"Wrong number of lines in {name} block")
            out_ax[idx[n]] = ax[k]
            out_par[idx[n]] = par[k]
            used[l] = n + 1
        return out_par, out_ax

    inp = McinInput(lat, labels, basis, moments, keywords, rows)
    if inp.kw("ham_singleion", False, bool):
        inp.si = axes("singleion_axes")
        if inp.si is None:
            raise McinError("ham_singleion is true but no singleion_axes block found")
    if inp.kw("ham_field", False, bool):
        inp.field = axes("field_axes")
        if inp.field is None:
            raise McinError("ham_field is true but no field_axes block found")
    if not inp.kw("ham_bij", False, bool):
        rows.pop("B", None)
    if not inp.kw("ham_dij", False, bool):
        rows.pop("D", None)
    return inp


def read(path: str) -> McinInput:
    with open(path) as f:
        return parse(f.read())


def _spin_glass(inp, L, ob, J, builder):
    """jij_spin_glass (jij.F90:506-556) on the J table: the sigma of each slot is the sig= of the
    row that made it (a table of row numbers over the same rows, zero-J rows dropped like the J
    table), over k_B (jij.F90:164); the disorder comes from spin_glass_seed."""
    from .setup_gen import spin_glass
    rows = inp.rows["J"]
    seed = inp.kw("spin_glass_seed", None, int)
    if seed is None:
        raise McinError("spin_glass needs spin_glass_seed here (the stream must be reproducible)")
    if (rows.sigma < 0).any():
        raise McinError("sig in jij_parameters block must be positive")
    ids = np.where(rows.value != 0.0, np.arange(1, rows.n + 1), 0).astype(float)
    nb = len(inp.labels)
    t = build_table(nb, inp.basis_frac, np.ones(nb), L, Rows(rows.f1, rows.f2, ids), 1, ob, builder)
    if t.z != J.z:
        raise McinError("spin glass: row table does not match the J table")
    idx = np.rint(t.val).astype(int)
    sigma = np.where(idx > 0, rows.sigma[np.maximum(idx, 1) - 1] / K_B, 0.0)
    return spin_glass(J, sigma, seed)


def to_model(inp: McinInput, supercell=None, builder=None, open_boundary=None) -> Model:
    """Build the reference's tables (internal units) for this input."""
    L = tuple(supercell or inp.supercell)
    nb = len(inp.labels)
    ob = inp.kw("boundary", "periodic").startswith("open") if open_boundary is None else open_boundary
    N = nb * L[0] * L[1] * L[2]
    mag = np.tile(inp.moments, L[0] * L[1] * L[2])

    def conv(r: Rows):
        return Rows(r.f1, r.f2, r.value / K_B, r.dvec)

    J = build_table(nb, inp.basis_frac, inp.moments, L, conv(inp.rows["J"]), 1, ob, builder) \
        if "J" in inp.rows else Table(0, np.zeros(N, np.int32), np.zeros((0, N), np.int32),
                                      np.zeros((0, N)))
    if "J" in inp.rows and inp.kw("spin_glass", ".false.").lstrip(".").startswith("t"):
        J = _spin_glass(inp, L, ob, J, builder)
    m = Model("mcin", L, nb, inp.basis_frac, mag, J, rows=dict(inp.rows))
    if "B" in inp.rows:
        m.B = build_table(nb, inp.basis_frac, inp.moments, L, conv(inp.rows["B"]), 1, ob, builder)
    if "D" in inp.rows:
        m.D = build_table(nb, inp.basis_frac, inp.moments, L, conv(inp.rows["D"]), 3, ob, builder)
    if inp.si is not None:
        m.si_par = inp.si[0] / K_B
        m.si_axis = inp.si[1]
    if inp.field is not None:
        m.h_par = inp.field[0] * BOHR_MAGN / K_B
        m.h_axis = inp.field[1]
    return m
