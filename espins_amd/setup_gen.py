"""Input generation for lattices beyond the built-in cubic ones (host only, no GPU).

* ``neighbor_shells``: the reference's neighbour-shell finder (neighbors_get,
  neighbors.F90:48-168) through ``esp_neighbor_shells``: shells of every atom of an arbitrary
  cell by increasing distance, multiplicities and neighbour lists in the reference's order.
* ``coupling_rows``: the coupling rows the reference's .mcin writer emits for per-kind-pair,
  per-shell parameters (mcin_write, neighbors.F90:647-1192): one row per neighbour of every
  atom, ``f1`` the atom, ``f2`` the neighbour's fractional position, the shell label ``sh``.
* ``format_rows`` / ``mcin_block``: the rows as ``f1=..:f2=..:jij=..:sh=..`` lines of a
  ``Jij_parameters`` (or ``Bij``/``Dij``) block that ``mcin.read`` and the Fortran driver read.
* ``spin_glass``: the spin-glass disorder of an exchange table (jij_spin_glass,
  jij.F90:506-556) through ``esp_spin_glass``.
"""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np

from . import _lib as E
from .lattice import Rows, Table


@dataclasses.dataclass
class Shell:
    radius: float
    neighbours: list          # [(atom b, (l, m, n))] in the reference's order


def neighbor_shells(lattice, basis_frac, nshell, tol=1e-5, cap_multi=512):
    """lattice[3][3] (rows = lattice vectors, Cartesian), basis_frac[nb][3]; nshell per atom
    (int or list).  Returns shells[a] = [Shell, ...] for every atom a."""
    lat = E.f64(lattice)
    bf = np.asarray(basis_frac, float)
    nb = len(bf)
    pos = E.f64(bf @ lat)
    ns = E.i32(np.broadcast_to(np.asarray(nshell, np.int32), (nb,)))
    cap_shell = max(1, int(ns.max()))
    dnn = np.zeros((nb, cap_shell))
    multi = np.zeros((nb, cap_shell), np.int32)
    lst = np.zeros((nb, cap_shell, cap_multi, 4), np.int32)
    E.check(E.lib.esp_neighbor_shells(nb, E.ptr(lat, C.c_double), E.ptr(pos, C.c_double),
                                      E.ptr(ns, C.c_int), float(tol), cap_shell, cap_multi,
                                      E.ptr(dnn, C.c_double), E.ptr(multi, C.c_int),
                                      E.ptr(lst, C.c_int)))
    out = []
    for a in range(nb):
        sh = []
        for s in range(int(ns[a])):
            m = int(multi[a, s])
            sh.append(Shell(float(dnn[a, s]), [(int(lst[a, s, k, 0]), tuple(int(x) for x in lst[a, s, k, 1:]))
                                                 for k in range(m)]))
        out.append(sh)
    return out


def coupling_rows(basis_frac, kinds, shells, params, dvec=None) -> Rows:
    """Rows for parameters params[s][kind a][kind b] (shell s, 0-based) on every neighbour of
    every atom; zero parameters give no row.  With dvec(a, b, cell) -> unit vector the rows
    carry DM vectors (Dij blocks)."""
    bf = np.asarray(basis_frac, float)
    f1, f2, val, sh, dv = [], [], [], [], []
    for a, atom_shells in enumerate(shells):
        for s, shell in enumerate(atom_shells):
            if s >= len(params):
                break
            for b, cell in shell.neighbours:
                v = float(params[s][kinds[a]][kinds[b]])
                if v == 0.0:
                    continue
                f1.append(bf[a])
                f2.append(bf[b] + np.array(cell, float))
                val.append(v)
                sh.append(s + 1)
                if dvec is not None:
                    dv.append(dvec(a, b, cell))
    r = Rows(np.array(f1, float).reshape(-1, 3), np.array(f2, float).reshape(-1, 3), np.array(val, float))
    r.shell = np.array(sh, np.int32)
    if dvec is not None:
        r.dvec = np.array(dv, float).reshape(-1, 3)
    return r


def format_rows(rows: Rows, key="jij", sigma=None):
    """``f1=x,y,z:f2=x,y,z:<key>=v:sh=s[:sig=w]`` lines (param_get_parameters,
    parameters.F90:2558-2846)."""
    lines = []
    shell = getattr(rows, "shell", None)
    for k in range(rows.n):
        a, b = rows.f1[k], rows.f2[k]
        t = (f"f1={a[0]:.8f},{a[1]:.8f},{a[2]:.8f}:f2={b[0]:.8f},{b[1]:.8f},{b[2]:.8f}:"
             f"{key}={rows.value[k]:.10g}")
        if shell is not None:
            t += f":sh={int(shell[k])}"
        if sigma is not None:
            t += f":sig={float(sigma[k]):.10g}"
        lines.append(t)
    return lines


def mcin_block(name, rows: Rows, key="jij", sigma=None):
    return "\n".join([f" Begin {name}"] + ["  " + l for l in format_rows(rows, key, sigma)] + [f" End {name}"])


def spin_glass(table: Table, sigma, seed: int) -> Table:
    """A copy of the exchange table with the reference's Gaussian disorder: sigma[z][N] per
    slot (internal units), MT19937 stream of seed."""
    val = np.array(table.val, np.float64, copy=True)
    sg = E.f64(sigma)
    cnt, nbr = E.i32(table.cnt), E.i32(table.nbr)
    bad = np.zeros(2, np.int32)
    rc = E.lib.esp_spin_glass(len(cnt), table.z, E.ptr(cnt, C.c_int), E.ptr(nbr, C.c_int),
                              E.ptr(val, C.c_double), E.ptr(sg, C.c_double), int(seed), E.ptr(bad, C.c_int))
    if rc != E.ESP_OK:
        raise ValueError(f"spin glass: asymmetric pair ({bad[0]}, {bad[1]})" if rc == -5 else f"rc {rc}")
    return Table(table.z, table.cnt, table.nbr, val)
