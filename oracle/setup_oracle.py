"""CPU restatement of the reference's input-generation and disorder routines -- TEST
INFRASTRUCTURE ONLY (tests/ use it as the checker; nothing in espins_amd/ imports it).

  shells      neighbors_get + neighbors_supercell_sort + internal_maxloc
              (neighbors.F90:48-168, 1266-1360), plain loops
  spin_glass  jij_spin_glass (jij.F90:506-556) with mtprng_normal (mtprng.f90:381-407) on the
              oracle's MT19937 stream (pinned to the reference's own mtprng.f90 build)
Pinned by closed forms in tests/test_setup_gen.py (cubic shells, symmetry, first normals).
"""
from __future__ import annotations

import math

import numpy as np

from oracle import oracle as O


def sorted_cells(lat, ns=5):
    cells = [(0, 0, 0)]
    for l in range(-ns, ns + 1):
        for m in range(-ns, ns + 1):
            for n in range(-ns, ns + 1):
                if (l, m, n) != (0, 0, 0):
                    cells.append((l, m, n))
    dist = [math.sqrt(sum((c[0] * lat[0][k] + c[1] * lat[1][k] + c[2] * lat[2][k]) ** 2 for k in range(3)))
            for c in cells]
    out = [None] * len(cells)
    d = list(dist)
    for pos in range(len(cells) - 1, -1, -1):
        g = max(range(len(d)), key=lambda k: (d[k], -k))          # first maximum
        lo = min(k for k in range(len(d)) if abs(d[k] - d[g]) < 1e-8)
        out[pos] = cells[lo]
        d[lo] = -1.0
    return out


def shells(lat, pos, nshell, tol=1e-5):
    lat = [list(map(float, r)) for r in lat]
    cells = sorted_cells(lat)
    img = [[c[0] * lat[0][k] + c[1] * lat[1][k] + c[2] * lat[2][k] for k in range(3)] for c in cells]
    nb = len(pos)

    def dist(a, b, q):
        return math.sqrt(sum((pos[a][k] - (pos[b][k] + img[q][k])) ** 2 for k in range(3)))

    out = []
    for a in range(nb):
        d0, d1 = 0.0, 99999999.0
        res = []
        for s in range(nshell[a]):
            counter = 0
            for q in range(len(cells)):
                for b in range(nb):
                    r = dist(a, b, q)
                    if r > tol and r > d0 + tol:
                        if r < d1 - tol:
                            d1, counter = r, 0
                        if d1 - tol < r < d1 + tol:
                            counter += 1
            res.append((d1, counter))
            d0, d1 = d1, 99999999.0
        full = []
        for d1, m in res:
            lst = []
            for q in range(len(cells)):
                for b in range(nb):
                    if len(lst) < m and d1 - tol <= dist(a, b, q) <= d1 + tol:
                        lst.append((b, cells[q]))
            full.append((d1, lst))
        out.append(full)
    return out


def normals(seed, n):
    """The first n numbers of mtprng_normal on a fresh MT19937 stream of seed."""
    u = O.mt_stream(seed, 8 * n + 64).astype(np.float64) * (1.0 / 4294967295.0)
    out, k = [], 0
    while len(out) < n:
        u1 = -1.0 + 2.0 * u[k]
        u2 = -1.0 + 2.0 * u[k + 1]
        k += 2
        rad = u1 * u1 + u2 * u2
        if not (0.0 < rad < 1.0):
            continue
        f = math.sqrt(-2.0 * math.log(rad) / rad)
        out += [u2 * f, u1 * f]
    return out[:n]


def spin_glass(cnt, nbr, val, sigma, seed):
    N = len(cnt)
    val = np.array(val, float, copy=True)
    npair = sum(1 for i in range(1, N + 1) for l in range(cnt[i - 1]) if i < nbr[l][i - 1])
    g = iter(normals(seed, npair))
    for i in range(1, N + 1):
        for l in range(cnt[i - 1]):
            j = nbr[l][i - 1]
            if i >= j:
                continue
            val[l][i - 1] = next(g) * sigma[l][i - 1] + val[l][i - 1]
            for lj in range(cnt[j - 1]):
                if nbr[lj][j - 1] == i:
                    val[lj][j - 1] = val[l][i - 1]
                    break
            else:
                raise ValueError("asymmetric pair")
    return val
