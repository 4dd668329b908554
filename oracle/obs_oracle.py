"""CPU restatement of the reference's optional observables -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use this module, as
the checker of the HIP engine; nothing in espins_amd/ imports it.

Each function follows one reference routine on one configuration, sites in the reference
order (1-based site j has basis atom mod(j, nb) with 0 -> nb, i.e. 0-based index j mod nb):
  orderparam      orderparam_get           get_quant.F90:396-422
  staggered       staggered_get            get_quant.F90:366-392
  pm_counts       Pm_get                   get_quant.F90:425-493 (grid mcarlo.F90:432-456)
  shell_kind_norm shell_neighbors_get      jij.F90:449-503 (shell_num_nbors_kind)
  spincorr        spincorr_get             get_quant.F90:496-540
  sfactor         mcarlo_structure_factor  get_quant.F90:649-717
  binning         the binning analysis of mcarlo.F90:582-590 (per measurement) and
                  mcarlo.F90:635-656 (at the end of the run)
Pinned by closed forms in tests/test_obs_oracle.py (ordered states, Neel states, reciprocal
lattice vectors, box means of linear series).
"""
from __future__ import annotations

import math

import numpy as np


def pm_grid(mag_basis):
    """mcarlo.F90:432-445: m_max = sum |m_a| over the basis, m_min = -m_max, step 0.01."""
    m_max = float(sum(abs(float(x)) for x in mag_basis))
    m_min = -m_max
    nfreq = int(round((m_max - m_min) / 0.01)) + 1
    if nfreq == 1:
        nfreq = 2
    d_m = (m_max - m_min) / (nfreq - 1)
    m_array = np.array([m_min + k * d_m for k in range(nfreq)])
    return m_min, m_max, nfreq, m_array


def _nint(x):
    """Fortran nint: nearest integer, halves away from zero."""
    return int(math.floor(x + 0.5)) if x >= 0 else -int(math.floor(-x + 0.5))


def orderparam(s, nb, axes):
    """OP = sum_j S_j . axes(basis j) / N (signed); the reference accumulates |OP|."""
    N = len(s)
    op = 0.0
    for j in range(N):
        op += float(np.dot(s[j], axes[j % nb])) / N
    return op


def staggered(s, nb, coeff):
    """|sum_j coeff(basis j) S_j / N|."""
    N = len(s)
    v = np.zeros(3)
    for j in range(N):
        v = v + s[j] * coeff[j % nb] / N
    return float(math.sqrt(float(np.dot(v, v))))


def pm_counts(s, nb, mag_basis, axes=None, coeff=None):
    """Counts per bin of one configuration: rows m (|cell moment|), then the cell order
    parameter (signed) if axes, then the cell staggered moment if coeff (Pm_get adds 1/d_m per
    count; the caller scales)."""
    m_min, m_max, nfreq, _ = pm_grid(mag_basis)
    rows = 1 + (axes is not None) + (coeff is not None)
    h = np.zeros((rows, nfreq), np.int64)
    N = len(s)

    def put(row, x):
        k = max(_nint((x - m_min) / (m_max - m_min) * (nfreq - 1)) + 1, 1)
        h[row, min(k, nfreq) - 1] += 1

    for j in range(0, N, nb):
        cell = s[j:j + nb]
        m = cell.sum(axis=0)
        put(0, math.sqrt(float(np.dot(m, m))))
        r = 1
        if axes is not None:
            put(r, sum(float(np.dot(cell[i], axes[i])) for i in range(nb)))
            r += 1
        if coeff is not None:
            v = sum(cell[i] * coeff[i] for i in range(nb))
            put(r, math.sqrt(float(np.dot(v, v))))
    return h


def shell_kind_norm(species, cnt, nbr, nspecies):
    """shell_num_nbors_kind(shell, kind of neighbour, kind of site) as shell_neighbors_get
    leaves it: for every site in order, the per-kind neighbour counts overwrite the column of
    the site's kind (so the last site of each kind wins).  nbr is 1-based."""
    nshell, N = cnt.shape
    T = np.zeros((nshell, nspecies, nspecies))
    for i in range(N):
        for s in range(nshell):
            ck = np.zeros(nspecies)
            for l in range(cnt[s, i]):
                ck[species[nbr[s, l, i] - 1]] += 1
            T[s, :, species[i]] = ck
    return T


def spincorr(s, species, cnt, nbr, mag, nspecies):
    """spincorr_tmp and spincorr_abs_tmp [nshell][nsp][nsp] of one configuration; the
    reference accumulates tmp, |tmp| and abs_tmp."""
    T = shell_kind_norm(species, cnt, nbr, nspecies)
    nshell, N = cnt.shape
    tmp = np.zeros((nshell, nspecies, nspecies))
    tab = np.zeros((nshell, nspecies, nspecies))
    for sh in range(nshell):
        for i in range(N):
            a = species[i]
            for l in range(cnt[sh, i]):
                j = nbr[sh, l, i] - 1
                b = species[j]
                den = T[sh, a, b] * mag[i] * mag[j]
                d = float(np.dot(s[i], s[j]))
                tmp[sh, a, b] += d / den
                tab[sh, a, b] += abs(d) / den
    return tmp, tab


def sfactor(s, pos, qpts, polar=None):
    """Per q: (|F|^2, |F|^2 - NSF^2, NSF^2) with F = sum_j (S_j - (S_j.q/q^2) q) e^{i q.r_j} and
    NSF the real part of sum_j polar.(S_j - ...) e^{i q.r_j} / |polar| (a real variable
    assigned a complex sum keeps the real part); q = 0 contributes zeros."""
    out = np.zeros((3, len(qpts)))
    for k, q in enumerate(qpts):
        q = np.asarray(q, float)
        if not q.any():
            continue
        q2 = float(np.dot(q, q))
        F = np.zeros(3, complex)
        nsf = 0.0
        for j in range(len(s)):
            fac = complex(math.cos(float(np.dot(pos[j], q))), math.sin(float(np.dot(pos[j], q))))
            perp = s[j] - float(np.dot(s[j], q)) / q2 * q
            F = F + perp * fac
            if polar is not None:
                nsf = nsf + (float(np.dot(polar, perp)) * fac / math.sqrt(float(np.dot(polar, polar)))).real
        f2 = float(np.real(np.vdot(F, F)))
        out[0, k] = f2
        out[1, k] = f2 - nsf * nsf
        out[2, k] = nsf * nsf
    return out


def binning(series, levels):
    """Binning error estimate of one temperature: series[n][k] of the per-measurement values
    (E/N, |m|, ...), levels (binning_level).  Returns var[k][levels] as the reference leaves
    local_binning before output (the output prints sqrt(var / N))."""
    x = np.asarray(series, float)
    n, K = x.shape
    L = len(levels)
    box = np.zeros((K, L))
    acc = np.zeros((K, L))
    total = np.zeros(K)
    for step in range(1, n + 1):
        v = x[step - 1]
        total = total + v
        box = box + v[:, None]
        for l, lev in enumerate(levels):
            if step % (2 ** lev) == 0:
                box[:, l] = box[:, l] / 2 ** lev
                acc[:, l] = acc[:, l] + box[:, l] ** 2
                box[:, l] = 0.0
    out = np.zeros((K, L))
    for l, lev in enumerate(levels):
        m_l = n // 2 ** lev
        mean = (total - box[:, l]) / 2 ** lev / m_l
        out[:, l] = (acc[:, l] / m_l - mean ** 2) / (m_l - 1)
    return out
