"""Pins of the CPU restatement of the optional observables (oracle/obs_oracle.py) by closed
forms: ordered and Neel states, reciprocal-lattice vectors, box means of linear series.  Each
check is one a plausible slip fails (a dropped normalisation, a wrong basis index, a missing
projection perpendicular to q, the transposed kind table, an incomplete box kept)."""
import math

import numpy as np
import pytest

from espins_amd import lattice as Lt
from oracle import obs_oracle as B
from oracle import oracle as O


def ferro(N, mag=1.0):
    s = np.zeros((N, 3))
    s[:, 2] = mag
    return s


def sc_neel(L):
    p = Lt.site_positions("sc", L).astype(int)
    s = np.zeros((L ** 3, 3))
    s[:, 2] = np.where(p.sum(axis=1) % 2 == 0, 1.0, -1.0)
    return s


def bcc_neel(L):
    s = np.zeros((2 * L ** 3, 3))
    s[0::2, 2] = 1.0
    s[1::2, 2] = -1.0
    return s


def test_pm_grid_sizes():
    assert B.pm_grid([1.0])[2] == 201                     # nint(2 / 0.01) + 1
    assert B.pm_grid([1.0] * 4)[2] == 801
    m_min, m_max, n, arr = B.pm_grid([2.5, 2.5])
    assert (m_min, m_max, n) == (-5.0, 5.0, 1001)
    assert arr[0] == -5.0 and abs(arr[-1] - 5.0) < 1e-12 and abs(arr[500]) < 1e-12


def test_order_and_staggered_closed_forms():
    s = ferro(64)
    assert B.orderparam(s, 1, [(0, 0, 1)]) == pytest.approx(1.0, abs=1e-14)
    assert B.orderparam(s, 1, [(1, 0, 0)]) == 0.0
    assert B.staggered(s, 1, [1.0]) == pytest.approx(1.0, abs=1e-14)
    n = bcc_neel(3)
    assert B.orderparam(n, 2, [(0, 0, 1), (0, 0, -1)]) == pytest.approx(1.0, abs=1e-14)
    assert B.orderparam(n, 2, [(0, 0, 1), (0, 0, 1)]) == pytest.approx(0.0, abs=1e-14)
    assert B.staggered(n, 2, [1.0, -1.0]) == pytest.approx(1.0, abs=1e-14)
    assert B.staggered(n, 2, [1.0, 1.0]) == pytest.approx(0.0, abs=1e-14)
    assert B.staggered(n, 2, [0.5, -0.5]) == pytest.approx(0.5, abs=1e-14)


def test_pm_counts_closed_forms():
    # ferro, one atom per cell: every cell at |m| = 1, the last bin
    h = B.pm_counts(ferro(64), 1, [1.0])
    assert h.shape == (1, 201) and h[0, 200] == 64 and h.sum() == 64
    # Neel BCC: cell moment 0 (middle bin of the 401-point grid), cell OP 2 and cell
    # staggered moment 2 (the last bin: the grid ends at sum |m_a| = 2)
    h = B.pm_counts(bcc_neel(2), 2, [1.0, 1.0], axes=[(0, 0, 1), (0, 0, -1)], coeff=[1.0, -1.0])
    assert h.shape == (3, 401)
    assert h[0, 200] == 8 and h[1, 400] == 8 and h[2, 400] == 8
    # OP of the opposite sign lands in the first bin (signed, not |OP|)
    h = B.pm_counts(bcc_neel(2), 2, [1.0, 1.0], axes=[(0, 0, -1), (0, 0, 1)])
    assert h[1, 0] == 8


def test_spincorr_closed_forms():
    L = 4
    cnt, nbr = Lt.shell_tables("sc", L, 2, builder=O.build_table_fn)
    N = L ** 3
    sp = np.zeros(N, int)
    mag = np.ones(N)
    assert (cnt[0] == 6).all() and (cnt[1] == 12).all()
    t, ta = B.spincorr(ferro(N), sp, cnt, nbr, mag, 1)
    assert np.allclose(t[:, 0, 0], N, rtol=1e-14) and np.allclose(ta[:, 0, 0], N, rtol=1e-14)
    t, ta = B.spincorr(sc_neel(L), sp, cnt, nbr, mag, 1)
    assert t[0, 0, 0] == pytest.approx(-N, rel=1e-14)       # nearest neighbours antiparallel
    assert t[1, 0, 0] == pytest.approx(N, rel=1e-14)        # face diagonals parallel
    assert ta[0, 0, 0] == pytest.approx(N, rel=1e-14)
    # moments enter as 1/(m_i m_j): spins of length 2 with moments 2 give the same numbers
    t2, _ = B.spincorr(2 * sc_neel(L), sp, cnt, nbr, 2 * mag, 1)
    assert np.allclose(t2, t, rtol=1e-14)


def test_spincorr_kind_table_orientation():
    """BCC with one kind per basis atom: every nearest neighbour is of the other kind, so only
    the off-diagonal entries of the kind table are non-zero; a transposed or diagonal read
    would divide by zero."""
    L = 2
    cnt, nbr = Lt.shell_tables("bcc", L, 1, builder=O.build_table_fn)
    N = 2 * L ** 3
    kinds = np.arange(N) % 2
    T = B.shell_kind_norm(kinds, cnt, nbr, 2)
    assert T[0, 0, 1] == 8 and T[0, 1, 0] == 8 and T[0, 0, 0] == 0 and T[0, 1, 1] == 0
    t, _ = B.spincorr(ferro(N), kinds, cnt, nbr, np.ones(N), 2)
    assert t[0, 0, 1] == pytest.approx(N / 2, rel=1e-14) and t[0, 1, 0] == pytest.approx(N / 2, rel=1e-14)
    assert t[0, 0, 0] == 0.0 and t[0, 1, 1] == 0.0


def test_sfactor_closed_forms():
    L = 4
    N = L ** 3
    pos = Lt.site_positions("sc", L)
    tp = 2 * math.pi
    q = [(tp, 0, 0), (0, 0, tp), (tp / 4, 0, 0), (0, 0, 0)]
    sf = B.sfactor(ferro(N), pos, q, polar=(0, 0, 1))
    assert sf[0, 0] == pytest.approx(N * N, rel=1e-12)      # q on the reciprocal lattice, q perp S
    assert sf[2, 0] == pytest.approx(N * N, rel=1e-12)      # all of it along the polarisation
    assert abs(sf[1, 0]) < 1e-6
    assert abs(sf[0, 1]) < 1e-6                             # q parallel to S: no perpendicular part
    assert abs(sf[0, 2]) < 1e-6                             # sum of exp(i pi x / 2) over x = 0..3
    assert (sf[:, 3] == 0).all()                            # q = 0 skipped
    # Neel at the zone corner: F0 = N z, perpendicular part N (z - (q_z / q^2) q)
    sf = B.sfactor(sc_neel(L), pos, [(math.pi, math.pi, math.pi)])
    assert sf[0, 0] == pytest.approx(2 * N * N / 3, rel=1e-12)


def test_binning_box_means():
    x = np.arange(1, 17, dtype=float)[:, None]
    v = B.binning(x, [1])
    assert v[0, 0] == pytest.approx(3.0, rel=1e-14)          # boxes 1.5 .. 15.5, spread 21, / 7
    x = np.arange(1, 18, dtype=float)[:, None]              # 17: the incomplete box is dropped
    v = B.binning(x, [2])
    assert v[0, 0] == pytest.approx(20.0 / 3.0, rel=1e-14)
    v = B.binning(np.full((32, 2), 0.25), [1, 2, 3])
    assert np.allclose(v, 0.0, atol=1e-16)
