"""Input generation (host only): the neighbour-shell finder (esp_neighbor_shells, neighbors.F90:
48-168), coupling rows and .mcin blocks from per-shell parameters, and the spin-glass disorder
(esp_spin_glass, jij.F90:506-556), against the oracle restatement (oracle/setup_oracle.py) and
closed forms: shell radii and multiplicities of SC/BCC/FCC, the cubic tables the library builds
from its own shell rows, the symmetry J_ij = J_ji and the polar-method normals of MT19937."""
import math

import numpy as np
import pytest

from espins_amd import lattice as Lt
from espins_amd import mcin
from espins_amd import setup_gen as G
from oracle import oracle as O
from oracle import setup_oracle as S

CUBIC = {
    "sc": [(1.0, 6), (math.sqrt(2), 12), (math.sqrt(3), 8), (2.0, 6)],
    "bcc": [(math.sqrt(3) / 2, 8), (1.0, 6), (math.sqrt(2), 12), (math.sqrt(11) / 2, 24)],
    "fcc": [(1 / math.sqrt(2), 12), (1.0, 6), (math.sqrt(1.5), 24), (math.sqrt(2), 12)],
}


@pytest.mark.parametrize("kind", ["sc", "bcc", "fcc"])
def test_cubic_shells_closed_forms(kind):
    sh = G.neighbor_shells(np.eye(3), Lt.BASIS[kind], 4)
    for a in range(len(Lt.BASIS[kind])):
        got = [(s.radius, len(s.neighbours)) for s in sh[a]]
        for (r, m), (r0, m0) in zip(got, CUBIC[kind]):
            assert r == pytest.approx(r0, abs=1e-12) and m == m0
        # every listed neighbour is at the shell radius
        bf = np.array(Lt.BASIS[kind])
        for s in sh[a]:
            for b, cell in s.neighbours:
                assert np.linalg.norm(bf[b] + cell - bf[a]) == pytest.approx(s.radius, abs=1e-12)


def test_shells_match_oracle_order_triclinic():
    lat = np.array([[2.1, 0.0, 0.0], [0.4, 1.9, 0.0], [0.3, -0.2, 2.6]])
    bf = np.array([[0.0, 0.0, 0.0], [0.37, 0.52, 0.41]])
    mine = G.neighbor_shells(lat, bf, [4, 3])
    ref = S.shells(lat, bf @ lat, [4, 3])
    for a in range(2):
        assert len(mine[a]) == len(ref[a])
        for s, (r, lst) in zip(mine[a], ref[a]):
            assert s.radius == pytest.approx(r, abs=1e-12)
            assert s.neighbours == [(b, tuple(c)) for b, c in lst]


def _site_sets(t):
    N = len(t.cnt)
    return [sorted((int(t.nbr[l, i]), round(float(t.val[l, i]), 12)) for l in range(int(t.cnt[i])))
            for i in range(N)]


@pytest.mark.parametrize("kind,J", [("bcc", (1.0, 0.3)), ("fcc", (-1.0, 0.2, 0.05))])
def test_generated_rows_give_the_cubic_tables(kind, J):
    bf = Lt.BASIS[kind]
    sh = G.neighbor_shells(np.eye(3), bf, len(J))
    rows = G.coupling_rows(bf, [0] * len(bf), sh, [[[j]] for j in J])
    L = 4
    t = Lt.build_table(len(bf), np.array(bf, float), np.ones(len(bf)), (L, L, L), rows, 1,
                       builder=O.build_table_fn)
    ref = Lt.cubic_model(kind, L, J=J, builder=O.build_table_fn).J
    assert _site_sets(t) == _site_sets(ref)
    assert sorted(set(rows.shell.tolist())) == list(range(1, len(J) + 1))


def test_mcin_block_round_trip():
    bf = Lt.BASIS["bcc"]
    sh = G.neighbor_shells(np.eye(3), bf, 2)
    rows = G.coupling_rows(bf, [0, 1], sh, [[[0.0, 0.01], [0.01, 0.0]], [[0.002, 0.0], [0.0, 0.003]]])
    text = f"""
 Begin Unit_Cell_Cart
   1.0 0.0 0.0
   0.0 1.0 0.0
   0.0 0.0 1.0
 End Unit_Cell_Cart
 Begin Atoms_Frac
  A 0.0 0.0 0.0 1.0
  B 0.5 0.5 0.5 1.0
 End Atoms_Frac
 supercell_size = 3 3 3
{G.mcin_block("Jij_parameters", rows)}
"""
    inp = mcin.parse(text)
    assert inp.rows["J"].n == rows.n == 2 * 8 + 2 * 6
    assert np.allclose(inp.rows["J"].f2, rows.f2, atol=1e-8)
    assert np.allclose(inp.rows["J"].value, rows.value, rtol=1e-9)      # eV, as written


def test_polar_normals_closed_form():
    """The first pair of mtprng_normal from the MT19937 stream itself."""
    u = O.mt_stream(1234, 8).astype(float) / 4294967295.0
    k = 0
    while True:
        u1, u2 = 2 * u[k] - 1, 2 * u[k + 1] - 1
        r = u1 * u1 + u2 * u2
        if 0 < r < 1:
            break
        k += 2
    f = math.sqrt(-2 * math.log(r) / r)
    assert S.normals(1234, 2) == [u2 * f, u1 * f]


def test_spin_glass_matches_oracle_and_is_symmetric():
    m = Lt.cubic_model("sc", 4, J=(1.0, 0.5), builder=O.build_table_fn)
    t = m.J
    sigma = np.where(np.arange(t.z)[:, None] < t.cnt[None, :], 0.7, 0.0) * (1 + 0.1 * (t.val > 0.7))
    got = G.spin_glass(t, sigma, 987).val
    want = S.spin_glass(t.cnt, t.nbr, t.val, sigma, 987)
    assert np.array_equal(got, want)
    N = m.N
    for i in range(N):
        for l in range(t.cnt[i]):
            j = t.nbr[l, i] - 1
            back = [lj for lj in range(t.cnt[j]) if t.nbr[lj, j] == i + 1]
            assert got[back[0], j] == got[l, i]


def test_spin_glass_statistics():
    m = Lt.cubic_model("sc", 8, J=(1.0,), builder=O.build_table_fn)
    t = m.J
    sigma = np.where(np.arange(t.z)[:, None] < t.cnt[None, :], 0.5, 0.0)
    got = G.spin_glass(t, sigma, 42).val
    N = m.N
    x = np.array([(got[l, i] - t.val[l, i]) / 0.5 for i in range(N) for l in range(t.cnt[i])
                  if i + 1 < t.nbr[l, i]])
    assert len(x) == 3 * N
    assert abs(x.mean()) < 4 / math.sqrt(len(x)) and abs(x.std() - 1) < 0.06


def test_spin_glass_rejects_asymmetric_tables():
    m = Lt.cubic_model("sc", 3, J=(1.0,), builder=O.build_table_fn)
    t = m.J
    nbr = t.nbr.copy()
    nbr[0, 0] = 5 if nbr[0, 0] != 5 else 6      # site 1 now points at a site that does not point back
    with pytest.raises(ValueError):
        G.spin_glass(Lt.Table(t.z, t.cnt, nbr, t.val), np.ones_like(t.val), 1)
