"""Pins of the CPU oracle (oracle/espins_oracle.c) against the reference's golden vectors,
known answers, closed forms, brute force and exact distributions.  CPU only."""
import json
import math
import os

import numpy as np
import pytest

from oracle import oracle as O
from espins_amd import lattice as Lt
from espins_amd import mcin
from tests.conftest import GOLDEN

K = Lt.K_B


# ---------------------------------------------------------------- RNG -----------------------
def test_mt19937_known_answer():
    # The C++ standard fixes the 10000th output of mt19937 seeded 5489 (init_genrand).
    s = O.mt_stream(5489, 10000)
    assert s[:3].tolist() == [3499211612, 581869302, 3890346734]
    assert int(s[-1]) == 4123659995


def test_mt19937_matches_reference_mtprng_golden():
    g = json.load(open(os.path.join(GOLDEN, "mt19937_ref.json")))
    for seed in ("5489", "12345", "2024", "1"):
        u = g[seed]["u32"]
        ours = O.mt_stream(int(seed), len(u))
        assert ours.tolist() == u
        st = O.mt_state(int(seed))
        r2 = [O.lib.orc_mt_real2(st) for _ in range(len(u))]
        assert np.array_equal(np.array(r2), np.array(g[seed]["real2"]))   # bit-exact


def test_mt19937_matches_numpy_legacy_seeding():
    # numpy's RandomState(int) seeds with init_genrand: an independent implementation
    rs = np.random.RandomState(987654)
    raw = rs._bit_generator.random_raw(2000).astype(np.uint64)
    assert np.array_equal(raw, O.mt_stream(987654, 2000).astype(np.uint64))


def test_philox4x32_10_known_answers():
    # Random123 known-answer vectors for philox4x32_10
    kat = [
        ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
        ((0xffffffff,) * 4, (0xffffffff, 0xffffffff), (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
        ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
         (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
    ]
    for ctr, key, want in kat:
        assert O.philox(ctr, key).tolist() == list(want)


# ---------------------------------------------------------------- energies ------------------
@pytest.mark.parametrize("name", ["ex01_cubic.mcin", "ex02_cubic.mcin", "ex03_fef3.mcin",
                                  "ex04_bmno.mcin", "ex05_snwo.mcin"])
def test_published_initial_energies(name):
    """Ferro initial energies printed by the reference's own example runs."""
    ref = json.load(open(os.path.join(GOLDEN, "examples", "published_energies.json")))[name]
    inp = mcin.read(os.path.join(GOLDEN, "examples", name))
    m = mcin.to_model(inp, builder=O.build_table_fn)
    E = O.Oracle(m).energy(Lt.ferro_spins(m.N, 1, m.mag)[0])
    e0 = ref["E0"]
    half_digit = 0.5 * 10 ** (math.floor(math.log10(abs(e0))) + 1 - ref["digits"])
    assert abs(E - e0) <= half_digit, (E, e0)


def _closed_ferro(kind, L, J, B=0.0, K_si=0.0):
    """E_ferro = -(N/2) sum_s z_s J_s + (N/2) z_1 B + N Delta (a.z)^2 with a = z."""
    N = len(Lt.BASIS[kind]) * L ** 3
    z = Lt.COORDINATION[kind]
    return -(N / 2) * sum(z[k] * J[k] for k in range(len(J))) + (N / 2) * z[0] * B + N * K_si


@pytest.mark.parametrize("kind,L,J,B,Ksi", [
    ("sc", 8, (1.0,), 0.0, 0.0),
    ("fcc", 4, (-1.0, 0.2, 0.05), 0.0, -0.1),
    ("sc", 6, (1.0,), 0.05, 0.02),
    ("fcc", 4, (-1.0, 0.1), 0.1, -0.05),
    ("bcc", 4, (1.0, -0.3, 0.1), 0.0, 0.0),
])
def test_ferro_closed_form(kind, L, J, B, Ksi):
    m = Lt.cubic_model(kind, L, J=J, B=B, D=0.3 if B else 0.0, K=Ksi, builder=O.build_table_fn)
    E = O.Oracle(m).energy(Lt.ferro_spins(m.N, 1)[0])
    want = _closed_ferro(kind, L, J, B, Ksi)
    assert abs(E - want) <= 1e-12 * max(1.0, abs(want)) * 50


def test_neel_flips_the_exchange_sign():
    L = 6
    m = Lt.cubic_model("sc", L, J=(1.0,), builder=O.build_table_fn)
    s = Lt.ferro_spins(m.N, 1)[0]
    idx = np.arange(m.N)
    x, y, z = idx % L, (idx // L) % L, idx // (L * L)
    s[:, 2] = np.where((x + y + z) % 2 == 0, 1.0, -1.0)
    assert O.Oracle(m).energy(s) == pytest.approx(+3.0 * m.N, rel=1e-14)


def test_dm_planar_spiral_closed_form():
    """Spins rotating by q per cell along x in the yz-plane: with D along the bond, each x bond
    contributes D (S_i x S_j).x = D sin(q) per bond direction and the total is N D sin(q)."""
    L, D, q = 8, 0.3, 2 * math.pi / 8
    m = Lt.cubic_model("sc", L, J=(0.0,), D=D, builder=O.build_table_fn)
    idx = np.arange(m.N)
    x = idx % L
    s = np.zeros((m.N, 3))
    s[:, 1] = np.cos(q * x)
    s[:, 2] = np.sin(q * x)
    E = O.Oracle(m).energy(s)
    assert E == pytest.approx(m.N * D * math.sin(q), rel=1e-12)


def _brute_energy(kind, L, J, B, D, Ksi, h, s):
    """O(N z) energy from lattice positions and wrapped displacements -- independent of the
    table builder's index arithmetic."""
    basis = np.array(Lt.BASIS[kind])
    nb = len(basis)
    sv = Lt.shell_vectors(kind, max(len(J), 1))
    E = 0.0
    for i in range(nb * L ** 3):
        a = i % nb
        c = i // nb
        cell = np.array([c % L, (c // L) % L, c // (L * L)])
        for k, shell in enumerate(sv[a]):
            for b, off, v in shell:
                cj = (cell + np.array(off)) % L
                j = ((cj[2] * L + cj[1]) * L + cj[0]) * nb + b
                d = s[i] @ s[j]
                if k < len(J):
                    E += -J[k] * d
                if k == 0:
                    E += B * d * d + D * (v / np.linalg.norm(v)) @ np.cross(s[i], s[j])
        E += 2 * Ksi * s[i][2] ** 2 + 2 * h * s[i][0]
    return E / 2


@pytest.mark.parametrize("kind,L", [("sc", 3), ("bcc", 2), ("fcc", 2)])
def test_energy_matches_brute_force(kind, L):
    J, B, D, Ksi, h = (1.0, -0.4), 0.15, 0.3, -0.2, 0.7
    m = Lt.cubic_model(kind, L, J=J, B=B, D=D, K=Ksi, field=h, field_axis=(1, 0, 0),
                       builder=O.build_table_fn)
    s = Lt.random_spins(m.N, 1, 5)[0]
    E = O.Oracle(m).energy(s)
    assert E == pytest.approx(_brute_energy(kind, L, J, B, D, Ksi, h, s), rel=1e-12, abs=1e-12)


@pytest.mark.parametrize("term", ["J", "B", "D", "K", "h", "all"])
def test_delta_energy_equals_energy_difference(term):
    kw = dict(J=(1.0, 0.3) if term in ("J", "all") else (0.0,))
    if term in ("B", "all"):
        kw["B"] = 0.2
    if term in ("D", "all"):
        kw["D"] = 0.4
    if term in ("K", "all"):
        kw["K"] = -0.3
    if term in ("h", "all"):
        kw["field"] = 0.5
    m = Lt.cubic_model("fcc", 2, builder=O.build_table_fn, **kw)
    o = O.Oracle(m)
    rng = np.random.default_rng(3)
    s = Lt.random_spins(m.N, 1, 11)[0]
    for _ in range(40):
        i = int(rng.integers(m.N))
        new = rng.normal(size=3)
        new /= np.linalg.norm(new)
        e0 = o.energy(s)
        t = s.copy()
        t[i] = new
        ds = t[i] - s[i]
        de = o.delta_energy(i, t, ds)
        assert de == pytest.approx(o.energy(t) - e0, abs=1e-11)


# ---------------------------------------------------------------- tables ---------------------
@pytest.mark.parametrize("kind", ["sc", "bcc", "fcc"])
def test_shell_multiplicities_and_symmetry(kind):
    m = Lt.cubic_model(kind, 4, J=(1.0, 0.5, 0.25), builder=O.build_table_fn)
    z = Lt.COORDINATION[kind]
    assert m.J.z == sum(z)
    assert (m.J.cnt == sum(z)).all()
    # symmetric bonds (jij.F90:324-371): j in nbr(i) <=> i in nbr(j), same value
    pairs = {}
    for l in range(m.J.z):
        for i in range(m.N):
            pairs.setdefault((i, m.J.nbr[l, i] - 1), []).append(m.J.val[l, i])
    for (i, j), v in pairs.items():
        assert sorted(pairs[(j, i)]) == sorted(v)


def test_product_table_builder_matches_oracle():
    for kind, kw in [("fcc", dict(J=(-1.0, 0.2, 0.05))), ("sc", dict(J=(1.0,), B=0.05, D=0.3))]:
        a = Lt.cubic_model(kind, 4, **kw)
        b = Lt.cubic_model(kind, 4, builder=O.build_table_fn, **kw)
        for t1, t2 in [(a.J, b.J), (a.B, b.B), (a.D, b.D)]:
            if t1 is None:
                assert t2 is None
                continue
            assert t1.z == t2.z
            assert np.array_equal(t1.cnt, t2.cnt) and np.array_equal(t1.nbr, t2.nbr)
            assert np.array_equal(t1.val, t2.val)


def test_open_boundary_drops_bonds():
    rows = Lt.shell_rows("sc", [1.0])
    L = (3, 3, 3)
    t_open = Lt.build_table(1, np.zeros((1, 3)), np.ones(1), L, rows, 1, open_boundary=True)
    t_oracle = Lt.build_table(1, np.zeros((1, 3)), np.ones(1), L, rows, 1, open_boundary=True,
                              builder=O.build_table_fn)
    assert np.array_equal(t_open.cnt, t_oracle.cnt) and np.array_equal(t_open.nbr, t_oracle.nbr)
    # corner site 0 keeps 3 bonds, centre site 13 keeps 6 (jij.F90:264-267)
    assert t_open.cnt[0] == 3 and t_open.cnt[13] == 6


# ---------------------------------------------------------------- Metropolis -----------------
def _one_site_model(field=0.0, Ksi=0.0):
    class M:
        pass
    m = M()
    m.N, m.nb, m.mag = 1, 1, np.ones(1)
    m.J = Lt.Table(0, np.zeros(1, np.int32), np.zeros((0, 1), np.int32), np.zeros((0, 1)))
    m.B = m.D = None
    m.si_par = np.array([Ksi]) if Ksi else None
    m.si_axis = np.array([[0.0, 0.0, 1.0]]) if Ksi else None
    m.h_par = np.array([field]) if field else None
    m.h_axis = np.array([[0.0, 0.0, 1.0]]) if field else None
    return m


def _dimer_model(J=0.0, B=0.0):
    class M:
        pass
    m = M()
    m.N, m.nb, m.mag = 2, 1, np.ones(2)
    mk = lambda v: Lt.Table(1, np.ones(2, np.int32), np.array([[2, 1]], np.int32), np.array([[v, v]]))
    m.J = mk(J)
    m.B = mk(B) if B else None
    m.D = None
    m.si_par = m.si_axis = m.h_par = m.h_axis = None
    return m


def _langevin(x):
    return 1.0 / math.tanh(x) - 1.0 / x


@pytest.mark.parametrize("mode", [0, 1])
def test_single_spin_in_field_exact_distribution(mode):
    """E = h S.z (energy_get's field term): <S.z> = -(coth(beta h) - 1/(beta h))."""
    h, beta, n = 1.0, 1.5, 200000
    o = O.Oracle(_one_site_model(field=h))
    s = np.array([[0.0, 0.0, 1.0]])
    st = O.mt_state(12345)
    xs = np.empty(n)
    for k in range(n // 100):
        for q in range(100):
            o.sweep_ref(s, beta, mode, 0.5 if mode else 0.125, st, 1, 0.0)
            xs[k * 100 + q] = s[0, 2]
    want = -_langevin(beta * h)
    blocks = xs.reshape(100, -1).mean(axis=1)
    err = blocks.std(ddof=1) / math.sqrt(len(blocks))
    assert abs(xs.mean() - want) < 4 * err + 1e-3, (xs.mean(), want, err)


def test_single_ion_exact_distribution():
    """E = Delta (S.z)^2: <(S.z)^2> = int x^2 e^{-b D x^2} / int e^{-b D x^2} on [-1,1]."""
    Dl, beta, n = -1.2, 2.0, 150000
    o = O.Oracle(_one_site_model(Ksi=Dl))
    s = np.array([[1.0, 0.0, 0.0]])
    st = O.mt_state(777)
    xs = np.empty(n)
    for k in range(n):
        o.sweep_ref(s, beta, 0, 0.125, st, 1, 0.0)
        xs[k] = s[0, 2] ** 2
    x = np.linspace(-1, 1, 200001)
    w = np.exp(-beta * Dl * x * x)
    want = np.trapezoid(x * x * w, x) / np.trapezoid(w, x)
    blocks = xs.reshape(100, -1).mean(axis=1)
    err = blocks.std(ddof=1) / math.sqrt(len(blocks))
    assert abs(xs.mean() - want) < 4 * err + 1e-3


def test_heisenberg_dimer_exact_distribution():
    """Two spins, E = -J S1.S2: <S1.S2> = L(beta J)."""
    J, beta, n = 1.0, 1.3, 150000
    o = O.Oracle(_dimer_model(J=J))
    s = np.array([[0.0, 0.0, 1.0], [0.0, 0.0, 1.0]])
    st = O.mt_state(99)
    xs = np.empty(n)
    e = o.energy(s)
    for k in range(n):
        e, _ = o.sweep_ref(s, beta, 0, 0.125, st, 1, e)
        xs[k] = s[0] @ s[1]
    assert e == pytest.approx(o.energy(s), abs=1e-9)   # running energy tracks the direct one
    blocks = xs.reshape(100, -1).mean(axis=1)
    err = blocks.std(ddof=1) / math.sqrt(len(blocks))
    assert abs(xs.mean() - _langevin(beta * J)) < 4 * err + 1e-3


def test_colour_order_philox_sweep_samples_the_same_law():
    """The colour-order / Philox update (the GPU's order) on a dimer: <S1.S2> = L(beta J)."""
    J, beta, n = -0.8, 2.0, 100000
    o = O.Oracle(_dimer_model(J=J))
    s = np.array([[0.0, 0.0, 1.0], [1.0, 0.0, 0.0]])
    xs = np.empty(n)
    for k in range(n):
        o.sweep_colour_philox(s, [0, 1, 2], [0, 1], beta, 0, 0.125, 42, k, 0, 0.0)
        xs[k] = s[0] @ s[1]
    blocks = xs.reshape(100, -1).mean(axis=1)
    err = blocks.std(ddof=1) / math.sqrt(len(blocks))
    assert abs(xs.mean() - _langevin(beta * J)) < 4 * err + 1e-3


def test_const_proposal_is_a_cone_around_the_old_spin():
    """'const' proposals: the angle to the old spin is tilt*acos(2u-1), so cos(angle) follows
    from u exactly, and the proposal has the old spin's length."""
    rng = np.random.default_rng(1)
    for _ in range(200):
        old = rng.normal(size=3)
        old *= 1.7 / np.linalg.norm(old)
        u1, u2 = rng.random(), rng.random()
        tilt = 0.3
        p = O.propose(1, tilt, 1.7, old, u1, u2)
        assert np.linalg.norm(p) == pytest.approx(1.7, rel=1e-12)
        c = p @ old / (1.7 * 1.7)
        assert c == pytest.approx(math.cos(tilt * math.acos(2 * u1 - 1)), abs=1e-12)


def test_pt_swap_acceptance_frequency():
    """pt_swap accepts with probability min(1, exp((b_i - b_j)(E_i - E_j)))."""
    tems = np.array([1.0, 1.25])
    energy = np.array([-10.0, -9.0])   # f = (1 - 0.8)(-1) = -0.2
    n = 40000
    st = O.mt_state(5)
    acc = 0
    for k in range(n):
        sr = np.array([0, 1], np.int32)
        ns = np.zeros((2, 2), np.int64)
        O.pt_swap(1, tems, energy, sr, ns, mt=st)
        acc += int(sr[0] == 1)
        assert ns[0, 0] == 1 and ns[1, 0] == 1
    p = math.exp(-0.2)
    assert abs(acc / n - p) < 4 * math.sqrt(p * (1 - p) / n)
    # favourable pair always swaps, without drawing
    sr = np.array([0, 1], np.int32)
    ns = np.zeros((2, 2), np.int64)
    O.pt_swap(1, tems, np.array([-9.0, -10.0]), sr, ns, mt=None, seed=1, step=0)
    assert sr.tolist() == [1, 0] and ns[:, 1].tolist() == [1, 1]


def test_binder_cumulant_and_estimators_known_values():
    """Two-point series E = +-a gives U_E = 2/3 and C = b^2 a^2 / N; a Gaussian series gives U ~ 0."""
    a, N, T = 3.0, 10, 2.0
    acc = np.zeros(6)
    for k in range(1000):
        e = a if k % 2 else -a
        O.lib.orc_accumulate(e, 0.25, O._p(acc, O._d))
    out = O.finalize(acc, 1000, T, N)
    assert out[4] == pytest.approx(2.0 / 3.0, rel=1e-14)
    assert out[2] == pytest.approx(a * a / T ** 2 / N, rel=1e-12)
    assert out[0] == pytest.approx(0.5, rel=1e-14) and out[3] == pytest.approx(0.0, abs=1e-12)
    rng = np.random.default_rng(0)
    acc = np.zeros(6)
    for e in rng.normal(0.0, 2.0, 200000):
        O.lib.orc_accumulate(float(e), 0.5, O._p(acc, O._d))
    out = O.finalize(acc, 200000, T, N)
    assert abs(out[4]) < 0.02
    assert out[2] == pytest.approx(4.0 / T ** 2 / N, rel=0.02)


# ---------------------------------------------------------------- magnetisation --------------
@pytest.mark.parametrize("mag", [1.0, 2.5])
def test_magnetization2_closed_forms(mag):
    """magnetization_get (get_quant.F90:340-363) returns |sum_j S_j / N|^2: ferro gives m^2, a
    two-sublattice Neel state 0, and a cone of half-angle alpha with evenly spaced azimuths
    (m cos alpha)^2.  A dropped 1/N, a sum of squares or a missing component fails one."""
    L = 6
    m = Lt.cubic_model("sc", L, J=(1.0,), builder=O.build_table_fn)
    m.mag = np.full(m.N, mag)
    o = O.Oracle(m)
    s = Lt.ferro_spins(m.N, 1, m.mag)[0]
    assert o.magnetization2(s) == pytest.approx(mag * mag, rel=1e-14)
    idx = np.arange(m.N)
    x, y, z = idx % L, (idx // L) % L, idx // (L * L)
    s[:, 2] = np.where((x + y + z) % 2 == 0, mag, -mag)
    assert o.magnetization2(s) == pytest.approx(0.0, abs=1e-28)
    alpha = 0.7
    phi = 2 * math.pi * idx / m.N
    cone = mag * np.stack([math.sin(alpha) * np.cos(phi), math.sin(alpha) * np.sin(phi),
                           np.full(m.N, math.cos(alpha))], 1)
    assert o.magnetization2(cone) == pytest.approx((mag * math.cos(alpha)) ** 2, rel=1e-12)
    # a uniform tilt of the whole ferro state keeps |m|^2
    tilt = np.tile(mag * np.array([0.6, 0.0, 0.8]), (m.N, 1))
    assert o.magnetization2(tilt) == pytest.approx(mag * mag, rel=1e-14)


# ---------------------------------------------------------------- colouring check -------------
def _checkerboard(L, N):
    idx = np.arange(N)
    return ((idx % L + (idx // L) % L + idx // (L * L)) % 2).astype(np.int32)


@pytest.mark.parametrize("term", ["J", "B", "D"])
def test_colouring_check_rejects_improper_colourings(term):
    """orc_check_colouring must accept a proper colouring and reject one that gives two bonded
    sites the same colour, for bonds that come from the J, the B or the D table alone.  The
    checkerboard of SC is proper for nearest neighbours and improper for the second shell."""
    L = 4
    base = Lt.cubic_model("sc", L, J=(1.0,), builder=O.build_table_fn)
    col = _checkerboard(L, base.N)
    assert O.Oracle(base).check_colouring(col)
    # one site moved to its neighbour's colour
    bad = col.copy()
    bad[0] = col[int(base.J.nbr[0, 0]) - 1]
    assert not O.Oracle(base).check_colouring(bad)
    # second-shell bonds present only in one table: the checkerboard is no longer proper
    nb, bf, bm = base.nb, base.basis_frac, np.ones(1)
    m = Lt.cubic_model("sc", L, J=(1.0,), builder=O.build_table_fn)
    if term == "J":
        m = Lt.cubic_model("sc", L, J=(1.0, 0.3), builder=O.build_table_fn)
    elif term == "B":
        m.B = Lt.build_table(nb, bf, bm, m.L, Lt.shell_rows("sc", [0.0, 0.2]), 1, builder=O.build_table_fn)
    else:
        m.D = Lt.build_table(nb, bf, bm, m.L, Lt.shell_rows("sc", [], dm=0.3, dm_shell=2), 3,
                             builder=O.build_table_fn)
    assert not O.Oracle(m).check_colouring(col)
    # a colouring of the sublattices modulo 2 in every axis (8 colours) separates shell 2
    idx = np.arange(m.N)
    col8 = ((idx % L) % 2 + 2 * (((idx // L) % L) % 2) + 4 * ((idx // (L * L)) % 2)).astype(np.int32)
    assert O.Oracle(m).check_colouring(col8)


def test_colouring_check_ignores_self_bonds():
    """A supercell one cell thick along an axis bonds a site to itself (the bond wraps onto
    the same site); such bonds involve only the updated site and do not constrain colours."""
    rows = Lt.shell_rows("sc", [1.0])
    t = Lt.build_table(1, np.zeros((1, 3)), np.ones(1), (1, 1, 2), rows, 1, builder=O.build_table_fn)
    m = Lt.Model("sc", (1, 1, 2), 1, np.zeros((1, 3)), np.ones(2), t)
    assert O.Oracle(m).check_colouring(np.array([0, 1], np.int32))
    assert not O.Oracle(m).check_colouring(np.array([0, 0], np.int32))


def test_exactly_summed_energy_pins():
    """orc_energy_sites + fsum: the energy_get terms summed without accumulated rounding.  With
    couplings that are exact binary fractions the ferro energy is exact in closed form and both
    sums hit it; with 0.1-type couplings (N = 131072) the exactly summed energy stays within a
    few ulp of the closed form while energy_get's long sequential sum drifts by its rounding."""
    for J, B, Ksi in (((-1.0, 0.125), 0.125, -0.0625), ((-1.0, 0.1), 0.1, -0.05)):
        m = Lt.cubic_model("fcc", 32, J=J, B=B, K=Ksi, builder=O.build_table_fn)
        o = O.Oracle(m)
        s = Lt.ferro_spins(m.N, 1)[0]
        want = _closed_ferro("fcc", 32, J, B, Ksi)
        if J[1] == 0.125:
            assert o.energy_exact(s) == want and o.energy(s) == want
        else:
            assert abs(o.energy_exact(s) - want) <= 1e-14 * abs(want)
            assert abs(o.energy(s) - want) < 64 * m.N * 2.3e-16 * abs(want)   # ~31 N terms added in turn
        r = Lt.random_spins(m.N, 1, 3)[0]
        assert o.energy_exact(r) == pytest.approx(o.energy(r), rel=64 * m.N * 2.3e-16, abs=1e-9)
    # a small case with every term: the grouping by site sums the same terms as energy_get
    ms = Lt.cubic_model("sc", 3, J=(1.0, 0.3), B=0.05, D=0.3, K=0.02, field=0.5, builder=O.build_table_fn)
    rs = Lt.random_spins(ms.N, 1, 4)[0]
    assert O.Oracle(ms).energy_exact(rs) == pytest.approx(O.Oracle(ms).energy(rs), rel=1e-14)
