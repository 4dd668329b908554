"""The optional observables of the HIP engine (order parameter, staggered moment, P(m), spin
correlations per shell, structure factor, their per-measurement trace) against the oracle
(oracle/obs_oracle.py, pinned by closed forms in test_obs_oracle.py) on the configurations
the engine measured: one measured MC step per run call, the spins read back after it, the
oracle's values accumulated by slot exactly as the reference accumulates them per measurement.
Bars: bin counts exact; sums 1e-12 relative (order of the floating-point sums differs)."""
import math

import numpy as np
import pytest

from espins_amd import lattice as Lt
from oracle import obs_oracle as B

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _setup(exchange_every=False):
    from espins_amd.engine import Engine
    L = 4
    m = Lt.cubic_model("bcc", L, J=(1.0, 0.3))
    N, nb = m.N, m.nb
    tems = Lt.ladder("log", 0.4, 3.0, 6)
    eng = Engine(0)
    eng.set_model(m)
    eng.set_replicas(tems, Lt.random_spins(N, len(tems), 5, m.mag))
    axes = np.array([(0.0, 0.0, 1.0), (0.6, 0.0, -0.8)])
    coeff = np.array([1.0, -1.0])
    eng.set_orderparam(axes)
    eng.set_staggered(coeff)
    nfreq = eng.set_pm(True)
    cnt, nbr = Lt.shell_tables("bcc", L, 2)
    kinds = np.arange(N) % 2
    eng.set_spincorr(kinds, cnt, nbr)
    pos = Lt.site_positions("bcc", L)
    tp = 2 * math.pi
    rng = np.random.default_rng(3)
    q = np.concatenate([[(0.0, 0.0, 0.0), (tp, 0, 0), (tp / 2, tp / 2, tp / 2)],
                        rng.uniform(-2 * tp, 2 * tp, size=(37, 3))])
    polar = np.array([0.3, -0.2, 0.9])
    eng.set_sfactor(q, pos, polar, every=2)
    return eng, m, dict(axes=axes, coeff=coeff, nfreq=nfreq, cnt=cnt, nbr=nbr, kinds=kinds, pos=pos,
                        q=q, polar=polar)


def _oracle_measure(s, m, c):
    op = B.orderparam(s, m.nb, c["axes"])
    sm = B.staggered(s, m.nb, c["coeff"])
    h = B.pm_counts(s, m.nb, m.mag[:m.nb], axes=c["axes"], coeff=c["coeff"])
    t, ta = B.spincorr(s, c["kinds"], c["cnt"], c["nbr"], m.mag, 2)
    return op, sm, h, t, ta


@pytest.mark.parametrize("exchange", [False, True])
def test_observables_match_oracle(exchange):
    eng, m, c = _setup()
    R = eng.R
    steps = 6
    eng.set_trace(steps)
    acc = np.zeros((R, 6))
    pm = np.zeros((R, 3, c["nfreq"]))
    sc = np.zeros((R, 3, 2, 2, 2))
    sf = np.zeros((R, 3, len(c["q"])))
    nsf = 0
    tr = np.zeros((steps, R, 2))
    for k in range(steps):
        say(f"step {k} exchange={exchange}")
        slot_of = np.argsort(eng.pt_map()[0])    # replica -> slot before the round
        eng.run(1, steps_per_round=1, steps_measure=1, measure=True, exchange=exchange,
                start_parity=1 + k % 2, seed=11, sweep0=k, step0=k)
        s_all = eng.spins(by_slot=False)         # by replica (exchanges move labels only)
        for r in range(R):
            sl = slot_of[r]
            op, sm, h, t, ta = _oracle_measure(s_all[r], m, c)
            acc[sl] += [abs(op), op * op, op ** 4, sm, sm * sm, sm ** 4]
            pm[sl] += h
            sc[sl, 0] += t
            sc[sl, 1] += np.abs(t)
            sc[sl, 2] += ta
            tr[k, sl] = [abs(op), sm]
            if (k + 1) % 2 == 0:
                sf[sl] += B.sfactor(s_all[r], c["pos"], c["q"], c["polar"])
        if (k + 1) % 2 == 0:
            nsf += 1
    got_op, n_op = eng.obs(1, (R, 3))
    got_sm, _ = eng.obs(2, (R, 3))
    got_pm, _ = eng.obs(4, (R, 3, c["nfreq"]))
    got_sc, _ = eng.obs(8, (R, 3, 2, 2, 2))
    got_sf, n_sf = eng.obs(16, (R, 3, len(c["q"])))
    got_tr = eng.obs_trace()
    say("op", got_op[:, 0], acc[:, 0])
    assert n_op == steps and n_sf == nsf
    assert np.allclose(got_op, acc[:, 0:3], rtol=1e-12, atol=0)
    assert np.allclose(got_sm, acc[:, 3:6], rtol=1e-12, atol=0)
    assert np.array_equal(got_pm, pm)
    assert np.allclose(got_sc, sc, rtol=1e-12, atol=1e-12 * m.N)
    scale = float(m.N) ** 2
    assert np.allclose(got_sf, sf, rtol=1e-10, atol=1e-12 * scale)
    assert got_tr.shape == (steps, R, 2)
    assert np.allclose(got_tr, tr, rtol=1e-12, atol=0)
    # the running energy is unaffected by the measurement kernels
    st = eng.stats()
    assert np.allclose(st["e_run"], eng.energy_get(), rtol=1e-10, atol=1e-10)
    eng.reset_stats()
    z, n = eng.obs(1, (R, 3))
    assert n == 0 and (z == 0).all()
    eng.close()


def test_observables_off_keeps_lds_path_and_errors_when_off():
    from espins_amd.engine import Engine
    from espins_amd._lib import EspError
    m = Lt.cubic_model("sc", 4, J=(1.0,))
    eng = Engine(0)
    eng.set_model(m)
    eng.set_replicas([1.0, 2.0], Lt.random_spins(m.N, 2, 1, m.mag))
    assert eng.model_info()["lds"]
    with pytest.raises(EspError):
        eng.obs(1, (2, 3))
    eng.set_pm(True)
    eng.run(1, steps_per_round=3, steps_measure=1, measure=True, exchange=False, seed=1)
    h, n = eng.obs(4, (2, 1, 201))
    assert n == 3 and h.sum() == 3 * 2 * m.N        # one count per cell and measurement
    eng.set_pm(False)
    with pytest.raises(EspError):
        eng.obs(4, (2, 1, 201))
    eng.close()


def _drop_bond(cnt, nbr, s, i, l):
    """Remove slot l of site i (0-based) from shell s of the 1-based tables."""
    n = cnt[s, i]
    nbr[s, l:n - 1, i] = nbr[s, l + 1:n, i].copy()
    nbr[s, n - 1, i] = 0
    cnt[s, i] = n - 1


@pytest.mark.parametrize("nspecies,tables", [(1, "full"), (1, "pairs"), (1, "ragged"), (1, "shuffled"),
                                             (2, "pairs"), (3, "ragged"), (3, "shuffled"), (4, "full"),
                                             (4, "ragged")])
def test_spincorr_kinds_moments_tables(nspecies, tables):
    """Spin correlations with 1-4 kinds (kind = basis atom mod nspecies), unequal moments and
    three shells of different widths (12, 6, 24 on FCC 3^3, where a site lists some partners
    more than once), on three kinds of tables: 'full' (symmetric: both directions of a bond
    come from one entry; the owner of a bond is the site whose slot for it is lower, half the
    slots of every site), 'shuffled' (symmetric, each site's list in random order: owners
    from closed trails), 'pairs' (symmetric with odd counts: both directions of a fifth of
    the bonds removed) and 'ragged' (not symmetric: single directions removed, one entry per
    directed bond).  The last site of each kind, whose counts give the shell norms, keeps its
    bonds.  Five replicas, so the last replica group is partly empty."""
    from espins_amd.engine import Engine
    L, R = 3, 5
    m = Lt.cubic_model("fcc", L, J=(1.0, 0.2))
    m.mag = np.tile([1.0, 1.5, 0.8, 2.0], L ** 3)
    cnt, nbr = Lt.shell_tables("fcc", L, 3)
    kinds = (np.arange(m.N) % m.nb % nspecies).astype(np.int32)
    rng = np.random.default_rng(nspecies)
    keep = np.zeros(m.N, bool)
    keep[[np.flatnonzero(kinds == x)[-1] for x in range(nspecies)]] = True
    if tables == "ragged":
        cut = (rng.random(cnt.shape) < 0.2) & ~keep[None, :]
        cnt = np.where(cut, rng.integers(0, cnt + 1), cnt).astype(np.int32)
    elif tables == "pairs":
        for s in range(3):
            for i in range(m.N):
                l = 0
                while l < cnt[s, i]:
                    j = nbr[s, l, i] - 1
                    if i < j and not keep[i] and not keep[j] and rng.random() < 0.2:
                        _drop_bond(cnt, nbr, s, i, l)
                        back = int(np.flatnonzero(nbr[s, :cnt[s, j], j] == i + 1)[0])
                        _drop_bond(cnt, nbr, s, j, back)
                    else:
                        l += 1
        assert (cnt % 2 == 1).any()
    elif tables == "shuffled":
        for s in range(3):
            for i in range(m.N):
                nbr[s, :cnt[s, i], i] = rng.permutation(nbr[s, :cnt[s, i], i])
    eng = Engine(0)
    eng.set_model(m)
    eng.set_replicas(Lt.ladder("lin", 0.5, 2.0, R), Lt.random_spins(m.N, R, 9, m.mag))
    eng.set_spincorr(kinds, cnt, nbr)
    ns = nspecies
    sc = np.zeros((R, 3, 3, ns, ns))
    for k in range(3):
        say(f"nspecies {ns} tables {tables} step {k}")
        slot_of = np.argsort(eng.pt_map()[0])
        eng.run(1, steps_per_round=1, steps_measure=1, measure=True, exchange=True,
                start_parity=1 + k % 2, seed=4, sweep0=k, step0=k)
        s_all = eng.spins(by_slot=False)
        for r in range(R):
            t, ta = B.spincorr(s_all[r], kinds, cnt, nbr, m.mag, ns)
            sc[slot_of[r], 0] += t
            sc[slot_of[r], 1] += np.abs(t)
            sc[slot_of[r], 2] += ta
    got, n = eng.obs(8, (R, 3, 3, ns, ns))
    eng.close()
    assert n == 3
    assert np.isfinite(sc).all()
    assert np.allclose(got, sc, rtol=1e-12, atol=1e-12 * m.N)


def test_sfactor_groups_chunks_and_partial_tiles():
    """S(q) with 17 replicas (three groups of eight, the last one partly empty), 130 q-points
    (three waves of 64, the last one partly empty, q = 0 among them) and 256 sites in four
    site chunks, evaluated at every measurement, against the oracle's per-configuration sums."""
    from espins_amd.engine import Engine
    L, R = 4, 17
    m = Lt.cubic_model("fcc", L, J=(1.0,))
    eng = Engine(0)
    eng.set_model(m)
    eng.set_replicas(Lt.ladder("log", 0.3, 3.0, R), Lt.random_spins(m.N, R, 21, m.mag))
    tp = 2 * math.pi
    rng = np.random.default_rng(8)
    q = np.concatenate([[(0.0, 0.0, 0.0), (tp, 0, 0)], rng.uniform(-3 * tp, 3 * tp, size=(128, 3))])
    pos = Lt.site_positions("fcc", L)
    polar = np.array([0.0, 0.6, -0.8])
    eng.set_sfactor(q, pos, polar, every=1)
    sf = np.zeros((R, 3, len(q)))
    for k in range(2):
        say(f"S(q) step {k}")
        slot_of = np.argsort(eng.pt_map()[0])
        eng.run(1, steps_per_round=1, steps_measure=1, measure=True, exchange=True,
                start_parity=1 + k % 2, seed=2, sweep0=k, step0=k)
        s_all = eng.spins(by_slot=False)
        for r in range(R):
            sf[slot_of[r]] += B.sfactor(s_all[r], pos, q, polar)
    got, n = eng.obs(16, (R, 3, len(q)))
    eng.close()
    assert n == 2
    assert np.allclose(got, sf, rtol=1e-10, atol=1e-12 * float(m.N) ** 2)


@pytest.mark.parametrize("positions", ["lattice", "jittered"])
def test_sfactor_sheared_supercell(positions):
    """S(q) on a BCC 6 x 4 x 3 supercell with a sheared cell (rows (1, 0, 0), (1/2, 1, 0),
    (1/4, -1/2, 1): every position exact in binary): 'lattice' positions take the per-q phase
    tables (three site chunks of (x1, x2) rows), 'jittered' ones (1e-3 random offsets) the
    sincos path; both against the oracle's per-configuration sums."""
    from espins_amd.engine import Engine
    Ls, R = (6, 4, 3), 9
    bf = np.array(Lt.BASIS["bcc"], float)
    nb = len(bf)
    N = Ls[0] * Ls[1] * Ls[2] * nb
    m = Lt.Model("bcc", Ls, nb, bf, np.ones(N),
                 Lt.build_table(nb, bf, np.ones(nb), Ls, Lt.shell_rows("bcc", [1.0]), 1))
    x2, x1, x0 = np.meshgrid(range(Ls[2]), range(Ls[1]), range(Ls[0]), indexing="ij")
    cells = np.stack([x0.ravel(), x1.ravel(), x2.ravel()], axis=1).astype(float)
    shear = np.array([[1.0, 0.0, 0.0], [0.5, 1.0, 0.0], [0.25, -0.5, 1.0]])
    pos = ((cells[:, None, :] + bf[None, :, :]).reshape(-1, 3)) @ shear
    rng = np.random.default_rng(5)
    if positions == "jittered":
        pos = pos + rng.uniform(-1e-3, 1e-3, pos.shape)
    tp = 2 * math.pi
    q = rng.uniform(-2 * tp, 2 * tp, size=(70, 3))
    eng = Engine(0)
    eng.set_model(m)
    eng.set_replicas(Lt.ladder("lin", 0.4, 2.5, R), Lt.random_spins(N, R, 13, m.mag))
    eng.set_sfactor(q, pos, None, every=1)
    sf = np.zeros((R, 3, len(q)))
    for k in range(2):
        say(f"S(q) {positions} step {k}")
        slot_of = np.argsort(eng.pt_map()[0])
        eng.run(1, steps_per_round=1, steps_measure=1, measure=True, exchange=True,
                start_parity=1 + k % 2, seed=6, sweep0=k, step0=k)
        s_all = eng.spins(by_slot=False)
        for r in range(R):
            sf[slot_of[r]] += B.sfactor(s_all[r], pos, q, None)
    got, n = eng.obs(16, (R, 3, len(q)))
    eng.close()
    assert n == 2
    assert np.allclose(got, sf, rtol=1e-10, atol=1e-12 * float(N) ** 2)
