"""GPU parity: libespins_hip.so (through the C ABI) against the CPU oracle on the same inputs."""
import math
import os
import sys

import numpy as np
import pytest

from oracle import oracle as O
from espins_amd import lattice as Lt
from espins_amd import mcin
from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def En():
    from espins_amd import engine
    return engine


def models():
    """(name, model) pairs covering every term of the Hamiltonian."""
    out = [
        ("sc_j1", Lt.cubic_model("sc", 6, J=(1.0,))),
        ("fcc_j123_si", Lt.cubic_model("fcc", 4, J=(-1.0, 0.2, 0.05), K=-0.1)),
        ("sc_jbd_si", Lt.cubic_model("sc", 6, J=(1.0,), B=0.05, D=0.3, K=0.02)),
        ("fcc_jb_si_h", Lt.cubic_model("fcc", 3, J=(-1.0, 0.1), B=0.1, K=-0.05, field=0.3,
                                       field_axis=(1, 1, 0))),
        ("bcc_j12", Lt.cubic_model("bcc", 4, J=(1.0, -0.25))),
    ]
    ex03 = mcin.to_model(mcin.read(os.path.join(GOLDEN, "examples", "ex03_fef3.mcin")), supercell=(2, 2, 2))
    out.append(("ex03_pyrochlore_jbd", ex03))
    ex04 = mcin.to_model(mcin.read(os.path.join(GOLDEN, "examples", "ex04_bmno.mcin")), supercell=(3, 3, 1))
    out.append(("ex04_selfbonds", ex04))
    return out


MODELS = models()


def colour_order(eng):
    ncol, col = eng.coloring()
    sites = np.argsort(col, kind="stable").astype(np.int32)
    off = np.zeros(ncol + 1, np.int32)
    off[1:] = np.cumsum(np.bincount(col, minlength=ncol))
    return off, sites, col


@pytest.mark.parametrize("name,m", MODELS, ids=[n for n, _ in MODELS])
def test_energy_matches_oracle(En, name, m):
    say("energy", name, m.N)
    R = 4
    spins = Lt.random_spins(m.N, R, 7, m.mag)
    spins[0] = Lt.ferro_spins(m.N, 1, m.mag)[0]
    with En.Engine() as eng:
        eng.set_model(m).set_replicas(np.linspace(0.5, 2.0, R), spins)
        E = eng.energy_get()
    o = O.Oracle(m)
    scale = np.abs(m.J.val).sum() + (np.abs(m.B.val).sum() if m.B else 0) + (np.abs(m.D.val).sum() if m.D else 0) + m.N
    for r in range(R):
        Eo = o.energy(spins[r])
        assert abs(E[r] - Eo) <= 1e-12 * max(abs(Eo), 1.0) + 1e-14 * scale, (r, E[r], Eo)


@pytest.mark.parametrize("path,indexing", [(1, 0), (1, 1), (2, 0)],
                         ids=["global-implicit", "global-tables", "lds"])
@pytest.mark.parametrize("name,m", MODELS, ids=[n for n, _ in MODELS])
def test_hostfed_decisions_bit_exact(En, name, m, path, indexing):
    """Deterministic mode: same proposals and uniforms in the same colour order give the same
    accept/reject decisions and bit-identical spins.  The global-memory phases run with partner
    indices computed from the lattice geometry (where the lattice allows it) and with the
    neighbour tables."""
    say("hostfed", name, path, indexing)
    R, nsw = 3, 4
    tems = np.array([0.3, 1.0, 3.0])
    spins = Lt.random_spins(m.N, R, 11, m.mag)
    rng = np.random.default_rng(21)
    o = O.Oracle(m)
    with En.Engine() as eng:
        eng.set_model(m)
        try:
            eng.set_path(path)
        except Exception:
            pytest.skip("replica does not fit in LDS")
        eng.set_indexing(indexing)
        if path == 1 and indexing == 0 and name in ("sc_j1", "fcc_j123_si", "sc_jbd_si", "bcc_j12"):
            assert eng.model_info()["layout"] == "implicit"
        if indexing == 1:
            assert eng.model_info()["layout"] != "implicit"
        eng.set_replicas(tems, spins)
        off, sites, col = colour_order(eng)
        assert o.check_colouring(col)
        ref = spins.copy()
        e_ref = [o.energy(ref[r]) for r in range(R)]
        rej_ref = np.zeros(R, np.int64)
        for k in range(nsw):
            prop = Lt.random_spins(m.N, R, 100 + k, m.mag)
            unif = rng.random((R, m.N))
            acc = eng.sweep_hostfed(prop, unif)
            for r in range(R):
                a, e_ref[r], nr = o.sweep_colour_hostfed(ref[r], off, sites, 1.0 / tems[r],
                                                         prop[r], unif[r], e_ref[r])
                rej_ref[r] += nr
                assert np.array_equal(a, acc[r]), (k, r, np.nonzero(a != acc[r]))
        got = eng.spins()
        assert np.array_equal(got, ref)
        st = eng.stats()
        assert np.array_equal(st["rejected"], rej_ref)
        for r in range(R):
            assert st["e_run"][r] == pytest.approx(e_ref[r], rel=1e-11, abs=1e-9)
        Ed = eng.energy_get()
        for r in range(R):
            assert st["e_run"][r] == pytest.approx(Ed[r], rel=1e-10, abs=1e-8)


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("path", [1, 2])
@pytest.mark.parametrize("name,m", MODELS[:4], ids=[n for n, _ in MODELS[:4]])
def test_philox_sweep_matches_oracle(En, name, m, mode, path):
    """Full on-device mode (Philox proposals and uniforms) against the oracle's colour-order
    Philox sweep: same decisions, spins equal to libm rounding."""
    say("philox", name, mode, path)
    R, nsw, seed = 3, 3, 0xC0FFEE1234
    tems = np.array([0.4, 1.1, 2.5])
    spins = Lt.random_spins(m.N, R, 3, m.mag)
    o = O.Oracle(m)
    with En.Engine() as eng:
        eng.set_model(m)
        try:
            eng.set_path(path)
        except Exception:
            pytest.skip("replica does not fit in LDS")
        eng.set_replicas(tems, spins)
        off, sites, _ = colour_order(eng)
        eng.sweep(nsw, mode=mode, tilt=0.3, seed=seed, sweep0=17)
        got = eng.spins()
        st = eng.stats()
    for r in range(R):
        s = spins[r].copy()
        e = o.energy(s)
        rej = 0
        for k in range(nsw):
            e, nr = o.sweep_colour_philox(s, off, sites, 1.0 / tems[r], mode, 0.3, seed, 17 + k, r, e)
            rej += nr
        assert st["rejected"][r] == rej
        np.testing.assert_allclose(got[r], s, rtol=0, atol=1e-12)
        assert st["e_run"][r] == pytest.approx(e, rel=1e-10, abs=1e-8)


@pytest.mark.parametrize("name", ["C2 FCC 16^3 J1-J3 Delta", "C4 SC 64^3 J1 B DM Delta", "C5 FCC 32^3 J1-J2 B Delta"])
def test_full_size_philox_sweep_matches_oracle(En, name):
    """The benchmark Hamiltonians at their full sizes: two on-device sweeps ('random', then
    'const') of a cold and a hot replica give the oracle's colour-order Philox sweep --
    the same rejections and spins to 1e-12."""
    if name.startswith("C2"):
        m, _ = Lt.config_c2(L=16, R=2)
    elif name.startswith("C4"):
        m, _ = Lt.config_c4(L=64, R=2)
    else:
        m, _ = Lt.config_c5(L=32, R=2)
    tems = np.array([0.1, 2.0])
    spins = Lt.random_spins(m.N, 2, 23, m.mag)
    seed = 0x5A5A5A5A
    with En.Engine() as eng:
        eng.set_model(m).set_replicas(tems, spins)
        off, sites, _ = colour_order(eng)
        say("full-size sweep", name, m.N, eng.model_info())
        eng.sweep(1, mode=0, seed=seed, sweep0=5)
        eng.sweep(1, mode=1, tilt=0.2, seed=seed, sweep0=6)
        got = eng.spins()
        st = eng.stats()
    o = O.Oracle(m)
    for r in range(2):
        s = spins[r].copy()
        e = o.energy(s)
        e, n0 = o.sweep_colour_philox(s, off, sites, 1.0 / tems[r], 0, 0.2, seed, 5, r, e)
        e, n1 = o.sweep_colour_philox(s, off, sites, 1.0 / tems[r], 1, 0.2, seed, 6, r, e)
        assert st["rejected"][r] == n0 + n1, (r, st["rejected"][r], n0, n1)
        np.testing.assert_allclose(got[r], s, rtol=0, atol=1e-12)
        assert st["e_run"][r] == pytest.approx(e, rel=1e-10, abs=1e-8)


def test_pt_exchange_matches_oracle(En):
    m = Lt.cubic_model("sc", 4, J=(1.0,))
    R = 9
    tems = Lt.ladder("log", 0.3, 3.0, R)
    spins = Lt.random_spins(m.N, R, 5, m.mag)
    seed = 77
    with En.Engine() as eng:
        eng.set_model(m).set_replicas(tems, spins)
        eng.sweep(2, seed=1)
        e_run = eng.stats()["e_run"]
        sr = np.arange(R, dtype=np.int32)
        ns = np.zeros((R, 2), np.int64)
        for step in range(12):
            start = 1 + step % 2
            eng.pt_swap(start, seed, step)
            O.pt_swap(start, tems, e_run, sr, ns, mt=None, seed=seed, step=step)
            got_sr, got_ns = eng.pt_map()
            assert np.array_equal(got_sr - 1, sr), step
            assert np.array_equal(got_ns, ns), step
        # labels moved, configurations did not
        assert np.array_equal(eng.spins(by_slot=False), eng.spins(by_slot=False))
        by_slot = eng.spins(by_slot=True)
        by_rep = eng.spins(by_slot=False)
        for t in range(R):
            assert np.array_equal(by_slot[t], by_rep[sr[t]])


def test_measure_accumulates_estimators(En):
    m = Lt.cubic_model("fcc", 3, J=(-1.0, 0.2), K=-0.1)
    R = 5
    tems = Lt.ladder("lin", 0.5, 2.0, R)
    with En.Engine() as eng:
        eng.set_model(m).set_replicas(tems, Lt.random_spins(m.N, R, 9, m.mag))
        acc = np.zeros((R, 6))
        for k in range(6):
            eng.sweep(1, seed=3, sweep0=k)
            eng.measure()
            s = eng.spins()
            e = eng.stats()["e_run"]
            for r in range(R):
                mv = s[r].mean(axis=0)
                m2 = mv @ mv
                acc[r] += [e[r], e[r] ** 2, e[r] ** 4, math.sqrt(m2), m2, m2 * m2]
        st = eng.stats()
        assert st["nmeas"] == 6
        np.testing.assert_allclose(st["acc"], acc, rtol=1e-12)


@pytest.mark.parametrize("path", [1, 2])
def test_run_graph_is_deterministic_and_equals_stepwise(En, path):
    m, tems = Lt.config_c2(L=4, R=8)
    spins = Lt.random_spins(m.N, 8, 1, m.mag)
    outs = []
    for graph in (False, True, True):
        with En.Engine() as eng:
            eng.set_model(m)
            eng.set_path(path)
            eng.set_replicas(tems, spins)
            eng.run(5, steps_per_round=2, steps_measure=2, measure=True, exchange=True,
                    seed=9, graph=graph)
            st = eng.stats()
            outs.append((eng.spins(), st["acc"], st["rejected"], eng.pt_map()))
    for o in outs[1:]:
        assert np.array_equal(o[0], outs[0][0])
        assert np.array_equal(o[1], outs[0][1])
        assert np.array_equal(o[2], outs[0][2])
        assert np.array_equal(o[3][0], outs[0][3][0]) and np.array_equal(o[3][1], outs[0][3][1])
    # stepwise with the public entry points reproduces the fused run
    with En.Engine() as eng:
        eng.set_model(m)
        eng.set_path(path)
        eng.set_replicas(tems, spins)
        sweep = 0
        for rnd in range(5):
            for s in range(2):
                eng.sweep(2, seed=9, sweep0=sweep)
                sweep += 2
                eng.measure()
            eng.pt_swap(1 + rnd % 2, 9, rnd)
        st = eng.stats()
        assert np.array_equal(eng.spins(), outs[0][0])
        np.testing.assert_allclose(st["acc"], outs[0][1], rtol=1e-13)
        assert np.array_equal(eng.pt_map()[0], outs[0][3][0])


@pytest.mark.parametrize("steps,M", [(2, 1), (3, 2)])
def test_lds_mid_launch_measurement_matches_global_path(En, steps, M):
    """Measurements in the middle of an LDS-resident launch leave the trajectory unchanged:
    the fused LDS run equals the global-memory path and the stepwise run bit for bit."""
    m, tems = Lt.config_c2(L=4, R=8)
    spins = Lt.random_spins(m.N, 8, 1, m.mag)
    outs = []
    for path in (2, 1):
        with En.Engine() as eng:
            eng.set_model(m)
            eng.set_path(path)
            eng.set_replicas(tems, spins)
            eng.run(1, steps_per_round=steps, steps_measure=M, measure=True, exchange=False,
                    seed=9, graph=False)
            st = eng.stats()
            outs.append((eng.spins(), st["rejected"], st["acc"]))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])
    np.testing.assert_allclose(outs[0][2], outs[1][2], rtol=1e-13)


@pytest.mark.parametrize("name", ["C2_fcc16_j123_si", "C4_sc64_jbd_si", "C5_fcc32_jb_si"])
def test_full_size_energies_match_oracle(En, name):
    """The benchmark Hamiltonians at their full sizes: direct energies of random and ferro
    configurations equal the oracle's energy_get restatement to 1e-12 relative."""
    from oracle import oracle as Oc
    if name.startswith("C2"):
        m, _ = Lt.config_c2(L=16, R=2)
    elif name.startswith("C4"):
        m, _ = Lt.config_c4(L=64, R=2)
    else:
        m, _ = Lt.config_c5(L=32, R=2)
    say("full-size energy", name, m.N)
    spins = Lt.random_spins(m.N, 3, 17, m.mag)
    spins[2] = Lt.ferro_spins(m.N, 1, m.mag)[0]
    with En.Engine() as eng:
        eng.set_model(m).set_replicas(np.array([0.5, 1.0, 2.0]), spins)
        E = eng.energy_get()
    o = Oc.Oracle(m)
    for r in range(3):
        # against the exactly summed energy_get terms: 1e-12 relative
        Ex = o.energy_exact(spins[r])
        assert abs(E[r] - Ex) <= 1e-12 * abs(Ex), (r, E[r], Ex)
        # energy_get's own sequential sum differs from it by at most its rounding (N eps)
        Eo = o.energy(spins[r])
        assert abs(E[r] - Eo) <= 64 * m.N * 2.3e-16 * abs(Eo), (r, E[r], Eo)


def test_batched_chain_matches_oracle_philox_sweep(En):
    """A chain of a batch advances exactly like the oracle's colour-order Philox sweep with the
    chain's stream base k << 16 (replica r draws stream (k << 16) + r)."""
    from espins_amd.engine import plan_model
    chains = [(Lt.cubic_model("fcc", 4, J=(1.0, 0.15)), np.array([0.6, 1.2, 2.0])),
              (Lt.cubic_model("sc", 8, J=(1.0,), B=0.05, D=0.3, K=0.02), np.array([0.4, 0.9])),
              (Lt.cubic_model("bcc", 6, J=(1.0, -0.2)), np.array([0.7, 1.4, 2.1, 3.0]))]
    init = [Lt.random_spins(m.N, len(t), 40 + k, m.mag) for k, (m, t) in enumerate(chains)]
    seed, nsw = 0xABCDEF12345, 3
    with En.Batch() as b:
        for (m, t), s in zip(chains, init):
            b.add_chain(m, t, s)
        for k in range(len(chains)):
            b.set_chain_config(k, 1, 1)          # global-memory colour phases
        b.commit()
        b.run(nsw, measure=False, exchange=False, seed=seed, graph=True)
        got = [b.spins(k) for k in range(len(chains))]
        rej = [b.stats(k)["rejected"] for k in range(len(chains))]
    for k, ((m, t), s0) in enumerate(zip(chains, init)):
        ncol, col, _ = plan_model(m)
        sites = np.argsort(col, kind="stable").astype(np.int32)
        off = np.concatenate([[0], np.cumsum(np.bincount(col, minlength=ncol))]).astype(np.int32)
        o = O.Oracle(m)
        for r in range(len(t)):
            s = s0[r].copy()
            e = o.energy(s)
            nr = 0
            for sw in range(nsw):
                e, x = o.sweep_colour_philox(s, off, sites, 1.0 / t[r], 0, 0.125, seed, sw,
                                             (k << 16) + r, e)
                nr += x
            assert rej[k][r] == nr, (k, r)
            np.testing.assert_allclose(got[k][r], s, rtol=0, atol=1e-12)
