"""Statistical parity on the GPU: exact distributions, the reference's published example01
observables, and parallel tempering against independent chains."""
import math
import os
import sys

import numpy as np
import pytest

from espins_amd import lattice as Lt
from espins_amd import mcin
from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def langevin(x):
    return 1.0 / math.tanh(x) - 1.0 / x


class _M:
    pass


def one_site(field=0.0, K=0.0):
    m = _M()
    m.kind, m.L, m.nb = "one", (1, 1, 1), 1
    m.mag = np.ones(1)
    m.J = Lt.Table(0, np.zeros(1, np.int32), np.zeros((0, 1), np.int32), np.zeros((0, 1)))
    m.B = m.D = None
    m.si_par = np.array([K]) if K else None
    m.si_axis = np.array([[0.0, 0.0, 1.0]]) if K else None
    m.h_par = np.array([field]) if field else None
    m.h_axis = np.array([[0.0, 0.0, 1.0]]) if field else None
    m.N = 1
    return m


def dimer(J):
    m = _M()
    m.kind, m.L, m.nb = "dimer", None, 1
    m.mag = np.ones(2)
    m.J = Lt.Table(1, np.ones(2, np.int32), np.array([[2, 1]], np.int32), np.array([[J, J]]))
    m.B = m.D = None
    m.si_par = m.si_axis = m.h_par = m.h_axis = None
    m.N = 2
    return m


def replica_means(eng, R, nmeas):
    st = eng.stats()
    assert st["nmeas"] == nmeas
    return st["acc"] / nmeas      # per slot: <E>, <E^2>, <E^4>, <|m|>, <m^2>, <m^4>


@pytest.mark.parametrize("mode", ["random", "const"])
def test_single_spin_in_field(mode):
    from espins_amd.engine import Engine
    h, T, R, n = 1.0, 0.7, 256, 4000
    with Engine() as eng:
        eng.set_model(one_site(field=h))
        eng.set_replicas(np.full(R, T), Lt.random_spins(1, R, 1))
        eng.run(1, steps_per_round=500, measure=False, exchange=False, mode=mode, tilt=0.5, seed=3)
        eng.reset_stats()
        eng.run(1, steps_per_round=n, measure=True, exchange=False, mode=mode, tilt=0.5, seed=3,
                sweep0=500)
        e = replica_means(eng, R, n)[:, 0]
    want = -h * langevin(h / T)                 # E = h S.z
    err = e.std(ddof=1) / math.sqrt(R)
    say("field", mode, e.mean(), want, err)
    assert abs(e.mean() - want) < 4 * err


def test_single_ion_and_dimer():
    from espins_amd.engine import Engine
    Dl, T, R, n = -1.0, 0.5, 256, 4000
    with Engine() as eng:
        eng.set_model(one_site(K=Dl))
        eng.set_replicas(np.full(R, T), Lt.random_spins(1, R, 2))
        eng.run(1, steps_per_round=300, measure=False, exchange=False, seed=5)
        eng.reset_stats()
        eng.run(1, steps_per_round=n, measure=True, exchange=False, seed=5, sweep0=300)
        e = replica_means(eng, R, n)[:, 0]      # E = Delta (S.z)^2
    x = np.linspace(-1, 1, 200001)
    w = np.exp(-Dl * x * x / T)
    want = Dl * np.trapezoid(x * x * w, x) / np.trapezoid(w, x)
    err = e.std(ddof=1) / math.sqrt(R)
    assert abs(e.mean() - want) < 4 * err, (e.mean(), want, err)

    J, T = 1.0, 0.8
    with Engine() as eng:
        eng.set_model(dimer(J))
        eng.set_replicas(np.full(R, T), Lt.random_spins(2, R, 7))
        eng.run(1, steps_per_round=300, measure=False, exchange=False, seed=9)
        eng.reset_stats()
        eng.run(1, steps_per_round=n, measure=True, exchange=False, seed=9, sweep0=300)
        e = replica_means(eng, R, n)[:, 0]      # E = -J S1.S2
    want = -J * langevin(J / T)
    err = e.std(ddof=1) / math.sqrt(R)
    assert abs(e.mean() - want) < 4 * err, (e.mean(), want, err)


def test_example01_observables_within_errors():
    """SC 4x4x4 Heisenberg ferromagnet, J = 3 meV, 20 temperatures 25-75 K, 'random'
    proposals, 2 sweeps per MC step (examples/example01/reference/cubic.mcin); compare <|m|>,
    <E>/N, C, chi, U_E, U_M with the reference's published cubic_mc.dat."""
    from espins_amd.engine import Engine, finalize
    inp = mcin.read(os.path.join(GOLDEN, "examples", "ex01_cubic.mcin"))
    m = mcin.to_model(inp)
    tems = inp.temperatures()
    ref = np.loadtxt(os.path.join(GOLDEN, "examples", "ex01_cubic_mc.dat"), comments="#")
    assert np.allclose(ref[:, 0], tems, atol=1e-4)
    copies, warm, n = 16, 5000, 25000
    R = len(tems) * copies
    T_all = np.repeat(tems, copies)
    with Engine() as eng:
        eng.set_model(m)
        eng.set_replicas(T_all, Lt.ferro_spins(m.N, R, m.mag))
        eng.run(1, steps_per_round=warm, steps_measure=2, measure=False, exchange=False, seed=11)
        eng.reset_stats()
        eng.run(1, steps_per_round=n, steps_measure=2, measure=True, exchange=False, seed=11,
                sweep0=2 * warm)
        st = eng.stats()
    obs = finalize(st["acc"], st["nmeas"], T_all, m.N)[:, :6]      # per replica
    obs = obs.reshape(len(tems), copies, 6)
    mean = obs.mean(axis=1)
    sd = obs.std(axis=1, ddof=1)
    se_mine = sd / math.sqrt(copies)
    # the reference ran 2e5 MC steps: its error is one copy's spread scaled to that length
    se_ref = sd * math.sqrt(n / 2.0e5)
    z = np.abs(mean - ref[:, 1:7]) / np.sqrt(se_mine ** 2 + se_ref ** 2)
    say("example01 |z| per T x (|m|, E/N, C, chi, U_E, U_M):\n", np.round(z, 2))
    within2 = float((z < 2.0).mean())
    say("fraction within 2 sigma:", within2)
    assert within2 >= 0.85
    assert z.max() < 5.0


def test_parallel_tempering_matches_independent_chains():
    """Canonical averages at each ladder temperature are the same with exchanges (PT) and
    without (independent chains): checks the label-swap exchange keeps detailed balance."""
    from espins_amd.engine import Engine, finalize
    m = Lt.cubic_model("fcc", 3, J=(-1.0, 0.2, 0.05), K=-0.1)
    tems = Lt.ladder("log", 0.6, 3.0, 8)
    copies, warm, n = 12, 2000, 8000
    res = {}
    for pt in (True, False):
        vals = []
        for c in range(copies):
            with Engine() as eng:
                eng.set_model(m)
                eng.set_replicas(tems, Lt.random_spins(m.N, len(tems), 100 + c))
                eng.run(warm, measure=False, exchange=pt, seed=1000 + c)
                eng.reset_stats()
                eng.run(n, measure=True, exchange=pt, seed=1000 + c, sweep0=warm, step0=warm)
                st = eng.stats()
                vals.append(finalize(st["acc"], st["nmeas"], tems, m.N)[:, [1, 2, 4]])
                if pt:
                    sr, ns = eng.pt_map()
                    assert (ns[:, 0] > 0).all() and (ns[:, 1] > 0).all()
        vals = np.array(vals)
        res[pt] = (vals.mean(axis=0), vals.std(axis=0, ddof=1) / math.sqrt(copies))
    z = np.abs(res[True][0] - res[False][0]) / np.sqrt(res[True][1] ** 2 + res[False][1] ** 2)
    say("PT vs independent |z| (E/N, C, |m|^2):\n", np.round(z, 2))
    assert float((z < 2.0).mean()) >= 0.8
    assert z.max() < 4.5
