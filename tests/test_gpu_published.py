"""Statistical parity with the reference's own published runs of its examples: the same input
files (Hamiltonians with J, biquadratic B, DM D, single-ion Delta; plain Metropolis and parallel
tempering) run on the engine, observables compared with the published _mc.dat tables
(examples/example0N/reference/*_mc.dat, copied as data into tests/golden/examples/).

Independent copies give the error of this run; the published run's error is one copy's spread
scaled to the published run length.  Each column is compared at combined error."""
import math
import os

import numpy as np
import pytest

from espins_amd import lattice as Lt
from espins_amd import mcin
from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu

COLS = ("|m|", "E/N", "C", "chi", "U_E", "U_M")


def say(*a):
    print(*a, flush=True)


def _load(name):
    inp = mcin.read(os.path.join(GOLDEN, "examples", f"{name}.mcin"))
    return inp, mcin.to_model(inp)


def _compare(name, ref, obs, n_mine, n_ref, cols, frac=0.85, zmax=6.0):
    """obs[T][copy][6]: z = |mean - ref| / sqrt(se_mine^2 + se_ref^2) per (T, column).

    With 16 copies the spread has 15 degrees of freedom: |z| > 6 has probability about 2e-5
    per entry (Student t), under 1 % over the 100-260 entries of one table; a dropped term or
    a wrong sign in the Hamiltonian moves whole columns by far more."""
    mean = obs.mean(axis=1)
    sd = obs.std(axis=1, ddof=1)
    se_mine = sd / math.sqrt(obs.shape[1])
    se_ref = sd * math.sqrt(n_mine / n_ref)
    z = np.abs(mean - ref[:, 1:7]) / np.maximum(np.sqrt(se_mine ** 2 + se_ref ** 2), 1e-300)
    z = z[:, cols]
    say(f"{name} |z| per T x {[COLS[c] for c in cols]}:\n", np.round(z, 2))
    for t, c in zip(*np.nonzero(z >= 3.0)):
        say(f"  T={ref[t, 0]:.4f} {COLS[cols[c]]}: mine {mean[t, cols[c]]:.6g} "
            f"+- {se_mine[t, cols[c]]:.3g}, published {ref[t, 1 + cols[c]]:.6g}")
    within = float((z < 2.0).mean())
    say(f"{name}: fraction within 2 sigma {within:.3f}, max {z.max():.2f}")
    assert within >= frac
    assert z.max() < zmax


def _plain(name, copies, warm, nmc):
    """Plain Metropolis (mcarlo.F90:490-633): every temperature x copies as independent chains
    in one context, the reference's steps_measure and proposal mode."""
    from espins_amd.engine import Engine, finalize
    inp, m = _load(name)
    tems = inp.temperatures()
    M = inp.kw("steps_measure", 2, int)
    mode = inp.kw("mcarlo_mode", "random")
    tilt = inp.kw("tilt_angles_max", 0.125, float)
    T_all = np.repeat(tems, copies)
    with Engine() as eng:
        eng.set_model(m)
        eng.set_replicas(T_all, Lt.ferro_spins(m.N, len(T_all), m.mag))
        eng.run(1, steps_per_round=warm, steps_measure=M, measure=False, exchange=False, mode=mode,
                tilt=tilt, seed=31)
        eng.reset_stats()
        eng.run(1, steps_per_round=nmc, steps_measure=M, measure=True, exchange=False, mode=mode,
                tilt=tilt, seed=31, sweep0=warm * M)
        st = eng.stats()
        assert np.allclose(st["e_run"], eng.energy_get(), rtol=1e-9, atol=1e-9)
    obs = finalize(st["acc"], st["nmeas"], T_all, m.N)[:, :6].reshape(len(tems), copies, 6)
    return tems, obs, inp


def _pt(name, copies, warm, nmc, start=None):
    """Parallel tempering (pt.F90:529-908) for `copies` independent ladders, advanced together
    as a batch: warmup1 plain, warmup2 with exchanges, main phase with exchanges and a
    measurement after every MC step.  start(m, tems, copy) gives the initial spins [R][N][3]
    (the reference's ferro start by default)."""
    from espins_amd.engine import Batch, PTSchedule, finalize
    inp, m = _load(name)
    tems = inp.temperatures()
    M = inp.kw("steps_measure", 2, int)
    pss = inp.kw("pt_steps_swap", 1, int)
    mode = inp.kw("mcarlo_mode", "random")
    tilt = inp.kw("tilt_angles_max", 0.125, float)
    sch = PTSchedule.of(warm, nmc, pss)
    kw = dict(steps_measure=M, mode=mode, tilt=tilt, seed=77)
    with Batch() as b:
        for c in range(copies):
            s0 = Lt.ferro_spins(m.N, len(tems), m.mag) if start is None else start(m, tems, c)
            b.add_chain(m, tems, s0)
        b.commit()
        b.run(1, steps_per_round=sch.warmup1, measure=False, exchange=False, graph=False, **kw)
        b.run(sch.n_swap_warmup, steps_per_round=pss, measure=False, exchange=True,
              sweep0=sch.warmup1 * M, **kw)
        b.reset_stats()
        b.run(sch.n_swap, steps_per_round=pss, measure=True, exchange=True,
              sweep0=2 * sch.warmup1 * M, step0=sch.n_swap_warmup, **kw)
        obs = []
        for c in range(copies):
            st = b.stats(c)
            assert st["nmeas"] == sch.steps_mc
            sr, ns = b.pt_map(c)
            assert sorted(sr.tolist()) == list(range(1, len(tems) + 1))
            assert (ns[:, 0] > 0).all()
            obs.append(finalize(st["acc"], st["nmeas"], tems, m.N)[:, :6])
            assert np.allclose(st["e_run"], b.energy_get(c), rtol=1e-9, atol=1e-9)
    return tems, np.transpose(np.array(obs), (1, 0, 2)), inp


def test_example03_pyrochlore_j_b_dm():
    """FeF3 pyrochlore (N = 256): J + biquadratic + Dzyaloshinskii-Moriya, 28 temperatures,
    plain Metropolis, 'random' proposals (examples/example03/reference/fef3_mc.dat)."""
    tems, obs, inp = _plain("ex03_fef3", copies=16, warm=40000, nmc=20000)
    ref = np.loadtxt(os.path.join(GOLDEN, "examples", "ex03_fef3_mc.dat"), comments="#")
    assert np.allclose(ref[:, 0], tems, atol=1e-4)
    _compare("example03", ref, obs, 20000, inp.kw("steps_mc", 0, int), cols=[0, 1, 2, 3, 5])


def test_example05_j_b_single_ion():
    """SnWO-type cell (N = 128): 18 exchange and 8 biquadratic bonds per site plus single-ion
    anisotropy, 48 temperatures, plain Metropolis (examples/example05/reference/snwo_mc.dat)."""
    tems, obs, inp = _plain("ex05_snwo", copies=16, warm=40000, nmc=20000)
    ref = np.loadtxt(os.path.join(GOLDEN, "examples", "ex05_snwo_mc.dat"), comments="#")
    assert np.allclose(ref[:, 0], tems, atol=1e-4)
    _compare("example05", ref, obs, 20000, inp.kw("steps_mc", 0, int), cols=[0, 1, 2, 3, 5])


def test_example02_parallel_tempering_const_proposals():
    """SC 14^3 (N = 2744) antiferromagnet J1 + J2 through its transition, 32 replicas 36-39.875 K,
    exchanges every 10 MC steps of 5 sweeps, 'const' proposals
    (examples/example02/reference/cubic_mc.dat).  With J2/J1 = 0.26 > 1/4 the ordered state is
    C-type (antiparallel in the xy-plane, parallel along z).  The transition is sharp (the
    published C peaks at 38.5 K), so a short run keeps the phase it starts in: replicas below
    38.5 K start C-type and the others from random directions, instead of the reference's ferro
    start, which its 10^6 warm-up steps forget and short runs do not."""
    def start(m, tems, c):
        i = np.arange(m.N)
        L = m.L[0]
        ct = np.zeros((m.N, 3))
        ct[:, 2] = np.where((i % L + (i // L) % L) % 2 == 0, 1.0, -1.0) * np.asarray(m.mag)
        rnd = Lt.random_spins(m.N, len(tems), seed=1000 + c, mag=m.mag)
        return np.where((np.asarray(tems) < 38.5)[:, None, None], ct[None], rnd)
    tems, obs, inp = _pt("ex02_cubic", copies=16, warm=20000, nmc=60000, start=start)
    ref = np.loadtxt(os.path.join(GOLDEN, "examples", "ex02_cubic_mc.dat"), comments="#")
    assert np.allclose(ref[:, 0], tems, atol=1e-4)
    _compare("example02", ref, obs, 60000, inp.kw("steps_mc", 0, int), cols=[0, 1, 2, 3])


def test_example04_parallel_tempering_seven_shells():
    """Layered cell 24 x 24 x 1 (N = 2304), 7 exchange shells (25 bonds per site), 64 replicas
    4-70 K, exchanges every 10 MC steps of 5 sweeps, 'random' proposals
    (examples/example04/reference/pt_with_uniform_temperatures/bmno_mc.dat)."""
    tems, obs, inp = _pt("ex04_bmno_pt", copies=16, warm=10000, nmc=30000)
    ref = np.loadtxt(os.path.join(GOLDEN, "examples", "ex04_bmno_pt_mc.dat"), comments="#")
    assert np.allclose(ref[:, 0], tems, atol=1e-4)
    _compare("example04", ref, obs, 30000, inp.kw("steps_mc", 0, int), cols=[0, 1, 2, 3])
