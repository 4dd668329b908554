"""Batched chains (esp_batch_*): every chain of a batch reproduces a single context run with the
same Philox stream base, kernel path and schedule."""
import numpy as np
import pytest

from espins_amd import lattice as Lt

pytestmark = pytest.mark.gpu


def chains():
    return [
        (Lt.cubic_model("sc", 6, J=(1.0, -0.2)), Lt.ladder("log", 0.5, 2.5, 8)),          # LDS
        (Lt.cubic_model("fcc", 8, J=(1.0, 0.15)), Lt.ladder("log", 0.6, 2.0, 6)),         # global
        (Lt.cubic_model("bcc", 5, J=(1.0, 0.1)), Lt.ladder("log", 0.5, 2.5, 5)),          # LDS
        (Lt.cubic_model("sc", 16, J=(1.0,), B=0.05, D=0.3, K=0.02), Lt.ladder("log", 0.3, 1.2, 4)),
        (Lt.cubic_model_fast("bcc", 12, J=(1.0, -0.25)), Lt.ladder("log", 0.5, 2.5, 7)),
    ]


def test_batch_chains_match_single_contexts():
    from espins_amd.engine import Batch, Engine, set_stream_base
    seed = 424242
    cs = chains()
    init = [Lt.random_spins(m.N, len(t), 10 + k, m.mag) for k, (m, t) in enumerate(cs)]
    sched = dict(steps_per_round=2, steps_measure=2, measure=True, exchange=True, seed=seed)
    with Batch() as b:
        for (m, t), s in zip(cs, init):
            b.add_chain(m, t, s)
        b.set_chain_config(1, 1, 1)     # global-memory path, 1 lane per site
        b.set_chain_config(3, 1, 2)     # global-memory path, 2 lanes per site
        b.commit()
        info = [b.info(k) for k in range(len(cs))]
        print(info, flush=True)
        assert any(i["lds"] for i in info) and any(not i["lds"] for i in info)
        assert len({i["variant"] for i in info}) >= 2
        b.run(6, **sched)
        res = [(b.spins(k), b.stats(k), b.pt_map(k)) for k in range(len(cs))]
    for k, ((m, t), s) in enumerate(zip(cs, init)):
        with Engine() as eng:
            eng.set_model(m)
            eng.set_path(2 if info[k]["lds"] else 1)
            eng.set_tps(info[k]["tps"])
            set_stream_base(eng, k << 16)
            eng.set_replicas(t, s)
            eng.run(6, graph=False, **sched)
            sp, st, mp = eng.spins(), eng.stats(), eng.pt_map()
        bs, bst, bmp = res[k]
        assert np.array_equal(bst["rejected"], st["rejected"]), k
        assert np.array_equal(bmp[0], mp[0]) and np.array_equal(bmp[1], mp[1]), k
        np.testing.assert_allclose(bs, sp, rtol=0, atol=1e-12)
        np.testing.assert_allclose(bst["acc"], st["acc"], rtol=1e-10)
        assert bst["nmeas"] == st["nmeas"] == 12


def test_batch_energies_and_graph_replay():
    from espins_amd.engine import Batch
    from oracle import oracle as O
    cs = chains()[:3]
    init = [Lt.random_spins(m.N, len(t), 30 + k, m.mag) for k, (m, t) in enumerate(cs)]
    outs = []
    for graph in (False, True):
        with Batch() as b:
            for (m, t), s in zip(cs, init):
                b.add_chain(m, t, s)
            b.commit()
            for k, ((m, t), s) in enumerate(zip(cs, init)):
                E = b.energy_get(k)
                o = O.Oracle(m)
                for r in range(len(t)):
                    assert E[r] == pytest.approx(o.energy(s[r]), rel=1e-12, abs=1e-10)
            b.run(5, measure=True, exchange=True, seed=3, graph=graph)
            outs.append([(b.spins(k), b.stats(k)["acc"]) for k in range(len(cs))])
            for k in range(len(cs)):   # running energy tracks the direct one
                assert np.allclose(b.stats(k)["e_run"], b.energy_get(k), rtol=1e-10, atol=1e-8)
    for a, c in zip(outs[0], outs[1]):
        assert np.array_equal(a[0], c[0]) and np.array_equal(a[1], c[1])


def test_autotune_picks_valid_shapes(tmp_path):
    from espins_amd.engine import Batch
    cs = chains()
    cache = str(tmp_path / "tune.txt")
    with Batch() as b:
        for k, (m, t) in enumerate(cs):
            b.add_chain(m, t, Lt.random_spins(m.N, len(t), k, m.mag))
        b.commit()
        b.autotune(nsweeps=3, cache=cache)
        for k in range(len(cs)):
            i = b.info(k)
            assert i["tps"] in (1, 2, 4) and i["chunk"] in (2, 4)
            assert (not i["lds"]) or i["lds_fits"]
        b.run(3, measure=True, exchange=True, seed=1)
        for k in range(len(cs)):
            assert b.stats(k)["nmeas"] == 3
    lines = open(cache).read().split("\n")
    assert len([l for l in lines if l.strip()]) == len(cs)


def test_autotune_leaves_chain_state_unchanged(tmp_path):
    """Tuning times real sweeps and exchanges on the chains' own buffers; afterwards spins,
    labels, swap counters and running energies are bit-identical to before, with and without
    a cache file, and a following run equals the same run of an untuned batch."""
    from espins_amd.engine import Batch
    cs = chains()
    cache = str(tmp_path / "tune.txt")
    init = [Lt.random_spins(m.N, len(t), 50 + k, m.mag) for k, (m, t) in enumerate(cs)]

    def state(b):
        return [(b.spins(k), b.pt_map(k), b.stats(k)["e_run"]) for k in range(len(cs))]

    runs = []
    for tune in (None, "fresh", "cached"):
        with Batch() as b:
            for (m, t), s in zip(cs, init):
                b.add_chain(m, t, s)
            b.commit()
            b.run(2, measure=True, exchange=True, seed=5)
            before = state(b)
            if tune:
                b.autotune(nsweeps=2, cache=cache)
                after = state(b)
                for x, y in zip(before, after):
                    assert np.array_equal(x[0], y[0])
                    assert np.array_equal(x[1][0], y[1][0])
                    assert np.array_equal(x[2], y[2])
            # the same kernel shapes for every run: the comparison is about the state only
            for k in range(len(cs)):
                b.set_chain_config(k, 1, 1)
            b.run(3, measure=True, exchange=True, seed=5, sweep0=4, step0=2)
            runs.append(state(b))
    for other in runs[1:]:
        for x, y in zip(runs[0], other):
            assert np.array_equal(x[0], y[0])
            assert np.array_equal(x[1][0], y[1][0])


def test_chunk_width_changes_nothing_but_speed():
    """Four partner gathers in flight per lane instead of two: the slots are still summed in
    order, so every chain's spins, rejections and labels are bit-identical; the estimators
    agree to 1e-12 (the compiler may contract the bond sums of the two kernel instances into
    different fused multiply-adds, which moves dE by an ulp)."""
    from espins_amd.engine import Batch
    cs = chains()
    init = [Lt.random_spins(m.N, len(t), 70 + k, m.mag) for k, (m, t) in enumerate(cs)]
    res = []
    for chunk in (2, 4):
        with Batch() as b:
            for (m, t), s in zip(cs, init):
                b.add_chain(m, t, s)
            for k in range(len(cs)):
                b.set_chain_config(k, 1, 1)
                b.set_chain_chunk(k, chunk)
                assert b.info(k)["chunk"] == chunk
            b.commit()
            b.run(3, measure=True, exchange=True, seed=9)
            res.append([(b.spins(k), b.stats(k), b.pt_map(k)) for k in range(len(cs))])
    for (s2, st2, m2), (s4, st4, m4) in zip(*res):
        assert np.array_equal(s2, s4)
        assert np.array_equal(st2["rejected"], st4["rejected"])
        np.testing.assert_allclose(st2["acc"], st4["acc"], rtol=1e-12)
        assert np.array_equal(m2[0], m4[0])
