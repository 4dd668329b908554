"""Learned proposal on the GPU (esp_learn_*, mode ESP_MODE_LEARNED) against the numpy
oracle: forward, gradient, Adam step and collected records; exactness of the
Metropolis-Hastings chain against closed-form distributions and plain Metropolis; and what a
trained network gains over the plain proposals."""
import math

import numpy as np
import pytest

from oracle import learn_oracle as LO
from espins_amd import lattice as Lt
from espins_amd import learn as Ln
from tests.test_gpu_stats import dimer, langevin, one_site
from tests.test_learn_oracle import random_records

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def weights32(seed, out_scale=0.4):
    return Ln.init_weights(seed, out_scale=out_scale)


@pytest.fixture(scope="module")
def Engine():
    from espins_amd.engine import Engine as E
    return E


@pytest.mark.parametrize("n", [1, 127, 1000])
def test_gradient_matches_oracle(Engine, n):
    say("grad", n)
    recs = random_records(n, 20 + n).astype(np.float32)
    w = weights32(21)
    with Engine() as eng:
        eng.learn_set_weights(w)
        eng.learn_put_records(recs)
        g, loss = eng.learn_grad()
    L, go = LO.loss_and_grad(w.astype(np.float64), recs.astype(np.float64))
    assert loss == pytest.approx(L, rel=2e-5, abs=2e-5)
    scale = np.abs(go).max()
    np.testing.assert_allclose(g, go, rtol=2e-3, atol=2e-4 * scale)


def test_eval_matches_oracle(Engine):
    recs = random_records(300, 22).astype(np.float32)
    w = weights32(23)
    with Engine() as eng:
        eng.learn_set_weights(w)
        eng.learn_put_records(recs)
        out = eng.learn_eval()
        assert np.array_equal(eng.learn_weights(), w)
    X, k0, gh, ax, s = LO.split_records(recs)
    o, _ = LO.forward(w.astype(np.float64), X)
    np.testing.assert_allclose(out[:, :6], o, rtol=1e-4, atol=1e-5)
    logw, kappa, _, _, _ = LO.mixture(o, k0, gh, ax)
    np.testing.assert_allclose(out[:, 6:8], kappa, rtol=1e-4)
    np.testing.assert_allclose(out[:, 9], np.exp(logw[:, 0]) / (1 - LO.UNIFORM), rtol=1e-4, atol=1e-6)
    lq = LO.log_q(o, k0, gh, ax, s)
    assert np.all(np.abs(out[:, 8] - lq) <= 1e-4 * (np.abs(lq) + np.maximum(1.0, k0)))


def test_sweep_network_matches_oracle(Engine):
    """The network the sweep runs (BF16 weights and hidden activations, matrix cores) against
    the oracle's restatement of that arithmetic, and within BF16 error of the FP32 network."""
    recs = random_records(1000, 26).astype(np.float32)   # 1000: a partial last wave
    w = weights32(27, out_scale=0.5)
    with Engine() as eng:
        eng.learn_set_weights(w)
        eng.learn_put_records(recs)
        out = eng.learn_eval(sweep_net=True)
        ref32 = eng.learn_eval(sweep_net=False)
    X = recs[:, :5].astype(np.float64)
    o = LO.forward_sweep_net(w.astype(np.float64), X)
    # agreement to FP32 rounding, except where an activation sits on a BF16 rounding
    # boundary and the two roundings differ by one BF16 step
    d = np.abs(out[:, :6] - o)
    scale = np.abs(ref32[:, :6]).max()
    assert np.mean(d > 1e-4 * (1 + np.abs(o))) < 2e-3, np.sort(d.ravel())[-10:]
    assert d.max() < 0.01 * scale
    assert np.abs(out[:, :6] - ref32[:, :6]).max() < 0.05 * scale
    # the mixture built from the sweep network is a normalised density as well
    X_, k0, gh, ax, s = LO.split_records(recs)
    lq = LO.log_q(out[:, :6], k0, gh, ax, s)
    assert np.all(np.abs(out[:, 8] - lq) <= 1e-4 * (np.abs(lq) + np.maximum(1.0, k0)))


def test_one_adam_step(Engine):
    """First Adam step from zero moments: w - lr * g / (|g| + eps)."""
    recs = random_records(512, 24).astype(np.float32)
    w = weights32(25)
    lr = 1e-3
    with Engine() as eng:
        eng.learn_set_weights(w)
        eng.learn_put_records(recs)
        g, _ = eng.learn_grad()
        loss = eng.learn_train(1, batch=512, lr=lr, seed=3)
        w1 = eng.learn_weights()
    assert loss[0] == pytest.approx(LO.loss(w.astype(np.float64), recs.astype(np.float64)), rel=1e-4)
    big = np.abs(g) > 1e-3 * np.abs(g).max()
    want = w.astype(np.float64) - lr * g / (np.abs(g) + 1e-8)
    np.testing.assert_allclose(w1[big], want[big], rtol=0, atol=2e-6)


def _collect_models():
    return [
        ("fcc_jb_si_h", Lt.cubic_model("fcc", 3, J=(-1.0, 0.1), B=0.1, K=-0.05, field=0.3,
                                       field_axis=(1, 1, 0))),
        ("sc_jbd_si", Lt.cubic_model("sc", 4, J=(1.0,), B=0.05, D=0.3, K=0.02, K_axis=(1, 0, 1))),
        ("bcc_j12", Lt.cubic_model("bcc", 4, J=(1.0, -0.25))),
    ]


@pytest.mark.parametrize("name,m", _collect_models(), ids=[n for n, _ in _collect_models()])
def test_collected_records_match_oracle(Engine, name, m):
    say("collect", name)
    R = 3
    tems = np.array([0.3, 0.9, 2.0])
    spins = Lt.random_spins(m.N, R, 31, m.mag)
    with Engine() as eng:
        eng.set_model(m).set_replicas(tems, spins)
        eng.learn_collect(400, seed=5)
        eng.learn_collect(100, seed=6)
        recs = eng.learn_records()
    assert recs.shape == (500, 16)
    ids = recs[:, 15].astype(np.int64)
    # draws spread over replicas and sites: 500 draws from R N ids leave
    # R N (1 - exp(-500 / (R N))) distinct ones on average
    RN = R * m.N
    assert len(set(ids.tolist())) > 0.85 * RN * (1 - math.exp(-500 / RN))
    for rec in recs[::5]:
        k = int(rec[15])
        r, i = divmod(k, m.N)
        want = LO.record(m, spins[r], r, i, 1.0 / tems[r])
        np.testing.assert_allclose(rec[:15], want[:15], rtol=2e-6, atol=2e-6)


def _exactness_cases():
    return [
        ("field", one_site(field=1.0), 0.7, lambda T: -1.0 * langevin(1.0 / T)),
        ("dimer", dimer(1.0), 0.8, lambda T: -1.0 * langevin(1.0 / T)),
    ]


@pytest.mark.parametrize("out_scale", [0.01, 0.8])
def test_learned_chain_samples_exact_distributions(Engine, out_scale):
    """Metropolis-Hastings with the learned proposal samples exp(-beta E) exactly, for a good
    and for a deliberately poor network: single spin in a field, single-ion axis, dimer."""
    w = weights32(40, out_scale=out_scale)
    R, n = 256, 3000
    for name, m, T, fn in _exactness_cases():
        with Engine() as eng:
            eng.set_model(m).set_replicas(np.full(R, T), Lt.random_spins(m.N, R, 2, m.mag))
            eng.learn_set_weights(w)
            eng.run(1, steps_per_round=200, measure=False, exchange=False, mode="learned", seed=7)
            eng.reset_stats()
            eng.run(1, steps_per_round=n, measure=True, exchange=False, mode="learned", seed=7,
                    sweep0=200)
            e = eng.stats()["acc"][:, 0] / n
        want = fn(T)
        err = e.std(ddof=1) / math.sqrt(R)
        say("exact", name, out_scale, e.mean(), want, err)
        assert abs(e.mean() - want) < 4 * err, (name, e.mean(), want, err)
    Dl, T = -1.0, 0.5                                     # E = Delta (S.z)^2
    with Engine() as eng:
        eng.set_model(one_site(K=Dl)).set_replicas(np.full(R, T), Lt.random_spins(1, R, 3))
        eng.learn_set_weights(w)
        eng.run(1, steps_per_round=200, measure=False, exchange=False, mode="learned", seed=8)
        eng.reset_stats()
        eng.run(1, steps_per_round=n, measure=True, exchange=False, mode="learned", seed=8,
                sweep0=200)
        e = eng.stats()["acc"][:, 0] / n
    x = np.linspace(-1, 1, 200001)
    wt = np.exp(-Dl * x * x / T)
    want = Dl * np.trapezoid(x * x * wt, x) / np.trapezoid(wt, x)
    err = e.std(ddof=1) / math.sqrt(R)
    say("exact single-ion", out_scale, e.mean(), want, err)
    assert abs(e.mean() - want) < 4 * err


def test_learned_run_graph_replay_and_running_energy(Engine):
    m, tems = Lt.config_c2(L=4, R=8)
    spins = Lt.random_spins(m.N, 8, 1, m.mag)
    w = weights32(41)
    outs = []
    for graph in (False, True, True):
        with Engine() as eng:
            eng.set_model(m).set_replicas(tems, spins)
            eng.learn_set_weights(w)
            eng.run(6, steps_per_round=2, steps_measure=1, measure=True, exchange=True, mode="learned",
                    seed=11, graph=graph)
            st = eng.stats()
            outs.append((eng.spins(), st["acc"], st["rejected"], eng.pt_map()[0]))
            np.testing.assert_allclose(st["e_run"], eng.energy_get(), rtol=1e-10, atol=1e-8)
            assert st["rejected"].sum() > 0
    for o in outs[1:]:
        for a, b in zip(o, outs[0]):
            assert np.array_equal(a, b)


def _frustrated(L=4):
    return Lt.cubic_model("fcc", L, J=(-1.0, 0.1), B=0.1, K=-0.05)


def test_learned_observables_match_plain_metropolis(Engine):
    """Same model, same temperatures, independent copies: <E> and <|m|> per temperature of
    the learned chain agree with plain Metropolis within errors."""
    m = _frustrated(3)
    temps = np.array([0.15, 0.4, 1.0])
    copies = 24
    T = np.repeat(temps, copies)
    R = T.size
    n = 3000
    res = {}
    for mode in ("random", "learned"):
        with Engine() as eng:
            eng.set_model(m).set_replicas(T, Lt.random_spins(m.N, R, 4, m.mag))
            eng.learn_set_weights(weights32(42, out_scale=0.05))
            eng.run(1, steps_per_round=1500, measure=False, exchange=False, mode=mode, seed=12)
            eng.reset_stats()
            eng.run(1, steps_per_round=n, measure=True, exchange=False, mode=mode, seed=12,
                    sweep0=1500)
            res[mode] = (eng.stats()["acc"] / n).reshape(len(temps), copies, 6)
    for k in (0, 3):
        a, b = res["random"][:, :, k], res["learned"][:, :, k]
        d = a.mean(1) - b.mean(1)
        err = np.sqrt(a.var(1, ddof=1) / copies + b.var(1, ddof=1) / copies)
        say("plain vs learned", k, a.mean(1), b.mean(1), err)
        assert np.all(np.abs(d) < 4.5 * err + 1e-12), (k, d, err)


def test_trained_proposal_gains_over_plain_metropolis(Engine):
    """Train on local conditionals from plain chains, then compare acceptance and effective
    samples of E per sweep (tau_int by Sokal windowing) with both plain proposals."""
    m = _frustrated(4)
    temps = np.array([0.05, 0.1, 0.2])
    copies = 8
    T = np.repeat(temps, copies)
    R = T.size
    warm, n = 600, 3000
    spins = Lt.random_spins(m.N, R, 5, m.mag)
    with Engine() as eng:
        eng.set_model(m).set_replicas(T, spins)
        eng.run(1, steps_per_round=warm, measure=False, exchange=False, mode="const", tilt=0.3, seed=13)
        losses = Ln.train_proposal(eng, n_records=1 << 16, epochs=4, batch=1024, lr=3e-3, seed=3,
                                   mode="const", sweep0=warm)
        say("losses", losses[0], losses[-1])
        assert losses[-1] < losses[0]
        w = eng.learn_weights()
        start = eng.spins()
    ess, acc = {}, {}
    for mode, tilt in (("random", 0.0), ("const", 0.3), ("learned", 0.0)):
        with Engine() as eng:
            eng.set_model(m).set_replicas(T, start)
            eng.learn_set_weights(w)
            eng.run(1, steps_per_round=300, measure=False, exchange=False, mode=mode, tilt=tilt, seed=14)
            eng.reset_stats()
            eng.set_trace(n)
            eng.run(1, steps_per_round=n, measure=True, exchange=False, mode=mode, tilt=tilt, seed=14,
                    sweep0=300)
            tr = eng.trace()[:, :, 0]
            acc[mode] = 1.0 - eng.stats()["rejected"] / (n * m.N)
        e, tau = Ln.ess_per_slot(tr)
        ess[mode] = e.reshape(len(temps), copies).mean(1) / n
        say(mode, "acceptance", acc[mode].reshape(len(temps), copies).mean(1), "ESS/sweep", ess[mode])
    a_l = acc["learned"].reshape(len(temps), copies).mean(1)
    a_r = acc["random"].reshape(len(temps), copies).mean(1)
    assert np.all(a_l > 0.5) and np.all(a_l > 3 * a_r)
    assert ess["learned"][0] > 3 * ess["random"][0]
    assert ess["learned"][0] > ess["const"][0]
