"""Pins of the learned-proposal oracle (oracle/learn_oracle.py): finite differences, sphere
quadrature, scipy's von Mises-Fisher, the closed-form CDF of Wood's sampler, the C oracle's
delta_energy_get and the heat-bath special case.  CPU only."""
import math
import os

import numpy as np
import pytest
from scipy import integrate, stats

from oracle import learn_oracle as LO
from oracle import oracle as O
from espins_amd import lattice as Lt
from espins_amd import learn as Ln

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def random_records(n, seed, kmax=30.0):
    rng = np.random.default_rng(seed)
    recs = np.zeros((n, 16))
    recs[:, 0:5] = rng.normal(0, 1, (n, 5))
    recs[:, 5] = rng.uniform(0.05, kmax, n)
    recs[:, 6:9] = _unit(rng.normal(size=(n, 3)))
    recs[:, 9:12] = _unit(rng.normal(size=(n, 3)))
    recs[:, 12:15] = _unit(rng.normal(size=(n, 3)))
    return recs


def random_weights(seed, out_scale=0.4):
    w = Ln.init_weights(seed, out_scale=out_scale).astype(np.float64)
    return w


def test_layout_matches_header():
    hdr = open(os.path.join(ROOT, "include", "espins_hip.h")).read()
    assert f"#define ESP_LEARN_NPARAMS {LO.NPARAMS}" in hdr
    assert "#define ESP_LEARN_RECORD  16" in hdr
    assert Ln.init_weights(0).size == LO.NPARAMS


def test_gradient_matches_central_differences():
    recs = random_records(48, 1)
    w = random_weights(2)
    L0, g = LO.loss_and_grad(w, recs)
    assert L0 == pytest.approx(LO.loss(w, recs), rel=1e-13)
    rng = np.random.default_rng(3)
    idx = np.concatenate([rng.choice(LO.NPARAMS, 200, replace=False),
                          np.arange(LO.NPARAMS - 6, LO.NPARAMS)])   # every output bias
    h = 1e-6
    for p in idx:
        e = np.zeros(LO.NPARAMS)
        e[p] = h
        fd = (LO.loss(w + e, recs) - LO.loss(w - e, recs)) / (2 * h)
        assert abs(fd - g[p]) <= 1e-6 + 1e-5 * abs(fd), (p, fd, g[p])
    # directional derivatives along random directions cover every parameter at once
    for k in range(5):
        d = rng.normal(size=LO.NPARAMS)
        d /= np.linalg.norm(d)
        fd = (LO.loss(w + h * d, recs) - LO.loss(w - h * d, recs)) / (2 * h)
        assert abs(fd - g @ d) <= 1e-6 + 1e-5 * abs(fd)


def test_gradient_through_clamped_log_scale_is_zero():
    recs = random_records(8, 4)
    w = random_weights(5)
    p = LO.unpack(w)
    p["b4"][0] = 50.0          # o[0] far above the clamp: kappa_0 no longer depends on it
    w = LO.pack(p)
    _, g = LO.loss_and_grad(w, recs)
    assert g[LO.NPARAMS - 6] == 0.0


@pytest.mark.parametrize("kappa", [1e-5, 0.3, 1.0, 7.0, 80.0, 900.0])
def test_vmf_density_is_normalised(kappa):
    mu = np.array([0.0, 0.0, 1.0])

    def f(w):
        x = np.array([math.sqrt(max(0.0, 1 - w * w)), 0.0, w])
        return math.exp(LO.log_vmf(kappa, mu, x))
    pts = [1 - 10.0 / kappa] if kappa > 20 else None
    tot, _ = integrate.quad(f, -1, 1, points=pts, limit=400, epsabs=1e-13, epsrel=1e-11)
    assert 2 * math.pi * tot == pytest.approx(1.0, rel=1e-8)


def test_vmf_matches_scipy():
    rng = np.random.default_rng(6)
    for kappa in (0.2, 3.0, 40.0, 300.0):
        mu = _unit(rng.normal(size=3))
        x = _unit(rng.normal(size=(20, 3)))
        x[0] = mu
        want = stats.vonmises_fisher(mu, kappa).logpdf(x)
        got = LO.log_vmf(kappa, mu, x)
        np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-10)


def test_mixture_density_is_normalised():
    """q integrates to one over the sphere (Fibonacci grid, moderate kappa)."""
    n = 400000
    k = np.arange(n) + 0.5
    z = 1 - 2 * k / n
    ph = math.pi * (1 + 5 ** 0.5) * k
    r = np.sqrt(1 - z * z)
    grid = np.stack([r * np.cos(ph), r * np.sin(ph), z], -1)
    rng = np.random.default_rng(7)
    for t in range(4):
        o = rng.normal(0, 1.0, (1, 6))
        k0 = np.array([rng.uniform(0.1, 6.0)])
        gh = _unit(rng.normal(size=(1, 3)))
        ax = _unit(rng.normal(size=(1, 3)))
        lq = LO.log_q(np.repeat(o, n, 0), np.repeat(k0, n), np.repeat(gh, n, 0), np.repeat(ax, n, 0), grid)
        assert 4 * math.pi * np.exp(lq).mean() == pytest.approx(1.0, rel=2e-4)


@pytest.mark.parametrize("kappa", [0.01, 1.0, 30.0, 500.0])
def test_wood_sampler_matches_vmf_cdf(kappa):
    rng = np.random.default_rng(int(kappa * 100) + 1)
    mu = _unit(rng.normal(size=3))
    n = 40000
    x = LO.sample_vmf(np.full(n, kappa), np.tile(mu, (n, 1)), rng.random(n), rng.random(n))
    np.testing.assert_allclose(np.linalg.norm(x, axis=1), 1.0, atol=1e-12)
    w = x @ mu

    def cdf(t):   # density of w = mu.x is proportional to exp(kappa w) on [-1, 1]
        return (np.exp(kappa * (t - 1)) - math.exp(-2 * kappa)) / (-math.expm1(-2 * kappa))
    assert stats.kstest(w, cdf).pvalue > 1e-3
    # azimuth about mu is uniform
    e1 = _unit(np.cross(mu, [0.3, -0.5, 0.8]))
    e2 = np.cross(mu, e1)
    ph = np.arctan2(x @ e2, x @ e1)
    keep = np.abs(w) < 0.999999
    assert stats.kstest(ph[keep], stats.uniform(-math.pi, 2 * math.pi).cdf).pvalue > 1e-3


def _models():
    return [
        ("fcc_jb_si_h", Lt.cubic_model("fcc", 3, J=(-1.0, 0.1), B=0.1, K=-0.05, field=0.3,
                                       field_axis=(1, 1, 0))),
        ("sc_jbd_si", Lt.cubic_model("sc", 4, J=(1.0,), B=0.05, D=0.3, K=0.02, K_axis=(1, 0, 1))),
    ]


@pytest.mark.parametrize("name,m", _models(), ids=[n for n, _ in _models()])
def test_site_energy_reproduces_delta_energy(name, m):
    """E_i(S') - E_i(S) from (G, Q) equals the C oracle's delta_energy_get for every site."""
    o = O.Oracle(m)
    rng = np.random.default_rng(8)
    s = Lt.random_spins(m.N, 1, 9, m.mag)[0]
    for i in range(0, m.N, max(1, m.N // 40)):
        G, QB, Delta, ax = LO.local_env(m, s, i)
        new = _unit(rng.normal(size=3)) * m.mag[i]
        old = s[i].copy()
        de_local = LO.site_energy(G, QB, Delta, ax, new) - LO.site_energy(G, QB, Delta, ax, old)
        s2 = s.copy()
        s2[i] = new
        de_ref = o.delta_energy(i, s2, new - old)
        assert de_local == pytest.approx(de_ref, rel=1e-11, abs=1e-12), i
        # and the total energy difference
        assert de_ref == pytest.approx(o.energy(s2) - o.energy(s), rel=1e-9, abs=1e-9)


def test_zero_output_layer_is_the_heat_bath_of_an_exchange_site():
    """W4 = b4 = 0: the network part of q is vMF(ghat, beta |m| |G|), the exact conditional
    exp(-beta E_i(x)) of a pure exchange site, so its log plus beta E_i(x) is the same for
    every x."""
    m = Lt.cubic_model("fcc", 3, J=(-1.0, 0.2))
    s = Lt.random_spins(m.N, 1, 10, m.mag)[0]
    p = LO.unpack(random_weights(11))
    p["W4"][:] = 0.0
    p["b4"][:] = 0.0
    w = LO.pack(p)
    beta = 1.7
    rng = np.random.default_rng(12)
    for i in (0, 5, 17):
        rec = LO.record(m, s, 0, i, beta)
        xs = _unit(rng.normal(size=(50, 3)))
        o, _ = LO.forward(w, np.tile(rec[:5], (50, 1)))
        logw, kappa, mu, _, _ = LO.mixture(o, np.full(50, rec[5]), np.tile(rec[6:9], (50, 1)),
                                           np.tile(rec[9:12], (50, 1)))
        np.testing.assert_allclose(np.exp(logw), (1 - LO.UNIFORM) / 2, rtol=1e-14)
        lq = LO.log_vmf(kappa[:, 0], mu[:, 0], xs)      # both network components coincide
        np.testing.assert_allclose(lq, LO.log_vmf(kappa[:, 1], mu[:, 1], xs), rtol=1e-14)
        G, QB, Delta, ax = LO.local_env(m, s, i)
        bE = np.array([beta * LO.site_energy(G, QB, Delta, ax, x * m.mag[i]) for x in xs])
        c = lq + bE
        np.testing.assert_allclose(c, c[0], rtol=0, atol=1e-5 * max(1.0, rec[5]))


def test_autocorrelation_time_of_ar1():
    """AR(1) with coefficient a has tau_int = (1 + a) / (2 (1 - a))."""
    rng = np.random.default_rng(13)
    for a in (0.0, 0.5, 0.9):
        n = 200000
        e = rng.normal(size=n)
        x = np.empty(n)
        x[0] = e[0]
        for t in range(1, n):
            x[t] = a * x[t - 1] + e[t]
        want = (1 + a) / (2 * (1 - a))
        got = Ln.autocorr_time(x)
        assert got == pytest.approx(want, rel=0.08), (a, got, want)


# ---------------------------------------------------------------- features -------------------
def test_features_closed_forms():
    """The five network inputs in closed-form cases: x0 = log(1 + beta m |G|), x1 = |ghat.a|
    (1 when G is along the axis, 0 when normal to it), x2 = log T, x3 = beta m^2 (ghat^T Q_B
    ghat - tr Q_B / 3) (zero for Q_B = 0 and for an isotropic Q_B = b I, (2/3) beta m^2 b for
    Q_B = b ghat ghat^T), x4 = beta m^2 Delta."""
    ax = np.array([0.0, 0.0, 1.0])
    beta, mag, g, b, Delta = 1.6, 1.5, 2.0, 0.35, -0.2
    x, k0, gh = LO.features(g * ax, np.zeros((3, 3)), Delta, ax, beta, mag)
    assert k0 == pytest.approx(beta * mag * g, rel=1e-15)
    assert x[0] == pytest.approx(math.log(1.0 + beta * mag * g), rel=1e-15)
    assert x[1] == 1.0 and x[3] == 0.0
    assert x[2] == pytest.approx(math.log(1.0 / beta), rel=1e-15)
    assert x[4] == pytest.approx(beta * mag * mag * Delta, rel=1e-15)
    np.testing.assert_allclose(gh, ax, rtol=0, atol=1e-16)
    # G normal to the axis, B along G
    G = np.array([0.0, -g, 0.0])
    x, _, gh = LO.features(G, b * np.outer(G / g, G / g), Delta, ax, beta, mag)
    assert x[1] == 0.0
    assert x[3] == pytest.approx(2.0 / 3.0 * beta * mag * mag * b, rel=1e-14)
    x, _, _ = LO.features(G, b * np.eye(3), Delta, ax, beta, mag)
    assert x[3] == pytest.approx(0.0, abs=1e-15)
    # the 45-degree case: |ghat.a| = 1/sqrt(2) for either sign of G
    for sgn in (1.0, -1.0):
        x, _, _ = LO.features(sgn * np.array([1.0, 0.0, 1.0]), np.zeros((3, 3)), 0.0, ax, beta, mag)
        assert x[1] == pytest.approx(1.0 / math.sqrt(2.0), rel=1e-15)


def test_record_of_a_ferro_site_in_closed_form():
    """local_env + features on the ferro state of an SC J1 + B1 + Delta model: every site has
    G = z J m z-hat and Q_B = z B m^2 z-hat z-hat^T, so x0 = log(1 + beta m z J m),
    x1 = 1, x3 = (2/3) beta m^2 z B m^2, x4 = beta m^2 Delta (z = 6)."""
    J, B, Delta, beta, mag = 1.0, 0.1, -0.05, 2.0, 1.0
    m = Lt.cubic_model("sc", 3, J=(J,), B=B, K=Delta)
    s = Lt.ferro_spins(m.N, 1, m.mag)[0]
    for i in (0, 13):
        rec = LO.record(m, s, 0, i, beta)
        assert rec[0] == pytest.approx(math.log1p(beta * mag * 6 * J * mag), rel=1e-14)
        assert rec[1] == 1.0
        assert rec[3] == pytest.approx(2.0 / 3.0 * beta * mag * mag * 6 * B * mag * mag, rel=1e-14)
        assert rec[4] == pytest.approx(beta * mag * mag * Delta, rel=1e-15)
        np.testing.assert_allclose(rec[6:9], [0, 0, 1], atol=1e-16)
        np.testing.assert_allclose(rec[12:15], [0, 0, 1], atol=1e-16)


# ---------------------------------------------------------------- sweep network --------------
def test_bf16_round_matches_torch_bfloat16():
    """Round to nearest even BF16, against torch's own float -> bfloat16 conversion, including
    exact ties (low 16 bits 0x8000) on both parities."""
    import torch
    rng = np.random.default_rng(21)
    a = np.concatenate([rng.normal(0, 3, 4000), rng.uniform(-1e-3, 1e-3, 1000)]).astype(np.float32)
    u = a.view(np.uint32)
    ties = ((u[:200] & 0xFFFF0000) | 0x8000).astype(np.uint32).view(np.float32)
    a = np.concatenate([a, ties])
    want = torch.from_numpy(a).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(LO.bf16_round(a), want)


def test_forward_sweep_net_hand_computed_case():
    """One path through the network with hand-chosen weights: x0 = 1 feeds hidden unit 0 of
    layer 1 with weight 1; layers 2 and 3 copy unit 0 with weight 1; output 0 reads unit 0 with
    weight 1 and output 1 is the bias 0.1.  With g the tanh-form GELU,
    g(x) = 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3))):
      a1 = bf16(g(1))               = bf16(0.84119199...) = 0.83984375 (= 215/256)
      a2 = bf16(g(0.83984375))      = bf16(0.67135989...) = 0.671875   (= 172/256)
      a3 = g(0.671875)              (full precision)
      o0 = a3, o1 = bf16(0.1) = 0.10009765625, others 0"""
    p = {k: np.zeros(s) for k, s in LO._SHAPES}
    p["W1"][0, 0] = 1.0
    p["W2"][0, 0] = 1.0
    p["W3"][0, 0] = 1.0
    p["W4"][0, 0] = 1.0
    p["b4"][1] = 0.1
    w = LO.pack(p)
    X = np.array([[1.0, 0.0, 0.0, 0.0, 0.0]])
    h = lambda t: 0.5 * t * (1.0 + math.tanh(math.sqrt(2.0 / math.pi) * (t + 0.044715 * t ** 3)))
    a1 = 215.0 / 256.0
    assert abs(h(1.0) - a1) < 2.0 ** -9        # within half a BF16 ulp at this magnitude
    a2 = 172.0 / 256.0
    assert abs(h(a1) - a2) < 2.0 ** -9
    a3 = h(a2)
    o = LO.forward_sweep_net(w, X)[0]
    assert o[0] == pytest.approx(a3, rel=1e-14)
    assert o[1] == 0.10009765625
    assert np.all(o[2:] == 0.0)
    # the full-precision network keeps the unrounded chain and differs from it
    of, _ = LO.forward(w, X)
    want = h(h(h(1.0)))
    assert of[0, 0] == pytest.approx(want, rel=1e-14)
    assert abs(of[0, 0] - o[0]) > 1e-4


def test_activation_is_torch_tanh_gelu():
    """The network's activation and its derivative against an independent implementation:
    torch's gelu(approximate='tanh') and its autograd gradient, in float64."""
    import torch
    x = torch.linspace(-12.0, 12.0, 4001, dtype=torch.float64, requires_grad=True)
    y = torch.nn.functional.gelu(x, approximate="tanh")
    (g,) = torch.autograd.grad(y.sum(), x)
    xn = x.detach().numpy()
    np.testing.assert_allclose(LO.act(xn), y.detach().numpy(), rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(LO.act_d(xn), g.numpy(), rtol=1e-12, atol=1e-14)
