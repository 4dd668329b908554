"""CPU restatement (numpy, FP64) of the learned Metropolis-Hastings proposal.

TEST INFRASTRUCTURE ONLY: imported by tests/ and nothing else.  The product computes all of
this on the GPU (espins_amd/csrc/esp_learn.hip); this module is the checker.

The learned proposal has no counterpart in the reference.  What it must agree with:
  * the site energy it is built on, E_i(S) = -G.S + S^T Q S, has to reproduce
    delta_energy_get (get_quant.F90:264-337) -- pinned in tests against the C oracle's
    orc_delta_energy, itself pinned to the reference's published energies;
  * the von Mises-Fisher density has to be normalised -- pinned by quadrature and against
    scipy.stats.vonmises_fisher;
  * the loss gradient has to be the derivative of the loss -- pinned by central differences;
  * Wood's sampler (Wood 1994, "Simulation of the von Mises Fisher distribution") has to draw
    from that density -- pinned by a Kolmogorov-Smirnov test against the closed-form CDF.

Layouts (identical to include/espins_hip.h):
  weights  W1[32][5] b1[32] W2[32][32] b2[32] W3[32][32] b3[32] W4[6][32] b4[6]
  record   x[5] kappa0 ghat[3] axis[3] s[3] id
  outputs  per component k: o[3k] log kappa scale, o[3k+1] weight logit, o[3k+2] tilt
"""
from __future__ import annotations

import math

import numpy as np

NIN, NH, NOUT = 5, 32, 6
UNIFORM = 1.0 / 32.0          # weight of the fixed uniform-sphere component
_SHAPES = [("W1", (NH, NIN)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)),
           ("W3", (NH, NH)), ("b3", (NH,)), ("W4", (NOUT, NH)), ("b4", (NOUT,))]
NPARAMS = sum(int(np.prod(s)) for _, s in _SHAPES)
assert NPARAMS == 2502


def unpack(w):
    w = np.asarray(w, np.float64)
    out, o = {}, 0
    for name, shp in _SHAPES:
        n = int(np.prod(shp))
        out[name] = w[o:o + n].reshape(shp)
        o += n
    return out


def pack(p):
    return np.concatenate([np.asarray(p[name], np.float64).ravel() for name, _ in _SHAPES])


GELU_C = 2.0 * math.sqrt(2.0 / math.pi)


def act(x):
    """GELU, tanh form (Hendrycks & Gimpel 2016): 0.5 x (1 + tanh(c (x + 0.044715 x^3) / 2))
    = x sigma(c (x + 0.044715 x^3)), c = 2 sqrt(2 / pi)."""
    x = np.asarray(x, np.float64)
    return 0.5 * x * (1.0 + np.tanh(0.5 * GELU_C * (x + 0.044715 * x ** 3)))


def act_d(x):
    x = np.asarray(x, np.float64)
    s = 0.5 * (1.0 + np.tanh(0.5 * GELU_C * (x + 0.044715 * x ** 3)))
    return s + x * s * (1.0 - s) * GELU_C * (1.0 + 3.0 * 0.044715 * x * x)


def forward(w, X):
    """Outputs [n][6] and the pre-activations for the backward pass."""
    p = unpack(w)
    z1 = X @ p["W1"].T + p["b1"]
    a1 = act(z1)
    z2 = a1 @ p["W2"].T + p["b2"]
    a2 = act(z2)
    z3 = a2 @ p["W3"].T + p["b3"]
    a3 = act(z3)
    o = a3 @ p["W4"].T + p["b4"]
    return o, (X, z1, a1, z2, a2, z3, a3)


def bf16_round(a):
    """Round to the nearest BF16 (ties to even), returned as float64."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward_sweep_net(w, X):
    """The network as the sweep evaluates it: BF16 weights; the inputs of the two 32 x 32
    layers rounded to BF16; sums and activations otherwise in full precision."""
    p = unpack(bf16_round(w))
    a1 = bf16_round(act(X @ p["W1"].T + p["b1"]))
    a2 = bf16_round(act(a1 @ p["W2"].T + p["b2"]))
    a3 = act(a2 @ p["W3"].T + p["b3"])
    return a3 @ p["W4"].T + p["b4"]


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def tilt_frame(ghat, ax):
    """Unit vector normal to ghat in the (ghat, axis) plane, toward the nearer pole."""
    c = np.sum(ghat * ax, axis=-1, keepdims=True)
    sgn = np.where(c < 0.0, -1.0, 1.0)
    e = sgn * ax - sgn * c * ghat
    en = np.linalg.norm(e, axis=-1, keepdims=True)
    alt = np.where(np.abs(ghat[..., :1]) < 0.9,
                   np.stack([np.zeros_like(ghat[..., 0]), -ghat[..., 2], ghat[..., 1]], -1),
                   np.stack([-ghat[..., 2], np.zeros_like(ghat[..., 0]), ghat[..., 0]], -1))
    e = np.where(en < 1e-12, alt, e)
    return e / np.linalg.norm(e, axis=-1, keepdims=True)


def mixture(o, kappa0, ghat, ax):
    """(log of (1 - UNIFORM) w_k [n,2], kappa [n,2], mu [n,2,3], tilt t [n,2], frame e [n,3])."""
    o = np.asarray(o, np.float64)
    a = o[:, [1, 4]]
    mx = a.max(axis=1, keepdims=True)
    logw = a - (mx + np.log(np.exp(a - mx).sum(axis=1, keepdims=True))) + math.log1p(-UNIFORM)
    s = np.clip(o[:, [0, 3]], -8.0, 8.0)
    kappa = np.asarray(kappa0, np.float64)[:, None] * np.exp(s) + 1e-6
    t = 0.5 * math.pi * np.tanh(o[:, [2, 5]])
    e = tilt_frame(ghat, ax)
    mu = np.cos(t)[..., None] * ghat[:, None, :] + np.sin(t)[..., None] * e[:, None, :]
    return logw, kappa, mu, t, e


def log_vmf(kappa, mu, x):
    """log of kappa / (4 pi sinh kappa) exp(kappa mu.x), stable for large kappa."""
    kappa = np.asarray(kappa, np.float64)
    d = np.sum(mu * x, axis=-1)
    return (np.log(kappa) - math.log(2.0 * math.pi) - np.log1p(-np.exp(-2.0 * kappa))
            + kappa * (d - 1.0))


def _with_uniform(lv):
    lu = np.full(lv.shape[:1] + (1,), math.log(UNIFORM / (4.0 * math.pi)))
    return np.concatenate([lv, lu], axis=1)


def log_q(o, kappa0, ghat, ax, x):
    """log of (1 - UNIFORM) sum_k w_k vMF(x; kappa_k, mu_k) + UNIFORM / (4 pi)."""
    logw, kappa, mu, _, _ = mixture(o, kappa0, ghat, ax)
    lv = _with_uniform(logw + log_vmf(kappa, mu, x[:, None, :]))
    mx = lv.max(axis=1, keepdims=True)
    return (mx + np.log(np.exp(lv - mx).sum(axis=1, keepdims=True)))[:, 0]


def split_records(recs):
    recs = np.asarray(recs, np.float64)
    return (recs[:, 0:5], recs[:, 5], _unit(recs[:, 6:9]), _unit(recs[:, 9:12]), _unit(recs[:, 12:15]))


def loss(w, recs):
    X, k0, gh, ax, s = split_records(recs)
    o, _ = forward(w, X)
    return float(-np.mean(log_q(o, k0, gh, ax, s)))


def loss_and_grad(w, recs):
    """Mean negative log-likelihood over the records and its gradient (backpropagation)."""
    X, k0, gh, ax, s = split_records(recs)
    n = X.shape[0]
    o, (X, z1, a1, z2, a2, z3, a3) = forward(w, X)
    logw, kappa, mu, t, e = mixture(o, k0, gh, ax)
    lv = _with_uniform(logw + log_vmf(kappa, mu, s[:, None, :]))
    mx = lv.max(axis=1, keepdims=True)
    lse = mx + np.log(np.exp(lv - mx).sum(axis=1, keepdims=True))
    r = np.exp(lv - lse)[:, :2]                            # responsibilities of the network part
    mus = np.sum(mu * s[:, None, :], axis=-1)
    with np.errstate(over="ignore"):   # 2 / expm1(2 kappa) -> 0 for large kappa
        dlk = 1.0 / kappa - 2.0 / np.expm1(2.0 * kappa) + (mus - 1.0)
    sraw = o[:, [0, 3]]
    dks = np.where((sraw > -8.0) & (sraw < 8.0), kappa - 1e-6, 0.0)
    th = np.tanh(o[:, [2, 5]])
    dmu = -np.sin(t)[..., None] * gh[:, None, :] + np.cos(t)[..., None] * e[:, None, :]
    g = np.zeros_like(o)
    g[:, [0, 3]] = -r * dlk * dks
    g[:, [1, 4]] = np.exp(logw) / (1.0 - UNIFORM) * r.sum(axis=1, keepdims=True) - r
    g[:, [2, 5]] = -r * kappa * np.sum(dmu * s[:, None, :], axis=-1) * 0.5 * math.pi * (1.0 - th * th)
    g /= n
    p = unpack(w)
    gr = {"W4": g.T @ a3, "b4": g.sum(0)}
    d3 = (g @ p["W4"]) * act_d(z3)
    gr["W3"], gr["b3"] = d3.T @ a2, d3.sum(0)
    d2 = (d3 @ p["W3"]) * act_d(z2)
    gr["W2"], gr["b2"] = d2.T @ a1, d2.sum(0)
    d1 = (d2 @ p["W2"]) * act_d(z1)
    gr["W1"], gr["b1"] = d1.T @ X, d1.sum(0)
    return float(-np.mean(lse)), pack(gr)


def sample_vmf(kappa, mu, v, u):
    """Wood's method on S^2: w = 1 + log(v + (1 - v) e^{-2 kappa}) / kappa, azimuth 2 pi u."""
    kappa = np.asarray(kappa, np.float64)
    w = 1.0 + np.log(v + (1.0 - v) * np.exp(-2.0 * kappa)) / kappa
    w = np.clip(w, -1.0, 1.0)
    st = np.sqrt(np.maximum(0.0, 1.0 - w * w))
    mu = _unit(mu)
    e1 = np.where(np.abs(mu[..., :1]) < 0.9,
                  np.stack([np.zeros_like(mu[..., 0]), -mu[..., 2], mu[..., 1]], -1),
                  np.stack([-mu[..., 2], np.zeros_like(mu[..., 0]), mu[..., 0]], -1))
    e1 = _unit(e1)
    e2 = np.cross(mu, e1)
    ph = 2.0 * math.pi * u
    return (w[..., None] * mu + st[..., None] * (np.cos(ph)[..., None] * e1 + np.sin(ph)[..., None] * e2))


# ---- the site energy of the reference's Hamiltonian as a function of one spin ------------------
def local_env(m, spins, i):
    """G[3], Q_B[3,3], Delta, axis[3], h-term for site i (0-based) of one configuration
    spins[N][3] of the model m (reference table layout): E_i(S) = -G.S + S^T (Q_B + Delta a a^T) S,
    the S-dependent part of the energy with every other spin fixed (get_quant.F90:264-337)."""
    G = np.zeros(3)
    QB = np.zeros((3, 3))
    for l in range(m.J.z):
        if l < m.J.cnt[i]:
            j = m.J.nbr[l, i] - 1
            G += m.J.val[l, i] * spins[j]
    if m.B is not None:
        for l in range(m.B.z):
            if l < m.B.cnt[i]:
                j = m.B.nbr[l, i] - 1
                QB += m.B.val[l, i] * np.outer(spins[j], spins[j])
    if m.D is not None:
        for l in range(m.D.z):
            if l < m.D.cnt[i]:
                j = m.D.nbr[l, i] - 1
                G -= np.cross(spins[j], m.D.val[l, i])
    b = i % m.nb
    if m.h_par is not None:
        G -= m.h_par[b] * _unit(m.h_axis[b])
    if m.si_par is not None:
        Delta, ax = float(m.si_par[b]), _unit(m.si_axis[b])
    else:
        Delta, ax = 0.0, np.array([0.0, 0.0, 1.0])
    return G, QB, Delta, ax


def site_energy(G, QB, Delta, ax, S):
    Q = QB + Delta * np.outer(ax, ax)
    return -G @ S + S @ Q @ S


def features(G, QB, Delta, ax, beta, mag):
    g = np.linalg.norm(G)
    gh = G / g if g > 1e-300 else ax
    kappa0 = beta * mag * g
    bm2 = beta * mag * mag
    x = np.array([math.log1p(kappa0), abs(gh @ ax), math.log(1.0 / beta),
                  bm2 * (gh @ QB @ gh - np.trace(QB) / 3.0), bm2 * Delta])
    return x, kappa0, gh


def record(m, spins, r, i, beta):
    """The record esp_learn_collect writes for replica r, site i (0-based), spins[N][3]."""
    G, QB, Delta, ax = local_env(m, spins, i)
    x, k0, gh = features(G, QB, Delta, ax, beta, float(m.mag[i]))
    s = spins[i] / np.linalg.norm(spins[i])
    return np.concatenate([x, [k0], gh, ax, s, [r * m.N + i]])
