"""Learned Metropolis-Hastings proposal: weight initialisation, the training loop on local
conditionals collected from the chains, and the effective-sample-size estimate used to judge
it.  Sampling, record collection, forward/backward and Adam all run in libespins_hip.so
(esp_learn_*); this module only sequences the calls.

The proposal for site i is a two-component von Mises-Fisher mixture whose concentration,
weight and tilt an MLP (5 -> 32 -> 32 -> 32 -> 6, GELU) computes from the site's environment.
Because it depends on the neighbours only, Metropolis-Hastings with the exact density ratio
keeps detailed balance for any weights; training maximises the likelihood of equilibrium
spins given their environment, which drives the proposal toward the heat-bath conditional.
"""
from __future__ import annotations

import math

import numpy as np

from . import _lib as E

NIN, NH, NOUT = 5, 32, 6


def init_weights(seed=0, out_scale=0.01):
    """He-normal hidden layers, a small output layer.  The output biases start component 0
    (weight 0.73) as vMF(ghat, beta |m| |G|) -- the exact heat-bath conditional of a pure
    exchange site -- and component 1 broader (kappa x 0.61) and tilted 0.73 rad toward the
    anisotropy axis, so that training can move the two apart."""
    rng = np.random.default_rng(seed)
    parts = [
        rng.normal(0.0, math.sqrt(2.0 / NIN), (NH, NIN)), np.zeros(NH),
        rng.normal(0.0, math.sqrt(2.0 / NH), (NH, NH)), np.zeros(NH),
        rng.normal(0.0, math.sqrt(2.0 / NH), (NH, NH)), np.zeros(NH),
        rng.normal(0.0, out_scale, (NOUT, NH)), np.array([0.0, 0.5, 0.0, -0.5, -0.5, 0.5]),
    ]
    w = np.concatenate([p.ravel() for p in parts]).astype(np.float32)
    assert w.size == E.LEARN_NPARAMS
    return w


def train_proposal(eng, n_records=1 << 20, collect_every=2, per_collect=None, epochs=4,
                   batch=2048, lr=2e-3, seed=1, weights=None, mode="random", sweep0=0,
                   lr_decay=0.5):
    """Collect n_records local conditionals from the engine's current chains (plain
    Metropolis sweeps between collections, so records are spread over the trajectory),
    then train with Adam.  Returns the per-epoch mean losses."""
    if weights is None:
        weights = init_weights(seed)
    eng.learn_set_weights(weights)
    eng.learn_clear()
    per = per_collect or max(1, min(n_records, eng.R * eng.N // 4))
    got, sw, k = 0, sweep0, 0
    while got < n_records:
        n = min(per, n_records - got)
        eng.learn_collect(n, seed=(seed << 20) + k)
        got += n
        k += 1
        eng.sweep(collect_every, mode=mode, seed=seed, sweep0=sw)
        sw += collect_every
    losses = []
    for e in range(epochs):
        losses += list(eng.learn_train(1, batch=batch, lr=lr * (lr_decay ** e), seed=seed + e))
    return np.array(losses)


def autocorr_time(x, c=6.0):
    """Integrated autocorrelation time tau_int = 1/2 + sum_t rho(t) with Sokal's automatic
    window (smallest W with W >= c tau_int(W)).  x: 1-D series."""
    x = np.asarray(x, np.float64)
    n = x.size
    if n < 4:
        return float("nan")
    y = x - x.mean()
    var = y @ y / n
    if var == 0.0:
        return 0.5
    f = np.fft.rfft(y, 2 * n)
    acf = np.fft.irfft(f * np.conj(f))[:n] / (n * var)
    tau = 0.5
    for W in range(1, n):
        tau += acf[W]
        if W >= c * tau:
            return float(max(tau, 0.5))
    return float(max(tau, 0.5))


def effective_samples(x, c=6.0):
    """n / (2 tau_int) for a 1-D series."""
    return x.size / (2.0 * autocorr_time(x, c))


def ess_per_slot(trace, c=6.0):
    """trace[n][R] -> (ESS[R], tau_int[R])."""
    tau = np.array([autocorr_time(trace[:, r], c) for r in range(trace.shape[1])])
    return trace.shape[0] / (2.0 * tau), tau
