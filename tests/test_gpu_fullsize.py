"""The benchmark's batched workload at its full size: 256 independent PT chains (SC/BCC/FCC,
L 8..48, 16-32 replicas; 472,136,384 spin updates per sweep, 11.3 GB of spins), built exactly
as bench.py builds it.

* One sweep of the largest chain (FCC 48^3, N = 442,368) equals the oracle's colour-order
  Philox sweep (the restatement of metropolis, get_quant.F90:52-182, in the colour order the
  library plans) bit for bit in the accept decisions and to 1e-12 in the spins, for the
  coldest and the hottest replica.
* After measured steps with replica exchanges, every chain's running energies (the per-wave
  accepted-dE sums) equal its direct energies (energy_get, get_quant.F90:186-261)."""
import numpy as np
import pytest

from espins_amd import lattice as Lt
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def test_full_batch_sweep_and_running_energies():
    from espins_amd.engine import Batch, plan_model
    chains = Lt.config_c3(n_chains=256, seed=7)
    models = []
    b = Batch(0)
    try:
        for i, (kind, L, J2, R, tems) in enumerate(chains):
            m = Lt.cubic_model_fast(kind, L, J=(1.0, J2))
            b.add_chain(m, tems, Lt.random_spins(m.N, R, 1000 + i))
            models.append(m)
        for k in range(len(chains)):
            b.set_chain_config(k, 1, 1)            # global-memory colour phases, as benchmarked
        b.commit()
        kbig = max(range(len(chains)), key=lambda k: models[k].N * len(chains[k][4]))
        m, tems = models[kbig], chains[kbig][4]
        say("full batch committed; largest chain", kbig, "N", m.N, "R", len(tems))
        s0 = Lt.random_spins(m.N, len(tems), 1000 + kbig)
        seed = 0x5EED5EED
        b.run(1, measure=False, exchange=False, seed=seed, graph=True)    # one sweep of every chain
        got = b.spins(kbig)
        rej = b.stats(kbig)["rejected"]
        ncol, col, _ = plan_model(m)
        sites = np.argsort(col, kind="stable").astype(np.int32)
        off = np.concatenate([[0], np.cumsum(np.bincount(col, minlength=ncol))]).astype(np.int32)
        o = O.Oracle(m)
        for r in (0, len(tems) - 1):
            s = s0[r].copy()
            e = o.energy(s)
            e, nr = o.sweep_colour_philox(s, off, sites, 1.0 / tems[r], 0, 0.125, seed, 0,
                                          (kbig << 16) + r, e)
            say("replica", r, "rejected", rej[r], nr)
            assert rej[r] == nr
            np.testing.assert_allclose(got[r], s, rtol=0, atol=1e-12)
        b.reset_stats()
        b.run(2, measure=True, exchange=True, seed=seed, sweep0=1, step0=1, graph=True)
        worst = 0.0
        for k in range(len(chains)):
            e_run = b.stats(k)["e_run"]
            e_dir = b.energy_get(k)
            np.testing.assert_allclose(e_run, e_dir, rtol=1e-10, atol=1e-8)
            worst = max(worst, float(np.max(np.abs(e_run - e_dir) / np.maximum(1.0, np.abs(e_dir)))))
        say("running vs direct energies of all chains: max relative difference", worst)
    finally:
        b.close()
