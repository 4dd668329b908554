"""Python mirror of the reference's hot-path interface, on the HIP engine.

Names follow the reference routines they replace:

==========================  =====================================================
reference (file:line)        here
==========================  =====================================================
energy_get (get_quant:186)   ``Engine.energy_get()``
metropolis (get_quant:52)    ``Engine.sweep()`` -- a whole sweep, colour order
magnetization_get (:340)     ``Engine.measure()`` (with the estimators, mcarlo:536)
pt_swap (pt.F90:1235)        ``Engine.pt_swap()``
mcarlo (mcarlo.F90:463)      ``mcarlo()``
pt_mcarlo (pt.F90:504)       ``pt_mcarlo()``
==========================  =====================================================

Every call goes through libespins_hip.so; errors raise ``EspError``.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math

import numpy as np

from . import _lib as E
from .lattice import Model

MODES = {"random": E.MODE_RANDOM, "rand": E.MODE_RANDOM, "const": E.MODE_CONST,
         "learned": E.MODE_LEARNED}


def _mode(m):
    if isinstance(m, str):
        for k, v in MODES.items():
            if m.lower().startswith(k):
                return v
        raise ValueError(f"mcarlo_mode not recognised: {m}")
    return int(m)


class Engine:
    """One engine context on one GPU."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        rc = E.lib.esp_create(C.byref(h), device)
        if rc != E.ESP_OK:
            raise E.EspError(rc, f"esp_create(device={device}) failed (is a GPU visible?)")
        self._h = h
        self.N = 0
        self.R = 0

    def close(self):
        if getattr(self, "_h", None):
            E.lib.esp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc):
        return E.check(rc, self._h)

    # ---- model ---------------------------------------------------------------------
    def set_model(self, m: Model):
        N = m.N
        mag = E.f64(m.mag)

        def tab(t, ncomp):
            if t is None or t.z == 0:
                return 0, None, None, None
            return t.z, E.i32(t.cnt), E.i32(t.nbr), E.f64(t.val)

        zJ, jc, jn, jv = tab(m.J, 1)
        zB, bc, bn, bv = tab(m.B, 1)
        zD, dc, dn, dv = tab(m.D, 3)
        sip = E.f64(m.si_par) if m.si_par is not None else None
        sia = E.f64(m.si_axis) if m.si_axis is not None else None
        hp = E.f64(m.h_par) if m.h_par is not None else None
        ha = E.f64(m.h_axis) if m.h_axis is not None else None
        d, i = C.c_double, C.c_int
        Ls = np.array(m.L, np.int32) if getattr(m, "L", None) is not None else None
        self._chk(E.lib.esp_set_supercell(self._h, E.ptr(Ls, i)))
        self._keep = (mag, jc, jn, jv, bc, bn, bv, dc, dn, dv, sip, sia, hp, ha)
        self._chk(E.lib.esp_set_model(
            self._h, N, m.nb, E.ptr(mag, d),
            zJ, E.ptr(jc, i), E.ptr(jn, i), E.ptr(jv, d),
            zB, E.ptr(bc, i), E.ptr(bn, i), E.ptr(bv, d),
            zD, E.ptr(dc, i), E.ptr(dn, i), E.ptr(dv, d),
            int(sip is not None), E.ptr(sip, d), E.ptr(sia, d),
            int(hp is not None), E.ptr(hp, d), E.ptr(ha, d)))
        self.N = N
        self.R = 0
        return self

    def set_path(self, path):
        """0 auto, 1 global-memory colour phases, 2 LDS-resident."""
        self._chk(E.lib.esp_set_path(self._h, int(path)))

    def set_tps(self, tps):
        """Lanes per site in the sweep kernels (1, 2, 4, 8; 0 = automatic)."""
        self._chk(E.lib.esp_set_tps(self._h, int(tps)))

    def set_indexing(self, mode):
        """0: partner indices computed from the lattice geometry when it allows; 1: read from the
        per-site neighbour tables."""
        self._chk(E.lib.esp_set_indexing(self._h, int(mode)))

    def coloring(self):
        n = C.c_int()
        col = np.zeros(self.N, np.int32)
        self._chk(E.lib.esp_get_coloring(self._h, C.byref(n), E.ptr(col, C.c_int)))
        return n.value, col

    def model_info(self):
        out = np.zeros(6, np.int32)
        self._chk(E.lib.esp_get_model_info(self._h, E.ptr(out, C.c_int)))
        return dict(N=int(out[0]), zmax=int(out[1]), ncol=int(out[2]), compressed=bool(out[3]),
                    layout=("bond", "basis", "implicit")[int(out[3])], lds=bool(out[4]),
                    tps=int(out[5]))

    # ---- replicas --------------------------------------------------------------------
    def set_replicas(self, tems, spins):
        tems = E.f64(tems)
        spins = E.f64(spins)
        R = len(tems)
        assert spins.shape == (R, self.N, 3), spins.shape
        self._chk(E.lib.esp_set_replicas(self._h, R, E.ptr(tems, C.c_double),
                                         E.ptr(spins, C.c_double)))
        self.R = R
        self.tems = tems.copy()
        return self

    def spins(self, by_slot=False):
        out = np.zeros((self.R, self.N, 3))
        self._chk(E.lib.esp_get_spins(self._h, int(by_slot), E.ptr(out, C.c_double)))
        return out

    def put_spins(self, spins):
        spins = E.f64(spins)
        self._chk(E.lib.esp_put_spins(self._h, E.ptr(spins, C.c_double)))

    # ---- hot path --------------------------------------------------------------------
    def energy_get(self):
        out = np.zeros(self.R)
        self._chk(E.lib.esp_energy(self._h, E.ptr(out, C.c_double)))
        return out

    def energy_noalloc(self, dev_ptr):
        """Enqueue the direct energies into device memory at dev_ptr (R doubles) without any
        device allocation (esp_energy_noalloc); call sync() before reading them."""
        self._chk(E.lib.esp_energy_noalloc(self._h, C.c_void_p(int(dev_ptr))))

    def sweep(self, nsweeps=1, mode="random", tilt=0.125, seed=1, sweep0=0):
        self._chk(E.lib.esp_sweep(self._h, int(nsweeps), _mode(mode), float(tilt),
                                  int(seed) & (2**64 - 1), int(sweep0)))

    def sweep_hostfed(self, prop, unif):
        prop = E.f64(prop)
        unif = E.f64(unif)
        acc = np.zeros((self.R, self.N), np.uint8)
        self._chk(E.lib.esp_sweep_hostfed(self._h, E.ptr(prop, C.c_double),
                                          E.ptr(unif, C.c_double), E.ptr(acc, C.c_uint8)))
        return acc

    def measure(self):
        self._chk(E.lib.esp_measure(self._h))

    def pt_swap(self, start, seed, step):
        self._chk(E.lib.esp_pt_exchange(self._h, int(start), int(seed) & (2**64 - 1), int(step)))

    def run(self, n_rounds, steps_per_round=1, steps_measure=1, measure=True, exchange=True,
            start_parity=1, mode="random", tilt=0.125, seed=1, sweep0=0, step0=0, graph=True):
        self._chk(E.lib.esp_run(self._h, int(n_rounds), int(steps_per_round), int(steps_measure),
                                int(measure), int(exchange), int(start_parity), _mode(mode),
                                float(tilt), int(seed) & (2**64 - 1), int(sweep0), int(step0),
                                int(graph)))

    # ---- optional observables (measured at every measurement, by slot) ---------------
    def set_orderparam(self, axes):
        """axes[nb][3] (orderparam_axes_cart) or None to switch off."""
        a = None if axes is None else E.f64(axes)
        self._chk(E.lib.esp_set_orderparam(self._h, E.ptr(a, C.c_double)))

    def set_staggered(self, coeff):
        """coeff[nb] (staggered_m_coeff) or None to switch off."""
        c = None if coeff is None else E.f64(coeff)
        self._chk(E.lib.esp_set_staggered(self._h, int(c is not None), E.ptr(c, C.c_double)))

    def set_pm(self, enable=True):
        """P(m) histograms of the cell moments; returns the grid size nfreq."""
        n = C.c_int(0)
        self._chk(E.lib.esp_set_pm(self._h, int(enable), C.byref(n)))
        return n.value

    def set_spincorr(self, species, cnt, nbr):
        """species[N] (0-based kinds), cnt[nshell][N], nbr[nshell][zs][N] (1-based ids), or
        species=None to switch off."""
        if species is None:
            self._chk(E.lib.esp_set_spincorr(self._h, 0, 0, None, 0, None, None))
            return
        sp, ct, nb = E.i32(species), E.i32(cnt), E.i32(nbr)
        nshell, zs = nb.shape[0], nb.shape[1]
        self._chk(E.lib.esp_set_spincorr(self._h, nshell, int(sp.max()) + 1, E.ptr(sp, C.c_int), zs,
                                         E.ptr(ct, C.c_int), E.ptr(nb, C.c_int)))

    def set_sfactor(self, q, pos, polar=None, every=1):
        """q[nq][3] and pos[N][3] (Cartesian), polar[3] or None; q=None switches off."""
        if q is None:
            self._chk(E.lib.esp_set_sfactor(self._h, 0, None, None, None, 1))
            return
        qv, pv = E.f64(q), E.f64(pos)
        pl = None if polar is None else E.f64(polar)
        self._chk(E.lib.esp_set_sfactor(self._h, len(qv), E.ptr(qv, C.c_double), E.ptr(pv, C.c_double),
                                        E.ptr(pl, C.c_double), int(every)))

    def obs(self, which, shape):
        """Accumulator of one observable (see esp_get_obs) as an array of `shape` (leading
        dimension R, by slot) and its measurement count."""
        out = np.zeros(shape)
        n = C.c_int64(0)
        self._chk(E.lib.esp_get_obs(self._h, int(which), E.ptr(out, C.c_double), C.byref(n)))
        return out, n.value

    def obs_trace(self):
        n = C.c_int(0)
        self._chk(E.lib.esp_get_obs_trace(self._h, None, C.byref(n)))
        out = np.zeros((n.value, self.R, 2))
        self._chk(E.lib.esp_get_obs_trace(self._h, E.ptr(out, C.c_double), C.byref(n)))
        return out

    def time_sweeps(self, nsweeps, mode="random", tilt=0.125, seed=1, sweep0=0):
        """HIP-event timing of the sweep kernels (see esp_time_sweeps)."""
        kms = C.c_double()
        nl = C.c_int()
        wall = C.c_double()
        upl = C.c_int64()
        self._chk(E.lib.esp_time_sweeps(self._h, int(nsweeps), _mode(mode), float(tilt),
                                        int(seed) & (2**64 - 1), int(sweep0), C.byref(kms),
                                        C.byref(nl), C.byref(wall), C.byref(upl)))
        return dict(kernel_ms=kms.value, nlaunch=nl.value, wall_ms=wall.value,
                    updates_per_launch=upl.value)

    def stats(self):
        acc = np.zeros((self.R, 6))
        rej = np.zeros(self.R, np.int64)
        erun = np.zeros(self.R)
        nm = np.zeros(1, np.int64)
        self._chk(E.lib.esp_get_stats(self._h, E.ptr(acc, C.c_double), E.ptr(rej, C.c_int64),
                                      E.ptr(erun, C.c_double), E.ptr(nm, C.c_int64)))
        return dict(acc=acc, rejected=rej, e_run=erun, nmeas=int(nm[0]))

    def pt_map(self):
        sr = np.zeros(self.R, np.int32)
        ns = np.zeros((self.R, 2), np.int64)
        self._chk(E.lib.esp_get_map(self._h, E.ptr(sr, C.c_int), E.ptr(ns, C.c_int64)))
        return sr, ns

    def reset_stats(self):
        self._chk(E.lib.esp_reset_stats(self._h))

    def set_trace(self, capacity):
        """Record E and |m| per slot at every measurement (0 = off)."""
        self._chk(E.lib.esp_set_trace(self._h, int(capacity)))
        return self

    def trace(self):
        """[n][R][2] by slot: E and |m| of each measurement since the last reset_stats."""
        n = C.c_int()
        self._chk(E.lib.esp_get_trace(self._h, None, C.byref(n)))
        out = np.zeros((n.value, self.R, 2))
        if n.value:
            self._chk(E.lib.esp_get_trace(self._h, E.ptr(out, C.c_double), C.byref(n)))
        return out

    def sync(self):
        self._chk(E.lib.esp_sync(self._h))

    def stream(self):
        return E.lib.esp_stream(self._h)

    def copy_probe(self, n_bytes, reps=10):
        """Mean ms per launch of a plain n_bytes device copy (counter calibration)."""
        ms = C.c_double()
        self._chk(E.lib.esp_copy_probe(self._h, int(n_bytes), int(reps), C.byref(ms)))
        return ms.value

    # ---- learned proposal -------------------------------------------------------------
    def learn_set_weights(self, w):
        w = E.f32(w)
        self._chk(E.lib.esp_learn_set_weights(self._h, E.ptr(w, C.c_float), w.size))
        return self

    def learn_weights(self):
        w = np.zeros(E.LEARN_NPARAMS, np.float32)
        self._chk(E.lib.esp_learn_get_weights(self._h, E.ptr(w, C.c_float), w.size))
        return w

    def learn_collect(self, n, seed=1):
        self._chk(E.lib.esp_learn_collect(self._h, int(n), int(seed)))

    def learn_clear(self):
        self._chk(E.lib.esp_learn_clear(self._h))

    def learn_num_records(self):
        n = C.c_int()
        self._chk(E.lib.esp_learn_num_records(self._h, C.byref(n)))
        return n.value

    def learn_records(self, first=0, n=None):
        n = self.learn_num_records() - first if n is None else n
        out = np.zeros((n, E.LEARN_RECORD), np.float32)
        self._chk(E.lib.esp_learn_get_records(self._h, int(first), int(n), E.ptr(out, C.c_float)))
        return out

    def learn_put_records(self, recs):
        recs = E.f32(recs).reshape(-1, E.LEARN_RECORD)
        self._chk(E.lib.esp_learn_put_records(self._h, E.ptr(recs, C.c_float), recs.shape[0]))

    def learn_grad(self, first=0, n=None):
        n = self.learn_num_records() - first if n is None else n
        g = np.zeros(E.LEARN_NPARAMS, np.float32)
        loss = C.c_double()
        self._chk(E.lib.esp_learn_grad(self._h, int(first), int(n), E.ptr(g, C.c_float), C.byref(loss)))
        return g, loss.value

    def learn_train(self, epochs=1, batch=1024, lr=1e-3, seed=1):
        out = np.zeros(max(epochs, 1))
        self._chk(E.lib.esp_learn_train(self._h, int(epochs), int(batch), float(lr), int(seed),
                                        E.ptr(out, C.c_double)))
        return out[:epochs]

    def learn_eval(self, first=0, n=None, sweep_net=False):
        n = self.learn_num_records() - first if n is None else n
        out = np.zeros((n, 10))
        self._chk(E.lib.esp_learn_eval(self._h, int(first), int(n), int(bool(sweep_net)),
                                       E.ptr(out, C.c_double)))
        return out


# ---- estimators --------------------------------------------------------------------------
def finalize(acc, nsteps, tems, N):
    """Final formulas of mcarlo.F90:713-733 / pt.F90:995-1011 per temperature:
    columns <|m|>, <E>/N, C, chi, U_E, U_M, Delta_E, Delta_M."""
    acc = np.asarray(acc, float) / float(nsteps)
    E_, E2, E4, M, M2, M4 = acc.T
    beta = 1.0 / np.asarray(tems, float)
    C_ = beta * beta * (E2 - E_ ** 2) / N
    chi = beta * (M2 - M ** 2) * N
    UE = 1.0 - (1.0 / 3.0) * (E4 / E2 ** 2)
    UM = 1.0 - (1.0 / 3.0) * (M4 / M2 ** 2)
    dE = np.sqrt(np.maximum(E2 - E_ ** 2, 0) / nsteps) / N
    dM = np.sqrt(np.maximum(M2 - M ** 2, 0) / nsteps)
    return np.stack([M, E_ / N, C_, chi, UE, UM, dE, dM], axis=1)


@dataclasses.dataclass
class PTSchedule:
    """pt_mcarlo's schedule rewrite (pt.F90:529-544)."""
    warmup1: int
    n_swap_warmup: int
    n_swap: int
    steps_mc: int

    @staticmethod
    def of(steps_warmup, steps_mc, pt_steps_swap):
        w = math.ceil(steps_warmup / 2.0)
        nsw = math.ceil(w / pt_steps_swap)
        w = nsw * pt_steps_swap
        ns = math.ceil(steps_mc / pt_steps_swap)
        return PTSchedule(w, nsw, ns, ns * pt_steps_swap)


def pt_mcarlo(eng: Engine, steps_warmup, steps_mc, steps_measure=1, pt_steps_swap=1,
              mode="random", tilt=0.125, seed=1, graph=True):
    """Warmup1 (plain), warmup2 (swaps), main phase (swaps + measurements), as
    pt.F90:557-908.  Returns (per-slot observables, schedule, stats, map)."""
    sch = PTSchedule.of(steps_warmup, steps_mc, pt_steps_swap)
    sweep = 0
    # warmup1: plain sweeps
    if sch.warmup1:
        eng.run(1, steps_per_round=sch.warmup1, steps_measure=steps_measure, measure=False,
                exchange=False, mode=mode, tilt=tilt, seed=seed, sweep0=sweep, graph=False)
        sweep += sch.warmup1 * steps_measure
    step = 0
    if sch.n_swap_warmup:
        eng.run(sch.n_swap_warmup, steps_per_round=pt_steps_swap, steps_measure=steps_measure,
                measure=False, exchange=True, start_parity=1, mode=mode, tilt=tilt, seed=seed,
                sweep0=sweep, step0=step, graph=graph)
        sweep += sch.n_swap_warmup * pt_steps_swap * steps_measure
        step += sch.n_swap_warmup
    eng.reset_stats()
    eng.run(sch.n_swap, steps_per_round=pt_steps_swap, steps_measure=steps_measure, measure=True,
            exchange=True, start_parity=1, mode=mode, tilt=tilt, seed=seed, sweep0=sweep,
            step0=step, graph=graph)
    st = eng.stats()
    obs = finalize(st["acc"], st["nmeas"], eng.tems, eng.N)
    return obs, sch, st, eng.pt_map()


def mcarlo(eng: Engine, steps_warmup, steps_mc, steps_measure=1, mode="random", tilt=0.125,
           seed=1, graph=True):
    """Plain Metropolis at every temperature concurrently (mcarlo.F90:490-633, each
    temperature an independent chain): warmup, then steps_mc measured MC steps."""
    if steps_warmup:
        eng.run(1, steps_per_round=steps_warmup, steps_measure=steps_measure, measure=False,
                exchange=False, mode=mode, tilt=tilt, seed=seed, sweep0=0, graph=False)
    st0 = eng.stats()
    eng.reset_stats()
    eng.run(1, steps_per_round=steps_mc, steps_measure=steps_measure, measure=True,
            exchange=False, mode=mode, tilt=tilt, seed=seed, sweep0=steps_warmup * steps_measure,
            graph=False)
    st = eng.stats()
    obs = finalize(st["acc"], st["nmeas"], eng.tems, eng.N)
    return obs, st0, st


def plan_model(m: Model):
    """Host-only: (ncol, colour[N], info) that esp_set_model would use for model m."""
    d, i = C.c_double, C.c_int

    def tab(t):
        if t is None or t.z == 0:
            return 0, None, None, None
        return t.z, E.i32(t.cnt), E.i32(t.nbr), E.f64(t.val)

    mag = E.f64(m.mag)
    zJ, jc, jn, jv = tab(m.J)
    zB, bc, bn, bv = tab(m.B)
    zD, dc, dn, dv = tab(m.D)
    Ls = np.array(m.L, np.int32) if getattr(m, "L", None) is not None else None
    n = C.c_int()
    col = np.zeros(m.N, np.int32)
    info = np.zeros(5, np.int32)
    E.check(E.lib.esp_plan_model(m.N, m.nb, E.ptr(mag, d), zJ, E.ptr(jc, i), E.ptr(jn, i),
                                 E.ptr(jv, d), zB, E.ptr(bc, i), E.ptr(bn, i), E.ptr(bv, d),
                                 zD, E.ptr(dc, i), E.ptr(dn, i), E.ptr(dv, d), E.ptr(Ls, i),
                                 C.byref(n), E.ptr(col, i), E.ptr(info, i)))
    return n.value, col, dict(N=int(info[0]), zmax=int(info[1]), ncol=int(info[2]),
                              compressed=bool(info[3]),
                              layout=("bond", "basis", "implicit")[int(info[3])])


class Batch:
    """Many independent PT chains advanced together (esp_batch_* in include/espins_hip.h)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        rc = E.lib.esp_batch_create(C.byref(h), device)
        if rc != E.ESP_OK:
            raise E.EspError(rc, f"esp_batch_create(device={device}) failed (is a GPU visible?)")
        self._h = h
        self.chains = []          # (N, R, tems)

    def close(self):
        if getattr(self, "_h", None):
            E.lib.esp_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc):
        if rc != E.ESP_OK:
            raise E.EspError(rc, E.lib.esp_batch_last_error(self._h).decode())

    def add_chain(self, m: Model, tems, spins):
        d, i = C.c_double, C.c_int

        def tab(t):
            if t is None or t.z == 0:
                return 0, None, None, None
            return t.z, E.i32(t.cnt), E.i32(t.nbr), E.f64(t.val)

        mag = E.f64(m.mag)
        zJ, jc, jn, jv = tab(m.J)
        zB, bc, bn, bv = tab(m.B)
        zD, dc, dn, dv = tab(m.D)
        sip = E.f64(m.si_par) if m.si_par is not None else None
        sia = E.f64(m.si_axis) if m.si_axis is not None else None
        hp = E.f64(m.h_par) if m.h_par is not None else None
        ha = E.f64(m.h_axis) if m.h_axis is not None else None
        Ls = np.array(m.L, np.int32) if getattr(m, "L", None) is not None else None
        tems = E.f64(tems)
        spins = E.f64(spins)
        assert spins.shape == (len(tems), m.N, 3)
        cid = C.c_int()
        self._chk(E.lib.esp_batch_add_chain(
            self._h, m.N, m.nb, E.ptr(mag, d),
            zJ, E.ptr(jc, i), E.ptr(jn, i), E.ptr(jv, d),
            zB, E.ptr(bc, i), E.ptr(bn, i), E.ptr(bv, d),
            zD, E.ptr(dc, i), E.ptr(dn, i), E.ptr(dv, d),
            int(sip is not None), E.ptr(sip, d), E.ptr(sia, d),
            int(hp is not None), E.ptr(hp, d), E.ptr(ha, d),
            E.ptr(Ls, i), len(tems), E.ptr(tems, d), E.ptr(spins, d), C.byref(cid)))
        self.chains.append((m.N, len(tems), tems.copy()))
        return cid.value

    def set_chain_config(self, k, path, tps):
        self._chk(E.lib.esp_batch_set_chain_config(self._h, int(k), int(path), int(tps)))

    def commit(self):
        self._chk(E.lib.esp_batch_commit(self._h))
        return self

    def autotune(self, nsweeps=4, cache=None):
        self._chk(E.lib.esp_batch_autotune(self._h, int(nsweeps),
                                           cache.encode() if cache else None))

    def run(self, n_rounds, steps_per_round=1, steps_measure=1, measure=True, exchange=True,
            start_parity=1, mode="random", tilt=0.125, seed=1, sweep0=0, step0=0, graph=True):
        self._chk(E.lib.esp_batch_run(self._h, int(n_rounds), int(steps_per_round),
                                      int(steps_measure), int(measure), int(exchange),
                                      int(start_parity), _mode(mode), float(tilt),
                                      int(seed) & (2**64 - 1), int(sweep0), int(step0), int(graph)))

    def time_sweeps(self, nsweeps, mode="random", seed=1, sweep0=0):
        """HIP-event timing of the batched colour phases (see esp_batch_time_sweeps)."""
        kms, wall = C.c_double(), C.c_double()
        nl, upd = C.c_int(), C.c_int64()
        self._chk(E.lib.esp_batch_time_sweeps(self._h, int(nsweeps), _mode(mode), int(seed) & (2**64 - 1),
                                              int(sweep0), C.byref(kms), C.byref(nl), C.byref(wall),
                                              C.byref(upd)))
        return dict(kernel_ms=kms.value, nlaunch=nl.value, wall_ms=wall.value, updates=upd.value)

    def set_chain_chunk(self, k, chunk):
        """Partner gathers in flight per lane on chain k's global-memory colour phases (2 or 4)."""
        self._chk(E.lib.esp_batch_set_chain_chunk(self._h, int(k), int(chunk)))

    def info(self, k):
        out = np.zeros(8, np.int32)
        self._chk(E.lib.esp_batch_info(self._h, int(k), E.ptr(out, C.c_int)))
        ch = C.c_int()
        self._chk(E.lib.esp_batch_get_chain_chunk(self._h, int(k), C.byref(ch)))
        return dict(N=int(out[0]), R=int(out[1]), ncol=int(out[2]), tps=int(out[3]),
                    lds=bool(out[4]), zmax=int(out[5]), variant=int(out[6]), lds_fits=bool(out[7]),
                    chunk=int(ch.value))

    def stats(self, k):
        N, R, _ = self.chains[k]
        acc = np.zeros((R, 6))
        rej = np.zeros(R, np.int64)
        erun = np.zeros(R)
        nm = np.zeros(1, np.int64)
        self._chk(E.lib.esp_batch_get_stats(self._h, int(k), E.ptr(acc, C.c_double),
                                            E.ptr(rej, C.c_int64), E.ptr(erun, C.c_double),
                                            E.ptr(nm, C.c_int64)))
        return dict(acc=acc, rejected=rej, e_run=erun, nmeas=int(nm[0]))

    def pt_map(self, k):
        N, R, _ = self.chains[k]
        sr = np.zeros(R, np.int32)
        ns = np.zeros((R, 2), np.int64)
        self._chk(E.lib.esp_batch_get_map(self._h, int(k), E.ptr(sr, C.c_int), E.ptr(ns, C.c_int64)))
        return sr, ns

    def spins(self, k, by_slot=False):
        N, R, _ = self.chains[k]
        out = np.zeros((R, N, 3))
        self._chk(E.lib.esp_batch_get_spins(self._h, int(k), int(by_slot), E.ptr(out, C.c_double)))
        return out

    def energy_get(self, k):
        N, R, _ = self.chains[k]
        out = np.zeros(R)
        self._chk(E.lib.esp_batch_energy(self._h, int(k), E.ptr(out, C.c_double)))
        return out

    def reset_stats(self):
        self._chk(E.lib.esp_batch_reset_stats(self._h))

    def sync(self):
        self._chk(E.lib.esp_batch_sync(self._h))


def set_stream_base(eng: Engine, base: int):
    eng._chk(E.lib.esp_set_stream_base(eng._h, int(base) & 0xFFFFFFFF))
