"""The allocation-free energy entry point (esp_energy_noalloc) against esp_energy, through the
C ABI, on two calls in a row -- the second with a smaller input."""
import ctypes as C

import numpy as np
import pytest

from espins_amd import lattice as Lt

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime libespins_hip.so is linked against (device buffers for the test)."""
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    return h


def _read(h, dev, n):
    out = np.zeros(n)
    assert h.hipMemcpy(out.ctypes.data, dev, 8 * n, 2) == 0      # device -> host
    return out


def test_energy_noalloc_matches_energy_on_two_calls():
    from espins_amd.engine import Engine
    h = _hip()
    dev = C.c_void_p()
    assert h.hipMalloc(C.byref(dev), 8 * 64) == 0
    try:
        with Engine() as eng:
            # first call: FCC 4^3 (N = 256), J1-J3 + single ion, 8 replicas
            m, tems = Lt.config_c2(L=4, R=8)
            eng.set_model(m).set_replicas(tems, Lt.random_spins(m.N, 8, 3, m.mag))
            eng.sweep(3, seed=1)
            ref = eng.energy_get()
            assert h.hipMemset(dev, 0xFF, 8 * 64) == 0
            eng.energy_noalloc(dev.value)
            eng.sync()
            got = _read(h, dev.value, 8)
            print("first call", got - ref, flush=True)
            assert np.array_equal(got, ref)
            # second call, smaller input: SC 3^3 with J + B + D + field, 3 replicas
            m2 = Lt.cubic_model("sc", 3, J=(1.0,), B=0.05, D=0.3, K=0.02, field=0.4)
            eng.set_model(m2).set_replicas(np.array([0.5, 1.0, 2.0]), Lt.random_spins(m2.N, 3, 4, m2.mag))
            ref2 = eng.energy_get()
            eng.energy_noalloc(dev.value)
            eng.sync()
            got2 = _read(h, dev.value, 3)
            print("second call", got2 - ref2, flush=True)
            assert np.array_equal(got2, ref2)
    finally:
        h.hipFree(dev)
