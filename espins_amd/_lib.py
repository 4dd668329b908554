"""ctypes binding of ``libespins_hip.so`` (declarations mirror include/espins_hip.h).

The library is the product path: importing this module raises if the built library is
missing, and every call checks the returned status and raises ``EspError`` with the
library's own message.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("ESPINS_HIP_LIB", os.path.join(_HERE, "libespins_hip.so"))

ESP_OK, ESP_EINVAL, ESP_EHIP, ESP_ESTATE, ESP_ENOMEM, ESP_ETABLE = 0, -1, -2, -3, -4, -5
MODE_RANDOM, MODE_CONST, MODE_LEARNED = 0, 1, 2
LEARN_NPARAMS, LEARN_RECORD = 2502, 16


class EspError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"espins_hip error {code}: {msg}")
        self.code = code


if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is not built: run `python -m espins_amd.build` (the HIP engine has no "
        "CPU fallback)")

lib = C.CDLL(LIB_PATH)

_p = C.c_void_p
_i = C.c_int
_d = C.c_double
_u64 = C.c_uint64
_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_u8p = C.POINTER(C.c_uint8)

_SIGS = {
    "esp_version": (_i, []),
    "esp_create": (_i, [C.POINTER(_p), _i]),
    "esp_destroy": (_i, [_p]),
    "esp_last_error": (C.c_char_p, [_p]),
    "esp_build_table": (_i, [_i, _dp, _dp, _ip, _i, _i, _dp, _dp, _dp, _dp, _i, _i, _ip, _ip, _dp, _ip]),
    "esp_set_model": (_i, [_p, _i, _i, _dp,
                           _i, _ip, _ip, _dp,
                           _i, _ip, _ip, _dp,
                           _i, _ip, _ip, _dp,
                           _i, _dp, _dp, _i, _dp, _dp]),
    "esp_set_supercell": (_i, [_p, _ip]),
    "esp_set_path": (_i, [_p, _i]),
    "esp_set_stream_base": (_i, [_p, C.c_uint32]),
    "esp_set_tps": (_i, [_p, _i]),
    "esp_set_indexing": (_i, [_p, _i]),
    "esp_plan_model": (_i, [_i, _i, _dp, _i, _ip, _ip, _dp, _i, _ip, _ip, _dp, _i, _ip, _ip, _dp,
                            _ip, _ip, _ip, _ip]),
    "esp_get_coloring": (_i, [_p, _ip, _ip]),
    "esp_get_model_info": (_i, [_p, _ip]),
    "esp_set_replicas": (_i, [_p, _i, _dp, _dp]),
    "esp_get_spins": (_i, [_p, _i, _dp]),
    "esp_put_spins": (_i, [_p, _dp]),
    "esp_energy": (_i, [_p, _dp]),
    "esp_energy_noalloc": (_i, [_p, C.c_void_p]),
    "esp_sweep": (_i, [_p, _i, _i, _d, _u64, _u64]),
    "esp_sweep_hostfed": (_i, [_p, _dp, _dp, _u8p]),
    "esp_measure": (_i, [_p]),
    "esp_pt_exchange": (_i, [_p, _i, _u64, _u64]),
    "esp_run": (_i, [_p, _i, _i, _i, _i, _i, _i, _i, _d, _u64, _u64, _u64, _i]),
    "esp_time_sweeps": (_i, [_p, _i, _i, _d, _u64, _u64, _dp, _ip, _dp, _i64p]),
    "esp_get_stats": (_i, [_p, _dp, _i64p, _dp, _i64p]),
    "esp_get_map": (_i, [_p, _ip, _i64p]),
    "esp_reset_stats": (_i, [_p]),
    "esp_set_trace": (_i, [_p, _i]),
    "esp_get_trace": (_i, [_p, _dp, _ip]),
    "esp_sync": (_i, [_p]),
    "esp_stream": (_p, [_p]),
    "esp_copy_probe": (_i, [_p, C.c_int64, _i, _dp]),
    "esp_neighbor_shells": (_i, [_i, _dp, _dp, _ip, _d, _i, _i, _dp, _ip, _ip]),
    "esp_spin_glass": (_i, [_i, _i, _ip, _ip, _dp, _dp, _i, _ip]),
    "esp_set_orderparam": (_i, [_p, _dp]),
    "esp_set_staggered": (_i, [_p, _i, _dp]),
    "esp_set_pm": (_i, [_p, _i, _ip]),
    "esp_set_spincorr": (_i, [_p, _i, _i, _ip, _i, _ip, _ip]),
    "esp_set_sfactor": (_i, [_p, _i, _dp, _dp, _dp, _i]),
    "esp_get_obs": (_i, [_p, _i, _dp, _i64p]),
    "esp_get_obs_trace": (_i, [_p, _dp, _ip]),
}

OBS_ORDER, OBS_STAGGERED, OBS_PM, OBS_SPINCORR, OBS_SFACTOR = 1, 2, 4, 8, 16

_fp = C.POINTER(C.c_float)
_SIGS.update({
    "esp_learn_set_weights": (_i, [_p, _fp, _i]),
    "esp_learn_get_weights": (_i, [_p, _fp, _i]),
    "esp_learn_collect": (_i, [_p, _i, _u64]),
    "esp_learn_clear": (_i, [_p]),
    "esp_learn_num_records": (_i, [_p, _ip]),
    "esp_learn_get_records": (_i, [_p, _i, _i, _fp]),
    "esp_learn_put_records": (_i, [_p, _fp, _i]),
    "esp_learn_grad": (_i, [_p, _i, _i, _fp, _dp]),
    "esp_learn_train": (_i, [_p, _i, _i, C.c_float, _u64, _dp]),
    "esp_learn_eval": (_i, [_p, _i, _i, _i, _dp]),
})

_SIGS.update({
    "esp_batch_create": (_i, [C.POINTER(_p), _i]),
    "esp_batch_destroy": (_i, [_p]),
    "esp_batch_last_error": (C.c_char_p, [_p]),
    "esp_batch_add_chain": (_i, [_p, _i, _i, _dp, _i, _ip, _ip, _dp, _i, _ip, _ip, _dp,
                                 _i, _ip, _ip, _dp, _i, _dp, _dp, _i, _dp, _dp,
                                 _ip, _i, _dp, _dp, _ip]),
    "esp_batch_set_chain_config": (_i, [_p, _i, _i, _i]),
    "esp_batch_set_chain_chunk": (_i, [_p, _i, _i]),
    "esp_batch_get_chain_chunk": (_i, [_p, _i, _ip]),
    "esp_batch_commit": (_i, [_p]),
    "esp_batch_autotune": (_i, [_p, _i, C.c_char_p]),
    "esp_batch_run": (_i, [_p, _i, _i, _i, _i, _i, _i, _i, _d, _u64, _u64, _u64, _i]),
    "esp_batch_num_chains": (_i, [_p]),
    "esp_batch_info": (_i, [_p, _i, _ip]),
    "esp_batch_time_sweeps": (_i, [_p, _i, _i, _u64, _u64, _dp, _ip, _dp, _i64p]),
    "esp_batch_get_stats": (_i, [_p, _i, _dp, _i64p, _dp, _i64p]),
    "esp_batch_get_map": (_i, [_p, _i, _ip, _i64p]),
    "esp_batch_get_spins": (_i, [_p, _i, _i, _dp]),
    "esp_batch_energy": (_i, [_p, _i, _dp]),
    "esp_batch_reset_stats": (_i, [_p]),
    "esp_batch_sync": (_i, [_p]),
})

OPTIONAL = set()


def _declare(sigs, optional=False):
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name, None)
        if fn is None:
            if optional:
                OPTIONAL.add(name)
                continue
            raise ImportError(f"{LIB_PATH} does not export {name}")
        fn.restype = res
        fn.argtypes = args


_declare(_SIGS)

EXPORTS = sorted(_SIGS)


def ptr(a, ctype):
    """Pointer to a contiguous numpy array (or NULL for None)."""
    if a is None:
        return C.cast(None, C.POINTER(ctype))
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data_as(C.POINTER(ctype))


def check(rc, ctx=None):
    if rc != ESP_OK:
        msg = lib.esp_last_error(ctx).decode() if ctx else ""
        raise EspError(rc, msg)
    return rc


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)
