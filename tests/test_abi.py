"""The C-ABI library loads without a GPU and exports every symbol include/espins_hip.h declares."""
import os
import re

import numpy as np

from tests.conftest import ROOT


def _declared():
    txt = open(os.path.join(ROOT, "include", "espins_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(esp_[a-z_0-9]+)\s*\(", txt)))


def test_every_declared_symbol_is_exported_and_bound():
    from espins_amd import _lib as E
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(E.lib, n), f"libespins_hip.so does not export {n}"
    # the Python binding declares every one of them with a prototype
    assert set(names) <= set(E.EXPORTS), set(names) - set(E.EXPORTS)


def test_version_and_error_paths_without_device():
    from espins_amd import _lib as E
    import ctypes as C
    assert E.lib.esp_version() >= 100
    # null output pointer -> EINVAL, never a crash
    assert E.lib.esp_create(None, 0) == E.ESP_EINVAL
    assert E.lib.esp_destroy(None) == E.ESP_EINVAL
    assert E.lib.esp_last_error(None) == b"null context"


def test_build_table_rejects_bad_arguments():
    from espins_amd import _lib as E
    import ctypes as C
    L = np.array([0, 2, 2], np.int32)
    bf = np.zeros(3)
    bm = np.ones(1)
    cnt = np.zeros(8, np.int32)
    nbr = np.zeros(8, np.int32)
    val = np.zeros(8)
    z = np.zeros(1, np.int32)
    rc = E.lib.esp_build_table(1, E.ptr(bf, C.c_double), E.ptr(bm, C.c_double), E.ptr(L, C.c_int),
                               0, 0, None, None, None, None, 1, 1, E.ptr(cnt, C.c_int),
                               E.ptr(nbr, C.c_int), E.ptr(val, C.c_double), E.ptr(z, C.c_int))
    assert rc == E.ESP_EINVAL
    # unknown atomic position in a row -> ETABLE (jij.F90:236-243 is fatal in the reference)
    L = np.array([2, 2, 2], np.int32)
    f1 = np.array([[0.3, 0.0, 0.0]])
    f2 = np.array([[1.0, 0.0, 0.0]])
    v = np.array([1.0])
    rc = E.lib.esp_build_table(1, E.ptr(bf, C.c_double), E.ptr(bm, C.c_double), E.ptr(L, C.c_int),
                               0, 1, E.ptr(f1, C.c_double), E.ptr(f2, C.c_double), E.ptr(v, C.c_double),
                               None, 1, 1, E.ptr(cnt, C.c_int), E.ptr(nbr, C.c_int),
                               E.ptr(val, C.c_double), E.ptr(z, C.c_int))
    assert rc == E.ESP_ETABLE
