"""ctypes wrapper of oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module,
as the checker or the CPU timing leg.  The product path never does.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} not built: run `make -C oracle`")
lib = C.CDLL(LIB_PATH)

_d, _i, _u32, _u64, _i64 = C.c_double, C.c_int, C.c_uint32, C.c_uint64, C.c_int64
_dp, _ip, _u32p, _i64p, _u8p = (C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_uint32),
                                C.POINTER(C.c_int64), C.POINTER(C.c_uint8))


class MT(C.Structure):
    _fields_ = [("mti", C.c_int32), ("mt", C.c_uint32 * 624)]


class OrcModel(C.Structure):
    _fields_ = [("N", _i), ("nb", _i), ("mag", _dp),
                ("zJ", _i), ("j_cnt", _ip), ("j_nbr", _ip), ("j_val", _dp),
                ("zB", _i), ("b_cnt", _ip), ("b_nbr", _ip), ("b_val", _dp),
                ("zD", _i), ("d_cnt", _ip), ("d_nbr", _ip), ("d_vec", _dp),
                ("have_si", _i), ("si_par", _dp), ("si_axis", _dp),
                ("have_field", _i), ("h_par", _dp), ("h_axis", _dp)]


_mp = C.POINTER(OrcModel)
_mtp = C.POINTER(MT)
for name, res, args in [
    ("orc_mt_init", None, [_mtp, C.c_int32]),
    ("orc_mt_u32", _u32, [_mtp]),
    ("orc_mt_real2", _d, [_mtp]),
    ("orc_mt_fill_u32", None, [C.c_int32, _i64, _u32p]),
    ("orc_philox4x32_10", None, [_u32p, _u32p, _u32p]),
    ("orc_energy", _d, [_mp, _dp]),
    ("orc_energy_sites", None, [_mp, _dp, _dp]),
    ("orc_delta_energy", _d, [_mp, _i, _dp, _dp]),
    ("orc_magnetization2", _d, [_mp, _dp]),
    ("orc_propose", None, [_i, _d, _d, _dp, _d, _d, _dp]),
    ("orc_sweep_ref", None, [_mp, _dp, _d, _i, _d, _mtp, _i64, _dp, _i64p]),
    ("orc_sweep_colour_philox", None, [_mp, _i, _ip, _ip, _dp, _d, _i, _d, _u64, _u64, _u32, _dp, _i64p]),
    ("orc_sweep_colour_hostfed", None, [_mp, _i, _ip, _ip, _dp, _d, _dp, _dp, _u8p, _dp, _i64p]),
    ("orc_pt_swap", None, [_i, _i, _dp, _dp, _ip, _i64p, _mtp, _u64, _u64, _u32]),
    ("orc_accumulate", None, [_d, _d, _dp]),
    ("orc_finalize", None, [_dp, _i64, _d, _i, _dp]),
    ("orc_build_table", _i, [_i, _dp, _dp, _ip, _i, _i, _dp, _dp, _dp, _dp, _i, _i, _ip, _ip, _dp]),
    ("orc_check_colouring", _i, [_mp, _ip]),
]:
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args


def _p(a, t):
    if a is None:
        return C.cast(None, C.POINTER(t))
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(t))


class Oracle:
    """Holds an OrcModel built from an espins_amd.lattice.Model (or any object with the
    same fields)."""

    def __init__(self, m):
        f = lambda a: None if a is None else np.ascontiguousarray(a, np.float64)
        g = lambda a: None if a is None else np.ascontiguousarray(a, np.int32)
        self.N, self.nb = m.N, m.nb
        self.keep = []

        def tab(t):
            if t is None or t.z == 0:
                return 0, None, None, None
            c, n, v = g(t.cnt), g(t.nbr), f(t.val)
            self.keep += [c, n, v]
            return t.z, c, n, v

        mag = f(m.mag)
        self.keep.append(mag)
        zJ, jc, jn, jv = tab(m.J)
        zB, bc, bn, bv = tab(m.B)
        zD, dc, dn, dv = tab(m.D)
        sip, sia, hp, ha = f(m.si_par), f(m.si_axis), f(m.h_par), f(m.h_axis)
        self.keep += [sip, sia, hp, ha]
        self.m = OrcModel(m.N, m.nb, _p(mag, _d), zJ, _p(jc, _i), _p(jn, _i), _p(jv, _d),
                          zB, _p(bc, _i), _p(bn, _i), _p(bv, _d),
                          zD, _p(dc, _i), _p(dn, _i), _p(dv, _d),
                          int(sip is not None), _p(sip, _d), _p(sia, _d),
                          int(hp is not None), _p(hp, _d), _p(ha, _d))

    def energy(self, s):
        s = np.ascontiguousarray(s, np.float64)
        return lib.orc_energy(C.byref(self.m), _p(s, _d))

    def energy_exact(self, s):
        """energy_get's terms summed exactly (math.fsum over the per-site groups): the correctly
        rounded energy, against which long sequential sums are judged."""
        import math
        s = np.ascontiguousarray(s, np.float64)
        e = np.zeros(self.N)
        lib.orc_energy_sites(C.byref(self.m), _p(s, _d), _p(e, _d))
        return math.fsum(e.tolist()) / 2.0

    def delta_energy(self, i0, s_new, ds):
        s_new = np.ascontiguousarray(s_new, np.float64)
        ds = np.ascontiguousarray(ds, np.float64)
        return lib.orc_delta_energy(C.byref(self.m), int(i0), _p(s_new, _d), _p(ds, _d))

    def magnetization2(self, s):
        s = np.ascontiguousarray(s, np.float64)
        return lib.orc_magnetization2(C.byref(self.m), _p(s, _d))

    def sweep_ref(self, s, beta, mode, tilt, mt, nsweeps, e_run):
        e = C.c_double(e_run)
        rej = C.c_int64(0)
        lib.orc_sweep_ref(C.byref(self.m), _p(s, _d), beta, mode, tilt, C.byref(mt), nsweeps,
                          C.byref(e), C.byref(rej))
        return e.value, rej.value

    def sweep_colour_philox(self, s, col_off, col_sites, beta, mode, tilt, seed, sweep, stream,
                            e_run):
        e = C.c_double(e_run)
        rej = C.c_int64(0)
        co = np.ascontiguousarray(col_off, np.int32)
        cs = np.ascontiguousarray(col_sites, np.int32)
        lib.orc_sweep_colour_philox(C.byref(self.m), len(co) - 1, _p(co, _i), _p(cs, _i),
                                    _p(s, _d), beta, mode, tilt, seed, sweep, stream,
                                    C.byref(e), C.byref(rej))
        return e.value, rej.value

    def sweep_colour_hostfed(self, s, col_off, col_sites, beta, prop, unif, e_run):
        e = C.c_double(e_run)
        rej = C.c_int64(0)
        co = np.ascontiguousarray(col_off, np.int32)
        cs = np.ascontiguousarray(col_sites, np.int32)
        prop = np.ascontiguousarray(prop, np.float64)
        unif = np.ascontiguousarray(unif, np.float64)
        acc = np.zeros(self.N, np.uint8)
        lib.orc_sweep_colour_hostfed(C.byref(self.m), len(co) - 1, _p(co, _i), _p(cs, _i),
                                     _p(s, _d), beta, _p(prop, _d), _p(unif, _d), _p(acc, C.c_uint8),
                                     C.byref(e), C.byref(rej))
        return acc, e.value, rej.value

    def check_colouring(self, colour):
        c = np.ascontiguousarray(colour, np.int32)
        return bool(lib.orc_check_colouring(C.byref(self.m), _p(c, _i)))


def mt_stream(seed, n):
    out = np.zeros(n, np.uint32)
    lib.orc_mt_fill_u32(seed, n, _p(out, C.c_uint32))
    return out


def mt_state(seed):
    st = MT()
    lib.orc_mt_init(C.byref(st), seed)
    return st


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    o = np.zeros(4, np.uint32)
    lib.orc_philox4x32_10(_p(c, C.c_uint32), _p(k, C.c_uint32), _p(o, C.c_uint32))
    return o


def propose(mode, tilt, mag, old, u1, u2):
    old = np.ascontiguousarray(old, np.float64)
    out = np.zeros(3)
    lib.orc_propose(mode, tilt, mag, _p(old, _d), u1, u2, _p(out, _d))
    return out


def pt_swap(start, tems, energy, slot_rep, nswap, mt=None, seed=0, step=0, stream=0):
    tems = np.ascontiguousarray(tems, np.float64)
    energy = np.ascontiguousarray(energy, np.float64)
    assert slot_rep.dtype == np.int32 and nswap.dtype == np.int64
    lib.orc_pt_swap(start, len(tems), _p(tems, _d), _p(energy, _d), _p(slot_rep, _i),
                    _p(nswap, C.c_int64), C.byref(mt) if mt is not None else None, seed, step, stream)


def finalize(acc, nsteps, T, N):
    acc = np.ascontiguousarray(acc, np.float64)
    out = np.zeros(7)
    lib.orc_finalize(_p(acc, _d), nsteps, T, N, _p(out, _d))
    return out


def build_table_fn(nb, bf, bm, Larr, open_boundary, f1, f2, v, dv, ncomp, zcap, cnt, nbr, val):
    """Drop-in `builder` for espins_amd.lattice.build_table using the oracle's restatement."""
    z = lib.orc_build_table(nb, _p(bf, _d), _p(bm, _d), _p(Larr, _i), int(open_boundary), len(v),
                            _p(f1, _d), _p(f2, _d), _p(v, _d), _p(dv, _d), ncomp, zcap,
                            _p(cnt, _i), _p(nbr, _i), _p(val, _d))
    if z < 0:
        raise ValueError(f"orc_build_table failed: {z}")
    return z
