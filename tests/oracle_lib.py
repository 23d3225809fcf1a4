"""ctypes access to the CPU oracle (oracle/liboracle*.so) for the tests.

The oracle is the checker: tests compare the product's results with it.
It works on CObjects buffers relocated to their own address, which is
how sixtracklib_amd.cobjects.CBuffer keeps them.
"""
import ctypes as ct
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")


class EbeConfig(ct.Structure):
    """layout of NS(ElemByElemConfig), output/elem_by_elem_config.h:39-53"""
    _fields_ = [(n, ct.c_int64) for n in (
        "order", "num_particles_to_store", "num_elements_to_store",
        "num_turns_to_store", "min_particle_id", "min_element_id",
        "min_turn", "max_particle_id", "max_element_id", "max_turn",
        "is_rolling")] + [("out_store_addr", ct.c_uint64)]


def _load(name):
    path = os.path.join(ORACLE_DIR, name)
    if not os.path.exists(path):
        import subprocess
        subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])
    lib = ct.CDLL(path)
    vp, i64, i32 = ct.c_void_p, ct.c_int64, ct.c_int32
    lib.oracle_remap.argtypes = [vp]
    lib.oracle_remap.restype = ct.c_int
    lib.oracle_track_until.argtypes = [vp, vp, i64, ct.c_int]
    lib.oracle_track_until.restype = i32
    lib.oracle_track_elem_by_elem.argtypes = [vp, vp, vp, i64]
    lib.oracle_track_elem_by_elem.restype = i32
    lib.oracle_track_line.argtypes = [vp, vp, i64, i64, ct.c_int]
    lib.oracle_track_line.restype = i32
    lib.oracle_track_element.argtypes = [vp, i64, vp, i64]
    lib.oracle_track_element.restype = i32
    lib.oracle_ebe_store_index.argtypes = [vp, i64, i64, i64]
    lib.oracle_ebe_store_index.restype = i64
    lib.oracle_monitor_store_index.argtypes = [vp, i64, i64]
    lib.oracle_monitor_store_index.restype = i64
    d = ct.c_double
    lib.oracle_cerrf.argtypes = [d, d, ct.POINTER(d), ct.POINTER(d)]
    lib.oracle_cerrf.restype = None
    lib.oracle_gauss_field.argtypes = [d, d, d, d, d, ct.c_int] + [ct.POINTER(d)] * 4
    lib.oracle_gauss_field.restype = None
    for f in ("oracle_sin", "oracle_cos", "oracle_exp"):
        getattr(lib, f).argtypes = [d]
        getattr(lib, f).restype = d
    lib.oracle_pow.argtypes = [d, d]
    lib.oracle_pow.restype = d
    return lib


_LIBS = {}


def oracle(libm=False):
    """liboracle_libm.so (reference transcendentals) if ``libm`` else
    liboracle.so (project transcendentals, the GPU's)."""
    key = bool(libm)
    if key not in _LIBS:
        _LIBS[key] = _load("liboracle_libm.so" if libm else "liboracle.so")
    return _LIBS[key]


def particles_ptr(pbuf, index=0):
    return pbuf.get_object_address(index)


def track_until(pbuf, ebuf, until_turn, libm=False, index=0, threads=1):
    return oracle(libm).oracle_track_until(particles_ptr(pbuf, index),
                                           ebuf.pointer, int(until_turn),
                                           int(threads))


def track_line(pbuf, ebuf, begin, end, finish_turn, libm=False, index=0):
    return oracle(libm).oracle_track_line(particles_ptr(pbuf, index),
                                          ebuf.pointer, int(begin), int(end),
                                          int(bool(finish_turn)))


def track_element(pbuf, ebuf, pidx, eidx, libm=False, index=0):
    return oracle(libm).oracle_track_element(particles_ptr(pbuf, index),
                                             int(pidx), ebuf.pointer, int(eidx))


def track_elem_by_elem(pbuf, ebuf, config, until_turn, libm=False, index=0):
    return oracle(libm).oracle_track_elem_by_elem(
        particles_ptr(pbuf, index), ct.addressof(config), ebuf.pointer,
        int(until_turn))


def cerrf(x, y, libm=False):
    re, im = ct.c_double(), ct.c_double()
    oracle(libm).oracle_cerrf(float(x), float(y), ct.byref(re), ct.byref(im))
    return re.value, im.value
