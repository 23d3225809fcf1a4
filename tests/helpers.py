"""Shared test helpers: particle sets from arrays, field comparison."""
import numpy as np

from sixtracklib_amd.cobjects import CBuffer
from sixtracklib_amd.particles import Particles, ALL_FIELDS, REAL_FIELDS, INDEX_FIELDS


def particles_from_arrays(arrays, cbuffer=None):
    n = len(arrays["x"])
    cb = cbuffer if cbuffer is not None else CBuffer()
    p = Particles(cbuffer=cb, num_particles=n)
    for f in ALL_FIELDS:
        if f in arrays:
            type(p).__dict__[f]._view(p)[:] = arrays[f]
    return cb, p


def particles_dict(p):
    return {f: np.array(getattr(p, f)) for f in ALL_FIELDS}


def assert_particles_equal(a, b, abs_tol=0.0, rel_tol=0.0, fields=REAL_FIELDS,
                           what="", check_index=True):
    """Real fields within |a-b| <= abs_tol + rel_tol*|b| (exact when both
    are 0); index fields exactly equal."""
    for f in fields:
        x = np.asarray(a[f]); y = np.asarray(b[f])
        if abs_tol == 0.0 and rel_tol == 0.0:
            ok = (x.view(np.uint64) == y.view(np.uint64)) | (np.isnan(x) & np.isnan(y))
            if not np.all(ok):
                i = int(np.argmin(ok))
                raise AssertionError("%s field %s differs at %d: %r vs %r (bits)"
                                     % (what, f, i, x[i], y[i]))
        else:
            d = np.abs(x - y)
            lim = abs_tol + rel_tol * np.abs(y)
            ok = (d <= lim) | (np.isnan(x) & np.isnan(y))
            if not np.all(ok):
                i = int(np.argmax(np.where(ok, 0, d - lim)))
                raise AssertionError("%s field %s differs at %d: %r vs %r (|d|=%g)"
                                     % (what, f, i, x[i], y[i], d[i]))
    for f in (INDEX_FIELDS if check_index else ()):
        x = np.asarray(a[f]); y = np.asarray(b[f])
        if not np.array_equal(x, y):
            i = int(np.argmax(x != y))
            raise AssertionError("%s index field %s differs at %d: %r vs %r"
                                 % (what, f, i, x[i], y[i]))
