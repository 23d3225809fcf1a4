"""The reference's LHC lattices rebuilt from the element tables in
tests/golden/ (written by tests/golden/make_fixtures.py from
lhc_no_bb/beam_elements.bin and beambeam/beam_elements.bin of the
reference's tests/testdata).

rebuild(table) makes a CObjects buffer whose objects equal the file's,
field for field; beam(n) the particles the golden results were computed
for.
"""
import os

import numpy as np

from sixtracklib_amd import beam_elements as be
from sixtracklib_amd import lattices
from sixtracklib_amd.cobjects import CBuffer

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# object type ids (internal/objects_type_id.h:20-41) -> element classes
_CLASSES = {2: be.Drift, 3: be.DriftExact, 4: be.Multipole, 5: be.Cavity, 6: be.XYShift,
            7: be.SRotation, 8: be.BeamBeam4D, 9: be.BeamBeam6D}


def load(name):
    """name: "lhc_no_bb" or "beambeam"."""
    with np.load(os.path.join(GOLDEN, name + "_lattice.npz")) as z:
        return {k: z[k] for k in z.files}


def rebuild(table):
    types = np.asarray(table["types"])
    n = len(types)
    eb = CBuffer(max_slots=len(types) * 8 + sum(v.size for v in table.values()) + 4096,
                 max_objects=n + 16, max_pointers=n + 16)
    pos = {}          # per type: next row
    off = {}          # per (type, field): next offset in a concatenated column
    for t in types:
        t = int(t)
        cls = _CLASSES[t]
        row = pos.get(t, 0)
        pos[t] = row + 1
        kw = {}
        prefix = "t%d_" % t
        for key, col in table.items():
            if not key.startswith(prefix) or key.endswith("_len"):
                continue
            field = key[len(prefix):]
            lens = table.get(key + "_len")
            if lens is None:
                kw[field] = col[row]
            else:
                o = off.get(key, 0)
                kw[field] = col[o:o + int(lens[row])]
                off[key] = o + int(lens[row])
        if t == 4:
            kw = {"bal": kw["bal"], "length": kw["length"], "hxl": kw["hxl"], "hyl": kw["hyl"]}
        elif t == 7:
            kw = {"cos_z": kw["cos_z"], "sin_z": kw["sin_z"]}
        cls(cbuffer=eb, **kw)
    return eb


def beam(n, seed=31):
    """n protons at 6.5 TeV with small transverse and longitudinal offsets
    (x, y ~ N(0, 1e-4) m, px, py ~ N(0, 1e-6), zeta ~ N(0, 1e-2) m,
    delta ~ N(0, 1e-5)), drawn from numpy default_rng(seed)."""
    pb, _ = lattices.gaussian_beam(n, seed=seed, sigma_x=1e-4, sigma_px=1e-6, sigma_y=1e-4,
                                   sigma_py=1e-6, sigma_zeta=1e-2, sigma_delta=1e-5)
    return pb
