"""Extract compact golden vectors from the reference's own test data.

Run once, here (the reference tree does not exist where the GPU tests
run):  python tests/golden/make_fixtures.py [/path/to/reference]

Sources (all in the reference's tests/testdata/):
* test_track_be_drift_data.bin / test_track_be_driftexact_data.bin --
  [initial Particles, 1000 Drift(Exact), final Particles] used by
  tests/sixtracklib/common/test_track_c99.cpp:298-357 (1 turn).  A seeded
  subset of particles is kept (all 1000 elements).
* bbsimple/{beam_elements,particles_dump}.bin -- element-by-element dump of
  a 15-element lattice with BB6D, BB4D, Drift, Multipole, Cavity used by
  test_track_c99.cpp:647-665 (1e-15).  Element parameters are stored per
  element, the dump as field arrays.
* test_faddeeva_errorfn_data.bin -- u64 n, then (x, y, Re w, Im w) on a
  100 x 100 grid (the cerrf known-answer grid); a seeded subset is kept.
* lhc_no_bb/beam_elements.bin (18,657 elements: Drift, Multipole of orders
  0-14, XYShift, SRotation, Cavity) and beambeam/beam_elements.bin (10,866
  elements with 72 BeamBeam4D and 2 BeamBeam6D) -- the lattices of the
  reference's LHC examples and benchmarks (tests/benchmark/common/
  benchmark_lhc_no_bb_c99.c) and of its GPU-vs-CPU test
  (tests/sixtracklib/cuda/track/test_track_job_track_until_c99.cpp).  Kept
  as element tables (the type sequence and, per type, its fields as
  arrays; pointer fields concatenated with their lengths), from which
  tests/reference_lattices.py rebuilds a CObjects buffer whose objects
  equal the file's field for field (checked here when the tables are
  written).  Beside each table: the oracle's result for 64 generated
  particles tracked 2 turns through the lattice read from the file itself
  (libm build, the reference's arithmetic).

Outputs are .npz files next to this script.  Nothing here runs at test
time.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from sixtracklib_amd.cobjects import CBuffer          # noqa: E402
from sixtracklib_amd.cobject import get_object        # noqa: E402
from sixtracklib_amd.particles import ALL_FIELDS      # noqa: E402
import sixtracklib_amd.beam_elements  # noqa: E402,F401  (registers classes)


def particles_arrays(p, sel=None):
    out = {}
    for f in ALL_FIELDS:
        a = np.array(getattr(p, f))
        out[f] = a if sel is None else a[sel]
    return out


def drift_fixture(ref, name, out, n_keep=128):
    cb = CBuffer.fromfile(os.path.join(ref, "tests/testdata", name))
    n = cb.n_objects
    init = get_object(cb, 0)
    final = get_object(cb, n - 1)
    lengths = np.array([get_object(cb, i).length for i in range(1, n - 1)])
    types = cb.typeids()[1:n - 1]
    rng = np.random.default_rng(12345)
    sel = np.sort(rng.choice(init.num_particles, size=n_keep, replace=False))
    d = {"lengths": lengths, "types": types}
    for k, v in particles_arrays(init, sel).items():
        d["init_" + k] = v
    for k, v in particles_arrays(final, sel).items():
        d["final_" + k] = v
    np.savez_compressed(os.path.join(HERE, out), **d)


def bbsimple_fixture(ref, out):
    be = CBuffer.fromfile(os.path.join(ref, "tests/testdata/bbsimple/beam_elements.bin"))
    pd = CBuffer.fromfile(os.path.join(ref, "tests/testdata/bbsimple/particles_dump.bin"))
    d = {"types": be.typeids()}
    for i in range(be.n_objects):
        o = get_object(be, i)
        for k, v in o.to_dict().items():
            d["e%d_%s" % (i, k)] = np.array(v)
    for j in range(pd.n_objects):
        for k, v in particles_arrays(get_object(pd, j)).items():
            d["p%d_%s" % (j, k)] = v
    d["n_dumps"] = np.array(pd.n_objects)
    np.savez_compressed(os.path.join(HERE, out), **d)


def faddeeva_fixture(ref, out, n_keep=2000):
    raw = np.fromfile(os.path.join(ref, "tests/testdata/test_faddeeva_errorfn_data.bin"),
                      dtype=np.uint64)
    n = int(raw[0])
    vals = raw[1:1 + 4 * n].view(np.float64).reshape(n, 4)
    rng = np.random.default_rng(2024)
    sel = np.sort(rng.choice(n, size=n_keep, replace=False))
    np.savez_compressed(os.path.join(HERE, out), xyw=vals[sel])


def element_table(cb):
    """Type sequence + per-type field arrays of a beam-element buffer."""
    types = cb.typeids().astype(np.uint64)
    cols = {}
    for i in range(cb.n_objects):
        t = int(types[i])
        for k, v in get_object(cb, i).to_dict().items():
            a = np.asarray(v)
            key = "t%d_%s" % (t, k)
            if a.ndim == 0:
                cols.setdefault(key, []).append(a[()])
            else:
                cols.setdefault(key, []).append(a.ravel())
                cols.setdefault(key + "_len", []).append(a.size)
    d = {"types": types}
    for key, vals in cols.items():
        if isinstance(vals[0], np.ndarray):
            d[key] = np.concatenate(vals)
        else:
            d[key] = np.array(vals)
    return d


def lattice_fixture(ref, rel, out, n_particles=64, turns=2):
    sys.path.insert(0, os.path.dirname(HERE))
    import oracle_lib
    import reference_lattices
    from sixtracklib_amd import lattices
    cb = CBuffer.fromfile(os.path.join(ref, "tests/testdata", rel))
    d = element_table(cb)
    rebuilt = reference_lattices.rebuild(d)
    assert rebuilt.n_objects == cb.n_objects
    for i in range(cb.n_objects):
        a, b = get_object(cb, i).to_dict(), get_object(rebuilt, i).to_dict()
        assert a.keys() == b.keys(), i
        for k in a:
            x, y = np.asarray(a[k]), np.asarray(b[k])
            assert x.shape == y.shape and x.tobytes() == y.tobytes(), (i, k)
    pb = reference_lattices.beam(n_particles)
    init = particles_arrays(get_object(pb, 0))
    assert oracle_lib.track_until(pb, cb, turns, libm=True) == 0
    final = particles_arrays(get_object(pb, 0))
    for k, v in init.items():
        d["init_" + k] = v
    for k, v in final.items():
        d["final_" + k] = v
    d["turns"] = np.array(turns)
    np.savez_compressed(os.path.join(HERE, out), **d)


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    drift_fixture(ref, "test_track_be_drift_data.bin", "drift_golden.npz")
    drift_fixture(ref, "test_track_be_driftexact_data.bin", "driftexact_golden.npz")
    bbsimple_fixture(ref, "bbsimple_golden.npz")
    faddeeva_fixture(ref, "faddeeva_golden.npz")
    lattice_fixture(ref, "lhc_no_bb/beam_elements.bin", "lhc_no_bb_lattice.npz")
    lattice_fixture(ref, "beambeam/beam_elements.bin", "beambeam_lattice.npz")
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
