"""C-ABI library checks that need no GPU: the library loads, every symbol
declared in include/sixtracklib_amd/*.h is exported, the host-side buffer
and element builders produce reference-format buffers, the lattice packer
flattens every element type, and the track job refuses to run without a
HIP device (no CPU fallback exists)."""
import ctypes as ct
import os
import re

import numpy as np
import pytest

from sixtracklib_amd.cobjects import CBuffer
from sixtracklib_amd.cobject import get_object
from sixtracklib_amd import lattices, beam_elements as be

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include", "sixtracklib_amd")


def _declared_symbols():
    names = set()
    for fn in os.listdir(INCLUDE):
        if not fn.endswith(".h"):
            continue
        txt = open(os.path.join(INCLUDE, fn)).read()
        for m in re.finditer(r"ST_API[^;{]*?NS\((\w+)\)\s*\(", txt, re.S):
            names.add("st_" + m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    from sixtracklib_amd import stcommon as st
    names = _declared_symbols()
    assert len(names) > 80
    missing = [n for n in sorted(names) if not hasattr(st.sixtracklib, n)]
    assert not missing, missing


def test_python_bindings_cover_every_declared_function():
    """stcommon binds every function the headers declare (the adjoint's in
    adjoint.py), each with as many argument types as the C parameters."""
    from sixtracklib_amd import stcommon as st
    from sixtracklib_amd import adjoint  # noqa: F401  (binds the st_Adjoint_* entry points)
    missing, wrong = [], []
    for fn in sorted(os.listdir(INCLUDE)):
        if not fn.endswith(".h"):
            continue
        txt = re.sub(r"/\*.*?\*/", "", open(os.path.join(INCLUDE, fn)).read(), flags=re.S)
        txt = re.sub(r"#define ST_CUDA_TRACK_JOB_ARG_GETTERS.*?#undef ST_CUDA_TRACK_JOB_ARG_GETTERS",
                     "", txt, flags=re.S)
        for m in re.finditer(r"ST_API[^;{]*?NS\((\w+)\)\s*\((.*?)\)\s*;", txt, re.S):
            name, args = "st_" + m.group(1), " ".join(m.group(2).split())
            n_args = 0 if args in ("", "void") else args.count(",") + 1
            f = getattr(st, name, None) if fn != "adjoint.h" else getattr(st.sixtracklib, name)
            if f is None:
                missing.append(name)
            elif f.argtypes is None or len(f.argtypes) != n_args:
                wrong.append((name, n_args, None if f.argtypes is None else len(f.argtypes)))
    assert not missing, missing
    assert not wrong, wrong
    for w in ("particles", "beam_elements", "output", "elem_by_elem_config", "debug_register",
              "particles_addr"):
        for form in ("has_%s_arg", "get_ptr_%s_arg", "get_ptr_const_%s_arg"):
            assert hasattr(st, "st_CudaTrackJob_" + form % w)


def test_c_buffer_roundtrip_with_python_buffer(tmp_path):
    from sixtracklib_amd import stcommon as st
    b = st.st_Buffer_new(0)
    for k in range(50):
        assert st.st_Drift_add(b, 0.5 + k)
    assert st.st_Multipole_add(b, 2, 0.0, 0.0, 0.0, 0)
    assert st.st_Cavity_add(b, 1e6, 4e8, 180.0)
    assert st.st_LimitRect_add(b, -0.1, 0.1, -0.2, 0.2)
    assert st.st_Particles_new(b, 17)
    path = str(tmp_path / "c.bin")
    assert st.st_Buffer_write_to_file(b, path.encode())
    st.st_Buffer_delete(b)
    cb = CBuffer.fromfile(path)
    assert cb.n_objects == 54
    assert [get_object(cb, k).length for k in range(50)] == [0.5 + k for k in range(50)]
    mp = get_object(cb, 50)
    assert mp.order == 2 and np.all(mp.bal == 0.0) and len(mp.bal) == 6
    assert get_object(cb, 51).frequency == 4e8
    assert get_object(cb, 53).num_particles == 17
    # Python-written buffer loads in C unchanged
    eb = lattices.kitchen_sink(0)
    p2 = str(tmp_path / "py.bin")
    eb.tofile(p2)
    b2 = st.st_Buffer_new_from_file(p2.encode())
    assert st.st_Buffer_get_num_of_objects(b2) == eb.n_objects
    st.st_Buffer_delete(b2)


def test_mapped_buffer_is_relocated_in_place():
    from sixtracklib_amd import stcommon as st
    eb = lattices.fodo_ring(n_cells=3)
    raw = CBuffer.frombytes(eb.tobytes())
    # relocate the image to another base (as if written by another
    # process), then map it: the mapping must relocate it back in place
    assert st.st_ManagedBuffer_remap_to(raw.pointer, raw.pointer + 4096) == 0
    assert int(raw.words[0]) == raw.pointer + 4096
    b = st.st_Buffer_new_on_memory(raw.pointer, raw.words.nbytes)
    assert b and not st.st_Buffer_needs_remapping(b)
    assert st.st_Buffer_get_num_of_objects(b) == eb.n_objects
    st.st_Buffer_delete(b)


def test_lattice_pack_all_element_types():
    from sixtracklib_amd import stcommon as st
    eb = lattices.kitchen_sink(0, with_monitor=True)
    n_el, n_mon = ct.c_int64(0), ct.c_int64(0)
    nw = st.st_Lattice_pack(eb.pointer, None, 0, None, ct.byref(n_el), None, ct.byref(n_mon))
    assert nw > 0 and n_el.value == eb.n_objects and n_mon.value == 1
    out = np.zeros(nw, dtype=np.uint64)
    offs = np.zeros(n_el.value + 1, dtype=np.uint32)
    assert st.st_Lattice_pack(eb.pointer, out.ctypes.data, nw, offs.ctypes.data,
                              ct.byref(n_el), None, ct.byref(n_mon)) == nw
    ops = [int(out[o]) & 0xff for o in offs[:-1]]
    nbytes = [int(out[o]) >> 32 for o in offs[:-1]]
    # records are contiguous, lengths in bytes, the element index is word 1
    assert all(np.diff(offs) * 8 == nbytes)
    assert [int(out[o + 1]) for o in offs[:-1]] == list(range(n_el.value))
    # end record closes the stream
    assert offs[-1] == nw - 2 and int(out[nw - 2]) & 0xff == 0xff
    assert int(out[nw - 2]) >> 32 == 16 and int(out[nw - 1]) == n_el.value
    # opcode per type id (lattice.h); Drift and Multipole are the only ones
    # without the 0x80 bit
    want = {2: 0x01, 3: 0x81, 4: 0x02, 5: 0x82, 6: 0x83, 7: 0x84, 11: 0x85, 12: 0x86,
            16: 0x87, 24: 0x88, 10: 0x89, 8: 0x8a, 9: 0x8b, 13: 0x8c, 14: 0x8d,
            257: 0x8e, 256: 0x8f}
    assert ops == [want[int(t)] for t in eb.typeids()]
    # cavity folding: (pi/180)*lag and ((2 pi)/c)*f
    k = list(eb.typeids()).index(5)
    cav = out[offs[k] + 2: offs[k] + 5].view(np.float64)
    assert cav[0] == 5e6
    assert cav[1] == (np.pi / 180.0) * 170.0
    assert cav[2] == ((2.0 * np.pi) / 299792458.0) * 200e6


def _sc_interpol_lattice(table_cb):
    eb = CBuffer()
    be.Drift(cbuffer=eb, length=1.0)
    be.SCInterpolatedProfile(cbuffer=eb, number_of_particles=1e11, sigma_x=1e-3,
                             sigma_y=1.2e-3, length=1.0, line_density_prof_fallback=0.5,
                             interpol_data_addr=table_cb.get_object_address(0))
    return eb


def _pack(eb, tables=()):
    from sixtracklib_amd import stcommon as st
    arr = (ct.c_void_p * max(1, len(tables)))(*[t.pointer for t in tables])
    n_el = ct.c_int64(0)
    nw = st.st_Lattice_pack_with_tables(eb.pointer, arr, len(tables), None, 0, None,
                                        ct.byref(n_el), None, None)
    assert nw > 0
    out = np.zeros(nw, dtype=np.uint64)
    offs = np.zeros(n_el.value + 1, dtype=np.uint32)
    assert st.st_Lattice_pack_with_tables(eb.pointer, arr, len(tables), out.ctypes.data, nw,
                                          offs.ctypes.data, ct.byref(n_el), None, None) == nw
    return out, offs


def test_lattice_pack_copies_tables_from_table_buffers():
    """An SCInterpolatedProfile whose table lies in another buffer: without
    that buffer the record selects the fallback density (method 255); with
    it the record carries method, size, z0, dz and both arrays."""
    tb = CBuffer()
    vals = np.linspace(1.0, 2.0, 9)
    be.LineDensityProfileData(cbuffer=tb, method=0, capacity=9, num_values=9, z0=-0.4, dz=0.1,
                              values=vals, derivatives=np.full(9, 10.0))
    eb = _sc_interpol_lattice(tb)
    out, offs = _pack(eb)
    rec = out[offs[1]: offs[2]]
    assert int(rec[2 + 10]) == 255 and int(rec[2 + 11]) == 0
    out, offs = _pack(eb, [tb])
    rec = out[offs[1]: offs[2]]
    assert int(rec[2 + 10]) == 0 and int(rec[2 + 11]) == 9
    assert rec[2 + 12: 2 + 14].view(np.float64).tolist() == [-0.4, 0.1]
    assert np.array_equal(rec[2 + 14: 2 + 23].view(np.float64), vals)
    assert np.all(rec[2 + 23: 2 + 32].view(np.float64) == 10.0)
    # a missing table-buffer list is an error
    from sixtracklib_amd import stcommon as st
    assert st.st_Lattice_pack_with_tables(eb.pointer, None, 1, None, 0, None, None, None,
                                          None) == -1


def test_line_density_interpolation_preparation():
    """Linear: forward differences; cubic: second derivatives of the natural
    spline (scipy CubicSpline bc_type='natural' as the independent check);
    interpol() evaluates both forms and is 0 outside the table."""
    from scipy.interpolate import CubicSpline
    z = -1.0 + 0.05 * np.arange(41)
    y = np.exp(-z ** 2) * (1.0 + 0.3 * z)
    cb = CBuffer()
    lin = be.LineDensityProfileData(cbuffer=cb, method="linear", capacity=41, z0=-1.0, dz=0.05,
                                    values=y)
    cub = be.LineDensityProfileData(cbuffer=cb, method="cubic", capacity=41, z0=-1.0, dz=0.05,
                                    values=y)
    lin.prepare_interpolation()
    cub.prepare_interpolation()
    assert np.array_equal(np.array(lin.derivatives)[:40], (y[1:] - y[:-1]) / 0.05)
    cs = CubicSpline(z, y, bc_type="natural")
    np.testing.assert_allclose(cub.derivatives, cs(z, 2), rtol=0, atol=1e-12)
    for zz in (-0.93, -0.2, 0.0, 0.4441, 0.97):
        assert abs(cub.interpol(zz) - cs(zz)) < 1e-14
        i = int((zz + 1.0) / 0.05)
        assert lin.interpol(zz) == y[i] + (zz - (-1.0 + 0.05 * i)) * lin.derivatives[i]
    assert lin.interpol(-1.5) == 0.0 and cub.interpol(1.01) == 0.0
    with pytest.raises(RuntimeError):
        be.LineDensityProfileData(cbuffer=cb, method="cubic", capacity=3, values=[1, 2, 3]
                                  ).prepare_interpolation()


def test_lattice_pack_rejects_unrelocated_buffer():
    from sixtracklib_amd import stcommon as st
    eb = lattices.fodo_ring(n_cells=2)
    raw = CBuffer.frombytes(eb.tobytes())
    assert st.st_ManagedBuffer_remap_to(raw.pointer, raw.pointer + 8) == 0
    assert st.st_Lattice_pack(raw.pointer, None, 0, None, None, None, None) == -1


def test_trackjob_rejects_cpu_arch():
    from sixtracklib_amd.trackjob import TrackJob
    eb = lattices.fodo_ring(n_cells=2)
    pb, _ = lattices.gaussian_beam(8)
    with pytest.raises(ValueError):
        TrackJob(eb, pb, arch="cpu")


def test_trackjob_required_output_sets():
    from sixtracklib_amd import stcommon as st
    eb = lattices.fodo_ring(n_cells=2)
    be.BeamMonitor(cbuffer=eb, num_stores=4, start=0, skip=1)
    pb, _ = lattices.gaussian_beam(10)
    cp = st.st_Buffer_new_on_memory(pb.pointer, pb.words.nbytes)
    ce = st.st_Buffer_new_on_memory(eb.pointer, eb.words.nbytes)
    sizes = (ct.c_int64 * 8)()
    has = ct.c_int64(0)
    n = st.st_TrackJob_required_output_sets(cp, 0, None, ce, 3, sizes, 8, ct.byref(has), None)
    assert n == 2 and has.value == 1
    assert sizes[0] == 10 * eb.n_objects * 3
    assert sizes[1] == 4 * 10
    # a store of 5 element passes per particle; a larger request keeps the full size
    n = st.st_TrackJob_required_output_sets(cp, 0, None, ce, 3, sizes, 8, ct.byref(has),
                                            b"fma=1 ebe_store_passes=5")
    assert n == 2 and sizes[0] == 10 * 5 and sizes[1] == 4 * 10
    n = st.st_TrackJob_required_output_sets(cp, 0, None, ce, 3, sizes, 8, ct.byref(has),
                                            b"ebe_store_passes=100000")
    assert sizes[0] == 10 * eb.n_objects * 3
    st.st_Buffer_delete(cp)
    st.st_Buffer_delete(ce)


def test_trackjob_without_device_raises():
    from sixtracklib_amd import stcommon as st
    if st.st_Hip_num_devices() > 0:
        pytest.skip("a device is present")
    from sixtracklib_amd.trackjob import TrackJob
    eb = lattices.fodo_ring(n_cells=2)
    pb, _ = lattices.gaussian_beam(8)
    with pytest.raises(RuntimeError):
        TrackJob(eb, pb)


def test_trackbatch_argument_and_device_errors():
    from sixtracklib_amd import stcommon as st
    from sixtracklib_amd.trackbatch import TrackBatch
    with pytest.raises(ValueError):
        TrackBatch([], [])
    eb = lattices.fodo_ring(n_cells=2)
    pb, _ = lattices.gaussian_beam(10)
    if st.st_Hip_num_devices() > 0:
        pytest.skip("a device is present: construction would succeed")
    with pytest.raises(RuntimeError):
        TrackBatch([eb], [pb])


def test_arch_and_device_id_strings():
    """track_job.h:123-135: the device id of a configuration string and the
    lower-case arch string (the reference's helpers)."""
    import ctypes as ct
    from sixtracklib_amd import stcommon as st
    buf = ct.create_string_buffer(16)
    assert st.st_TrackJob_extract_device_id_str(b"fma=1 device=3", buf, 16) == 0
    assert buf.value == b"3"
    assert st.st_TrackJob_extract_device_id_str(b"0.1;fma=0", buf, 16) == 0
    assert buf.value == b"0.1"
    assert st.st_TrackJob_extract_device_id_str(b"fma=1", buf, 16) == 0 and buf.value == b""
    assert st.st_TrackJob_extract_device_id_str(b"device=12345", buf, 3) == -1
    assert st.st_TrackJob_sanitize_arch_str(b" HiP ", buf, 16) == 0 and buf.value == b"hip"
    inplace = ct.create_string_buffer(b"  CUDA", 16)
    assert st.st_TrackJob_sanitize_arch_str_inplace(inplace, 16) == 0 and inplace.value == b"cuda"


def test_assign_items_search_removal_and_pairs():
    """Address-assignment bookkeeping (track_job.h:500-660) on a job made
    without buffers (no device needed): items of one (dest, src) pair stay
    next to each other in the order added; index, lookup by index, removal
    and the distinct pairs."""
    import ctypes as ct
    from sixtracklib_amd import stcommon as st
    job = st.st_TrackJob_create(b"hip", b"")
    assert job
    try:
        add = st.st_TrackJob_add_assign_address_item_detailed
        a = add(job, 257, 2, 5, 48, 258, 1024, 0, 0)
        b = add(job, 11, 3, 0, 8, 1, 1, 0, 0)
        c = add(job, 257, 2, 7, 48, 258, 1024, 1, 0)
        assert a and b and c
        assert st.st_TrackJob_total_num_assign_items(job) == 3
        assert st.st_TrackJob_num_assign_items(job, 2, 1024) == 2
        first = st.st_TrackJob_ptr_assign_address_item_by_index(job, 2, 1024, 0).contents
        second = st.st_TrackJob_ptr_assign_address_item_by_index(job, 2, 1024, 1).contents
        assert (first.dest_elem_index, second.dest_elem_index) == (5, 7)
        assert not st.st_TrackJob_ptr_assign_address_item_by_index(job, 2, 1024, 2)
        assert st.st_TrackJob_index_of_assign_address_item(job, ct.byref(second)) == 1
        assert st.st_TrackJob_num_distinct_available_assign_address_items_dest_src_pairs(job) == 2
        pairs = (st.st_AssignAddressItemDestSrcPair * 4)()
        assert st.st_TrackJob_available_assign_address_items_dest_src_pairs(job, 4, pairs) == 2
        assert [(p.dest_buffer_id, p.src_buffer_id) for p in pairs[:2]] == [(2, 1024), (3, 1)]
        beg = st.st_TrackJob_assign_items_dest_src_begin(job, 2, 1024)
        end = st.st_TrackJob_assign_items_dest_src_end(job, 2, 1024)
        size = ct.sizeof(st.st_AssignAddressItem)
        assert (ct.cast(end, ct.c_void_p).value - ct.cast(beg, ct.c_void_p).value) == 2 * size
        ob, oe = st.st_TrackJob_assign_items_begin(job, 2, 1024), st.st_TrackJob_assign_items_end(job, 2, 1024)
        assert (oe - ob) == 2 * 24                       # two NS(Object) descriptors
        item = st.st_AssignAddressItem()
        ct.memmove(ct.byref(item), ct.byref(first), size)
        assert st.st_TrackJob_remove_assign_address_item(job, ct.byref(item)) == 0
        assert st.st_TrackJob_num_assign_items(job, 2, 1024) == 1
        assert st.st_TrackJob_remove_assign_address_item(job, ct.byref(item)) == -1
    finally:
        st.st_TrackJob_delete(job)


def test_new_style_and_cuda_job_without_device():
    """NS(TrackJobNew_*) / NS(CudaTrackJob_*) on jobs made without buffers:
    "cuda" names the HIP job, the arch id is this library's, there are no
    controller or argument objects; debug mode is remembered."""
    from sixtracklib_amd import stcommon as st
    for arch in (b"cuda", b"hip"):
        job = st.st_TrackJobNew_create(arch, b"block=128")
        assert job
        assert st.st_TrackJobNew_get_arch_id(job) == st.st_ARCHITECTURE_HIP
        assert st.st_TrackJobNew_get_arch_string(job) == b"hip"
        assert not st.st_TrackJobNew_uses_controller(job) and not st.st_TrackJobNew_uses_arguments(job)
        assert not st.st_TrackJobNew_is_in_debug_mode(job)
        assert st.st_TrackJobNew_enable_debug_mode(job) == 0 and st.st_TrackJobNew_is_in_debug_mode(job)
        assert st.st_TrackJobNew_add_assign_address_item_detailed(job, 257, 2, 0, 48, 258, 1024, 0, 0)
        assert st.st_TrackJobNew_total_num_assign_items(job) == 1
        st.st_TrackJobNew_delete(job)
    assert not st.st_TrackJobNew_create(b"opencl", None)
    # the debug mode belongs to its job
    jobs = [st.st_TrackJobNew_create(b"hip", None) for _ in range(300)]
    assert all(jobs)
    for k, j in enumerate(jobs):
        if k % 3 == 0:
            assert st.st_TrackJobNew_enable_debug_mode(j) == 0
    assert [bool(st.st_TrackJobNew_is_in_debug_mode(j)) for j in jobs] == \
        [k % 3 == 0 for k in range(len(jobs))]
    assert st.st_TrackJobNew_disable_debug_mode(jobs[0]) == 0
    assert not st.st_TrackJobNew_is_in_debug_mode(jobs[0])
    for j in jobs:
        st.st_TrackJobNew_delete(j)
    cj = st.st_CudaTrackJob_create()
    assert cj and not st.st_CudaTrackJob_has_controller(cj)
    assert not st.st_CudaTrackJob_get_ptr_particles_arg(cj)
    assert st.st_CudaTrackJob_default_threads_per_block(cj) == 256
    st.st_TrackJobNew_delete(cj)


def test_python_buffer_handle(tmp_path):
    """python/sixtracklib/buffer.py ``Buffer``: mapped on a CBuffer (the C
    buffer sees the CBuffer's objects and capacities), read from a file,
    made empty with given capacities, grown by reserve."""
    from sixtracklib_amd.buffer import Buffer
    eb = lattices.fodo_ring()
    b = Buffer.from_cbuffer(eb)
    assert b.maps_to_cbuffer and b.cbuffer is eb and not b.owns_buffer
    assert b.num_objects == eb.n_objects == 402
    assert b.slot_size == 8 and b.size == eb.words.nbytes
    assert b.data_begin_addr == eb.pointer and not b.needs_remapping
    assert b.max_num_objects >= 402 and b.max_num_slots >= b.num_slots
    path = str(tmp_path / "fodo.bin")
    assert b.write_to_file(path)
    c = Buffer(path_to_file=path)
    assert c.num_objects == 402 and c.is_owner and not c.maps_to_cbuffer
    d = Buffer(num_objects=8, num_slots=64, num_dataptrs=4)
    assert d.num_objects == 0 and d.max_num_objects >= 8 and d.max_num_slots >= 64
    d.reserve(max_num_objects=100, max_num_slots=4000, max_num_dataptrs=40)
    assert d.max_num_objects >= 100 and d.max_num_slots >= 4000 and d.max_num_dataptrs >= 40
    assert Buffer(size=1024).num_objects == 0


def test_particles_set_lists_the_buffer_and_particles_addr():
    from sixtracklib_amd.particles import ParticlesSet, ParticlesAddr, particles_in
    from sixtracklib_amd.config_helper import supports
    ps = ParticlesSet()
    ps.Particles(num_particles=3)
    ps.Particles(num_particles=5)
    again = ParticlesSet(ps.cbuffer)
    assert [p.num_particles for p in again.particles] == [3, 5] and again.num_particles == 8
    assert particles_in(again, 1).num_particles == 5
    cb = CBuffer()
    a = ParticlesAddr(cbuffer=cb, num_particles=7, x=0x1000, state=0x2000)
    assert cb.get_object_typeid(0) == 512 and a.x == 0x1000 and a.state == 0x2000 and a.px == 0
    assert supports("hip") and supports("cuda") and not supports("cpu")


def test_cuda_trackjob_without_device_raises():
    from sixtracklib_amd.cuda import CudaTrackJob
    from sixtracklib_amd import stcommon as st
    if st.st_Hip_num_devices() > 0:
        pytest.skip("a HIP device is present")
    eb = lattices.fodo_ring()
    pb, _ = lattices.gaussian_beam(10, seed=0)
    with pytest.raises(RuntimeError):
        CudaTrackJob(eb, pb)


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus above 1 exits with status 2 and a
    message, before any device work."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    p = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300)
    assert p.returncode == 2
    assert p.stdout == ""
    assert "one GPU" in p.stderr
