"""The new-style job API (NS(TrackJobNew_*), common/track/track_job_base.h)
and the CUDA job's constructors (NS(CudaTrackJob_*), cuda/track_job.h)
through the C-ABI, against the oracle, and the particle addresses of a job
(track_job.h:206-290) read back through the HIP runtime.

Pattern: tests/sixtracklib/cuda/track/test_track_job_track_until_c99.cpp
(a CUDA job tracks the same particles as the CPU tracker; here bit for bit
in the exact build).
"""
import ctypes as ct

import numpy as np
import pytest

import oracle_lib
from helpers import particles_dict, assert_particles_equal
from sixtracklib_amd import lattices
from sixtracklib_amd.cobjects import CBuffer
from sixtracklib_amd.particles import particles_in, ALL_FIELDS, Particles

pytestmark = pytest.mark.gpu


def _copy(cb):
    return CBuffer.frombytes(cb.tobytes())


def _pdict(cb, idx=0):
    return particles_dict(particles_in(cb, idx))


def _mapped(cb):
    from sixtracklib_amd import stcommon as st
    ptr = st.st_Buffer_new_on_memory(cb.pointer, cb.words.nbytes)
    assert ptr
    return ptr


def _beam():
    return lattices.gaussian_beam(777, seed=9, sigma_x=4e-4, sigma_y=4e-4, sigma_px=4e-6,
                                  sigma_py=4e-6)[0]


def test_track_job_new_until_line_and_collect_match_oracle():
    from sixtracklib_amd import stcommon as st
    eb = lattices.fodo_ring()
    pb = _beam()
    ob = _copy(pb)
    assert oracle_lib.track_until(ob, eb, 20) == 0
    pbuf, ebuf = _mapped(pb), _mapped(eb)
    job = st.st_TrackJobNew_new(b"cuda", pbuf, ebuf, b"fma=0")
    assert job, st.last_error()
    try:
        assert st.st_TrackJobNew_get_total_num_of_particles(job) == 777
        assert st.st_TrackJobNew_track_until(job, 10) == 0
        # turns 10..19 as lines of the whole ring, each finishing its turn
        n = eb.n_objects
        for _ in range(10):
            assert st.st_TrackJobNew_track_line(job, 0, 200, False) == 0
            assert st.st_TrackJobNew_track_line(job, 200, n, True) == 0
        assert st.st_TrackJobNew_collect(job) == 0x05          # particles | output
        assert_particles_equal(_pdict(pb), _pdict(ob), what="TrackJobNew vs oracle")
    finally:
        st.st_TrackJobNew_delete(job)
        st.st_Buffer_delete(pbuf)
        st.st_Buffer_delete(ebuf)


def test_track_job_new_elem_by_elem_matches_oracle():
    from sixtracklib_amd import stcommon as st
    eb = lattices.fodo_ring(n_cells=10)
    E = eb.n_objects
    pb = _beam()
    n = particles_in(pb).num_particles
    ob = _copy(pb)
    pbuf, ebuf = _mapped(pb), _mapped(eb)
    job = st.st_TrackJobNew_new_with_output(b"hip", pbuf, ebuf, None, 3, b"fma=0")
    assert job, st.last_error()
    try:
        assert st.st_TrackJobNew_track_elem_by_elem(job, 3) == 0
        assert st.st_TrackJobNew_collect_output(job) == 0
        assert st.st_TrackJobNew_collect_particles(job) == 0
        obuf = st.st_TrackJobNew_get_output_buffer(job)
        off = st.st_TrackJobNew_get_elem_by_elem_output_buffer_offset(job)
        size = st.st_Buffer_get_size(obuf)
        raw = (ct.c_ubyte * size).from_address(st.st_Buffer_get_data_begin_addr(obuf))
        out = CBuffer.frombytes(bytes(raw))
        got = _pdict(out, off)
        ocb = CBuffer(max_slots=22 * n * E * 3 + 64)
        Particles(cbuffer=ocb, num_particles=n * E * 3)
        cfg = oracle_lib.EbeConfig(order=0, num_particles_to_store=n, num_elements_to_store=E,
                                   num_turns_to_store=3, min_particle_id=0, min_element_id=0,
                                   min_turn=0, max_particle_id=n - 1, max_element_id=E - 1,
                                   max_turn=2, is_rolling=1, out_store_addr=ocb.get_object_address(0))
        assert oracle_lib.track_elem_by_elem(ob, eb, cfg, 3) == 0
        assert_particles_equal(_pdict(pb), _pdict(ob), what="particles")
        assert_particles_equal(got, _pdict(ocb), what="TrackJobNew element-by-element store")
    finally:
        st.st_TrackJobNew_delete(job)
        st.st_Buffer_delete(pbuf)
        st.st_Buffer_delete(ebuf)


def test_cuda_track_job_nodes_and_tracking_match_oracle():
    from sixtracklib_amd import stcommon as st
    n_nodes = st.st_CudaTrackJob_get_num_available_nodes()
    assert n_nodes >= 1
    idx = (ct.c_uint32 * 8)()
    assert st.st_CudaTrackJob_get_available_node_indices_list(8, idx) == n_nodes
    assert list(idx[:n_nodes]) == list(range(n_nodes))
    eb = lattices.kitchen_sink(1)
    pb = lattices.gaussian_beam(300, seed=4, sigma_x=2e-3, sigma_y=2e-3, sigma_px=1e-4,
                                sigma_py=1e-4, sigma_zeta=0.05)[0]
    ob = _copy(pb)
    assert oracle_lib.track_until(ob, eb, 4) == 0
    pbuf, ebuf = _mapped(pb), _mapped(eb)
    job = st.st_CudaTrackJob_new(b"0.0", pbuf, ebuf)
    assert job, st.last_error()
    try:
        assert st.st_TrackJobNew_track_until(job, 4) == 0
        st.st_TrackJobNew_collect(job)
        assert_particles_equal(_pdict(pb), _pdict(ob), what="CudaTrackJob vs oracle")
    finally:
        st.st_TrackJobNew_delete(job)
        st.st_Buffer_delete(pbuf)
        st.st_Buffer_delete(ebuf)


def test_particle_addresses_point_at_the_device_copy():
    """fetch_particle_addresses gives the device addresses of the job's copy
    of each particle set: the arrays read there through the HIP runtime
    equal the collected particles, field for field."""
    from sixtracklib_amd import stcommon as st
    hip = ct.CDLL(st.hip_runtime_path() or "libamdhip64.so")
    hip.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
    hip.hipMemcpy.restype = ct.c_int
    eb = lattices.fodo_ring()
    pb = CBuffer()
    lattices.gaussian_beam(100, seed=1, cbuffer=pb)
    lattices.gaussian_beam(37, seed=2, cbuffer=pb)
    pbuf, ebuf = _mapped(pb), _mapped(eb)
    sets = (ct.c_uint64 * 2)(0, 1)
    job = st.st_TrackJobNew_new_detailed(b"hip", pbuf, 2, sets, ebuf, None, 0, b"fma=0")
    assert job, st.last_error()
    try:
        assert st.st_TrackJobNew_track_until(job, 5) == 0
        assert not st.st_TrackJobNew_get_particle_addresses(job, 0)     # not fetched yet
        assert st.st_TrackJobNew_fetch_particle_addresses(job) == 0
        st.st_TrackJobNew_collect(job)
        for k, n in ((0, 100), (1, 37)):
            a = st.st_TrackJobNew_get_particle_addresses(job, k).contents
            assert a.num_particles == n
            want = _pdict(pb, k)
            for f in ALL_FIELDS:
                name = "at_element_id" if f == "at_element" else f
                dev = getattr(a, name + "_addr")
                host = np.empty(n, dtype=want[f].dtype)
                assert hip.hipMemcpy(host.ctypes.data, dev, host.nbytes, 2) == 0   # D2H
                assert np.array_equal(host.view(np.uint64), want[f].view(np.uint64)), (k, f)
        assert st.st_TrackJob_clear_all_particle_addresses(job) == 0
        assert not st.st_TrackJobNew_get_particle_addresses(job, 0)
    finally:
        st.st_TrackJobNew_delete(job)
        st.st_Buffer_delete(pbuf)
        st.st_Buffer_delete(ebuf)


def test_python_trackjob_reference_methods():
    """The Python TrackJob's remaining methods of the reference binding
    (python/sixtracklib/trackjob.py): particle addresses, reset with new
    buffers, assign-item queries, identity and node queries, and the
    OpenCL-only features refused as the reference refuses them."""
    from sixtracklib_amd import stcommon as st
    from sixtracklib_amd.trackjob import TrackJob
    eb = lattices.fodo_ring()
    pb = _beam()
    job = TrackJob(eb, pb, config_str="fma=0")
    assert job.type() == job.arch_id
    assert TrackJob.enabled_archs() == ["hip"]
    assert TrackJob.num_all_nodes("hip") >= 1 and TrackJob.num_all_nodes("opencl") == 0
    assert TrackJob.num_available_nodes("cuda") == TrackJob.num_all_nodes("hip")
    assert not job.allows_add_program and not job.allows_enable_kernel
    with pytest.raises(RuntimeError):
        job.add_program("k.cl", "")
    with pytest.raises(RuntimeError):
        job.controller
    assert job.opencl_context_addr == 0
    # particle addresses: the x field of the job's device copy holds x
    assert job.can_fetch_particle_addresses
    job.fetch_particle_addresses()
    assert job.has_particle_addresses
    pa = job.get_particle_addresses(0).contents
    assert pa.num_particles == 777 and pa.x_addr != 0
    hip = ct.CDLL(st.hip_runtime_path() or "libamdhip64.so")
    hip.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
    hip.hipMemcpy.restype = ct.c_int
    host = np.empty(777, dtype=np.float64)
    assert hip.hipMemcpy(host.ctypes.data, pa.x_addr, host.nbytes, 2) == 0      # D2H
    np.testing.assert_array_equal(host, _pdict(pb)["x"])
    job.clear_all_particle_addresses()
    assert not job.get_particle_addresses(0)
    # reset onto other particles and another lattice, then track: the
    # oracle's result for the new buffers
    eb2 = lattices.lhc_like_ring(seed=2, n_cells=8, tune=8.31 * 8 / 40)
    pb2 = lattices.gaussian_beam(300, seed=4, sigma_x=2e-4, sigma_y=2e-4, sigma_px=2e-6,
                                 sigma_py=2e-6)[0]
    ob = _copy(pb2)
    assert oracle_lib.track_until(ob, eb2, 3) == 0
    job.reset(eb2, pb2)
    assert job.particles_buffer is pb2 and job.beam_elements_buffer is eb2
    job.track_until(3).collect()
    assert_particles_equal(_pdict(pb2), _pdict(ob), what="after reset")
    # assign-item queries
    item = st.st_AssignAddressItem(dest_elem_type_id=257, dest_buffer_id=2, dest_elem_index=0,
                                   dest_pointer_offset=48, src_elem_type_id=258,
                                   src_buffer_id=1024, src_elem_index=0, src_pointer_offset=0)
    assert not job.has_assign_item(item) and job.index_of_assign_address_item(item) is None
    job.add_assign_address_item(item)
    assert job.has_assign_item(item)
    assert job.has_assign_item(dest_buffer_id=2, src_buffer_id=1024, index=0)
    assert job.index_of_assign_address_item(item) == 0
    assert job.ptr_assign_address_item(dest_buffer_id=2, src_buffer_id=1024, index=0)
    del job


def test_python_cuda_trackjob_matches_oracle():
    """The reference's Python GPU job (python/sixtracklib/cuda.py
    ``CudaTrackJob``) on HIP: node "0.0", the exact build; track_until and
    element by element into its output ParticlesSet, bit-identical to the
    oracle; the new-style properties and camel-case spellings."""
    from sixtracklib_amd import CudaTrackJob, TrackJobBaseNew
    eb = lattices.fodo_ring()
    pb = _beam()
    ob = _copy(pb)
    assert oracle_lib.track_until(ob, eb, 7) == 0
    job = CudaTrackJob(eb, pb, node_id_str="0.0", config_str="fma=0")
    assert isinstance(job, TrackJobBaseNew)
    assert not job.has_controller and not job.has_particles_arg and not job.uses_controller
    assert job.num_particle_sets == 1 and job.total_num_particles == 777
    assert job.default_threads_per_block > 0
    assert not job.is_in_debug_mode
    assert job.enable_debug_mode().is_in_debug_mode
    job.disable_debug_mode()
    job.track_until(7).collectParticles()
    assert job.last_track_status_success
    assert_particles_equal(_pdict(pb), _pdict(ob), what="CudaTrackJob")
    # element by element: the job's output ParticlesSet holds the oracle's store
    eb2 = lattices.fodo_ring()
    pb2 = lattices.gaussian_beam(64, seed=13, sigma_x=4e-4, sigma_y=4e-4, sigma_px=4e-6,
                                 sigma_py=4e-6)[0]
    job2 = CudaTrackJob(eb2, pb2, until_turn_elem_by_elem=2, config_str="fma=0")
    assert job2.has_elem_by_elem_output and job2.has_elem_by_elem_config
    job2.track_elem_by_elem(2)
    job2.collect()
    out = job2.output
    off = job2.elem_by_elem_output_buffer_offset
    n_out = out.particles[off].num_particles
    assert n_out == 64 * 402 * 2
    ocb = CBuffer(max_slots=22 * n_out + 64)
    Particles(cbuffer=ocb, num_particles=n_out)
    ob2 = _copy(lattices.gaussian_beam(64, seed=13, sigma_x=4e-4, sigma_y=4e-4, sigma_px=4e-6,
                                       sigma_py=4e-6)[0])
    cfg = oracle_lib.EbeConfig(order=0, num_particles_to_store=64, num_elements_to_store=402,
                               num_turns_to_store=2, min_particle_id=0, min_element_id=0,
                               min_turn=0, max_particle_id=63, max_element_id=401, max_turn=1,
                               is_rolling=1, out_store_addr=ocb.get_object_address(0))
    assert oracle_lib.track_elem_by_elem(ob2, eb2, cfg, 2) == 0
    assert_particles_equal(_pdict(out.cbuffer, off), _pdict(ocb), what="CudaTrackJob ebe store")
    assert_particles_equal(_pdict(pb2), _pdict(ob2), what="CudaTrackJob ebe particles")
    del job, job2
