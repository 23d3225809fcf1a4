"""Batched multi-seed tracking: many (lattice, particle set) pairs in one
persistent launch per chunk of turns, the chunks of a call in one HIP graph
(C API in include/sixtracklib_amd/track_batch.h).

No counterpart in the reference, whose users run one TrackJob per seed one
after the other (python/sixtracklib/trackjob.py:1285-1560 per job).  Each
seed's result is bit-identical to TrackJob(lattice, particles).track_until
on that seed alone (same per-particle code and build flags).

    batch = TrackBatch([lat_0, ..., lat_63], [parts_0, ..., parts_63],
                       config_str="fma=1")
    batch.autotune(trial_turns=2)          # optional
    batch.track_until(1000)
    batch.collect()                        # results in parts_s
"""
from __future__ import annotations

import ctypes as ct

from . import stcommon as st


class TrackBatch:
    def __init__(self, lattices, particles, config_str=None):
        if len(lattices) != len(particles) or not lattices:
            raise ValueError("need one lattice per particle buffer (and at least one)")
        self._lattices = list(lattices)       # keep the CBuffers alive
        self._particles = list(particles)
        self._c_buffers = []
        self.ptr = None
        n = len(lattices)
        lat_arr = (st.st_Buffer_p * n)(*[self._map(b) for b in self._lattices])
        par_arr = (st.st_Buffer_p * n)(*[self._map(b) for b in self._particles])
        self.ptr = st.st_TrackBatch_new(n, lat_arr, par_arr,
                                        st.string_to_encoded_ctypes_str(config_str))
        if not self.ptr:
            self._free_buffers()
            raise RuntimeError("TrackBatch creation failed: %s" % st.last_error())

    def _map(self, cbuf):
        p = st.st_Buffer_new_on_memory(cbuf.pointer, cbuf.words.nbytes)
        if not p:
            raise RuntimeError("cannot map buffer")
        self._c_buffers.append(p)
        return p

    def _free_buffers(self, _buffer_delete=st.st_Buffer_delete):
        for p in self._c_buffers:
            _buffer_delete(p)
        self._c_buffers = []

    # library functions bound as defaults: module globals may be gone at exit
    def __del__(self, _batch_delete=st.st_TrackBatch_delete):
        if getattr(self, "ptr", None):
            _batch_delete(self.ptr)
            self.ptr = None
        if getattr(self, "_c_buffers", None):
            self._free_buffers()

    def track_until(self, until_turn):
        status = st.st_TrackBatch_track_until(self.ptr, int(until_turn))
        if status != 0:
            raise RuntimeError("TrackBatch.track_until failed with status %d %s"
                               % (status, st.last_error()))
        return status

    def collect(self):
        if st.st_TrackBatch_collect(self.ptr) != 0:
            raise RuntimeError("TrackBatch.collect failed: %s" % st.last_error())

    def push(self):
        if st.st_TrackBatch_push(self.ptr) != 0:
            raise RuntimeError("TrackBatch.push failed: %s" % st.last_error())

    def set_config(self, config_str):
        if st.st_TrackBatch_set_config(self.ptr, config_str.encode()) != 0:
            raise ValueError("bad TrackBatch config %r: %s" % (config_str, st.last_error()))

    @property
    def config(self):
        n = st.st_TrackBatch_get_config(self.ptr, None, 0)
        buf = ct.create_string_buffer(int(n) + 1)
        st.st_TrackBatch_get_config(self.ptr, buf, n + 1)
        return buf.value.decode()

    def autotune(self, trial_turns=2, skip_turns=None):
        """Time every candidate (block, waves_per_cu, queues, particles per
        lane, lattice staging) for ``trial_turns`` turns on this batch,
        starting after ``skip_turns`` turns (default 10 x trial_turns, past
        the early losses), keep the fastest and restore the batch state;
        returns the chosen configuration string."""
        if skip_turns is None:
            rc = st.st_TrackBatch_autotune(self.ptr, int(trial_turns))
        else:
            rc = st.st_TrackBatch_autotune_window(self.ptr, int(skip_turns), int(trial_turns))
        if rc != 0:
            raise RuntimeError("TrackBatch.autotune failed: %s" % st.last_error())
        return self.config

    @property
    def num_seeds(self):
        return int(st.st_TrackBatch_num_seeds(self.ptr))

    @property
    def total_particles(self):
        return int(st.st_TrackBatch_total_particles(self.ptr))

    @property
    def last_kernel_time_ms(self):
        return float(st.st_TrackBatch_last_kernel_time_ms(self.ptr))
