"""CObjects managed buffers without the third-party ``cobjects`` package.

The reference Python API builds lattices and particle sets with
``cobjects.CBuffer`` / ``cobjects.CObject`` (python/sixtracklib/
beam_elements.py:12, particles.py:1) and hands the raw memory to the C
library (``st_Buffer_new_mapped_on_cbuffer``, trackjob.py:1314-1327).
``cobjects`` is not vendored in the reference, so this module implements
the same binary format from its C definition:

* header: 8 x u64 ``[stored base address, total size, header size,
  address of the slots / objects / dataptrs / garbage sections, 0]``
  (sixtracklib/common/buffer/managed_buffer_minimal.h:396-466, 629-765);
* each section starts with ``{max_size in bytes, num_entities}``;
* objects are 24-byte ``{begin_addr, type_id, size}`` records
  (buffer/buffer_object.h:24-30); object payloads live in the slots
  section, 8-byte aligned; ``size`` covers the struct and every array it
  points to;
* the dataptrs section lists the address of every pointer field so the
  buffer can be relocated (buffer/managed_buffer_remap.h:644-681).

The buffer is kept relocated to the address of its own numpy storage, so
the bytes can be handed to C code as they are.  Files written by the
reference (``tests/testdata/*.bin``) load unchanged, and files written here
load in the reference.
"""
from __future__ import annotations

import numpy as np

SLOT = 8
HEADER_WORDS = 8


def _round_up(n, a=SLOT):
    return (int(n) + a - 1) // a * a


class CBuffer:
    """A growable CObjects buffer stored in a ``numpy.uint64`` array."""

    def __init__(self, max_slots=512, max_objects=64, max_pointers=64,
                 max_garbage=0, template=None):
        if template is not None:
            self._words = np.array(template, dtype=np.uint64, copy=True)
            self._remap()
            return
        self._alloc(int(max_slots), int(max_objects), int(max_pointers),
                    int(max_garbage))

    # -- construction -----------------------------------------------------
    def _alloc(self, n_slots, n_objs, n_ptrs, n_garb):
        sec_sizes = [16 + n_slots * 8, 16 + n_objs * 24, 16 + n_ptrs * 8,
                     16 + n_garb * 16]
        total = HEADER_WORDS * 8 + sum(sec_sizes)
        w = np.zeros(total // 8, dtype=np.uint64)
        base = w.ctypes.data
        w[0] = base
        w[1] = total
        w[2] = HEADER_WORDS * 8
        off = HEADER_WORDS * 8
        for k, sz in enumerate(sec_sizes):
            w[3 + k] = base + off
            w[off // 8] = sz
            w[off // 8 + 1] = 0
            off += sz
        self._words = w

    @classmethod
    def frombytes(cls, data):
        raw = np.frombuffer(bytes(data), dtype=np.uint64)
        return cls(template=raw)

    @classmethod
    def fromfile(cls, path):
        with open(path, "rb") as fh:
            return cls.frombytes(fh.read())

    def tobytes(self):
        return self._words[: self.size // 8].tobytes()

    def tofile(self, path):
        with open(path, "wb") as fh:
            fh.write(self.tobytes())

    # -- relocation ---------------------------------------------------------
    def _remap(self):
        """Relocate every stored address to the current numpy storage
        (buffer/managed_buffer_remap.h:139-178, 345-550)."""
        w = self._words
        base = w.ctypes.data
        stored = int(w[0])
        if stored == base:
            return
        off = np.uint64((base - stored) % (1 << 64))
        w[3:7] = w[3:7] + off
        w[0] = base
        objs = self._section(4)
        n = int(w[objs + 1])
        if n:
            idx = objs + 2 + 3 * np.arange(n)
            w[idx] = w[idx] + off
        dps = self._section(5)
        n = int(w[dps + 1])
        if n:
            idx = dps + 2 + np.arange(n)
            w[idx] = w[idx] + off
            slots = (w[idx] - np.uint64(base)) // np.uint64(8)
            slots = slots.astype(np.int64)
            nz = w[slots] != 0
            w[slots[nz]] = w[slots[nz]] + off

    # -- geometry -----------------------------------------------------------
    @property
    def base(self):
        return self._words.ctypes.data

    @property
    def size(self):
        return int(self._words[1])

    @property
    def words(self):
        return self._words

    @property
    def pointer(self):
        return self._words.ctypes.data

    def _section(self, sid):
        """word index of the start of section ``sid`` (3..6)"""
        return (int(self._words[sid]) - self.base) // 8

    def _sec_capacity(self, sid, entity_bytes):
        s = self._section(sid)
        return (int(self._words[s]) - 16) // entity_bytes

    @property
    def n_objects(self):
        return int(self._words[self._section(4) + 1])

    def _obj_rec(self, i):
        s = self._section(4)
        n = int(self._words[s + 1])
        if i < 0:
            i += n
        if not 0 <= i < n:
            raise IndexError("object index %d out of range" % i)
        r = s + 2 + 3 * i
        return r

    def get_object_typeid(self, i):
        return int(self._words[self._obj_rec(i) + 1])

    def get_object_size(self, i):
        return int(self._words[self._obj_rec(i) + 2])

    def get_object_word_offset(self, i):
        """word index of object ``i``'s struct inside ``words``"""
        return (int(self._words[self._obj_rec(i)]) - self.base) // 8

    def get_object_address(self, i):
        return int(self._words[self._obj_rec(i)])

    def typeids(self):
        s = self._section(4)
        n = int(self._words[s + 1])
        return self._words[s + 2: s + 2 + 3 * n].reshape(n, 3)[:, 1].copy()

    # -- adding objects -----------------------------------------------------
    def _grow(self, need_slots, need_objs, need_ptrs):
        w = self._words
        sl, ob, dp, ga = (self._section(k) for k in (3, 4, 5, 6))
        used_slots, n_objs = int(w[sl + 1]), int(w[ob + 1])
        n_ptrs, n_garb = int(w[dp + 1]), int(w[ga + 1])
        cap_slots = self._sec_capacity(3, 8)
        cap_objs = self._sec_capacity(4, 24)
        cap_ptrs = self._sec_capacity(5, 8)
        cap_garb = self._sec_capacity(6, 16)
        new_slots = max(cap_slots, 64)
        while new_slots < used_slots + need_slots:
            new_slots *= 2
        new_objs = max(cap_objs, 16)
        while new_objs < n_objs + need_objs:
            new_objs *= 2
        new_ptrs = max(cap_ptrs, 16)
        while new_ptrs < n_ptrs + need_ptrs:
            new_ptrs *= 2
        old = self._words
        old_base = self.base
        self._alloc(new_slots, new_objs, new_ptrs, cap_garb)
        nw = self._words
        new_base = self.base
        shift = np.uint64((new_base - old_base) % (1 << 64))
        nsl = self._section(3)
        nob = self._section(4)
        ndp = self._section(5)
        # slots: copy payload, then shift addresses (slot payload moves by
        # a different amount than the header base when sections grow)
        nw[nsl + 1] = used_slots
        nw[nsl + 2: nsl + 2 + used_slots] = old[sl + 2: sl + 2 + used_slots]
        slot_shift = np.int64(((nsl - sl) * 8) + (new_base - old_base))
        nw[nob + 1] = n_objs
        if n_objs:
            rec = old[ob + 2: ob + 2 + 3 * n_objs].reshape(n_objs, 3).copy()
            rec[:, 0] = (rec[:, 0].astype(np.int64) + slot_shift).astype(np.uint64)
            nw[nob + 2: nob + 2 + 3 * n_objs] = rec.ravel()
        nw[ndp + 1] = n_ptrs
        if n_ptrs:
            ptr_addr = old[dp + 2: dp + 2 + n_ptrs].astype(np.int64) + slot_shift
            nw[ndp + 2: ndp + 2 + n_ptrs] = ptr_addr.astype(np.uint64)
            widx = (ptr_addr - np.int64(new_base)) // 8
            vals = nw[widx].astype(np.int64)
            nz = vals != 0
            vals[nz] += slot_shift
            nw[widx] = vals.astype(np.uint64)
        del shift
        _ = n_garb

    def add_object(self, type_id, struct_words, arrays=()):
        """Append an object.

        ``struct_words``: sequence of 8-byte values (numpy uint64 view of the
        POD struct).  ``arrays``: list of ``(field_word_index, ndarray)``; each
        array is placed after the struct (8-byte aligned) and the struct word
        at ``field_word_index`` receives its address (a dataptr entry).
        Returns the object index.
        """
        sw = np.asarray(struct_words).view(np.uint64).ravel()
        blobs = []
        for fidx, arr in arrays:
            a = np.ascontiguousarray(arr)
            b = np.frombuffer(a.tobytes() + b"\0" * (_round_up(a.nbytes) - a.nbytes),
                              dtype=np.uint64)
            blobs.append((int(fidx), b))
        nslots = len(sw) + sum(len(b) for _, b in blobs)
        w = self._words
        sl, ob, dp = self._section(3), self._section(4), self._section(5)
        if (int(w[sl + 1]) + nslots > self._sec_capacity(3, 8) or
                int(w[ob + 1]) + 1 > self._sec_capacity(4, 24) or
                int(w[dp + 1]) + len(blobs) > self._sec_capacity(5, 8)):
            self._grow(nslots, 1, len(blobs))
            w = self._words
            sl, ob, dp = self._section(3), self._section(4), self._section(5)
        base = self.base
        used = int(w[sl + 1])
        start = sl + 2 + used
        w[start: start + len(sw)] = sw
        cur = start + len(sw)
        nptr = int(w[dp + 1])
        for fidx, b in blobs:
            w[cur: cur + len(b)] = b
            w[start + fidx] = base + cur * 8
            w[dp + 2 + nptr] = base + (start + fidx) * 8
            nptr += 1
            cur += len(b)
        w[dp + 1] = nptr
        w[sl + 1] = used + nslots
        nobj = int(w[ob + 1])
        r = ob + 2 + 3 * nobj
        w[r] = base + start * 8
        w[r + 1] = int(type_id)
        w[r + 2] = nslots * 8
        w[ob + 1] = nobj + 1
        # total size stays the allocated length (the reference stores the
        # buffer's full extent in header[1])
        return nobj

    def __len__(self):
        return self.n_objects
