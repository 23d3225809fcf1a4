"""Typed objects stored in a :class:`~sixtracklib_amd.cobjects.CBuffer`.

Mirrors the small part of the third-party ``cobjects`` API that the
reference's Python package uses to declare its element and particle
classes (``CObject`` subclasses with ``CField`` descriptors, e.g.
python/sixtracklib/beam_elements.py:42-66, particles.py:6-225): scalar
fields are 8-byte words of the object's POD struct, pointer fields are
arrays stored after the struct whose address is kept in the struct.

Attribute access always goes through the buffer, so values written here
are what the C library and the GPU see after ``TrackJob.push_*``.
"""
from __future__ import annotations

import numpy as np

from .cobjects import CBuffer

_DTYPES = {
    "real": np.float64, "float64": np.float64, "f8": np.float64,
    "int64": np.int64, "i8": np.int64,
    "uint64": np.uint64, "u8": np.uint64,
}


class CField:
    """Descriptor of one 8-byte struct word (scalar or pointer)."""

    def __init__(self, index, ftype="real", default=0, const=False,
                 pointer=False, length=None, alignment=8, setter=None):
        self.index = int(index)
        self.dtype = _DTYPES[ftype]
        self.default = default
        self.const = const
        self.pointer = pointer
        self.length = length
        self.setter = setter
        self.name = None

    def __set_name__(self, owner, name):
        self.name = name

    def _view(self, obj):
        buf = obj._buffer
        w = buf.words
        off = buf.get_object_word_offset(obj._index)
        if not self.pointer:
            return w[off + self.index: off + self.index + 1].view(self.dtype)
        addr = int(w[off + self.index])
        n = obj._array_length(self)
        start = (addr - buf.base) // 8
        return w[start: start + n].view(self.dtype)

    def __get__(self, obj, owner=None):
        if obj is None:
            return self
        v = self._view(obj)
        return v if self.pointer else v[0]

    def __set__(self, obj, value):
        v = self._view(obj)
        if self.pointer:
            v[:] = value
        else:
            v[0] = value
        if self.setter is not None:
            self.setter(obj)


class CObject:
    """Base class: one object in a CBuffer."""

    _typeid = 0
    _registry = {}

    def __init_subclass__(cls, **kw):
        super().__init_subclass__(**kw)
        fields = {}
        for klass in reversed(cls.__mro__):
            for name, val in vars(klass).items():
                if isinstance(val, CField):
                    fields[name] = val
        cls._cfields = sorted(fields.values(), key=lambda f: f.index)
        if cls._typeid:
            CObject._registry.setdefault(int(cls._typeid), cls)

    def __init__(self, cbuffer=None, **kwargs):
        if cbuffer is None:
            cbuffer = CBuffer()
        self._buffer = cbuffer
        nwords = max((f.index for f in self._cfields), default=-1) + 1
        words = np.zeros(nwords, dtype=np.uint64)
        scalars = {}
        for f in self._cfields:
            if f.pointer:
                continue
            val = kwargs.get(f.name, f.default)
            if val is None:
                val = 0
            words[f.index: f.index + 1].view(f.dtype)[0] = val
            scalars[f.name] = words[f.index: f.index + 1].view(f.dtype)[0]
        arrays = []
        for f in self._cfields:
            if not f.pointer:
                continue
            n = int(eval(str(f.length), {}, dict(scalars)))
            arr = np.zeros(n, dtype=f.dtype)
            val = kwargs.get(f.name, None)
            if val is None:
                val = f.default
            arr[:] = val
            arrays.append((f.index, arr))
        self._index = cbuffer.add_object(self._typeid, words, arrays)
        self._lengths = None

    def _array_length(self, field):
        scal = {f.name: getattr(self, f.name) for f in self._cfields
                if not f.pointer}
        return int(eval(str(field.length), {}, scal))

    @classmethod
    def _from_buffer(cls, cbuffer, index):
        obj = cls.__new__(cls)
        obj._buffer = cbuffer
        obj._index = int(index)
        obj._lengths = None
        return obj

    @property
    def cbuffer(self):
        return self._buffer

    @property
    def index(self):
        return self._index

    def to_dict(self):
        return {f.name: (np.array(getattr(self, f.name)) if f.pointer
                         else getattr(self, f.name)) for f in self._cfields}

    def __repr__(self):
        items = ", ".join("%s=%r" % (f.name, getattr(self, f.name))
                          for f in self._cfields if not f.pointer)
        return "%s(%s)" % (type(self).__name__, items)


def get_object(cbuffer, index):
    """Wrap object ``index`` of ``cbuffer`` in its registered class."""
    tid = cbuffer.get_object_typeid(index)
    cls = CObject._registry.get(tid)
    if cls is None:
        raise TypeError("no Python class registered for type id %d" % tid)
    return cls._from_buffer(cbuffer, index)
