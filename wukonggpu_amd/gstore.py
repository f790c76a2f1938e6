"""Host gstore: triples, the synthetic generator and the image builder.

Thin wrappers over libwkg_gstore.so (include/wukong_gstore.h). Arrays are
exposed as zero-copy numpy views of the native buffers.
"""
import ctypes as C

import numpy as np

from . import _native as N

VID_BASE = 1 << 17  # first normal vertex id (datagen/generate_data.cpp:88-98)
TYPE_ID = 1         # rdf:type
PREDICATE_ID = 0    # __PREDICATE__
IN, OUT = 0, 1


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc}): {N.gstore_lib().wks_last_error().decode()}")


def _view(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).view(dtype)


class Triples:
    """An owned set of (s, p, o) id triples (u32 each)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)

    @classmethod
    def generate(cls, seed, num_entities, num_preds, num_triples, zipf_s, degree_cap=0,
                 type_sizes=(), num_hubs=0, hub_degree=0, num_linked=0, link_frac=0.25, siblings=()):
        """Seeded synthetic graph (wks_generate). num_linked pairs predicate
        2+2k+1 to 2+2k (a share link_frac of its lists taken from the source
        predicate's list of the same subject); siblings is a list of
        (source pid, planted pid): consecutive objects of every source list
        get a planted edge."""
        L = N.gstore_lib()
        ts = (C.c_uint64 * max(len(type_sizes), 1))(*[int(x) for x in type_sizes])
        ss = (C.c_uint32 * max(len(siblings), 1))(*[int(a) for a, _ in siblings])
        sd = (C.c_uint32 * max(len(siblings), 1))(*[int(b) for _, b in siblings])
        prm = N.GenParams(seed, num_entities, num_preds, len(type_sizes), num_triples, zipf_s,
                          degree_cap, C.cast(ts, C.POINTER(C.c_uint64)), num_hubs, hub_degree,
                          num_linked, link_frac, len(siblings), C.cast(ss, C.POINTER(C.c_uint32)),
                          C.cast(sd, C.POINTER(C.c_uint32)))
        h = C.c_void_p()
        _check(L.wks_generate(C.byref(prm), C.byref(h)), "generate")
        return cls(h.value)

    @classmethod
    def from_arrays(cls, s, p, o):
        L = N.gstore_lib()
        s = np.ascontiguousarray(s, dtype=np.uint32)
        p = np.ascontiguousarray(p, dtype=np.uint32)
        o = np.ascontiguousarray(o, dtype=np.uint32)
        u = C.POINTER(C.c_uint32)
        h = C.c_void_p()
        _check(L.wks_triples_from_arrays(s.ctypes.data_as(u), p.ctypes.data_as(u), o.ctypes.data_as(u),
                                         len(s), C.byref(h)), "triples_from_arrays")
        return cls(h.value)

    @classmethod
    def read(cls, path):
        h = C.c_void_p()
        _check(N.gstore_lib().wks_triples_read_file(str(path).encode(), C.byref(h)), "read triples")
        return cls(h.value)

    @classmethod
    def read_dataset(cls, path):
        """A dataset directory in the reference's ID format (id_* files,
        str_index, str_normal): (Triples, info dict)."""
        h = C.c_void_p()
        info = N.DatasetInfo()
        _check(N.gstore_lib().wks_dataset_read(str(path).encode(), C.byref(h), C.byref(info)), "read dataset")
        return cls(h.value), {k: int(getattr(info, k)) for k, _ in N.DatasetInfo._fields_}

    def write_dataset(self, path, files=1):
        _check(N.gstore_lib().wks_dataset_write(self._h, str(path).encode(), int(files)), "write dataset")

    def write(self, path):
        _check(N.gstore_lib().wks_triples_write_file(self._h, str(path).encode()), "write triples")

    def __len__(self):
        return int(N.gstore_lib().wks_triples_count(self._h))

    def arrays(self):
        L, n = N.gstore_lib(), len(self)
        return (_view(L.wks_triples_s(self._h), n, np.uint32), _view(L.wks_triples_p(self._h), n, np.uint32),
                _view(L.wks_triples_o(self._h), n, np.uint32))

    def close(self):
        if self._h:
            N.gstore_lib().wks_triples_free(self._h)
            self._h = None

    __del__ = close


class GStore:
    """A built gstore image (core/gstore.hpp layout)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)

    @classmethod
    def build(cls, triples, main_load=0.75, num_threads=0):
        h = C.c_void_p()
        bp = N.BuildParams(main_load, num_threads)
        _check(N.gstore_lib().wks_build(triples._h, C.byref(bp), C.byref(h)), "build")
        return cls(h.value)

    @classmethod
    def load(cls, path):
        h = C.c_void_p()
        _check(N.gstore_lib().wks_gstore_load(str(path).encode(), C.byref(h)), "load image")
        return cls(h.value)

    def save(self, path):
        _check(N.gstore_lib().wks_gstore_save(self._h, str(path).encode()), "save image")

    def close(self):
        if self._h:
            N.gstore_lib().wks_gstore_free(self._h)
            self._h = None

    __del__ = close

    @property
    def num_slots(self):
        return int(N.gstore_lib().wks_num_slots(self._h))

    @property
    def num_buckets(self):
        return int(N.gstore_lib().wks_num_buckets(self._h))

    @property
    def num_buckets_ext(self):
        return int(N.gstore_lib().wks_num_buckets_ext(self._h))

    @property
    def num_keys(self):
        return int(N.gstore_lib().wks_num_keys(self._h))

    @property
    def num_edges(self):
        return int(N.gstore_lib().wks_num_edges(self._h))

    @property
    def num_metas(self):
        return int(N.gstore_lib().wks_num_metas(self._h))

    def slots(self):
        """(num_slots, 2) u64 array of (key, ptr)."""
        L = N.gstore_lib()
        return _view(L.wks_slots(self._h), 2 * self.num_slots, np.uint64).reshape(-1, 2)

    def edges(self):
        return _view(N.gstore_lib().wks_edges(self._h), self.num_edges, np.uint32)

    def metas(self):
        """(num_metas, 8) u64 array in pred_meta_t field order."""
        L = N.gstore_lib()
        p = C.cast(L.wks_metas(self._h), C.POINTER(C.c_uint64))
        return _view(p, 8 * self.num_metas, np.uint64).reshape(-1, 8)

    def metas_ptr(self):
        return N.gstore_lib().wks_metas(self._h)

    def image_nbytes(self):
        return self.num_slots * 16 + self.num_edges * 4

    def degree(self, vid, pid, d):
        """Length of the edge list of key (vid, pid, d) (0 when absent), without a copy."""
        L = N.gstore_lib()
        ptr = C.POINTER(C.c_uint32)()
        n = C.c_uint64()
        _check(L.wks_get_edges(self._h, vid, pid, d, C.byref(ptr), C.byref(n)), "get_edges")
        return int(n.value)

    def get_edges(self, vid, pid, d):
        """Edge list of key (vid, pid, d) as a numpy copy (empty when absent)."""
        L = N.gstore_lib()
        ptr = C.POINTER(C.c_uint32)()
        n = C.c_uint64()
        _check(L.wks_get_edges(self._h, vid, pid, d, C.byref(ptr), C.byref(n)), "get_edges")
        return _view(ptr, n.value, np.uint32).copy()

    def oracle_args(self):
        """Arguments (slots, num_slots, edges, num_edges, metas, num_metas) for the oracle."""
        L = N.gstore_lib()
        return (L.wks_slots(self._h), self.num_slots, L.wks_edges(self._h), self.num_edges,
                C.cast(L.wks_metas(self._h), C.c_void_p), self.num_metas)
