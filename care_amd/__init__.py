"""care_amd: MI355X-native CARE read-correction engine.

Python bindings (ctypes) over the C-ABI declared in include/care_amd.h
(engine: read store in HBM, device minhash tables, fused correction kernel)
and include/care_host.h (read packing, the care-gpu command line).

The native libraries are built in-tree by care_amd.build (or
__graft_entry__.build()).  Nothing here falls back to a CPU path: if the HIP
engine is missing, loading fails with an error.
"""

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD_DIR = os.path.join(HERE, "_build")
ENGINE_LIB = os.path.join(BUILD_DIR, "libcare_amd.so")
HOST_LIB = os.path.join(BUILD_DIR, "libcare_host.so")
CARE_GPU = os.path.join(BUILD_DIR, "care-gpu")
READSIM = os.path.join(BUILD_DIR, "readsim")

CARE_OK = 0


class CareOptions(ctypes.Structure):
    """care_options (include/care_amd.h); defaults of include/options.hpp:36-98."""

    _fields_ = [
        ("kmer_length", ctypes.c_int),
        ("num_hash_maps", ctypes.c_int),
        ("coverage", ctypes.c_float),
        ("paired_end", ctypes.c_int),
        ("use_quality", ctypes.c_int),
        ("exclude_ambiguous", ctypes.c_int),
        ("correct_candidates", ctypes.c_int),
        ("new_columns_to_correct", ctypes.c_int),
        ("min_overlap", ctypes.c_int),
        ("max_mismatch_ratio", ctypes.c_float),
        ("min_overlap_ratio", ctypes.c_float),
        ("error_factor", ctypes.c_float),
        ("coverage_factor", ctypes.c_float),
        ("paired_filter_threshold", ctypes.c_float),
        ("correction_type", ctypes.c_int),
        ("correction_type_cands", ctypes.c_int),
        ("threshold_anchor", ctypes.c_float),
        ("threshold_cands", ctypes.c_float),
        ("candidate_limit", ctypes.c_int),
        ("max_refinements", ctypes.c_int),
    ]


class BatchResult(ctypes.Structure):
    _fields_ = [
        ("anchor_status", ctypes.POINTER(ctypes.c_uint8)),
        ("anchor_sequences", ctypes.POINTER(ctypes.c_char)),
        ("sequence_pitch", ctypes.c_int),
        ("candidate_read_ids", ctypes.POINTER(ctypes.c_uint32)),
        ("candidate_sequences", ctypes.POINTER(ctypes.c_char)),
        ("candidate_capacity", ctypes.c_int),
        ("num_candidates", ctypes.c_int),
    ]


class BatchStats(ctypes.Structure):
    _fields_ = [
        ("query_ms", ctypes.c_double),
        ("correct_ms", ctypes.c_double),
        ("total_ms", ctypes.c_double),
        ("raw_hits", ctypes.c_uint64),
        ("candidates", ctypes.c_uint64),
        ("launches", ctypes.c_uint32),
    ]


class CareTuning(ctypes.Structure):
    """Launch shapes of the query and correction kernels (care_tuning);
    no shape changes a result."""
    _fields_ = [
        ("batch_anchors", ctypes.c_int),
        ("correct_threads", ctypes.c_int),
        ("correct_blocks_per_cu", ctypes.c_int),
        ("msa_chunk", ctypes.c_int),
        ("query_lds_hits", ctypes.c_int),
        ("query_register_hits", ctypes.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class MapView(ctypes.Structure):
    """One minhash map in key-grouped form (care_map_view)."""
    _fields_ = [
        ("num_keys", ctypes.c_uint64),
        ("keys", ctypes.c_void_p),
        ("offsets", ctypes.c_void_p),
        ("counts", ctypes.c_void_p),
        ("num_values", ctypes.c_uint64),
        ("values", ctypes.c_void_p),
    ]


MLP_FEATURES = 32
MLP_PARAMS = 4737


class MlpTrainOptions(ctypes.Structure):
    _fields_ = [
        ("epochs", ctypes.c_int),
        ("batch", ctypes.c_int),
        ("lr", ctypes.c_float),
        ("beta1", ctypes.c_float),
        ("beta2", ctypes.c_float),
        ("eps", ctypes.c_float),
        ("seed", ctypes.c_uint64),
    ]


# exported C symbols and their signatures
ENGINE_SYMBOLS = {
    "care_options_default": (None, [ctypes.POINTER(CareOptions)]),
    "care_last_error": (ctypes.c_char_p, []),
    "care_context_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(CareOptions), ctypes.POINTER(ctypes.c_void_p)]),
    "care_context_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "care_readstore_upload": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p,
                                             ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_int, ctypes.c_int]),
    "care_readstore_gather": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                             ctypes.c_int, ctypes.c_void_p]),
    "care_minhash_build": (ctypes.c_int, [ctypes.c_void_p]),
    "care_minhash_import": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(MapView), ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int]),
    "care_minhash_geometry": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                             ctypes.POINTER(ctypes.c_int)]),
    "care_minhash_map_size": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64),
                                             ctypes.POINTER(ctypes.c_uint64)]),
    "care_minhash_export_map": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                               ctypes.c_void_p, ctypes.c_void_p]),
    "care_minhash_signatures": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p]),
    "care_minhash_query": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_uint64]),
    "care_correct_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int,
                                          ctypes.POINTER(BatchResult)]),
    "care_correct_batch_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int,
                                                  ctypes.POINTER(BatchResult)]),
    "care_correct_batch_graph": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int,
                                                ctypes.POINTER(BatchResult)]),
    "care_set_launch_config": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "care_set_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(CareTuning)]),
    "care_get_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(CareTuning)]),
    "care_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(CareTuning)]),
    "care_forest_load": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_uint32]),
    "care_last_batch_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(BatchStats)]),
    "care_host_buffer_alloc": (ctypes.c_int, [ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)]),
    "care_host_buffer_free": (ctypes.c_int, [ctypes.c_void_p]),
    "care_mlp_load": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "care_mlp_gradient": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "care_mlp_train": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64,
                                      ctypes.POINTER(MlpTrainOptions), ctypes.c_void_p, ctypes.c_void_p]),
    "care_mlp_infer": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p,
                                      ctypes.c_void_p]),
    "care_gate_columns": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]),
    "care_gate_columns_candidates": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p,
                                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                    ctypes.c_void_p, ctypes.c_uint64,
                                                    ctypes.POINTER(ctypes.c_uint64)]),
    "care_gate_columns_plain": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p,
                                               ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                               ctypes.c_void_p, ctypes.c_uint64,
                                               ctypes.POINTER(ctypes.c_uint64)]),
    "care_mlp_load_candidates": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "care_readstore_gather_sequences_async": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                                             ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
    "care_readstore_gather_lengths_async": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                                           ctypes.c_void_p, ctypes.c_void_p]),
    "care_readstore_gather_qualities_async": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                                             ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
    "care_readstore_ambiguous_async": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                                      ctypes.c_void_p, ctypes.c_void_p]),
    "care_minhash_max_candidates": (ctypes.c_int, [ctypes.c_void_p]),
    "care_minhash_query_workspace_bytes": (ctypes.c_size_t, [ctypes.c_void_p, ctypes.c_int]),
    "care_minhash_query_async": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                                ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]),
    "care_device_alloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)]),
    "care_device_free": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "care_device_copy": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                        ctypes.c_int]),
    "care_stream_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]),
    "care_stream_destroy": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "care_stream_synchronize": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
}

HOST_SYMBOLS = {
    "care_host_main": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p)]),
    "care_host_last_error": (ctypes.c_char_p, []),
    "care_readset_load": (ctypes.c_void_p, [ctypes.POINTER(ctypes.c_char_p), ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int]),
    "care_readset_free": (None, [ctypes.c_void_p]),
    "care_readset_size": (ctypes.c_uint64, [ctypes.c_void_p]),
    "care_readset_max_length": (ctypes.c_int, [ctypes.c_void_p]),
    "care_readset_seq_pitch": (ctypes.c_int, [ctypes.c_void_p]),
    "care_readset_qual_pitch": (ctypes.c_int, [ctypes.c_void_p]),
    "care_readset_seqs": (ctypes.c_void_p, [ctypes.c_void_p]),
    "care_readset_lengths": (ctypes.c_void_p, [ctypes.c_void_p]),
    "care_readset_ambiguous": (ctypes.c_void_p, [ctypes.c_void_p]),
    "care_readset_quals": (ctypes.c_void_p, [ctypes.c_void_p]),
    "care_readset_save": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]),
    "care_readset_load_preprocessed": (ctypes.c_void_p, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int]),
    "care_minhash_save_file": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_float]),
    "care_minhash_load_file": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int,
                                              ctypes.POINTER(ctypes.c_int)]),
}

_engine = None
_host = None


def _bind(lib, table):
    for name, (res, args) in table.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def engine():
    """The HIP engine library; raises if it has not been built."""
    global _engine
    if _engine is None:
        if not os.path.exists(ENGINE_LIB):
            raise ImportError("care_amd HIP engine not built (%s missing); run __graft_entry__.build()" % ENGINE_LIB)
        _engine = _bind(ctypes.CDLL(ENGINE_LIB, mode=ctypes.RTLD_GLOBAL), ENGINE_SYMBOLS)
    return _engine


def host():
    global _host
    if _host is None:
        engine()
        if not os.path.exists(HOST_LIB):
            raise ImportError("care_amd host library not built (%s missing)" % HOST_LIB)
        _host = _bind(ctypes.CDLL(HOST_LIB), HOST_SYMBOLS)
    return _host


class CareError(RuntimeError):
    pass


def pinned_empty(shape, dtype, lib=None):
    """numpy array over page-locked host memory (care_host_buffer_alloc),
    freed when the array is collected; result copies into it need no
    staging."""
    import weakref
    lib = lib if lib is not None else engine()
    dt = np.dtype(dtype)
    count = int(np.prod(shape)) if shape else 1
    nbytes = max(1, count * dt.itemsize)
    p = ctypes.c_void_p()
    if lib.care_host_buffer_alloc(nbytes, ctypes.byref(p)) != CARE_OK:
        raise CareError(lib.care_last_error().decode())
    buf = (ctypes.c_char * nbytes).from_address(p.value)
    arr = np.frombuffer(buf, dtype=dt, count=count).reshape(shape)
    weakref.finalize(buf, lib.care_host_buffer_free, ctypes.c_void_p(p.value))
    return arr


def _check(rc):
    if rc != CARE_OK:
        raise CareError("care_amd error %d: %s" % (rc, engine().care_last_error().decode()))


def default_options(**kw):
    opt = CareOptions()
    engine().care_options_default(ctypes.byref(opt))
    for k, v in kw.items():
        if not hasattr(opt, k):
            raise KeyError(k)
        setattr(opt, k, v)
    return opt


class ReadSet:
    """Packed reads (2-bit, N rule, binned qualities) parsed by the C++ host."""

    def __init__(self, files, paired_end=False, qual_bits=0, _handle=None):
        h = host()
        if _handle is None:
            arr = (ctypes.c_char_p * len(files))(*[f.encode() for f in files])
            self._p = h.care_readset_load(arr, len(files), int(paired_end), int(qual_bits))
        else:
            self._p = _handle
        if not self._p:
            raise CareError("read loading failed: %s" % h.care_host_last_error().decode())
        self.n = int(h.care_readset_size(self._p))
        self.max_length = int(h.care_readset_max_length(self._p))
        self.pitch = int(h.care_readset_seq_pitch(self._p))
        self.qpitch = int(h.care_readset_qual_pitch(self._p))
        self.qual_bits = int(qual_bits) if self.qpitch else 0

    def _array(self, fn, dtype, count):
        ptr = fn(self._p)
        if not ptr or count == 0:
            return np.zeros(0, dtype=dtype)
        buf = (ctypes.c_char * (count * np.dtype(dtype).itemsize)).from_address(ptr)
        return np.frombuffer(buf, dtype=dtype).copy()

    @property
    def seqs(self):
        return self._array(host().care_readset_seqs, np.uint32, self.n * self.pitch).reshape(self.n, self.pitch)

    @property
    def lengths(self):
        return self._array(host().care_readset_lengths, np.int32, self.n)

    @property
    def ambiguous(self):
        return self._array(host().care_readset_ambiguous, np.uint8, self.n)

    @property
    def quals(self):
        return self._array(host().care_readset_quals, np.uint32, self.n * self.qpitch).reshape(self.n, max(self.qpitch, 1))

    def save(self, path, qual_bits=8):
        """Writes the reference's preprocessed-reads file (care_readset_save)."""
        if host().care_readset_save(self._p, path.encode(), qual_bits) != 0:
            raise CareError(host().care_host_last_error().decode())

    @classmethod
    def load_preprocessed(cls, path, use_quality=True, qual_bits=8):
        """Reads a preprocessed-reads file (care_readset_load_preprocessed)."""
        p = host().care_readset_load_preprocessed(path.encode(), int(use_quality), int(qual_bits))
        if not p:
            raise CareError(host().care_host_last_error().decode())
        return cls(None, qual_bits=qual_bits if use_quality else 0, _handle=p)

    def close(self):
        if self._p:
            host().care_readset_free(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Engine:
    """One device context of the HIP engine."""

    def __init__(self, options, device=0, lib=None):
        """lib: an already bound engine library (default: the built HIP
        engine)."""
        self.lib = lib if lib is not None else engine()
        self.options = options
        ctx = ctypes.c_void_p()
        self._ck(self.lib.care_context_create(device, ctypes.byref(options), ctypes.byref(ctx)))
        self.ctx = ctx
        self.n = 0
        self.max_length = 0

    def _ck(self, rc):
        if rc != CARE_OK:
            raise CareError("care_amd error %d: %s" % (rc, self.lib.care_last_error().decode()))

    def close(self):
        if self.ctx:
            self.lib.care_context_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def upload(self, seqs, lengths, ambiguous=None, quals=None, qual_bits=8, max_length=None):
        seqs = np.ascontiguousarray(seqs, dtype=np.uint32)
        lengths = np.ascontiguousarray(lengths, dtype=np.int32)
        n, pitch = seqs.shape
        amb = None if ambiguous is None else np.ascontiguousarray(ambiguous, dtype=np.uint8)
        q = None if quals is None else np.ascontiguousarray(quals, dtype=np.uint32)
        ml = int(max_length if max_length is not None else lengths.max())
        self._ck(self.lib.care_readstore_upload(
            self.ctx, n, ml, seqs.ctypes.data, pitch, lengths.ctypes.data,
            None if amb is None else amb.ctypes.data, None if q is None else q.ctypes.data,
            0 if q is None else q.shape[1], qual_bits))
        self.n = n
        self.max_length = ml

    def upload_readset(self, rs):
        quals = rs.quals if (self.options.use_quality and rs.qpitch) else None
        self.upload(rs.seqs, rs.lengths, rs.ambiguous, quals, rs.qual_bits or 8, rs.max_length)

    def build(self):
        self._ck(self.lib.care_minhash_build(self.ctx))

    def export_map(self, m):
        """Map m in key-grouped form: (keys ascending, offsets, counts, values)."""
        nk, nv = ctypes.c_uint64(), ctypes.c_uint64()
        self._ck(self.lib.care_minhash_map_size(self.ctx, m, ctypes.byref(nk), ctypes.byref(nv)))
        keys = np.zeros(nk.value, dtype=np.uint64)
        offs = np.zeros(nk.value, dtype=np.uint32)
        cnts = np.zeros(nk.value, dtype=np.uint32)
        vals = np.zeros(nv.value, dtype=np.uint32)
        self._ck(self.lib.care_minhash_export_map(self.ctx, m, keys.ctypes.data, offs.ctypes.data, cnts.ctypes.data,
                                                  vals.ctypes.data))
        return keys, offs, cnts, vals

    def import_maps(self, maps, kmer_length, results_per_map):
        """Installs prebuilt maps [(keys, offsets, counts, values), ...] in
        place of build()."""
        arrs = []
        views = (MapView * len(maps))()
        for i, (k, o, c, v) in enumerate(maps):
            k = np.ascontiguousarray(k, dtype=np.uint64)
            o = np.ascontiguousarray(o, dtype=np.uint32)
            c = np.ascontiguousarray(c, dtype=np.uint32)
            v = np.ascontiguousarray(v, dtype=np.uint32)
            arrs.append((k, o, c, v))
            views[i] = MapView(len(k), k.ctypes.data, o.ctypes.data, c.ctypes.data, len(v), v.ctypes.data)
        self._ck(self.lib.care_minhash_import(self.ctx, views, len(maps), kmer_length, results_per_map))
        self.options.num_hash_maps = len(maps)
        self.options.kmer_length = kmer_length

    def set_launch_config(self, batch=0, threads=0):
        self._ck(self.lib.care_set_launch_config(self.ctx, batch, threads))

    def set_tuning(self, **shape):
        """Sets launch shapes (care_tuning fields; missing fields are 0 = automatic)."""
        t = CareTuning(**shape)
        self._ck(self.lib.care_set_tuning(self.ctx, ctypes.byref(t)))

    def tuning(self):
        t = CareTuning()
        self._ck(self.lib.care_get_tuning(self.ctx, ctypes.byref(t)))
        return t.as_dict()

    def autotune(self, first=0, n=65536):
        """Times candidate launch shapes on anchors [first, first+n) and keeps the fastest."""
        t = CareTuning()
        self._ck(self.lib.care_autotune(self.ctx, first, n, ctypes.byref(t)))
        return t.as_dict()

    def signatures(self, first, n):
        out = np.zeros((n, self.options.num_hash_maps), dtype=np.uint64)
        self._ck(self.lib.care_minhash_signatures(self.ctx, first, n, out.ctypes.data))
        return out

    def query(self, first, n, capacity=None):
        cap = capacity if capacity is not None else max(1, n) * 64
        while True:
            ids = np.zeros(cap, dtype=np.uint32)
            offs = np.zeros(n + 1, dtype=np.int32)
            rc = self.lib.care_minhash_query(self.ctx, first, n, ids.ctypes.data, offs.ctypes.data, cap)
            if rc == 4 and capacity is None:
                cap *= 4
                continue
            self._ck(rc)
            return ids[: offs[n]], offs

    def correct(self, first, n, candidate_capacity=None, out=None, noalloc=False, graph=False):
        """Corrects anchors [first, first+n).  out: arrays reused across
        calls, (status, seqs) or (status, seqs, cand_ids, cand_seqs):
        uint8 [n], uint8 [n][pitch], uint32 [cap], uint8 [cap][pitch].
        noalloc: through care_correct_batch_noalloc (no device allocation);
        graph: through care_correct_batch_graph (out in page-locked memory,
        pinned_empty)."""
        pitch = (self.max_length + 15) // 16 * 16
        ccap = candidate_capacity if candidate_capacity is not None else (n * 64 if self.options.correct_candidates else 0)
        if out is not None:
            status, seqs = out[0], out[1]
            assert status.shape[0] >= n and seqs.shape[0] >= n and seqs.shape[1] == pitch
        else:
            status = np.zeros(n, dtype=np.uint8)
            seqs = np.zeros((n, pitch), dtype=np.uint8)
        if out is not None and len(out) == 4:
            cids, cseqs = out[2], out[3]
            ccap = min(ccap, cids.shape[0], cseqs.shape[0])
            assert cseqs.shape[1] == pitch
        else:
            cids = np.zeros(max(ccap, 1), dtype=np.uint32)
            cseqs = np.zeros((max(ccap, 1), pitch), dtype=np.uint8)
        r = BatchResult()
        r.anchor_status = status.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
        r.anchor_sequences = seqs.ctypes.data_as(ctypes.POINTER(ctypes.c_char))
        r.sequence_pitch = pitch
        r.candidate_read_ids = cids.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
        r.candidate_sequences = cseqs.ctypes.data_as(ctypes.POINTER(ctypes.c_char))
        r.candidate_capacity = ccap
        fn = (self.lib.care_correct_batch_noalloc if noalloc else
              self.lib.care_correct_batch_graph if graph else self.lib.care_correct_batch)
        self._ck(fn(self.ctx, first, n, ctypes.byref(r)))
        nc = r.num_candidates
        return status, seqs, cids[:nc], cseqs[:nc]

    def stats(self):
        s = BatchStats()
        self._ck(self.lib.care_last_batch_stats(self.ctx, ctypes.byref(s)))
        return s

    # ---- device buffers and streams (stream-ordered interface) ----
    def device_array(self, nbytes):
        """A device allocation of the context's device; freed by device_free."""
        p = ctypes.c_void_p()
        self._ck(self.lib.care_device_alloc(self.ctx, int(nbytes), ctypes.byref(p)))
        return p.value

    def device_free(self, p):
        self._ck(self.lib.care_device_free(self.ctx, p))

    def to_device(self, arr):
        a = np.ascontiguousarray(arr)
        p = self.device_array(max(1, a.nbytes))
        self._ck(self.lib.care_device_copy(self.ctx, p, a.ctypes.data, a.nbytes, 1))
        return p

    def from_device(self, p, shape, dtype):
        out = np.zeros(shape, dtype=dtype)
        self._ck(self.lib.care_device_copy(self.ctx, out.ctypes.data, p, out.nbytes, 0))
        return out

    def stream_create(self):
        s = ctypes.c_void_p()
        self._ck(self.lib.care_stream_create(self.ctx, ctypes.byref(s)))
        return s.value

    def stream_destroy(self, s):
        self._ck(self.lib.care_stream_destroy(self.ctx, s))

    def stream_sync(self, s):
        self._ck(self.lib.care_stream_synchronize(self.ctx, s))

    # ---- learned gate ----
    def mlp_load(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.shape == (MLP_PARAMS,)
        self._ck(self.lib.care_mlp_load(self.ctx, p.ctypes.data))

    def mlp_gradient(self, features, labels, params):
        x = np.ascontiguousarray(features, dtype=np.float32)
        y = np.ascontiguousarray(labels, dtype=np.uint8)
        p = np.ascontiguousarray(params, dtype=np.float32)
        g = np.zeros(MLP_PARAMS, dtype=np.float32)
        loss = ctypes.c_float()
        self._ck(self.lib.care_mlp_gradient(self.ctx, x.ctypes.data, y.ctypes.data, len(y), p.ctypes.data,
                                          g.ctypes.data, ctypes.byref(loss)))
        return g, loss.value

    def mlp_train(self, features, labels, params, epochs=5, batch=4096, lr=1e-3, beta1=0.9, beta2=0.999,
                  eps=1e-8, seed=1):
        x = np.ascontiguousarray(features, dtype=np.float32)
        y = np.ascontiguousarray(labels, dtype=np.uint8)
        p = np.array(params, dtype=np.float32)
        o = MlpTrainOptions(epochs, batch, lr, beta1, beta2, eps, seed)
        losses = np.zeros(epochs, dtype=np.float32)
        self._ck(self.lib.care_mlp_train(self.ctx, x.ctypes.data, y.ctypes.data, len(y), ctypes.byref(o),
                                       p.ctypes.data, losses.ctypes.data))
        return p, losses

    def mlp_infer(self, features, params):
        x = np.ascontiguousarray(features, dtype=np.float32)
        p = np.ascontiguousarray(params, dtype=np.float32)
        z = np.zeros(len(x), dtype=np.float32)
        self._ck(self.lib.care_mlp_infer(self.ctx, x.ctypes.data, len(x), p.ctypes.data, z.ctypes.data))
        return z

    def mlp_load_candidates(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.shape == (MLP_PARAMS,)
        self._ck(self.lib.care_mlp_load_candidates(self.ctx, p.ctypes.data))

    def gate_columns(self, first, n, capacity=None, candidates=False, plain=False):
        """Decision columns of the gated mode: features [c][32], read ids,
        positions, original and consensus bases (candidates: those of the
        candidate corrections, forward-strand positions and bases; plain:
        over the plain method's MSAs, every candidate and refinement)."""
        fn = (self.lib.care_gate_columns_candidates if candidates else
              self.lib.care_gate_columns_plain if plain else self.lib.care_gate_columns)
        cap = capacity if capacity is not None else max(1, n) * 8
        while True:
            f = np.zeros((cap, MLP_FEATURES), dtype=np.float32)
            ids = np.zeros(cap, dtype=np.uint32)
            pos = np.zeros(cap, dtype=np.uint16)
            orig = np.zeros(cap, dtype=np.uint8)
            cons = np.zeros(cap, dtype=np.uint8)
            cnt = ctypes.c_uint64()
            rc = fn(self.ctx, first, n, f.ctypes.data, ids.ctypes.data, pos.ctypes.data, orig.ctypes.data,
                    cons.ctypes.data, cap, ctypes.byref(cnt))
            if rc == 4 and capacity is None:
                cap *= 4
                continue
            self._ck(rc)
            c = cnt.value
            return f[:c], ids[:c], pos[:c], orig[:c], cons[:c]


def run_care_gpu(args, **kw):
    """Runs the care-gpu command line (a child process)."""
    if not os.path.exists(CARE_GPU):
        raise ImportError("care-gpu not built (%s missing)" % CARE_GPU)
    return subprocess.run([CARE_GPU] + list(args), **kw)


def simulate(profile, seed, out, truth, genome=None, coverage=None, length=None):
    """Seeded synthetic reads (care_amd/sim/readsim.cpp)."""
    cmd = [READSIM, "gen", "--profile", profile, "--seed", str(seed), "--out", out, "--truth", truth]
    if genome is not None:
        cmd += ["--genome", str(int(genome))]
    if coverage is not None:
        cmd += ["--coverage", str(coverage)]
    if length is not None:
        cmd += ["--length", str(int(length))]
    subprocess.run(cmd, check=True)


def evaluate(orig, corrected, truth):
    """Base-level TP/FP/FN/TN (evaluationtool/arteval.cpp:183-227 definitions)."""
    import json
    p = subprocess.run([READSIM, "eval", "--orig", orig, "--corr", corrected, "--truth", truth],
                       check=True, stdout=subprocess.PIPE, text=True)
    return json.loads(p.stdout)
