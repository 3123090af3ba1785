"""Python view of libparasail_hip.so (the hip family of parasail on MI355X).

The library's C ABI is declared in include/parasail_hip.h.  This module only
loads it through ctypes and wraps a few calls for tests and the benchmark;
there is no Python or CPU fallback: when the shared library or the device is
missing, calls raise.
"""
import ctypes
import time
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libparasail_hip.so")

FLAG_NW, FLAG_SG, FLAG_SW = 1 << 0, 1 << 1, 1 << 2
FLAG_SG_S1_BEG, FLAG_SG_S1_END = 1 << 3, 1 << 4
FLAG_SATURATED = 1 << 6
FLAG_BANDED = 1 << 7
FLAG_SCAN, FLAG_STRIPED = 1 << 10, 1 << 11
FLAG_SG_S2_BEG, FLAG_SG_S2_END = 1 << 14, 1 << 15
FLAG_STATS, FLAG_TRACE = 1 << 16, 1 << 19
FLAG_BITS_8, FLAG_BITS_16, FLAG_BITS_32 = 1 << 20, 1 << 21, 1 << 22
FLAG_LANES_64 = 1 << 30

ALGS = ["sw", "nw", "sg", "sg_qb", "sg_qe", "sg_qx", "sg_db", "sg_de", "sg_dx",
        "sg_qb_de", "sg_qe_db", "sg_qb_db", "sg_qe_de"]
# (s1_beg, s1_end, s2_beg, s2_end) of every sg name (src/sg_helper.h:53-63)
SG_FLAGS = {"sg": (1, 1, 1, 1), "sg_qb": (1, 0, 0, 0), "sg_qe": (0, 1, 0, 0),
            "sg_qx": (1, 1, 0, 0), "sg_db": (0, 0, 1, 0), "sg_de": (0, 0, 0, 1),
            "sg_dx": (0, 0, 1, 1), "sg_qb_de": (1, 0, 0, 1), "sg_qe_db": (0, 1, 1, 0),
            "sg_qb_db": (1, 0, 1, 0), "sg_qe_de": (0, 1, 0, 1)}


class Result(ctypes.Structure):
    _fields_ = [("score", ctypes.c_int), ("end_query", ctypes.c_int),
                ("end_ref", ctypes.c_int), ("flag", ctypes.c_int),
                ("extra", ctypes.c_void_p)]


class StatsExtra(ctypes.Structure):
    _fields_ = [("matches", ctypes.c_int), ("similar", ctypes.c_int),
                ("length", ctypes.c_int), ("extra", ctypes.c_void_p)]


class TraceExtra(ctypes.Structure):
    _fields_ = [("trace_table", ctypes.c_void_p), ("trace_ins_table", ctypes.c_void_p),
                ("trace_del_table", ctypes.c_void_p)]


class Matrix(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("matrix", ctypes.POINTER(ctypes.c_int)),
                ("mapper", ctypes.POINTER(ctypes.c_int)), ("size", ctypes.c_int),
                ("max", ctypes.c_int), ("min", ctypes.c_int),
                ("user_matrix", ctypes.POINTER(ctypes.c_int)), ("type", ctypes.c_int),
                ("length", ctypes.c_int), ("alphabet", ctypes.c_char_p),
                ("query", ctypes.c_char_p)]


class Cigar(ctypes.Structure):
    _fields_ = [("seq", ctypes.POINTER(ctypes.c_uint32)), ("len", ctypes.c_int),
                ("beg_query", ctypes.c_int), ("beg_ref", ctypes.c_int)]


class Hit(ctypes.Structure):
    _fields_ = [("score", ctypes.c_int), ("end_query", ctypes.c_int), ("end_ref", ctypes.c_int),
                ("target", ctypes.c_int)]

    def as_tuple(self):
        return (self.score, self.end_query, self.end_ref, self.target)


class BatchResult(ctypes.Structure):
    _fields_ = [("score", ctypes.POINTER(ctypes.c_int)), ("end_query", ctypes.POINTER(ctypes.c_int)),
                ("end_ref", ctypes.POINTER(ctypes.c_int)), ("matches", ctypes.POINTER(ctypes.c_int)),
                ("similar", ctypes.POINTER(ctypes.c_int)), ("length", ctypes.POINTER(ctypes.c_int)),
                ("flag", ctypes.POINTER(ctypes.c_int)),
                ("cigars", ctypes.POINTER(ctypes.POINTER(Cigar)))]


RP = ctypes.POINTER(Result)
MP = ctypes.POINTER(Matrix)
FN = ctypes.CFUNCTYPE(RP, ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                      ctypes.c_int, ctypes.c_int, MP)

_lib = None


def lib():
    """Load libparasail_hip.so; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libparasail_hip.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        L.parasail_result_free.argtypes = [RP]
        L.parasail_lookup_function.restype = ctypes.c_void_p
        L.parasail_lookup_function.argtypes = [ctypes.c_char_p]
        L.parasail_lookup_pfunction.restype = ctypes.c_void_p
        L.parasail_lookup_pfunction.argtypes = [ctypes.c_char_p]
        L.parasail_lookup_pcreator.restype = ctypes.c_void_p
        L.parasail_lookup_pcreator.argtypes = [ctypes.c_char_p]
        L.parasail_matrix_lookup.restype = MP
        L.parasail_matrix_lookup.argtypes = [ctypes.c_char_p]
        L.parasail_matrix_create.restype = MP
        L.parasail_matrix_create.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
        L.parasail_matrix_pssm_create.restype = MP
        L.parasail_matrix_pssm_create.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int),
                                                  ctypes.c_int]
        L.parasail_matrix_convert_square_to_pssm.restype = MP
        L.parasail_matrix_convert_square_to_pssm.argtypes = [MP, ctypes.c_char_p, ctypes.c_int]
        L.parasail_matrix_free.argtypes = [MP]
        L.parasail_result_get_cigar.restype = ctypes.POINTER(Cigar)
        L.parasail_result_get_cigar.argtypes = [RP, ctypes.c_char_p, ctypes.c_int,
                                                ctypes.c_char_p, ctypes.c_int, MP]
        L.parasail_cigar_decode.restype = ctypes.c_void_p
        L.parasail_cigar_decode.argtypes = [ctypes.POINTER(Cigar)]
        L.parasail_cigar_free.argtypes = [ctypes.POINTER(Cigar)]
        L.parasail_nw_banded_hip.restype = RP
        L.parasail_nw_banded_hip.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p,
                                             ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_int, MP]
        L.parasail_hip_align_batch.restype = ctypes.c_int
        L.parasail_hip_align_batch.argtypes = [
            ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p,
            ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int), ctypes.c_char_p,
            ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int), ctypes.c_int,
            ctypes.c_int, ctypes.c_int, MP, ctypes.POINTER(BatchResult)]
        L.parasail_hip_last_kernel_ms.restype = ctypes.c_double
        I64, I32 = ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int)
        L.parasail_hip_db_create.restype = ctypes.c_void_p
        L.parasail_hip_db_create.argtypes = [ctypes.c_char_p, I64, I32, ctypes.c_int, MP]
        L.parasail_hip_db_free.argtypes = [ctypes.c_void_p]
        L.parasail_hip_db_search.restype = ctypes.c_int
        L.parasail_hip_db_search.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, MP,
                                             ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.POINTER(Hit)]
        L.parasail_hip_server_create.restype = ctypes.c_void_p
        L.parasail_hip_server_create.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
        L.parasail_hip_server_submit.restype = ctypes.c_int
        L.parasail_hip_server_submit.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_char_p,
                                                 ctypes.c_int, MP, ctypes.c_int, ctypes.c_int,
                                                 ctypes.c_int, ctypes.c_double]
        L.parasail_hip_server_poll.restype = ctypes.c_int
        L.parasail_hip_server_poll.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64),
                                               ctypes.POINTER(Hit), ctypes.POINTER(ctypes.c_int),
                                               ctypes.POINTER(ctypes.c_double), ctypes.c_int]
        L.parasail_hip_server_free.argtypes = [ctypes.c_void_p]
        L.parasail_hip_server_stats.restype = ctypes.c_int
        L.parasail_hip_server_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64),
                                                ctypes.POINTER(ctypes.c_int64),
                                                ctypes.POINTER(ctypes.c_int)]
        L.parasail_hip_nw_banded_set_segment.restype = ctypes.c_int
        L.parasail_hip_nw_banded_set_segment.argtypes = [ctypes.c_int]
        L.parasail_hip_now.restype = ctypes.c_double
        L.parasail_hip_last_main_kernel_ms.restype = ctypes.c_double
        L.parasail_hip_last_error.restype = ctypes.c_char_p
        L.parasail_hip_available.restype = ctypes.c_int
        L.parasail_time.restype = ctypes.c_double
        libc = ctypes.CDLL(None)
        libc.free.argtypes = [ctypes.c_void_p]
        L._libc_free = libc.free
        _lib = L
    return _lib


def available():
    return bool(lib().parasail_hip_available())


def blosum62():
    return lib().parasail_matrix_lookup(b"blosum62")


def matrix_create(alphabet, match, mismatch):
    return lib().parasail_matrix_create(alphabet.encode(), match, mismatch)


def function(name):
    ptr = lib().parasail_lookup_function(name.encode())
    if not ptr:
        raise KeyError(name)
    return FN(ptr)


def _b(s):
    return s if isinstance(s, bytes) else s.encode()


def align(name, s1, s2, open_, gap, matrix):
    """Call one per-pair drop-in by name; returns a dict with the result
    fields (and 'cigar' for trace functions)."""
    f = function(name)
    a, b = _b(s1) if s1 is not None else None, _b(s2)
    r = f(a, 0 if a is None else len(a), b, len(b), open_, gap, matrix)
    if not r:
        return None
    res = r.contents
    out = {"score": res.score, "end_query": res.end_query, "end_ref": res.end_ref,
           "flag": res.flag}
    if res.flag & FLAG_STATS:
        st = ctypes.cast(res.extra, ctypes.POINTER(StatsExtra)).contents
        out.update(matches=st.matches, similar=st.similar, length=st.length)
    if res.flag & FLAG_TRACE and not res.flag & FLAG_SATURATED:
        c = lib().parasail_result_get_cigar(r, a, len(a), b, len(b), matrix)
        out["cigar"] = cigar_tuple(c)
    lib().parasail_result_free(r)
    return out


def cigar_tuple(c):
    """(cigar string, beg_query, beg_ref) from a parasail_cigar_t*, freed;
    the string is parasail_cigar_decode's."""
    if not c:
        return None
    cc = c.contents
    p = lib().parasail_cigar_decode(c)
    s = ctypes.string_at(p).decode() if p else ""
    if p:
        lib()._libc_free(p)
    out = (s, cc.beg_query, cc.beg_ref)
    lib().parasail_cigar_free(c)
    return out


_last_call_s = 0.0


def last_call_ms():
    """Wall time of the last align_arrays / align_batch library call (ms),
    without the conversion of its results in Python."""
    return _last_call_s * 1e3


def align_arrays(alg, kind, width, batch, open_, gap, matrix, want_cigar=False):
    """Batched alignment of a workloads.PairBatch (concatenated residues with
    offsets and lengths) through parasail_hip_align_batch; results as numpy
    arrays (and, for traceback with want_cigar, a list of CIGAR tuples)."""
    n = len(batch)
    res = {k: np.zeros(n, dtype=np.int32) for k in
           ("score", "end_query", "end_ref", "matches", "similar", "length", "flag")}
    br = BatchResult()
    for k in res:
        setattr(br, k, res[k].ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    cig = None
    if want_cigar:
        cig = (ctypes.POINTER(Cigar) * max(n, 1))()
        br.cigars = ctypes.cast(cig, ctypes.POINTER(ctypes.POINTER(Cigar)))
    I64 = ctypes.POINTER(ctypes.c_int64)
    I32 = ctypes.POINTER(ctypes.c_int)
    qoff = np.ascontiguousarray(batch.qoff, dtype=np.int64)
    toff = np.ascontiguousarray(batch.toff, dtype=np.int64)
    qlen = np.ascontiguousarray(batch.qlen, dtype=np.int32)
    tlen = np.ascontiguousarray(batch.tlen, dtype=np.int32)
    global _last_call_s
    t0 = time.perf_counter()
    rc = lib().parasail_hip_align_batch(
        alg.encode(), kind, width, batch.q, qoff.ctypes.data_as(I64), qlen.ctypes.data_as(I32),
        batch.t, toff.ctypes.data_as(I64), tlen.ctypes.data_as(I32), n, open_, gap, matrix,
        ctypes.byref(br))
    _last_call_s = time.perf_counter() - t0
    if rc != 0:
        raise RuntimeError("parasail_hip_align_batch failed: %s"
                           % lib().parasail_hip_last_error().decode(errors="replace"))
    if want_cigar:
        res["cigar"] = [cigar_tuple(cig[k]) for k in range(n)]
    return res


def align_batch(alg, kind, width, queries, targets, open_, gap, matrix, want_cigar=False):
    """Batched alignment through parasail_hip_align_batch.  queries/targets
    are lists of str/bytes (queries may be None for a PSSM).  Returns a dict
    of numpy arrays (and a list of cigars for kind 2)."""
    n = len(targets)
    tb = [_b(t) for t in targets]
    tl = np.array([len(t) for t in tb], dtype=np.int32)
    to = np.zeros(n, dtype=np.int64)
    if n:
        to[1:] = np.cumsum(tl, dtype=np.int64)[:-1]
    tcat = b"".join(tb)
    if queries is not None:
        qb = [_b(q) for q in queries]
        ql = np.array([len(q) for q in qb], dtype=np.int32)
        qo = np.zeros(n, dtype=np.int64)
        if n:
            qo[1:] = np.cumsum(ql, dtype=np.int64)[:-1]
        qcat = b"".join(qb)
    res = {k: np.zeros(n, dtype=np.int32) for k in
           ("score", "end_query", "end_ref", "matches", "similar", "length", "flag")}
    br = BatchResult()
    for k in res:
        setattr(br, k, res[k].ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    cig = None
    if want_cigar:
        cig = (ctypes.POINTER(Cigar) * max(n, 1))()
        br.cigars = ctypes.cast(cig, ctypes.POINTER(ctypes.POINTER(Cigar)))
    I64 = ctypes.POINTER(ctypes.c_int64)
    I32 = ctypes.POINTER(ctypes.c_int)
    rc = lib().parasail_hip_align_batch(
        alg.encode(), kind, width,
        qcat if queries is not None else None,
        qo.ctypes.data_as(I64) if queries is not None else None,
        ql.ctypes.data_as(I32) if queries is not None else None,
        tcat, to.ctypes.data_as(I64), tl.ctypes.data_as(I32), n, open_, gap, matrix,
        ctypes.byref(br))
    if rc != 0:
        raise RuntimeError("parasail_hip_align_batch failed: %s"
                           % lib().parasail_hip_last_error().decode(errors="replace"))
    if want_cigar:
        res["cigar"] = [cigar_tuple(cig[k]) for k in range(n)]
    return res


def last_kernel_ms():
    """Device time (ms) of the last call's kernels."""
    return lib().parasail_hip_last_kernel_ms()


def last_main_kernel_ms():
    """Device time (ms) of the last call's dominant kernel."""
    return lib().parasail_hip_last_main_kernel_ms()


class Database:
    """An HBM-resident target database (parasail_hip_db_*): a
    workloads.PairBatch-like object's targets, or concatenated residues with
    offsets and lengths."""

    def __init__(self, residues, offsets, lengths, matrix):
        self._off = np.ascontiguousarray(offsets, dtype=np.int64)
        self._len = np.ascontiguousarray(lengths, dtype=np.int32)
        self.n = len(self._len)
        self.cells_per_row = int(self._len.astype(np.int64).sum())
        self.h = lib().parasail_hip_db_create(
            residues, self._off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
            self._len.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), self.n, matrix)
        if not self.h:
            raise RuntimeError("parasail_hip_db_create failed: %s"
                               % lib().parasail_hip_last_error().decode(errors="replace"))

    def search(self, query, matrix, open_, gap, k):
        """SW of query (str/bytes, or None with a PSSM) against every target:
        list of (score, end_query, end_ref, target), best first."""
        hits = (Hit * k)()
        q = _b(query) if query is not None else None
        m = len(q) if q is not None else 0
        n = lib().parasail_hip_db_search(self.h, q, m, matrix, open_, gap, k, hits)
        if n < 0:
            raise RuntimeError("parasail_hip_db_search failed: %s"
                               % lib().parasail_hip_last_error().decode(errors="replace"))
        return [hits[i].as_tuple() for i in range(n)]

    def close(self):
        if self.h:
            lib().parasail_hip_db_free(self.h)
            self.h = None


class Server:
    """The serving loop over a Database (parasail_hip_server_*)."""

    def __init__(self, db, mem_budget=0):
        self.db = db
        self.h = lib().parasail_hip_server_create(db.h, mem_budget)
        if not self.h:
            raise RuntimeError("parasail_hip_server_create failed")

    def submit(self, rid, query, matrix, open_, gap, k, arrival=0.0):
        q = _b(query) if query is not None else None
        m = len(q) if q is not None else 0
        if lib().parasail_hip_server_submit(self.h, rid, q, m, matrix, open_, gap, k, arrival) != 0:
            raise RuntimeError("parasail_hip_server_submit failed")

    def poll(self, k, wait=True):
        """(id, hits, latency_ms) of the next finished request, or None."""
        rid = ctypes.c_int64()
        hits = (Hit * k)()
        nh = ctypes.c_int()
        lat = ctypes.c_double()
        rc = lib().parasail_hip_server_poll(self.h, ctypes.byref(rid), hits, ctypes.byref(nh),
                                            ctypes.byref(lat), 1 if wait else 0)
        if rc == 0:
            return None
        if rc < 0:
            raise RuntimeError("request %d failed" % rid.value)
        return rid.value, [hits[i].as_tuple() for i in range(nh.value)], lat.value

    def stats(self):
        """(completed, deferred at admission, most in flight at once)."""
        c, d, a = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int()
        lib().parasail_hip_server_stats(self.h, ctypes.byref(c), ctypes.byref(d), ctypes.byref(a))
        return c.value, d.value, a.value

    def close(self):
        if self.h:
            lib().parasail_hip_server_free(self.h)
            self.h = None


def now():
    """The serving loop's clock (seconds)."""
    return lib().parasail_hip_now()


PSSM_ALPHABET = b"ARNDCQEGHILKMFPSTWYVBZX"


def blosum62_pssm(query, noise=None):
    """A PSSM (parasail_matrix_pssm_create) whose row i is the BLOSUM62 row
    of query[i], plus noise[i] (ints, len(query) x 23) when given, over the
    alphabet ARNDCQEGHILKMFPSTWYVBZX (the BLOSUM62 mapper).  Returns the
    matrix and the values array, which must outlive nothing (the matrix
    copies them) but is returned for inspection."""
    b = blosum62().contents
    q = _b(query)
    idx = np.array([b.mapper[c] for c in PSSM_ALPHABET])
    tab = np.ctypeslib.as_array(b.matrix, shape=(b.size * b.size,)).reshape(b.size, b.size)
    rows = np.array([b.mapper[c] for c in q])
    vals = tab[rows][:, idx].astype(np.int32)
    if noise is not None:
        vals = vals + np.asarray(noise, dtype=np.int32)
    v = np.ascontiguousarray(vals.reshape(-1))
    m = lib().parasail_matrix_pssm_create(PSSM_ALPHABET,
                                          v.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), len(q))
    if not m:
        raise RuntimeError("parasail_matrix_pssm_create failed")
    return m, v


def nw_banded_batch(queries, targets, open_, gap, k, matrix):
    """parasail_hip_nw_banded_batch over lists of str/bytes: score,
    end_query, end_ref and flag arrays."""
    n = len(queries)
    qb = [_b(q) for q in queries]
    tb = [_b(t) for t in targets]
    ql = np.array([len(x) for x in qb], dtype=np.int32)
    tl = np.array([len(x) for x in tb], dtype=np.int32)
    qo = np.zeros(n, dtype=np.int64)
    to = np.zeros(n, dtype=np.int64)
    if n:
        qo[1:] = np.cumsum(ql, dtype=np.int64)[:-1]
        to[1:] = np.cumsum(tl, dtype=np.int64)[:-1]
    return _nw_banded(b"".join(qb), qo, ql, b"".join(tb), to, tl, open_, gap, k, matrix)


def nw_banded_arrays(batch, open_, gap, k, matrix):
    """parasail_hip_nw_banded_batch over a workloads.PairBatch."""
    return _nw_banded(batch.q, np.ascontiguousarray(batch.qoff, dtype=np.int64),
                      np.ascontiguousarray(batch.qlen, dtype=np.int32), batch.t,
                      np.ascontiguousarray(batch.toff, dtype=np.int64),
                      np.ascontiguousarray(batch.tlen, dtype=np.int32), open_, gap, k, matrix)


def _nw_banded(qcat, qo, ql, tcat, to, tl, open_, gap, k, matrix):
    n = len(ql)
    res = {k_: np.zeros(n, dtype=np.int32) for k_ in ("score", "end_query", "end_ref", "flag")}
    br = BatchResult()
    for k_, v in res.items():
        setattr(br, k_, v.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    I64, I32 = ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int)
    L = lib()
    L.parasail_hip_nw_banded_batch.restype = ctypes.c_int
    L.parasail_hip_nw_banded_batch.argtypes = [ctypes.c_char_p, I64, I32, ctypes.c_char_p, I64, I32,
                                               ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_int, MP, ctypes.POINTER(BatchResult)]
    rc = L.parasail_hip_nw_banded_batch(qcat, qo.ctypes.data_as(I64), ql.ctypes.data_as(I32),
                                        tcat, to.ctypes.data_as(I64), tl.ctypes.data_as(I32),
                                        n, open_, gap, k, matrix, ctypes.byref(br))
    if rc != 0:
        raise RuntimeError("parasail_hip_nw_banded_batch failed: %s"
                           % lib().parasail_hip_last_error().decode(errors="replace"))
    return res


def nw_banded_set_segment(columns):
    """Columns a segment of a long banded pair owns (0: no segments);
    returns the previous setting."""
    L = lib()
    L.parasail_hip_nw_banded_set_segment.restype = ctypes.c_int
    L.parasail_hip_nw_banded_set_segment.argtypes = [ctypes.c_int]
    return L.parasail_hip_nw_banded_set_segment(int(columns))


def nw_banded_last_stats():
    """(segmented pairs, segments, pairs aligned again whole) of the last
    banded call."""
    v = [ctypes.c_int(0) for _ in range(3)]
    lib().parasail_hip_nw_banded_last_stats(*[ctypes.byref(x) for x in v])
    return tuple(x.value for x in v)
