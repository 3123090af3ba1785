"""ctypes view of the CPU oracle (oracle/oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the ``cpu_baseline`` leg of ``bench.py`` as the checker.  The product
library never imports or links it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

SW, NW, SG = 0, 1, 2
SCORE, STATS, TRACE = 0, 1, 2
NEG_INF = -(2 ** 31) // 2


class OrcResult(ctypes.Structure):
    _fields_ = [("score", ctypes.c_int), ("end_query", ctypes.c_int),
                ("end_ref", ctypes.c_int), ("matches", ctypes.c_int),
                ("similar", ctypes.c_int), ("length", ctypes.c_int)]

    def as_tuple(self):
        return (self.score, self.end_query, self.end_ref,
                self.matches, self.similar, self.length)


class OrcMatrix(ctypes.Structure):
    _fields_ = [("size", ctypes.c_int), ("length", ctypes.c_int),
                ("type", ctypes.c_int), ("max", ctypes.c_int),
                ("min", ctypes.c_int), ("mapper", ctypes.c_int * 256),
                ("matrix", ctypes.POINTER(ctypes.c_int)),
                ("alphabet", ctypes.c_char * 257)]


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        P = ctypes.POINTER
        _lib.orc_align.argtypes = [ctypes.c_int] * 6 + [
            P(ctypes.c_int), ctypes.c_int, P(ctypes.c_int), ctypes.c_int,
            ctypes.c_int, ctypes.c_int, P(OrcMatrix), P(OrcResult),
            P(ctypes.c_int8)]
        _lib.orc_align_tables.argtypes = [ctypes.c_int] * 6 + [
            P(ctypes.c_int), ctypes.c_int, P(ctypes.c_int), ctypes.c_int,
            ctypes.c_int, ctypes.c_int, P(OrcMatrix), P(OrcResult),
            P(ctypes.c_int8), P(P(ctypes.c_int))]
        _lib.orc_nw_banded.argtypes = [
            P(ctypes.c_int), ctypes.c_int, P(ctypes.c_int), ctypes.c_int,
            ctypes.c_int, ctypes.c_int, ctypes.c_int, P(OrcMatrix), P(OrcResult)]
        _lib.orc_cigar.argtypes = [
            ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
            ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
            ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
            ctypes.c_char_p, P(ctypes.c_uint32), ctypes.c_int, P(ctypes.c_int),
            P(ctypes.c_int)]
        _lib.orc_matrix_create.argtypes = [P(OrcMatrix), ctypes.c_char_p,
                                           ctypes.c_int, ctypes.c_int, ctypes.c_int]
        _lib.orc_matrix_pssm_create.argtypes = [P(OrcMatrix), ctypes.c_char_p,
                                                P(ctypes.c_int), ctypes.c_int,
                                                ctypes.c_int]
        _lib.orc_matrix_blosum62.argtypes = [P(OrcMatrix)]
        _lib.orc_matrix_square_from_values.argtypes = [
            P(OrcMatrix), ctypes.c_char_p, P(ctypes.c_int), ctypes.c_int,
            ctypes.c_int, ctypes.c_int]
        _lib.orc_matrix_convert_square_to_pssm.argtypes = [
            P(OrcMatrix), P(OrcMatrix), ctypes.c_char_p, ctypes.c_int]
        _lib.orc_matrix_free.argtypes = [P(OrcMatrix)]
        _lib.orc_striped_unwind.argtypes = [ctypes.c_void_p, ctypes.c_int,
                                            ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, P(ctypes.c_int)]
        _lib.orc_batch.restype = ctypes.c_double
        _lib.orc_batch.argtypes = [ctypes.c_int] * 7 + [
            P(ctypes.c_int), P(ctypes.c_int64), P(ctypes.c_int),
            P(ctypes.c_int), P(ctypes.c_int64), P(ctypes.c_int),
            ctypes.c_int, ctypes.c_int, P(OrcMatrix), P(OrcResult), ctypes.c_int]
        _lib.orc_striped_sat.argtypes = [ctypes.c_int] * 5 + [
            P(ctypes.c_int), ctypes.c_int, P(ctypes.c_int), ctypes.c_int,
            ctypes.c_int, ctypes.c_int, P(OrcMatrix), P(OrcResult)]
        _lib.orc_striped_batch.restype = ctypes.c_double
        _lib.orc_striped_batch.argtypes = [ctypes.c_int] * 6 + [
            P(ctypes.c_int), P(ctypes.c_int64), P(ctypes.c_int),
            P(ctypes.c_int), P(ctypes.c_int64), P(ctypes.c_int),
            ctypes.c_int, ctypes.c_int, P(OrcMatrix), P(OrcResult), P(ctypes.c_int),
            ctypes.c_int, ctypes.c_int]
    return _lib


class Matrix:
    """Owns an OrcMatrix; mirrors parasail_matrix_t's fields."""

    def __init__(self, m):
        self._m = m

    def __del__(self):
        try:
            lib().orc_matrix_free(ctypes.byref(self._m))
        except Exception:
            pass

    @property
    def ptr(self):
        return ctypes.byref(self._m)

    @property
    def size(self):
        return self._m.size

    @property
    def length(self):
        return self._m.length

    @property
    def is_pssm(self):
        return self._m.type == 1

    @property
    def max(self):
        return self._m.max

    @property
    def min(self):
        return self._m.min

    @property
    def mapper(self):
        return np.array(self._m.mapper[:], dtype=np.int32)

    def values(self):
        n = self._m.size * self._m.length
        return np.ctypeslib.as_array(self._m.matrix, shape=(n,)).reshape(
            self._m.length, self._m.size).copy()

    def map(self, s):
        b = s.encode() if isinstance(s, str) else bytes(s)
        mp = np.array(self._m.mapper[:], dtype=np.int32)
        return np.ascontiguousarray(mp[np.frombuffer(b, dtype=np.uint8)])


def blosum62():
    m = OrcMatrix()
    assert lib().orc_matrix_blosum62(ctypes.byref(m)) == 0
    return Matrix(m)


def matrix_from_values(alphabet, values, mx, mn):
    m = OrcMatrix()
    v = np.ascontiguousarray(values, dtype=np.int32)
    assert lib().orc_matrix_square_from_values(
        ctypes.byref(m), alphabet.encode(),
        v.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), len(alphabet), mx, mn) == 0
    return Matrix(m)


def matrix_create(alphabet, match, mismatch, case_sensitive=False):
    m = OrcMatrix()
    assert lib().orc_matrix_create(ctypes.byref(m), alphabet.encode(), match,
                                   mismatch, int(case_sensitive)) == 0
    return Matrix(m)


def matrix_pssm_create(alphabet, values, length, case_sensitive=False):
    m = OrcMatrix()
    v = np.ascontiguousarray(values, dtype=np.int32)
    assert lib().orc_matrix_pssm_create(
        ctypes.byref(m), alphabet.encode(),
        v.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), length,
        int(case_sensitive)) == 0
    return Matrix(m)


def convert_square_to_pssm(sq, s1):
    m = OrcMatrix()
    b = s1.encode() if isinstance(s1, str) else s1
    assert lib().orc_matrix_convert_square_to_pssm(ctypes.byref(m), sq.ptr, b,
                                                   len(b)) == 0
    return Matrix(m)


def _iptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def align(alg, kind, s1, s2, open_, gap, matrix, sg_flags=(0, 0, 0, 0)):
    """Run one serial alignment; s1/s2 are str or bytes (s1 may be None for
    a PSSM).  Returns (OrcResult tuple, trace ndarray or None)."""
    q = None if s1 is None else matrix.map(s1)
    t = matrix.map(s2)
    m = matrix.length if matrix.is_pssm else len(q)
    res = OrcResult()
    tr = np.zeros(m * len(t), dtype=np.int8) if kind == TRACE else None
    rc = lib().orc_align(alg, kind, *sg_flags,
                         None if q is None else _iptr(q), 0 if q is None else len(q),
                         _iptr(t), len(t), open_, gap, matrix.ptr,
                         ctypes.byref(res),
                         None if tr is None else tr.ctypes.data_as(
                             ctypes.POINTER(ctypes.c_int8)))
    if rc != 0:
        return None, None
    return res.as_tuple(), (None if tr is None else tr.reshape(m, len(t)))


def align_tables(alg, kind, s1, s2, open_, gap, matrix, sg_flags=(0, 0, 0, 0)):
    """orc_align with the PARASAIL_TABLE outputs: (result tuple, tables) where
    tables is an (m, n) int32 array of every cell's H, or for STATS a list of
    four (H, matches, similar, length)."""
    q = None if s1 is None else matrix.map(s1)
    t = matrix.map(s2)
    m = matrix.length if matrix.is_pssm else len(q)
    planes = 4 if kind == STATS else 1
    tabs = [np.zeros(m * len(t), dtype=np.int32) for _ in range(planes)]
    ptrs = (ctypes.POINTER(ctypes.c_int) * 4)(*([_iptr(x) for x in tabs] +
                                                 [None] * (4 - planes)))
    res = OrcResult()
    rc = lib().orc_align_tables(alg, kind, *sg_flags,
                                None if q is None else _iptr(q), 0 if q is None else len(q),
                                _iptr(t), len(t), open_, gap, matrix.ptr, ctypes.byref(res),
                                None, ptrs)
    if rc != 0:
        return None, None
    tabs = [x.reshape(m, len(t)) for x in tabs]
    return res.as_tuple(), (tabs if kind == STATS else tabs[0])


def nw_banded(s1, s2, open_, gap, k, matrix):
    q = matrix.map(s1)
    t = matrix.map(s2)
    res = OrcResult()
    rc = lib().orc_nw_banded(_iptr(q), len(q), _iptr(t), len(t), open_, gap, k,
                             matrix.ptr, ctypes.byref(res))
    return None if rc != 0 else res.as_tuple()


def cigar(seqA, seqB, is_sg, end_query, end_ref, trace, striped=False, bits=8,
          lanes=1, case_sensitive=False, aliases=None):
    a = seqA.encode() if isinstance(seqA, str) else seqA
    b = seqB.encode() if isinstance(seqB, str) else seqB
    cap = len(a) + len(b) + 8
    seq = np.zeros(cap, dtype=np.uint32)
    bq = ctypes.c_int()
    br = ctypes.c_int()
    tr = np.ascontiguousarray(trace)
    n = lib().orc_cigar(a, len(a), b, len(b), int(is_sg), end_query, end_ref,
                        tr.ctypes.data_as(ctypes.c_void_p), int(striped), bits,
                        lanes, int(case_sensitive),
                        None if aliases is None else aliases.encode(),
                        seq.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), cap,
                        ctypes.byref(bq), ctypes.byref(br))
    if n < 0:
        return None
    return seq[:n].copy(), bq.value, br.value


def cigar_string(ops):
    return "".join("%d%s" % (int(v) >> 4, "MIDNSHP=XB"[int(v) & 0xF]) for v in ops)


def batch(alg, kind, queries, targets, open_, gap, matrix, sg_flags=(0, 0, 0, 0),
          threads=0):
    """OpenMP-over-pairs serial oracle (test_openmp.c pattern).  queries and
    targets are lists of mapped int32 arrays.  Returns (results[n,6], seconds)."""
    n = len(queries)
    qlen = np.array([len(x) for x in queries], dtype=np.int32)
    tlen = np.array([len(x) for x in targets], dtype=np.int32)
    qoff = np.zeros(n, dtype=np.int64)
    toff = np.zeros(n, dtype=np.int64)
    if n:
        qoff[1:] = np.cumsum(qlen)[:-1]
        toff[1:] = np.cumsum(tlen)[:-1]
    q = np.ascontiguousarray(np.concatenate(queries) if n else np.zeros(1), dtype=np.int32)
    t = np.ascontiguousarray(np.concatenate(targets) if n else np.zeros(1), dtype=np.int32)
    out = (OrcResult * max(n, 1))()
    secs = lib().orc_batch(alg, kind, *sg_flags, n, _iptr(q),
                           qoff.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                           _iptr(qlen), _iptr(t),
                           toff.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                           _iptr(tlen), open_, gap, matrix.ptr, out, threads)
    res = np.array([out[k].as_tuple() for k in range(n)], dtype=np.int64).reshape(n, 6)
    return res, secs


def striped_available():
    """Whether the host has AVX2 for the striped baseline."""
    return bool(lib().orc_striped_available())


def striped_sat(alg, s1, s2, open_, gap, matrix, sg_flags=(0, 0, 0, 0)):
    """One pair through the AVX2 striped 8 -> 16 -> 32-bit chain (score
    only).  Returns ((score, end_query, end_ref), width that held)."""
    a = np.ascontiguousarray(matrix.map(s1) if isinstance(s1, (str, bytes)) else s1, dtype=np.int32)
    b = np.ascontiguousarray(matrix.map(s2) if isinstance(s2, (str, bytes)) else s2, dtype=np.int32)
    res = OrcResult()
    w = lib().orc_striped_sat(alg, *sg_flags, _iptr(a), len(a), _iptr(b), len(b), open_, gap,
                              matrix.ptr, ctypes.byref(res))
    if w < 0:
        return None, w
    return (res.score, res.end_query, res.end_ref), w


def striped_batch(alg, queries, targets, open_, gap, matrix, sg_flags=(0, 0, 0, 0), threads=0,
                  start=8):
    """The AVX2 striped chain from `start` bits (8, or 16 as the _16
    functions) over every pair, OpenMP over pairs.  queries and targets are
    lists of mapped int32 arrays.  Returns (results[n,3], widths[n],
    seconds)."""
    n = len(queries)
    qlen = np.array([len(x) for x in queries], dtype=np.int32)
    tlen = np.array([len(x) for x in targets], dtype=np.int32)
    qoff = np.zeros(n, dtype=np.int64)
    toff = np.zeros(n, dtype=np.int64)
    if n:
        qoff[1:] = np.cumsum(qlen)[:-1]
        toff[1:] = np.cumsum(tlen)[:-1]
    q = np.ascontiguousarray(np.concatenate(queries) if n else np.zeros(1), dtype=np.int32)
    t = np.ascontiguousarray(np.concatenate(targets) if n else np.zeros(1), dtype=np.int32)
    out = (OrcResult * max(n, 1))()
    width = np.zeros(max(n, 1), dtype=np.int32)
    secs = lib().orc_striped_batch(alg, *sg_flags, n, _iptr(q),
                                   qoff.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                   _iptr(qlen), _iptr(t),
                                   toff.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                   _iptr(tlen), open_, gap, matrix.ptr, out, _iptr(width),
                                   threads, start)
    res = np.array([out[k].as_tuple()[:3] for k in range(n)], dtype=np.int64).reshape(n, 3)
    return res, width[:n], secs
