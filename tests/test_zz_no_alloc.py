"""parasail_hip_align_batch_noalloc, the entry point that allocates nothing,
against parasail_hip_align_batch at width 32 (the entry point it is a
variant of), through the C ABI: two calls in a row, the second smaller,
score and stats, on inputs already in device memory."""
import ctypes
import random

import numpy as np
import pytest

import parasail_amd as P

pytestmark = pytest.mark.gpu

AA = "ARNDCQEGHILKMFPSTWYV"


@pytest.fixture(scope="module", autouse=True)
def need_device():
    if not P.available():
        pytest.skip("no HIP device")


def _api():
    L = P.lib()
    I64, I32 = ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int)
    L.parasail_hip_align_batch_noalloc.restype = ctypes.c_int
    L.parasail_hip_align_batch_noalloc.argtypes = [
        ctypes.c_char_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
        ctypes.c_int, ctypes.c_int, P.MP, ctypes.POINTER(P.BatchResult), ctypes.c_void_p]
    return L, I64, I32


def _pairs(rng, n, lo, hi):
    qs = ["".join(rng.choice(AA) for _ in range(rng.randint(lo, hi))) for _ in range(n)]
    ts = ["".join(rng.choice(AA) for _ in range(rng.randint(lo, hi))) for _ in range(n)]
    for k in range(0, n, 7):
        ts[k] = qs[k][rng.randint(0, len(qs[k]) // 2):]  # related pairs
    return qs, ts


class _Hip:
    """The HIP runtime the library itself uses (libamdhip64.so.7, already
    loaded with it), for the test's own device buffers."""

    def __init__(self):
        P.lib()
        self.h = ctypes.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                     ctypes.c_int]
        self.h.hipFree.argtypes = [ctypes.c_void_p]
        self.live = []

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), max(arr.nbytes, 1)) == 0
        self.live.append(p.value)
        assert self.h.hipMemcpy(p, arr.ctypes.data, arr.nbytes, 1) == 0  # host to device
        return p.value

    def get(self, ptr, n, dtype=np.int32):
        out = np.zeros(n, dtype=dtype)
        assert self.h.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0  # device to host
        return out

    def sync(self):
        assert self.h.hipDeviceSynchronize() == 0

    def free(self):
        for p in self.live:
            self.h.hipFree(p)
        self.live = []


def _device_call(L, hip, alg, kind, qs, ts, max_len, pm):
    def cat(seqs):
        b = np.frombuffer("".join(seqs).encode(), dtype=np.uint8)
        ln = np.array([len(s) for s in seqs], dtype=np.int32)
        off = np.zeros(len(seqs), dtype=np.int64)
        off[1:] = np.cumsum(ln, dtype=np.int64)[:-1]
        return hip.put(b), hip.put(off), hip.put(ln)

    q, qo, ql = cat(qs)
    t, to, tl = cat(ts)
    n = len(qs)
    keys = ("score", "end_query", "end_ref", "matches", "similar", "length", "flag")
    outs = {k: hip.put(np.full(n, -7, dtype=np.int32)) for k in keys}
    br = P.BatchResult()
    for k, v in outs.items():
        setattr(br, k, ctypes.cast(ctypes.c_void_p(v), ctypes.POINTER(ctypes.c_int)))
    rc = L.parasail_hip_align_batch_noalloc(alg.encode(), kind, q, qo, ql, t, to, tl, n, max_len,
                                            11, 1, pm, ctypes.byref(br), None)
    hip.sync()
    got = {k: hip.get(v, n) for k, v in outs.items()}
    hip.free()
    return rc, got


def test_noalloc_equals_align_batch_width32():
    L, _, _ = _api()
    hip = _Hip()
    pm = P.blosum62()
    rng = random.Random(5)
    big = _pairs(rng, 700, 20, 900)
    small = _pairs(rng, 90, 5, 300)
    for alg in ("sw", "nw", "sg", "sg_qb_de"):
        for kind in (0, 1):
            for qs, ts in (big, small):
                max_len = max(len(t) for t in ts)
                rc, got = _device_call(L, hip, alg, kind, qs, ts, max_len, pm)
                assert rc == 0
                exp = P.align_batch(alg, kind, 32, qs, ts, 11, 1, pm)
                keys = ["score", "end_query", "end_ref", "flag"]
                if kind == 1:
                    keys += ["matches", "similar", "length"]
                for k in keys:
                    assert (got[k] == exp[k]).all(), (alg, kind, len(qs), k)


def test_noalloc_limits():
    L, _, _ = _api()
    hip = _Hip()
    pm = P.blosum62()
    rng = random.Random(6)
    qs, ts = _pairs(rng, 40, 10, 200)
    # a pair longer than the caller's bound is reported invalid, the others
    # are still computed
    ts[3] = "W" * 260
    rc, got = _device_call(L, hip, "sw", 0, qs, ts, 250, pm)
    assert rc == 0
    exp = P.align_batch("sw", 0, 32, qs, ts, 11, 1, pm)
    assert got["flag"][3] == ctypes.c_int(0x80000020).value
    ok = np.arange(len(qs)) != 3
    assert (got["score"][ok] == exp["score"][ok]).all()
    # bounds past the static scratch array fail before any device work
    rc, _ = _device_call(L, hip, "sw", 1, qs, ts, 3_000_000, pm)
    assert rc == -1
    rc, _ = _device_call(L, hip, "sw", 0, qs, ts, 9_000_000, pm)
    assert rc == -1
