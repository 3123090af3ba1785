"""Database search and the serving loop (parasail_hip_db_*, parasail_hip_server_*)
against the CPU oracle.

A search is SW of one query against every target of an HBM-resident
database; its hits must equal the oracle's parasail_sw results sorted by
score (desc) then target index (asc), bit for bit (score, end_query,
end_ref, target).  The server must return, for every request of a mix that
runs concurrently, exactly what the same request returns alone.
"""
import ctypes
import random

import numpy as np
import pytest

import oracle as O
import parasail_amd as P
from parasail_amd import workloads as WL

pytestmark = pytest.mark.gpu

AA = "ARNDCQEGHILKMFPSTWYV"
PSSM_ALPHA = b"ARNDCQEGHILKMFPSTWYVBZX"


@pytest.fixture(scope="module", autouse=True)
def need_device():
    if not P.available():
        pytest.skip("no HIP device")


@pytest.fixture(scope="module")
def db_case():
    """1,500 targets: log-normal lengths around 350, a few longer than 2,000
    (whole-wave chains), one W-rich target whose self score leaves int16."""
    rng = np.random.default_rng(21)
    lens = WL.lognormal_lengths(rng, 1500, 350, 0.7, 30, 1500)
    lens[[7, 300, 901]] = [2100, 3500, 2600]
    seqs = [WL.protein_bytes(rng, int(n)) for n in lens]
    r5 = random.Random(5)
    wrich = bytes(r5.choice(b"WWWWWWWWWC") for _ in range(3100))
    seqs[1200] = wrich
    lens[1200] = len(wrich)
    res = b"".join(seqs)
    off = np.zeros(len(seqs), dtype=np.int64)
    off[1:] = np.cumsum([len(s) for s in seqs])[:-1]
    pm = P.blosum62()
    db = P.Database(res, off, [len(s) for s in seqs], pm)
    yield {"db": db, "seqs": seqs, "pm": pm, "om": O.blosum62(), "wrich": wrich}
    db.close()


def oracle_hits(seqs, q, om, k, pssm=None):
    if pssm is None:
        qm = om.map(q)
        res, _ = O.batch(O.SW, O.SCORE, [qm] * len(seqs), [om.map(t) for t in seqs], 11, 1, om)
        rows = [(int(res[i, 0]), int(res[i, 1]), int(res[i, 2]), i) for i in range(len(seqs))]
    else:
        rows = []
        for i, t in enumerate(seqs):
            r, _ = O.align(O.SW, O.SCORE, None, t, 11, 1, pssm)
            rows.append((r[0], r[1], r[2], i))
    rows.sort(key=lambda x: (-x[0], x[3]))
    return rows[:k]


@pytest.mark.parametrize("m", [20, 77, 300, 330, 1000])
def test_db_search_vs_oracle(db_case, m):
    rng = np.random.default_rng(100 + m)
    q = WL.protein_bytes(rng, m)
    got = db_case["db"].search(q, db_case["pm"], 11, 1, 50)
    exp = oracle_hits(db_case["seqs"], q, db_case["om"], 50)
    assert got == exp


def test_db_search_pssm_and_int32_tier(db_case):
    L = P.lib()
    pm, om = db_case["pm"], db_case["om"]
    # a PSSM from a query (src/memory.c:567-621) ranks like that query
    q = WL.protein_bytes(np.random.default_rng(7), 250)
    pp = L.parasail_matrix_convert_square_to_pssm(pm, q, len(q))
    op = O.convert_square_to_pssm(om, q)
    got = db_case["db"].search(None, pp, 11, 1, 20)
    assert got == oracle_hits(db_case["seqs"], None, om, 20, pssm=op)
    L.parasail_matrix_free(pp)
    # the W-rich self hit scores above the int16 window: exact int32 re-run
    w = db_case["wrich"]
    got = db_case["db"].search(w, pm, 11, 1, 5)
    exp = oracle_hits(db_case["seqs"], w, om, 5)
    assert got == exp and got[0][0] > 32767 and got[0][3] == 1200


def test_db_search_rejects_other_alphabets(db_case):
    dna = P.matrix_create("ACGT", 2, -3)
    with pytest.raises(RuntimeError):
        db_case["db"].search(b"ACGT", dna, 5, 2, 10)


def noisy_pssm(rng, q):
    """Per-position BLOSUM62 rows of q with uniform integer noise in [-2, 2]."""
    b = P.blosum62().contents
    vals = np.zeros((len(q), len(PSSM_ALPHA)), dtype=np.int32)
    for i, c in enumerate(q):
        row = b.mapper[c]
        for j, a in enumerate(PSSM_ALPHA):
            vals[i, j] = b.matrix[row * b.size + b.mapper[a]] + int(rng.integers(-2, 3))
    v = np.ascontiguousarray(vals.reshape(-1))
    return P.lib().parasail_matrix_pssm_create(
        PSSM_ALPHA, v.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), len(q)), v


def test_server_results_alone_equal_in_mix(db_case):
    """Eight requests in flight together (query lengths 40..1200, two PSSMs)
    each return exactly what they return searched alone."""
    rng = np.random.default_rng(33)
    db, pm = db_case["db"], db_case["pm"]
    reqs = []
    keep = []
    for i, m in enumerate([40, 300, 1200, 90, 510, 300, 700, 150]):
        q = WL.protein_bytes(rng, m)
        if i in (2, 5):
            mx, v = noisy_pssm(rng, q)
            keep.append(v)
            reqs.append((None, mx))
        else:
            reqs.append((q, pm))
    alone = [db.search(q, mx, 11, 1, 50) for q, mx in reqs]
    srv = P.Server(db)
    try:
        t0 = P.now()
        for i, (q, mx) in enumerate(reqs):
            srv.submit(i, q, mx, 11, 1, 50, t0)
        got = {}
        for _ in reqs:
            rid, hits, lat = srv.poll(50)
            assert lat > 0
            got[rid] = hits
        completed, deferred, max_active = srv.stats()
    finally:
        srv.close()
    for i in range(len(reqs)):
        assert got[i] == alone[i], i
    # the default budget holds them all: they ran side by side
    assert completed == len(reqs) and deferred == 0 and max_active > 1, (completed, deferred, max_active)
    for q, mx in reqs:
        if q is None:
            P.lib().parasail_matrix_free(mx)


def test_server_admission_waits_for_memory(db_case):
    """A memory budget smaller than two requests' footprints: each request
    waits at admission until the one in flight finishes (one at a time), and
    every request still returns what it returns searched alone."""
    rng = np.random.default_rng(34)
    db, pm = db_case["db"], db_case["pm"]
    # the first request is long, so the others arrive while it runs
    qs = [WL.protein_bytes(rng, m) for m in (4000, 60, 900, 130, 400)]
    alone = [db.search(q, pm, 11, 1, 20) for q in qs]
    for budget in (1, 0):
        srv = P.Server(db, mem_budget=budget)
        try:
            t0 = P.now()
            for i, q in enumerate(qs):
                srv.submit(i, q, pm, 11, 1, 20, t0)
            got, order = {}, []
            for _ in qs:
                rid, hits, _ = srv.poll(20)
                got[rid] = hits
                order.append(rid)
            completed, deferred, max_active = srv.stats()
        finally:
            srv.close()
        for i in range(len(qs)):
            assert got[i] == alone[i], (budget, i)
        assert completed == len(qs)
        if budget == 1:
            # one request on the device at a time, admitted oldest first
            assert max_active == 1 and deferred >= 1, (max_active, deferred)
            assert order == list(range(len(qs))), order
        else:
            assert deferred == 0 and 1 < max_active <= 4, (max_active, deferred)
