"""The _table and _rowcol forms of the hip functions against the oracle's
DP tables, as tests/test_verify_tables.c and tests/test_verify_rowcols.c
compare every function's tables (score, and for _stats the matches, similar
and length tables) and last row/column with the serial ones, on the
reference's FASTA fixtures with its gap sets, plus random pairs past the
int32 lane-chain kernel's range (queries over 1,280 rows) and PSSMs.  A
result that reports saturation is not compared, as there."""
import ctypes
import itertools
import random

import numpy as np
import pytest

import oracle as O
import parasail_amd as P

pytestmark = pytest.mark.gpu

GAPS = [(10, 1), (10, 2), (14, 2), (40, 2)]
SMALL = ["test_small_2.fasta", "test_P0_3.fasta", "test_case.fasta", "test_64.fasta",
         "test_p29.fasta"]


@pytest.fixture(scope="module", autouse=True)
def need_device():
    if not P.available():
        pytest.skip("no HIP device")


def _getters():
    L = P.lib()
    names = ["score_table", "matches_table", "similar_table", "length_table", "score_row",
             "matches_row", "similar_row", "length_row", "score_col", "matches_col",
             "similar_col", "length_col"]
    for nm in names:
        f = getattr(L, "parasail_result_get_" + nm)
        f.restype = ctypes.POINTER(ctypes.c_int)
        f.argtypes = [P.RP]
    L.parasail_result_is_table.argtypes = [P.RP]
    L.parasail_result_is_rowcol.argtypes = [P.RP]
    L.parasail_result_is_stats_table.argtypes = [P.RP]
    L.parasail_result_is_stats_rowcol.argtypes = [P.RP]
    return L


def _arr(p, n):
    return np.ctypeslib.as_array(p, shape=(n,)).copy()


def oracle_alg(name):
    if name == "sw":
        return O.SW, (0, 0, 0, 0)
    if name == "nw":
        return O.NW, (0, 0, 0, 0)
    return O.SG, P.SG_FLAGS[name]


def check_tables(L, name, stats, tab, strat, w, s1, s2, o, g, pm, om, profile=None):
    """One _table/_rowcol call against the oracle's tables; True if compared."""
    kind = O.STATS if stats else O.SCORE
    fn = "parasail_%s%s_%s_%s%s_hip_64_%s" % (name, "_stats" if stats else "", tab, strat,
                                            "_profile" if profile else "", w)
    a, b = s1.encode() if s1 is not None else None, s2.encode()
    if profile is None:
        r = P.function(fn)(a, 0 if a is None else len(a), b, len(b), o, g, pm)
    else:
        ptr = L.parasail_lookup_pfunction(fn.encode())
        assert ptr, fn
        r = ctypes.CFUNCTYPE(P.RP, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int,
                             ctypes.c_int)(ptr)(profile, b, len(b), o, g)
    assert r, fn
    try:
        alg, flags = oracle_alg(name)
        exp, etabs = O.align_tables(alg, kind, s1, s2, o, g, om, flags)
        planes = etabs if stats else [etabs]
        m, n = planes[0].shape
        res = r.contents
        if res.flag & P.FLAG_SATURATED:
            return False
        ctx = (fn, o, g, m, n)
        assert (res.score, res.end_query, res.end_ref) == exp[:3], ctx
        if tab == "table":
            assert L.parasail_result_is_table(r) and bool(L.parasail_result_is_stats_table(r)) == stats
            fields = ["score", "matches", "similar", "length"][:len(planes)]
            for f, e in zip(fields, planes):
                got = _arr(getattr(L, "parasail_result_get_%s_table" % f)(r), m * n)
                assert (got.reshape(m, n) == e).all(), ctx + (f,)
        else:
            assert L.parasail_result_is_rowcol(r) and bool(L.parasail_result_is_stats_rowcol(r)) == stats
            fields = ["score", "matches", "similar", "length"][:len(planes)]
            for f, e in zip(fields, planes):
                row = _arr(getattr(L, "parasail_result_get_%s_row" % f)(r), n)
                col = _arr(getattr(L, "parasail_result_get_%s_col" % f)(r), m)
                assert (row == e[m - 1, :]).all(), ctx + (f, "row")
                assert (col == e[:, n - 1]).all(), ctx + (f, "col")
        return True
    finally:
        L.parasail_result_free(r)


@pytest.fixture(scope="module")
def mats(fixtures):
    L = P.lib()
    L.parasail_matrix_set_value.argtypes = [P.MP, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    d = fixtures["matrices"]["pam50"]
    pam = P.matrix_create(d["alphabet"][:-1], 0, 0)
    for r in range(d["size"]):
        for c in range(d["size"]):
            L.parasail_matrix_set_value(pam, r, c, d["values"][r * d["size"] + c])
    op = O.matrix_from_values(d["alphabet"], d["values"], d["max"], d["min"])
    return {"blosum62": (P.blosum62(), O.blosum62()), "pam50": (pam, op)}


@pytest.mark.parametrize("stats", [False, True])
def test_tables_and_rowcols_on_reference_pairs(stats, fixtures, mats):
    L = _getters()
    pairs = []
    for fn in SMALL:
        seqs = [s for _, s in fixtures["fasta"][fn]]
        pairs += [(seqs[a], seqs[b]) for a, b in itertools.combinations(range(len(seqs)), 2)]
    compared = 0
    for mk in ("blosum62", "pam50"):
        pm, om = mats[mk]
        for (o, g) in GAPS:
            for s1, s2 in pairs:
                for name in P.ALGS:
                    for tab in ("table", "rowcol"):
                        for strat, w in (("striped", "sat"), ("scan", "16")):
                            compared += check_tables(L, name, stats, tab, strat, w, s1, s2,
                                                     o, g, pm, om)
    assert compared > 0


def test_tables_long_query_profile_and_pssm(mats):
    """Queries past the lane-chain kernel's 1,280 rows (the one-lane int32
    tier), the profile forms, and a PSSM."""
    L = _getters()
    L.parasail_profile_create_stats_hip_64_sat.restype = ctypes.c_void_p
    L.parasail_profile_create_stats_hip_64_sat.argtypes = [ctypes.c_char_p, ctypes.c_int, P.MP]
    L.parasail_profile_free.argtypes = [ctypes.c_void_p]
    rng = random.Random(3)
    pm, om = mats["blosum62"]
    aa = "ARNDCQEGHILKMFPSTWYV"
    q = "".join(rng.choice(aa) for _ in range(1400))
    t = q[200:600] + "".join(rng.choice(aa) for _ in range(90))
    for name in ("sw", "nw", "sg", "sg_qx"):
        for stats in (False, True):
            for tab in ("table", "rowcol"):
                assert check_tables(L, name, stats, tab, "striped", "32", q, t, 11, 1, pm, om)
    # profile forms
    q2 = "".join(rng.choice(aa) for _ in range(300))
    qb = q2.encode()  # the profile borrows s1
    prof = L.parasail_profile_create_stats_hip_64_sat(qb, len(qb), pm)
    assert prof
    for name in ("sw", "sg_de"):
        for stats in (False, True):
            t2 = "".join(rng.choice(aa) for _ in range(rng.randint(50, 500)))
            assert check_tables(L, name, stats, "table", "striped", "sat", q2, t2, 11, 1, pm, om,
                                profile=prof)
    L.parasail_profile_free(prof)
    # a PSSM (square matrix rows per position, convert_square_to_pssm)
    L.parasail_matrix_convert_square_to_pssm.restype = P.MP
    pssm = L.parasail_matrix_convert_square_to_pssm(pm, q2.encode(), len(q2))
    opssm = O.convert_square_to_pssm(om, q2)
    t3 = "".join(rng.choice(aa) for _ in range(250))
    for name in ("sw", "nw"):
        for tab in ("table", "rowcol"):
            assert check_tables(L, name, False, tab, "striped", "sat", None, t3, 11, 1, pssm, opssm)
    L.parasail_matrix_free(pssm)
