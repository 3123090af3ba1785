"""The reference's own differential runs, on its own data, through the hip ABI.

Mirrors tests/test_verify.c:93-290 (scores, ends and stats of every function
against serial), tests/test_verify_cigars.c:91-293 (CIGARs) and
tests/test_verify_traces.c:79-102 (whole trace tables, striped results
unwound with parasail_striped_unwind) on every pair (a < b) of each FASTA
the reference's tests use (tests/golden/fixtures.json, extracted from
data/*.fasta), with BLOSUM62 and PAM50 and the reference's gap sets
(10/1, 10/2, 14/2, 40/2; 9/1 as well for traces).  As there, a result that
reports saturation is not compared (test_verify.c:196-203); the oracle
(tests/test_oracle.py pins it) stands in for the serial functions.
"""
import concurrent.futures
import ctypes
import itertools

import numpy as np
import pytest

import oracle as O
import parasail_amd as P

pytestmark = pytest.mark.gpu

GAPS = [(10, 1), (10, 2), (14, 2), (40, 2)]
TRACE_GAPS = [(9, 1)] + GAPS
KINDS = {"score": ("", O.SCORE, 0), "stats": ("_stats", O.STATS, 1), "trace": ("_trace", O.TRACE, 2)}
# the files whose pairs run through every function; test_P0_10 (45 pairs,
# up to 2,005 residues) runs through the "sat" striped functions, per call
# and batched
SMALL = ["test_small_2.fasta", "test_P0_5.fasta", "test_case.fasta", "test_case2.fasta",
         "test_64.fasta", "test_64r.fasta", "test_p29.fasta"]


@pytest.fixture(scope="module", autouse=True)
def need_device():
    if not P.available():
        pytest.skip("no HIP device")


@pytest.fixture(scope="module")
def mats(fixtures):
    L = P.lib()
    L.parasail_matrix_set_value.argtypes = [P.MP, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    d = fixtures["matrices"]["pam50"]
    pam = P.matrix_create(d["alphabet"][:-1], 0, 0)
    for r in range(d["size"]):
        for c in range(d["size"]):
            L.parasail_matrix_set_value(pam, r, c, d["values"][r * d["size"] + c])
    assert (pam.contents.max, pam.contents.min) == (d["max"], d["min"])
    ob = O.blosum62()
    op = O.matrix_from_values(d["alphabet"], d["values"], d["max"], d["min"])
    return {"blosum62": (P.blosum62(), ob), "pam50": (pam, op)}


def pairs_of(fixtures, files):
    out = []
    for fn in files:
        seqs = [s for _, s in fixtures["fasta"][fn]]
        for a, b in itertools.combinations(range(len(seqs)), 2):
            out.append((fn, a, b, seqs[a], seqs[b]))
    return out


def oracle_alg(name):
    if name == "sw":
        return O.SW, (0, 0, 0, 0)
    if name == "nw":
        return O.NW, (0, 0, 0, 0)
    return O.SG, P.SG_FLAGS[name]


def oracle_all(cases):
    """cases: list of (name, kind, s1, s2, o, g, om) -> (result, trace) each,
    on host threads (the oracle's ctypes calls release the GIL)."""
    def one(c):
        name, kind, s1, s2, o, g, om = c
        alg, flags = oracle_alg(name)
        return O.align(alg, KINDS[kind][1], s1, s2, o, g, om, flags)
    with concurrent.futures.ThreadPoolExecutor(8) as ex:
        return list(ex.map(one, cases))


ALG_FLAG = {"sw": P.FLAG_SW, "nw": P.FLAG_NW}


def call(fn_name, s1, s2, o, g, pm, profile=None):
    """One per-call function (or pfunction over `profile`); returns the
    result pointer (caller frees)."""
    L = P.lib()
    a, b = s1.encode(), s2.encode()
    if profile is None:
        f = P.function(fn_name)
        return f(a, len(a), b, len(b), o, g, pm), a, b
    ptr = L.parasail_lookup_pfunction(fn_name.encode())
    assert ptr, fn_name
    pf = ctypes.CFUNCTYPE(P.RP, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int,
                          ctypes.c_int)(ptr)
    return pf(profile, b, len(b), o, g), a, b


def check_result(ctx, name, kind, r, a, b, exp, tr, pm, om, want_table):
    """Compare one result object with the oracle; True if compared (not
    saturated)."""
    L = P.lib()
    assert r, ctx
    res = r.contents
    flag = res.flag
    exp_alg = ALG_FLAG.get(name, P.FLAG_SG)
    assert flag & exp_alg, ctx  # test_verify.c:81-90
    if flag & P.FLAG_SATURATED:
        return False
    assert (res.score, res.end_query, res.end_ref) == tuple(exp[:3]), (ctx, exp)
    if kind == "stats":
        st = ctypes.cast(res.extra, ctypes.POINTER(P.StatsExtra)).contents
        assert (st.matches, st.similar, st.length) == tuple(exp[3:]), (ctx, exp)
    if kind == "trace":
        alg = oracle_alg(name)[0]
        c = O.cigar(a, b, alg == O.SG, exp[1], exp[2], tr)
        expc = None if c is None else (O.cigar_string(c[0]), c[1], c[2])
        got = P.cigar_tuple(L.parasail_result_get_cigar(r, a, len(a), b, len(b), pm))
        assert got == expc, ctx
        if want_table:
            te = ctypes.cast(res.extra, ctypes.POINTER(P.TraceExtra)).contents
            un = L.parasail_striped_unwind(len(a), len(b), r, te.trace_table)
            assert un, ctx
            got_t = np.ctypeslib.as_array(un, shape=(len(a) * len(b),)).copy()
            P.lib()._libc_free(ctypes.cast(un, ctypes.c_void_p))
            assert (got_t.reshape(len(a), len(b)) == tr.astype(np.int32)).all(), ctx
    return True


def _setup_unwind():
    L = P.lib()
    L.parasail_striped_unwind.restype = ctypes.POINTER(ctypes.c_int)
    L.parasail_striped_unwind.argtypes = [ctypes.c_int, ctypes.c_int, P.RP, ctypes.c_void_p]
    for nm in ("parasail_profile_create_hip_64_8", "parasail_profile_create_hip_64_16",
               "parasail_profile_create_hip_64_32", "parasail_profile_create_hip_64_sat",
               "parasail_profile_create_stats_hip_64_8", "parasail_profile_create_stats_hip_64_16",
               "parasail_profile_create_stats_hip_64_32",
               "parasail_profile_create_stats_hip_64_sat"):
        getattr(L, nm).restype = ctypes.c_void_p
        getattr(L, nm).argtypes = [ctypes.c_char_p, ctypes.c_int, P.MP]
    L.parasail_profile_free.argtypes = [ctypes.c_void_p]


@pytest.mark.parametrize("kind", ["score", "stats", "trace"])
def test_every_function_on_reference_pairs(kind, fixtures, mats):
    """Every hip function of `kind` (13 algorithms x striped/scan x widths
    8/16/32/sat, plain and profile) on the pairs of the reference's small
    FASTA fixtures, both matrices, the reference's gap sets."""
    _setup_unwind()
    L = P.lib()
    pairs = pairs_of(fixtures, SMALL)
    gaps = TRACE_GAPS if kind == "trace" else GAPS
    cases = [(name, kind, s1, s2, o, g, mats[mk][1])
             for mk in ("blosum62", "pam50") for (o, g) in gaps
             for (_, _, _, s1, s2) in pairs for name in P.ALGS]
    exps = oracle_all(cases)
    suffix = KINDS[kind][0]
    compared = skipped = 0
    for (name, _, s1, s2, o, g, om), (exp, tr) in zip(cases, exps):
        mk = "blosum62" if om is mats["blosum62"][1] else "pam50"
        pm = mats[mk][0]
        for strat in ("striped", "scan"):
            for w in ("8", "16", "32", "sat"):
                fn = "parasail_%s%s_%s_hip_64_%s" % (name, suffix, strat, w)
                r, a, b = call(fn, s1, s2, o, g, pm)
                ctx = (fn, mk, o, g, s1[:20], s2[:20])
                ok = check_result(ctx, name, kind, r, a, b, exp, tr, pm, om,
                                  want_table=(w == "sat"))
                L.parasail_result_free(r)
                compared += ok
                skipped += not ok
        # the profile form of the sat function
        creator = "parasail_profile_create%s_hip_64_sat" % ("_stats" if kind == "stats" else "")
        qb = s1.encode()  # the profile borrows s1 (src/memory.c:673): keep it alive
        prof = getattr(L, creator)(qb, len(qb), pm)
        assert prof
        fn = "parasail_%s%s_striped_profile_hip_64_sat" % (name, suffix)
        r, a, b = call(fn, s1, s2, o, g, pm, profile=prof)
        assert check_result((fn, mk, o, g), name, kind, r, a, b, exp, tr, pm, om, False)
        L.parasail_result_free(r)
        L.parasail_profile_free(prof)
    # the sat functions never report saturation; most fixed widths hold
    assert compared > 3 * skipped


@pytest.mark.parametrize("kind", ["score", "stats", "trace"])
def test_long_reference_pairs_per_call_and_batched(kind, fixtures, mats):
    """Every fixture pair, test_P0_10's 45 pairs of 127-2,005 residues
    among them, through parasail_hip_align_batch (width sat, device CIGARs
    for traces) with both matrices and every gap set (the first and last
    for _stats and _trace), and test_P0_10's pairs through the sat striped
    function of every algorithm per call (first gap set), against the
    oracle on every pair."""
    _setup_unwind()
    L = P.lib()
    pairs = pairs_of(fixtures, sorted(fixtures["fasta"]))
    gaps = TRACE_GAPS if kind == "trace" else GAPS
    if kind != "score":
        # queries past 1,280 rows take the one-lane int32 tier for _stats
        # and _trace: two gap sets keep the run short
        gaps = [gaps[0], gaps[-1]]
    suffix = KINDS[kind][0]
    for mk in ("blosum62", "pam50"):
        pm, om = mats[mk]
        for (o, g) in gaps:
            cases = [(name, kind, s1, s2, o, g, om) for name in P.ALGS
                     for (_, _, _, s1, s2) in pairs]
            exps = oracle_all(cases)
            k = 0
            for name in P.ALGS:
                q = [p[3] for p in pairs]
                t = [p[4] for p in pairs]
                got = P.align_batch(name, KINDS[kind][2], 0, q, t, o, g, pm,
                                    want_cigar=(kind == "trace"))
                for i, (fa, a_i, b_i, s1, s2) in enumerate(pairs):
                    exp, tr = exps[k + i]
                    ctx = (name, kind, mk, o, g, fa, a_i, b_i)
                    assert not got["flag"][i] & P.FLAG_SATURATED, ctx
                    assert (got["score"][i], got["end_query"][i], got["end_ref"][i]) == \
                        tuple(exp[:3]), ctx
                    if kind == "stats":
                        assert (got["matches"][i], got["similar"][i], got["length"][i]) == \
                            tuple(exp[3:]), ctx
                    if kind == "trace":
                        alg = oracle_alg(name)[0]
                        c = O.cigar(s1, s2, alg == O.SG, exp[1], exp[2], tr)
                        expc = None if c is None else (O.cigar_string(c[0]), c[1], c[2])
                        assert got["cigar"][i] == expc, ctx
                    if fa == "test_P0_10.fasta" and (o, g) == gaps[0]:
                        fn = "parasail_%s%s_striped_hip_64_sat" % (name, suffix)
                        r, a, b = call(fn, s1, s2, o, g, pm)
                        assert check_result((fn,) + ctx, name, kind, r, a, b, exp, tr, pm, om,
                                            want_table=True)
                        L.parasail_result_free(r)
                k += len(pairs)
