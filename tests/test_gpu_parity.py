"""Parity of the hip family (through its C ABI) with the CPU oracle.

Bit-exact on every pair: score, end_query, end_ref, matches/similar/length
and CIGAR, for SW, NW and the eleven SG variants, per call and batched,
square matrices and PSSMs, plus the reference's known answers
(tests/test_basic.c:20-85).  The oracle is pinned in tests/test_oracle.py.
"""
import ctypes
import random

import numpy as np
import pytest

import oracle as O
import parasail_amd as P

pytestmark = pytest.mark.gpu

AA = "ARNDCQEGHILKMFPSTWYV"
# Robinson & Robinson (1991) background frequencies, per 1000
RR = [78.05, 51.29, 44.87, 53.64, 19.25, 42.64, 62.95, 73.77, 21.99, 51.42,
      90.19, 57.44, 22.43, 38.56, 52.03, 71.20, 58.41, 13.30, 32.16, 64.41]


def rseq(rng, n, alpha=AA, weights=None):
    return "".join(rng.choices(alpha, weights=weights, k=n))


def mutate(rng, s, alpha, sub=0.1, indel=0.03):
    out = []
    for c in s:
        r = rng.random()
        if r < sub:
            out.append(rng.choice(alpha))
        elif r < sub + indel / 2:
            continue
        elif r < sub + indel:
            out.append(c)
            out.append(rng.choice(alpha))
        else:
            out.append(c)
    return "".join(out) or alpha[0]


@pytest.fixture(scope="module", autouse=True)
def need_device():
    if not P.available():
        pytest.skip("no HIP device")


@pytest.fixture(scope="module")
def mats():
    return {"blosum62": (P.blosum62(), O.blosum62()),
            "dna": (P.matrix_create("ACGT", 2, -3), O.matrix_create("ACGT", 2, -3))}


def oracle_alg(name):
    if name == "sw":
        return O.SW, (0, 0, 0, 0)
    if name == "nw":
        return O.NW, (0, 0, 0, 0)
    return O.SG, P.SG_FLAGS[name]


def check_one(name, kind, s1, s2, o, g, pm, om, width="sat"):
    suffix = {"score": "", "stats": "_stats", "trace": "_trace"}[kind]
    fn = "parasail_%s%s_striped_hip_64_%s" % (name, suffix, width)
    got = P.align(fn, s1, s2, o, g, pm)
    assert got is not None, fn
    alg, flags = oracle_alg(name)
    okind = {"score": O.SCORE, "stats": O.STATS, "trace": O.TRACE}[kind]
    exp, tr = O.align(alg, okind, s1, s2, o, g, om, flags)
    if got["flag"] & P.FLAG_SATURATED:
        assert width != "sat"
        return got
    ctx = (fn, s1, s2, o, g)
    assert (got["score"], got["end_query"], got["end_ref"]) == exp[:3], ctx
    if kind == "stats":
        assert (got["matches"], got["similar"], got["length"]) == exp[3:], ctx
    if kind == "trace":
        c = O.cigar(s1, s2, alg == O.SG, exp[1], exp[2], tr)
        expc = None if c is None else (O.cigar_string(c[0]), c[1], c[2])
        assert got.get("cigar") == expc, ctx
    return got


def test_known_answers(fixtures):
    L = P.lib()
    b = P.blosum62()
    d = fixtures["matrices"]["pam50"]
    pam = P.matrix_create(d["alphabet"][:-1], 0, 0)
    L.parasail_matrix_set_value.argtypes = [P.MP, P.ctypes.c_int, P.ctypes.c_int, P.ctypes.c_int]
    for r in range(24):
        for c in range(24):
            L.parasail_matrix_set_value(pam, r, c, d["values"][r * 24 + c])
    acgt = P.matrix_create("acgt", 1, -1)
    mats = {"blosum62": b, "pam50": pam}
    for kat in fixtures["kat"]:
        m = acgt if isinstance(kat["matrix"], list) else mats[kat["matrix"]]
        for fn in ("sw_striped_hip_64_16", "sw_striped_hip_64_8", "sw_scan_hip_64_sat",
                   "sw_striped_hip_64_32"):
            r = P.align(fn, kat["s1"], kat["s2"], kat["open"], kat["gap"], m)
            assert r["score"] == kat["score"], (kat["src"], fn)


@pytest.mark.parametrize("name", P.ALGS)
def test_per_call_all_algs(name, mats):
    rng = random.Random(hash(name) & 0xffff)
    for trial in range(12):
        key = "blosum62" if trial % 2 == 0 else "dna"
        pm, om = mats[key]
        alpha = AA if key == "blosum62" else "ACGT"
        s1 = rseq(rng, rng.randint(1, 180), alpha)
        s2 = mutate(rng, s1, alpha) if trial % 3 == 0 else rseq(rng, rng.randint(1, 180), alpha)
        o, g = rng.choice([(11, 1), (10, 1), (5, 2), (14, 2), (3, 1)])
        for kind in ("score", "stats", "trace"):
            check_one(name, kind, s1, s2, o, g, pm, om, width="sat")


def test_per_call_from_many_threads(mats):
    """Per-call functions called from 16 host threads at once (the OpenMP
    pattern of tests/test_openmp.c:353-369): calls made together are served
    by shared launches, grouped by scoring; every result is checked against
    the oracle.  The tasks mix algorithms, kinds, widths, matrices and gaps
    so that groups of several scorings are queued together."""
    from concurrent.futures import ThreadPoolExecutor
    rng = random.Random(77)
    tasks = []
    for k in range(480):
        key = "blosum62" if k % 3 else "dna"
        alpha = AA if key == "blosum62" else "ACGT"
        s1 = rseq(rng, rng.randint(1, 400), alpha)
        s2 = mutate(rng, s1, alpha) if k % 2 else rseq(rng, rng.randint(1, 400), alpha)
        name = rng.choice(["sw", "sw", "nw", "sg", "sg_qb_de"])
        kind = rng.choice(["score", "score", "stats", "trace"])
        width = rng.choice(["sat", "16", "32"])
        o, g = rng.choice([(11, 1), (10, 1), (5, 2)])
        tasks.append((name, kind, width, key, s1, s2, o, g))

    def run(part):
        out = []
        for name, kind, width, key, s1, s2, o, g in part:
            suffix = {"score": "", "stats": "_stats", "trace": "_trace"}[kind]
            fn = "parasail_%s%s_striped_hip_64_%s" % (name, suffix, width)
            out.append(P.align(fn, s1, s2, o, g, mats[key][0]))
        return out

    parts = [tasks[k::16] for k in range(16)]
    with ThreadPoolExecutor(16) as pool:
        got = list(pool.map(run, parts))
    for part, res in zip(parts, got):
        for (name, kind, width, key, s1, s2, o, g), r in zip(part, res):
            ctx = (name, kind, width, key, s1, s2, o, g)
            assert r is not None, ctx
            want = check_one(name, kind, s1, s2, o, g, mats[key][0], mats[key][1], width=width)
            assert r == want, ctx


@pytest.mark.parametrize("width", ["8", "16", "32"])
def test_fixed_widths_and_saturation(width, mats):
    rng = random.Random(int(width))
    pm, om = mats["blosum62"]
    # a W-rich self alignment passes the 8-bit and (for length >= 3000) 16-bit windows
    long_w = "".join(rng.choice("WWWWC") for _ in range(3100))
    cases = [(rseq(rng, 50), rseq(rng, 60)), ("W" * 30, "W" * 30), (long_w, long_w)]
    for s1, s2 in cases:
        for name in ("sw", "nw", "sg"):
            got = check_one(name, "score", s1, s2, 11, 1, pm, om, width=width)
            sat = bool(got["flag"] & P.FLAG_SATURATED)
            exp, _ = O.align(oracle_alg(name)[0], O.SCORE, s1, s2, 11, 1, om,
                             oracle_alg(name)[1])
            if name == "sw":
                assert sat == (exp[0] > (1 << int(width)) - 11 - 2) if width != "32" else not sat
                if sat:
                    assert got["score"] == (1 << int(width)) - 1 and got["end_query"] == 0
            elif sat:
                assert got["score"] == 0 and got["end_ref"] == 0
            bits = {"8": P.FLAG_BITS_8, "16": P.FLAG_BITS_16, "32": P.FLAG_BITS_32}[width]
            assert got["flag"] & bits


def test_sat_reports_narrowest_width(mats):
    pm, om = mats["blosum62"]
    r = P.align("parasail_sw_striped_hip_64_sat", "W" * 10, "W" * 10, 11, 1, pm)
    assert r["score"] == 110 and r["flag"] & P.FLAG_BITS_8
    r = P.align("parasail_sw_striped_hip_64_sat", "W" * 30, "W" * 30, 11, 1, pm)
    assert r["score"] == 330 and r["flag"] & P.FLAG_BITS_16
    assert not r["flag"] & P.FLAG_SATURATED


def test_batch_sat_sw_reports_narrowest_width(mats):
    """A `_sat` SW batch reports, pair by pair, the width the per-call
    `_sat` function reports (src/satcheck.c:41520-41537 tries 8, then 16):
    the 8-bit window holds W^10 (score 110) and random protein pairs, not
    W^30 (score 330)."""
    pm, _ = mats["blosum62"]
    prng = random.Random(31)
    qs = ["W" * 10, "W" * 30, "W" * 25] + [rseq(prng, prng.randint(20, 400)) for _ in range(40)]
    ts = ["W" * 10, "W" * 30, "W" * 23] + [rseq(prng, prng.randint(20, 400)) for _ in range(40)]
    got = P.align_batch("sw", 0, 0, qs, ts, 11, 1, pm)
    bits = P.FLAG_BITS_8 | P.FLAG_BITS_16 | P.FLAG_BITS_32
    widths = set()
    for k, (q, t) in enumerate(zip(qs, ts)):
        r = P.align("parasail_sw_striped_hip_64_sat", q, t, 11, 1, pm)
        assert got["score"][k] == r["score"], k
        assert int(got["flag"][k]) & bits == r["flag"] & bits, (k, r["score"])
        widths.add(r["flag"] & bits)
    assert widths == {P.FLAG_BITS_8, P.FLAG_BITS_16}


def test_profile_functions(mats):
    L = P.lib()
    pm, om = mats["blosum62"]
    rng = random.Random(4)
    q = rseq(rng, 300)
    L.parasail_profile_free.argtypes = [P.ctypes.c_void_p]
    for w in ("16", "sat", "8"):
        create = P.ctypes.CFUNCTYPE(P.ctypes.c_void_p, P.ctypes.c_char_p, P.ctypes.c_int, P.MP)(
            L.parasail_lookup_pcreator(("parasail_sw_striped_profile_hip_64_" + w).encode()))
        pf = P.ctypes.CFUNCTYPE(P.RP, P.ctypes.c_void_p, P.ctypes.c_char_p, P.ctypes.c_int,
                                P.ctypes.c_int, P.ctypes.c_int)(
            L.parasail_lookup_pfunction(("parasail_sw_striped_profile_hip_64_" + w).encode()))
        qb = q.encode()  # the profile borrows s1 (src/memory.c:673): keep it alive
        prof = create(qb, len(qb), pm)
        assert prof
        for _ in range(5):
            t = rseq(rng, rng.randint(50, 1000))
            r = pf(prof, t.encode(), len(t), 11, 1)
            exp, _ = O.align(O.SW, O.SCORE, q, t, 11, 1, om)
            rc = r.contents
            # SW striped saturates when the score passes 2^w - max - 2 on the
            # biased scale (src/sw_striped_avx2_256_8.c), and then reports
            # score 2^w - 1 with zero ends
            sat = w == "8" and exp[0] > (1 << 8) - 11 - 2
            assert bool(rc.flag & P.FLAG_SATURATED) == sat, (w, t)
            if sat:
                assert (rc.score, rc.end_query, rc.end_ref) == ((1 << 8) - 1, 0, 0)
            else:
                assert (rc.score, rc.end_query, rc.end_ref) == exp[:3]
            L.parasail_result_free(r)
        L.parasail_profile_free(prof)


def test_profile_device_copy(mats):
    """A hip profile over a square matrix holds a device copy of its query
    codes and matrix, made once by the creator and freed by the profile's
    `free` (src/memory.c:698-746).  Every profile call reads that copy: a
    score call after the caller's s1 buffer is overwritten still aligns the
    query the profile was made from, as the reference's striped profile,
    built at creation (src/memory_avx2.c:81-132), does.  Score and stats
    calls of every width equal the oracle."""
    L = P.lib()
    pm, om = mats["blosum62"]
    rng = random.Random(12)
    q = rseq(rng, 240)

    class PData(P.ctypes.Structure):
        _fields_ = [("score", P.ctypes.c_void_p), ("matches", P.ctypes.c_void_p),
                    ("similar", P.ctypes.c_void_p)]

    class Profile(P.ctypes.Structure):
        _fields_ = [("s1", P.ctypes.c_void_p), ("s1Len", P.ctypes.c_int),
                    ("matrix", P.ctypes.c_void_p), ("p8", PData), ("p16", PData),
                    ("p32", PData), ("p64", PData), ("free", P.ctypes.c_void_p),
                    ("stop", P.ctypes.c_int)]

    class HipData(P.ctypes.Structure):
        _fields_ = [("codes", P.ctypes.c_void_p), ("dev", P.ctypes.c_void_p)]

    L.parasail_profile_free.argtypes = [P.ctypes.c_void_p]
    targets = [rseq(rng, rng.randint(30, 700)) for _ in range(6)]
    for kind, w, slot in (("", "16", "p16"), ("", "8", "p8"), ("", "32", "p32"),
                          ("", "sat", "p16"), ("_stats", "16", "p16"), ("_stats", "sat", "p16")):
        fn = "parasail_sw%s_striped_profile_hip_64_%s" % (kind, w)
        create = P.ctypes.CFUNCTYPE(P.ctypes.c_void_p, P.ctypes.c_void_p, P.ctypes.c_int, P.MP)(
            L.parasail_lookup_pcreator(fn.encode()))
        pf = P.ctypes.CFUNCTYPE(P.RP, P.ctypes.c_void_p, P.ctypes.c_char_p, P.ctypes.c_int,
                                P.ctypes.c_int, P.ctypes.c_int)(
            L.parasail_lookup_pfunction(fn.encode()))
        buf = P.ctypes.create_string_buffer(q.encode())
        prof = create(P.ctypes.addressof(buf), len(q), pm)
        assert prof, fn
        pr = Profile.from_address(prof)
        data = getattr(pr, slot)
        assert data.score and not data.matches and not data.similar, fn
        hd = HipData.from_address(data.score)
        assert hd.codes and hd.dev, fn
        if not kind:
            # the profile no longer reads the caller's residues
            P.ctypes.memmove(buf, ("W" * len(q)).encode(), len(q))
        for t in targets:
            r = pf(prof, t.encode(), len(t), 11, 1)
            assert r, fn
            alg_kind = O.STATS if kind else O.SCORE
            exp, _ = O.align(O.SW, alg_kind, q, t, 11, 1, om)
            rc = r.contents
            if rc.flag & P.FLAG_SATURATED:
                assert w == "8" and exp[0] > (1 << 8) - 11 - 2, (fn, t)
            else:
                assert (rc.score, rc.end_query, rc.end_ref) == exp[:3], (fn, t)
                if kind:
                    st = P.ctypes.cast(rc.extra, P.ctypes.POINTER(P.StatsExtra)).contents
                    assert (st.matches, st.similar, st.length) == exp[3:6], (fn, t)
            L.parasail_result_free(r)
        L.parasail_profile_free(prof)


def test_config0_all_targets():
    """The per-pair workload in full: the 300-residue query against all
    2,000 targets through parasail_sw_striped_profile_hip_64_16 (and its
    stats form), called from 16 host threads as tests/test_openmp.c:353-369
    calls the per-pair functions, every result against the oracle."""
    from concurrent.futures import ThreadPoolExecutor
    from parasail_amd import workloads as WL
    L = P.lib()
    L.parasail_profile_free.argtypes = [P.ctypes.c_void_p]
    pm, om = P.blosum62(), O.blosum62()
    q, b = WL.config0()
    targets = [b.target(k) for k in range(len(b))]
    assert len(targets) == 2000
    for fn in ("parasail_sw_striped_profile_hip_64_16", "parasail_sw_stats_striped_profile_hip_64_16"):
        create = P.ctypes.CFUNCTYPE(P.ctypes.c_void_p, P.ctypes.c_char_p, P.ctypes.c_int, P.MP)(
            L.parasail_lookup_pcreator(fn.encode()))
        pf = P.ctypes.CFUNCTYPE(P.RP, P.ctypes.c_void_p, P.ctypes.c_char_p, P.ctypes.c_int,
                                P.ctypes.c_int, P.ctypes.c_int)(
            L.parasail_lookup_pfunction(fn.encode()))
        prof = create(q, len(q), pm)
        assert prof

        def run(ks):
            out = []
            for k in ks:
                r = pf(prof, targets[k], len(targets[k]), 11, 1)
                rc = r.contents
                row = (rc.score, rc.end_query, rc.end_ref, rc.flag & P.FLAG_SATURATED)
                if rc.flag & P.FLAG_STATS:
                    st = P.ctypes.cast(rc.extra, P.ctypes.POINTER(P.StatsExtra)).contents
                    row += (st.matches, st.similar, st.length)
                L.parasail_result_free(r)
                out.append((k, row))
            return out

        with ThreadPoolExecutor(16) as pool:
            got = dict(x for part in pool.map(run, [range(k, 2000, 16) for k in range(16)])
                       for x in part)
        L.parasail_profile_free(prof)
        stats = "stats" in fn
        for k in range(2000):
            exp, _ = O.align(O.SW, O.STATS if stats else O.SCORE, q.decode(),
                             targets[k].decode(), 11, 1, om)
            want = tuple(exp[:3]) + (0,) + (tuple(exp[3:]) if stats else ())
            assert got[k] == want, (fn, k)


def test_pssm(mats):
    L = P.lib()
    pm, om = mats["blosum62"]
    rng = random.Random(8)
    for _ in range(4):
        s1 = rseq(rng, rng.randint(5, 120))
        s2 = rseq(rng, rng.randint(5, 160))
        pp = L.parasail_matrix_convert_square_to_pssm(pm, s1.encode(), len(s1))
        op = O.convert_square_to_pssm(om, s1)
        for name in ("sw", "nw", "sg", "sg_qb_de"):
            alg, flags = oracle_alg(name)
            exp, _ = O.align(alg, O.SCORE, None, s2, 11, 1, op, flags)
            got = P.align("parasail_%s_striped_hip_64_16" % name, None, s2, 11, 1, pp)
            assert (got["score"], got["end_query"], got["end_ref"]) == exp[:3]
            # stats with a PSSM still use the consensus (sw_stats.c:62-64)
            exps, _ = O.align(alg, O.STATS, s1, s2, 11, 1, op, flags)
            gots = P.align("parasail_%s_stats_striped_hip_64_16" % name, s1, s2, 11, 1, pp)
            assert (gots["score"], gots["matches"], gots["similar"], gots["length"]) == \
                (exps[0],) + exps[3:]
        L.parasail_matrix_free(pp)


def test_edge_cases(mats):
    pm, om = mats["blosum62"]
    cases = [("A", "A"), ("A", "W"), ("W", "AAAAAAAAAA"), ("AAAAAAAAAA", "W"),
             ("MKT", "MKT"), ("XXXJJZ", "BZX*J"), ("a" * 40, "A" * 39)]
    for s1, s2 in cases:
        for name in P.ALGS:
            for kind in ("score", "stats", "trace"):
                check_one(name, kind, s1, s2, 11, 1, pm, om)
        check_one("sw", "score", s1, s2, 0, 0, pm, om)
        check_one("nw", "score", s1, s2, 1, 0, pm, om)


def test_invalid_arguments_return_null(mats):
    pm, _ = mats["blosum62"]
    f = P.function("parasail_sw_striped_hip_64_16")
    assert not f(b"AC", 2, b"AC", 2, -1, 1, pm)
    assert not f(b"AC", 2, b"AC", 0, 1, 1, pm)
    assert not f(None, 2, b"AC", 2, 1, 1, pm)
    assert not f(b"AC", 2, b"AC", 2, 1, 1, None)


def lognormal_lengths(rng, n, median, sigma, lo, hi):
    return np.clip(np.rint(median * np.exp(sigma * rng.standard_normal(n))), lo, hi).astype(int)


@pytest.mark.parametrize("name", ["sw", "nw", "sg", "sg_qx", "sg_de", "sg_qb_db"])
def test_batch_vs_oracle(name, mats):
    """Config-1 shaped batch: log-normal lengths around 300, BLOSUM62 11/1."""
    rng = np.random.default_rng(1)
    prng = random.Random(2)
    pm, om = mats["blosum62"]
    n = 600
    ql = lognormal_lengths(rng, n, 300, 0.6, 20, 5000)
    tl = lognormal_lengths(rng, n, 300, 0.6, 20, 5000)
    qs = [rseq(prng, int(k), AA, RR) for k in ql]
    ts = [rseq(prng, int(k), AA, RR) for k in tl]
    got = P.align_batch(name, 0, 0, qs, ts, 11, 1, pm)
    alg, flags = oracle_alg(name)
    exp, _ = O.batch(alg, O.SCORE, [om.map(q) for q in qs], [om.map(t) for t in ts],
                     11, 1, om, flags)
    bad = np.nonzero((got["score"] != exp[:, 0]) | (got["end_query"] != exp[:, 1]) |
                     (got["end_ref"] != exp[:, 2]))[0]
    assert bad.size == 0, [(int(k), int(ql[k]), int(tl[k]), got["score"][k], exp[k, :3].tolist())
                           for k in bad[:5]]


def test_batch_long_and_ragged(mats):
    """Multi-band queries (> 512 rows), ragged target lengths, short pairs."""
    prng = random.Random(3)
    pm, om = mats["blosum62"]
    qs, ts = [], []
    for k in range(60):
        m = prng.choice([1, 2, 7, 63, 64, 65, 511, 512, 513, 900, 1500, 2600])
        q = rseq(prng, m)
        t = mutate(prng, q, AA) if k % 2 else rseq(prng, prng.randint(1, 1400))
        qs.append(q)
        ts.append(t)
    for name in ("sw", "nw", "sg"):
        got = P.align_batch(name, 0, 0, qs, ts, 11, 1, pm)
        alg, flags = oracle_alg(name)
        exp, _ = O.batch(alg, O.SCORE, [om.map(q) for q in qs], [om.map(t) for t in ts],
                         11, 1, om, flags)
        for k in range(len(qs)):
            assert (got["score"][k], got["end_query"][k], got["end_ref"][k]) == \
                tuple(exp[k, :3]), (name, k, len(qs[k]), len(ts[k]))


def test_batch_stats_trace_dna(mats):
    """Config-2 shaped: 250 bp reads in 400 bp windows, ACGT 2/-3, 5/2."""
    prng = random.Random(5)
    pm, om = mats["dna"]
    qs, ts = [], []
    for _ in range(200):
        w = rseq(prng, 400, "ACGT")
        off = prng.randint(0, 150)
        qs.append(mutate(prng, w[off:off + 250], "ACGT", sub=0.04, indel=0.01))
        ts.append(w)
    for name in ("sg", "sg_qx", "sw", "nw"):
        alg, flags = oracle_alg(name)
        st = P.align_batch(name, 1, 16, qs, ts, 5, 2, pm)
        exp, _ = O.batch(alg, O.STATS, [om.map(q) for q in qs], [om.map(t) for t in ts],
                         5, 2, om, flags)
        for col, key in enumerate(("score", "end_query", "end_ref", "matches", "similar", "length")):
            assert (st[key] == exp[:, col]).all(), (name, key)
        tr = P.align_batch(name, 2, 16, qs, ts, 5, 2, pm, want_cigar=True)
        for k in range(len(qs)):
            e, t = O.align(alg, O.TRACE, qs[k], ts[k], 5, 2, om, flags)
            c = O.cigar(qs[k], ts[k], alg == O.SG, e[1], e[2], t)
            expc = None if c is None else (O.cigar_string(c[0]), c[1], c[2])
            assert tr["cigar"][k] == expc, (name, k)


@pytest.mark.parametrize("name", ["sg", "sg_qx", "sg_de", "sg_qb_db", "sw", "nw"])
def test_batch_stats_trace_wavefront(name, mats):
    """_stats and _trace through the int32 wavefront kernel (no saturation
    check), queries over every lane shape and past its range (the general
    tier in the same batch), mixed-case residues, CIGARs from the device
    walk."""
    prng = random.Random(21)
    pm, om = mats["dna"]
    qs, ts = [], []
    for k in range(90):
        m = prng.choice([1, 3, 17, 64, 250, 257, 320, 321, 640, 641, 1000, 1280, 1281])
        q = rseq(prng, m, "ACGT")
        t = mutate(prng, q, "ACGT", sub=0.05, indel=0.02) if k % 3 else \
            rseq(prng, prng.randint(1, 700), "ACGT")
        if k % 5 == 0:
            q = q.lower()
        qs.append(q)
        ts.append(t)
    alg, flags = oracle_alg(name)
    st = P.align_batch(name, 1, 0, qs, ts, 5, 2, pm)
    exp, _ = O.batch(alg, O.STATS, [om.map(q) for q in qs], [om.map(t) for t in ts],
                     5, 2, om, flags)
    for col, key in enumerate(("score", "end_query", "end_ref", "matches", "similar", "length")):
        bad = np.nonzero(st[key] != exp[:, col])[0]
        assert bad.size == 0, (name, key, [(int(k), len(qs[k]), len(ts[k])) for k in bad[:4]])
    tr = P.align_batch(name, 2, 0, qs, ts, 5, 2, pm, want_cigar=True)
    for k in range(len(qs)):
        e, t = O.align(alg, O.TRACE, qs[k], ts[k], 5, 2, om, flags)
        assert (tr["score"][k], tr["end_query"][k], tr["end_ref"][k]) == e[:3], (name, k)
        c = O.cigar(qs[k], ts[k], alg == O.SG, e[1], e[2], t)
        expc = None if c is None else (O.cigar_string(c[0]), c[1], c[2])
        assert tr["cigar"][k] == expc, (name, k, len(qs[k]), len(ts[k]))


def test_banded(mats):
    L = P.lib()
    pm, om = mats["dna"]
    prng = random.Random(6)
    for _ in range(10):
        s1 = rseq(prng, prng.randint(1, 400), "ACGT")
        s2 = mutate(prng, s1, "ACGT", sub=0.1, indel=0.04)
        for k in (1, 3, 16, 256):
            r = L.parasail_nw_banded_hip(s1.encode(), len(s1), s2.encode(), len(s2), 5, 2, k, pm)
            exp = O.nw_banded(s1, s2, 5, 2, k, om)
            rc = r.contents
            assert (rc.score, rc.end_query, rc.end_ref) == exp[:3]
            assert rc.flag & P.FLAG_BANDED and rc.flag & P.FLAG_NW
            L.parasail_result_free(r)


def _oracle_pairs(batch, idx, name, om):
    alg, flags = oracle_alg(name)
    qs = [om.map(batch.query(int(k))) for k in idx]
    ts = [om.map(batch.target(int(k))) for k in idx]
    exp, _ = O.batch(alg, O.SCORE, qs, ts, 11, 1, om, flags)
    return exp


def test_batch_composition_independent(mats):
    """A pair's result does not depend on the rest of its batch: the planner
    picks rows per lane, chain width and queue order from the whole batch."""
    from parasail_amd import workloads as WL
    pm, om = mats["blosum62"]
    b = WL.config1(seed=11, npairs=3000)
    sub = np.sort(np.random.default_rng(12).choice(len(b), 300, replace=False))
    bs = b.subset(sub)
    for name in ("sw", "nw", "sg"):
        full = P.align_arrays(name, 0, 0, b, 11, 1, pm)
        part = P.align_arrays(name, 0, 0, bs, 11, 1, pm)
        for key in ("score", "end_query", "end_ref"):
            assert (full[key][sub] == part[key]).all(), (name, key)
        exp = _oracle_pairs(bs, range(len(bs)), name, om)
        got = np.stack([part["score"], part["end_query"], part["end_ref"]], axis=1)
        bad = np.nonzero((got != exp[:, :3]).any(axis=1))[0]
        assert bad.size == 0, (name, [(int(sub[k]), got[k].tolist(), exp[k, :3].tolist())
                                      for k in bad[:5]])


def test_config1_full_size_properties(mats):
    """The full 200,000-pair protein batch (log-normal lengths around 300,
    BLOSUM62 11/1): a random 2,000-pair sample equals the serial oracle bit
    for bit, and every pair equals the oracle's AVX2 striped restatement
    (test_oracle.py pins it to the serial oracle); on every pair SW >= SG
    (all ends free) >= NW, SW >= 0, the end positions lie inside the
    sequences, and a second pass reproduces the first."""
    from parasail_amd import workloads as WL
    pm, om = mats["blosum62"]
    b = WL.config1()
    n = len(b)
    res = {name: P.align_arrays(name, 0, 0, b, 11, 1, pm) for name in ("sw", "nw", "sg")}
    sw, nw, sg = res["sw"]["score"], res["nw"]["score"], res["sg"]["score"]
    assert (sw >= 0).all() and (sw >= sg).all() and (sg >= nw).all()
    for name in ("sw", "nw", "sg"):
        r = res[name]
        assert ((r["end_query"] >= 0) & (r["end_query"] < b.qlen)).all(), name
        assert ((r["end_ref"] >= 0) & (r["end_ref"] < b.tlen)).all(), name
        assert not (r["flag"] & P.FLAG_SATURATED).any(), name
    again = P.align_arrays("sw", 0, 0, b, 11, 1, pm)
    for key in ("score", "end_query", "end_ref"):
        assert (again[key] == res["sw"][key]).all()
    idx = np.sort(np.random.default_rng(13).choice(n, 2000, replace=False))
    for name in ("sw", "nw", "sg"):
        exp = _oracle_pairs(b, idx, name, om)
        r = res[name]
        got = np.stack([r["score"][idx], r["end_query"][idx], r["end_ref"][idx]], axis=1)
        bad = np.nonzero((got != exp[:, :3]).any(axis=1))[0]
        assert bad.size == 0, (name, [(int(idx[k]), got[k].tolist(), exp[k, :3].tolist())
                                      for k in bad[:5]])
    if not O.striped_available():
        return
    qs = [om.map(b.query(k)) for k in range(n)]
    ts = [om.map(b.target(k)) for k in range(n)]
    for name, alg, fl in (("sw", O.SW, (0, 0, 0, 0)), ("nw", O.NW, (0, 0, 0, 0)),
                          ("sg", O.SG, (1, 1, 1, 1))):
        exp, _, secs = O.striped_batch(alg, qs, ts, 11, 1, om, fl, threads=16)
        assert secs >= 0
        r = res[name]
        got = np.stack([r["score"], r["end_query"], r["end_ref"]], axis=1)
        bad = np.nonzero((got != exp).any(axis=1))[0]
        assert bad.size == 0, (name, [(int(k), got[k].tolist(), exp[k].tolist())
                                      for k in bad[:5]])


def test_banded_batch(mats):
    """Many pairs in one launch: bands of every width bucket (2..32 slots per
    lane), |m - n| beyond the band, a band wider than 64 x 32 slots (the
    single-lane loop) and pairs of ~20,000 bp with k = 256."""
    pm, om = mats["dna"]
    prng = random.Random(16)
    qs, ts, ks = [], [], []
    for i in range(48):
        m = prng.choice([1, 2, 5, 40, 200, 700, 1500, 3000])
        s1 = rseq(prng, m, "ACGT")
        s2 = mutate(prng, s1, "ACGT", sub=0.1, indel=0.04) if i % 3 else rseq(prng, prng.randint(1, 3000), "ACGT")
        qs.append(s1)
        ts.append(s2)
    for k in (1, 3, 16, 256):
        got = P.nw_banded_batch(qs, ts, 5, 2, k, pm)
        for i in range(len(qs)):
            exp = O.nw_banded(qs[i], ts[i], 5, 2, k, om)
            assert (got["score"][i], got["end_query"][i], got["end_ref"][i]) == exp[:3], (k, i)
    big_q = [rseq(prng, 20000 + 997 * i, "ACGT") for i in range(4)]
    big_t = [mutate(prng, q, "ACGT", sub=0.1, indel=0.02) for q in big_q]
    got = P.nw_banded_batch(big_q, big_t, 5, 2, 256, pm)
    for i in range(4):
        assert got["score"][i] == O.nw_banded(big_q[i], big_t[i], 5, 2, 256, om)[0], i


def test_banded_rows(mats):
    """Pairs of >= 512 query rows take the row-band kernel when segments are
    off, and the segment kernel for bands of 384 to 2,048 slots when they
    are on: bands from 3 to 1,800 columns (every rows-per-lane choice), the
    query longer and shorter than the target, bands reaching past both ends
    of the query, targets of 1 to 40 columns, zero gap costs, scores in
    registers (DNA 2/-3, 5/2; a PSSM of it) and in an LDS table (BLOSUM62
    11/1; DNA with a gap extension of 4 above an open of 2)."""
    prev = P.nw_banded_set_segment(0)
    try:
        _banded_rows_cases(mats)
        P.nw_banded_set_segment(1)
        _banded_rows_cases(mats)
    finally:
        P.nw_banded_set_segment(prev)


def _banded_rows_cases(mats):
    L = P.lib()
    prng = random.Random(21)
    shapes = [(512, 530, 5), (700, 640, 40), (3000, 3100, 20), (2500, 2400, 200),
              (9000, 9050, 256), (6000, 6000, 700), (4000, 3000, 300), (1500, 1400, 800),
              (5000, 5200, 100), (2000, 2002, 450)]
    for key, alpha, o, g in (("dna", "ACGT", 5, 2), ("blosum62", AA, 11, 1), ("dna", "ACGT", 2, 4)):
        pm, om = mats[key]
        qs, ts, ks = [], [], []
        for (m, n, k) in shapes:
            s1 = rseq(prng, m, alpha)
            s2 = mutate(prng, s1, alpha, sub=0.1, indel=0.02)
            s2 = (s2 + rseq(prng, max(0, n - len(s2)), alpha))[:n]
            qs.append(s1)
            ts.append(s2)
            ks.append(k)
        for k in sorted(set(ks)):
            idx = [i for i in range(len(qs)) if ks[i] == k]
            got = P.nw_banded_batch([qs[i] for i in idx], [ts[i] for i in idx], o, g, k, pm)
            for x, i in enumerate(idx):
                exp = O.nw_banded(qs[i], ts[i], o, g, k, om)
                assert (got["score"][x], got["end_query"][x], got["end_ref"][x]) == exp[:3], \
                    (key, o, g, len(qs[i]), len(ts[i]), k)
    # targets of 1 to 40 columns, k = 1, free gaps, a 512-row query
    pm, om = mats["dna"]
    edge = [(600, 1, 1, 5, 2), (700, 3, 2, 5, 2), (512, 40, 9, 5, 2), (900, 880, 1, 5, 2),
            (1000, 1010, 30, 0, 0), (512, 512, 256, 3, 0)]
    for (m, n, k, o, g) in edge:
        s1 = rseq(prng, m, "ACGT")
        s2 = rseq(prng, n, "ACGT")
        got = P.nw_banded_batch([s1], [s2], o, g, k, pm)
        exp = O.nw_banded(s1, s2, o, g, k, om)
        assert (got["score"][0], got["end_query"][0], got["end_ref"][0]) == exp[:3], (m, n, k, o, g)
    for (m, n, k) in ((3000, 2950, 64), (800, 900, 300)):
        s1 = rseq(prng, m, "ACGT")
        s2 = mutate(prng, s1, "ACGT", sub=0.1, indel=0.02)[:n]
        pp = L.parasail_matrix_convert_square_to_pssm(pm, s1.encode(), len(s1))
        op = O.convert_square_to_pssm(om, s1)
        r = L.parasail_nw_banded_hip(s1.encode(), len(s1), s2.encode(), len(s2), 5, 2, k, pp)
        exp = O.nw_banded(s1, s2, 5, 2, k, op)
        assert r.contents.score == exp[0], (m, n, k)
        L.parasail_result_free(r)
        L.parasail_matrix_free(pp)


def test_banded_segments(mats):
    """Long banded pairs cut into segments that run at the same time and
    are joined by the rank-convergence check: mutated DNA pairs (query longer
    and shorter than the target, several segment lengths, segments off), a
    protein pair over BLOSUM62 (LDS table), and unrelated pairs whose
    segments do not join and are aligned again whole -- every score equal to
    the oracle's banded DP."""
    prng = random.Random(31)
    pm, om = mats["dna"]
    qs, ts = [], []
    for i, m in enumerate((24000, 31000, 41000, 52000)):
        s1 = rseq(prng, m, "ACGT")
        s2 = mutate(prng, s1, "ACGT", sub=0.1, indel=0.02)
        if i % 2:
            s1, s2 = s2, s1
        qs.append(s1)
        ts.append(s2)
    for k in (256, 200):
        exp = [O.nw_banded(qs[i], ts[i], 5, 2, k, om)[:3] for i in range(len(qs))]
        prev = P.nw_banded_set_segment(8192)
        try:
            for cols in (8192, 2048, 0):
                P.nw_banded_set_segment(cols)
                got = P.nw_banded_batch(qs, ts, 5, 2, k, pm)
                pairs, segs, redone = P.nw_banded_last_stats()
                if cols:
                    assert pairs == len(qs) and segs > 2 * pairs and redone == 0, (cols, pairs, segs, redone)
                else:
                    assert pairs == 0
                for i in range(len(qs)):
                    assert (got["score"][i], got["end_query"][i], got["end_ref"][i]) == exp[i], \
                        (k, cols, i, len(qs[i]), len(ts[i]))
        finally:
            P.nw_banded_set_segment(prev)
    # unrelated pairs: the states never line up, the pairs are redone whole;
    # a protein pair beside them joins
    pb, ob = mats["blosum62"]
    u = [(rseq(prng, 26000, "ACGT"), rseq(prng, 26100, "ACGT")) for _ in range(2)]
    got = P.nw_banded_batch([a for a, _ in u], [b for _, b in u], 5, 2, 256, pm)
    assert P.nw_banded_last_stats()[2] >= 1
    for i, (a, b) in enumerate(u):
        assert got["score"][i] == O.nw_banded(a, b, 5, 2, 256, om)[0], i
    s1 = rseq(prng, 30000, AA)
    s2 = mutate(prng, s1, AA, sub=0.15, indel=0.02)
    got = P.nw_banded_batch([s1], [s2], 11, 1, 300, pb)
    assert P.nw_banded_last_stats()[:2][0] == 1
    assert got["score"][0] == O.nw_banded(s1, s2, 11, 1, 300, ob)[0]


def test_long_nw_int32(mats):
    """Global scores of pairs too long for int16 (the multi-wave int32
    kernel): one to three groups of bands, mutated and unrelated pairs,
    through the batch and the per-call scan entry point."""
    pm, om = mats["dna"]
    prng = random.Random(17)
    shapes = [(20000, 3000), (9000, 9000), (16000, 16100), (40000, 2500), (1500, 20000)]
    qs, ts = [], []
    for i, (m, n) in enumerate(shapes):
        q = rseq(prng, m, "ACGT")
        t = mutate(prng, q, "ACGT", sub=0.1, indel=0.02)[:n] if i % 2 == 0 else rseq(prng, n, "ACGT")
        qs.append(q)
        ts.append(t)
    got = P.align_batch("nw", 0, 32, qs, ts, 5, 2, pm)
    exp, _ = O.batch(O.NW, O.SCORE, [om.map(q) for q in qs], [om.map(t) for t in ts], 5, 2, om)
    for i in range(len(qs)):
        assert (got["score"][i], got["end_query"][i], got["end_ref"][i]) == tuple(exp[i, :3]), i
    r = P.align("parasail_nw_scan_hip_64_32", qs[1], ts[1], 5, 2, pm)
    assert r["score"] == exp[1, 0]


def test_long_nw_config3_extremes(mats):
    """Long-pair lengths at both ends of the 20,000-200,000 bp range, 13
    groups of bands for the longest, where the oracle is too slow for the
    full DP: a copy with one block of k residues deleted or inserted scores
    exactly 2(n - k') - open - (k - 1) gap under match 2 / open 5 / extend 2
    (at most min(m, n) matched columns of at most 2 each, at least k
    residues in gaps, one gap the cheapest way to leave them), which the
    plain alignment reaches.  Banded NW with k = 256 on the same pairs is
    compared with the oracle's banded DP (200,000 x 513 cells), and the
    20,000 bp mutated pair with the oracle's full DP."""
    pm, om = mats["dna"]
    prng = random.Random(31)
    q1 = rseq(prng, 200000, "ACGT")
    t1 = q1[:123457] + q1[123457 + 77:]             # 77 deleted
    q2 = rseq(prng, 150000, "ACGT")
    t2 = q2[:40000] + rseq(prng, 33, "ACGT") + q2[40000:]  # 33 inserted
    q3 = rseq(prng, 20000, "ACGT")
    t3 = mutate(prng, q3, "ACGT", sub=0.1, indel=0.02)
    qs, ts = [q1, q2, q3], [t1, t2, t3]
    got = P.align_batch("nw", 0, 32, qs, ts, 5, 2, pm)
    want = [2 * 199923 - 5 - 76 * 2, 2 * 150000 - 5 - 32 * 2]
    for i in range(2):
        assert (got["score"][i], got["end_query"][i], got["end_ref"][i]) == \
            (want[i], len(qs[i]) - 1, len(ts[i]) - 1), i
    exp, _ = O.batch(O.NW, O.SCORE, [om.map(q3)], [om.map(t3)], 5, 2, om)
    assert (got["score"][2], got["end_query"][2], got["end_ref"][2]) == tuple(exp[0, :3])
    band = P.nw_banded_batch(qs, ts, 5, 2, 256, pm)
    for i in range(3):
        e = O.nw_banded(qs[i], ts[i], 5, 2, 256, om)
        assert (band["score"][i], band["end_query"][i], band["end_ref"][i]) == e[:3], i
    # the block's offset (77, 33) lies inside the band: banded equals global
    assert band["score"][0] == want[0] and band["score"][1] == want[1]


def test_long_nw_int32_table_scores(mats):
    """The long-pair kernel with scores read from the LDS table (alphabets
    past 8 codes, or W + open outside 4 bits): protein pairs with BLOSUM62
    11/1 and DNA with open 20."""
    bp, bo = mats["blosum62"]
    dp_, do = mats["dna"]
    prng = random.Random(18)
    # m + n past the int16 window for BLOSUM62 11/1, so both take this kernel
    q = rseq(prng, 20000)
    qs = [q, q[:1200]]
    ts = [mutate(prng, q, AA)[:14000], rseq(prng, 31000)]
    got = P.align_batch("nw", 0, 32, qs, ts, 11, 1, bp)
    exp, _ = O.batch(O.NW, O.SCORE, [bo.map(x) for x in qs], [bo.map(x) for x in ts], 11, 1, bo)
    for i in range(2):
        assert got["score"][i] == exp[i, 0], i
    qd = rseq(prng, 17000, "ACGT")
    td = mutate(prng, qd, "ACGT", sub=0.1, indel=0.02)[:3000]
    got = P.align_batch("nw", 0, 32, [qd], [td], 20, 2, dp_)
    exp, _ = O.batch(O.NW, O.SCORE, [do.map(qd)], [do.map(td)], 20, 2, do)
    assert got["score"][0] == exp[0, 0]


def _resident_api():
    L = P.lib()
    I64, I32 = ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int)
    L.parasail_hip_batch_create.restype = ctypes.c_void_p
    L.parasail_hip_batch_create.argtypes = [ctypes.c_char_p, I64, I32, ctypes.c_char_p, I64, I32,
                                            ctypes.c_int, P.MP]
    L.parasail_hip_batch_align.restype = ctypes.c_int
    L.parasail_hip_batch_align.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int,
                                           ctypes.c_int, ctypes.POINTER(P.BatchResult)]
    L.parasail_hip_batch_align_multi.restype = ctypes.c_int
    L.parasail_hip_batch_align_multi.argtypes = [ctypes.c_void_p, ctypes.c_int,
                                                 ctypes.POINTER(ctypes.c_char_p), ctypes.c_int,
                                                 ctypes.c_int, ctypes.POINTER(P.BatchResult)]
    L.parasail_hip_last_pass_ms.restype = ctypes.c_int
    L.parasail_hip_last_pass_ms.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                            ctypes.POINTER(ctypes.c_double)]
    L.parasail_hip_batch_free.argtypes = [ctypes.c_void_p]
    return L


def _bufs(n):
    arrs = {k: np.zeros(n, dtype=np.int32) for k in ("score", "end_query", "end_ref", "flag")}
    br = P.BatchResult()
    for k, v in arrs.items():
        setattr(br, k, v.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return arrs, br


@pytest.mark.gpu
def test_resident_batch_single_and_multi(mats):
    """The HBM-resident batch (parasail_hip_batch_create/align) returns what
    the per-call batch entry point returns (itself checked against the
    oracle above), and parasail_hip_batch_align_multi returns, pass by pass,
    what the single-algorithm calls return, whatever the order of its
    algorithms and on repeated calls."""
    from parasail_amd import workloads as WL
    pm, om = mats["blosum62"]
    b = WL.config1(seed=21, npairs=3000)
    n = len(b)
    L = _resident_api()
    I64, I32 = ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int)
    qoff = np.ascontiguousarray(b.qoff, dtype=np.int64)
    toff = np.ascontiguousarray(b.toff, dtype=np.int64)
    qlen = np.ascontiguousarray(b.qlen, dtype=np.int32)
    tlen = np.ascontiguousarray(b.tlen, dtype=np.int32)
    h = L.parasail_hip_batch_create(b.q, qoff.ctypes.data_as(I64), qlen.ctypes.data_as(I32), b.t,
                                    toff.ctypes.data_as(I64), tlen.ctypes.data_as(I32), n, pm)
    assert h
    try:
        single = {}
        for name in ("sw", "nw", "sg", "sg_qb_de"):
            arrs, br = _bufs(n)
            assert L.parasail_hip_batch_align(h, name.encode(), 11, 1, ctypes.byref(br)) == 0
            ref = P.align_arrays(name, 0, 0, b, 11, 1, pm)
            for key in ("score", "end_query", "end_ref"):
                assert (arrs[key] == ref[key]).all(), (name, key)
            _resident_flags_agree(name, arrs["flag"], ref["flag"])
            single[name] = arrs
        # spot check against the oracle
        for i in range(0, n, 300):
            q = bytes(b.q[b.qoff[i]:b.qoff[i] + b.qlen[i]])
            t = bytes(b.t[b.toff[i]:b.toff[i] + b.tlen[i]])
            for name in ("sw", "nw", "sg"):
                alg, fl = oracle_alg(name)
                exp, _ = O.align(alg, O.SCORE, q, t, 11, 1, om, fl)
                assert (single[name]["score"][i], single[name]["end_query"][i],
                        single[name]["end_ref"][i]) == tuple(exp[:3]), (name, i)
        for order in (("sw", "nw", "sg"), ("sg_qb_de", "sw"), ("nw", "sg", "sw", "sg_qb_de")):
            outs = (P.BatchResult * len(order))()
            arrs = []
            for k in range(len(order)):
                a, br = _bufs(n)
                arrs.append(a)
                outs[k] = br
            names = (ctypes.c_char_p * len(order))(*[x.encode() for x in order])
            assert L.parasail_hip_batch_align_multi(h, len(order), names, 11, 1, outs) == 0
            for k, name in enumerate(order):
                for key in ("score", "end_query", "end_ref", "flag"):
                    assert (arrs[k][key] == single[name][key]).all(), (order, name, key)
                km, mm = ctypes.c_double(), ctypes.c_double()
                assert L.parasail_hip_last_pass_ms(k, ctypes.byref(km), ctypes.byref(mm)) == 0
                assert km.value > 0 and 0 < mm.value <= km.value + 1e-3
            assert L.parasail_hip_last_pass_ms(len(order), None, None) == -1
    finally:
        L.parasail_hip_batch_free(h)


def test_autotuned_planner_same_results(mats):
    """parasail_hip_autotune measures a step cost for every rows-per-lane
    choice (SW and NW) and later plans use them: the shapes may change, the
    results may not -- every pair equals its untuned result, which equals
    the oracle on a sample."""
    from parasail_amd import workloads as WL
    pm, om = mats["blosum62"]
    L = P.lib()
    L.parasail_hip_autotune.restype = ctypes.c_int
    L.parasail_hip_autotune.argtypes = [P.MP, ctypes.c_int, ctypes.c_int]
    L.parasail_hip_autotune_table.restype = ctypes.c_int
    L.parasail_hip_autotune_table.argtypes = [ctypes.POINTER(ctypes.c_int),
                                              ctypes.POINTER(ctypes.c_double),
                                              ctypes.POINTER(ctypes.c_double), ctypes.c_int]
    b = WL.config1(seed=23, npairs=4000)
    before = {name: P.align_arrays(name, 0, 0, b, 11, 1, pm) for name in ("sw", "nw", "sg")}
    assert L.parasail_hip_autotune_table(None, None, None, 0) == 0
    try:
        assert L.parasail_hip_autotune(pm, 11, 1) == 0
        rows = (ctypes.c_int * 16)()
        swc, nwc = (ctypes.c_double * 16)(), (ctypes.c_double * 16)()
        k = L.parasail_hip_autotune_table(rows, swc, nwc, 16)
        assert k == 9 and [rows[i] for i in range(k)] == [4, 6, 8, 10, 12, 14, 16, 18, 20]
        for i in range(k):
            assert swc[i] > 0 and nwc[i] > 0, (i, swc[i], nwc[i])
        # more rows per lane, more work per step
        assert swc[k - 1] > swc[0] and nwc[k - 1] > nwc[0]
        for name in ("sw", "nw", "sg"):
            after = P.align_arrays(name, 0, 0, b, 11, 1, pm)
            for key in ("score", "end_query", "end_ref", "flag"):
                assert (after[key] == before[name][key]).all(), (name, key)
    finally:
        L.parasail_hip_autotune_reset()
    assert L.parasail_hip_autotune_table(None, None, None, 0) == 0
    for i in range(0, len(b), 400):
        q = bytes(b.q[b.qoff[i]:b.qoff[i] + b.qlen[i]])
        t = bytes(b.t[b.toff[i]:b.toff[i] + b.tlen[i]])
        exp, _ = O.align(O.SW, O.SCORE, q, t, 11, 1, om)
        assert (before["sw"]["score"][i], before["sw"]["end_query"][i],
                before["sw"]["end_ref"][i]) == tuple(exp[:3]), i


def test_resident_batch_graph_replays(mats):
    """A multi-pass call repeating the previous one is captured into a HIP
    graph and later ones replay it: every call's results equal the eager
    single-algorithm calls', with eager work on the same batch and device
    (single passes, a per-call alignment) between the replays."""
    from parasail_amd import workloads as WL
    pm, om = mats["blosum62"]
    b = WL.config1(seed=22, npairs=3000)
    n = len(b)
    L = _resident_api()
    L.parasail_hip_batch_graph_stats.restype = ctypes.c_int
    L.parasail_hip_batch_graph_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int),
                                                 ctypes.POINTER(ctypes.c_int)]
    L.parasail_hip_batch_set_graph.restype = ctypes.c_int
    L.parasail_hip_batch_set_graph.argtypes = [ctypes.c_void_p, ctypes.c_int]
    I64, I32 = ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int)
    qoff = np.ascontiguousarray(b.qoff, dtype=np.int64)
    toff = np.ascontiguousarray(b.toff, dtype=np.int64)
    qlen = np.ascontiguousarray(b.qlen, dtype=np.int32)
    tlen = np.ascontiguousarray(b.tlen, dtype=np.int32)
    h = L.parasail_hip_batch_create(b.q, qoff.ctypes.data_as(I64), qlen.ctypes.data_as(I32), b.t,
                                    toff.ctypes.data_as(I64), tlen.ctypes.data_as(I32), n, pm)
    assert h
    order = ("sw", "nw", "sg")
    try:
        single = {}
        for name in order:
            arrs, br = _bufs(n)
            assert L.parasail_hip_batch_align(h, name.encode(), 11, 1, ctypes.byref(br)) == 0
            single[name] = arrs
        names = (ctypes.c_char_p * len(order))(*[x.encode() for x in order])
        prev_times = None
        for call in range(6):
            outs = (P.BatchResult * len(order))()
            arrs = []
            for k in range(len(order)):
                a, br = _bufs(n)
                arrs.append(a)
                outs[k] = br
            assert L.parasail_hip_batch_align_multi(h, len(order), names, 11, 1, outs) == 0
            for k, name in enumerate(order):
                for key in ("score", "end_query", "end_ref", "flag"):
                    assert (arrs[k][key] == single[name][key]).all(), (call, name, key)
            # each call's pass events are recorded anew (a replay whose events
            # kept an earlier call's state would let the host read results
            # before their download ended)
            times = []
            for k in range(len(order)):
                km, mm = ctypes.c_double(), ctypes.c_double()
                assert L.parasail_hip_last_pass_ms(k, ctypes.byref(km), ctypes.byref(mm)) == 0
                assert km.value > 0 and 0 < mm.value <= km.value + 1e-3, (call, k, km.value, mm.value)
                times.append((km.value, mm.value))
            assert times != prev_times, (call, times)
            prev_times = times
            # eager work between replays: a single pass over the same batch
            # and a per-call alignment
            a, br = _bufs(n)
            assert L.parasail_hip_batch_align(h, b"nw", 11, 1, ctypes.byref(br)) == 0
            assert (a["score"] == single["nw"]["score"]).all(), call
            r = P.align("parasail_sw_striped_hip_64_16", "HEAGAWGHEE", "PAWHEAE", 10, 1, pm)
            assert r["score"] == O.align(O.SW, O.SCORE, "HEAGAWGHEE", "PAWHEAE", 10, 1, om)[0][0]
        cap, rep = ctypes.c_int(), ctypes.c_int()
        assert L.parasail_hip_batch_graph_stats(h, ctypes.byref(cap), ctypes.byref(rep)) == 0
        # call 0 repeats no call, call 1 is captured, calls 2-5 replay
        assert (cap.value, rep.value) == (1, 4), (cap.value, rep.value)
        assert L.parasail_hip_batch_set_graph(h, 0) == 1
    finally:
        L.parasail_hip_batch_free(h)


# ------------------------------------------------ saturation windows
#
# The reference's fixed-width kernels report saturation by predicates of
# their own (align.c restates them); here they are applied, independently,
# to the exact DP values of a plain Python DP over every cell.

def _dp_values(alg, flags, q, t, M, o, g):
    """(min of H/E/F over cells and boundaries, max cell H, min cell H,
    max length count of a cell's H) -- the recurrences of src/sw_stats.c,
    src/sg_stats.c, src/nw_stats.c cell by cell."""
    m, n = len(q), len(t)
    NEGI = -(1 << 30)
    s1b, s1e, s2b, s2e = flags if alg == O.SG else (0, 0, 0, 0)
    sw = alg == O.SW
    Hp = [0] + [0 if (sw or s2b) else -o - (j - 1) * g for j in range(1, n + 1)]
    Fp = [NEGI] * (n + 1)
    Lp = [0] * (n + 1)
    FLp = [0] * (n + 1)
    vmin = min(Hp)
    vmax = vminh = vmaxl = 0
    for i in range(1, m + 1):
        Hc = [0 if (sw or s1b) else -o - (i - 1) * g] + [0] * n
        vmin = min(vmin, Hc[0])
        Lc = [0] * (n + 1)
        FLc = [0] * (n + 1)
        Fc = [NEGI] * (n + 1)
        E, EL = NEGI, 0
        for j in range(1, n + 1):
            f_opn, f_ext = Hp[j] - o, Fp[j] - g
            F = max(f_opn, f_ext)
            e_opn, e_ext = Hc[j - 1] - o, E - g
            E = max(e_opn, e_ext)
            hd = Hp[j - 1] + int(M[q[i - 1], t[j - 1]])
            H = max(hd, E, F, 0) if sw else max(hd, E, F)
            fl = (Lp[j] if f_opn > f_ext else FLp[j]) + 1
            EL = (Lc[j - 1] if e_opn > e_ext else EL) + 1
            L = Lp[j - 1] + 1 if H == hd else (fl if H == F else EL)
            if sw and H <= 0:
                L = 0
            Hc[j], Fc[j], Lc[j], FLc[j] = H, F, L, fl
            vmin = min(vmin, H, E, F)
            vmax = max(vmax, H)
            vminh = min(vminh, H)
            vmaxl = max(vmaxl, L)
        Hp, Fp, Lp, FLp = Hc, Fc, Lc, FLc
    return vmin, vmax, vminh, vmaxl


def _ref_saturates(alg, kind, scan, w, o, mx_max, mx_min, score, vals):
    vmin, vmax, vminh, vmaxl = vals
    pos = (1 << (w - 1)) - 1 - mx_max - 1
    neg = ((-(1 << (w - 1)) + o) if -o < mx_min else (-(1 << (w - 1)) - mx_min)) + 1
    if alg == O.SW:
        if not scan and kind != "trace":
            # biased by INT_MIN(w) (sw_striped_avx2_256_16.c:205-210,350-362);
            # a count clamps at 2^w - 1 (sw_stats_striped_avx2_256_16.c:440-442)
            return score > (1 << w) - mx_max - 2 or (kind == "stats" and vmaxl >= (1 << w) - 1)
        return score > pos or (kind == "stats" and vmaxl > pos)
    if kind == "stats":
        return vminh < neg or vmax > pos or vmaxl > pos
    if scan and kind == "score" and alg == O.NW:
        return vminh < neg or vmax > pos
    return vmin < neg or vmax > pos


def test_saturation_windows_by_kind_and_strategy(mats):
    """Every fixed width (8, 16) and sat of SW, NW and SG, score/_stats/
    _trace, striped and scan: saturation and the saturated outputs follow
    the reference's kernel of that variant (stats counts included:
    sg_stats_striped_avx2_256_8.c:474-477, sw_stats_striped_avx2_256_8.c:
    473-475,645); unsaturated results equal the oracle."""
    rng = random.Random(77)
    pm, om = mats["blosum62"]
    M = om.values()
    a150 = rseq(rng, 150, weights=RR)
    b150 = rseq(rng, 150, weights=RR)
    r300 = rseq(rng, 300, weights=RR)
    cases = [(a150, b150), (r300, mutate(rng, r300, AA, sub=0.45, indel=0.02)),
             (rseq(rng, 40), rseq(rng, 45)), ("W" * 14, "W" * 14)]
    for s1, s2 in cases:
        q, t = om.map(s1), om.map(s2)
        for name in ("sw", "nw", "sg"):
            alg, flags = oracle_alg(name)
            vals = _dp_values(alg, flags, q, t, M, 11, 1)
            for kind in ("score", "stats", "trace"):
                okind = {"score": O.SCORE, "stats": O.STATS, "trace": O.TRACE}[kind]
                exp, _ = O.align(alg, okind, s1, s2, 11, 1, om, flags)
                suffix = {"score": "", "stats": "_stats", "trace": "_trace"}[kind]
                for scan in (0, 1):
                    strat = "scan" if scan else "striped"
                    sat_w = [w for w in (8, 16) if not _ref_saturates(
                        alg, kind, scan, w, 11, om.max, om.min, exp[0], vals)]
                    for w in ("8", "16", "sat"):
                        fn = "parasail_%s%s_%s_hip_64_%s" % (name, suffix, strat, w)
                        got = P.align(fn, s1, s2, 11, 1, pm)
                        ctx = (fn, len(s1), len(s2), vals, exp)
                        if w == "sat":
                            ew = sat_w[0] if sat_w else 32
                            bits = {8: P.FLAG_BITS_8, 16: P.FLAG_BITS_16, 32: P.FLAG_BITS_32}[ew]
                            assert got["flag"] & bits and not got["flag"] & P.FLAG_SATURATED, ctx
                            assert got["score"] == exp[0], ctx
                            continue
                        sat = int(w) not in sat_w
                        assert bool(got["flag"] & P.FLAG_SATURATED) == sat, ctx
                        if not sat:
                            assert (got["score"], got["end_query"], got["end_ref"]) == exp[:3], ctx
                            continue
                        wi = int(w)
                        if alg == O.SW and not scan and kind == "score":
                            sv = (1 << wi) - 1
                        elif alg == O.SW and not scan and kind == "stats":
                            sv = 1 << (wi - 1)
                        else:
                            sv = 0
                        assert (got["score"], got["end_query"], got["end_ref"]) == (sv, 0, 0), ctx
                        if kind == "stats":
                            assert (got["matches"], got["similar"], got["length"]) == (sv,) * 3, ctx
    # the case the reference's rule is about: a ~150 aa pair saturates the
    # 8-bit sg _stats kernel through its counts while its score fits
    assert _dp_values(O.SG, (1, 1, 1, 1), om.map(a150), om.map(b150), M, 11, 1)[3] > 115


def test_batch_stats_sat_widths_follow_counts(mats):
    """Batched SW/NW/SG _stats and _trace with width sat report, per pair,
    the width the per-call sat function reports (counts included)."""
    rng = random.Random(78)
    pm, om = mats["blosum62"]
    qs, ts = [], []
    for k in range(24):
        a = rseq(rng, rng.randint(20, 320), weights=RR)
        b = mutate(rng, a, AA, sub=0.4, indel=0.02) if k % 2 else rseq(rng, rng.randint(20, 320))
        qs.append(a)
        ts.append(b)
    for name in ("sw", "nw", "sg", "sg_qx"):
        for kind, kname in ((1, "_stats"), (2, "_trace")):
            got = P.align_batch(name, kind, 0, qs, ts, 11, 1, pm)
            for k in range(len(qs)):
                one = P.align("parasail_%s%s_striped_hip_64_sat" % (name, kname), qs[k], ts[k],
                              11, 1, pm)
                assert int(got["flag"][k]) == one["flag"], (name, kind, k)
                assert int(got["score"][k]) == one["score"]
                if kind == 1:
                    assert (int(got["length"][k]), int(got["matches"][k])) == \
                        (one["length"], one["matches"])


def test_batch_sat_nw_sg_reports_narrowest_width(mats):
    """A `_sat` NW/SG batch reports, pair by pair, the width of the per-call
    `_sat` function (src/satcheck.c tries 8, then 16): 8 where every H/E/F
    of the table and its boundary stays in the 8-bit window (the reference's
    nw/sg_striped_*_8 test, sg_striped_avx2_256_8.c:317-320), which the
    oracle's table decides here for the short pairs.  The batch's short
    pairs run on the packed kernel's extrema variant, the long ones are
    settled by their boundary.  A fixed-width 8 batch saturates the same
    pairs as the per-call `_8` function."""
    rng = random.Random(91)
    pm, om = mats["blosum62"]
    M = om.values()
    qs, ts = [], []
    for k in range(48):
        a = rseq(rng, rng.randint(8, 100), weights=RR)
        b = mutate(rng, a, AA, sub=0.2, indel=0.03) if k % 3 == 0 else rseq(rng, rng.randint(8, 100))
        qs.append(a)
        ts.append(b)
    qs += ["W" * 12, "W" * 9, rseq(rng, 300), rseq(rng, 104), rseq(rng, 40)]
    ts += ["W" * 12, "W" * 9, rseq(rng, 280), rseq(rng, 30), rseq(rng, 110)]
    bits = P.FLAG_BITS_8 | P.FLAG_BITS_16 | P.FLAG_BITS_32
    for name in ("nw", "sg", "sg_qx", "sg_db"):
        alg, flags = oracle_alg(name)
        got = P.align_batch(name, 0, 0, qs, ts, 11, 1, pm)
        got8 = P.align_batch(name, 0, 8, qs, ts, 11, 1, pm)
        exp, _ = O.batch(alg, O.SCORE, [om.map(q) for q in qs], [om.map(t) for t in ts],
                         11, 1, om, flags)
        widths = set()
        for k, (q, t) in enumerate(zip(qs, ts)):
            ctx = (name, k, len(q), len(t))
            one = P.align("parasail_%s_striped_hip_64_sat" % name, q, t, 11, 1, pm)
            assert int(got["score"][k]) == one["score"] == exp[k, 0], ctx
            assert int(got["flag"][k]) & bits == one["flag"] & bits, ctx
            if max(len(q), len(t)) <= 120:
                vals = _dp_values(alg, flags, om.map(q), om.map(t), M, 11, 1)
                sat8 = _ref_saturates(alg, "score", False, 8, 11, 11, -4, exp[k, 0], vals)
                want = P.FLAG_BITS_16 if sat8 else P.FLAG_BITS_8
                assert int(got["flag"][k]) & bits == want, ctx
            one8 = P.align("parasail_%s_striped_hip_64_8" % name, q, t, 11, 1, pm)
            assert int(got8["flag"][k]) == one8["flag"], ctx
            assert int(got8["score"][k]) == one8["score"], ctx
            widths.add(int(got["flag"][k]) & bits)
        assert widths == {P.FLAG_BITS_8, P.FLAG_BITS_16}, name


def test_batch_gap_above_open(mats):
    """Extension dearer than opening (open 2, extend 3): NW and SG batches
    equal the oracle (the global boundary row -open - (j-1) gap is then not
    the cheapest run of gaps from the corner)."""
    rng = random.Random(92)
    pm, om = mats["blosum62"]
    qs = [rseq(rng, rng.randint(1, 400), weights=RR) for _ in range(120)]
    ts = [rseq(rng, rng.randint(1, 400), weights=RR) for _ in range(120)]
    for name in ("nw", "sg", "sg_qb", "sg_db"):
        alg, flags = oracle_alg(name)
        got = P.align_batch(name, 0, 0, qs, ts, 2, 3, pm)
        exp, _ = O.batch(alg, O.SCORE, [om.map(q) for q in qs], [om.map(t) for t in ts],
                         2, 3, om, flags)
        for k in range(len(qs)):
            assert (int(got["score"][k]), int(got["end_query"][k]), int(got["end_ref"][k])) == \
                tuple(int(v) for v in exp[k, :3]), (name, k, len(qs[k]), len(ts[k]))


def _resident_flags_agree(name, res, batch):
    """The resident batch's flags against parasail_hip_align_batch's at
    width sat: the same for SW; for NW/SG the same but for the width, which
    the resident batch reports as its tier's (16, or 32 for the int32 tier;
    never below the sat function's)."""
    bits = P.FLAG_BITS_8 | P.FLAG_BITS_16 | P.FLAG_BITS_32
    if name == "sw":
        assert (res == batch).all()
        return
    assert ((res & ~bits) == (batch & ~bits)).all(), name
    rb, bb = res & bits, batch & bits
    assert np.isin(rb, (P.FLAG_BITS_16, P.FLAG_BITS_32)).all(), name
    assert (rb[bb == P.FLAG_BITS_32] == P.FLAG_BITS_32).all(), name


def _concat(seqs):
    bs = [s.encode() for s in seqs]
    ln = np.array([len(x) for x in bs], dtype=np.int32)
    off = np.zeros(len(bs), dtype=np.int64)
    off[1:] = np.cumsum(ln, dtype=np.int64)[:-1]
    return b"".join(bs), off, ln


def test_batch_int16_overflow_redo(mats):
    """Config-1-shaped batches with pairs the packed-int16 kernel cannot
    hold.  SW: with a match score of 120, self pairs of 300 and 700 residues
    score past the kernel's overflow limit 0x7BFF - (max + open + 1) (its
    rows' maximum is taken on binary16 bit patterns, exact below 0x7C00) and
    are redone by the int32 tier after the run-time test of
    src/sw_striped_avx2_256_16.c:350-362; the 300 one fits the biased 16-bit
    window (2^16 - max - 2), the 700 one only 32 bits, and width 16 reports
    it saturated (src/sw_striped_avx2_256_16.c:374-382).  NW/SG: a pair past
    the provable int16 bound runs on an int32 tier (the resident batch's
    flag reads 32).  Scores and ends equal the oracle through
    parasail_hip_align_batch and the resident batch; the batch's sat flags
    are the per-call sat function's."""
    rng = random.Random(90)
    pb, ob = mats["blosum62"]
    alpha = "ACDEFGHIKLMNPQRSTVWY"
    pbig, obig = P.matrix_create(alpha, 120, -10), O.matrix_create(alpha, 120, -10)
    qs, ts = [], []
    for _ in range(300):
        a = rseq(rng, int(min(1200, max(20, rng.lognormvariate(np.log(300), 0.6)))), weights=RR)
        b = rseq(rng, int(min(1200, max(20, rng.lognormvariate(np.log(300), 0.6)))), weights=RR)
        qs.append(a)
        ts.append(b)
    s300, s700 = rseq(rng, 300), rseq(rng, 700)
    lng = rseq(rng, 1700, weights=RR)
    L = _resident_api()
    I64, I32 = ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int)
    # SW over the big-match pairs; NW and SG past the int16 bound with a
    # 1,700-residue pair at extend 10 (every value provably inside int16
    # needs (m + n) * extend below ~32,000)
    for name, gap, pm, om in (("sw", 1, pbig, obig), ("nw", 10, pb, ob), ("sg", 10, pb, ob)):
        alg, fl = oracle_alg(name)
        qn, tn = list(qs), list(ts)
        if name == "sw":
            qn[40], tn[40] = s300, s300
            qn[170], tn[170] = s700, s700
            # just below the limit (int16), between the limit and int16's
            # own bound, and at that bound
            for k, ln in ((80, 263), (120, 265), (200, 272)):
                qn[k], tn[k] = s300[:ln], s300[:ln]
        qn[260], tn[260] = lng, mutate(rng, lng, AA, 0.1, 0.02)
        qcat, qoff, qlen = _concat(qn)
        tcat, toff, tlen = _concat(tn)
        h = L.parasail_hip_batch_create(qcat, qoff.ctypes.data_as(I64), qlen.ctypes.data_as(I32),
                                        tcat, toff.ctypes.data_as(I64), tlen.ctypes.data_as(I32),
                                        len(qn), pm)
        assert h
        try:
            got = P.align_batch(name, 0, 0, qn, tn, 11, gap, pm)
            arrs, br = _bufs(len(qn))
            assert L.parasail_hip_batch_align(h, name.encode(), 11, gap, ctypes.byref(br)) == 0
        finally:
            L.parasail_hip_batch_free(h)
        bits = P.FLAG_BITS_8 | P.FLAG_BITS_16 | P.FLAG_BITS_32
        for key in ("score", "end_query", "end_ref") + (("flag",) if name == "sw" else ()):
            assert (arrs[key] == got[key]).all(), (name, key)
        for k in range(len(qn)):
            exp, _ = O.align(alg, O.SCORE, qn[k], tn[k], 11, gap, om, fl)
            assert (got["score"][k], got["end_query"][k], got["end_ref"][k]) == exp[:3], \
                (name, k, len(qn[k]), len(tn[k]))
        if name != "sw":
            # the resident batch reports its tier's width (int32 for the
            # long pair); the batch the per-call sat function's, whose
            # windows follow the table's values
            res_w = np.where(np.arange(len(qn)) == 260, P.FLAG_BITS_32, P.FLAG_BITS_16)
            assert ((arrs["flag"] & bits) == res_w).all()
            assert ((arrs["flag"] & ~bits) == (got["flag"] & ~bits)).all()
            for k in (0, 1, 2, 260):
                one = P.align("parasail_%s_striped_hip_64_sat" % name, qn[k], tn[k], 11, gap, pm)
                assert int(got["flag"][k]) & bits == one["flag"] & bits, (name, k)
            continue
        ovf = 0x7BFF - (120 + 11 + 1)
        assert ovf < got["score"][40] <= 65535 - 122 < got["score"][170]
        assert got["score"][80] == 263 * 120 <= ovf < 265 * 120 == got["score"][120]
        assert got["score"][200] == 272 * 120 > 32767 - (120 + 11 + 1)
        for k in (40, 170, 260):
            one = P.align("parasail_sw_striped_hip_64_sat", qn[k], tn[k], 11, gap, pm)
            assert int(got["flag"][k]) & bits == one["flag"] & bits, k
        assert got["flag"][40] & P.FLAG_BITS_16 and got["flag"][170] & P.FLAG_BITS_32
        g16 = P.align_batch(name, 0, 16, qn, tn, 11, gap, pm)
        assert g16["flag"][170] & P.FLAG_SATURATED
        assert (g16["score"][170], g16["end_query"][170], g16["end_ref"][170]) == (65535, 0, 0)
        assert not g16["flag"][40] & P.FLAG_SATURATED and g16["score"][40] == got["score"][40]


def test_file_to_resident_batch(tmp_path, fixtures, mats):
    """FASTA records (one file plain, one gzip) read by
    parasail_sequences_from_file, packed by parasail_sequences_pack, split
    at the '$' separators into offsets and lengths, made an HBM-resident
    batch (record i of one file against record i of the other) and aligned:
    scores and ends equal the oracle's."""
    import gzip
    pm, om = mats["blosum62"]
    recs = fixtures["fasta"]["test_P0_10.fasta"]
    fa = tmp_path / "a.fa"
    fb = tmp_path / "b.fa.gz"
    fa.write_text("".join(">%s\n%s\n" % (nm, s) for nm, s in recs))
    fb.write_bytes(gzip.compress("".join(">%s x\n%s\n" % (nm, s[::-1][:300])
                                         for nm, s in recs).encode()))
    L = _resident_api()

    class Str(ctypes.Structure):
        _fields_ = [("l", ctypes.c_size_t), ("s", ctypes.c_char_p)]

    class Seq(ctypes.Structure):
        _fields_ = [("name", Str), ("comment", Str), ("seq", Str), ("qual", Str)]

    class Seqs(ctypes.Structure):
        _fields_ = [("seqs", ctypes.POINTER(Seq)), ("l", ctypes.c_size_t),
                    ("characters", ctypes.c_size_t), ("shortest", ctypes.c_size_t),
                    ("longest", ctypes.c_size_t), ("mean", ctypes.c_float),
                    ("stddev", ctypes.c_float)]

    L.parasail_sequences_from_file.restype = ctypes.POINTER(Seqs)
    L.parasail_sequences_from_file.argtypes = [ctypes.c_char_p]
    L.parasail_sequences_pack.restype = ctypes.c_void_p
    L.parasail_sequences_pack.argtypes = [ctypes.POINTER(Seqs), ctypes.POINTER(ctypes.c_size_t)]
    L.parasail_sequences_free.argtypes = [ctypes.POINTER(Seqs)]

    def packed(path):
        p = L.parasail_sequences_from_file(str(path).encode())
        assert p
        n = ctypes.c_size_t()
        buf = L.parasail_sequences_pack(p, ctypes.byref(n))
        data = ctypes.string_at(buf, n.value)
        L._libc_free(buf)
        L.parasail_sequences_free(p)
        # record i starts after the i-th '$'
        ends = [i for i, c in enumerate(data) if c == ord("$")]
        starts = [0] + [e + 1 for e in ends[:-1]]
        return data, np.array(starts, dtype=np.int64), \
            np.array([e - s for s, e in zip(starts, ends)], dtype=np.int32)

    qd, qoff, qlen = packed(fa)
    td, toff, tlen = packed(fb)
    assert len(qlen) == len(tlen) == len(recs) and (qlen > 0).all() and (tlen > 0).all()
    I64, I32 = ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int)
    h = L.parasail_hip_batch_create(qd, qoff.ctypes.data_as(I64), qlen.ctypes.data_as(I32), td,
                                    toff.ctypes.data_as(I64), tlen.ctypes.data_as(I32),
                                    len(qlen), pm)
    assert h
    try:
        for name in ("sw", "nw", "sg"):
            alg, fl = oracle_alg(name)
            arrs, br = _bufs(len(qlen))
            assert L.parasail_hip_batch_align(h, name.encode(), 11, 1, ctypes.byref(br)) == 0
            for k, (_, s) in enumerate(recs):
                exp, _ = O.align(alg, O.SCORE, s, s[::-1][:300], 11, 1, om, fl)
                assert (arrs["score"][k], arrs["end_query"][k], arrs["end_ref"][k]) == exp[:3], \
                    (name, k)
    finally:
        L.parasail_hip_batch_free(h)


def test_pssm_from_matrix_file(tmp_path, fixtures):
    """A PSSM read by parasail_matrix_from_file from the reference's
    data/cd02249_matrix_safe.txt, aligned per call and batched, equals the
    oracle with the same position rows (parasail_matrix_pssm_create)."""
    L = P.lib()
    L.parasail_matrix_from_file.restype = P.MP
    L.parasail_matrix_from_file.argtypes = [ctypes.c_char_p]
    text = fixtures["matrix_files"]["cd02249_matrix_safe.txt"]
    path = tmp_path / "pssm.txt"
    path.write_text(text)
    pm = L.parasail_matrix_from_file(str(path).encode())
    assert pm and pm.contents.type == 1
    mc = pm.contents
    alphabet = mc.alphabet.decode()
    vals = np.array([mc.matrix[k] for k in range(mc.size * mc.length)], dtype=np.int32)
    om = O.matrix_pssm_create(alphabet, vals, mc.length)
    rng = random.Random(41)
    ts = ["".join(rng.choice(alphabet) for _ in range(rng.randint(20, 300))) for _ in range(24)]
    for name in ("sw", "nw", "sg", "sg_qb_de"):
        alg, fl = oracle_alg(name)
        got = P.align_batch(name, 0, 0, None, ts, 11, 1, pm)
        for k, t in enumerate(ts):
            exp, _ = O.align(alg, O.SCORE, None, t, 11, 1, om, fl)
            assert (got["score"][k], got["end_query"][k], got["end_ref"][k]) == exp[:3], (name, k)
            if k < 6:
                one = P.align("parasail_%s_striped_hip_64_sat" % name, None, t, 11, 1, pm)
                assert (one["score"], one["end_query"], one["end_ref"]) == exp[:3], (name, k)
    L.parasail_matrix_free(pm)
