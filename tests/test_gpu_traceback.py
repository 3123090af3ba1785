"""Alignment strings, the printed alignment and the SSW emulation on hip
trace results, against the oracle.

parasail_result_get_traceback[_extra] (src/traceback_template.c:13-214) is
checked against the strings the oracle's CIGAR spells out (a diagonal run
prints both residues and '|', ':' or '.', an I run the query against '-',
a D run '-' against the target; SW leaves out the run a walk takes along
the table's edge), parasail_traceback_generic_extra2's text (:216-430)
against those strings laid out in blocks, and parasail_ssw
(src/pssw.c:45-183) against the oracle's forward SW, its SW over the
reversed prefixes and its CIGAR of the aligned stretch."""
import ctypes
import os
import random

import pytest

import oracle as O
import parasail_amd as P

pytestmark = pytest.mark.gpu

AA = "ARNDCQEGHILKMFPSTWYV"


@pytest.fixture(scope="module", autouse=True)
def need_device():
    if not P.available():
        pytest.skip("no HIP device")


class Traceback(ctypes.Structure):
    _fields_ = [("query", ctypes.c_char_p), ("comp", ctypes.c_char_p), ("ref", ctypes.c_char_p)]


class SswResult(ctypes.Structure):
    _fields_ = [("score1", ctypes.c_uint16), ("ref_begin1", ctypes.c_int32),
                ("ref_end1", ctypes.c_int32), ("read_begin1", ctypes.c_int32),
                ("read_end1", ctypes.c_int32), ("cigar", ctypes.POINTER(ctypes.c_uint32)),
                ("cigarLen", ctypes.c_int32)]


def _api():
    L = P.lib()
    L.parasail_result_get_traceback_extra.restype = ctypes.POINTER(Traceback)
    L.parasail_result_get_traceback_extra.argtypes = [
        P.RP, ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int, P.MP, ctypes.c_char,
        ctypes.c_char, ctypes.c_char, ctypes.c_int, ctypes.c_char_p]
    L.parasail_traceback_free.argtypes = [ctypes.POINTER(Traceback)]
    L.parasail_traceback_generic_extra2.restype = None
    L.parasail_traceback_generic_extra2.argtypes = [
        ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p,
        ctypes.c_char_p, P.MP, P.RP, ctypes.c_char, ctypes.c_char, ctypes.c_char, ctypes.c_int,
        ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p]
    L.parasail_ssw.restype = ctypes.POINTER(SswResult)
    L.parasail_ssw.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                               ctypes.c_int, ctypes.c_int, P.MP]
    L.parasail_result_ssw_free.argtypes = [ctypes.POINTER(SswResult)]
    return L


def oracle_alg(name):
    if name == "sw":
        return O.SW, (0, 0, 0, 0)
    if name == "nw":
        return O.NW, (0, 0, 0, 0)
    return O.SG, P.SG_FLAGS[name]


def expected_strings(alg, s1, s2, exp, tr, M, om, aliases=None, case_sensitive=False):
    c = O.cigar(s1, s2, alg == O.SG, exp[1], exp[2], tr)
    ops, bq, br = c
    runs = [(int(v) >> 4, "MIDNSHP=XB"[int(v) & 0xF]) for v in ops]
    i, j = bq, br
    if alg == O.SW and runs and runs[0][1] in "ID":
        # the run the walk takes along the table's edge (cigar_template.c:
        # 145-164) is not part of the printed local alignment
        cnt, op = runs.pop(0)
        i, j = (i + cnt, j) if op == "I" else (i, j + cnt)
    q, cmp, d = [], [], []
    qm, tm = om.map(s1), om.map(s2)
    for cnt, op in runs:
        for _ in range(cnt):
            if op in "=X":
                a, b = s1[i], s2[j]
                eq = a == b if case_sensitive else a.upper() == b.upper()
                if aliases:
                    al = aliases if case_sensitive else aliases.upper()
                    for k in range(len(al) - 1):
                        if al[k] == a:
                            eq |= al[k + 1] == b
                        elif al[k + 1] == a:
                            eq |= al[k] == b
                ch = "|" if eq else (":" if M[qm[i], tm[j]] > 0 else ".")
                q.append(a)
                d.append(b)
                cmp.append(ch)
                i += 1
                j += 1
            elif op == "I":
                q.append(s1[i])
                d.append("-")
                cmp.append(" ")
                i += 1
            else:
                q.append("-")
                d.append(s2[j])
                cmp.append(" ")
                j += 1
    return "".join(q), "".join(cmp), "".join(d)


def expected_print(qs, cs, ds, nameA, nameB, width, name_width, int_width, sw, score,
                   end_query, end_ref, lena, lenb):
    int_width = max(int_width, len(str(lena + lenb)))
    n = len(cs)
    c_ins, c_del = qs.count("-"), ds.count("-")
    qp = end_query + 1 - n + c_ins if sw else 0
    dp = end_ref + 1 - n + c_del if sw else 0
    out = []
    mch = sim = gaps = 0
    for at in range(0, n, width):
        blk = slice(at, min(n, at + width))
        out.append("\n")
        d0 = dp
        dp += sum(1 for ch in ds[blk] if ch != "-")
        out.append("%-*s %*d %s %*d\n" % (name_width, nameB[:name_width], int_width, d0 + 1,
                                          ds[blk], int_width, dp))
        out.append(" " * (name_width + 1 + int_width + 1) + cs[blk] + "\n")
        for ch in cs[blk]:
            mch += ch == "|"
            sim += ch in "|:"
            gaps += ch == " "
        q0 = qp
        qp += sum(1 for ch in qs[blk] if ch != "-")
        out.append("%-*s %*d %s %*d\n" % (name_width, nameA[:name_width], int_width, q0 + 1,
                                          qs[blk], int_width, qp))
    out.append("\nLength: %d\n" % n)
    out.append("Identity:   %*d/%d (%4.1f%%)\n" % (int_width, mch, n, 100.0 * mch / n))
    out.append("Similarity: %*d/%d (%4.1f%%)\n" % (int_width, sim, n, 100.0 * sim / n))
    out.append("Gaps:       %*d/%d (%4.1f%%)\n" % (int_width, gaps, n, 100.0 * gaps / n))
    out.append("Score: %d\n" % score)
    return "".join(out)


def test_traceback_strings_and_printed_form(tmp_path, fixtures):
    L = _api()
    libc = ctypes.CDLL(None)
    libc.fopen.restype = ctypes.c_void_p
    libc.fopen.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    libc.fclose.argtypes = [ctypes.c_void_p]
    pm, om = P.blosum62(), O.blosum62()
    M = om.values()
    rng = random.Random(11)
    pairs = [(s[0][1], s[1][1]) for s in (fixtures["fasta"]["test_small_2.fasta"],
                                           fixtures["fasta"]["test_case.fasta"],
                                           fixtures["fasta"]["test_P0_2.fasta"])]
    for _ in range(8):
        a = "".join(rng.choice(AA) for _ in range(rng.randint(5, 200)))
        b = a[rng.randint(0, len(a) // 2):] + "".join(rng.choice(AA) for _ in range(rng.randint(0, 30)))
        pairs.append((a, b))
    for name in ("sw", "nw", "sg", "sg_qb", "sg_de", "sg_qx"):
        alg, flags = oracle_alg(name)
        for k, (s1, s2) in enumerate(pairs):
            for strat in ("striped", "scan"):
                fn = "parasail_%s_trace_%s_hip_64_sat" % (name, strat)
                a, b = s1.encode(), s2.encode()
                r = P.function(fn)(a, len(a), b, len(b), 10, 1, pm)
                assert r
                try:
                    exp, tr = O.align(alg, O.TRACE, s1, s2, 10, 1, om, flags)
                    ctx = (fn, k)
                    for aliases, cs in ((None, 0), (b"AT", 0), (None, 1)):
                        tb = L.parasail_result_get_traceback_extra(r, a, len(a), b, len(b), pm,
                                                                   b"|", b":", b".", cs, aliases)
                        assert tb, ctx
                        got = (tb.contents.query.decode(), tb.contents.comp.decode(),
                               tb.contents.ref.decode())
                        L.parasail_traceback_free(tb)
                        want = expected_strings(alg, s1, s2, exp, tr, M, om,
                                                aliases.decode() if aliases else None, cs)
                        assert got == want, ctx + (aliases, cs)
                    if strat == "striped":
                        path = str(tmp_path / ("tb_%s_%d.txt" % (name, k)))
                        f = libc.fopen(path.encode(), b"w")
                        assert f
                        L.parasail_traceback_generic_extra2(
                            a, len(a), b, len(b), b"query_name", b"ref", pm, r, b"|", b":", b".",
                            50, 8, 1, 3, f, 0, None)
                        libc.fclose(f)
                        text = open(path).read()
                        qs, cs_, ds = expected_strings(alg, s1, s2, exp, tr, M, om)
                        want = expected_print(qs, cs_, ds, "query_name", "ref", 50, 8, 3,
                                              alg == O.SW, exp[0], exp[1], exp[2], len(s1),
                                              len(s2))
                        assert text == want, ctx
                        os.unlink(path)
                finally:
                    L.parasail_result_free(r)


def test_ssw_emulation():
    L = _api()
    pm, om = P.blosum62(), O.blosum62()
    rng = random.Random(12)
    for trial in range(30):
        a = "".join(rng.choice(AA) for _ in range(rng.randint(10, 400)))
        if trial % 2:
            core = a[rng.randint(0, len(a) // 3): rng.randint(len(a) // 2, len(a))]
            b = "".join(rng.choice(AA) for _ in range(rng.randint(0, 50))) + core + \
                "".join(rng.choice(AA) for _ in range(rng.randint(0, 50)))
        else:
            b = "".join(rng.choice(AA) for _ in range(rng.randint(10, 400)))
        r = L.parasail_ssw(a.encode(), len(a), b.encode(), len(b), 11, 1, pm)
        fwd, _ = O.align(O.SW, O.SCORE, a, b, 11, 1, om)
        l1, l2 = fwd[1] + 1, fwd[2] + 1
        rev, _ = O.align(O.SW, O.SCORE, a[:l1][::-1], b[:l2][::-1], 11, 1, om)
        off1, off2 = fwd[1] - rev[1], fwd[2] - rev[2]
        sa, sb = a[off1:off1 + rev[1] + 1], b[off2:off2 + rev[2] + 1]
        fin, tr = O.align(O.SW, O.TRACE, sa, sb, 11, 1, om)
        ops = O.cigar(sa, sb, False, fin[1], fin[2], tr)[0]
        assert r, trial
        rc = r.contents
        got = (rc.score1, rc.ref_begin1, rc.ref_end1, rc.read_begin1, rc.read_end1,
               [rc.cigar[k] for k in range(rc.cigarLen)])
        L.parasail_result_ssw_free(r)
        assert got == (fwd[0], off2, fwd[2], off1, fwd[1], [int(x) for x in ops]), trial
    # past the 16-bit windows the emulation returns NULL, as src/pssw.c:83-
    # 163 does: the forward pass's (biased, above 2^16 - max - 2) or the
    # traceback's (unbiased, above 2^15 - max - 2)
    big = P.matrix_create("ACDEFGHIKLMNPQRSTVWY", 120, -10)
    s = "".join(rng.choice(AA) for _ in range(700))
    assert not L.parasail_ssw(s.encode(), len(s), s.encode(), len(s), 11, 1, big)
    assert not L.parasail_ssw(s.encode(), 300, s.encode(), 300, 11, 1, big)
    r = L.parasail_ssw(s.encode(), 250, s.encode(), 250, 11, 1, big)
    assert r and r.contents.score1 == 250 * 120 and r.contents.cigarLen == 1
    assert (r.contents.read_begin1, r.contents.read_end1) == (0, 249)
    L.parasail_result_ssw_free(r)
