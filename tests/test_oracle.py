"""Pins the CPU oracle (oracle/oracle.c) before anything is compared with it.

The reference's sources cannot be compiled here, so the oracle is pinned by:
  * the reference's known-answer tests (tests/test_basic.c:20-85, via
    tests/golden/fixtures.json);
  * brute-force enumeration of every alignment path on tiny inputs, which
    fixes scores and, through closed-form argmin rules, the end positions;
  * brute force of every cell's trace code from path maxima (H source,
    E/F open against extend, the SW zero mask) for SW, NW and all 16 SG
    end-flag sets;
  * brute force of banded NW over the paths that stay inside the band of
    nw_banded.c:92-97, boundary moves included;
  * invariants: a CIGAR re-scored with the matrix and gaps gives the score,
    the _stats counts of SW, NW and every SG flag set equal the counts along
    the trace path into the stats end cell, PSSM from
    convert_square_to_pssm equals the square matrix, banded NW with a band
    covering the whole matrix equals NW, SG >= NW.
Each check is chosen so that a plausible slip (wrong boundary, swapped E/F,
wrong tie order, off-by-one end) fails it.
"""
import itertools
import random

import numpy as np
import pytest

import oracle as O

NEG = O.NEG_INF


# ---------------------------------------------------------------- helpers

def path_shapes(maxd):
    """All op strings (D=diag, U=consume s1, L=consume s2) per displacement."""
    shapes = {}
    for di in range(maxd + 1):
        for dj in range(maxd + 1):
            out = []

            def rec(i, j, ops):
                if i == di and j == dj:
                    out.append(ops)
                    return
                if i < di and j < dj:
                    rec(i + 1, j + 1, ops + "D")
                if i < di:
                    rec(i + 1, j, ops + "U")
                if j < dj:
                    rec(i, j + 1, ops + "L")
            rec(0, 0, "")
            shapes[(di, dj)] = out
    return shapes


SHAPES = path_shapes(4)


def path_score(ops, a, b, q, t, M, open_, gap):
    s, i, j, prev = 0, a, b, None
    for o in ops:
        if o == "D":
            s += M[q[i], t[j]]
            i += 1
            j += 1
        elif o == "U":
            s -= gap if prev == "U" else open_
            i += 1
        else:
            s -= gap if prev == "L" else open_
            j += 1
        prev = o
    return s


def brute_H(alg, q, t, M, open_, gap, flags=(0, 0, 0, 0)):
    """H[i][j] (1-based DP cells) = best score of a path from an allowed start
    to (i, j).  SW starts anywhere (and floors at 0), NW at (0,0), SG at (0,0)
    plus (a,0) if s1_beg and (0,b) if s2_beg."""
    m, n = len(q), len(t)
    H = np.full((m + 1, n + 1), NEG, dtype=np.int64)
    for i in range(1, m + 1):
        for j in range(1, n + 1):
            starts = []
            if alg == O.SW:
                starts = [(a, b) for a in range(i + 1) for b in range(j + 1)]
            else:
                starts = [(0, 0)]
                if alg == O.SG and flags[0]:
                    starts += [(a, 0) for a in range(1, i + 1)]
                if alg == O.SG and flags[2]:
                    starts += [(0, b) for b in range(1, j + 1)]
            best = 0 if alg == O.SW else NEG
            for (a, b) in starts:
                for ops in SHAPES[(i - a, j - b)]:
                    if not ops:
                        continue
                    best = max(best, path_score(ops, a, b, q, t, M, open_, gap))
            H[i, j] = best
    return H


def expected_ends(alg, H, flags):
    """Closed-form end rules (derived from sw.c:150-158, sg.c:171-235)."""
    m, n = H.shape[0] - 1, H.shape[1] - 1
    if alg == O.SW:
        S = H[1:, 1:].max()
        cells = [(i, j) for i in range(1, m + 1) for j in range(1, n + 1) if H[i, j] == S]
        jmin = min(j for _, j in cells)
        imin = min(i for i, j in cells if j == jmin)
        return S, imin - 1, jmin - 1
    if alg == O.NW:
        return H[m, n], m - 1, n - 1
    s1b, s1e, s2b, s2e = flags
    col = [(H[i, n], i) for i in range(1, m + 1)]
    row = [(H[m, j], j) for j in range(1, n + 1)]
    if s1e and s2e:
        S = max(max(v for v, _ in col), max(v for v, _ in row))
        rj = [j for v, j in row if v == S and j < n]
        if rj:
            return S, m - 1, min(rj) - 1
        ci = [i for v, i in col if v == S and i < m]
        if ci:
            return S, min(ci) - 1, n - 1
        return S, m - 1, n - 1
    if s1e:
        S = max(v for v, _ in col)
        return S, min(i for v, i in col if v == S) - 1, n - 1
    if s2e:
        S = max(v for v, _ in row)
        return S, m - 1, min(j for v, j in row if v == S) - 1
    return H[m, n], m - 1, n - 1


SG_FLAGS = [f for f in itertools.product((0, 1), repeat=4)]


def small_matrix(rng):
    kind = rng.randrange(3)
    if kind == 0:
        return O.matrix_create("ACG", rng.randint(1, 3), -rng.randint(1, 3)), "ACG"
    if kind == 1:
        return O.matrix_create("ACGT", 2, -3), "ACGT"
    return O.blosum62(), "ARNDCWK"


# ---------------------------------------------------------------- KATs

def _matrix_from_fixture(fixtures, spec):
    if isinstance(spec, list):
        return O.matrix_create(spec[0], spec[1], spec[2])
    d = fixtures["matrices"][spec]
    return O.matrix_from_values(d["alphabet"], d["values"], d["max"], d["min"])


def test_known_answers(fixtures):
    for kat in fixtures["kat"]:
        mx = _matrix_from_fixture(fixtures, kat["matrix"])
        res, _ = O.align(O.SW, O.SCORE, kat["s1"], kat["s2"], kat["open"], kat["gap"], mx)
        assert res[0] == kat["score"], kat["src"]


def test_blosum62_table_matches_reference(fixtures):
    d = fixtures["matrices"]["blosum62"]
    b = O.blosum62()
    assert b.values().ravel().tolist() == d["values"]
    assert (b.max, b.min, b.size) == (d["max"], d["min"], d["size"])
    ref = O.matrix_from_values(d["alphabet"], d["values"], d["max"], d["min"])
    assert (b.mapper == ref.mapper).all()
    assert b.mapper[ord("a")] == 0 and b.mapper[ord("J")] == 23 and b.mapper[ord("*")] == 23


# ---------------------------------------------------------------- brute force

@pytest.mark.parametrize("alg", [O.SW, O.NW])
def test_brute_force_sw_nw(alg):
    rng = random.Random(1000 + alg)
    for _ in range(60):
        mx, alpha = small_matrix(rng)
        M = mx.values()
        m, n = rng.randint(1, 4), rng.randint(1, 4)
        s1 = "".join(rng.choice(alpha) for _ in range(m))
        s2 = "".join(rng.choice(alpha) for _ in range(n))
        o, g = rng.choice([(0, 0), (1, 1), (3, 1), (5, 2), (2, 0)])
        q, t = mx.map(s1), mx.map(s2)
        H = brute_H(alg, q, t, M, o, g)
        exp = expected_ends(alg, H, None)
        for kind in (O.SCORE, O.STATS, O.TRACE):
            res, _ = O.align(alg, kind, s1, s2, o, g, mx)
            assert res[:3] == tuple(int(x) for x in exp), (alg, kind, s1, s2, o, g)


def test_brute_force_sg_all_flags():
    rng = random.Random(7)
    for flags in SG_FLAGS:
        for _ in range(12):
            mx, alpha = small_matrix(rng)
            M = mx.values()
            m, n = rng.randint(1, 4), rng.randint(1, 4)
            s1 = "".join(rng.choice(alpha) for _ in range(m))
            s2 = "".join(rng.choice(alpha) for _ in range(n))
            o, g = rng.choice([(1, 1), (3, 1), (5, 2), (2, 0)])
            q, t = mx.map(s1), mx.map(s2)
            H = brute_H(O.SG, q, t, M, o, g, flags)
            exp = expected_ends(O.SG, H, flags)
            for kind in (O.SCORE, O.STATS, O.TRACE):
                res, _ = O.align(O.SG, kind, s1, s2, o, g, mx, flags)
                assert res[:3] == tuple(int(x) for x in exp), (flags, kind, s1, s2, o, g)


# ---------------------------------------------------------------- invariants

def rescore_cigar(ops, bq, br, q, t, M, open_, gap, free_lead=(0, 0), free_trail=(0, 0)):
    """Score an encoded CIGAR; 'I' consumes the query (s1), 'D' the target."""
    runs = [(int(v) >> 4, "MIDNSHP=XB"[int(v) & 0xF]) for v in ops]
    s, i, j = 0, bq, br
    for k, (cnt, op) in enumerate(runs):
        if op in "=X":
            for _ in range(cnt):
                s += M[q[i], t[j]]
                i += 1
                j += 1
        else:
            lead = (k == 0)
            trail = (k == len(runs) - 1)
            free = (op == "I" and ((lead and free_lead[0]) or (trail and free_trail[0]))) or \
                   (op == "D" and ((lead and free_lead[1]) or (trail and free_trail[1])))
            if not free:
                s -= open_ + (cnt - 1) * gap
            if op == "I":
                i += cnt
            else:
                j += cnt
    return s, i, j


def rand_seq(rng, alpha, n):
    return "".join(rng.choice(alpha) for _ in range(n))


@pytest.mark.parametrize("alg", [O.SW, O.NW, O.SG])
def test_cigar_rescores_to_score(alg):
    rng = random.Random(31 + alg)
    b = O.blosum62()
    dna = O.matrix_create("ACGT", 2, -3)
    for trial in range(80):
        mx, alpha = (b, "ARNDCQEGHILKMFPSTWYV") if trial % 2 else (dna, "ACGT")
        M = mx.values()
        s1 = rand_seq(rng, alpha, rng.randint(1, 40))
        s2 = rand_seq(rng, alpha, rng.randint(1, 40))
        if trial % 5 == 0:
            s2 = s1[: max(1, len(s1) - 3)] + rand_seq(rng, alpha, 3)
        o, g = rng.choice([(11, 1), (10, 1), (5, 2), (14, 2), (3, 3)])
        flags = rng.choice(SG_FLAGS) if alg == O.SG else (0, 0, 0, 0)
        res, tr = O.align(alg, O.TRACE, s1, s2, o, g, mx, flags)
        c = O.cigar(s1, s2, alg == O.SG, res[1], res[2], tr)
        if c is None:
            # sg cigar is only defined when one sequence ends at its last residue
            assert alg == O.SG and res[1] + 1 != len(s1) and res[2] + 1 != len(s2)
            continue
        ops, bq, br = c
        q, t = mx.map(s1), mx.map(s2)
        lead = (flags[0], flags[2])
        if alg == O.SW and len(ops):
            # cigar_template.c:145-164: once the walk leaves the matrix it
            # emits the rest of the other sequence as one run from (0,0);
            # for a local alignment that run lies outside the alignment.
            first = "MIDNSHP=XB"[int(ops[0]) & 0xF]
            lead = (first == "I" and br == 0, first == "D" and bq == 0)
        sc, ei, ej = rescore_cigar(ops, bq, br, q, t, M, o, g,
                                   free_lead=lead,
                                   free_trail=(flags[1], flags[3]))
        assert sc == res[0], (alg, s1, s2, o, g, flags, O.cigar_string(ops))
        if alg == O.NW:
            assert (bq, br, ei, ej) == (0, 0, len(s1), len(s2))
        if alg == O.SW:
            assert (ei, ej) == (res[1] + 1, res[2] + 1)
            # SW stats are the counts along the same path
            st, _ = O.align(alg, O.STATS, s1, s2, o, g, mx)
            runs = [(int(v) >> 4, "MIDNSHP=XB"[int(v) & 0xF]) for v in ops]
            i, j, sim = bq, br, 0
            if lead[0] or lead[1]:
                cnt, op = runs.pop(0)
                i, j = (i + cnt, j) if op == "I" else (i, j + cnt)
            for cnt, op in runs:
                for _ in range(cnt):
                    if op in "=X":
                        sim += M[q[i], t[j]] > 0
                        i += 1
                        j += 1
                    elif op == "I":
                        i += 1
                    else:
                        j += 1
            assert st[:3] == res[:3]
            assert st[3] == sum(c for c, op in runs if op == "=")
            assert st[4] == sim
            assert st[5] == sum(c for c, _ in runs)


def test_pssm_conversion_equals_square():
    rng = random.Random(5)
    b = O.blosum62()
    alpha = "ARNDCQEGHILKMFPSTWYV"
    for _ in range(20):
        s1 = rand_seq(rng, alpha, rng.randint(1, 60))
        s2 = rand_seq(rng, alpha, rng.randint(1, 60))
        p = O.convert_square_to_pssm(b, s1)
        assert p.is_pssm and p.length == len(s1)
        for alg in (O.SW, O.NW, O.SG):
            r1, _ = O.align(alg, O.SCORE, s1, s2, 11, 1, b, (1, 1, 1, 1))
            r2, _ = O.align(alg, O.SCORE, None, s2, 11, 1, p, (1, 1, 1, 1))
            assert r1 == r2
            # stats still need s1 (sw_stats.c:62-64)
            r3, _ = O.align(alg, O.STATS, s1, s2, 11, 1, b, (1, 1, 1, 1))
            r4, _ = O.align(alg, O.STATS, s1, s2, 11, 1, p, (1, 1, 1, 1))
            assert r3 == r4
        assert O.align(O.SW, O.STATS, None, s2, 11, 1, p)[0] is None


def test_pssm_create_layout():
    vals = np.arange(3 * 4, dtype=np.int32).reshape(3, 4) - 5
    p = O.matrix_pssm_create("ACGT", vals.ravel(), 3)
    v = p.values()
    assert v.shape == (3, 5)
    assert (v[:, :4] == vals).all() and (v[:, 4] == vals.min()).all()
    assert (p.max, p.min) == (vals.max(), vals.min())


def test_matrix_create_layout():
    m = O.matrix_create("acgt", 2, -3)
    v = m.values()
    assert v.shape == (5, 5)
    assert (np.diag(v)[:4] == 2).all() and v[0, 1] == -3
    assert (v[4, :] == 0).all() and (v[:, 4] == 0).all()
    assert m.mapper[ord("A")] == 0 and m.mapper[ord("t")] == 3 and m.mapper[ord("N")] == 4


def test_banded_full_band_equals_nw():
    rng = random.Random(9)
    dna = O.matrix_create("ACGT", 2, -3)
    for _ in range(40):
        s1 = rand_seq(rng, "ACGT", rng.randint(1, 50))
        s2 = rand_seq(rng, "ACGT", rng.randint(1, 50))
        k = max(len(s1), len(s2)) + 2
        rb = O.nw_banded(s1, s2, 5, 2, k, dna)
        rn, _ = O.align(O.NW, O.SCORE, s1, s2, 5, 2, dna)
        assert rb[:3] == rn[:3]


def test_banded_narrow_band_is_a_lower_bound():
    """A band restricts the paths, so its score can only drop, and it reaches
    the NW score once the band holds the optimal path's diagonals."""
    rng = random.Random(10)
    dna = O.matrix_create("ACGT", 2, -3)
    for _ in range(40):
        s1 = rand_seq(rng, "ACGT", rng.randint(5, 60))
        s2 = s1[: rng.randint(1, len(s1))] + rand_seq(rng, "ACGT", rng.randint(0, 5))
        rn, _ = O.align(O.NW, O.SCORE, s1, s2, 5, 2, dna)
        prev = None
        for k in (1, 2, 4, 8, 64):
            rb = O.nw_banded(s1, s2, 5, 2, k, dna)
            assert rb[0] <= rn[0]
            if prev is not None:
                assert rb[0] >= prev
            prev = rb[0]


def test_sg_at_least_nw_and_sw_nonnegative():
    rng = random.Random(11)
    b = O.blosum62()
    for _ in range(40):
        s1 = rand_seq(rng, "ARNDCQEGHILKMFPSTWYV", rng.randint(1, 30))
        s2 = rand_seq(rng, "ARNDCQEGHILKMFPSTWYV", rng.randint(1, 30))
        nw, _ = O.align(O.NW, O.SCORE, s1, s2, 11, 1, b)
        for flags in SG_FLAGS:
            sg, _ = O.align(O.SG, O.SCORE, s1, s2, 11, 1, b, flags)
            assert sg[0] >= nw[0]
            if flags == (0, 0, 0, 0):
                assert sg == nw
        sw, _ = O.align(O.SW, O.SCORE, s1, s2, 11, 1, b)
        assert sw[0] >= 0


def test_identical_sequences_self_score():
    b = O.blosum62()
    M = b.values()
    s = "MKTAYIAKQRQISFVKSHFSRQ"
    q = b.map(s)
    self_score = int(sum(M[x, x] for x in q))
    for alg in (O.SW, O.NW):
        r, _ = O.align(alg, O.SCORE, s, s, 11, 1, b)
        assert r[0] == self_score
    st, _ = O.align(O.NW, O.STATS, s, s, 11, 1, b)
    assert st[3:] == (len(s), len(s), len(s))


def test_invalid_arguments_return_null():
    b = O.blosum62()
    assert O.align(O.SW, O.SCORE, "AC", "AC", -1, 1, b)[0] is None
    assert O.align(O.SW, O.SCORE, "AC", "AC", 1, -1, b)[0] is None
    assert O.align(O.SW, O.SCORE, "AC", "", 1, 1, b)[0] is None
    assert O.align(O.NW, O.SCORE, "", "AC", 1, 1, b)[0] is None


def test_striped_unwind_roundtrip():
    rng = np.random.default_rng(3)
    m, n, lanes = 37, 11, 16
    row = rng.integers(0, 127, size=(m, n)).astype(np.int16)
    seglen = (m + lanes - 1) // lanes
    st = np.zeros(n * seglen * lanes, dtype=np.int16)
    for j in range(n):
        for i in range(m):
            st[j * seglen * lanes + (i % seglen) * lanes + i // seglen] = row[i, j]
    out = np.zeros(m * n, dtype=np.int32)
    O.lib().orc_striped_unwind(st.ctypes.data, 16, lanes, m, n,
                               out.ctypes.data_as(O.ctypes.POINTER(O.ctypes.c_int)))
    assert (out.reshape(m, n) == row).all()
    # the striped CIGAR walk equals the row-major one on the same table
    b = O.blosum62()
    s1 = "".join(rng.choice(list("ARNDCQEGHIL"), size=m))
    s2 = "".join(rng.choice(list("ARNDCQEGHIL"), size=n))
    res, tr = O.align(O.NW, O.TRACE, s1, s2, 11, 1, b)
    st8 = np.zeros(n * seglen * lanes, dtype=np.int16)
    for j in range(n):
        for i in range(m):
            st8[j * seglen * lanes + (i % seglen) * lanes + i // seglen] = tr[i, j]
    c1 = O.cigar(s1, s2, False, res[1], res[2], tr)
    c2 = O.cigar(s1, s2, False, res[1], res[2], st8, striped=True, bits=16, lanes=lanes)
    assert O.cigar_string(c1[0]) == O.cigar_string(c2[0]) and c1[1:] == c2[1:]


# ------------------------------------------------- trace codes and stats
#
# Every DP cell's trace code is fixed by path maxima: H is the best path
# into the cell, H_dag the best one whose last move is diagonal, E (F) the
# best one whose last move consumes s2 (s1), and E opens (DIAG_E) exactly
# when the best such path whose move before the last one is not also an
# s2 move beats every path that extends an s2 run (sg_trace.c:144-150,
# sw_trace.c:145-147).  Enumerating paths restates none of the recurrence
# (no max over neighbours), so a swapped E/F, a wrong open/extend tie or a
# wrong H-source priority fails here.  Needs gap <= open: reopening a gap
# right after one is then never better than extending it, and a path of
# moves has one score.

def cell_path_maxima(alg, q, t, M, open_, gap, flags=(0, 0, 0, 0)):
    """For every 1-based DP cell: best score over paths into it, grouped by
    (last move, whether the move before it is the same gap move)."""
    m, n = len(q), len(t)
    if alg == O.SW:
        starts = [(a, b) for a in range(m + 1) for b in range(n + 1)]
    else:
        starts = [(0, 0)]
        if alg == O.SG and flags[0]:
            starts += [(a, 0) for a in range(1, m + 1)]
        if alg == O.SG and flags[2]:
            starts += [(0, b) for b in range(1, n + 1)]
    best = {}

    def note(cell, key, s):
        d = best.setdefault(cell, {})
        if key not in d or s > d[key]:
            d[key] = s

    def rec(i, j, s, prev, prev2):
        if i >= 1 and j >= 1:
            note((i, j), (prev, prev2 == prev), s)
        if i < m and j < n:
            rec(i + 1, j + 1, s + M[q[i], t[j]], "D", prev)
        if i < m:
            rec(i + 1, j, s - (gap if prev == "U" else open_), "U", prev)
        if j < n:
            rec(i, j + 1, s - (gap if prev == "L" else open_), "L", prev)

    for (a, b) in starts:
        # NW/SG: the boundary row and column are reached only by the moves
        # that start at (0, 0) or at a free start, as the DP initialises them
        rec(a, b, 0, None, None)
    return best


def expected_trace(alg, q, t, M, open_, gap, flags=(0, 0, 0, 0)):
    m, n = len(q), len(t)
    best = cell_path_maxima(alg, q, t, M, open_, gap, flags)
    codes = np.zeros((m, n), dtype=np.int64)
    ninf = -10 ** 9
    for i in range(1, m + 1):
        for j in range(1, n + 1):
            d = best.get((i, j), {})
            g = lambda k: d.get(k, ninf)  # noqa: E731
            hdag = max(g(("D", False)), g(("D", True)))
            f_opn, f_ext = g(("U", False)), g(("U", True))
            e_opn, e_ext = g(("L", False)), g(("L", True))
            F, E = max(f_opn, f_ext), max(e_opn, e_ext)
            H = max(hdag, E, F, 0 if alg == O.SW else ninf)
            code = (32 if f_opn > f_ext else 64) | (8 if e_opn > e_ext else 16)
            code |= 4 if H == hdag else (2 if H == F else 1)
            if alg == O.SW and H == 0:
                code &= 120
            codes[i - 1, j - 1] = code
    return codes


def _tiny_cases(rng, count):
    for _ in range(count):
        mx, alpha = small_matrix(rng)
        m, n = rng.randint(1, 4), rng.randint(1, 4)
        s1 = "".join(rng.choice(alpha) for _ in range(m))
        s2 = "".join(rng.choice(alpha) for _ in range(n))
        o, g = rng.choice([(1, 1), (3, 1), (5, 2), (2, 0), (2, 2), (4, 1)])
        yield mx, s1, s2, o, g


@pytest.mark.parametrize("alg", [O.SW, O.NW])
def test_trace_table_brute_force_sw_nw(alg):
    rng = random.Random(2000 + alg)
    for mx, s1, s2, o, g in _tiny_cases(rng, 80):
        q, t = mx.map(s1), mx.map(s2)
        exp = expected_trace(alg, q, t, mx.values(), o, g)
        _, tr = O.align(alg, O.TRACE, s1, s2, o, g, mx)
        assert (tr.astype(np.int64) == exp).all(), (alg, s1, s2, o, g, tr, exp)


def test_trace_table_brute_force_sg_all_flags():
    rng = random.Random(2002)
    for flags in SG_FLAGS:
        for mx, s1, s2, o, g in _tiny_cases(rng, 10):
            q, t = mx.map(s1), mx.map(s2)
            exp = expected_trace(O.SG, q, t, mx.values(), o, g, flags)
            _, tr = O.align(O.SG, O.TRACE, s1, s2, o, g, mx, flags)
            assert (tr.astype(np.int64) == exp).all(), (flags, s1, s2, o, g, tr, exp)


def counts_along_cigar(ops, bq, br, end_query, end_ref, q, t, M):
    """(matches, similar, length) over the moves of a CIGAR that enter a DP
    cell of the matrix (row and column >= 1), up to the end cell.  Moves
    along the DP's boundary row or column carry no statistics: the
    boundary's counts are 0 (sg_stats.c:146-170,189-201, nw_stats.c,
    sw_stats.c), and end gaps past the end cell are not part of the DP."""
    i, j = bq, br
    mt = sm = ln = 0
    for v in ops:
        cnt, op = int(v) >> 4, "MIDNSHP=XB"[int(v) & 0xF]
        for _ in range(cnt):
            if (i, j) == (end_query + 1, end_ref + 1):
                return mt, sm, ln
            if op in "=X":
                mt += int(q[i] == t[j])
                sm += int(M[q[i], t[j]] > 0)
                i += 1
                j += 1
            elif op == "I":
                i += 1
            else:
                j += 1
            if i >= 1 and j >= 1:
                ln += 1
    return mt, sm, ln


def _stats_vs_path(alg, s1, s2, o, g, mx, flags):
    st, _ = O.align(alg, O.STATS, s1, s2, o, g, mx, flags)
    sc, tr = O.align(alg, O.TRACE, s1, s2, o, g, mx, flags)
    # the stats functions' end rules may pick another cell than the score
    # function's (sg_stats.c:252-309 against sg.c:171-235); the path into
    # the stats end cell comes from the same trace table
    c = O.cigar(s1, s2, alg == O.SG, st[1], st[2], tr)
    assert c is not None, (alg, flags, s1, s2, st)
    ops, bq, br = c
    q, t = mx.map(s1), mx.map(s2)
    exp = counts_along_cigar(ops, bq, br, st[1], st[2], q, t, mx.values())
    assert st[3:] == exp, (alg, flags, s1, s2, o, g, st, exp, O.cigar_string(ops), bq, br)
    if alg != O.SG:
        assert st[:3] == sc[:3]


@pytest.mark.parametrize("alg", [O.SW, O.NW])
def test_stats_are_counts_along_the_trace_path(alg):
    rng = random.Random(3000 + alg)
    b = O.blosum62()
    dna = O.matrix_create("ACGT", 2, -3)
    for trial in range(150):
        mx, alpha = (b, "ARNDCQEGHILKMFPSTWYV") if trial % 2 else (dna, "ACGT")
        s1 = rand_seq(rng, alpha, rng.randint(1, 30))
        s2 = rand_seq(rng, alpha, rng.randint(1, 30))
        if trial % 3 == 0:
            s2 = s1[rng.randint(0, len(s1) - 1):] + rand_seq(rng, alpha, rng.randint(0, 4))
        o, g = rng.choice([(11, 1), (10, 1), (5, 2), (14, 2), (3, 3), (1, 1)])
        _stats_vs_path(alg, s1, s2, o, g, mx, (0, 0, 0, 0))


def test_stats_are_counts_along_the_trace_path_sg_all_flags():
    rng = random.Random(3002)
    b = O.blosum62()
    dna = O.matrix_create("ACGT", 2, -3)
    for flags in SG_FLAGS:
        for trial in range(25):
            mx, alpha = (b, "ARNDCQEGHILKMFPSTWYV") if trial % 2 else (dna, "ACGT")
            s1 = rand_seq(rng, alpha, rng.randint(1, 30))
            s2 = rand_seq(rng, alpha, rng.randint(1, 30))
            if trial % 3 == 0:
                s2 = s1[rng.randint(0, len(s1) - 1):] + rand_seq(rng, alpha, rng.randint(0, 4))
            o, g = rng.choice([(11, 1), (10, 1), (5, 2), (14, 2), (3, 3), (1, 1)])
            _stats_vs_path(O.SG, s1, s2, o, g, mx, flags)


# ------------------------------------------------------ banded brute force
#
# The band of nw_banded.c:92-97 in diagonal terms: cell (r, j) (0-based,
# r = -1 / j = -1 the DP's boundary) lies in it when
#   min(0, m - n) - k <= r - j <= max(0, m - n) + k,
# i.e. the start and end diagonals with k diagonals of slack.  Every path
# from (-1, -1) to (m-1, n-1) that stays inside the band, boundary moves
# included (the top row is set only while row -1 is in a column's band,
# nw_banded.c:134-136; the left column's initial slots, :118-124), competes;
# the banded score is the best of them.

def banded_brute(q, t, M, open_, gap, k):
    m, n = len(q), len(t)
    lo, hi = min(0, m - n) - k, max(0, m - n) + k
    best = [None]

    def inband(r, j):
        return lo <= r - j <= hi

    def rec(r, j, s, prev):
        if (r, j) == (m - 1, n - 1):
            if best[0] is None or s > best[0]:
                best[0] = s
            return
        if r + 1 < m and j + 1 < n and inband(r + 1, j + 1):
            rec(r + 1, j + 1, s + M[q[r + 1], t[j + 1]], "D")
        if r + 1 < m and inband(r + 1, j):
            rec(r + 1, j, s - (gap if prev == "U" else open_), "U")
        if j + 1 < n and inband(r, j + 1):
            rec(r, j + 1, s - (gap if prev == "L" else open_), "L")

    rec(-1, -1, 0, None)
    return best[0]


def test_banded_brute_force():
    rng = random.Random(12)
    for trial in range(400):
        mx, alpha = small_matrix(rng)
        m, n = rng.randint(1, 6), rng.randint(1, 6)
        if trial % 3 == 0:
            n = m
        s1 = "".join(rng.choice(alpha) for _ in range(m))
        s2 = "".join(rng.choice(alpha) for _ in range(n))
        o, g = rng.choice([(1, 1), (3, 1), (5, 2), (2, 0), (4, 1), (2, 2)])
        k = rng.choice([1, 1, 2, 3, 4])
        exp = banded_brute(mx.map(s1), mx.map(s2), mx.values(), o, g, k)
        got = O.nw_banded(s1, s2, o, g, k, mx)
        assert exp is not None
        assert got[0] == exp, (s1, s2, o, g, k, got, exp)
        assert got[1:3] == (m - 1, n - 1)


# ------------------------------------------------------------ DP tables
#
# The _table outputs (src/sw.c:159-170, sg_stats.c:258-263): every cell's H
# equals the best path into it (brute force), and every cell's statistics
# the counts along the trace path into that cell.

def test_score_tables_brute_force():
    rng = random.Random(4000)
    for alg in (O.SW, O.NW, O.SG):
        for trial in range(40):
            mx, alpha = small_matrix(rng)
            m, n = rng.randint(1, 4), rng.randint(1, 4)
            s1 = "".join(rng.choice(alpha) for _ in range(m))
            s2 = "".join(rng.choice(alpha) for _ in range(n))
            o, g = rng.choice([(1, 1), (3, 1), (5, 2), (2, 0)])
            flags = rng.choice(SG_FLAGS) if alg == O.SG else (0, 0, 0, 0)
            H = brute_H(alg, mx.map(s1), mx.map(s2), mx.values(), o, g, flags)
            for kind in (O.SCORE, O.STATS):
                res, tabs = O.align_tables(alg, kind, s1, s2, o, g, mx, flags)
                tab = tabs[0] if kind == O.STATS else tabs
                assert (tab == H[1:, 1:]).all(), (alg, kind, flags, s1, s2, o, g)
                assert res == O.align(alg, kind, s1, s2, o, g, mx, flags)[0]


def test_stats_tables_are_counts_along_trace_paths():
    rng = random.Random(4001)
    b = O.blosum62()
    dna = O.matrix_create("ACGT", 2, -3)
    for alg in (O.SW, O.NW, O.SG):
        for trial in range(12):
            mx, alpha = (b, "ARNDCQEGHILKMFPSTWYV") if trial % 2 else (dna, "ACGT")
            s1 = rand_seq(rng, alpha, rng.randint(1, 14))
            s2 = rand_seq(rng, alpha, rng.randint(1, 14))
            o, g = rng.choice([(11, 1), (5, 2), (3, 3), (1, 1)])
            flags = rng.choice(SG_FLAGS) if alg == O.SG else (0, 0, 0, 0)
            _, tabs = O.align_tables(alg, O.STATS, s1, s2, o, g, mx, flags)
            _, tr = O.align(alg, O.TRACE, s1, s2, o, g, mx, flags)
            q, t = mx.map(s1), mx.map(s2)
            for i in range(len(s1)):
                for j in range(len(s2)):
                    # the walk from cell (i, j) (cigar_template.c:143-251)
                    c = O.cigar(s1, s2, False, i, j, tr)
                    exp = counts_along_cigar(c[0], c[1], c[2], i, j, q, t, mx.values())
                    got = (tabs[1][i, j], tabs[2][i, j], tabs[3][i, j])
                    assert got == exp, (alg, flags, s1, s2, o, g, i, j)


def test_striped_baseline_equals_serial():
    """The CPU baseline bench.py times (the AVX2 striped 8 -> 16 -> 32-bit
    chain, oracle/striped.c) returns the serial recurrence's score and
    ends for SW, NW and SG under all 16 end-flag sets, at every width tier:
    query lengths around the 32- and 16-lane segment boundaries, the
    reference's gap sets (open > gap, tests/test_verify.c:130-190) and the
    benchmarks' 11/1 and 5/2, pairs that hold at 8 bits, need 16, and
    (a W-rich self pair past 32,767) need 32."""
    if not O.striped_available():
        pytest.skip("host without AVX2")
    rng = random.Random(77)
    b = O.blosum62()
    dna = O.matrix_create("ACGT", 2, -3)
    aa = "ARNDCQEGHILKMFPSTWYV"
    seen = set()
    for trial in range(900):
        mx, alpha = (dna, "ACGT") if trial % 5 == 0 else (b, aa)
        m = rng.choice([1, 2, 15, 16, 17, 31, 32, 33, 63, 65, 200, 700])
        n = rng.choice([1, 3, 40, 300, 900])
        s1 = rand_seq(rng, alpha, m)
        s2 = rand_seq(rng, alpha, n)
        if rng.random() < 0.3:
            s2 = s1[:n] + s2[: max(0, n - m)]
        alg = rng.choice([O.SW, O.NW, O.SG])
        flags = rng.choice(SG_FLAGS) if alg == O.SG else (0, 0, 0, 0)
        o, g = rng.choice([(10, 1), (10, 2), (14, 2), (40, 2), (11, 1), (5, 2)])
        exp, _ = O.align(alg, O.SCORE, s1, s2, o, g, mx, flags)
        got, w = O.striped_sat(alg, s1, s2, o, g, mx, flags)
        assert got == exp[:3], (alg, flags, m, n, o, g, w)
        seen.add((alg, w))
    # SW's 16-bit lanes are biased by -32768 and hold scores up to
    # 2^16 - max - 2 = 65,523 (sw_striped_avx2_256_16.c:205-210,350-362);
    # NW and SG's signed ones up to 32,767 - max - 1
    for alg, length in ((O.SW, 3100), (O.SW, 6000), (O.NW, 3100), (O.SG, 3100)):
        s = "W" * length
        exp, _ = O.align(alg, O.SCORE, s, s, 11, 1, b, (1, 1, 1, 1))
        got, w = O.striped_sat(alg, s, s, 11, 1, b, (1, 1, 1, 1))
        assert exp[0] == 11 * length
        assert (got, w) == (exp[:3], 16 if (alg, length) == (O.SW, 3100) else 32)
        seen.add((alg, w))
    for alg in (O.SW, O.NW, O.SG):
        assert {(alg, 8), (alg, 16), (alg, 32)} <= seen, (alg, seen)
