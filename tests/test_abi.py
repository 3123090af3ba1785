"""C-ABI checks that need no GPU: the library loads, exports every entry
point include/parasail_hip.h declares, resolves the reference's name grammar
through its lookup tables, and refuses (NULL, no CPU fallback) when no
device is usable."""
import ctypes
import os
import re
import subprocess

import pytest

import parasail_amd as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "parasail_hip.h")


def declared_symbols():
    pre = subprocess.run(["gcc", "-E", "-P", "-x", "c", HEADER], check=True,
                         stdout=subprocess.PIPE, text=True).stdout
    pre = re.sub(r"typedef[^;]*;", "", pre, flags=re.S)
    names = set(re.findall(r"\b(parasail_\w+)\s*\(", pre))
    return names


def exported_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", P.LIB_PATH], check=True,
                         stdout=subprocess.PIPE, text=True).stdout
    return {l.split()[-1] for l in out.splitlines() if " T " in l}


def test_every_declared_symbol_is_exported():
    decl = declared_symbols()
    assert len(decl) > 600  # 13 algs x 3 kinds x 2 strategies x 2 (profile) x 4 widths
    missing = decl - exported_symbols()
    assert not missing, sorted(missing)[:20]


def test_lookup_follows_reference_grammar():
    L = P.lib()
    for alg in P.ALGS:
        for kind in ("", "_stats", "_trace"):
            for strat in ("striped", "scan"):
                for w in ("8", "16", "32", "sat"):
                    name = "parasail_%s%s_%s_hip_64_%s" % (alg, kind, strat, w)
                    assert L.parasail_lookup_function(name.encode()), name
                    # the reference also accepts the name without its prefix
                    assert L.parasail_lookup_function(name[len("parasail_"):].encode())
                    pname = "parasail_%s%s_%s_profile_hip_64_%s" % (alg, kind, strat, w)
                    assert L.parasail_lookup_pfunction(pname.encode()), pname
                    assert L.parasail_lookup_pcreator(pname.encode()), pname
    assert not L.parasail_lookup_function(b"parasail_sw_striped_avx2_256_16")
    assert not L.parasail_lookup_function(b"nonsense")
    assert not L.parasail_lookup_function(None)


def test_matrix_objects_cpu_side():
    L = P.lib()
    b = P.blosum62()
    assert b.contents.size == 24 and b.contents.max == 11 and b.contents.min == -4
    assert b.contents.mapper[ord("a")] == 0 and b.contents.mapper[ord("J")] == 23
    assert not L.parasail_matrix_lookup(b"pam50")
    m = P.matrix_create("ACGT", 2, -3)
    mc = m.contents
    assert mc.size == 5 and mc.alphabet == b"ACGT*" and mc.max == 2 and mc.min == -3
    assert mc.matrix[0] == 2 and mc.matrix[1] == -3 and mc.matrix[4] == 0 and mc.matrix[20] == 0
    assert mc.mapper[ord("t")] == 3 and mc.mapper[ord("N")] == 4
    L.parasail_matrix_free(m)
    p = L.parasail_matrix_convert_square_to_pssm(b, b"ARN", 3)
    pc = p.contents
    assert pc.type == 1 and pc.length == 3 and pc.size == 24
    assert [pc.matrix[k] for k in (0, 24 + 1, 48 + 2)] == [4, 5, 6]
    L.parasail_matrix_free(p)


def test_cigar_codec_cpu_side():
    L = P.lib()
    L.parasail_cigar_encode_string.restype = ctypes.POINTER(P.Cigar)
    L.parasail_cigar_encode_string.argtypes = [ctypes.c_char_p]
    c = L.parasail_cigar_encode_string(b"3=2I2=1X4D14=")
    s = ctypes.string_at(L.parasail_cigar_decode(c)).decode()
    assert s == "3=2I2=1X4D14="
    assert c.contents.seq[1] == (2 << 4) | 1
    L.parasail_cigar_free(c)


@pytest.mark.skipif(P.available(), reason="a HIP device is present")
def test_no_device_returns_null():
    """No CPU fallback: without a device every entry point returns NULL."""
    r = P.align("sw_striped_hip_64_16", "ASDF", "ASDF", 10, 1, P.blosum62())
    assert r is None
    with pytest.raises(RuntimeError):
        P.align_batch("sw", 0, 16, ["ASDF"], ["ASDF"], 10, 1, P.blosum62())


def test_multi_pass_rejects_bad_arguments_without_a_device():
    """parasail_hip_batch_align_multi checks its arguments before it touches
    the device, and a pass index with no timing reports -1."""
    L = P.lib()
    L.parasail_hip_batch_align_multi.restype = ctypes.c_int
    L.parasail_hip_batch_align_multi.argtypes = [ctypes.c_void_p, ctypes.c_int,
                                                 ctypes.POINTER(ctypes.c_char_p), ctypes.c_int,
                                                 ctypes.c_int, ctypes.POINTER(P.BatchResult)]
    L.parasail_hip_last_pass_ms.restype = ctypes.c_int
    L.parasail_hip_last_pass_ms.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                            ctypes.POINTER(ctypes.c_double)]
    names = (ctypes.c_char_p * 2)(b"sw", b"nw")
    outs = (P.BatchResult * 2)()
    assert L.parasail_hip_batch_align_multi(None, 2, names, 11, 1, outs) == -1
    fake = ctypes.c_void_p(1)  # never dereferenced: the arguments fail first
    assert L.parasail_hip_batch_align_multi(fake, 5, names, 11, 1, outs) == -1
    assert L.parasail_hip_batch_align_multi(fake, 0, names, 11, 1, outs) == -1
    assert L.parasail_hip_batch_align_multi(fake, 2, names, -1, 1, outs) == -1
    bad = (ctypes.c_char_p * 2)(b"sw", b"xx")
    assert L.parasail_hip_batch_align_multi(fake, 2, bad, 11, 1, outs) == -1
    assert L.parasail_hip_last_pass_ms(-1, None, None) == -1
    assert L.parasail_hip_last_pass_ms(7, None, None) == -1


def test_sequences_from_file_and_pack(tmp_path):
    """FASTA and FASTQ records as src/parser.c:48-221 reads them through
    contrib/kseq.h:177-218: name up to the first whitespace, the rest of the
    header line as comment, wrapped and CRLF lines joined, empty lines
    skipped, FASTQ quality read to the sequence's length; length statistics
    of parasail/stats.h; packing of src/parser.c:235-259."""
    import ctypes
    from parasail_amd import lib

    class Str(ctypes.Structure):
        _fields_ = [("l", ctypes.c_size_t), ("s", ctypes.c_char_p)]

    class Seq(ctypes.Structure):
        _fields_ = [("name", Str), ("comment", Str), ("seq", Str), ("qual", Str)]

    class Seqs(ctypes.Structure):
        _fields_ = [("seqs", ctypes.POINTER(Seq)), ("l", ctypes.c_size_t),
                    ("characters", ctypes.c_size_t), ("shortest", ctypes.c_size_t),
                    ("longest", ctypes.c_size_t), ("mean", ctypes.c_float),
                    ("stddev", ctypes.c_float)]

    L = lib()
    L.parasail_sequences_from_file.restype = ctypes.POINTER(Seqs)
    L.parasail_sequences_from_file.argtypes = [ctypes.c_char_p]
    L.parasail_sequences_pack.restype = ctypes.c_void_p
    L.parasail_sequences_pack.argtypes = [ctypes.POINTER(Seqs), ctypes.POINTER(ctypes.c_size_t)]
    L.parasail_sequences_free.argtypes = [ctypes.POINTER(Seqs)]
    libc = ctypes.CDLL(None)
    libc.free.argtypes = [ctypes.c_void_p]

    def read(text):
        f = tmp_path / "in.txt"
        f.write_bytes(text)
        p = L.parasail_sequences_from_file(str(f).encode())
        assert p
        s = p.contents
        recs = [tuple(getattr(s.seqs[i], k).s or b"" for k in ("name", "comment", "seq", "qual"))
                for i in range(s.l)]
        for i in range(s.l):
            for k in ("name", "comment", "seq", "qual"):
                x = getattr(s.seqs[i], k)
                assert x.l == len(x.s or b"")
        stats = (s.characters, s.shortest, s.longest, s.mean, s.stddev)
        n = ctypes.c_size_t()
        packed = L.parasail_sequences_pack(p, ctypes.byref(n))
        pk = ctypes.string_at(packed, n.value + 1)
        libc.free(packed)
        L.parasail_sequences_free(p)
        return recs, stats, pk

    recs, stats, pk = read(b"junk line\n>sp|P1 first protein\nMKV\nLLA\n\n>P2\r\nAC\r\nD\r\n"
                           b">P3\tx y \n" + b"W" * 9 + b"\n")
    assert recs == [(b"sp|P1", b"first protein", b"MKVLLA", b""), (b"P2", b"", b"ACD", b""),
                    (b"P3", b"x y ", b"W" * 9, b"")]
    assert stats[:3] == (18, 3, 9)
    assert abs(stats[3] - 6.0) < 1e-6 and abs(stats[4] - 3.0) < 1e-6  # sample stddev of 6, 3, 9
    assert pk == b"MKVLLA$ACD$WWWWWWWWW$\0"
    # FASTQ: a quality line may start with '@'; wrapped quality; second record
    recs, stats, pk = read(b"@r1 c\nACGT\nAC\n+r1\n@@II\nII\n@r2\nGG\n+\n##\n")
    assert recs == [(b"r1", b"c", b"ACGTAC", b"@@IIII"), (b"r2", b"", b"GG", b"##")]
    assert stats[:3] == (8, 2, 6) and pk == b"ACGTAC$GG$\0"
    # a quality shorter than its sequence ends the parse after the records before it
    recs, _, _ = read(b"@a\nAC\n+\nII\n@b\nACGT\n+\nI\n")
    assert recs == [(b"a", b"", b"AC", b"II")]
    # a header character as the file's last byte starts no record (kseq's
    # ks_getuntil returns -1 when it reads nothing at the end)
    recs, stats, pk = read(b">a\nACGT\n>")
    assert recs == [(b"a", b"", b"ACGT", b"")] and stats[:3] == (4, 4, 4)
    assert abs(stats[3] - 4.0) < 1e-6 and pk == b"ACGT$\0"
    # gzip input is decompressed (the reference's HAVE_ZLIB build,
    # src/parser.c:86-114), giving the records of the plain file
    import gzip
    text = b">sp|P1 first protein\nMKV\nLLA\n>P2\nACD\n"
    plain = read(text)
    f = tmp_path / "in.fa.gz"
    f.write_bytes(gzip.compress(text))
    p = L.parasail_sequences_from_file(str(f).encode())
    assert p and p.contents.l == 2
    assert [p.contents.seqs[i].seq.s for i in range(2)] == [r[2] for r in plain[0]]
    L.parasail_sequences_free(p)
    L.parasail_sequences_free(None)
    assert not L.parasail_sequences_from_file(str(tmp_path / "missing.fa").encode())


class _FileStat(ctypes.Structure):
    _fields_ = [("sequences", ctypes.c_ulong), ("characters", ctypes.c_ulong),
                ("shortest", ctypes.c_ulong), ("longest", ctypes.c_ulong),
                ("mean", ctypes.c_float), ("stddev", ctypes.c_float)]


class _File(ctypes.Structure):
    _fields_ = [("fd", ctypes.c_int), ("size", ctypes.c_long), ("buf", ctypes.c_void_p)]


def _file_api():
    from parasail_amd import lib
    L = lib()
    L.parasail_open.restype = ctypes.POINTER(_File)
    L.parasail_open.argtypes = [ctypes.c_char_p]
    L.parasail_close.argtypes = [ctypes.POINTER(_File)]
    for nm in ("parasail_stat", "parasail_stat_fasta", "parasail_stat_fastq"):
        getattr(L, nm).restype = ctypes.POINTER(_FileStat)
        getattr(L, nm).argtypes = [ctypes.POINTER(_File)]
    for nm in ("parasail_stat_buffer", "parasail_stat_fasta_buffer", "parasail_stat_fastq_buffer"):
        getattr(L, nm).restype = ctypes.POINTER(_FileStat)
        getattr(L, nm).argtypes = [ctypes.c_char_p, ctypes.c_long]
    for nm in ("parasail_read", "parasail_pack", "parasail_pack_fasta", "parasail_pack_fastq"):
        getattr(L, nm).restype = ctypes.c_void_p
        getattr(L, nm).argtypes = [ctypes.POINTER(_File), ctypes.POINTER(ctypes.c_long)]
    L.parasail_pack_buffer.restype = ctypes.c_void_p
    L.parasail_pack_buffer.argtypes = [ctypes.c_char_p, ctypes.c_long, ctypes.POINTER(ctypes.c_long)]
    L.parasail_is_fasta.argtypes = [ctypes.POINTER(_File)]
    L.parasail_is_fastq.argtypes = [ctypes.POINTER(_File)]
    return L


def test_file_layer_stat_pack_read(tmp_path):
    """parasail_open/stat/pack/read/close and their _buffer forms
    (src/io.c:53-800): FASTA records counted by their letters with '$' after
    each; FASTQ four-line records, whose statistics sample a record's length
    when read and again when the next record starts, counting no characters
    (src/io.c:401-496)."""
    L = _file_api()
    libc = ctypes.CDLL(None)
    libc.free.argtypes = [ctypes.c_void_p]
    fa = tmp_path / "a.fa"
    text = b">one first\nMKV\nLL\r\n>two\nACDE\n\n>three\nW\n"
    fa.write_bytes(text)
    pf = L.parasail_open(str(fa).encode())
    assert pf and pf.contents.size == len(text)
    assert L.parasail_is_fasta(pf) == 1 and L.parasail_is_fastq(pf) == 0
    st = L.parasail_stat(pf).contents
    assert (st.sequences, st.characters, st.shortest, st.longest) == (3, 10, 1, 5)
    assert abs(st.mean - 10 / 3) < 1e-5 and abs(st.stddev - 2.0816659) < 1e-5
    n = ctypes.c_long()
    p = L.parasail_pack(pf, ctypes.byref(n))
    assert ctypes.string_at(p, n.value + 1) == b"MKVLL$ACDE$W$\0"
    libc.free(p)
    p = L.parasail_read(pf, ctypes.byref(n))
    assert n.value == len(text) and ctypes.string_at(p, n.value) == text
    libc.free(p)
    L.parasail_close(pf)
    p = L.parasail_pack_buffer(text, len(text), ctypes.byref(n))
    assert ctypes.string_at(p, n.value) == b"MKVLL$ACDE$W$"
    libc.free(p)
    # errors: non-alpha residues, an empty last record, no header
    assert not L.parasail_stat_fasta_buffer(b">a\nAC1D\n", 8)
    assert not L.parasail_stat_fasta_buffer(b">a\nACD\n>b\n", 10)
    assert not L.parasail_stat_fasta_buffer(b"ACD\n", 4)
    # FASTQ
    fq = b"@r1\nACGT\n+\nIIII\n@r2\nGG\n+r2\n##\n"
    st = L.parasail_stat_buffer(fq, len(fq)).contents
    assert (st.sequences, st.characters, st.shortest, st.longest) == (2, 0, 2, 4)
    assert abs(st.mean - 10 / 3) < 1e-5  # lengths 4, 4, 2 sampled
    p = L.parasail_pack_buffer(fq, len(fq), ctypes.byref(n))
    assert ctypes.string_at(p, n.value) == b"ACGT$GG$"
    libc.free(p)
    assert not L.parasail_stat_fastq_buffer(b"@r1\nACGT\nIIII\n", 14)
    assert not L.parasail_open(str(tmp_path / "missing.fa").encode())


def _parse_pssm_text(text):
    """The values of a PSSM text file, split by whitespace (a plain reading
    of data/cd02249_matrix_*.txt)."""
    lines = [ln.split() for ln in text.splitlines() if ln.strip() and not ln.startswith("#")]
    alphabet = "".join(lines[0])
    rows, query = [], ""
    for ln in lines[1:]:
        if ln[0].lstrip("-").isdigit():
            query += "*"
            rows.append([int(v) for v in ln])
        else:
            query += ln[0]
            rows.append([int(v) for v in ln[1:]])
    return alphabet, query, rows


def test_matrix_from_file(tmp_path, fixtures):
    """parasail_matrix_from_file (src/io.c:921-1178) on the reference's PSSM
    files (with and without a query letter per row) and on a square matrix
    written out as text: alphabet, query, values, extrema, type, length and
    a mapper that sends unknown letters to the last index."""
    from parasail_amd import lib, Matrix, MP
    L = lib()
    L.parasail_matrix_from_file.restype = MP
    L.parasail_matrix_from_file.argtypes = [ctypes.c_char_p]
    L.parasail_matrix_from_file_case_sensitive.restype = MP
    L.parasail_matrix_from_file_case_sensitive.argtypes = [ctypes.c_char_p]
    L.parasail_matrix_free.argtypes = [MP]
    for fn, text in fixtures["matrix_files"].items():
        path = tmp_path / fn
        path.write_text(text)
        m = L.parasail_matrix_from_file(str(path).encode())
        assert m, fn
        mc = m.contents
        alphabet, query, rows = _parse_pssm_text(text)
        assert mc.type == 1 and mc.size == len(alphabet) and mc.length == len(rows)
        assert mc.alphabet.decode() == alphabet and mc.query.decode() == query
        vals = [mc.matrix[k] for k in range(mc.size * mc.length)]
        assert vals == [v for r in rows for v in r]
        assert (mc.max, mc.min) == (max(vals), min(vals))
        assert mc.mapper[ord("A")] == alphabet.index("A") == mc.mapper[ord("a")]
        assert mc.mapper[ord("Z")] == mc.size - 1
        L.parasail_matrix_free(m)
    # a square matrix in the same text form (comment, blank line, CRLF)
    d = fixtures["matrices"]["blosum62"]
    a = d["alphabet"]
    body = "# BLOSUM62\n   " + "  ".join(a) + "\r\n\n" + "".join(
        a[r] + " " + " ".join("%3d" % d["values"][r * 24 + c] for c in range(24)) + "\n"
        for r in range(24))
    path = tmp_path / "blosum62.txt"
    path.write_text(body)
    for cs in (0, 1):
        f = L.parasail_matrix_from_file_case_sensitive if cs else L.parasail_matrix_from_file
        m = f(str(path).encode())
        assert m
        mc = m.contents
        assert mc.type == 0 and mc.size == 24 and mc.length == 24
        assert [mc.matrix[k] for k in range(576)] == d["values"]
        assert (mc.max, mc.min) == (d["max"], d["min"])
        assert mc.mapper[ord("w")] == (23 if cs else a.index("W"))
        L.parasail_matrix_free(m)
    # a row short of values, and a row letter outside the alphabet line's form
    path.write_text("A C\nA 1 2\nC 3\n")
    assert not L.parasail_matrix_from_file(str(path).encode())
