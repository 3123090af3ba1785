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
