"""Seeded synthetic workloads of the benchmark configurations.

Protein residues are drawn iid over the 20 standard amino acids with the
Robinson & Robinson background frequencies (Robinson A.B., Robinson L.R.,
"Distribution of glutamine and asparagine residues and their near
neighbors in peptides and proteins", PNAS 88:8880-8884, 1991).
"""
import numpy as np

AA = "ARNDCQEGHILKMFPSTWYV"
RR_PER_1000 = np.array([78.05, 51.29, 44.87, 53.64, 19.25, 42.64, 62.95, 73.77, 21.99, 51.42,
                        90.19, 57.44, 22.43, 38.56, 52.03, 71.20, 58.41, 13.30, 32.16, 64.41])
_AA_BYTES = np.frombuffer(AA.encode(), dtype=np.uint8)
_DNA_BYTES = np.frombuffer(b"ACGT", dtype=np.uint8)


def lognormal_lengths(rng, n, median, sigma, lo, hi):
    x = np.rint(median * np.exp(sigma * rng.standard_normal(n)))
    return np.clip(x, lo, hi).astype(np.int64)


def protein_bytes(rng, total):
    p = RR_PER_1000 / RR_PER_1000.sum()
    return _AA_BYTES[rng.choice(20, size=int(total), p=p)].tobytes()


def concat_offsets(lengths):
    off = np.zeros(len(lengths), dtype=np.int64)
    if len(lengths):
        off[1:] = np.cumsum(lengths)[:-1]
    return off


class PairBatch:
    """Concatenated query and target residues with offsets and lengths."""

    def __init__(self, qbytes, qlen, tbytes, tlen):
        self.q = qbytes
        self.qlen = np.asarray(qlen, dtype=np.int32)
        self.qoff = concat_offsets(self.qlen)
        self.t = tbytes
        self.tlen = np.asarray(tlen, dtype=np.int32)
        self.toff = concat_offsets(self.tlen)

    def __len__(self):
        return len(self.qlen)

    @property
    def cells(self):
        return int(np.dot(self.qlen.astype(np.int64), self.tlen.astype(np.int64)))

    def query(self, k):
        o = int(self.qoff[k])
        return self.q[o:o + int(self.qlen[k])]

    def target(self, k):
        o = int(self.toff[k])
        return self.t[o:o + int(self.tlen[k])]

    def subset(self, idx):
        return PairBatch(b"".join(self.query(k) for k in idx), self.qlen[idx],
                         b"".join(self.target(k) for k in idx), self.tlen[idx])


def config0(seed=0, ntargets=2000):
    """One 300-residue query against targets of length U[50, 1000]."""
    rng = np.random.default_rng(seed)
    q = protein_bytes(rng, 300)
    tl = rng.integers(50, 1001, size=ntargets)
    t = protein_bytes(rng, int(tl.sum()))
    return q, PairBatch(q * ntargets, np.full(ntargets, 300), t, tl)


def config1(seed=1, npairs=200000):
    """Independent protein pairs, lengths log-normal (median 300, sigma 0.6,
    clipped to [20, 5000])."""
    rng = np.random.default_rng(seed)
    ql = lognormal_lengths(rng, npairs, 300, 0.6, 20, 5000)
    tl = lognormal_lengths(rng, npairs, 300, 0.6, 20, 5000)
    return PairBatch(protein_bytes(rng, ql.sum()), ql, protein_bytes(rng, tl.sum()), tl)


def _geom(rng, mean):
    return int(rng.geometric(1.0 / mean))


def mutate_dna(rng, s, sub, ins, dele, mean_indel):
    """Substitutions at rate sub, insertions and deletions at the given
    rates with geometric lengths of the given mean."""
    out = bytearray()
    i, n = 0, len(s)
    while i < n:
        r = rng.random()
        if r < dele:
            i += _geom(rng, mean_indel)
            continue
        if r < dele + ins:
            out += _DNA_BYTES[rng.integers(0, 4, size=_geom(rng, mean_indel))].tobytes()
        c = s[i]
        if rng.random() < sub:
            c = int(_DNA_BYTES[(int(np.nonzero(_DNA_BYTES == c)[0][0]) + rng.integers(1, 4)) % 4])
        out.append(c)
        i += 1
    return bytes(out) if out else b"A"


def config2(seed=2, npairs=100000):
    """250 bp reads (4% substitutions, 0.5% insertions, 0.5% deletions,
    geometric indel lengths of mean 2) against their 400 bp windows."""
    rng = np.random.default_rng(seed)
    win = _DNA_BYTES[rng.integers(0, 4, size=(npairs, 400))]
    reads, rl = [], []
    for k in range(npairs):
        off = int(rng.integers(0, 151))
        r = mutate_dna(rng, win[k, off:off + 250].tobytes(), 0.04, 0.005, 0.005, 2)
        reads.append(r)
        rl.append(len(r))
    return PairBatch(b"".join(reads), rl, win.tobytes(), np.full(npairs, 400))


def protein_bytes_fast(rng, total):
    """protein_bytes for database-sized draws: residues from a 65,536-entry
    table holding each amino acid in proportion to its background frequency
    (frequencies rounded to 1/65,536)."""
    p = RR_PER_1000 / RR_PER_1000.sum()
    counts = np.floor(p * 65536).astype(np.int64)
    counts[np.argmax(counts)] += 65536 - counts.sum()
    table = np.repeat(_AA_BYTES, counts)
    return table[rng.integers(0, 65536, size=int(total), dtype=np.uint16)].tobytes()


def config4_db(seed=4, ntargets=1000000):
    """The serving database: proteins with log-normal lengths (median 350,
    sigma 0.7, clipped to [30, 35000])."""
    rng = np.random.default_rng(seed)
    tl = lognormal_lengths(rng, ntargets, 350, 0.7, 30, 35000)
    return protein_bytes_fast(rng, int(tl.sum())), concat_offsets(tl), tl.astype(np.int32)


def config4_queries(seed=5, nreq=64, qps=10.0, pssm_frac=0.1):
    """Serving requests: (arrival offset s, query bytes, pssm values or None)
    with Poisson arrivals at qps, log-normal lengths (median 300, sigma 0.6,
    clipped to [20, 5000]); a pssm_frac share carries a PSSM of per-position
    BLOSUM62 rows with uniform integer noise in [-2, 2], given as the int
    values of the 23-letter alphabet ARNDCQEGHILKMFPSTWYVBZX per position."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / qps, size=nreq)
    arr = np.cumsum(gaps) - gaps[0]
    ql = lognormal_lengths(rng, nreq, 300, 0.6, 20, 5000)
    out = []
    for k in range(nreq):
        q = protein_bytes(rng, int(ql[k]))
        pssm = None
        if rng.random() < pssm_frac:
            pssm = rng.integers(-2, 3, size=(int(ql[k]), 23)).astype(np.int32)  # noise
        out.append((float(arr[k]), q, pssm))
    return out


def mutate_fast(rng, s, sub, indel_rate, mean_len):
    """s (uint8 ACGT array) with substitutions at rate sub (to a different
    base) and indels at rate indel_rate per base, insertion or deletion with
    equal odds, geometric lengths of the given mean."""
    s = s.copy()
    subs = rng.random(len(s)) < sub
    codes = np.searchsorted(_DNA_BYTES, s[subs])
    s[subs] = _DNA_BYTES[(codes + rng.integers(1, 4, size=int(subs.sum()))) % 4]
    ev = np.nonzero(rng.random(len(s)) < indel_rate)[0]
    lens = rng.geometric(1.0 / mean_len, size=len(ev))
    ins = rng.random(len(ev)) < 0.5
    out, prev = [], 0
    for p, ln, isins in zip(ev.tolist(), lens.tolist(), ins.tolist()):
        if p < prev:
            continue
        out.append(s[prev:p])
        if isins:
            out.append(_DNA_BYTES[rng.integers(0, 4, size=ln)])
            prev = p
        else:
            prev = p + ln
    out.append(s[prev:])
    r = np.concatenate(out) if out else s[:0]
    return r if len(r) else _DNA_BYTES[:1]


def config3(seed=3, npairs=256):
    """Long DNA pairs: lengths uniform in [20000, 200000], the second a copy
    of the first with 10% substitutions and 2% indels (geometric lengths,
    mean 3)."""
    rng = np.random.default_rng(seed)
    ql = rng.integers(20000, 200001, size=npairs)
    qs, ts = [], []
    for k in range(npairs):
        q = _DNA_BYTES[rng.integers(0, 4, size=int(ql[k]))]
        qs.append(q)
        ts.append(mutate_fast(rng, q, 0.10, 0.02, 3))
    tl = np.array([len(t) for t in ts])
    return PairBatch(np.concatenate(qs).tobytes(), ql, np.concatenate(ts).tobytes(), tl)


def banded_cells(m, n, k):
    """Cells parasail_nw_banded computes for an m x n pair with band k: per
    column, the band's rows that fall inside the query."""
    low = (n - m - k) if m > n else (m - n - k)
    col_len = k - low + 1
    off0 = -k if m > n else low
    j = np.arange(n, dtype=np.int64)
    lo = np.maximum(0, off0 + j)
    hi = np.minimum(m, off0 + j + col_len)
    return int(np.maximum(0, hi - lo).sum())


def dp32_trace_bytes(m, n):
    """HBM bytes of one pair's trace in the wavefront kernel's layout: T
    lanes (16, 32 or 64, the fewest whose band of <= 20 rows per lane covers
    the query), R rows per lane (a multiple of 4) taking R rounded up to 8
    4-bit fields, (n + T + 1) steps."""
    for T in (16, 32, 64):
        if m <= T * 20:
            R = max(4, ((m + T - 1) // T + 3) & ~3)
            return ((n + T + 1) * T * ((R + 7) & ~7) + 1) // 2
    return m * n
