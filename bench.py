"""Benchmark of the hip family on one MI355X.

A step is one pass of the hot path over the batched workload: 200,000
independent protein pairs (lengths log-normal, median 300, sigma 0.6,
clipped to [20, 5000], Robinson-Robinson residues, BLOSUM62, open 11
extend 1), aligned score-only with SW, NW and SG (all four ends free), each
in the int16-first tiering of parasail_*_striped_profile_sat, inputs resident
in HBM.  Prints one JSON line (GCUPS over the three algorithms), with the
plain one-launch-per-call kernel (1 query of 300 vs 2,000 targets of
U[50,1000], parasail_sw_striped_profile_16 semantics) beside it, the
packed-int16 VALU roofline of the dominant kernel and the CPU oracle timed
on the host cores over a bounded sample.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--pairs N]
"""
import argparse
import ctypes
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import parasail_amd as P  # noqa: E402
from parasail_amd import workloads as WL  # noqa: E402

# VALU ops per cell, fixed from the recurrence: SW 10, NW/SG 9
OPS = {"sw": 10, "nw": 9, "sg": 9}
ALGS = ("sw", "nw", "sg")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_resident(batch, matrix):
    L = P.lib()
    L.parasail_hip_batch_create.restype = ctypes.c_void_p
    L.parasail_hip_batch_create.argtypes = [
        ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int),
        ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int),
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
    L.parasail_hip_batch_set_graph.restype = ctypes.c_int
    L.parasail_hip_batch_set_graph.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.parasail_hip_batch_graph_stats.restype = ctypes.c_int
    L.parasail_hip_batch_graph_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int),
                                                 ctypes.POINTER(ctypes.c_int)]
    I64, I32 = ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int)
    h = L.parasail_hip_batch_create(batch.q, batch.qoff.ctypes.data_as(I64),
                                    batch.qlen.ctypes.data_as(I32), batch.t,
                                    batch.toff.ctypes.data_as(I64),
                                    batch.tlen.ctypes.data_as(I32), len(batch), matrix)
    if not h:
        raise RuntimeError("parasail_hip_batch_create failed: %s"
                           % L.parasail_hip_last_error().decode(errors="replace"))
    return h


def run_resident(h, alg, n, keep=None):
    L = P.lib()
    br = P.BatchResult()
    out = keep if keep is not None else {k: np.zeros(n, dtype=np.int32)
                                         for k in ("score", "end_query", "end_ref", "flag")}
    for k, v in out.items():
        setattr(br, k, v.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    if L.parasail_hip_batch_align(h, alg.encode(), 11, 1, ctypes.byref(br)) != 0:
        raise RuntimeError("parasail_hip_batch_align(%s) failed" % alg)
    return out, P.last_kernel_ms(), P.last_main_kernel_ms()


class MultiPass:
    """One parasail_hip_batch_align_multi call per step: every algorithm's
    pass over the resident batch, each pass's results into its own arrays."""

    def __init__(self, h, algs, bufs):
        self.h, self.algs = h, algs
        self.outs = (P.BatchResult * len(algs))()
        for k, alg in enumerate(algs):
            for key, v in bufs[alg].items():
                setattr(self.outs[k], key, v.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        self.names = (ctypes.c_char_p * len(algs))(*[x.encode() for x in algs])

    def run(self):
        L = P.lib()
        if L.parasail_hip_batch_align_multi(self.h, len(self.algs), self.names, 11, 1,
                                            self.outs) != 0:
            raise RuntimeError("parasail_hip_batch_align_multi failed")
        ms = []
        for k in range(len(self.algs)):
            km, mm = ctypes.c_double(), ctypes.c_double()
            L.parasail_hip_last_pass_ms(k, ctypes.byref(km), ctypes.byref(mm))
            ms.append((km.value, mm.value))
        return ms


def per_call_gcups(reps=1, threads=1):
    """Config 0: parasail_sw_striped_profile_hip_64_16 per call, from one
    host thread (one launch per call) or from `threads` threads over slices
    of the targets, as tests/test_openmp.c:353-369 runs it under OpenMP
    (calls made together share a launch)."""
    L = P.lib()
    q, b = WL.config0()
    m = P.blosum62()
    create = ctypes.CFUNCTYPE(ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, P.MP)(
        L.parasail_lookup_pcreator(b"parasail_sw_striped_profile_hip_64_16"))
    pf = ctypes.CFUNCTYPE(P.RP, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int,
                          ctypes.c_int)(
        L.parasail_lookup_pfunction(b"parasail_sw_striped_profile_hip_64_16"))
    L.parasail_profile_free.argtypes = [ctypes.c_void_p]
    prof = create(q, len(q), m)
    targets = [b.target(k) for k in range(len(b))]
    for t in targets[:20]:
        L.parasail_result_free(pf(prof, t, len(t), 11, 1))
    def run(part):
        for t in part:
            L.parasail_result_free(pf(prof, t, len(t), 11, 1))

    best = None
    pool = ThreadPoolExecutor(threads) if threads > 1 else None
    for _ in range(reps):
        t0 = time.perf_counter()
        if pool is None:
            run(targets)
        else:
            list(pool.map(run, [targets[k::threads] for k in range(threads)]))
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    if pool is not None:
        pool.shutdown()
    L.parasail_profile_free(prof)
    return b.cells / best / 1e9, b.cells, best


def per_call_driver(threads=(1, 16)):
    """Config 0 through parasail_sw_striped_profile_hip_64_16 from C threads,
    as parasail's OpenMP driver calls it (tests/test_openmp.c:353-369):
    the query and the 2,000 targets written as FASTA, timed by the built
    parasail_amd/percall_omp.  Returns {threads: (GCUPS, seconds)}."""
    import subprocess
    import tempfile
    exe = os.path.join(ROOT, "parasail_amd", "percall_omp")
    if not os.path.exists(exe):
        raise RuntimeError("per-call driver not built: run __graft_entry__.build()")
    q, b = WL.config0()
    with tempfile.TemporaryDirectory() as d:
        fq, fd = os.path.join(d, "q.fa"), os.path.join(d, "db.fa")
        with open(fq, "wb") as f:
            f.write(b">q\n" + q + b"\n")
        with open(fd, "wb") as f:
            for k in range(len(b)):
                f.write(b">t%d\n" % k + b.target(k) + b"\n")
        out = subprocess.run([exe, "sw_striped_profile_hip_64_16", fq, fd, "11", "1"] +
                             [str(t) for t in threads], check=True, capture_output=True,
                             text=True, timeout=300).stdout
    res = {}
    for line in out.split("\n"):
        f = line.split()
        if len(f) == 4:
            res[int(f[0])] = (float(f[2]), float(f[1]))
    return res


def per_call_cpu(threads):
    """Config 0 on the host: the 300-residue query against the 2,000 targets
    through the restated AVX2 striped SW at 16 bits (the
    parasail_sw_striped_profile_16 semantics; the query profile is built per
    pair), OpenMP over the targets on `threads` cores.  Returns (GCUPS,
    seconds, equal to the serial oracle) or None without AVX2."""
    import oracle as O
    om = O.blosum62()
    if not O.striped_available():
        return None
    q, b = WL.config0()
    qs = [om.map(q)] * len(b)
    ts = [om.map(b.target(k)) for k in range(len(b))]
    cells = sum(len(x) for x in ts) * len(q)
    best, res = None, None
    for _ in range(3):
        res, _w, secs = O.striped_batch(O.SW, qs, ts, 11, 1, om, threads=threads, start=16)
        best = secs if best is None else min(best, secs)
    ref, _ = O.batch(O.SW, O.SCORE, qs[:200], ts[:200], 11, 1, om, threads=threads)
    return cells / best / 1e9, best, bool((ref[:, :3] == res[:200]).all())


def cpu_baseline(batch, gpu=None, target_s=15.0):
    """The reference's dispatched CPU path restated: AVX2 Farrar striped
    query profile, one profile per pair, 8-bit lanes first, then 16-bit,
    then the serial recurrence (the `_striped_sat` chain), OpenMP over pairs
    (tests/test_openmp.c:353-369), on a bounded sample of the batch.  gpu:
    the GPU's results of the batch, {alg: {"score", "end_query", "end_ref"}},
    compared with the CPU's on the sample."""
    import oracle as O
    om = O.blosum62()
    # the host's share of cores per GPU is 16, whatever the OS reports
    threads = max(1, min(os.cpu_count() or 1, 16))
    rng = np.random.default_rng(9)
    algs = {"sw": (O.SW, (0, 0, 0, 0)), "nw": (O.NW, (0, 0, 0, 0)), "sg": (O.SG, (1, 1, 1, 1))}
    if not O.striped_available():
        return None

    def run(ids):
        qs = [om.map(batch.query(k)) for k in ids]
        ts = [om.map(batch.target(k)) for k in ids]
        cells = sum(len(a) * len(b) for a, b in zip(qs, ts))
        secs, same, widths = 0.0, True, {}
        for name, (alg, flags) in algs.items():
            res, w, s = O.striped_batch(alg, qs, ts, 11, 1, om, flags, threads=threads)
            if s < 0:
                raise RuntimeError("striped baseline failed")
            secs += s
            widths[name] = {str(b): int(np.count_nonzero(w == b)) for b in (8, 16, 32)}
            if gpu is not None and name in gpu:
                g = np.stack([gpu[name][k][ids] for k in ("score", "end_query", "end_ref")], 1)
                same &= bool((g == res).all())
        return 3 * cells, secs, same, widths

    idx = rng.choice(len(batch), size=min(len(batch), 400), replace=False)
    _, secs, _, _ = run(idx)
    scale = max(1, int(target_s / max(secs, 1e-3)))
    n2 = min(len(batch), len(idx) * scale)
    idx = rng.choice(len(batch), size=n2, replace=False)
    cells, secs, same, widths = run(idx)
    return {"value": round(cells / secs / 1e9, 3), "unit": "GCUPS", "cores": threads,
            "kind": "port",
            "sample": "%d random pairs of the batch x (SW, NW, SG), AVX2 striped query profile "
                      "per pair, int8 then int16 then int32 (the reference's dispatched "
                      "_striped_sat chain restated), OpenMP over pairs, %.1f s" % (n2, secs),
            "widths_held": widths,
            "equals_gpu_on_sample": same if gpu is not None else None}


def serve_load(db, pm, nres, nreq, qps):
    """nreq Poisson requests at qps through one serving loop: latency of
    each from its arrival to its hits in host memory."""
    reqs = WL.config4_queries(nreq=nreq, qps=qps)
    mats = []
    for _, q, noise in reqs:
        mats.append(P.blosum62_pssm(q, noise)[0] if noise is not None else pm)
    srv = P.Server(db)
    lat = {}
    start = P.now() + 0.05
    i = 0
    while len(lat) < len(reqs):
        now = P.now()
        while i < len(reqs) and start + reqs[i][0] <= now:
            _, q, noise = reqs[i]
            srv.submit(i, None if noise is not None else q, mats[i], 11, 1, 50, start + reqs[i][0])
            i += 1
        r = srv.poll(50, wait=i >= len(reqs))
        if r is not None:
            lat[r[0]] = r[2]
        elif i < len(reqs):
            time.sleep(max(0.0, min(0.002, start + reqs[i][0] - P.now())))
    end = P.now()
    _, deferred, max_active = srv.stats()
    srv.close()
    for k, (_, _, noise) in enumerate(reqs):
        if noise is not None:
            P.lib().parasail_matrix_free(mats[k])
    cells = sum(len(q) for _, q, _ in reqs) * nres
    lt = np.array([lat[k] for k in range(len(reqs))])
    return {"qps": qps, "requests": len(reqs),
            "pssm_requests": sum(1 for r in reqs if r[2] is not None),
            "p50_ms": round(float(np.percentile(lt, 50)), 1),
            "p99_ms": round(float(np.percentile(lt, 99)), 1),
            "mean_ms": round(float(lt.mean()), 1),
            "gcups": round(cells / (end - start) / 1e9, 1),
            "offered_gcups": round(cells / (reqs[-1][0] if reqs[-1][0] > 0 else 1.0) / 1e9, 1),
            "max_in_flight": max_active, "admission_waits": deferred}


def serving(nreq, qps, sweep):
    """Config 4: a database of 1,000,000 proteins resident in HBM; nreq
    Poisson requests at qps (10% PSSM), each SW of its query against the
    whole database returning the top 50; latency from arrival to hits in
    host memory.  The same at each rate of sweep (latency against load),
    and one request of 300 residues searched alone."""
    pm = P.blosum62()
    t0 = time.perf_counter()
    res, off, tl = WL.config4_db()
    tgen = time.perf_counter() - t0
    t0 = time.perf_counter()
    db = P.Database(res, off, tl, pm)
    tup = time.perf_counter() - t0
    del res
    nres = int(tl.astype(np.int64).sum())
    log("serving: database of %d targets, %.3e residues (generated %.1f s, resident in %.1f s)"
        % (len(tl), nres, tgen, tup))
    q300 = WL.protein_bytes(np.random.default_rng(77), 300)
    db.search(q300[:60], pm, 11, 1, 50)  # warm-up
    t0 = time.perf_counter()
    db.search(q300, pm, 11, 1, 50)
    alone_s = time.perf_counter() - t0
    main = serve_load(db, pm, nres, nreq, qps)
    log("serving: %.0f QPS p50 %.1f ms p99 %.1f ms, %.0f GCUPS"
        % (qps, main["p50_ms"], main["p99_ms"], main["gcups"]))
    points = []
    for x in sweep:
        pt = main if x == qps else serve_load(db, pm, nres, nreq, x)
        log("serving: %.0f QPS p50 %.1f ms p99 %.1f ms, %.0f GCUPS, %d in flight at most"
            % (x, pt["p50_ms"], pt["p99_ms"], pt["gcups"], pt["max_in_flight"]))
        points.append(pt)
    db.close()
    out = {"workload": "1,000,000-protein database (log-normal 350, 0.7, [30, 35000]) in HBM; "
                       "Poisson queries (log-normal 300, 0.6, [20, 5000]), 10% PSSM; SW 11/1, "
                       "top 50",
           "db_residues": nres}
    out.update(main)
    out["search_alone_300aa"] = {"ms": round(alone_s * 1e3, 1),
                                 "gcups": round(300 * nres / alone_s / 1e9, 1)}
    out["qps_sweep"] = points
    return out


def stats_trace(peak32, npairs=100000):
    """DNA reads against their windows: 100,000 pairs of a 250 bp read (4%
    substitutions, 0.5% insertions and 0.5% deletions of geometric length,
    mean 2) and its 400 bp window, match 2 mismatch -3, open 5 extend 2;
    semi-global (all four ends free) with statistics, then with traceback
    and CIGAR strings, in parasail_sg_stats_striped_16 and
    parasail_sg_trace_striped_16 semantics (width 16)."""
    dm = P.matrix_create("ACGT", 2, -3)
    t0 = time.perf_counter()
    b = WL.config2(npairs=npairs)
    log("stats/trace: %d pairs, %.3e cells (generated in %.1f s)"
        % (len(b), b.cells, time.perf_counter() - t0))
    # one untimed call of each at full size first: kernels loaded, and the
    # library's pinned staging buffer grown to the batch (kept for the process)
    P.align_arrays("sg", 1, 16, b, 5, 2, dm)
    P.align_arrays("sg", 2, 16, b, 5, 2, dm, want_cigar=True)
    rs = P.align_arrays("sg", 1, 16, b, 5, 2, dm)
    wall_s = P.last_call_ms() / 1e3
    kms_s = P.last_kernel_ms()
    log("stats/trace: stats %.1f ms kernel, %.1f ms call" % (kms_s, wall_s * 1e3))
    t0 = time.perf_counter()
    rt = P.align_arrays("sg", 2, 16, b, 5, 2, dm, want_cigar=True)
    wall_py = time.perf_counter() - t0
    wall_t = P.last_call_ms() / 1e3
    kms_t = P.last_kernel_ms()
    log("stats/trace: trace+cigar %.1f ms kernel, %.1f ms call, %.1f ms with Python strings"
        % (kms_t, wall_t * 1e3, wall_py * 1e3))
    P.lib().parasail_matrix_free(dm)
    ncig = sum(1 for c in rt["cigar"] if c)
    # int32 VALU instructions per cell of the wavefront kernel: 8 for the
    # scores, 4 compares, and for _stats 4 selects + 3 adds of the packed
    # statistics word, for _trace 5 to build the direction code + 1 to pack
    ops_s, ops_t = 19, 18
    # the fixed per-cell counts of the recurrence (9 for the scores, +24 for
    # the three statistics, +6 for the trace), comparable across kernels
    fix_s, fix_t = 33, 15
    trace_bytes = sum(WL.dp32_trace_bytes(int(b.qlen[i]), int(b.tlen[i])) for i in range(len(b)))
    return {"workload": "100,000 DNA pairs: 250 bp reads (4% sub, 0.5% ins, 0.5% del, geom "
                        "mean 2) vs 400 bp windows, +2/-3, open 5 extend 2, sg all ends free",
            "pairs": len(b), "cells": b.cells,
            "sg_stats": {"gcups_kernel": round(b.cells / (kms_s / 1e3) / 1e9, 1),
                         "roofline": {"bound": "valu", "unit": "TOP/s", "ops_per_cell": ops_s,
                                      "achieved": round(b.cells * ops_s / (kms_s / 1e3) / 1e12, 2),
                                      "peak": round(peak32 / 1e12, 2),
                                      "frac": round(b.cells * ops_s / (kms_s / 1e3) / peak32, 4),
                                      "ops_per_cell_fixed": fix_s,
                                      "frac_fixed": round(b.cells * fix_s / (kms_s / 1e3) / peak32, 4),
                                      "frac_fixed_note": "the kernel keeps matches, similar and "
                                                         "length in one packed word (3 selects "
                                                         "and 1 add per cell where the fixed "
                                                         "count prices 3 recurrences), so it "
                                                         "can exceed 1"},
                         "gcups_call": round(b.cells / wall_s / 1e9, 1),
                         "kernel_ms": round(kms_s, 1), "call_ms": round(wall_s * 1e3, 1),
                         "checksum": int(rs["score"].astype(np.int64).sum()
                                         + rs["matches"].astype(np.int64).sum()
                                         + rs["length"].astype(np.int64).sum())},
            "sg_trace_cigar": {"gcups_kernel": round(b.cells / (kms_t / 1e3) / 1e9, 1),
                               "roofline": {"bound": "valu", "unit": "TOP/s", "ops_per_cell": ops_t,
                                            "achieved": round(b.cells * ops_t / (kms_t / 1e3) / 1e12, 2),
                                            "peak": round(peak32 / 1e12, 2),
                                            "frac": round(b.cells * ops_t / (kms_t / 1e3) / peak32, 4),
                                            "ops_per_cell_fixed": fix_t,
                                            "frac_fixed": round(b.cells * fix_t / (kms_t / 1e3) / peak32, 4),
                                            "kernel_ms_includes": "trace DP and the CIGAR walk"},
                               "trace_bytes_hbm": trace_bytes,
                               "trace_write_gbs": round(trace_bytes / (kms_t / 1e3) / 1e9, 1),
                               "gcups_call": round(b.cells / wall_t / 1e9, 1),
                               "kernel_ms": round(kms_t, 1), "call_ms": round(wall_t * 1e3, 1),
                               "call_ms_includes": "upload, planning, kernels, 100,000 "
                                                   "parasail_cigar_t objects",
                               "with_python_strings_ms": round(wall_py * 1e3, 1),
                               "cigars": ncig}}


def long_pairs(peak32):
    """Long DNA pairs: 256 pairs of U[20000, 200000] bp, the second a mutated
    copy of the first (10% substitutions, 2% indels of geometric length,
    mean 3), match 2 mismatch -3, open 5 extend 2.  Global score in int32
    (parasail_nw_scan_32 semantics; the multi-wave long-pair kernel) and
    banded global (parasail_nw_banded, k = 256) of the same pairs, inputs
    uploaded before the timed kernels."""
    dm = P.matrix_create("ACGT", 2, -3)
    t0 = time.perf_counter()
    b = WL.config3()
    log("long pairs: %d pairs, %.3e cells (generated in %.1f s)"
        % (len(b), b.cells, time.perf_counter() - t0))
    small = b.subset(np.argsort(b.qlen)[:2])
    P.align_arrays("nw", 0, 32, small, 5, 2, dm)
    P.nw_banded_arrays(small, 5, 2, 256, dm)
    t0 = time.perf_counter()
    r = P.align_arrays("nw", 0, 32, b, 5, 2, dm)
    wall = time.perf_counter() - t0
    kms = P.last_kernel_ms()
    log("long pairs: NW int32 %.1f ms kernel, %.1f ms call" % (kms, wall * 1e3))
    kb = int(np.count_nonzero(r["flag"] & P.FLAG_BITS_32))
    t0 = time.perf_counter()
    rb = P.nw_banded_arrays(b, 5, 2, 256, dm)
    wall_b = time.perf_counter() - t0
    kms_b = P.last_kernel_ms()
    seg_pairs, seg_count, seg_redone = P.nw_banded_last_stats()
    # the same pairs without segments: one pair per workgroup (row bands)
    prev = P.nw_banded_set_segment(0)
    rb0 = P.nw_banded_arrays(b, 5, 2, 256, dm)
    kms_b0 = P.last_kernel_ms()
    P.nw_banded_set_segment(prev)
    assert (rb0["score"] == rb["score"]).all(), "banded: segmented and whole-pair scores differ"
    bcells = sum(WL.banded_cells(int(b.qlen[i]), int(b.tlen[i]), 256) for i in range(len(b)))
    log("long pairs: banded k=256 %.1f ms kernel, %.1f ms call (%d segments; %.1f ms without)"
        % (kms_b, wall_b * 1e3, seg_count, kms_b0))
    P.lib().parasail_matrix_free(dm)
    ach = b.cells * 7 / (kms / 1e3)
    return {"workload": "256 DNA pairs U[20000, 200000] bp, mutated copies (10% sub, 2% indel "
                        "geom mean 3), +2/-3, open 5 extend 2",
            "pairs": len(b), "cells": b.cells,
            "nw_int32": {"gcups": round(b.cells / (kms / 1e3) / 1e9, 1),
                         "kernel_ms": round(kms, 1), "call_ms": round(wall * 1e3, 1),
                         "pairs_int32": kb, "kernel": "k_nwlong<16> (16 bands per workgroup)",
                         "roofline": {"bound": "valu", "achieved": round(ach / 1e12, 2),
                                      "peak": round(peak32 / 1e12, 2), "unit": "TOP/s",
                                      "frac": round(ach / peak32, 4),
                                      "ops_per_cell": 7, "ops_per_cell_fixed": 9,
                                      "frac_fixed": round(b.cells * 9 / (kms / 1e3) / peak32, 4)}},
            "nw_banded_k256": {"gcups": round(bcells / (kms_b / 1e3) / 1e9, 1), "cells": bcells,
                               "kernel_ms": round(kms_b, 1), "call_ms": round(wall_b * 1e3, 1),
                               "kernel": "k_banded_seg (one-wave lane chains over segments of the "
                                         "band, joined exactly by rank convergence) + k_seg_join",
                               "segments": {"pairs": seg_pairs, "segments": seg_count,
                                            "pairs_redone_whole": seg_redone,
                                            "without_segments_kernel_ms": round(kms_b0, 1)},
                               "roofline": {"bound": "valu",
                                            "achieved": round(bcells * 9 / (kms_b / 1e3) / 1e12, 2),
                                            "peak": round(peak32 / 1e12, 2), "unit": "TOP/s",
                                            "frac": round(bcells * 9 / (kms_b / 1e3) / peak32, 4),
                                            "ops_per_cell": 9, "ops_per_cell_fixed": 9,
                                            "frac_fixed": round(bcells * 9 / (kms_b / 1e3) / peak32, 4)},
                               "score_checksum": int(rb["score"].astype(np.int64).sum())},
            "nw_score_checksum": int(r["score"].astype(np.int64).sum())}


def copy_calibration(mib):
    """One plain device-to-device copy of a known size: under rocprofv3
    --pmc FETCH_SIZE / WRITE_SIZE its counters give the scale from counter
    KiB to HBM bytes for the kernels of this library."""
    import torch
    n = mib << 20
    src = torch.ones(n // 4, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n}), flush=True)


def traffic_from_profile():
    """HBM bytes per launch of the SW wavefront kernel, as measured by the
    committed counter profile (two rocprofv3 --pmc passes, FETCH_SIZE and
    WRITE_SIZE, scaled by a plain copy of known size); None if absent."""
    path = os.path.join(ROOT, "profiles", "k_wave_sw_config1_hbm_traffic.txt")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch"):
                    return float(line.split("=")[1])
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--pairs", type=int, default=200000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-percall", action="store_true")
    ap.add_argument("--uniform", default=None,
                    help="M,N: every pair M x N residues instead of the log-normal mix")
    ap.add_argument("--algs", default="sw,nw,sg")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-long", action="store_true")
    ap.add_argument("--no-stats", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="plan config 1 with the instruction-count step costs")
    ap.add_argument("--serve-requests", type=int, default=64)
    ap.add_argument("--qps", type=float, default=8.0)
    ap.add_argument("--qps-sweep", default="4,8,16,24",
                    help="request rates of the serving sweep (comma-separated; empty: none)")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device once (HBM counter calibration)")
    a = ap.parse_args()
    if a.copy_calib:
        copy_calibration(a.copy_calib)
        return
    if a.gpus != 1:
        print("bench.py measures one GPU: every configuration of this path fits one MI355X and "
              "no multi-GPU figure is produced (pairs are independent; see the sharding "
              "design).", file=sys.stderr)
        sys.exit(2)
    if not P.available():
        raise RuntimeError("no HIP device: %s" % P.lib().parasail_hip_last_error())
    cus, clk = ctypes.c_int(), ctypes.c_int()
    P.lib().parasail_hip_device_info(ctypes.byref(cus), ctypes.byref(clk))
    peak_ops = cus.value * clk.value * 1e3 * 128  # packed int16: 64 lanes x 2 halves
    log("device: %d CUs at %.0f MHz; packed-int16 VALU peak %.2f TOP/s"
        % (cus.value, clk.value / 1e3, peak_ops / 1e12))

    t0 = time.perf_counter()
    if a.uniform:
        um, un = (int(x) for x in a.uniform.split(","))
        rng = np.random.default_rng(5)
        batch = WL.PairBatch(WL.protein_bytes(rng, um * a.pairs), np.full(a.pairs, um),
                             WL.protein_bytes(rng, un * a.pairs), np.full(a.pairs, un))
    else:
        batch = WL.config1(npairs=a.pairs)
    algs = tuple(a.algs.split(","))
    cells = batch.cells
    log("workload: %d pairs, %.3e cells per algorithm (generated in %.1f s)"
        % (len(batch), cells, time.perf_counter() - t0))
    m = P.blosum62()
    bufs = {alg: {k: np.zeros(len(batch), dtype=np.int32)
                  for k in ("score", "end_query", "end_ref", "flag")} for alg in algs}
    # the planner's step costs measured on this device (parasail_hip_autotune)
    # against the instruction count it uses untuned: a few steps of each plan
    tune = None
    if not a.no_autotune:
        L = P.lib()
        L.parasail_hip_autotune.restype = ctypes.c_int
        L.parasail_hip_autotune.argtypes = [P.MP, ctypes.c_int, ctypes.c_int]
        L.parasail_hip_autotune_table.restype = ctypes.c_int
        L.parasail_hip_autotune_table.argtypes = [ctypes.POINTER(ctypes.c_int),
                                                  ctypes.POINTER(ctypes.c_double),
                                                  ctypes.POINTER(ctypes.c_double), ctypes.c_int]
        h0 = make_resident(batch, m)
        mp0 = MultiPass(h0, algs, bufs)
        for _ in range(2):
            mp0.run()
        t0 = time.perf_counter()
        for _ in range(3):
            mp0.run()
        untuned = (time.perf_counter() - t0) / 3 * 1e3
        L.parasail_hip_batch_free(h0)
        t0 = time.perf_counter()
        if L.parasail_hip_autotune(m, 11, 1) != 0:
            raise RuntimeError("parasail_hip_autotune failed")
        tune_s = time.perf_counter() - t0
        rows = (ctypes.c_int * 16)()
        swc, nwc = (ctypes.c_double * 16)(), (ctypes.c_double * 16)()
        k = L.parasail_hip_autotune_table(rows, swc, nwc, 16)
        tune = {"seconds": round(tune_s, 2), "untuned_ms_per_step": round(untuned, 2),
                "rows_per_lane": [rows[i] for i in range(k)],
                "sw_ns_per_wave_step": [round(swc[i], 4) for i in range(k)],
                "nw_ns_per_wave_step": [round(nwc[i], 4) for i in range(k)]}
        log("autotune: %.2f s; untuned step %.2f ms" % (tune_s, untuned))
    h = make_resident(batch, m)
    # a step: SW, NW and SG over the resident batch in one library call, each
    # pass's download and conversion overlapping the next pass's kernels
    mp = MultiPass(h, algs, bufs)
    for w in range(a.warmup):
        mp.run()
        log("warmup %d done" % w)
    kms = {alg: [] for alg in algs}
    mms = {alg: [] for alg in algs}
    t0 = time.perf_counter()
    for s in range(a.steps):
        for alg, (ms, mm) in zip(algs, mp.run()):
            kms[alg].append(ms)
            mms[alg].append(mm)
        log("step %d done" % s)
    elapsed = time.perf_counter() - t0
    # the multi-pass call is one HIP graph launch after its capture; the
    # same steps as separate launches, for what the graph saves
    L = P.lib()
    gcap, grep_ = ctypes.c_int(), ctypes.c_int()
    L.parasail_hip_batch_graph_stats(h, ctypes.byref(gcap), ctypes.byref(grep_))
    L.parasail_hip_batch_set_graph(h, 0)
    t0 = time.perf_counter()
    for s in range(a.steps):
        mp.run()
    elapsed_ng = time.perf_counter() - t0
    log("multi-pass step: %.2f ms as a graph replay, %.2f ms as launches"
        % (elapsed / a.steps * 1e3, elapsed_ng / a.steps * 1e3))
    P.lib().parasail_hip_batch_free(h)
    # the same batch from host arrays in one parasail_hip_align_batch call
    # per algorithm: encoding, upload, planning, kernels and download timed
    cold = {}
    for alg in algs:
        t1 = time.perf_counter()
        rc_ = P.align_arrays(alg, 0, 0, batch, 11, 1, m)
        dt = time.perf_counter() - t1
        same = all((rc_[k] == bufs[alg][k]).all() for k in ("score", "end_query", "end_ref"))
        cold[alg] = {"call_ms": round(dt * 1e3, 1), "gcups_call": round(cells / dt / 1e9, 1),
                     "kernel_ms": round(P.last_kernel_ms(), 2), "equals_resident": bool(same)}
        log("cold batch %s: %.1f ms call" % (alg, dt * 1e3))
    ms_per_step = elapsed / a.steps * 1e3
    value = len(algs) * cells / (elapsed / a.steps) / 1e9
    per_alg = {alg: round(cells / (np.mean(kms[alg]) / 1e3) / 1e9, 1) for alg in algs}
    a0 = "sw" if "sw" in algs else algs[0]
    # roofline of the dominant kernel: the SW wavefront kernel, one launch
    # per pass, timed with HIP events on its own stream
    main_s = np.mean(mms[a0]) / 1e3
    achieved = cells * OPS[a0] / main_s
    w32 = {alg: int(np.count_nonzero(bufs[alg]["flag"] & P.FLAG_BITS_32)) for alg in algs}
    line = {
        "metric": "GCUPS batched vs per-pair kernel on 1 MI355X, bit-exact; serving p50/p99 ms",
        "value": round(value, 1), "unit": "GCUPS", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(ms_per_step, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int16", "data": "synthetic",
        "config": {"workload": "200k protein pairs, log-normal(300, 0.6) lengths, "
                               "BLOSUM62 11/1, SW+NW+SG score-only, int16-first sat",
                   "pairs": len(batch), "cells_per_alg": cells},
        "gcups_kernel_by_alg": per_alg,
        # each pass's main kernel (HIP events) at its fixed ops per cell
        "frac_by_alg": {alg: round(cells * OPS[alg] / (np.mean(mms[alg]) / 1e3) / peak_ops, 4)
                        for alg in algs},
        "graph": {"captures": gcap.value, "replays": grep_.value,
                  "ms_per_step_without_graph": round(elapsed_ng / a.steps * 1e3, 2)},
        "autotune": dict(tune, tuned_ms_per_step=round(ms_per_step, 2)) if tune else None,
        "cold_batch": dict(cold, includes="residues from host arrays: encoding, upload, "
                                          "planning, kernels, download; first call of "
                                          "each algorithm"),
        "pairs_int32_tier": w32,
        "roofline": {"bound": "valu", "achieved": round(achieved / 1e12, 2),
                     "peak": round(peak_ops / 1e12, 2), "unit": "TOP/s",
                     "frac": round(achieved / peak_ops, 4),
                     "traffic": traffic_from_profile() if a0 == "sw" and not a.uniform else None,
                     "kernel": "k_wave<%s> (%s wavefront, one launch per pass)"
                               % ("true" if a0 == "sw" else "false", a0.upper()),
                     "kernel_ms": round(main_s * 1e3, 3), "ops_per_cell": OPS[a0]},
    }
    if not a.no_percall:
        g, c0, s0 = per_call_gcups()
        g16, _, s16 = per_call_gcups(threads=16)
        line["per_pair_kernel"] = {"gcups": round(g, 3), "cells": c0, "seconds": round(s0, 3),
                                   "gcups_16_threads": round(g16, 3),
                                   "seconds_16_threads": round(s16, 3),
                                   "function": "parasail_sw_striped_profile_hip_64_16",
                                   "workload": "1 x 300 aa query vs 2,000 targets U[50,1000]",
                                   "threads_from": "Python threads (ctypes)"}
        drv = per_call_driver((1, 16, 32))
        line["per_pair_kernel"]["c_openmp"] = {
            "gcups_by_threads": {str(t): round(v[0], 3) for t, v in sorted(drv.items())},
            "driver": "parasail_amd/percall_omp: OpenMP over the targets, one profile, each "
                      "result freed (tests/test_openmp.c:353-369), best of 3"}
        if "sw" in per_alg:
            line["batched_over_per_pair_sw"] = round(per_alg["sw"] / g, 1)
        if not a.no_cpu:
            c1, c16 = per_call_cpu(1), per_call_cpu(min(os.cpu_count() or 1, 16))
            if c1 and c16:
                line["per_pair_kernel"]["cpu_avx2_striped_16"] = {
                    "gcups_1_core": round(c1[0], 3), "gcups_16_cores": round(c16[0], 3),
                    "cores": min(os.cpu_count() or 1, 16), "equals_serial": c1[2] and c16[2],
                    "kind": "port", "note": "the same 2,000 calls through the restated AVX2 "
                    "striped SW at 16 bits on the host, best of 3"}
    if not a.no_stats:
        line["stats_trace"] = stats_trace(cus.value * clk.value * 1e3 * 64)
    if not a.no_long:
        line["long_pairs"] = long_pairs(cus.value * clk.value * 1e3 * 64)
    if not a.no_serve:
        line["serving"] = serving(a.serve_requests, a.qps,
                                  [float(x) for x in a.qps_sweep.split(",") if x])
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(batch, gpu=bufs)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
