#!/usr/bin/env python3
"""Throughput of the gfx950 exact k-mer count on one MI355X.

One step = the 24-sample cohort (k=31 canonical, 256 partitions; a 20 Mbp
base genome plus 2% SNVs per sample, 1% read errors, log-uniform 2e5..8e6
reads of 150 bp per sample) counted from pinned host memory through the
pipelined batch: H2D of the FASTQ, count, DVLE chunk encode on the device,
D2H of the chunk payloads a .kite file holds.  `value` = input k-mers of
all 24 samples per step / time per step.  Beside it: the plain per-sample
path on the same cohort ("batched vs plain"), the dominant kernel's
roofline (HIP events over the timed steps) and the PCIe rate against the
PCIe 5.0 x16 link (and the pinned-copy rate measured in the same run).

More measurements ride on the same line: Pearson + Spearman of the
cohort's union at --min 2 (split-bf16 MFMA Gram against the FP64 Gram and
the FP64 oracle), the single-sample k=31 count with its FASTQ resident in
HBM (the count kernels' own rate), `kite dump`, BGZF inflate, and the CPU
baseline (the reference's count loop restated in C, on a bounded sample of
the cohort).

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, AMD published


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(arrs, names, threads: int, reps: int = 3) -> dict:
    """Time the reference's count loop restated in C (`kite_oracle
    count-ref`: workers taking record-aligned chunks in turn, inline rolling
    k-mers into 64-entry write-combining partition slots, then per worker and
    partition one run -- MSD/LSD byte radix, RLE, DVLE encode; output
    discarded) on a bounded sample of the cohort: whole samples, each its own
    count run as the reference would do them, summed; median of `reps`
    passes."""
    exe = ROOT / "oracle" / "build" / "kite_oracle"
    if not exe.exists():
        return {"value": None, "unit": "k-mers/s", "cores": 0, "kind": "port", "sample": "oracle binary missing"}
    tmpdir = "/dev/shm" if os.path.isdir("/dev/shm") else tempfile.gettempdir()
    paths = []
    try:
        for a in arrs:
            p = Path(tempfile.mkstemp(dir=tmpdir, suffix=".fq")[1])
            a.tofile(p)
            paths.append(p)
        rates, kmers = [], 0
        for _ in range(reps):
            secs, kmers = 0.0, 0
            for p in paths:
                # chunk size as the reference planner picks it for a large budget: 64 MiB
                cmd = [str(exe), "count-ref", str(p), "/dev/null", "31", "1", "256", str(64 << 20), str(threads)]
                r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
                if r.returncode != 0:
                    raise RuntimeError(r.stderr)
                res = json.loads(r.stdout.strip().splitlines()[-1])
                secs += res["count_s"]
                kmers += res["kmers"]
            rates.append(kmers / secs)
        return {"value": float(np.median(rates)), "unit": "k-mers/s", "cores": threads, "kind": "port",
                "host_nproc": os.cpu_count(),
                "sample": f"cohort samples {', '.join(names)} ({kmers} k-mers), each its own count run, "
                          f"median of {reps} passes, {threads} threads, output /dev/null",
                "note": "the reference's count loop restated in C (count/mod.rs:512-540, worker.rs:56-449, "
                        "sorter.rs:67-272, 514-530): write-combined partition scatter, MSD/LSD byte radix "
                        "with comparison-sort leaves, saturating RLE and DVLE encode per worker run; without "
                        "its AVX-512 histogram; 64 MiB chunks, the reference planner's choice on a "
                        "large-memory host",
                "rates": rates}
    finally:
        for p in paths:
            p.unlink(missing_ok=True)


def traffic_from_profile(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` on `workload`, from the committed
    rocprofv3 --pmc summary (FETCH_SIZE and WRITE_SIZE, copy-calibrated)."""
    path = ROOT / "profiles" / f"hbm_traffic_{workload}.md"
    if not path.exists():
        return None, None
    for line in path.read_text().splitlines():
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 2 and cells[0] == f"traffic_bytes_per_launch:{kernel}":
            return float(cells[1]), str(path.relative_to(ROOT))
    return None, None


def make_batch24(scale: float):
    """The 24-sample cohort's FASTQ in pinned host memory (one PinnedArray
    per sample) and its workloads."""
    import ctypes as C

    from kite_amd import PinnedArray, synth_reads
    from kite_amd._native import load_synth
    from kite_amd.workloads import batch24

    ws = batch24(24, scale=scale)
    t0 = time.time()
    pins = []
    lib = load_synth()
    for w in ws:
        g = w.genome()
        spec = w.spec()
        need = lib.kite_synth_reads(C.byref(spec), g.ctypes.data, 0, w.n_reads, None, 0, 1)
        pa = PinnedArray(need)
        synth_reads(spec, g, 0, w.n_reads, out=pa.array)
        pins.append(pa)
    total = sum(p.nbytes for p in pins)
    log(f"batch24: {len(pins)} samples, {sum(w.n_reads for w in ws)} reads, {total / 1e9:.2f} GB "
        f"pinned FASTQ (generated in {time.time() - t0:.1f}s)")
    return ws, pins


# MI355X's host link: PCIe 5.0 x16, 32 GT/s per lane with 128b/130b
# encoding, 63.0 GB/s each way, 126.0 GB/s with both directions busy.
PCIE5_X16_BOTH_GBPS = 2 * 32e9 * 16 * 128 / 130 / 8 / 1e9


def pcie_peak(ctx) -> dict:
    """Page-locked host <-> device copy rates on this box through the
    engine's own runtime (kite_gpu_copy_peak): H2D alone, D2H alone, both
    at once on two streams; 1 GiB copies, best of 5 after one untimed copy
    each way."""
    out = ctx.copy_peak(1 << 30)
    log(f"pcie: H2D {out['h2d_GBps']:.1f} GB/s, D2H {out['d2h_GBps']:.1f} GB/s, "
        f"both at once {out['both_GBps']:.1f} GB/s (page-locked, 1 GiB copies)")
    return out


def settle_link(ctx, limit_s: float = 60.0) -> dict:
    """Waits, before the timed passes, until device-to-host copies run at
    nearly the host-to-device rate (256 MiB page-locked copies, checked
    every half second, at most limit_s).  Right after a process frees many
    GB of device memory -- this one's plain-mode passes, or an earlier
    process on the box -- D2H copies run at about half rate for several
    seconds while H2D copies do not."""
    t0 = time.perf_counter()
    checks = []
    while True:
        o = ctx.copy_peak(256 << 20)
        checks.append([round(o["h2d_GBps"], 1), round(o["d2h_GBps"], 1)])
        ok = o["d2h_GBps"] >= 0.9 * o["h2d_GBps"]
        if ok or time.perf_counter() - t0 > limit_s:
            break
        time.sleep(0.5)
    waited = time.perf_counter() - t0
    if len(checks) > 1 or not ok:
        log(f"host link settled={ok} after {waited:.1f} s ({len(checks)} checks, first H2D/D2H GB/s "
            f"{checks[0]}, last {checks[-1]})")
    return {"settled": ok, "waited_s": waited, "checks": len(checks), "first_gbps": checks[0],
            "last_gbps": checks[-1]}


def size_class(nbytes: int) -> int:
    """The engine's input size classes for tuned settings (below 64 MiB,
    512 MiB, 4 GiB, above)."""
    return 0 if nbytes < 64 << 20 else (1 if nbytes < 512 << 20 else (2 if nbytes < 4 << 30 else 3))


TUNED_KEYS = ("split_target_div", "max_digit_bits", "hash_classes", "extract_subs", "split_above_div")


def apply_tuned(ctx, spec: str, k: int, P: int) -> dict:
    """--tuned "CLASS:div,bits,classes,subs,above;...": set the count
    settings of those size classes (kite_gpu_set_tuned) instead of timing
    candidates, e.g. to profile with the choices an earlier run made."""
    out = {}
    for part in filter(None, (x.strip() for x in spec.split(";"))):
        cls, vals = part.split(":")
        chosen = dict(zip(TUNED_KEYS, (int(v) for v in vals.split(","))))
        if len(chosen) != len(TUNED_KEYS):
            raise SystemExit(f"--tuned: class {cls} needs {len(TUNED_KEYS)} values")
        ctx.set_tuned(k, P, int(cls), chosen)
        out[str(int(cls))] = {"chosen": chosen, "default_ms": None, "tuned_ms": None, "sample": None,
                              "trials": 0, "fixed": "--tuned"}
        log(f"tuned settings for size class {cls}: {chosen}")
    return out


def autotune_cohort(ctx, arrs, names, k: int, canon: bool, P: int) -> dict:
    """kite_gpu_autotune on the largest sample of every size class in the
    cohort; the context keeps the choices for the batch that follows."""
    by_cls: dict[int, list[int]] = {}
    for i, a in enumerate(arrs):
        by_cls.setdefault(size_class(a.size), []).append(i)
    out = {}
    for cls, idx in sorted(by_cls.items()):
        i = max(idx, key=lambda j: arrs[j].size)
        t = time.perf_counter()
        rep = ctx.autotune(arrs[i], k, canon, P)
        rep["sample"] = names[i]
        rep["samples_in_class"] = len(idx)
        rep["seconds"] = time.perf_counter() - t
        log(f"autotune class {cls} on {names[i]} ({arrs[i].size / 1e9:.2f} GB): {rep['chosen']} "
            f"{rep['default_ms']:.2f} -> {rep['tuned_ms']:.2f} ms in {rep['trials']} trials ({rep['seconds']:.1f} s)")
        out[str(cls)] = rep
    return out


def kernel_passes(ctx, arrs, k: int, canon: bool, P: int, passes: int) -> dict:
    """Kernel timing passes over the cohort (untimed): the launches the
    batch's graphs hold, made one by one on the engine stream with HIP
    events around each (a captured graph holds no timing) -- the same
    kernels, grids and device-side sizes.  Per-kernel time, launches and
    algorithmic bytes summed over the passes."""
    from kite_amd import BATCH_PIPELINED

    def nothing_enc(i, st, codec, data, boff, rows, mm):
        return 0

    kern: dict[str, dict] = {}
    ctx.count_profile(True)
    try:
        for _ in range(passes):
            ctx.count_batch(arrs, k, canon, P, mode=BATCH_PIPELINED, on_result=nothing_enc, encoded=True)
            for name, v in ctx.count_kernel_stats().items():
                a = kern.setdefault(name, {"ms": 0.0, "launches": 0, "bytes": 0})
                for key in a:
                    a[key] += v[key]
    finally:
        ctx.count_profile(False)
    return kern


def bench_batch24(ctx, pins, steps: int, warmup: int, plain_reps: int) -> tuple[dict, list, dict, int]:
    """The headline: the 24-sample cohort from pinned host memory through
    the pipelined batch (H2D of the FASTQ, count, DVLE chunk encode on the
    device, D2H of the chunks), timed per pass over all 24 samples, against
    the plain per-sample path (allocation, H2D, count, encode, D2H, sync)."""
    from kite_amd import BATCH_PIPELINED, BATCH_PLAIN, BATCH_SERIAL

    arrs = [p.array for p in pins]
    k, canon, P = 31, True, 256

    # untimed: size the ring buffers, keep the raw tables for analyze, and
    # check that both modes give identical tables and identical chunk bytes
    def digest(i, st, keys, counts, off):
        return (st["kmers"], st["unique"], int(np.bitwise_xor.reduce(keys.ravel())) if keys.size else 0,
                int(counts.astype(np.uint64).sum()), int(off[-1]))

    tables = []

    def keep(i, st, keys, counts, off):
        tables.append((keys[:, 0].copy(), counts.copy(), off.copy()))
        return digest(i, st, keys, counts, off)

    d_pipe = ctx.count_batch(arrs, k, canon, P, mode=BATCH_PIPELINED, on_result=keep)
    kmers = sum(d[0] for d in d_pipe)
    unique = sum(d[1] for d in d_pipe)

    def cdigest(i, st, codec, data, boff, rows, mm):
        n8 = data.size // 8 * 8
        words = data[:n8].view(np.uint64)
        return (int(boff[-1]), int(np.bitwise_xor.reduce(words)) if words.size else 0,
                int(words.sum(dtype=np.uint64)) if words.size else 0, data[n8:].tobytes(),
                int(np.bitwise_xor.reduce(mm.ravel())))

    e_pipe = ctx.count_batch(arrs, k, canon, P, mode=BATCH_PIPELINED, on_result=cdigest, encoded=True)
    enc_bytes = sum(d[0] for d in e_pipe)
    fastq_bytes = sum(a.size for a in arrs)

    def nothing_enc(i, st, codec, data, boff, rows, mm):
        return 0

    def nothing(i, st, keys, counts, off):
        return 0

    def graph_flag(i, st, codec, data, boff, rows, mm):
        return st["graph"]

    link = settle_link(ctx)
    warm = []
    for i in range(warmup):
        t = time.perf_counter()
        ctx.count_batch(arrs, k, canon, P, mode=BATCH_PIPELINED, on_result=nothing_enc, encoded=True)
        warm.append(time.perf_counter() - t)
        log(f"batch24 warmup {i + 1}/{warmup}: {warm[-1] * 1e3:.1f} ms")
    # more untimed passes until two in a row agree within 3% of the fastest
    # so far, at most 30
    extra = 0
    while warmup and extra < 30 and (len(warm) < 2 or max(warm[-2:]) > 1.03 * min(warm)):
        t = time.perf_counter()
        ctx.count_batch(arrs, k, canon, P, mode=BATCH_PIPELINED, on_result=nothing_enc, encoded=True)
        warm.append(time.perf_counter() - t)
        extra += 1
    if extra:
        log(f"batch24 settled after {extra} more untimed passes: {[round(x * 1e3) for x in warm[warmup:]]} ms")

    # timed region: every sample's count and chunk encode is one launch of a
    # captured HIP graph (every batch call returns after its last D2H and
    # callback)
    walls = []
    t_start = time.perf_counter()
    for i in range(steps):
        t = time.perf_counter()
        got = ctx.count_batch(arrs, k, canon, P, mode=BATCH_PIPELINED, on_result=graph_flag, encoded=True)
        walls.append(time.perf_counter() - t)
    wall = time.perf_counter() - t_start
    graph_samples = int(sum(got))
    log(f"batch24 timed {steps} steps: {wall / steps * 1e3:.1f} ms/step, {kmers * steps / wall / 1e9:.2f} G k-mers/s "
        f"(steps {min(walls) * 1e3:.1f}..{max(walls) * 1e3:.1f} ms; {graph_samples} of {len(arrs)} samples "
        f"as one graph launch each)")
    kern_passes = 3
    kern = kernel_passes(ctx, arrs, k, canon, P, kern_passes)
    # the classic pipeline for comparison: kernels launched one by one after
    # the host read-backs of each sample (KITE_BATCH_GRAPH=0)
    os.environ["KITE_BATCH_GRAPH"] = "0"
    try:
        classic = []
        for i in range(kern_passes):
            t = time.perf_counter()
            ctx.count_batch(arrs, k, canon, P, mode=BATCH_PIPELINED, on_result=nothing_enc, encoded=True)
            classic.append(time.perf_counter() - t)
    finally:
        del os.environ["KITE_BATCH_GRAPH"]
    classic_med = float(np.median(classic))
    log(f"batch24 classic pipeline (kernels launched one by one): median {classic_med * 1e3:.1f} ms of "
        f"{[round(x * 1e3, 1) for x in classic]}")
    # two untimed passes: one with per-sample stage events on the three
    # streams (where the time goes), one with no events at all (what the
    # kernel events cost)
    ctx.count_profile(False, timeline=True)
    t = time.perf_counter()
    ctx.count_batch(arrs, k, canon, P, mode=BATCH_PIPELINED, on_result=nothing_enc, encoded=True)
    timeline_pass = time.perf_counter() - t
    tl = ctx.batch_timeline()
    ctx.count_profile(False)
    t = time.perf_counter()
    ctx.count_batch(arrs, k, canon, P, mode=BATCH_PIPELINED, on_result=nothing_enc, encoded=True)
    bare_pass = time.perf_counter() - t
    log(f"batch24 untimed passes: with stage events {timeline_pass * 1e3:.1f} ms, no events {bare_pass * 1e3:.1f} ms")

    # the per-sample plain path gives the same tables and chunks (checked
    # after the timed steps: its per-sample device allocations and frees
    # slow the host link for seconds after)
    d_plain = ctx.count_batch(arrs, k, canon, P, mode=BATCH_PLAIN, on_result=digest)
    assert d_pipe == d_plain, "plain and pipelined batch tables differ"
    e_plain = ctx.count_batch(arrs, k, canon, P, mode=BATCH_PLAIN, on_result=cdigest, encoded=True)
    assert e_pipe == e_plain, "plain and pipelined encoded chunks differ"

    # the plain per-sample path, median of plain_reps
    plain = []
    for _ in range(plain_reps):
        t = time.perf_counter()
        ctx.count_batch(arrs, k, canon, P, mode=BATCH_PLAIN, on_result=nothing_enc, encoded=True)
        plain.append(time.perf_counter() - t)
    plain_med = float(np.median(plain))
    log(f"batch24 plain (encoded): median {plain_med * 1e3:.1f} ms of {[round(x * 1e3, 1) for x in plain]}")
    # one sample at a time in persistent buffers: no allocation, no overlap
    serial = []
    for _ in range(plain_reps):
        t = time.perf_counter()
        ctx.count_batch(arrs, k, canon, P, mode=BATCH_SERIAL, on_result=nothing_enc, encoded=True)
        serial.append(time.perf_counter() - t)
    serial_med = float(np.median(serial))
    log(f"batch24 serial (encoded, persistent buffers): median {serial_med * 1e3:.1f} ms of "
        f"{[round(x * 1e3, 1) for x in serial]}")
    # the raw-table modes (rows instead of chunk payloads cross PCIe), once each
    raw = {}
    for name, mode in (("pipelined", BATCH_PIPELINED), ("plain", BATCH_PLAIN)):
        t = time.perf_counter()
        ctx.count_batch(arrs, k, canon, P, mode=mode, on_result=nothing)
        dt = time.perf_counter() - t
        raw[name] = {"wall_s": dt, "kmers_per_s": kmers / dt}
        log(f"batch24 raw tables {name}: {dt * 1e3:.1f} ms")

    # stage occupancy of the last timed step, from the per-sample events
    span = float(tl[:, 5].max()) if len(tl) else 0.0
    busy = {name: float(np.sum(tl[:, b] - tl[:, a])) for name, a, b in
            (("h2d_ms", 0, 1), ("engine_ms", 2, 3), ("d2h_ms", 4, 5))}
    out = {
        "samples": len(arrs), "fastq_bytes": int(fastq_bytes), "input_kmers": int(kmers),
        "unique_kmers_total": int(unique), "chunk_bytes": int(enc_bytes), "wall_s": wall, "steps": steps,
        "step_ms": [w * 1e3 for w in walls],
        "plain": {"wall_s": plain_med, "kmers_per_s": kmers / plain_med, "reps_s": plain,
                  "includes": "per sample: device allocation, H2D, count, chunk encode, D2H, synchronise, free"},
        "batched_vs_plain": plain_med / (wall / steps),
        "serial": {"wall_s": serial_med, "kmers_per_s": kmers / serial_med, "reps_s": serial,
                   "includes": "per sample, in persistent device buffers: H2D, count, chunk encode, D2H, "
                               "synchronise (nothing overlapped)"},
        "batched_vs_serial": serial_med / (wall / steps),
        "graph": {"samples_per_step_as_one_graph_launch": graph_samples,
                  "classic_pipeline_ms": classic_med * 1e3, "classic_reps_ms": [x * 1e3 for x in classic],
                  "note": "timed steps: each sample's count and its chunk encode are one hipGraphLaunch each, no "
                          "host read between samples; kernel timing passes: the same launches made one by one "
                          "with HIP events; classic: the pipeline with kernels launched after host read-backs "
                          "(KITE_BATCH_GRAPH=0)"},
        "kernel_timing_passes": kern_passes,
        "link_settle": link,
        "warmup_ms": [x * 1e3 for x in warm],
        "raw_tables": {**raw, "speedup": raw["plain"]["wall_s"] / raw["pipelined"]["wall_s"],
                       "bytes_d2h": int(unique * 12)},
        "bit_exact_plain_vs_pipelined": True,
        "stages": {"span_ms": span, **busy,
                   "h2d_frac": busy["h2d_ms"] / span if span else None,
                   "d2h_frac": busy["d2h_ms"] / span if span else None,
                   "pass_ms_with_stage_events": timeline_pass * 1e3, "pass_ms_without_events": bare_pass * 1e3,
                   "per_sample_ms": [[round(float(x), 3) for x in row] for row in tl],
                   "columns": ["h2d0", "h2d1", "run0", "run1_encoded", "d2h0", "d2h1"]},
        "ring": int(os.environ.get("KITE_BATCH_RING", "3")),
    }
    return out, tables, kern, kmers


def bench_analyze(ctx, tables: list, min_count: int, cpu_rows_frac_parts: int) -> dict:
    """Correlation of the union of the cohort tables (counts >= min_count)."""
    from kite_amd._native import CORR_FP64_BASELINE
    filt = []
    for keys, counts, off in tables:
        m = counts >= min_count
        cs = np.concatenate([np.zeros(1, dtype=np.uint64), np.cumsum(m, dtype=np.uint64)])
        filt.append((keys[m], counts[m].astype(np.uint64), cs[off.astype(np.int64)]))
    t = time.perf_counter()
    r = ctx.analyze_tables(filt, 1, methods=3, flags=CORR_FP64_BASELINE)
    wall = time.perf_counter() - t
    st = r["stats"]
    err_p = float(np.max(np.abs(r["pearson"] - r["pearson_fp64"])))
    err_s = float(np.max(np.abs(r["spearman"] - r["spearman_fp64"])))
    log(f"analyze: {st['rows']} union rows x {st['cols']} ({st['code_bytes']} B codes); Pearson + Spearman "
        f"split-bf16 Gram {st['gram_ms_total']:.2f} ms, FP64 Gram {st['fp64_ms_total']:.2f} ms "
        f"({st['passes']} / {st['fp64_passes']} passes); max |bf16 - fp64| pearson {err_p:.2e} spearman {err_s:.2e}")
    out = {
        "union_rows": int(st["rows"]), "samples": int(st["cols"]), "code_bytes": int(st["code_bytes"]),
        "corr_ms_split_bf16": st["gram_ms_total"],
        "corr_ms_fp64_gpu": st["fp64_ms_total"],
        "speedup_vs_fp64_gpu": st["fp64_ms_total"] / max(1e-9, st["gram_ms_total"]),
        "passes": {"split_bf16": int(st["passes"]), "fp64": int(st["fp64_passes"])},
        "includes": "Gram of the Pearson and Spearman values + FP64 reduction; the split-bf16 Gram computes "
                    "both in one pass over the matrix, the FP64 Gram one pass per method (its faster form)",
        "max_abs_err_vs_fp64_gpu": {"pearson": err_p, "spearman": err_s},
        "analyze_wall_s": wall,
    }
    # FP64 oracle on a row sample: the union rows of the first partitions
    P = len(filt[0][2]) - 1
    sub = []
    for keys, counts, off in filt:
        e = int(off[cpu_rows_frac_parts])
        o = off[:cpu_rows_frac_parts + 1].copy()
        sub.append((keys[:e], counts[:e],
                    np.concatenate([o, np.full(P - cpu_rows_frac_parts, o[-1], dtype=np.uint64)])))
    uk, um, uo = ctx.union_tables(sub, 1)
    sys.path.insert(0, str(ROOT / "oracle"))
    import corr_oracle as co
    t = time.perf_counter()
    x = co.transform(um)
    p_cpu = co.pearson(x)
    s_cpu = co.spearman(um.astype(np.float64))
    cpu_s = time.perf_counter() - t
    g = ctx.corr_counts(um, methods=3, flags=CORR_FP64_BASELINE)
    out["cpu_fp64_oracle"] = {
        "rows": int(um.shape[0]), "seconds": cpu_s, "rows_per_s": um.shape[0] / cpu_s, "cores": 1,
        "note": "numpy FP64 (one thread except the BLAS Gram)",
        "sample": f"union rows of partitions [0, {cpu_rows_frac_parts}) of {P}",
        "max_abs_err_split_bf16": {"pearson": float(np.max(np.abs(g["pearson"] - p_cpu))),
                                   "spearman": float(np.max(np.abs(g["spearman"] - s_cpu)))},
        "max_abs_err_fp64_gpu": {"pearson": float(np.max(np.abs(g["pearson_fp64"] - p_cpu))),
                                 "spearman": float(np.max(np.abs(g["spearman_fp64"] - s_cpu)))},
    }
    log(f"analyze oracle sample: {um.shape[0]} rows in {cpu_s:.1f}s; max err bf16 vs oracle "
        f"{out['cpu_fp64_oracle']['max_abs_err_split_bf16']}")
    return out


def merge_inputs(dirpath: Path, reads: int) -> list[Path]:
    """Two k=63 canonical samples of the 100 Mbp genome (P=1024) with
    different read seeds, each counted by `kite count` on the GPU into
    dirpath/s0.kite and dirpath/s1.kite."""
    import dataclasses
    import subprocess
    from kite_amd.workloads import K63_CANON_100MBP
    base = K63_CANON_100MBP.scaled(reads)
    samples = [dataclasses.replace(base, seed=3003, genome_seed=3003),
               dataclasses.replace(base, seed=3004, genome_seed=3003, name=base.name + "-second")]
    genome = samples[0].genome()
    kite_bin = ROOT / "kite_amd" / "lib" / "kite"
    out = []
    for i, w in enumerate(samples):
        fq_path = dirpath / f"s{i}.fq"
        w.fastq(genome=genome).tofile(fq_path)
        kp = dirpath / f"s{i}.kite"
        r = subprocess.run([str(kite_bin), "count", "-i", str(fq_path), "-o", str(kp), "-k", "63", "--canonical",
                            "true", "--partitions", "1024"], capture_output=True, text=True, timeout=600)
        if r.returncode:
            raise RuntimeError(f"kite count failed: {r.stderr[-2000:]}")
        fq_path.unlink()
        out.append(kp)
    return out


def bench_merge(reads: int = 1_000_000, settle=None) -> dict:
    """`kite merge` of two k=63 samples (merge_inputs): the device engine
    (kite_gpu_merge_runs: the stored payloads to the device, run decode,
    header filter, merge-path union with u64 sums, u32 check and DVLE
    encode there, the chunk payloads back to the writer) against the host
    engine (`--engine cpu`: partition decode, union and DVLE encode on this
    host's worker threads), byte-identical outputs, best of two runs each.
    `merge_s`: the device engine's kite_gpu_merge_runs call (everything
    between reading the inputs' indexes and finalizing the output); the
    host engine's unions alone.  `settle`, when given, runs before every
    timed device-engine run (the host link's recovery after the previous
    process freed its device memory; see settle_link)."""
    import re
    import subprocess
    import tempfile
    kite_bin = ROOT / "kite_amd" / "lib" / "kite"
    tmp = Path(tempfile.mkdtemp(prefix="kite_merge_", dir="/dev/shm" if os.path.isdir("/dev/shm") else None))
    try:
        t0 = time.perf_counter()
        ins = merge_inputs(tmp, reads)
        log(f"merge inputs: 2 x {reads} reads (k=63, P=1024) counted in {time.perf_counter() - t0:.1f} s, "
            f"{sum(p.stat().st_size for p in ins) / 1e9:.2f} GB .kite")
        res = {}
        for engine in ("gpu", "cpu"):
            best = None
            for _ in range(2):
                outp = tmp / f"m_{engine}.kite"
                if settle and engine == "gpu":
                    settle()
                t = time.perf_counter()
                r = subprocess.run([str(kite_bin), "merge", "--engine", engine, "-i", *map(str, ins), "-o", str(outp)],
                                   capture_output=True, text=True, timeout=600)
                wall = time.perf_counter() - t
                if r.returncode:
                    raise RuntimeError(f"kite merge failed: {r.stderr[-2000:]}")
                m = re.search(r"wrote (\d+) rows .* merge ([\d.]+)s", r.stderr)
                cur = {"wall_s": wall, "merge_s": float(m.group(2)), "rows": int(m.group(1))}
                g = re.search(r"in ([\d.]+) ms, device decode ([\d.]+) ms, merge ([\d.]+) ms, "
                              r"encode ([\d.]+) ms, out ([\d.]+) ms", r.stderr)
                if g:
                    cur["stages_ms"] = dict(zip(("payloads_in", "device_decode", "device_merge", "device_encode",
                                                 "chunks_out_and_write"), map(float, g.groups())))
                if best is None or cur["wall_s"] < best["wall_s"]:
                    best = cur
            res[engine] = best
            log(f"kite merge --engine {engine}: {best['wall_s']:.2f} s wall, merge {best['merge_s'] * 1e3:.1f} ms, "
                f"{best['rows']} rows" + (f", stages {best['stages_ms']}" if "stages_ms" in best else ""))
        same = (tmp / "m_gpu.kite").read_bytes() == (tmp / "m_cpu.kite").read_bytes()
        in_rows = []
        for p in ins:
            r = subprocess.run([str(kite_bin), "info", "-i", str(p)], capture_output=True, text=True, timeout=120)
            mm = re.search(r"Total Rows:\s*(\d+)", r.stdout)
            in_rows.append(int(mm.group(1)) if mm else None)
        return {"workload": f"two k63-canon-100mbp samples of {reads} reads (seeds 3003, 3004), P=1024",
                "input_kite_bytes": [p.stat().st_size for p in ins], "input_rows": in_rows,
                "gpu": res["gpu"], "cpu": res["cpu"], "bit_exact_gpu_vs_cpu": same,
                "wall_speedup": res["cpu"]["wall_s"] / res["gpu"]["wall_s"],
                "includes": "wall: process start, .kite reads, partition decode, union, chunk encode, .kite write "
                            "(/dev/shm)"}
    finally:
        import shutil
        shutil.rmtree(tmp, ignore_errors=True)


def bench_dump(w, fq, settle=None) -> dict:
    """`kite dump` of the workload's .kite (written by `kite count` on the
    GPU): the device engine against the host engine (partition workers on
    this host's cores), wall time to /dev/null, and a hash of each engine's
    full text (bit-exact check at full size).  `settle`, when given, runs
    before every timed run of the CLI on the device: each CLI process frees
    its device memory at exit, and device-to-host copies run at about half
    rate for seconds after such a free (see settle_link)."""
    import re
    import subprocess
    import tempfile
    try:
        import xxhash
        mkhash = xxhash.xxh3_128
    except ImportError:  # pragma: no cover
        import hashlib
        mkhash = hashlib.sha256
    kite_bin = ROOT / "kite_amd" / "lib" / "kite"
    tmp = Path(tempfile.mkdtemp(prefix="kite_dump_", dir="/tmp"))
    try:
        fq_path = tmp / "r.fq"
        fq.tofile(fq_path)
        kp = tmp / "c.kite"
        # the drop-in CLI end to end: FASTQ file (page cache) -> .kite file
        if settle:
            settle()
        t0 = time.perf_counter()
        r = subprocess.run([str(kite_bin), "count", "-i", str(fq_path), "-o", str(kp), "-k", str(w.k),
                            "--canonical", "true" if w.canonical else "false", "--partitions", str(w.partitions)],
                           capture_output=True, text=True, timeout=600)
        count_wall = time.perf_counter() - t0
        if r.returncode:
            raise RuntimeError(f"kite count failed: {r.stderr[-2000:]}")
        done = [ln for ln in r.stderr.splitlines() if "done:" in ln]
        log(f"kite count CLI: {count_wall:.2f} s wall for {fq.size / 1e9:.3f} GB FASTQ -> "
            f"{kp.stat().st_size / 1e9:.3f} GB .kite ({done[-1].split('done: ')[-1] if done else ''})")
        # --mode compact: the same count, chunks zstd-compressed on the host
        # where the reference's policy keeps them
        kz = tmp / "z.kite"
        if settle:
            settle()
        t0 = time.perf_counter()
        r = subprocess.run([str(kite_bin), "count", "-i", str(fq_path), "-o", str(kz), "-k", str(w.k),
                            "--canonical", "true" if w.canonical else "false", "--partitions", str(w.partitions),
                            "--mode", "compact"], capture_output=True, text=True, timeout=600)
        compact_wall = time.perf_counter() - t0
        if r.returncode:
            raise RuntimeError(f"kite count --mode compact failed: {r.stderr[-2000:]}")
        compact_bytes = kz.stat().st_size
        kz.unlink()
        log(f"kite count CLI --mode compact: {compact_wall:.2f} s wall -> {compact_bytes / 1e9:.3f} GB .kite")
        fq_path.unlink()
        threads = min(16, os.cpu_count() or 1)
        out = {"kite_bytes": kp.stat().st_size, "host_threads": threads,
               "cli_count": {"wall_s": count_wall,
                             "includes": "process start, device init, FASTQ read from the page cache, count, "
                                         "DVLE encode, .kite write"},
               "cli_count_compact": {"wall_s": compact_wall, "kite_bytes": compact_bytes,
                                     "includes": "as cli_count, plus zstd level 5-6 of the chunks on the host "
                                                 "where the chunk codec policy keeps it"}}
        for engine in ("gpu", "host"):
            env = dict(os.environ, KITE_DUMP_ENGINE=engine)
            cmd = [str(kite_bin), "dump", "-i", str(kp), "-t", str(threads)]
            best = None
            for _ in range(2):
                if settle and engine == "gpu":
                    settle()
                t0 = time.perf_counter()
                r = subprocess.run(cmd + ["-o", "/dev/null"], capture_output=True, text=True, env=env, timeout=600)
                wall = time.perf_counter() - t0
                if r.returncode:
                    raise RuntimeError(f"kite dump ({engine}) failed: {r.stderr[-2000:]}")
                best = wall if best is None else min(best, wall)
            e = {"wall_s": best}
            m = re.search(r"gpu: (\d+) runs, (\d+) rows, (\d+) lines, (\d+) bytes in (\d+) group.*decode ([\d.]+) ms, "
                          r"merge ([\d.]+) ms, format ([\d.]+) ms; text ([\d.]+) ms", r.stderr)
            if m:
                e.update({"rows": int(m.group(2)), "lines": int(m.group(3)), "text_bytes": int(m.group(4)),
                          "groups": int(m.group(5)), "decode_ms": float(m.group(6)), "merge_ms": float(m.group(7)),
                          "format_ms": float(m.group(8)), "text_ms": float(m.group(9))})
            # full text hash through a pipe
            h = mkhash()
            p = subprocess.Popen(cmd + ["-o", "-"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, env=env)
            n = 0
            while True:
                b = p.stdout.read(16 << 20)
                if not b:
                    break
                n += len(b)
                h.update(b)
            if p.wait(timeout=600):
                raise RuntimeError(f"kite dump ({engine}) to stdout failed")
            e["hash"] = h.hexdigest()
            e["stdout_bytes"] = n
            out[engine] = e
            log(f"dump {engine}: {best:.2f} s wall" + (f", device {e['decode_ms'] + e['merge_ms'] + e['format_ms']:.1f} ms"
                                                      if m else ""))
        g, hst = out["gpu"], out["host"]
        out["bit_exact_gpu_vs_host"] = g["hash"] == hst["hash"] and g["stdout_bytes"] == hst["stdout_bytes"]
        rows = g.get("rows", 0)
        out["rows_per_s"] = {"gpu_wall": rows / g["wall_s"], "host_wall": rows / hst["wall_s"],
                             "gpu_device": rows / ((g["decode_ms"] + g["merge_ms"] + g["format_ms"]) * 1e-3)
                             if rows else None}
        out["speedup_wall"] = hst["wall_s"] / g["wall_s"]
        out["includes"] = ("wall: process start, .kite read, decode, merge, format, text to /dev/null; "
                           "device: decode + merge + format kernels; each device run after the host link settled")
        return out
    finally:
        import shutil
        shutil.rmtree(tmp, ignore_errors=True)


def bench_bgzf(ctx, w, fq, threads: int) -> dict:
    """BGZF input: the workload's FASTQ as BGZF members (64 KiB pieces,
    zlib level 6 on this host's cores), inflated on the device
    (kite_gpu_inflate_bgzf: the inflate kernel, all members resident in
    HBM) against zlib on the host cores, plus a whole count from the BGZF
    bytes (device inflate + count) with the same result as from plain."""
    import concurrent.futures as cf
    import zlib
    piece = 0xFF00
    raw = fq.tobytes()

    def member(i):
        d = raw[i:i + piece]
        c = zlib.compressobj(6, zlib.DEFLATED, -15, 9)
        z = c.compress(d) + c.flush()
        head = (bytes([0x1F, 0x8B, 8, 4, 0, 0, 0, 0, 0, 0xFF, 6, 0, 0x42, 0x43, 2, 0])
                + (12 + 6 + len(z) + 8 - 1).to_bytes(2, "little"))
        return head + z + (zlib.crc32(d) & 0xFFFFFFFF).to_bytes(4, "little") + len(d).to_bytes(4, "little")

    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(threads) as ex:
        members = list(ex.map(member, range(0, len(raw), piece)))
    comp = b"".join(members)
    log(f"bgzf: {len(members)} members, {len(comp) / 1e9:.3f} GB compressed in {time.perf_counter() - t0:.1f}s")
    # device inflate, twice (the first call sizes the buffers)
    out = None
    best = None
    for _ in range(2):
        out, ms = ctx.inflate_bgzf(comp)
        best = ms if best is None else min(best, ms)
    ok = out == raw
    del out
    # zlib on the host cores, member by member (the reference's thread pool)
    bounds = []
    pos = 0
    for m in members:
        bounds.append((pos, len(m)))
        pos += len(m)

    def inflate_host(b):
        o, n = b
        return len(zlib.decompress(comp[o + 18:o + n - 8], -15))

    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(threads) as ex:
        host_bytes = sum(ex.map(inflate_host, bounds))
    host_s = time.perf_counter() - t0
    # a whole count from the BGZF bytes
    t0 = time.perf_counter()
    ctx.count_begin(w.k, w.canonical, w.partitions)
    used = ctx.count_feed_bgzf(comp)
    st = ctx.count_finish()
    count_s = time.perf_counter() - t0
    res = {"members": len(members), "compressed_bytes": len(comp), "fastq_bytes": len(raw),
           "device_inflate_ms": best, "device_inflate_GBps": len(raw) / (best * 1e-3) / 1e9,
           "inflate_bit_exact": ok,
           "host_zlib": {"threads": threads, "seconds": host_s, "GBps": host_bytes / host_s / 1e9},
           "count_from_bgzf": {"wall_s": count_s, "kmers": int(st.kmers), "unique": int(st.unique),
                               "consumed_all": used == len(comp), "batches": int(st.batches),
                               "includes": "host->device copy of the compressed bytes, device inflate, count"}}
    log(f"bgzf: device inflate {best:.1f} ms ({res['device_inflate_GBps']:.1f} GB/s, exact={ok}); "
        f"host zlib {threads} threads {host_s:.2f} s; count from bgzf {count_s:.3f} s")
    return res


class DeviceBytes:
    """A device copy of host bytes, through the HIP runtime the engine
    library loaded (hipMalloc / hipMemcpy / hipFree)."""

    def __init__(self, arr: np.ndarray):
        import ctypes as C
        from kite_amd import load_gpu
        load_gpu()  # the engine's runtime is the one resolved below
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        p = C.c_void_p()
        if self.hip.hipMalloc(C.byref(p), max(arr.size, 1)) != 0:
            raise RuntimeError("hipMalloc failed")
        self.ptr, self.nbytes = p.value, arr.size
        if self.hip.hipMemcpy(C.c_void_p(self.ptr), arr.ctypes.data, arr.size, 1) != 0:  # host to device
            raise RuntimeError("hipMemcpy failed")

    def free(self) -> None:
        import ctypes as C
        if self.ptr:
            self.hip.hipFree(C.c_void_p(self.ptr))
            self.ptr = None


def calibrate_copy(nbytes: int = 1 << 30) -> None:
    """One device-to-device copy of a known size (the runtime's copy kernel):
    the reference point that scales FETCH_SIZE / WRITE_SIZE readings."""
    import ctypes as C
    src = DeviceBytes(np.zeros(nbytes, dtype=np.uint8))
    dst = DeviceBytes(np.zeros(1, dtype=np.uint8))
    p = C.c_void_p()
    if src.hip.hipMalloc(C.byref(p), nbytes) != 0:
        raise RuntimeError("hipMalloc failed")
    if src.hip.hipMemcpy(p, C.c_void_p(src.ptr), nbytes, 3) != 0:  # device to device
        raise RuntimeError("hipMemcpy failed")
    src.hip.hipDeviceSynchronize()
    src.hip.hipFree(p)
    src.free()
    dst.free()
    log(f"calibration copy: {nbytes >> 20} MiB read + {nbytes >> 20} MiB written")


def bench_single(ctx, w, steps: int, warmup: int, autotune: bool) -> tuple[dict, np.ndarray]:
    """One sample resident in HBM (the FASTQ already on the device, the
    tables left there): the count pipeline's own rate and per-kernel
    rooflines, without PCIe.  Autotuned on its own sample first (its genome
    differs from the cohort's)."""
    fq = w.fastq()
    tune = None
    if autotune:
        tune = ctx.autotune(fq, w.k, w.canonical, w.partitions)
        log(f"autotune {w.name}: {tune['chosen']} {tune['default_ms']:.2f} -> {tune['tuned_ms']:.2f} ms")
    dev = DeviceBytes(fq)
    ptr, nbytes = dev.ptr, dev.nbytes

    def step():
        ctx.count_begin(w.k, w.canonical, w.partitions)
        ctx.count_feed_device(ptr, nbytes)
        return ctx.count_finish()

    for i in range(warmup):
        st = step()
    ctx.count_profile(True)
    kern: dict[str, dict] = {}
    t0 = time.perf_counter()
    stats = []
    for i in range(steps):
        stats.append(step())  # returns after the device work
        for name, v in ctx.count_kernel_stats().items():
            a = kern.setdefault(name, {"ms": 0.0, "launches": 0, "bytes": 0})
            for key in a:
                a[key] += v[key]
    wall = time.perf_counter() - t0
    ctx.count_profile(False)
    kmers = stats[-1].kmers
    assert all(s.kmers == kmers and s.unique == stats[0].unique for s in stats)
    dev.free()
    rate = kmers * steps / wall
    log(f"single sample {w.name} (HBM-resident): {wall / steps * 1e3:.3f} ms/step, {rate / 1e9:.2f} G k-mers/s")
    out = {"workload": w.name, "k": w.k, "canonical": w.canonical, "partitions": w.partitions,
           "reads": w.n_reads, "fastq_bytes": int(nbytes), "input_kmers": int(kmers),
           "unique_kmers": int(stats[-1].unique), "steps": steps, "ms_per_step": wall / steps * 1e3,
           "kmers_per_s": rate, "roofline": roofline_of(kern, w.name), "kernels": kernel_table(kern, steps),
           "autotune": tune,
           "includes": "FASTQ resident in HBM; count kernels only; results stay in HBM"}
    return out, fq


def kernel_table(kern: dict, steps: int) -> dict:
    return {n: {"ms_per_step": v["ms"] / steps, "launches_per_step": v["launches"] / steps,
                "GBps": (v["bytes"] / (v["ms"] * 1e-3) / 1e9) if v["ms"] and v["bytes"] else None}
            for n, v in sorted(kern.items(), key=lambda kv: -kv[1]["ms"])}


def roofline_of(kern: dict, workload: str) -> dict:
    """Roofline of the dominant kernel (largest total time): algorithmic bytes
    per launch over its average launch time (HIP events)."""
    dom_name, dom = max(kern.items(), key=lambda kv: kv[1]["ms"])
    per_launch_s = dom["ms"] / dom["launches"] * 1e-3
    per_launch_bytes = dom["bytes"] / dom["launches"]
    achieved = per_launch_bytes / per_launch_s / 1e9
    traffic, traffic_src = traffic_from_profile(dom_name, workload)
    return {"bound": "hbm", "kernel": dom_name, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
            "algorithmic_bytes_per_launch": per_launch_bytes, "avg_launch_ms": per_launch_s * 1e3,
            "launches": dom["launches"]}


def single_only(ctx, args) -> int:
    """--no-batch: the single-sample leg alone, printed as the line (for
    profiling commands; not the headline configuration)."""
    from kite_amd.workloads import WORKLOADS
    w = WORKLOADS[args.single_workload]
    if args.tuned:
        apply_tuned(ctx, args.tuned, w.k, w.partitions)
    single, fq = bench_single(ctx, w, args.single_steps, 3, not (args.no_autotune or args.tuned))
    line = {"metric": "input k-mers/s (single sample, FASTQ resident in HBM; profiling run)",
            "value": single["kmers_per_s"], "unit": "k-mers/s", "n_gpus": 1, "steps": single["steps"],
            "warmup": 3, "ms_per_step": single["ms_per_step"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "u64", "data": "synthetic (seeded generator, regenerated on the box)",
            "config": {"workload": w.name, "k": w.k, "canonical": w.canonical, "partitions": w.partitions},
            "roofline": single["roofline"], "kernels": single["kernels"], "autotune": single["autotune"]}
    print(json.dumps(line), flush=True)
    ctx.close()
    return 0


def bgzf_only(ctx, args) -> int:
    """--bgzf-only: the BGZF inflate leg alone on the single-sample workload,
    printed as the line (for profiling commands; not the headline)."""
    from kite_amd.workloads import WORKLOADS
    w = WORKLOADS[args.single_workload]
    threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    bgz = bench_bgzf(ctx, w, w.fastq(), threads)
    print(json.dumps({"bgzf_only": {"workload": w.name, **bgz}}), flush=True)
    ctx.close()
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch-scale", type=float, default=1.0, help="scale of the 24 read counts")
    ap.add_argument("--plain-reps", type=int, default=3)
    ap.add_argument("--single-workload", default="k31-canon-50mbp-5m")
    ap.add_argument("--single-steps", type=int, default=10)
    ap.add_argument("--cpu-samples", default="0,1,2", help="cohort samples the CPU baseline counts")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="count with the built-in default settings")
    ap.add_argument("--tuned", default="",
                    help='fixed count settings per input size class instead of the autotuner, '
                         '"CLASS:split_div,digit_bits,hash_classes,extract_subs,split_above;..." '
                         '(e.g. the choices of an earlier run, for profiling it)')
    ap.add_argument("--no-batch", action="store_true",
                    help="profiling runs: only the HBM-resident single-sample leg, printed as the line")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run one 1 GiB device-to-device copy (a known-size kernel that scales "
                         "FETCH_SIZE / WRITE_SIZE counter readings)")
    ap.add_argument("--kernel-passes-only", action="store_true",
                    help="for profiler runs: only the cohort's kernel timing passes (the graphs' launches made "
                         "one by one between HIP events; no graph launch, no timed steps); prints their "
                         "per-kernel table and roofline instead of the bench line")
    ap.add_argument("--no-analyze", action="store_true")
    ap.add_argument("--no-single", action="store_true", help="skip the HBM-resident single-sample leg")
    ap.add_argument("--no-dump", action="store_true", help="skip the kite dump leg")
    ap.add_argument("--no-bgzf", action="store_true", help="skip the BGZF inflate leg")
    ap.add_argument("--bgzf-only", action="store_true", help="run the BGZF inflate leg alone (profiling)")
    ap.add_argument("--no-merge", action="store_true", help="skip the kite merge leg")
    ap.add_argument("--merge-reads", type=int, default=1_000_000, help="reads of each k=63 merge input")
    ap.add_argument("--merge-inputs", default="",
                    help="profiling runs: only write the two k=63 merge inputs into this directory and exit")
    ap.add_argument("--analyze-cpu-parts", type=int, default=8,
                    help="partitions (of 256) whose union rows the CPU oracle sample covers")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(--gpus must be 1)", file=sys.stderr)
        return 2

    from kite_amd import GpuContext, KiteError
    from kite_amd.workloads import WORKLOADS

    threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    try:
        ctx = GpuContext(0)
    except KiteError as e:
        print(f"bench.py needs a HIP device: {e}", file=sys.stderr)
        return 1
    if args.merge_inputs:
        d = Path(args.merge_inputs)
        d.mkdir(parents=True, exist_ok=True)
        ctx.close()
        for p in merge_inputs(d, args.merge_reads):
            log(f"wrote {p} ({p.stat().st_size / 1e9:.2f} GB)")
        return 0
    if args.calibrate_copy:
        calibrate_copy()
    if args.no_batch:
        return single_only(ctx, args)
    if args.bgzf_only:
        return bgzf_only(ctx, args)
    # (an earlier process's device memory frees slow D2H copies for a while)
    settle_link(ctx)
    pcie = pcie_peak(ctx)
    ws, pins = make_batch24(args.batch_scale)
    tune = None
    if args.tuned:
        tune = apply_tuned(ctx, args.tuned, 31, 256)
    elif not args.no_autotune:
        tune = autotune_cohort(ctx, [p.array for p in pins], [w_.name for w_ in ws], 31, True, 256)
    workload = "batch24-k31-canon-20mbp-snv2"
    if args.kernel_passes_only:
        # one untimed pass of the same launches first (buffers sized, first use)
        kernel_passes(ctx, [p.array for p in pins], 31, True, 256, 1)
        kern = kernel_passes(ctx, [p.array for p in pins], 31, True, 256, 3)
        roof = roofline_of(kern, workload)
        log(f"kernel passes only: {roof['kernel']} {roof['avg_launch_ms']:.4f} ms per launch over "
            f"{roof['launches']} launches, {roof['achieved']:.0f} GB/s")
        print(json.dumps({"kernel_passes_only": {"workload": workload, "passes": 3, "roofline": roof,
                                                 "kernels": kernel_table(kern, 3)}}), flush=True)
        for p in pins:
            p.free()
        ctx.close()
        return 0
    batch, tables, kern, kmers = bench_batch24(ctx, pins, args.steps, args.warmup, args.plain_reps)
    wall = batch["wall_s"]
    value = kmers * args.steps / wall
    ms_per_step = wall / args.steps * 1e3
    roof = roofline_of(kern, workload)
    kernels = kernel_table(kern, batch["kernel_timing_passes"])
    log("batch24 kernels: " + json.dumps({n: round(v["ms_per_step"], 3) for n, v in kernels.items()}))
    moved = batch["fastq_bytes"] + batch["chunk_bytes"]
    achieved = moved / (wall / args.steps) / 1e9
    pcie_line = {"h2d_bytes": batch["fastq_bytes"], "d2h_bytes": batch["chunk_bytes"],
                 "achieved_GBps": achieved, "peak_GBps": PCIE5_X16_BOTH_GBPS,
                 "frac": achieved / PCIE5_X16_BOTH_GBPS,
                 "peak_source": "PCIe 5.0 x16, both directions (2 x 63.0 GB/s)",
                 "measured_frac": achieved / pcie["both_GBps"],
                 "measured_source": "kite_gpu_copy_peak in this run: page-locked 1 GiB copies, H2D and D2H at once",
                 **{f"measured_{k}": v for k, v in pcie.items()}}
    log(f"batch24 PCIe: {achieved:.1f} GB/s = {pcie_line['frac']:.2f} of the link's {PCIE5_X16_BOTH_GBPS:.1f}, "
        f"{pcie_line['measured_frac']:.2f} of the {pcie['both_GBps']:.1f} copies measured; "
        f"plain {batch['plain']['kmers_per_s'] / 1e9:.2f} G k-mers/s, "
        f"batched/plain {batch['batched_vs_plain']:.2f}x")

    analyze = None
    if not args.no_analyze:
        analyze = bench_analyze(ctx, tables, 2, args.analyze_cpu_parts)
    del tables

    cpu = None
    if not args.no_cpu_baseline:
        idx = [int(x) for x in args.cpu_samples.split(",") if x != ""]
        log(f"cpu baseline: the reference's count loop in C on cohort samples {idx}, {threads} threads")
        cpu = cpu_baseline([pins[i].array for i in idx], [ws[i].name for i in idx], threads)
        log(f"cpu baseline: {cpu['value'] / 1e6 if cpu['value'] else 0:.1f} M k-mers/s")
    for p in pins:
        p.free()
    del pins
    # the batch legs' device buffers go before the single-sample and CLI
    # legs: a `kite` process sizes its device batches from the device
    # memory free when it opens
    ctx.close()
    ctx = GpuContext(0)

    single = fq = None
    w = WORKLOADS[args.single_workload]
    if not args.no_single:
        single, fq = bench_single(ctx, w, args.single_steps, 3, not (args.no_autotune or args.tuned))
    if not args.no_dump:
        fq = w.fastq() if fq is None else fq
        single = single or {}
        single["dump"] = bench_dump(w, fq, settle=lambda: settle_link(ctx))
    merge = None
    if not args.no_merge:
        merge = bench_merge(args.merge_reads, settle=lambda: settle_link(ctx))
    if not args.no_bgzf:
        fq = w.fastq() if fq is None else fq
        bgz = bench_bgzf(ctx, w, fq, threads)
        if single and "input_kmers" in single and bgz["count_from_bgzf"]["kmers"] != single["input_kmers"]:
            raise SystemExit("bgzf count differs from the plain count")
        single = single or {}
        single["bgzf"] = bgz

    line = {
        "metric": "input k-mers/s, batched vs plain GPU (bit-exact); corr-matrix time @1e-5 err",
        "value": value,
        "unit": "k-mers/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded generator, regenerated on the box)",
        "config": {"workload": workload, "samples": 24, "k": 31, "canonical": True, "partitions": 256,
                   "reads": int(sum(w_.n_reads for w_ in ws)), "read_len": 150,
                   "fastq_bytes": batch["fastq_bytes"], "input_kmers": int(kmers),
                   "unique_kmers_total": batch["unique_kmers_total"], "chunk_bytes": batch["chunk_bytes"],
                   "step": "all 24 samples from pinned host memory: H2D of the FASTQ, count, DVLE chunk "
                           "encode on the device, D2H of the chunks (the .kite payloads)",
                   "batch_scale": args.batch_scale, "parallelism": "single GPU",
                   "autotune": ({c: {"chosen": r["chosen"], "default_ms": r["default_ms"], "tuned_ms": r["tuned_ms"],
                                     "sample": r["sample"], "trials": r["trials"]} for c, r in tune.items()}
                                if tune else "off (built-in defaults)")},
        "batched_vs_plain": batch["batched_vs_plain"],
        "plain": batch["plain"],
        "batched_vs_serial": batch["batched_vs_serial"],
        "roofline": roof,
        "pcie": pcie_line,
        "kernels": kernels,
        "batch24": batch,
        "autotune": tune,
        "analyze": analyze,
        "single_sample_hbm": single,
        "merge": merge,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
