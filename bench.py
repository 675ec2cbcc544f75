#!/usr/bin/env python3
"""Throughput of the gfx950 exact k-mer count on one MI355X.

One step = one pass of the count pipeline over the whole workload: FASTQ
bytes resident in HBM -> per-partition sorted, fully reduced (k-mer, count)
runs resident in HBM.  The default workload is the single-sample k=31
canonical count (50 Mbp genome with 10% diverged repeats, 5,000,000 reads
of 150 bp, 1% substitution errors, 0.01% N, 256 partitions).

Two more measurements ride on the same line:

* ``batch24``: 24 independent samples (20 Mbp base genome + 2% SNVs per
  sample, log-uniform 2e5..8e6 reads) counted from pinned host memory, once
  one at a time ("plain": per-sample device allocation, H2D, kernels, D2H,
  synchronise) and once through the pipelined batch (persistent buffers,
  copies overlapped with kernels).  Rate = total input k-mers / wall time,
  both including the H2D of the FASTQ and the D2H of the reduced tables.
* ``analyze``: Pearson + Spearman of the union count matrix of those 24
  tables (--min 2), split-bf16 MFMA Gram against the FP64 MFMA Gram, and
  the FP64 oracle on a row sample.

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


def cpu_baseline(w, n_reads: int, threads: int) -> dict:
    """Time the oracle (the C restatement of the reference `kite count`) on
    the first n_reads reads of the same workload, on this host's cores."""
    exe = ROOT / "oracle" / "build" / "kite_oracle"
    if not exe.exists():
        return {"value": None, "unit": "k-mers/s", "cores": 0, "kind": "port",
                "sample": "oracle binary missing"}
    fq = w.fastq(n_reads=n_reads)
    tmpdir = "/dev/shm" if os.path.isdir("/dev/shm") else tempfile.gettempdir()
    path = Path(tempfile.mkstemp(dir=tmpdir, suffix=".fq")[1])
    try:
        fq.tofile(path)
        del fq
        # chunk size as the reference planner picks it for a large budget: 64 MiB
        cmd = [str(exe), "count", str(path), "/dev/null", str(w.k), str(int(w.canonical)),
               str(w.partitions), str(64 << 20), str(threads)]
        best = None
        for _ in range(3):  # median of 3 in the reference plan; keep the bounded sample cheap
            r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
            if r.returncode != 0:
                raise RuntimeError(r.stderr)
            res = json.loads(r.stdout.strip().splitlines()[-1])
            best = res if best is None or res["count_s"] < best["count_s"] else best
        return {"value": best["kmers"] / best["count_s"], "unit": "k-mers/s", "cores": threads,
                "kind": "port",
                "sample": f"first {n_reads} reads of the workload ({best['kmers']} k-mers), "
                          f"best of 3, {threads} threads, output /dev/null"}
    finally:
        path.unlink(missing_ok=True)


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


def bench_batch24(ctx, scale: float, reps: int) -> tuple[dict, list]:
    """Plain vs pipelined count of the 24-sample cohort from pinned memory."""
    import ctypes as C

    from kite_amd import BATCH_PIPELINED, BATCH_PLAIN, PinnedArray, synth_reads
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
    arrs = [p.array for p in pins]
    total_bytes = sum(a.size for a in arrs)
    log(f"batch24: {len(arrs)} samples, {sum(w.n_reads for w in ws)} reads, {total_bytes / 1e9:.2f} GB "
        f"pinned FASTQ (generated in {time.time() - t0:.1f}s)")
    k, canon, P = ws[0].k, ws[0].canonical, ws[0].partitions

    # untimed: size the persistent buffers, keep the tables, and check that
    # both modes give identical tables (checksums per sample)
    def digest(i, st, keys, counts, off):
        return (st["kmers"], st["unique"], int(np.bitwise_xor.reduce(keys.ravel())) if keys.size else 0,
                int(counts.astype(np.uint64).sum()), int(off[-1]))

    tables = []

    def keep(i, st, keys, counts, off):
        tables.append((keys[:, 0].copy(), counts.copy(), off.copy()))
        return digest(i, st, keys, counts, off)

    d_pipe = ctx.count_batch(arrs, k, canon, P, mode=BATCH_PIPELINED, on_result=keep)
    d_plain = ctx.count_batch(arrs, k, canon, P, mode=BATCH_PLAIN, on_result=digest)
    assert d_pipe == d_plain, "plain and pipelined batch tables differ"
    kmers = sum(d[0] for d in d_pipe)

    def nothing(i, st, keys, counts, off):
        return 0

    # encoded mode: chunk payloads built on the device; the same bytes from both modes
    def cdigest(i, st, codec, data, boff, rows, mm):
        n8 = data.size // 8 * 8
        words = data[:n8].view(np.uint64)
        return (int(boff[-1]), int(np.bitwise_xor.reduce(words)) if words.size else 0,
                int(words.sum(dtype=np.uint64)) if words.size else 0, data[n8:].tobytes())

    e_pipe = ctx.count_batch(arrs, k, canon, P, mode=BATCH_PIPELINED, on_result=cdigest, encoded=True)
    e_plain = ctx.count_batch(arrs, k, canon, P, mode=BATCH_PLAIN, on_result=cdigest, encoded=True)
    assert e_pipe == e_plain, "plain and pipelined encoded chunks differ"
    enc_bytes = sum(d[0] for d in e_pipe)

    def nothing_enc(i, st, codec, data, boff, rows, mm):
        return 0

    res = {}
    for name, mode, enc in (("plain", BATCH_PLAIN, False), ("pipelined", BATCH_PIPELINED, False),
                            ("plain_encoded", BATCH_PLAIN, True), ("pipelined_encoded", BATCH_PIPELINED, True)):
        best = None
        for _ in range(reps):
            t = time.perf_counter()
            ctx.count_batch(arrs, k, canon, P, mode=mode, on_result=nothing_enc if enc else nothing, encoded=enc)
            dt = time.perf_counter() - t
            best = dt if best is None or dt < best else best
        res[name] = {"wall_s": best, "kmers_per_s": kmers / best}
        log(f"batch24 {name}: {best * 1e3:.1f} ms, {kmers / best / 1e9:.2f} G k-mers/s (best of {reps})")
    out = {
        "samples": len(arrs), "reads": int(sum(w.n_reads for w in ws)), "fastq_bytes": int(total_bytes),
        "input_kmers": int(kmers), "unique_kmers_total": int(sum(d[1] for d in d_pipe)),
        "plain": res["plain"], "pipelined": res["pipelined"],
        "speedup": res["plain"]["wall_s"] / res["pipelined"]["wall_s"],
        "includes": "H2D from pinned host memory, kernels, D2H of the reduced tables",
        "encoded": {"plain": res["plain_encoded"], "pipelined": res["pipelined_encoded"],
                    "speedup": res["plain_encoded"]["wall_s"] / res["pipelined_encoded"]["wall_s"],
                    "chunk_bytes": int(enc_bytes),
                    "includes": "H2D of the FASTQ, kernels, DVLE chunk encoding on the device, D2H of the chunks"},
        "bit_exact_plain_vs_pipelined": True, "scale": scale,
    }
    for p in pins:
        p.free()
    return out, tables


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


def bench_dump(w, fq) -> dict:
    """`kite dump` of the workload's .kite (written by `kite count` on the
    GPU): the device engine against the host engine (partition workers on
    this host's cores), wall time to /dev/null, and a hash of each engine's
    full text (bit-exact check at full size)."""
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
                           "device: decode + merge + format kernels")
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


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="k31-canon-50mbp-5m")
    ap.add_argument("--reads", type=int, default=None, help="override the read count")
    ap.add_argument("--cpu-sample-reads", type=int, default=1_000_000)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the 24-sample batch and analyze legs")
    ap.add_argument("--no-analyze", action="store_true")
    ap.add_argument("--no-dump", action="store_true", help="skip the kite dump leg")
    ap.add_argument("--no-bgzf", action="store_true", help="skip the BGZF inflate leg")
    ap.add_argument("--batch-scale", type=float, default=1.0, help="scale of the 24 read counts")
    ap.add_argument("--batch-reps", type=int, default=2)
    ap.add_argument("--analyze-cpu-parts", type=int, default=8,
                    help="partitions (of 256) whose union rows the CPU oracle sample covers")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run one 1 GiB device-to-device copy (a known-size kernel that "
                         "scales FETCH_SIZE / WRITE_SIZE counter readings)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(--gpus must be 1)", file=sys.stderr)
        return 2

    import torch

    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        return 1

    from kite_amd import GpuContext
    from kite_amd.workloads import WORKLOADS

    w = WORKLOADS[args.workload]
    if args.reads:
        w = w.scaled(args.reads)
    t0 = time.time()
    genome = w.genome()
    fq = w.fastq(genome=genome)
    log(f"workload {w.name}: {w.n_reads} reads, {fq.size / 1e9:.3f} GB FASTQ "
        f"(generated in {time.time() - t0:.1f}s)")
    dev = torch.from_numpy(fq).to("cuda")
    torch.cuda.synchronize()
    ptr, nbytes = dev.data_ptr(), dev.numel()

    if args.calibrate_copy:
        src = torch.empty(1 << 30, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        log("calibration copy: 1 GiB read + 1 GiB written")

    ctx = GpuContext(0)

    def step():
        ctx.count_begin(w.k, w.canonical, w.partitions)
        ctx.count_feed_device(ptr, nbytes)
        return ctx.count_finish()

    for i in range(args.warmup):
        st = step()
        log(f"warmup {i + 1}/{args.warmup}: {st.kmers} k-mers, {st.unique} unique, "
            f"device {st.device_ms:.2f} ms")

    # timed region: kernel events recorded on the engine's stream during the same steps
    ctx.count_profile(True)
    kern: dict[str, dict] = {}
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    stats = []
    for i in range(args.steps):
        st = step()
        stats.append(st)
        for name, v in ctx.count_kernel_stats().items():
            a = kern.setdefault(name, {"ms": 0.0, "launches": 0, "bytes": 0})
            for key in a:
                a[key] += v[key]
    torch.cuda.synchronize()
    t_end = time.perf_counter()
    ctx.count_profile(False)
    wall = t_end - t_start
    kmers = stats[-1].kmers
    assert all(s.kmers == kmers and s.unique == stats[0].unique for s in stats)
    ms_per_step = wall / args.steps * 1e3
    value = kmers * args.steps / wall
    log(f"timed {args.steps} steps: {ms_per_step:.3f} ms/step, {value / 1e9:.2f} G k-mers/s")

    dom_name, dom = max(kern.items(), key=lambda kv: kv[1]["ms"])
    per_launch_s = dom["ms"] / dom["launches"] * 1e-3
    per_launch_bytes = dom["bytes"] / dom["launches"]
    achieved = per_launch_bytes / per_launch_s / 1e9
    kernels = {n: {"ms_per_step": v["ms"] / args.steps,
                   "GBps": (v["bytes"] / (v["ms"] * 1e-3) / 1e9) if v["ms"] else None}
               for n, v in sorted(kern.items(), key=lambda kv: -kv[1]["ms"])}
    log("kernels: " + json.dumps(kernels))

    traffic, traffic_src = traffic_from_profile(dom_name, w.name)

    dump = None
    if not args.no_dump:
        dump = bench_dump(w, fq)

    bgz = None
    if not args.no_bgzf:
        bgz = bench_bgzf(ctx, w, fq, args.cpu_threads or min(16, os.cpu_count() or 1))
        if bgz["count_from_bgzf"]["kmers"] != kmers:
            raise SystemExit("bgzf count differs from the plain count")

    batch = analyze = None
    if not args.no_batch:
        del dev
        torch.cuda.empty_cache()
        batch, tables = bench_batch24(ctx, args.batch_scale, args.batch_reps)
        if not args.no_analyze:
            analyze = bench_analyze(ctx, tables, 2, args.analyze_cpu_parts)
        del tables

    cpu = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log(f"cpu baseline: oracle count on {args.cpu_sample_reads} reads, {threads} threads")
        cpu = cpu_baseline(w, min(args.cpu_sample_reads, w.n_reads), threads)
        log(f"cpu baseline: {cpu['value'] / 1e6 if cpu['value'] else 0:.1f} M k-mers/s")

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
        "dtype": "u64" if stats[-1].words == 1 else f"u64x{stats[-1].words}",
        "data": "synthetic (seeded generator, regenerated on the box)",
        "config": {"workload": w.name, "k": w.k, "canonical": w.canonical,
                   "partitions": w.partitions, "reads": w.n_reads, "read_len": w.read_len,
                   "fastq_bytes": int(nbytes), "input_kmers": int(kmers),
                   "unique_kmers": int(stats[-1].unique), "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom_name, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "traffic_source": traffic_src,
                     "algorithmic_bytes_per_launch": per_launch_bytes,
                     "avg_launch_ms": per_launch_s * 1e3},
        "kernels": kernels,
        "batch24": batch,
        "analyze": analyze,
        "dump": dump,
        "bgzf": bgz,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
