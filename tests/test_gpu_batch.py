"""Batched multi-sample count (kite_gpu_count_batch) against single-sample
counts and the oracle: both modes must return every sample's table
bit-exactly, in sample order."""
import numpy as np
import pytest

import oracle_lib as ol
from kite_amd import BATCH_PIPELINED, BATCH_PLAIN, KiteError, PinnedArray
from kite_amd.workloads import batch24

pytestmark = pytest.mark.gpu


def _same(a, b, label):
    assert np.array_equal(a.part_off, b.part_off), label
    assert np.array_equal(a.keys, b.keys), label
    assert np.array_equal(a.counts, b.counts), label


def _samples(n_reads_scale=0.004):
    ws = batch24(8, scale=n_reads_scale)
    g = ws[0].genome()
    out = []
    for w in ws:
        fq = w.fastq(genome=g)
        pa = PinnedArray(fq.size)
        pa.array[:] = fq
        out.append(pa)
    return ws, out


def test_batch_modes_match_single_counts(gpu_ctx):
    ws, pins = _samples()
    arrs = [p.array for p in pins]
    # edge cases: an empty sample and one that does not start with '@'
    bad = np.frombuffer(b"xr\nACGTACGTACGTACGTACGTACGTACGTACGTAA\n+\n" + b"I" * 34 + b"\n", dtype=np.uint8)
    arrs = arrs[:3] + [np.zeros(0, dtype=np.uint8), bad] + arrs[3:]
    single = [gpu_ctx.count(a, 31, True, 256)[0] for a in arrs]
    for mode in (BATCH_PLAIN, BATCH_PIPELINED, BATCH_PIPELINED):
        res = gpu_ctx.count_batch(arrs, 31, True, 256, mode=mode)
        assert len(res) == len(arrs)
        for i, (r, st) in enumerate(res):
            _same(r, single[i], f"mode={mode} sample={i}")
            assert st["unique"] == len(single[i].counts)
            assert st["input_bytes"] == arrs[i].size
        assert res[4][1]["skipped_chunks"] == 1 and len(res[4][0].counts) == 0
        assert res[3][1]["kmers"] == 0
    for p in pins:
        p.free()


def test_batch_against_oracle(gpu_ctx):
    ws, pins = _samples(0.002)
    arrs = [p.array for p in pins[:4]]
    res = gpu_ctx.count_batch(arrs, 31, True, 256, mode=BATCH_PIPELINED)
    for i, a in enumerate(arrs):
        tab = ol.count_fastq(a, 31, True, 256, chunk_size=1 << 40, threads=8)
        r, st = res[i]
        assert st["kmers"] == tab.n_kmers
        assert np.array_equal(r.keys, tab.keys)
        assert np.array_equal(r.counts, np.minimum(tab.counts, 0xFFFFFFFF).astype(np.uint32))
    for p in pins:
        p.free()


def test_batch_growing_samples_and_other_widths(gpu_ctx):
    """Samples of increasing size force the persistent buffers to grow
    inside the pipeline; k = 63 exercises 2-word keys."""
    from kite_amd.workloads import K63_CANON_100MBP as w
    g = w.genome()
    arrs = [w.fastq(n_reads=n, genome=g, first_read=7 * n) for n in (2000, 9000, 30000, 500)]
    single = [gpu_ctx.count(a, 63, True, 64)[0] for a in arrs]
    res = gpu_ctx.count_batch(arrs, 63, True, 64, mode=BATCH_PIPELINED)
    for i in range(len(arrs)):
        _same(res[i][0], single[i], f"k63 sample={i}")


def test_batch_callback_abort(gpu_ctx):
    ws, pins = _samples(0.001)
    arrs = [p.array for p in pins[:3]]
    seen = []

    def cb(i, st, keys, counts, off):
        seen.append(i)
        if i == 1:
            raise RuntimeError("stop here")
        return int(counts.sum())

    with pytest.raises(RuntimeError, match="stop here"):
        gpu_ctx.count_batch(arrs, 31, True, 256, mode=BATCH_PIPELINED, on_result=cb)
    assert seen[:2] == [0, 1]
    # the context stays usable
    res = gpu_ctx.count_batch(arrs, 31, True, 256, mode=BATCH_PIPELINED)
    assert len(res) == 3
    with pytest.raises(KiteError):
        gpu_ctx.count_batch(arrs, 31, True, 300, mode=BATCH_PIPELINED)
    for p in pins:
        p.free()


@pytest.mark.parametrize("mode", [BATCH_PLAIN, BATCH_PIPELINED])
def test_batch_encoded_chunks(gpu_ctx, mode):
    """Chunk payloads of every sample, encoded on the device, equal the
    oracle encoder's bytes on the single-sample tables (k = 31 and 63)."""
    ws, pins = _samples(0.003)
    arrs = [p.array for p in pins[:5]] + [np.zeros(0, dtype=np.uint8)]
    for k, P in ((31, 256), (63, 16)):
        single = [gpu_ctx.count(a, k, True, P)[0] for a in arrs]
        res = gpu_ctx.count_batch(arrs, k, True, P, mode=mode, encoded=True)
        for i, (codec, data, boff, rows, mm, st) in enumerate(res):
            r = single[i]
            assert codec == 2
            assert np.array_equal(rows, r.part_off)
            assert st["unique"] == len(r.counts)
            for p in range(P):
                a, b = int(r.part_off[p]), int(r.part_off[p + 1])
                chunk = data[int(boff[p]):int(boff[p + 1])]
                if a == b:
                    assert len(chunk) == 0, f"k={k} sample={i} partition={p}"
                    continue
                # a boolean keeps pytest from diffing megabytes on failure
                same = chunk == ol.dvle_encode(r.keys[a:b], r.counts[a:b], r.words)
                assert same, f"k={k} sample={i} partition={p}"
                assert [int(x) for x in mm[p, :r.words]] == [int(x) for x in r.keys[a]]
                assert [int(x) for x in mm[p, 4:4 + r.words]] == [int(x) for x in r.keys[b - 1]]
    for p in pins:
        p.free()


@pytest.mark.parametrize("ring", ["2", "3", "4"])
def test_batch_ring_depths_and_serial_mode(gpu_ctx, monkeypatch, ring):
    """Every ring depth of the pipelined batch, with more samples than slots
    (each slot reused several times, sizes up and down), and the serial mode
    (persistent buffers, no overlap): raw tables and chunk payloads equal the
    single-sample counts."""
    from kite_amd import BATCH_SERIAL
    monkeypatch.setenv("KITE_BATCH_RING", ring)
    ws, pins = _samples(0.002)
    arrs = [p.array for p in pins] + [pins[0].array[:0], pins[5].array]
    single = [gpu_ctx.count(a, 31, True, 256)[0] for a in arrs]
    for mode in (BATCH_PIPELINED, BATCH_SERIAL):
        res = gpu_ctx.count_batch(arrs, 31, True, 256, mode=mode)
        for i, (r, st) in enumerate(res):
            _same(r, single[i], f"ring={ring} mode={mode} sample={i}")
        enc = gpu_ctx.count_batch(arrs, 31, True, 256, mode=mode, encoded=True)
        for i, (codec, data, boff, rows, mm, st) in enumerate(enc):
            r = single[i]
            assert np.array_equal(rows, r.part_off)
            for p in (0, 17, 255):
                a, b = int(r.part_off[p]), int(r.part_off[p + 1])
                same = data[int(boff[p]):int(boff[p + 1])] == ol.dvle_encode(r.keys[a:b], r.counts[a:b], 1)
                assert same, f"ring={ring} mode={mode} sample={i} partition={p}"
    for p in pins:
        p.free()


def test_batch_stage_timeline_order(gpu_ctx):
    """The per-sample stage events of a profiled pipelined batch follow the
    pipeline's dependencies: each count starts after its own H2D ended, each
    D2H after its count, and the D2H copies run in sample order."""
    ws, pins = _samples(0.002)
    arrs = [p.array for p in pins]
    gpu_ctx.count_profile(False, timeline=True)
    try:
        gpu_ctx.count_batch(arrs, 31, True, 256, mode=BATCH_PIPELINED, encoded=True)
        tl = gpu_ctx.batch_timeline()
    finally:
        gpu_ctx.count_profile(False)
    assert tl.shape == (len(arrs), 6)
    eps = 0.05  # ms: events of different streams around one dependency
    for i, (h0, h1, r0, r1, d0, d1) in enumerate(tl):
        assert 0 <= h0 <= h1 <= r0 + eps and r0 <= r1 and d0 <= d1, i
        assert d0 >= r1 - eps, i
        if i:
            assert d0 >= tl[i - 1][5] - eps and h0 >= tl[i - 1][1] - eps, i


def test_batch_graph_path_matches_classic(gpu_ctx, monkeypatch):
    """The pipelined batch of encoded chunks runs each sample's count and
    encode as one captured HIP graph launch; its chunks, row offsets, key
    ranges and counts equal the classic pipeline's (KITE_BATCH_GRAPH=0,
    kernels launched one by one after host read-backs), for one- and
    two-word keys, across repeated batches (the graphs are kept) and after
    the settings change (the graphs are captured again)."""
    ws, pins = _samples(0.003)
    arrs = [p.array for p in pins] + [np.zeros(0, dtype=np.uint8)]
    for k, P in ((31, 256), (45, 64)):
        monkeypatch.setenv("KITE_BATCH_GRAPH", "0")
        ref = gpu_ctx.count_batch(arrs, k, True, P, mode=BATCH_PIPELINED, encoded=True)
        monkeypatch.delenv("KITE_BATCH_GRAPH")
        for rep in range(3):
            if rep == 2:
                monkeypatch.setenv("KITE_SPLIT_DIV", "8")  # other settings: captured again
            got = gpu_ctx.count_batch(arrs, k, True, P, mode=BATCH_PIPELINED, encoded=True)
            for i, (g, r) in enumerate(zip(got, ref)):
                assert g[5]["graph"] == 1 and r[5]["graph"] == 0, (k, rep, i)
                assert g[0] == r[0] and g[1] == r[1], (k, rep, i)
                for q in (2, 3, 4):
                    assert np.array_equal(g[q], r[q]), (k, rep, i, q)
                for f in ("kmers", "unique", "input_bytes", "skipped_chunks"):
                    assert g[5][f] == r[5][f], (k, rep, i, f)
        monkeypatch.delenv("KITE_SPLIT_DIV")
    for p in pins:
        p.free()


def test_batch_graph_falls_back_past_its_capacity(monkeypatch):
    """A sample whose node list would exceed the planned count's capacity
    (lowered here to one node past the partitions) stops the graphs there:
    that sample and the ones after it run through the classic pipeline, in
    order, with the same results; the small samples before it (no partition
    large enough to split) ran as graphs."""
    from kite_amd import GpuContext
    ws, pins = _samples(0.01)
    big = sorted((p.array for p in pins), key=lambda a: a.size)[-2:]  # about 60 k reads: partitions split
    assert big[0].size > 10_000_000
    # the first 400 records: no partition large enough to split
    small = [a[:int(np.flatnonzero(a == 10)[4 * 400 - 1]) + 1] for a in (pins[0].array, pins[1].array, big[0])]
    arrs = small[:2] + [big[0], small[2], big[1]]
    monkeypatch.setenv("KITE_PLAN_NODES", "257")
    ctx = GpuContext(0)
    try:
        got = ctx.count_batch(arrs, 31, True, 256, mode=BATCH_PIPELINED, encoded=True)
        monkeypatch.setenv("KITE_BATCH_GRAPH", "0")
        ref = ctx.count_batch(arrs, 31, True, 256, mode=BATCH_PIPELINED, encoded=True)
    finally:
        ctx.close()
    assert [g[5]["graph"] for g in got] == [1, 1, 0, 0, 0]
    for i, (g, r) in enumerate(zip(got, ref)):
        assert g[1] == r[1] and np.array_equal(g[2], r[2]) and np.array_equal(g[3], r[3]), i
        assert g[5]["unique"] == r[5]["unique"] and g[5]["kmers"] == r[5]["kmers"], i
    for p in pins:
        p.free()


def test_batch_profiled_planned_launches(gpu_ctx):
    """With kernel profiling on, the batch makes the graphs' launches one by
    one between HIP events: the same chunks, and per-kernel device time with
    each kernel's algorithmic bytes."""
    ws, pins = _samples(0.002)
    arrs = [p.array for p in pins[:5]]
    ref = gpu_ctx.count_batch(arrs, 31, True, 256, mode=BATCH_PIPELINED, encoded=True)
    gpu_ctx.count_profile(True)
    try:
        got = gpu_ctx.count_batch(arrs, 31, True, 256, mode=BATCH_PIPELINED, encoded=True)
        ks = gpu_ctx.count_kernel_stats()
    finally:
        gpu_ctx.count_profile(False)
    for i, (g, r) in enumerate(zip(got, ref)):
        assert r[5]["graph"] == 1 and g[5]["graph"] == 0, i
        assert g[1] == r[1] and np.array_equal(g[2], r[2]) and np.array_equal(g[3], r[3]), i
    kmers = sum(r[5]["kmers"] for r in ref)
    unique = sum(r[5]["unique"] for r in ref)
    assert ks["k_bucket_sort"]["bytes"] == kmers * 8 + unique * 12
    for name in ("k_nl_count", "k_extract_hist", "k_extract_scatter", "k_split_hist", "k_split_scatter",
                 "k_bucket_sort", "k_compact"):
        assert ks[name]["ms"] > 0 and ks[name]["launches"] >= len(arrs), name
    for p in pins:
        p.free()


@pytest.mark.parametrize("ring", [2, 3])
def test_batch_graph_d2h_on_chosen_engines(gpu_ctx, monkeypatch, ring):
    """The graph batch queues each sample's payload D2H on an SDMA engine it
    picks (one engine bit per non-empty sample in the stats); the chunks
    equal those of the same batch with the engine left to the HIP runtime
    (KITE_D2H_ENGINES=0), over repeated batches and ring depths 2 and 3
    (the engine thread waits for the copy of sample i - R on the host)."""
    ws, pins = _samples(0.003)
    arrs = [p.array for p in pins] + [np.zeros(0, dtype=np.uint8)] + [p.array for p in pins[:3]]
    monkeypatch.setenv("KITE_BATCH_RING", str(ring))
    monkeypatch.setenv("KITE_D2H_ENGINES", "0")
    ref = gpu_ctx.count_batch(arrs, 31, True, 256, mode=BATCH_PIPELINED, encoded=True)
    monkeypatch.delenv("KITE_D2H_ENGINES")
    for rep in range(2):
        got = gpu_ctx.count_batch(arrs, 31, True, 256, mode=BATCH_PIPELINED, encoded=True)
        for i, (g, r) in enumerate(zip(got, ref)):
            assert g[5]["graph"] == 1 and r[5]["graph"] == 1, (rep, i)
            assert r[5]["d2h_engine"] == 0, (rep, i)
            e = g[5]["d2h_engine"]
            if g[1]:
                assert e != 0 and e & (e - 1) == 0, (rep, i, e)
            assert g[0] == r[0] and g[1] == r[1], (rep, i)
            for q in (2, 3, 4):
                assert np.array_equal(g[q], r[q]), (rep, i, q)
    for p in pins:
        p.free()
