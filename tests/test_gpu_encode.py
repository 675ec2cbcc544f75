"""Chunk payloads encoded on the device (kite_gpu_count_emit) against the
oracle's encoder, for every key width, and across several device batches."""
import random

import numpy as np
import pytest

import kite_file as kf
import oracle_lib as ol
from kite_amd import GpuContext

pytestmark = pytest.mark.gpu


def _reads(n, L, seed, alphabet="ACGTACGTACGTN"):
    rng = random.Random(seed)
    recs = []
    for i in range(n):
        s = "".join(rng.choice(alphabet) for _ in range(L))
        recs.append(f"@e{i}\n{s}\n+\n{'I' * L}\n")
    return "".join(recs).encode()


@pytest.mark.parametrize("k,canon,P", [(21, False, 16), (31, True, 64), (47, True, 8), (63, False, 32),
                                       (77, True, 4), (128, False, 2)])
def test_emitted_chunks_match_oracle_encoder(gpu_ctx, k, canon, P):
    fq = _reads(2500, 220, k)
    res, st = gpu_ctx.count(fq, k, canon, P)
    chunks = gpu_ctx.count_emit()
    W = res.words
    seen = set()
    for part, codec, rows, raw, kmin, kmax, data in chunks:
        a, b = int(res.part_off[part]), int(res.part_off[part + 1])
        assert rows == b - a and rows > 0
        assert codec == (1 if W == 4 else 2)
        assert raw == len(data)
        same = data == ol.dvle_encode(res.keys[a:b], res.counts[a:b], W)
        assert same, f"partition {part}"
        assert list(kmin[:W]) == [int(x) for x in res.keys[a]] and all(v == 0 for v in kmin[W:])
        assert list(kmax[:W]) == [int(x) for x in res.keys[b - 1]] and all(v == 0 for v in kmax[W:])
        seen.add(part)
    assert seen == {p for p in range(P) if res.part_off[p + 1] > res.part_off[p]}


@pytest.mark.parametrize("k,canon", [(25, True), (55, False), (90, True)])
def test_multi_batch_emission(k, canon):
    """A tiny device budget splits the input into several batches; each
    batch is encoded on the device and emitted; the decoded runs summed per
    partition equal the oracle's table."""
    fq = _reads(4000, 180, 1000 + k, alphabet="ACGT")
    arr = np.frombuffer(fq, dtype=np.uint8)
    chunks = ol.plain_chunks(arr, 48 << 10)
    P = 8
    with GpuContext(0, hbm_budget_bytes=4 << 20) as ctx:
        ctx.count_begin(k, canon, P)
        for off, ln in chunks:
            ctx.count_feed(arr[off:off + ln])
        st = ctx.count_finish()
        assert st.batches > 1
        emitted = ctx.count_emit()
    tab = ol.count_fastq(arr, k, canon, P, chunk_size=48 << 10, threads=4)
    W = ol.words_for(k)
    kw = {1: 64, 2: 128, 4: 256}[W]
    sums = [dict() for _ in range(P)]
    for part, codec, rows, raw, kmin, kmax, data in emitted:
        for key, c in kf.decode_run(codec, data, rows, kw):
            sums[part][key] = sums[part].get(key, 0) + c
    for p in range(P):
        a, b = int(tab.part_off[p]), int(tab.part_off[p + 1])
        exp = {ol.words_to_int(tab.keys[i]): int(tab.counts[i]) for i in range(a, b)}
        assert sums[p] == exp, f"partition {p}"
