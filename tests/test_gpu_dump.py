"""`kite dump` on the device (kite_gpu_dump) against the oracle's dump and
the host engine: byte-identical text for every key width, for files with
many runs per partition (oracle counts with small chunks and several
threads), with header filters, u32 saturation, zstd-free multi-group runs,
and the errors malformed runs raise."""
import os
import random
import subprocess
from pathlib import Path

import numpy as np
import pytest

import kite_file as kf
import oracle_lib as ol
from kite_amd import KiteError

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
KITE = ROOT / "kite_amd" / "lib" / "kite"


def kite(*args, engine="gpu", env_extra=None):
    env = dict(os.environ, KITE_DUMP_ENGINE=engine, **(env_extra or {}))
    r = subprocess.run([str(KITE), *map(str, args)], capture_output=True, text=True, timeout=300, env=env)
    if r.returncode != 0:
        raise AssertionError(f"kite {args} failed ({r.returncode}):\n{r.stderr}")
    return r


def _reads(n, L, seed, alphabet="ACGTACGTACGTN"):
    rng = random.Random(seed)
    recs = []
    for i in range(n):
        s = "".join(rng.choice(alphabet) for _ in range(L))
        recs.append(f"@d{i}\n{s}\n+\n{'I' * L}\n")
    return "".join(recs).encode()


def _same_file(a: Path, b: Path, label: str):
    x, y = a.read_bytes(), b.read_bytes()
    same = x == y  # a boolean keeps pytest from diffing megabytes
    assert same, f"{label}: {len(x)} vs {len(y)} bytes"


def _split_runs(t, seed):
    """Every partition of an oracle table as three sorted runs; a key may
    sit in two runs with its count split, so the dump must sum them."""
    rng = random.Random(seed)
    runs = []
    for p in range(t.partitions):
        a, b = int(t.part_off[p]), int(t.part_off[p + 1])
        parts = [[], [], []]
        for i in range(a, b):
            v, c = ol.words_to_int(t.keys[i]), int(t.counts[i])
            r = rng.randrange(3)
            if c >= 2 and rng.random() < 0.5:
                c1 = rng.randrange(1, c)
                parts[r].append((v, c1))
                parts[(r + 1) % 3].append((v, c - c1))
            else:
                parts[r].append((v, c))
        for rows in parts:
            if rows:
                runs.append((p, sorted(rows)))
    rng.shuffle(runs)  # file order; the index groups them by partition
    return runs


@pytest.mark.parametrize("k,canon,P", [(21, False, 16), (31, True, 64), (45, True, 8), (63, False, 32),
                                       (77, True, 4), (128, False, 2)])
def test_dump_multirun_files_match_oracle(k, canon, P, tmp_path):
    fq = _reads(1500, 200, 7 * k + P, alphabet="ACGTACGTACGTACGTACGTN")
    t = ol.count_fastq(fq, k, canon, P, chunk_size=40 << 10, threads=4)
    kw = 64 * t.words
    f1 = tmp_path / "one.kite"
    ol.write_table(f1, t)
    ol.kite_dump(f1, tmp_path / "o1.tsv")
    kite("dump", "-i", f1, "-o", tmp_path / "g1.tsv")
    _same_file(tmp_path / "g1.tsv", tmp_path / "o1.tsv", f"one run k={k}")
    f = tmp_path / "multi.kite"
    kf.write(f, k, canon, P, kw, _split_runs(t, k))
    assert len(kf.read(f).chunks) > P
    ol.kite_dump(f, tmp_path / "o.tsv")
    r = kite("dump", "-i", f, "-o", tmp_path / "g.tsv")
    assert "engine: gpu" in r.stderr
    _same_file(tmp_path / "g.tsv", tmp_path / "o.tsv", f"runs k={k}")
    _same_file(tmp_path / "g.tsv", tmp_path / "o1.tsv", f"runs vs one run k={k}")
    kite("dump", "-i", f, "-o", tmp_path / "h.tsv", engine="host")
    _same_file(tmp_path / "h.tsv", tmp_path / "o.tsv", f"host k={k}")


def test_dump_filters_and_saturation(tmp_path):
    rng = random.Random(11)
    runs = []
    for _ in range(40):
        p = rng.randrange(8)
        keys = sorted(rng.sample(range(1 << 20), 60))
        runs.append((p, [(x, rng.choice([1, 2, 3, 9, 0x7FFFFFFF, 0xFFFFFFF0])) for x in keys]))
    for mn, mx in [(None, None), (2, None), (None, 9), (3, 0xFFFFFFFF), (0xFFFFFFFF, None)]:
        f = tmp_path / "m.kite"
        kf.write(f, 11, False, 8, 64, runs, min_count=mn, max_count=mx)
        ol.kite_dump(f, tmp_path / "o.tsv")
        kite("dump", "-i", f, "-o", tmp_path / "g.tsv")
        _same_file(tmp_path / "g.tsv", tmp_path / "o.tsv", f"filter {mn},{mx}")


@pytest.mark.parametrize("kw,k", [(128, 33), (128, 64), (256, 65), (256, 100), (256, 128)])
def test_dump_wide_keys_multirun(kw, k, tmp_path):
    rng = random.Random(kw * 1000 + k)
    runs = []
    for _ in range(12):
        keys = sorted({rng.randrange(1 << (2 * k)) for _ in range(80)})
        # overlapping runs: shared keys must sum
        keys = sorted(set(keys) | set(rng.sample(range(1 << 10), 10)))
        runs.append((rng.randrange(4), [(v, rng.randrange(1, 9)) for v in keys]))
    f = tmp_path / "w.kite"
    kf.write(f, k, True, 4, kw, runs)
    ol.kite_dump(f, tmp_path / "o.tsv")
    kite("dump", "-i", f, "-o", tmp_path / "g.tsv")
    _same_file(tmp_path / "g.tsv", tmp_path / "o.tsv", f"kw={kw} k={k}")


def test_dump_of_gpu_count_and_groups(tmp_path):
    """`kite count` on the GPU, then the dump in one group and in many small
    groups (group size forced by KITE_DUMP_GROUP_ROWS) equal the oracle's."""
    fq = _reads(6000, 150, 5, alphabet="ACGT")
    (tmp_path / "r.fq").write_bytes(fq)
    kite("count", "-i", tmp_path / "r.fq", "-o", tmp_path / "c.kite", "-k", 25, "--canonical", "true",
         "--partitions", 32)
    ol.kite_dump(tmp_path / "c.kite", tmp_path / "o.tsv")
    kite("dump", "-i", tmp_path / "c.kite", "-o", tmp_path / "g.tsv")
    _same_file(tmp_path / "g.tsv", tmp_path / "o.tsv", "one group")
    r = kite("dump", "-i", tmp_path / "c.kite", "-o", tmp_path / "g2.tsv", env_extra={"KITE_DUMP_GROUP_ROWS": "5000"})
    assert " group(s)" in r.stderr and " in 1 group(s)" not in r.stderr
    _same_file(tmp_path / "g2.tsv", tmp_path / "o.tsv", "many groups")
    out = kite("dump", "-i", tmp_path / "c.kite", "-o", "-").stdout.encode()
    assert out == (tmp_path / "o.tsv").read_bytes()


def _dvle(rows, words):
    keys = np.array([ol.key_to_words(v, words) for v, _ in rows], dtype=np.uint64).reshape(-1, words)
    counts = np.array([c for _, c in rows], dtype=np.uint32)
    return ol.dvle_encode(keys, counts, words)


def test_dump_abi_runs_and_errors(gpu_ctx):
    rows_a = [(3, 5), (10, 1), (1000, 2), (1 << 40, 7)]
    rows_b = [(10, 4), (11, 1), (1 << 40, 0xFFFFFFFF)]
    runs = [(0, 2, len(rows_a), _dvle(rows_a, 1)), (0, 2, len(rows_b), _dvle(rows_b, 1)),
            (2, 2, 1, _dvle([(77, 3)], 1))]
    text, st = gpu_ctx.dump_runs(21, 64, 4, runs)
    exp = b""
    tot = {}
    for v, c in rows_a + rows_b:
        tot[v] = min(tot.get(v, 0) + c, 0xFFFFFFFF)
    for v in sorted(tot):
        exp += f"{ol.decode_key(v, 1, 21)}\t{tot[v]}\n".encode()
    exp += f"{ol.decode_key(77, 1, 21)}\t3\n".encode()
    assert text == exp
    assert st["rows"] == len(tot) + 1 and st["lines"] == len(tot) + 1 and st["bytes"] == len(exp)
    # filters from the caller
    text, _ = gpu_ctx.dump_runs(21, 64, 4, runs, min_count=4, max_count=7)
    assert text == b"".join(l + b"\n" for l in exp.split(b"\n")[:-1] if 4 <= int(l.split(b"\t")[1]) <= 7)
    # nothing to dump
    assert gpu_ctx.dump_runs(21, 64, 4, [])[0] == b""
    # a payload cut before its last row
    bad = _dvle(rows_a, 1)
    with pytest.raises(KiteError, match="ends before its rows"):
        gpu_ctx.dump_runs(21, 64, 4, [(1, 2, len(rows_a), bad[:-2])])
    # a varint without its end byte
    over = bytes(_dvle(rows_a[:1], 1)) + b"\x80" * 11 + b"\x01\x01"
    with pytest.raises(KiteError, match="varint"):
        gpu_ctx.dump_runs(21, 64, 4, [(1, 2, 2, over)])
    # 256-bit runs must be ascending
    import struct
    raw = b"".join(struct.pack("<QQQQI", v, 0, 0, 0, 1) for v in (9, 4))
    with pytest.raises(KiteError, match="ascending"):
        gpu_ctx.dump_runs(70, 256, 2, [(0, 1, 2, raw)])
    # argument checks
    with pytest.raises(KiteError, match="non-decreasing"):
        gpu_ctx.dump_runs(21, 64, 4, runs[::-1])
    with pytest.raises(KiteError, match="zstd"):
        gpu_ctx.dump_runs(21, 64, 4, [(0, 6, 1, _dvle([(1, 1)], 1))])


def test_histo_matches_oracle(tmp_path):
    """`kite histo` on the device over several inputs (multi-run files,
    totals above the LDS bin range, header filters) equals the oracle's."""
    rng = random.Random(21)
    paths = []
    for fi, (k, kw, P, mn, mx) in enumerate([(21, 64, 8, None, None), (33, 128, 4, 2, None),
                                             (70, 256, 2, None, 500000), (15, 64, 16, 3, 9000)]):
        runs = []
        for _ in range(30):
            keys = sorted({rng.randrange(1 << min(2 * k, 40)) for _ in range(120)})
            runs.append((rng.randrange(P), [(v, rng.choice([1, 1, 2, 3, 7, 8191, 8192, 70000, 400000]))
                                            for v in keys]))
        f = tmp_path / f"in{fi}.kite"
        kf.write(f, k, True, P, kw, runs, min_count=mn, max_count=mx)
        paths.append(f)
    args = []
    for p in paths:
        args += ["-i", p]
    ol.kite_histo(paths, tmp_path / "o.tsv")
    env = dict(os.environ, KITE_HISTO_ENGINE="gpu")
    r = subprocess.run([str(KITE), "histo", *map(str, args), "-o", str(tmp_path / "g.tsv")], capture_output=True,
                       text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stderr
    assert "engine: gpu" in r.stderr
    _same_file(tmp_path / "g.tsv", tmp_path / "o.tsv", "histo")


def test_histo_abi(gpu_ctx):
    rows_a = [(3, 5), (10, 1), (1000, 2), (1 << 40, 9000)]
    rows_b = [(10, 4), (11, 1), (1 << 40, 0xFFFFFFFF)]
    runs = [(0, 2, len(rows_a), _dvle(rows_a, 1)), (0, 2, len(rows_b), _dvle(rows_b, 1)),
            (3, 2, 1, _dvle([(77, 5)], 1))]
    tot = {}
    for v, c in rows_a + rows_b + [(77, 5)]:
        tot[(v, v == 77)] = tot.get((v, v == 77), 0) + c
    exp = {}
    for c in tot.values():
        exp[c] = exp.get(c, 0) + 1
    bins, st = gpu_ctx.histo_runs(64, 4, runs)
    assert bins == exp  # includes 9000 + 0xFFFFFFFF, a u64 total
    assert st["lines"] == len(tot)
    bins, _ = gpu_ctx.histo_runs(64, 4, runs, min_count=5, max_count=5)
    assert bins == {5: 3}
    assert gpu_ctx.histo_runs(64, 4, [])[0] == {}
