"""`kite merge` on the device (kite_gpu_merge_runs) against a Python
restatement over the independent .kite run codec (tests/kite_file.py):
per input, the runs of a partition summed per key (u64 saturating); the
input's header filter on each total; the inputs summed; one chunk per
non-empty partition (DVLE for 64- and 128-bit keys, RAW for 256-bit),
compared byte for byte with its rows and first/last keys.  Also: partition
groups of a few rows, empty inputs and partitions, the u32 overflow error,
and the CLI's merge of compact-mode (zstd) and filtered inputs against the
host engine and the oracle."""
import os
import random
import subprocess
from pathlib import Path

import pytest

import kite_file as kf
import oracle_lib as ol
from kite_amd import KiteError

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
KITE = ROOT / "kite_amd" / "lib" / "kite"
U64 = 2**64 - 1


def kite(*args, env_extra=None):
    env = dict(os.environ, **(env_extra or {}))
    r = subprocess.run([str(KITE), *map(str, args)], capture_output=True, text=True, timeout=600, env=env)
    if r.returncode != 0:
        raise AssertionError(f"kite {args} failed ({r.returncode}):\n{r.stderr}")
    return r


def make_input(rng, P, key_width, pool, n_per, max_count=50, big=False):
    """Stored runs of one input: every partition's keys (drawn from `pool`,
    so inputs overlap) in one to three sorted runs, a key sometimes split
    over two runs.  Returns (runs for the ABI, {partition: [(key, count)]}
    of the stored rows)."""
    runs, rows_of = [], {}
    for p in range(P):
        if rng.random() < 0.15:
            continue  # a partition without runs
        keys = sorted(set(rng.choice(pool[p]) for _ in range(rng.randrange(1, n_per))))
        nr = rng.randrange(1, 4)
        parts = [[] for _ in range(nr)]
        for k in keys:
            c = rng.randrange(1, max_count)
            if big and rng.random() < 0.1:
                c = 2**32 - 1 - rng.randrange(3)
            r = rng.randrange(nr)
            if nr > 1 and c >= 2 and rng.random() < 0.4:
                c1 = rng.randrange(1, c)
                parts[r].append((k, c1))
                parts[(r + 1) % nr].append((k, c - c1))
            else:
                parts[r].append((k, c))
        for rows in parts:
            if not rows:
                continue
            rows.sort()
            codec, data = kf.encode_run(rows, key_width)
            runs.append((p, codec, len(rows), data))
            rows_of.setdefault(p, []).append(rows)
        if rng.random() < 0.1:
            runs.append((p, 2 if key_width < 256 else 1, 0, b""))  # an empty run
    return runs, rows_of


def expected(inputs, P, key_width):
    out = []
    for p in range(P):
        acc = {}
        for rows_of, mn, mx in inputs:
            tot = {}
            for rows in rows_of.get(p, []):
                for k, c in rows:
                    tot[k] = min(tot.get(k, 0) + c, U64)
            for k, c in tot.items():
                if (mn is None or c >= mn) and (mx is None or c <= mx):
                    acc[k] = min(acc.get(k, 0) + c, U64)
        if acc:
            rows = sorted(acc.items())
            codec, data = kf.encode_run(rows, key_width)
            out.append((p, codec, len(rows), data, rows[0][0], rows[-1][0]))
    return out


def words_int(w) -> int:
    return sum(int(x) << (64 * i) for i, x in enumerate(w))


def check(got, exp):
    assert [g[0] for g in got] == [e[0] for e in exp]
    for g, e in zip(got, exp):
        part, codec, rows, raw, kmin, kmax, data = g
        assert (part, codec, rows) == e[:3]
        assert raw == len(data)
        assert data == e[3], f"partition {part}: payload differs"
        assert (words_int(kmin), words_int(kmax)) == (e[4], e[5])


@pytest.mark.parametrize("key_width,n_in", [(64, 2), (64, 4), (128, 3), (256, 2)])
def test_merge_runs_matches_python(gpu_ctx, key_width, n_in):
    rng = random.Random(key_width * 7 + n_in)
    P = 16
    bits = {64: 62, 128: 126, 256: 250}[key_width]
    pool = [[rng.getrandbits(bits) for _ in range(400)] for _ in range(P)]
    ins, abi = [], []
    for i in range(n_in):
        runs, rows_of = make_input(rng, P, key_width, pool, 300)
        mn = [None, 3, None, 2][i % 4]
        mx = [None, None, 60, 70][i % 4]
        ins.append((rows_of, mn, mx))
        abi.append((runs, mn, mx))
    got, st = gpu_ctx.merge_runs(key_width, P, abi)
    check(got, expected(ins, P, key_width))
    assert st["rows"] == sum(g[2] for g in got) and st["groups"] >= 1
    assert st["bytes"] == sum(len(g[6]) for g in got)


def test_merge_runs_groups_and_empty_inputs(gpu_ctx):
    """Partition groups of a few rows each (every group decoded, merged and
    encoded on its own), an input without runs, counts near u32::MAX whose
    per-input totals pass u32 and are filtered out by a max bound."""
    rng = random.Random(11)
    P, key_width = 32, 128
    pool = [[rng.getrandbits(120) for _ in range(300)] for _ in range(P)]
    r0, o0 = make_input(rng, P, key_width, pool, 200)
    r1, o1 = make_input(rng, P, key_width, pool, 200, big=True)
    ins = [(o0, None, None), (o1, None, 1000), ({}, None, None)]
    abi = [(r0, None, None), (r1, None, 1000), ([], None, None)]
    exp = expected(ins, P, key_width)
    os.environ["KITE_MERGE_GROUP_ROWS"] = "150"
    try:
        got, st = gpu_ctx.merge_runs(key_width, P, abi)
    finally:
        del os.environ["KITE_MERGE_GROUP_ROWS"]
    assert st["groups"] > 4
    check(got, exp)
    got1, st1 = gpu_ctx.merge_runs(key_width, P, abi)
    assert st1["groups"] == 1 and got1 == got


@pytest.mark.parametrize("key_width", [64, 128, 256])
def test_merge_runs_count_overflow(gpu_ctx, key_width):
    """A merged total above u32::MAX is the reference's error
    (serialize_partition_chunk), naming the key for 64/128-bit keys."""
    key = (1 << 40) + 12345 if key_width == 64 else (7 << 100) + 99
    if key_width == 256:
        key = (5 << 200) + 3
    ra = [(3, *kf.encode_run([(key - 1, 4), (key, 2**32 - 1)], key_width))]
    rb = [(3, *kf.encode_run([(key, 1)], key_width))]
    ra = [(p, codec, 2, data) for p, codec, data in ra]
    rb = [(p, codec, 1, data) for p, codec, data in rb]
    with pytest.raises(KiteError) as ei:
        gpu_ctx.merge_runs(key_width, 8, [(ra, None, None), (rb, None, None)])
    assert ei.value.code == -5
    msg = str(ei.value)
    if key_width == 256:
        assert "merged count exceeds u32" in msg
    else:
        assert f"merged count exceeds u32 for key {key}" in msg
    # the same inputs with a filter that drops the key merge fine
    got, _ = gpu_ctx.merge_runs(key_width, 8, [(ra, None, 10), (rb, None, None)])
    assert [(g[0], g[2]) for g in got] == [(3, 2)]


def _reads(n, L, seed, genome):
    rng = random.Random(seed)
    recs = []
    for i in range(n):
        s = rng.randrange(len(genome) - L)
        r = list(genome[s:s + L])
        for j in range(L):
            if rng.random() < 0.01:
                r[j] = rng.choice("ACGT")
        recs.append(f"@r{i}\n{''.join(r)}\n+\n{'I' * L}\n")
    return "".join(recs)


@pytest.mark.parametrize("k,canon,P", [(31, True, 64), (55, False, 16), (99, True, 8)])
def test_cli_merge_compact_and_filtered_inputs(tmp_path, k, canon, P):
    """Three inputs of one genome (most keys shared): one counted in compact
    mode (zstd chunks), one with --min 2, one with a header max; the device
    engine's merge equals the host engine's byte for byte, and its dump the
    dump of the oracle's merge of fast-mode counts of the same reads."""
    rng = random.Random(k)
    genome = "".join(rng.choice("ACGT") for _ in range(20000))
    fast, mixed = [], []
    for i in range(3):
        fq = tmp_path / f"in{i}.fq"
        fq.write_text(_reads(1500, 150, 100 * k + i, genome))
        flags = ["-k", k, "--canonical", str(canon).lower(), "--partitions", P]
        f = tmp_path / f"f{i}.kite"
        kite("count", "-i", fq, "-o", f, *flags)
        fast.append(f)
        m = tmp_path / f"m{i}.kite"
        extra = {0: ["--mode", "compact"], 1: ["--min", 2], 2: []}[i]
        kite("count", "-i", fq, "-o", m, *flags, *extra)
        mixed.append(m)
    kite("info", "-i", mixed[2], "--set-max", 40)
    kite("info", "-i", fast[1], "--set-min", 2)
    kite("info", "-i", fast[2], "--set-max", 40)
    kite("merge", "-i", *mixed, "-o", tmp_path / "g.kite")
    kite("merge", "-i", *mixed, "-o", tmp_path / "h.kite", "--engine", "cpu")
    assert (tmp_path / "g.kite").read_bytes() == (tmp_path / "h.kite").read_bytes()
    ol.kite_merge(fast, tmp_path / "o.kite", lanes=2)
    ol.kite_dump(tmp_path / "o.kite", tmp_path / "o.tsv")
    text = subprocess.run([str(KITE), "dump", "-i", str(tmp_path / "g.kite"), "-o", "-"], capture_output=True,
                          timeout=300, check=True).stdout
    assert text == (tmp_path / "o.tsv").read_bytes()
