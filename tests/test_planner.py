"""The auto partition planner of `kite count` (derive_heuristic_plan,
src/count/mod.rs:624-864, helpers :866-1016) on the CPU.

`kite count --test-mode` prints the plan; KITE_COUNT_PLAN_ONLY=1 stops the
CLI there, before any device is opened.  Inputs are sparse files whose
first 40 MiB are real FASTQ (the planner samples at most 32 MiB / 200k
reads, src/count/mod.rs:1018-1079) and whose size is the workload's, so
the planner sees the configs' input sizes without writing them.

Every expected value below is worked out by hand from the reference's
formulas and constants (AUTO_MIN/MAX_PARTITIONS 256/2048, WORKER_SCRATCH
48 MiB, thread stack reserve 8 MiB, src/count/mod.rs:40-43), with the
arithmetic in the comments, not recomputed by a second implementation.
"""
import os
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
KITE = ROOT / "kite_amd" / "lib" / "kite"
GIB = 1 << 30


@pytest.fixture(scope="module", autouse=True)
def _need_binary():
    if not KITE.exists():
        pytest.skip("kite host binary not built")


def sparse_fastq(path: Path, read_len: int, size: int) -> Path:
    rec = ("@r\n" + "ACGT" * (read_len // 4) + "ACGT"[:read_len % 4] + "\n+\n" + "I" * read_len + "\n").encode()
    n = (40 << 20) // len(rec) + 1
    with open(path, "wb") as f:
        f.write(rec * n)
        f.truncate(size)
    return path


def plan(fq: Path, k: int, threads: int, memory: str, extra=()) -> dict:
    env = dict(os.environ, KITE_COUNT_PLAN_ONLY="1")
    r = subprocess.run([str(KITE), "count", "-i", str(fq), "-o", str(fq.parent / "o.kite"), "-k", str(k),
                        "--canonical", "true", "-t", str(threads), "--memory", memory, "--test-mode", *extra],
                       capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stderr
    m = re.search(r"\[auto\] P=(\d+), chunk=(\d+)MB, S=(\d+)MB", r.stderr)
    assert m, r.stderr
    rl = re.search(r"read_len=([\d.]+)bp", r.stderr)
    return {"P": int(m.group(1)), "chunk_mb": int(m.group(2)), "staging_mb": int(m.group(3)),
            "read_len": float(rl.group(1)), "stderr": r.stderr}


def test_config2_size_k31_gives_256(tmp_path):
    # 1,568,888,890 B of 150 bp reads, k=31 (u64, E=1, target 192 MiB):
    # expansion = 150/10 = 15; est = 1568888890 * 15 = 23,533,333,350;
    # ceil(est / 201,326,592) = 117 -> next power of two 128 -> clamped up to 256.
    # 16 threads, 64 GiB: 14 workers, chunk = 64 GiB / (14 * 96) = 48 MiB (floor),
    # staging 16 MiB fits (reserved 27.9 GiB, 2.58 GiB per worker >= 1250 MiB).
    p = plan(sparse_fastq(tmp_path / "c2.fq", 150, 1_568_888_890), 31, 16, "64G")
    assert p["read_len"] == 150.0
    assert (p["P"], p["chunk_mb"], p["staging_mb"]) == (256, 48, 16)


def test_config3_size_k63_gives_1024_at_64g(tmp_path):
    # 5,150,000,000 B of 250 bp reads, k=63 (u128, E=2, target 448-64 = 384 MiB):
    # expansion = min(25, 24) = 24; est = 5.15e9 * 24 * 2 = 247,200,000,000;
    # ceil(est / 402,653,184) = 614 -> 1024.
    # Cap at 64 GiB, 16 threads: (16 GiB) / (8 MiB + 14 * 5 KiB = 8,460,288) = 2030 -> 1024.
    fq = sparse_fastq(tmp_path / "c3.fq", 250, 5_150_000_000)
    p = plan(fq, 63, 16, "64G")
    assert (p["P"], p["chunk_mb"], p["staging_mb"]) == (1024, 48, 16)
    # At 32 GiB the cap is 8 GiB / 8,460,288 = 1015 -> 512 (floor power of two).
    # u128 workers need 2 GiB each; untracked = 32/6 GiB + 19 * 8 MiB + 32 MiB
    # = 5.51 GiB.  14 workers (chunk 24.4 MiB, 8 deep): staging 16/8/4/2 MiB
    # leave 1.26/1.55/1.69/1.76 GiB per worker; 13 workers at 2 MiB leave
    # 1.90 GiB; 12 workers: chunk = 32 GiB / 1152 = 28.4 MiB, staging
    # 16/8/4 MiB leave 1.48/1.81/1.98 GiB, 2 MiB leaves (32 - 0.222 - 1
    # - 0.5625 - 5.51) / 12 = 2.06 GiB.
    p = plan(fq, 63, 16, "32G")
    assert (p["P"], p["chunk_mb"], p["staging_mb"]) == (512, 28, 2)


def test_memory_cap_on_partitions(tmp_path):
    # 8 GiB of 150 bp reads at k=31: est = 8 GiB * 15 = 120 GiB -> 640 x 192 MiB -> 1024.
    fq = sparse_fastq(tmp_path / "big.fq", 150, 8 * GIB)
    p = plan(fq, 31, 16, "64G")
    assert p["P"] == 1024
    # 24 GiB, 10 threads (8 workers): (6 GiB) / (8 MiB + 40 KiB = 8,429,568) = 764
    # -> 512 caps the derived 1024.  chunk = 24 GiB / 768 = 32 MiB; staging
    # 16 MiB: reserved = 0.25 + 8 + 0.375 + 4.13 = 12.76 GiB, 1.40 GiB per worker.
    p = plan(fq, 31, 10, "24G")
    assert (p["P"], p["chunk_mb"], p["staging_mb"]) == (512, 32, 16)
    # 4 GiB, 4 threads (2 workers): (1 GiB) / 8,398,848 = 127 -> raised to 256.
    # With 2 workers no staging size leaves 1250 MiB per worker; with one
    # worker chunk = 4 GiB / 96 = 42.67 MiB and staging 4 MiB leaves
    # 4 - (0.33 + 1 + 0.047 + 1.33) = 1.29 GiB >= 1250 MiB (1.22 GiB).
    p = plan(fq, 31, 4, "4G")
    assert (p["P"], p["chunk_mb"], p["staging_mb"]) == (256, 42, 4)


def test_explicit_partitions_win(tmp_path):
    fq = sparse_fastq(tmp_path / "big.fq", 150, 8 * GIB)
    p = plan(fq, 31, 16, "64G", extra=("--partitions", "64"))
    assert p["P"] == 64
    assert "Using user-defined partitions=64" in p["stderr"]


def test_small_input_and_key_widths_floor_at_256(tmp_path):
    # a 40 MiB file: est is far below one target partition for every key width
    fq = sparse_fastq(tmp_path / "s.fq", 100, 40 << 20)
    for k in (21, 31, 45, 63, 101):
        assert plan(fq, k, 8, "64G")["P"] == 256, k
