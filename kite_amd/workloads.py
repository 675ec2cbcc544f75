"""Seeded synthetic workloads for the count / merge / analyze benchmarks.

Each workload fixes the genome model, the read model and the count flags.
Reads are generated by libkite_synth.so; any read range is reproducible on
its own, so a bounded prefix of a workload is the start of the full file.

Where the stated genome model is ambiguous the interpretation is recorded on
the workload: ``k31-canon-50mbp`` places 1000 diverged copies (40 per
segment) of 25 random 5 kbp segments, i.e. 5 Mbp = 10% of the genome.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from ._native import SynthSpec, synth_genome, synth_reads


@dataclass(frozen=True)
class Workload:
    name: str
    k: int
    canonical: bool
    partitions: int
    genome_len: int
    n_reads: int
    read_len: int
    both_strands: bool
    err_rate: float
    n_rate: float
    seed: int
    min_count: int | None = None
    repeat_segments: int = 0
    repeat_seg_len: int = 0
    repeat_copies: int = 0
    repeat_div: float = 0.0
    variant_seed: int = 0
    snv_rate: float = 0.0
    genome_seed: int = 0  # 0: the genome uses `seed` too
    note: str = ""

    def spec(self) -> SynthSpec:
        return SynthSpec(self.seed, self.genome_len, self.read_len, int(self.both_strands),
                         self.err_rate, self.n_rate, self.repeat_segments, self.repeat_seg_len,
                         self.repeat_copies, 0, self.repeat_div, self.variant_seed, self.snv_rate)

    def genome(self) -> np.ndarray:
        g = self.spec()
        if self.genome_seed:
            g.seed = self.genome_seed
        return synth_genome(g)

    def fastq(self, n_reads: int | None = None, genome: np.ndarray | None = None,
              first_read: int = 0) -> np.ndarray:
        g = self.genome() if genome is None else genome
        n = self.n_reads if n_reads is None else n_reads
        return synth_reads(self.spec(), g, first_read, first_read + n)

    def expected_kmers(self) -> float:
        return self.n_reads * max(0, self.read_len - self.k + 1) * (1 - self.n_rate) ** self.k

    def scaled(self, n_reads: int, genome_len: int | None = None) -> "Workload":
        d = dict(self.__dict__)
        d["n_reads"] = n_reads
        if genome_len is not None:
            d["genome_len"] = genome_len
        d["name"] = f"{self.name}@{n_reads}"
        return Workload(**d)


K21_1MBP = Workload(
    name="k21-1mbp-200k", k=21, canonical=False, partitions=256, genome_len=1_000_000,
    n_reads=200_000, read_len=150, both_strands=False, err_rate=0.005, n_rate=0.0001, seed=1001)

K31_CANON_50MBP = Workload(
    name="k31-canon-50mbp-5m", k=31, canonical=True, partitions=256, genome_len=50_000_000,
    n_reads=5_000_000, read_len=150, both_strands=True, err_rate=0.01, n_rate=0.0001, seed=2002,
    min_count=2, repeat_segments=25, repeat_seg_len=5000, repeat_copies=1000, repeat_div=0.01,
    note="1000 copies (40 per segment) of 25 x 5 kbp segments = 10% of the genome")

K63_CANON_100MBP = Workload(
    name="k63-canon-100mbp-10m", k=63, canonical=True, partitions=1024, genome_len=100_000_000,
    n_reads=10_000_000, read_len=250, both_strands=True, err_rate=0.008, n_rate=0.0002, seed=3003)

BATCH_BASE_SEED = 4004


def batch24(n_samples: int = 24, scale: float = 1.0) -> list[Workload]:
    """24 samples: shared 20 Mbp base genome + per-sample 2% SNV layer,
    read counts log-uniform in [2e5, 8e6] (scaled by `scale`), 150 bp reads."""
    rng = np.random.default_rng(BATCH_BASE_SEED)
    counts = np.exp(rng.uniform(math.log(2e5), math.log(8e6), size=n_samples))
    out = []
    for i, c in enumerate(counts):
        out.append(Workload(
            name=f"batch24-s{i:02d}", k=31, canonical=True, partitions=256,
            genome_len=20_000_000, n_reads=max(1, int(c * scale)), read_len=150, both_strands=True,
            err_rate=0.01, n_rate=0.0001, seed=BATCH_BASE_SEED * 1000 + i, min_count=2,
            variant_seed=BATCH_BASE_SEED * 7919 + i, snv_rate=0.02, genome_seed=BATCH_BASE_SEED))
    return out


WORKLOADS = {w.name: w for w in (K21_1MBP, K31_CANON_50MBP, K63_CANON_100MBP)}
