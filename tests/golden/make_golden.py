"""Write the small golden fixtures under tests/golden/.

The expected outputs come from the brute-force restatement in tests/brute.py
(string k-mers, string reverse complements, collections.Counter), not from
the oracle, so the fixtures pin the oracle and the engine independently.

    python tests/golden/make_golden.py
"""
import json
import random
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
import brute  # noqa: E402

CASES = [
    # name, k, canonical, partitions
    ("k5_plain_p4", 5, False, 4),
    ("k9_canon_p8", 9, True, 8),
    ("k21_canon_p16", 21, True, 16),
    ("k33_canon_p4", 33, True, 4),
    ("k67_plain_p2", 67, False, 2),
]


def make_fastq(rng: random.Random, n: int) -> bytes:
    recs = []
    for i in range(n):
        L = rng.randint(0, 140)
        seq = "".join(rng.choice("ACGTACGTACGTNacgt") for _ in range(L))
        recs.append(f"@read{i} extra ACGT\n{seq}\n+\n{'I' * L}\n")
    return "".join(recs).encode()


def main() -> None:
    rng = random.Random(20261203)
    fq = make_fastq(rng, 120)
    (HERE / "small.fastq").write_bytes(fq)
    meta = {}
    for name, k, canon, P in CASES:
        out = brute.dump(fq, k, canon, P)
        (HERE / f"small_{name}.tsv").write_bytes(out)
        meta[name] = {"k": k, "canonical": canon, "partitions": P, "rows": out.count(b"\n")}
    (HERE / "small.json").write_text(json.dumps(meta, indent=1) + "\n")


if __name__ == "__main__":
    main()
