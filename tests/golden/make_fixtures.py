"""Write the expected per-item results of the large workloads (run by hand).

The inputs are regenerated from their seeds by modlin-rs_amd/workloads.py; each
fixture stores a SHA-256 digest of every input next to the expected result so
that a change in a generator is caught before a result is compared.  Expected
results come from the CPU oracle (oracle/oracle.c, a restatement of src/fp.rs),
itself pinned against the reference's test vectors by tests/test_oracle_*.py.

    python tests/golden/make_fixtures.py            # all fixtures
    python tests/golden/make_fixtures.py c2 c4      # a subset

Fixtures:
  c2_coranks.json   batched rank, 64 matrices, p = 2^61-1: corank and the
                    algorithmic F_p multiply-accumulate count W of each matrix
  c4_lc.json        batched Berlekamp-Massey, 1024 sequences, p = 2^61-1:
                    linear complexity, algorithmic bytes and multiplications
  c5_rank_coranks.json  256 matrices 2048 x 2048, p = 65521, u32 storage
  c5_bm.json        one 200,000-element MIXMAX-256 sequence, p = 2^61-1
"""
import hashlib
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "modlin-rs_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

import oracle as O  # noqa: E402
import workloads as W  # noqa: E402

M61 = (1 << 61) - 1


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:16]


def write(name: str, obj: dict) -> None:
    path = os.path.join(HERE, name)
    with open(path, "w") as fh:
        json.dump(obj, fh, indent=None, separators=(",", ":"))
        fh.write("\n")
    print(f"wrote {path}", flush=True)


def c2() -> None:
    mats = W.c2()
    cor, work, dig = [], [], []
    for i, m in enumerate(mats):
        n = m.shape[0]
        t = time.time()
        r, _, w = O.rank(M61, m, n, with_trace=True)
        cor.append(n - r)
        work.append(w)
        dig.append(digest(m))
        print(f"c2 {i} n={n} corank={n - r} {time.time() - t:.1f}s", flush=True)
    write("c2_coranks.json", {"p": M61, "sides": [m.shape[0] for m in mats], "coranks": cor, "work": work,
                              "sha256_16": dig})


def _lc_one(args):
    p, s = args
    return O.lc(p, s, with_stats=True)


def c4() -> None:
    seqs, specs = W.c4()
    with ThreadPoolExecutor(max_workers=os.cpu_count()) as ex:
        res = list(ex.map(_lc_one, [(M61, s) for s in seqs]))
    write("c4_lc.json", {"p": M61, "specs": [list(x) for x in specs], "lc": [r[0] for r in res],
                         "bytes": [r[1]["bytes"] for r in res], "mulmods": [r[1]["mulmods"] for r in res],
                         "sha256_16": [digest(s) for s in seqs]})


def c5() -> None:
    s = W.c5_bm()
    t = time.time()
    lc, st = O.lc(M61, s, with_stats=True)
    print(f"c5 bm lc={lc} {time.time() - t:.1f}s", flush=True)
    write("c5_bm.json", {"p": M61, "n": int(s.size), "lc": lc, "bytes": st["bytes"], "mulmods": st["mulmods"],
                         "sha256_16": digest(s)})
    p = 65521
    cor, dig = [], []
    mats = W.c5_rank(count=256, n=2048, p=p)
    for i, m in enumerate(mats):
        r = O.rank(p, m.astype(np.uint64), 2048)
        cor.append(2048 - r)
        dig.append(digest(m))
        if i % 16 == 0:
            print(f"c5 rank {i} corank={2048 - r}", flush=True)
    write("c5_rank_coranks.json", {"p": p, "n": 2048, "coranks": cor, "sha256_16": dig})


if __name__ == "__main__":
    which = sys.argv[1:] or ["c2", "c4", "c5"]
    for w in which:
        globals()[w]()
