"""The modlin command-line tool on the GPU: the README transcripts, stdout byte
for byte (tests/golden/readme_transcripts.json), and a multi-rep run against the
oracle fed with the same generator stretches (main.rs:82-105)."""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import modlin, oracle

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "modlin-rs_amd", "modlin")
with open(os.path.join(ROOT, "tests", "golden", "readme_transcripts.json"), encoding="utf-8") as fh:
    RUNS = json.load(fh)["runs"]


@pytest.mark.parametrize("run", RUNS, ids=[r["readme_lines"] for r in RUNS])
def test_readme_transcript(run):
    print("cli", run["args"], flush=True)
    r = subprocess.run([CLI, *run["args"]], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.splitlines() == run["stdout"]


def test_reps_use_disjoint_stretches():
    p, n, reps = 65521, 70, 3
    r = subprocess.run([CLI, "-G", "mixmax17", "-R", str(n), str(reps), "-p", str(p), "-S", "5"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    data = modlin.prng_fill(modlin.PRNG_MIXMAX17, 5, reps * n * n) % np.uint64(p)
    want = []
    for i in range(reps):
        corank = n - oracle.rank(p, data[i * n * n:(i + 1) * n * n], n)
        want.append(f"Matrix {i + 1}/{reps}\tcorank={corank}\tp={oracle.pretty_p_value(oracle.corank_tail_pvalue(p, n, corank))}")
    assert r.stdout.splitlines()[3:] == want


@pytest.mark.parametrize("args", [
    ["-G", "mixmax", "--lux", "-R", "300", "4", "-p", "1000000007", "-S", "9"],
    ["-G", "xoroshiro128pp", "-R", "257", "3", "-p", "65521", "-S", "0x1f"],
    ["-G", "mixmax256", "-L", "3000", "5", "-p", "2305843009213693951", "-S", "2"],
    ["-G", "mixmax17", "-R", "600", "2", "-p", "2147483647", "-S", "1", "--batch-gib", "0.001"],
])
def test_device_generation_matches_host(args):
    # the default device fill and --host-gen must print the same lines
    print("cli", args, flush=True)
    a = subprocess.run([CLI, *args], capture_output=True, text=True, timeout=120)
    b = subprocess.run([CLI, *args, "--host-gen"], capture_output=True, text=True, timeout=120)
    assert a.returncode == 0 and b.returncode == 0, (a.stderr, b.stderr)
    assert a.stdout == b.stdout
