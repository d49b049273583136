"""The modlin command-line tool: argument handling and output lines that need no
GPU (src/cli.rs:115-190, src/main.rs:19-60).

The README transcripts are replayed on the GPU by tests/test_gpu_cli.py.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "modlin-rs_amd", "modlin")
M61 = str((1 << 61) - 1)

pytestmark = pytest.mark.skipif(not os.path.exists(CLI), reason="modlin CLI not built")


def run(*args, timeout=60):
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=timeout)


@pytest.mark.parametrize("args,msg", [
    (["-R", "1", "-p", "7"], "SIDE/LENGTH must be at least 2"),
    (["-L", "0", "-p", "7"], "SIDE/LENGTH must be at least 2"),
    (["-R", "10", "0", "-p", "7"], "reps must be at least 1"),
    (["-R", "10", "-p", "1"], "modulus must be at least 2"),
    (["-R", "10", "-p", str(1 << 63)], "modulus must be below 2⁶³"),
    (["-R", "10", "-p", "9"], "modulus must be prime"),
    (["-R", "10", "-p", str((1 << 61) + 1)], "modulus must be prime"),
    (["-R", "10", "-p", "3215031751"], "modulus must be prime"),       # strong pseudoprime to bases 2, 3, 5, 7
    (["-R", "10"], "--modulus"),
    (["-p", "7"], "--rank"),
    (["-R", "10", "-L", "10", "-p", "7"], "cannot be used with"),
    (["-R", "x", "-p", "7"], "invalid value"),
])
def test_validation_exits_2(args, msg):
    r = run(*args)
    assert r.returncode == 2
    assert msg in r.stderr
    assert r.stdout == ""


def test_zero_seed_mixmax_exits_1():
    r = run("-G", "mixmax17", "-R", "10", "-p", "7")
    assert r.returncode == 1
    assert r.stdout.splitlines() == ["Generator: MIXMAX (TRandomMixMax17, N=17)", "Seed: 0x0000000000000000"]
    assert "Error initializing the generator: the zero seed is not allowed for MIXMAX" in r.stderr


@pytest.mark.parametrize("seed,shown", [("0x1_0", "0x0000000000000010"), ("0b101", "0x0000000000000005"),
                                        ("0O17", "0x000000000000000f"), ("1_000", "0x00000000000003e8"),
                                        ("18446744073709551615", "0xffffffffffffffff")])
def test_seed_parsing(seed, shown):
    # header lines come before any GPU work (main.rs:22-27, 72-78)
    r = run("-G", "xoroshiro128pp", "-R", "3", "2", "-p", "5", "-S", seed)
    lines = r.stdout.splitlines()
    assert lines[:3] == ["Generator: xoroshiro128++", f"Seed: {shown}",
                         "Running a modular rank test using 0.000 MiB of RAM: 2 3×3 matrices over the field of size 5"]


def test_memory_line_units():
    r = run("-G", "xoroshiro128pp", "-L", "100000000", "3", "-p", "5")
    assert r.stdout.splitlines()[2] == ("Running a modular linear-complexity test using 2.980 GiB of RAM: "
                                        "3 sequences of length 100000000 over the field of size 5")


def test_generator_names():
    for g, name in [("mixmax", "MIXMAX (TRandomMixMax, N=240, s=487013230256099140)"),
                    ("mixmax240", "MIXMAX (TRandomMixMax, N=240, s=487013230256099140)"),
                    ("mixmax17", "MIXMAX (TRandomMixMax17, N=17)"),
                    ("mixmax256", "MIXMAX (TRandomMixMax256, N=256, skip=2)")]:
        r = run("-G", g, "-R", "3", "-p", "5", "-S", "1")
        assert r.stdout.splitlines()[0] == f"Generator: {name}"
        r = run("-G", g, "--lux", "-R", "3", "-p", "5", "-S", "1")
        assert r.stdout.splitlines()[0] == f"Generator: {name}, lux: keep 5 drop 9"
