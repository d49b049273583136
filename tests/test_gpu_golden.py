"""GPU results on the full large workloads against the committed fixtures.

Bit-exact bar: every corank and linear complexity equals the CPU oracle's
(tests/golden/*.json, written by tests/golden/make_fixtures.py).
"""
import json
import os

import pytest

from conftest import modlin, workloads

pytestmark = pytest.mark.gpu
HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
M61 = (1 << 61) - 1


def _load(name):
    with open(os.path.join(HERE, name)) as fh:
        return json.load(fh)


@pytest.mark.parametrize("method", [modlin.METHOD_PLAIN, modlin.METHOD_MFMA])
def test_c2_batch(ctx, method):
    fx = _load("c2_coranks.json")
    mats = workloads.c2()
    print(f"c2 batch method={method}", flush=True)
    ranks = ctx.rank_batch(modlin.Field(M61), mats, method=method)
    assert [m.shape[0] - r for m, r in zip(mats, ranks)] == fx["coranks"]


def test_c4_batch(ctx):
    fx = _load("c4_lc.json")
    seqs, _ = workloads.c4()
    print("c4 batch", flush=True)
    assert ctx.lc_batch(modlin.Field(M61), seqs) == fx["lc"]


@pytest.mark.parametrize("method", [modlin.METHOD_PLAIN, modlin.METHOD_MFMA])
def test_c5_rank_batch_u32(ctx, method):
    fx = _load("c5_rank_coranks.json")
    mats = workloads.c5_rank(count=64)
    print(f"c5 rank u32 x64 method={method}", flush=True)
    ranks = ctx.rank_batch(modlin.Field(fx["p"]), mats, method=method)
    assert [2048 - r for r in ranks] == fx["coranks"][:64]
