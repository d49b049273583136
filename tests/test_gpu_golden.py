"""GPU results on the full large workloads against the committed fixtures.

Bit-exact bar: every corank and linear complexity equals the CPU oracle's
(tests/golden/*.json, written by tests/golden/make_fixtures.py).
"""
import json
import os

import numpy as np
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
    """All 256 matrices of the small-prime batch (u32 entries generated on the
    device, the same as workloads.c5_rank), together with the planted-deficiency
    members of workloads.c5_planted (coranks 1 .. 1000, oracle-confirmed by
    tests/test_golden.py), in one batch."""
    fx = _load("c5_rank_coranks.json")
    buf, views = workloads.c5_rank_device(ctx)
    try:
        planted = workloads.c5_planted()
        print(f"c5 rank u32 x{len(views)} + {len(planted)} planted, method={method}", flush=True)
        ranks = ctx.rank_batch(modlin.Field(fx["p"]), views + planted, method=method)
        coranks = [2048 - r for r in ranks]
        assert coranks[:256] == fx["coranks"]
        assert coranks[256:] == list(workloads.C5_PLANTED_DEFICIENCY)
    finally:
        buf.free()


def test_c5_device_batch_equals_host(ctx):
    buf, views = workloads.c5_rank_device(ctx, count=3, n=64)
    try:
        host = workloads.c5_rank(count=3, n=64)
        for v, h in zip(views, host):
            assert np.array_equal(v.to_host().reshape(64, 64), h)
    finally:
        buf.free()
