"""Tuning-table selection (modlin_tuning_select), on the CPU.

A call takes the entry of its kind (rank or bm), field mode (m61, m31, gL) and
batch size (1 or many) whose size bucket is nearest in log2 to the call's
largest side; with no entry for its batch size it falls back to the other; with
no entry for its mode it gets zeros (the library's defaults).  Malformed tables
are rejected.
"""
import pytest

from conftest import modlin

M61 = (1 << 61) - 1
TABLE = """# kind mode side count far near gather every panel bm_coop bm_threads
rank m61 4096 many 4 2 1024 8 64 0 0
rank m61 8192 1 8 1 512 16 64 0 0
rank m61 16384 1 2 2 2048 8 48 0 0
rank g3 2048 many 6 2 1024 8 64 0 0 1 1   # 65521
rank m31 16384 1 4 2 1024 8 64 0 0 2 1
bm m61 262144 1 0 0 0 0 0 128 0
bm m61 16384 many 0 0 0 0 0 0 512
"""


@pytest.fixture
def table(tmp_path):
    p = tmp_path / "tuning.txt"
    p.write_text(TABLE)
    return str(p)


def test_exact_and_nearest(table):
    t = modlin.tuning_select(table, M61, 3781, 64)
    assert (t["far_per_cu"], t["near_per_cu"], t["check_every"]) == (4, 2, 8)
    t = modlin.tuning_select(table, M61, 8192, 1)
    assert (t["far_per_cu"], t["gather_wgs"], t["check_every"]) == (8, 512, 16)
    t = modlin.tuning_select(table, M61, 30000, 1)      # bucket 32768: nearest is 16384
    assert (t["far_per_cu"], t["panel"]) == (2, 48)
    t = modlin.tuning_select(table, M61, 5000, 1)       # bucket 8192
    assert t["far_per_cu"] == 8


def test_count_fallback_and_modes(table):
    t = modlin.tuning_select(table, M61, 8192, 4)       # no "many" entry near 8192: 4096 many
    assert t["far_per_cu"] == 4
    t = modlin.tuning_select(table, 65521, 2048, 256)   # 3 digits -> g3
    assert t["far_per_cu"] == 6
    t = modlin.tuning_select(table, (1 << 31) - 1, 300, 1)  # only a far m31 entry: nearest still
    assert t["far_per_cu"] == 4
    t = modlin.tuning_select(table, (1 << 31) - 5, 8192, 1)  # no g4 entry: defaults
    assert all(v == 0 for v in t.values())


def test_far_update_form_fields(table):
    """The two optional trailing fields: panels per far update (K depth) and the
    far update's quarters per B fragment (limb-class grouping)."""
    t = modlin.tuning_select(table, 65521, 2048, 256)
    assert (t["far_depth"], t["far_quarters"]) == (1, 1)
    t = modlin.tuning_select(table, (1 << 31) - 1, 16384, 1)
    assert (t["far_depth"], t["far_quarters"]) == (2, 1)
    t = modlin.tuning_select(table, M61, 4096, 64)      # seven-field entry: zeros (defaults)
    assert (t["far_depth"], t["far_quarters"]) == (0, 0)


def test_bm_entries(table):
    t = modlin.tuning_select(table, M61, 200000, 1, bm=True)
    assert t["bm_coop_wgs"] == 128
    assert modlin.tuning_select(table, M61, 200000, 1)["bm_coop_wgs"] == 0
    t = modlin.tuning_select(table, M61, 20000, 1024, bm=True)
    assert (t["bm_threads"], t["bm_coop_wgs"]) == (512, 0)


@pytest.mark.parametrize("bad", ["rank m61 4096 many 4 2 1024 8 64 0",        # missing field
                                 "rank m61 3000 1 4 2 1024 8 64 0 0",          # side not a power of two
                                 "rank m61 4096 1 4 2 1024 7 64 0 0",          # odd check_every
                                 "rank m61 4096 some 4 2 1024 8 64 0 0",       # count
                                 "bm m61 4096 many 0 0 0 0 0 0 384",           # bm_threads
                                 "rank m61 4096 1 4 2 1024 8 64 0 0 1",        # one optional field only
                                 "rank m61 4096 1 4 2 1024 8 64 0 0 3 1",      # far_depth
                                 "rank m61 4096 1 4 2 1024 8 64 0 0 1 3",      # far_quarters
                                 "rank m61 4096 1 4 2 1024 8 64 0 0 1 1 1",    # extra field
                                 "tile m61 4096 1 4 2 1024 8 64 0 0"])         # kind
def test_malformed_tables_rejected(tmp_path, bad):
    p = tmp_path / "bad.txt"
    p.write_text(bad + "\n")
    with pytest.raises(ValueError):
        modlin.tuning_select(str(p), M61, 4096, 1)
