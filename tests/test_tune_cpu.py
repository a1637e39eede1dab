"""CPU checks of the fused kernel's tile-shape table (host only, no device): the cache-file
format round trip, rejection of malformed entries, and clearing."""
import os

import pytest


@pytest.fixture()
def serving():
    from rbergomi_amd import serving
    serving.tune_clear()
    yield serving
    serving.tune_clear()


def test_cache_round_trip(serving, tmp_path):
    src = tmp_path / "tune.txt"
    src.write_text("# Np mode mc pb\n128 0 0 2\n96 1 0 4\n512 2 1 1\n")
    serving.tune_load(src)
    assert sorted(serving.tune_table()) == [(96, 1, 0, 4), (128, 0, 0, 2), (512, 2, 1, 1)]
    out = tmp_path / "saved.txt"
    serving.tune_save(out)
    serving.tune_clear()
    assert serving.tune_table() == []
    serving.tune_load(out)
    assert sorted(serving.tune_table()) == [(96, 1, 0, 4), (128, 0, 0, 2), (512, 2, 1, 1)]
    # a later file overrides entries with the same key
    src.write_text("128 0 0 1\n")
    serving.tune_load(src)
    assert (128, 0, 0, 1) in serving.tune_table()


@pytest.mark.parametrize("line", ["100 0 0 2", "128 3 0 2", "128 0 2 2", "128 0 0 3", "128 0 0"])
def test_malformed_entries_rejected(serving, tmp_path, line):
    import rbergomi_amd as rb
    f = tmp_path / "bad.txt"
    f.write_text("64 0 0 2\n" + line + "\n")
    with pytest.raises(rb.RbError) as e:
        serving.tune_load(f)
    assert "line 2" in str(e.value)
    assert serving.tune_table() == []  # nothing of a rejected file is kept


def test_missing_file(serving, tmp_path):
    import rbergomi_amd as rb
    with pytest.raises(rb.RbError):
        serving.tune_load(tmp_path / "absent.txt")
    assert not os.path.exists(tmp_path / "absent.txt")
