"""The batched engine's tuning table as the library reads it (tgp_batch_tuning;
no GPU needed): entries, mean-degree classes, the override variable and the
errors a malformed table gives."""
import pytest

import thegraphpartitioning_amd as tgp


def table(tmp_path, text):
    p = tmp_path / "tune.txt"
    p.write_text(text)
    return str(p)


def test_table_entries_and_classes(tmp_path, monkeypatch):
    monkeypatch.delenv("TGP_BATCH_SETTING", raising=False)
    monkeypatch.setenv("TGP_BATCH_TUNE", table(tmp_path, """# a comment
lanes 8
sort_tile 2048   # trailing comment
heavy_default 77

heavy 0 12 48
heavy 12 20 192
heavy 20 inf 384
"""))
    assert tgp.batch_tuning(3.0) == (8, 48, 2048)
    assert tgp.batch_tuning(11.99) == (8, 48, 2048)
    assert tgp.batch_tuning(12.0) == (8, 192, 2048)     # a class holds its lower bound
    assert tgp.batch_tuning(19.5) == (8, 192, 2048)
    assert tgp.batch_tuning(20.0) == (8, 384, 2048)
    assert tgp.batch_tuning(1e6) == (8, 384, 2048)


def test_first_matching_class_and_default(tmp_path, monkeypatch):
    monkeypatch.delenv("TGP_BATCH_SETTING", raising=False)
    monkeypatch.setenv("TGP_BATCH_TUNE", table(tmp_path, """heavy 5 10 40
heavy 0 100 50
heavy_default 60
"""))
    assert tgp.batch_tuning(7.0) == (4, 40, 4096)       # lanes and tile built in
    assert tgp.batch_tuning(2.0) == (4, 50, 4096)
    assert tgp.batch_tuning(150.0) == (4, 60, 4096)


def test_override_variable(tmp_path, monkeypatch):
    monkeypatch.setenv("TGP_BATCH_TUNE", table(tmp_path, "lanes 8\nheavy 0 inf 48\n"))
    monkeypatch.setenv("TGP_BATCH_SETTING", "2,300,8192")
    assert tgp.batch_tuning(10.0) == (2, 300, 8192)
    for bad in ["3,96,4096", "4,96,1000", "4,-1,4096", "4,96", "4,96,4096x"]:
        monkeypatch.setenv("TGP_BATCH_SETTING", bad)
        with pytest.raises(tgp.TGPError, match="TGP_BATCH_SETTING"):
            tgp.batch_tuning(10.0)


@pytest.mark.parametrize("text,line,what", [
    ("lanes 3\n", 1, "lanes takes"),
    ("lanes 4\nsort_tile 1024\n", 2, "sort_tile takes"),
    ("heavy 12 12 96\n", 1, "heavy takes"),
    ("heavy 0 12\n", 1, "heavy takes"),
    ("# ok\nheavy_default -5\n", 2, "heavy_default takes"),
    ("lanes 4 4\n", 1, "lanes takes"),
    ("\nbogus 1\n", 2, "unknown entry 'bogus'"),
])
def test_malformed_table_names_file_and_line(tmp_path, monkeypatch, text, line, what):
    monkeypatch.delenv("TGP_BATCH_SETTING", raising=False)
    p = table(tmp_path, text)
    monkeypatch.setenv("TGP_BATCH_TUNE", p)
    with pytest.raises(tgp.TGPError) as e:
        tgp.batch_tuning(10.0)
    assert p in str(e.value) and f"line {line}" in str(e.value) and what in str(e.value)


def test_missing_table_is_an_error(tmp_path, monkeypatch):
    monkeypatch.delenv("TGP_BATCH_SETTING", raising=False)
    monkeypatch.setenv("TGP_BATCH_TUNE", str(tmp_path / "absent.txt"))
    with pytest.raises(tgp.TGPError, match="cannot open the batch tuning table"):
        tgp.batch_tuning(10.0)


def test_shipped_table_parses(monkeypatch):
    """the table beside the library (written by the tuner on an MI355X), if
    present, is well formed and its settings are ones the engine is built for"""
    monkeypatch.delenv("TGP_BATCH_TUNE", raising=False)
    monkeypatch.delenv("TGP_BATCH_SETTING", raising=False)
    for d in (1.0, 10.0, 15.0, 25.0, 100.0):
        lanes, heavy, tile = tgp.batch_tuning(d)
        assert lanes in (2, 4, 8) and tile in (2048, 4096, 8192) and heavy >= 0
