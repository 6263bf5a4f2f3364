"""The launch-shape tuning table: the library's reader (through the C ABI,
no device needed), the autotuner's resolution classes, and the committed
gfx950 table."""
import json

import pytest


def write(tmp_path, obj, name="t.json"):
    p = tmp_path / name
    p.write_text(obj if isinstance(obj, str) else json.dumps(obj))
    return p


TABLE = {"arch": "gfx950", "generated_by": "test", "entries": [
    {"class": "1080p", "n": 7, "mbs": 4, "waves": 6, "wgs_per_cu": 0, "ms": 0.06},
    {"class": "1080p", "n": 11, "mbs": 2, "waves": 5, "wgs_per_cu": 0, "ms": 0.1},
    {"class": "cif", "n": 5, "mbs": 8, "waves": 2, "wgs_per_cu": 0, "ms": 0.01},
    {"class": "batch", "n": 0, "mbs": 2, "waves": 4, "wgs_per_cu": 3, "ms": 1.5},
]}


def test_lookup_by_class_and_nearest_length(tmp_path):
    import vp8g
    p = write(tmp_path, TABLE)
    assert vp8g.tuning_lookup(p, 1920, 1080, 7) == {"mbs": 4, "waves": 6, "wgs_per_cu": 0}
    assert vp8g.tuning_lookup(p, 1920, 1080, 11) == {"mbs": 2, "waves": 5, "wgs_per_cu": 0}
    assert vp8g.tuning_lookup(p, 1920, 1080, 10)["waves"] == 5   # nearest n
    assert vp8g.tuning_lookup(p, 1920, 1080, 9)["waves"] == 6    # tie -> shorter
    assert vp8g.tuning_lookup(p, 1920, 1088, 3)["waves"] == 6    # still 1080p
    assert vp8g.tuning_lookup(p, 352, 288, 5) == {"mbs": 8, "waves": 2, "wgs_per_cu": 0}
    assert vp8g.tuning_lookup(p, 640, 360, 5) == {"mbs": 0, "waves": 0, "wgs_per_cu": 0}
    assert vp8g.tuning_lookup(p, 1, 1, 1, batch=True) == {"mbs": 2, "waves": 4, "wgs_per_cu": 3}


def test_classes_agree_between_library_and_autotuner(tmp_path):
    import vp8g
    from vp8g.autotune import resolution_class
    sizes = [(1, 1), (176, 144), (177, 144), (352, 288), (353, 288), (640, 360), (641, 360),
             (1280, 720), (1281, 720), (1920, 1088), (1920, 1089), (3840, 2160)]
    entries = [{"class": c, "n": 5, "mbs": 1 + i % 8, "waves": 1 + i, "wgs_per_cu": 0}
               for i, c in enumerate(("qcif", "cif", "360p", "720p", "1080p", "4k"))]
    p = write(tmp_path, {"entries": entries})
    by_class = {e["class"]: e["waves"] for e in entries}
    for w, h in sizes:
        assert vp8g.tuning_lookup(p, w, h, 5)["waves"] == by_class[resolution_class(w, h)], (w, h)


@pytest.mark.parametrize("bad", [
    '{"entries": [{"class": "cif", "n": 5, "waves": 2}]}',           # no mbs
    '{"entries": [{"class": "cif", "n": 5, "mbs": 9, "waves": 2}]}',  # mbs out of range
    '{"entries": [{"class": "cif", "n": 5, "mbs": 4, "waves": 17}]}',
    '{"entries": [{"class": "cif", "mbs": 4, "waves": 2, "wgs_per_cu": 40}]}',
    '{"entries": [{"class": "cif", "mbs": 4, "waves": 2',             # unterminated
])
def test_malformed_tables_are_rejected(tmp_path, bad):
    import vp8g
    with pytest.raises(vp8g.Vp8gError):
        vp8g.tuning_lookup(write(tmp_path, bad), 352, 288, 5)


def test_missing_file_is_rejected(tmp_path):
    import vp8g
    with pytest.raises(vp8g.Vp8gError):
        vp8g.tuning_lookup(tmp_path / "absent.json", 352, 288, 5)


def test_committed_table_is_readable():
    import vp8g
    if not vp8g.TUNING_PATH.exists():
        pytest.skip("no tuning table committed")
    table = json.loads(vp8g.TUNING_PATH.read_text())
    assert table["arch"] == "gfx950"
    for e in table["entries"]:
        got = vp8g.tuning_lookup(vp8g.TUNING_PATH, 1920, 1080, 7, batch=e["class"] == "batch")
        assert got["mbs"] in (0, 1, 2, 4, 8) and 0 <= got["waves"] <= 16
        assert 1 <= e["mbs"] <= 8 and 1 <= e["waves"] <= 16
