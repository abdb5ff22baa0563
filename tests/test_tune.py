"""The LDE tile-shape table (host-side, no GPU needed): export lists every
tunable entry, import applies valid text atomically and rejects malformed or
out-of-range lines without changing anything, set/get agree."""
import pytest

import pico_amd
from pico_amd._lib import PgError


@pytest.fixture
def table():
    saved = pico_amd.tune_export()
    yield saved
    pico_amd.tune_import(saved)


def _entries(text):
    return [tuple(int(x) for x in line.split()) for line in text.splitlines() if line and not line.startswith("#")]


def test_export_lists_every_tunable_entry(table):
    rows = _entries(table)
    # passes with a second shape: log sizes 7..11, 4 kinds, 2 fields, 4 width buckets
    assert len(rows) == 2 * 4 * 5 * 4
    assert {r[2] for r in rows} == {7, 8, 9, 10, 11}
    assert all(r[4] in (0, 1) for r in rows)


def test_import_set_get(table):
    pico_amd.tune_import("# comment\n1 2 10 3 1\n0 1 8 0 1   # trailing comment\n\n")
    assert pico_amd.tune_get(1, 2, 10, 300) == 1
    assert pico_amd.tune_get(1, 2, 10, 100) == 0  # another width bucket
    assert pico_amd.tune_get(0, 1, 8, 8) == 1
    pico_amd.tune_set(0, 1, 8, 0, 0)
    assert pico_amd.tune_get(0, 1, 8, 5) == 0
    rows = _entries(pico_amd.tune_export())
    assert (1, 2, 10, 3, 1) in rows


@pytest.mark.parametrize("bad", ["0 1 8 0 2\n", "2 1 8 0 1\n", "0 1 5 0 1\n", "0 4 8 0 0\n", "0 1 8 4 0\n",
                                 "0 1 8\n", "0 1 8 0 1 7\n", "x y z w v\n"])
def test_import_rejects_bad_lines_atomically(table, bad):
    before = pico_amd.tune_export()
    with pytest.raises(PgError):
        pico_amd.tune_import("1 2 10 3 1\n" + bad)  # the valid first line must not be applied either
    assert pico_amd.tune_export() == before
    with pytest.raises(PgError):
        pico_amd.tune_set(0, 1, 5, 0, 1)  # 2^5 has one shape only
