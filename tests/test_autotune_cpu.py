"""Launch-shape table of the library (no GPU needed: the table lives on the host) and the
autotuner's table construction."""
import pytest

from grim_20261007_amd import autotune, native


@pytest.fixture
def lib():
    l = native.load()
    l.grim_clear_launch_shapes()
    yield l
    l.grim_clear_launch_shapes()


def test_table_lookup_ranges(lib):
    assert lib.grim_launch_shape(0, 2000) == 0                 # empty: built-in rule
    assert lib.grim_set_launch_shape(0, 0, 8) == 0
    assert lib.grim_set_launch_shape(0, 4000, 32) == 0
    assert lib.grim_set_launch_shape(0, 1000, 4) == 0           # inserted between
    assert [lib.grim_launch_shape(0, n) for n in (1, 999, 1000, 3999, 4000, 10 ** 6)] == [8, 8, 4, 4, 32, 32]
    assert lib.grim_set_launch_shape(0, 1000, 16) == 0          # replaced
    assert lib.grim_launch_shape(0, 2000) == 16
    assert lib.grim_set_launch_shape(0, 1000, 0) == 0           # removed
    assert lib.grim_launch_shape(0, 2000) == 8
    assert lib.grim_launch_shape(1, 2000) == 0                  # knobs are independent


def test_table_rejects_bad_arguments(lib):
    assert lib.grim_set_launch_shape(0, 0, 5) == 3              # not a far-pass shape
    assert lib.grim_set_launch_shape(0, 0, 64) == 3
    assert lib.grim_set_launch_shape(1, 0, 8) == 3              # not a triangular-solve shape
    assert lib.grim_set_launch_shape(2, 0, 8) == 3              # not a block-solve shape
    assert lib.grim_set_launch_shape(3, 0, 1) == 3              # unknown knob
    assert lib.grim_set_launch_shape(0, -1, 4) == 3
    for i in range(16):
        assert lib.grim_set_launch_shape(1, 100 * i, 1 + (i % 2)) == 0
    assert lib.grim_set_launch_shape(1, 5000, 4) == 1           # full
    assert lib.grim_set_launch_shape(1, 300, 4) == 0            # replacing still works
    assert lib.grim_clear_launch_shapes() == 0
    assert lib.grim_launch_shape(1, 300) == 0


def test_table_from_measurements_merges_and_splits():
    times = {500: {4: 1.0, 8: 2.0}, 2000: {4: 3.0, 8: 4.0}, 8000: {4: 9.0, 8: 7.0}}
    assert autotune._table_from(times) == [[0, 4], [4000, 8]]
    assert autotune._boundaries([500, 2000, 8000]) == [0, 1000, 4000]


def test_apply_loads_a_table(lib):
    autotune.apply({"far_columns": [[0, 4], [4000, 16]], "trsm_columns": [[0, 1], [3000, 2]],
                    "solve_columns": [[0, 2], [1000, 1]]})
    assert autotune.shape(0, 100) == 4 and autotune.shape(0, 5000) == 16
    assert autotune.shape(1, 2999) == 1 and autotune.shape(1, 3000) == 2
    assert autotune.shape(2, 999) == 2 and autotune.shape(2, 1000) == 1
    autotune.apply({"far_columns": [[0, 8]]})                   # a table replaces every knob
    assert autotune.shape(1, 3000) == 0 and autotune.shape(2, 0) == 0


def test_shipped_table_is_valid(lib):
    table = autotune.resolve("mi355x")
    for knob, name in autotune.KNOB_NAMES.items():
        entries = table[name]
        assert entries and entries[0][0] == 0
        assert [e[0] for e in entries] == sorted(e[0] for e in entries)
        assert all(v in autotune.CHOICES[knob] for _, v in entries)
    autotune.apply(table)
    for n in (1, 2000, 10 ** 6):
        want = [v for start, v in table["far_columns"] if start <= n][-1]
        assert autotune.shape(0, n) == want
