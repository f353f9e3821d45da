"""
CPU tests of the launch-shape tuner's host side: the shape table (lookup by
mode and padded length, JSON round trip, rejection of invalid shapes), the
candidate list, and how KernelParams turn shapes, explicit or from an
activated table, into the C-ABI flags declared in include/d4amd.h.
"""

from pathlib import Path

import pytest

from tad_dftd4_amd import _lib, runtime, tune

ROOT = Path(__file__).resolve().parents[1]

TABLE = tune.ShapeTable(
    device="test",
    entries={"energy": [[64, 16, 16], [128, 32, 8]], "forces": [[200, 64, 16]]},
)


def _rows(flags):
    return flags & _lib.D4_FLAG_ROWS_MASK


def test_header_flag_values():
    hdr = (ROOT / "include" / "d4amd.h").read_text()
    for name in ("ROWS_16", "ROWS_32", "ROWS_64", "ROWS_MASK", "ATM8"):
        assert f"#define D4_FLAG_{name} {getattr(_lib, 'D4_FLAG_' + name)}" in hdr


def test_lookup_first_bucket_at_least_as_long():
    assert TABLE.lookup(False, 10) == (16, 16)
    assert TABLE.lookup(False, 64) == (16, 16)
    assert TABLE.lookup(False, 65) == (32, 8)
    assert TABLE.lookup(False, 129) == (0, 0)  # beyond the tuned range: built-in rules
    assert TABLE.lookup(True, 120) == (64, 16)
    assert TABLE.lookup(True, 201) == (0, 0)
    assert tune.ShapeTable().lookup(True, 50) == (0, 0)


def test_json_round_trip_and_validation(tmp_path):
    path = tmp_path / "shapes.json"
    TABLE.save(path)
    back = tune.ShapeTable.load(path)
    assert back.entries == TABLE.entries and back.device == "test"
    with pytest.raises(ValueError):
        tune.ShapeTable.from_json('{"entries": {"energy": [[64, 24, 8]]}}')
    with pytest.raises(ValueError):
        tune.ShapeTable.from_json('{"entries": {"energy": [[64, 16, 12]]}}')
    with pytest.raises(ValueError):
        tune.ShapeTable.from_json('{"entries": {"hessian": [[64, 16, 8]]}}')
    with pytest.raises(ValueError):
        tune.ShapeTable.from_json('{"entries": {"energy": [[0, 16, 8]]}}')


def test_argument_checks():
    with pytest.raises(TypeError):
        tune.activate({"energy": [[64, 16, 8]]})
    assert tune.active() is None
    with pytest.raises(ValueError):
        tune.tune("cpu", modes=("hessian",))


def test_shipped_table_is_valid():
    path = ROOT / "tad_dftd4_amd" / "shapes_gfx950.json"
    table = tune.ShapeTable.load(path)
    for mode in tune.MODES:
        ent = table.entries[mode]
        assert [n for n, _, _ in ent] == sorted(tune.BUCKETS)
        for n_max, rows, tile in ent:
            assert (rows, tile) in tune.candidates(n_max)
            assert table.timings[mode][str(n_max)][f"{rows}x{tile}"] == min(
                table.timings[mode][str(n_max)].values())


def test_candidates():
    c = tune.candidates(120)
    assert c[0] == (0, 0) and len(c) == 1 + len(tune.ROWS) * len(tune.TILES)
    assert set(tune.candidates(384)) == {(0, 0), (16, 0), (32, 0), (64, 0)}


def test_explicit_shapes_to_flags():
    kp = runtime.KernelParams()
    f = kp.c(100, True).flags
    assert _rows(f) == 0 and not f & _lib.D4_FLAG_ATM8 and not f & _lib.D4_FLAG_ATM16
    for rl, bits in ((16, _lib.D4_FLAG_ROWS_16), (32, _lib.D4_FLAG_ROWS_32),
                     (64, _lib.D4_FLAG_ROWS_64)):
        assert _rows(runtime.KernelParams(row_lanes=rl).c().flags) == bits
    assert runtime.KernelParams(atm_tile=8).c().flags & _lib.D4_FLAG_ATM8
    f16 = runtime.KernelParams(atm_tile=16).c().flags
    assert f16 & _lib.D4_FLAG_ATM16 and not f16 & _lib.D4_FLAG_ATM8
    with pytest.raises(ValueError):
        runtime.KernelParams(row_lanes=8).c()
    with pytest.raises(ValueError):
        runtime.KernelParams(atm_tile=4).c()


def test_activated_table_drives_default_shapes():
    tune.activate(TABLE)
    try:
        kp = runtime.KernelParams()
        f = kp.c(100, False).flags
        assert _rows(f) == _lib.D4_FLAG_ROWS_32 and f & _lib.D4_FLAG_ATM8
        f = kp.c(100, True).flags
        assert _rows(f) == _lib.D4_FLAG_ROWS_64 and f & _lib.D4_FLAG_ATM16
        assert _rows(kp.c(300, True).flags) == 0  # outside the table
        # explicit shapes win over the table
        f = runtime.KernelParams(row_lanes=16).c(100, True).flags
        assert _rows(f) == _lib.D4_FLAG_ROWS_16 and not f & _lib.D4_FLAG_ATM16
    finally:
        tune.activate(None)
    assert tune.active() is None
    assert _rows(runtime.KernelParams().c(100, False).flags) == 0
