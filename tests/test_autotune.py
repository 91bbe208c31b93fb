"""Launch-variant table and tile colouring, on the CPU.

The variant calls are host-only (they set the table the launches read), so they run
without a GPU.  The colouring rule of the 3D tiled deposit (colors_for / tile_of_color in
mpf_binned.hip) is checked by brute force: on every periodic axis of up to nine tiles, for
every tile length, last-tile length and stencil, tiles of one colour must deposit to
disjoint node sets, so that each node receives at most one addition per launch.
"""

from __future__ import annotations

import ctypes
import json

import pytest

from mpifield_amd import _lib, autotune


def test_variant_shapes_and_table_round_trip(tmp_path):
    lib = _lib.load()
    shapes = autotune.variant_shapes(lib)
    assert shapes == [((8, 8, 16), 256), ((8, 8, 32), 256), ((8, 8, 16), 512), ((8, 8, 32), 512)]
    table = {"tiled": [{"op": "scatter", "kind": 1, "degree": 2, "ncomp_group": 3, "variant": 3},
                       {"op": "gather", "kind": 2, "degree": 3, "ncomp_group": 1, "variant": 2}]}
    old = [lib.mpf_tiled_variant(0, 1, 2, 3, -1), lib.mpf_tiled_variant(1, 2, 3, 1, -1)]
    try:
        assert autotune.apply_table(lib, table) == 2
        assert lib.mpf_tiled_variant(0, 1, 2, 3, -1) == 3
        assert lib.mpf_tiled_variant(1, 2, 3, 1, -1) == 2
        # the degree is ignored for linear and quadratic (one stencil each)
        assert lib.mpf_tiled_variant(0, 1, 0, 3, -1) == 3
    finally:
        lib.mpf_tiled_variant(0, 1, 2, 3, old[0])
        lib.mpf_tiled_variant(1, 2, 3, 1, old[1])


@pytest.mark.parametrize("args", [(2, 1, 2, 3, 0), (0, 3, 2, 3, 0), (0, 1, 2, 2, 0), (0, 1, 2, 3, 4)])
def test_variant_arguments_are_checked(args):
    lib = _lib.load()
    assert lib.mpf_tiled_variant(*args) < 0
    assert b"tiled_variant" in _lib.load().mpf_last_error()
    tile, block = (ctypes.c_int * 3)(), ctypes.c_int()
    assert lib.mpf_tiled_variant_shape(4, tile, ctypes.byref(block)) != 0


def test_committed_table_names_known_shapes():
    if not autotune.TABLE.exists():
        pytest.skip("no tuned table committed")
    table = json.loads(autotune.TABLE.read_text())
    shapes = autotune.variant_shapes(_lib.load())
    for e in table["tiled"]:
        assert e["op"] in autotune.OPS and 0 <= e["variant"] < len(shapes)
        assert tuple(e["tile"]) == shapes[e["variant"]][0] and e["block"] == shapes[e["variant"]][1]
    assert 0 < table["binned"]["binned_min"] < 1 << 31


def colouring(nt: int, num: int, tile: int, span: int) -> list[int]:
    """The colour of each tile along a periodic axis, as the deposit launches them: two
    alternating colours, or three (tile 0 alone, tiles 1 .. nt - 1 alternating) for an odd
    tile count or a last tile of at most span - 2 cells."""
    last = num - (nt - 1) * tile
    if nt % 2 == 1 or last <= span - 2:
        return [2] + [(k - 1) % 2 for k in range(1, nt)]
    return [k % 2 for k in range(nt)]


@pytest.mark.parametrize("tile", [8, 16, 32])
@pytest.mark.parametrize("span", [1, 2, 3, 4])
def test_same_colour_tiles_deposit_to_disjoint_nodes(tile, span):
    for nt in range(2, 10):
        for last in range(1, tile + 1):
            num = (nt - 1) * tile + last
            if num < tile + span - 1:  # the tiles need a periodic axis a footprint longer than a tile
                continue
            colours = colouring(nt, num, tile, span)
            for off in range(span):
                nodes = []
                for k in range(nt):
                    cells = range(k * tile, min((k + 1) * tile, num))
                    nodes.append({(c - off + s) % num for c in cells for s in range(span)})
                for a in range(nt):
                    for b in range(a + 1, nt):
                        if colours[a] == colours[b]:
                            assert not nodes[a] & nodes[b], (nt, num, tile, span, off, a, b)


def test_two_colours_alone_would_race_on_a_short_last_tile():
    """The case three colours exist for: an even tile count and a one-cell last tile, where
    tiles 0 and nt - 2 (both colour 0 of an alternation) meet below the seam."""
    tile, span, nt = 16, 3, 4
    num = (nt - 1) * tile + 1
    nodes = [{(c - 1 + s) % num for c in range(k * tile, min((k + 1) * tile, num)) for s in range(span)}
             for k in range(nt)]
    assert nodes[0] & nodes[2]
    assert colouring(nt, num, tile, span)[0] != colouring(nt, num, tile, span)[2]
