"""Autotuning of the tiled kernels: the launch variant of the float32 3D deposit and gather
per (kind, degree, components per group), and the batch size from which mpf_scatter /
mpf_gather bin inside the call.

A launch variant (mpf_tiled_variant_shape) is a workgroup tile of 8 x 8 x 16 or 8 x 8 x 32
cells (one or two of the bins' tiles along z; the bins are the same) and a workgroup of 256
or 512 work-items.  The results are the same in every variant, bit for bit; only the speed
differs with the stencil, the component count and the grid.  The reference has one fixed
thread policy per kernel (scatter_quadratic_3d_pyccel.py:22-25, 151-157); this replaces it
with measured choices.

    python -m mpifield_amd.autotune            # on the MI355X: tune the standard shapes and
                                              # write tuned_tiled.json next to this file

``tune_tiled`` tunes one shape and sets its picks in the library; ``apply_table`` sets the
picks of a table (done when the library loads, from tuned_tiled.json if it exists).
"""

from __future__ import annotations

import ctypes
import json
import sys
from pathlib import Path

TABLE = Path(__file__).resolve().parent / "tuned_tiled.json"
OPS = {"scatter": 0, "gather": 1}


def variant_shapes(lib) -> list[tuple[tuple[int, int, int], int]]:
    """(workgroup tile cells, workgroup size) of every launch variant."""
    shapes = []
    v = 0
    while True:
        tile = (ctypes.c_int * 3)()
        block = ctypes.c_int(0)
        if lib.mpf_tiled_variant_shape(v, tile, ctypes.byref(block)) != 0:
            return shapes
        shapes.append((tuple(tile), block.value))
        v += 1


def apply_table(lib, table: dict | None = None) -> int:
    """Set the launch variants of ``table`` (default: tuned_tiled.json when present) in the
    library; returns how many entries were set."""
    if table is None:
        if not TABLE.exists():
            return 0
        table = json.loads(TABLE.read_text())
    count = 0
    for e in table.get("tiled", []):
        if lib.mpf_tiled_variant(OPS[e["op"]], int(e["kind"]), int(e["degree"]), int(e["ncomp_group"]),
                                 int(e["variant"])) >= 0:
            count += 1
    return count


def _time(torch, fn, repeats: int) -> float:
    fn()
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(repeats):
        fn()
    stop.record()
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / repeats


def tune_tiled(kind: int, degree: int, ncomp_group: int, nodes=(256, 256, 256), n: int = 1 << 24,
               repeats: int = 5, seed: int = 0, rounds: int = 3) -> list[dict]:
    """Time every launch variant of the deposit and the gather of ``ncomp_group`` (1 or 3)
    components of weighting (kind, degree) on a periodic grid of ``nodes`` with ``n``
    uniform float32 markers; set the fastest of each in the library and return one entry per
    operation (its timings in ms per call, by variant)."""
    import torch

    from mpifield_amd import _lib, binning

    lib = _lib.load()
    geometry = (0.0, 0.0, 0.0, *(1.0 / k for k in nodes), *nodes, 1, 1, 1)
    gen = torch.Generator(device="cuda").manual_seed(seed)
    pos = torch.rand((3, n), generator=gen, device="cuda", dtype=torch.float32)
    val = torch.randn((ncomp_group, n), generator=gen, device="cuda", dtype=torch.float32)
    bins = binning.bin_markers(pos, val, (kind, degree), geometry)
    field = torch.zeros((*nodes, ncomp_group), dtype=torch.float64, device="cuda")
    out = torch.empty((ncomp_group, n), dtype=torch.float32, device="cuda")
    runs = {"scatter": lambda: binning.scatter_binned(bins, field),
            "gather": lambda: binning.gather_binned(bins, field, out)}
    shapes = variant_shapes(lib)
    entries = []
    for op, fn in runs.items():
        code = OPS[op]
        # every variant warmed up first, then timed in interleaved rounds (the median of
        # each): timing the variants one after the other in a single pass favoured the later
        # ones while the clocks ramped up
        for v in range(len(shapes)):
            lib.mpf_tiled_variant(code, kind, degree, ncomp_group, v)
            _time(torch, fn, 2)
        samples = {v: [] for v in range(len(shapes))}
        for _ in range(rounds):
            for v in range(len(shapes)):
                lib.mpf_tiled_variant(code, kind, degree, ncomp_group, v)
                samples[v].append(_time(torch, fn, repeats))
        times = {v: sorted(t)[len(t) // 2] for v, t in samples.items()}
        best = min(times, key=times.get)
        lib.mpf_tiled_variant(code, kind, degree, ncomp_group, best)
        entries.append({"op": op, "kind": kind, "degree": degree, "ncomp_group": ncomp_group, "variant": best,
                        "tile": list(shapes[best][0]), "block": shapes[best][1],
                        "ms_by_variant": {str(v): round(t, 4) for v, t in times.items()},
                        "workload": {"nodes": list(nodes), "markers": n}})
    return entries


def tune_binned_min(kind: int = 1, degree: int = 2, ncomp: int = 3, nodes=(128, 128, 128),
                    sizes=(1 << 14, 1 << 15, 1 << 16, 1 << 17, 1 << 18, 1 << 19, 1 << 20), repeats: int = 3) -> dict:
    """The smallest batch from which binning inside mpf_scatter + mpf_gather (fixed point)
    beats the direct kernels on this shape; sets it in the library (mpf_set_binned_min)."""
    import torch

    from mpifield_amd import _lib
    from mpifield_amd import kernels as K

    lib = _lib.load()
    geometry = (0.0, 0.0, 0.0, *(1.0 / k for k in nodes), *nodes, 1, 1, 1)
    field = torch.zeros((*nodes, ncomp), dtype=torch.float64, device="cuda")
    timings = {}
    crossover = None
    for n in sizes:
        pos = torch.rand((3, n), device="cuda", dtype=torch.float32)
        val = torch.randn((ncomp, n), device="cuda", dtype=torch.float32)
        res = torch.empty((ncomp, n), dtype=torch.float32, device="cuda")

        def step():
            K.scatter(kind, degree, pos, val, field, geometry, mode=_lib.MODE_FIXED)
            K.gather(kind, degree, pos, field, res, geometry)

        old = lib.mpf_set_binned_min(0)
        binned = _time(torch, step, repeats)
        lib.mpf_set_binned_min(1 << 62)
        direct = _time(torch, step, repeats)
        lib.mpf_set_binned_min(old)
        timings[n] = {"binned_ms": round(binned, 4), "direct_ms": round(direct, 4)}
        if crossover is None and binned < direct:
            crossover = n
    chosen = crossover if crossover is not None else max(sizes) * 2
    lib.mpf_set_binned_min(chosen)
    return {"binned_min": chosen, "timings": timings, "shape": {"kind": kind, "degree": degree, "ncomp": ncomp,
                                                                "nodes": list(nodes)}}


#: the shapes tuned for the committed table: (kind, degree, components per group, nodes, markers)
STANDARD = [
    (1, 2, 3, (256, 256, 256), 1 << 26),   # quadratic, three components: the headline shape
    (1, 2, 1, (256, 256, 256), 1 << 25),
    (0, 1, 3, (256, 256, 256), 1 << 25),   # linear
    (0, 1, 1, (256, 256, 256), 1 << 25),
    (2, 3, 1, (128, 128, 128), 1 << 24),   # cubic
    (2, 2, 3, (256, 256, 256), 1 << 25),   # spline degree 2
    (2, 1, 1, (256, 256, 256), 1 << 25),
    (2, 0, 1, (256, 256, 256), 1 << 25),
]


def main(argv=None) -> None:
    import argparse

    import torch

    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", type=Path, default=TABLE, help="where to write the table (default: tuned_tiled.json "
                    "next to this file, which the library applies when it loads)")
    args = ap.parse_args(argv)
    if not torch.cuda.is_available():
        print("autotune needs the MI355X", file=sys.stderr)
        sys.exit(2)
    table = {"device": torch.cuda.get_device_name(0), "tiled": []}
    for kind, degree, c, nodes, n in STANDARD:
        entries = tune_tiled(kind, degree, c, nodes, n)
        table["tiled"] += entries
        for e in entries:
            print(json.dumps(e), flush=True)
        torch.cuda.empty_cache()
    table["binned"] = tune_binned_min()
    print(json.dumps(table["binned"]), flush=True)
    args.out.parent.mkdir(parents=True, exist_ok=True)
    args.out.write_text(json.dumps(table, indent=1) + "\n")
    print(f"wrote {args.out}", flush=True)


if __name__ == "__main__":
    main()
