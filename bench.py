"""Benchmark: 3D quadratic B-spline scatter + gather of 64M float32 markers on a 256^3
periodic grid with 3 components, on one MI355X.

    python bench.py [--steps K] [--warmup W] [--gpus 1]

One step = one deposit of every marker's 3 values into the float64 field (adds) followed
by one gather of the field at every marker into float32 results: the hot path once over
the whole batch, inputs already resident in HBM. Markers (numpy default_rng(2)): positions
uniform in [0, 1)^3, values N(1, 0.1).

Headline (``value``): markers kept sorted by stencil cell between steps, the way a particle
code keeps them -- the on-device counting sort (mpf_bin_markers) runs once before timing;
the step is the tiled deposit (int64 fixed point, bit-reproducible) + the tiled gather.
Also reported: the same step with the sort inside it (shuffled markers, results returned in
input order) and the plain float64-atomic one-thread-per-marker form on the same markers.

Prints ONE JSON line: throughput (marker scatter+gather pairs per second), the roofline of
the dominant kernel (algorithmic bytes / its HIP-event-timed launches), the error against
the float64 oracle, and the CPU oracle's throughput on a bounded sample.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "markers/s, 3D quadratic scatter+gather 64M; % HBM peak; fixed-point vs FP64"


def parse() -> argparse.Namespace:
    p = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--nodes", type=int, default=256)
    p.add_argument("--markers", type=int, default=1 << 26)
    p.add_argument("--components", type=int, default=3)
    p.add_argument("--cpu-sample", type=int, default=1 << 21, help="markers in the CPU oracle's timed sample")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--only-headline", action="store_true", help="time the headline step only (for rocprofv3)")
    p.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                   help="only copy MIB MiB device-to-device (the counter calibration for rocprofv3 --pmc)")
    return p.parse_args()


def algorithmic_bytes(n: int, nodes: int, comps: int, ndim: int = 3, marker_bytes: int = 4) -> dict:
    """scatter: read positions + values once, read-modify-write the float64 field once;
    gather: read positions, write results, read the field once.  Binning traffic excluded."""
    g = nodes**ndim
    scatter = n * (marker_bytes * ndim + marker_bytes * comps) + 2 * g * comps * 8
    gather = n * (marker_bytes * ndim + marker_bytes * comps) + g * comps * 8
    return {"scatter": scatter, "gather": gather, "pair": scatter + gather}


TRAFFIC_SUMMARY = ROOT / "profiles" / "quadratic3d_256cubed_64M_traffic.json"


def measured_traffic(kernel: str, n: int, nodes: int, comps: int) -> tuple[float | None, str | None]:
    """HBM bytes per launch of `kernel` at the default workload, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (counters scaled by a plain copy of known size); None for
    other workloads or when the summary is absent."""
    if (n, nodes, comps) != (1 << 26, 256, 3) or not TRAFFIC_SUMMARY.exists():
        return None, None
    summary = json.loads(TRAFFIC_SUMMARY.read_text())
    for name, value in summary.get("traffic_bytes_per_launch", {}).items():
        if name.startswith(kernel) or (kernel == "gather_tiled" and name.startswith("gather_items")):
            return float(value), str(TRAFFIC_SUMMARY.relative_to(ROOT))
    return None, None


def log(msg: str) -> None:
    print(f"# {msg}", file=sys.stderr, flush=True)


def cpu_baseline(pos: np.ndarray, val: np.ndarray, nodes: int, sample: int) -> dict:
    """The oracle (C, float64, the reference's OpenMP policy) on `sample` markers."""
    from oracle import oracle as orc

    threads = os.cpu_count() or 1
    geometry = (0.0, 0.0, 0.0, 1.0 / nodes, 1.0 / nodes, 1.0 / nodes, nodes, nodes, nodes, 1, 1, 1)
    p = pos[:, :sample].astype(np.float64)
    v = val[:, :sample].astype(np.float64)
    field = np.zeros((nodes, nodes, nodes, v.shape[0]))
    t0 = time.perf_counter()
    orc.scatter("quadratic", p, v, field, *geometry, threads=threads)  # serial: 256^3 x 3 > MAX_BUFFER_VALUES
    t1 = time.perf_counter()
    out = np.zeros_like(v)
    orc.gather("quadratic", p, field, out, *geometry, threads=threads)
    t2 = time.perf_counter()
    return {
        "value": sample / (t2 - t0),
        "unit": "markers/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {sample} of the same 64M markers (input order) on the 256^3 x 3 grid: scatter {t1 - t0:.2f} s "
        f"(serial, the reference's MAX_BUFFER_VALUES policy) + gather {t2 - t1:.2f} s on {threads} threads",
    }


class Profile:
    """The library's HIP-event timing of its named kernels (mpf_profile_*)."""

    def __init__(self, lib):
        self.lib = lib

    def reset(self, on: bool) -> None:
        self.lib.mpf_profile_reset()
        self.lib.mpf_profile_enable(1 if on else 0)

    def query(self, name: str) -> tuple[float, int]:
        total = ctypes.c_double(0.0)
        count = ctypes.c_int64(0)
        self.lib.mpf_profile_query(name.encode(), ctypes.addressof(total), ctypes.addressof(count))
        return total.value, count.value


def time_steps(torch, step, steps: int, warmup: int, stream) -> tuple[float, float]:
    """(ms per step by HIP events, wall ms per step) over `steps` after `warmup`."""
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record(stream)
    for _ in range(steps):
        step()
    stop.record(stream)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    return start.elapsed_time(stop) / steps, wall


def main() -> None:
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the path runs on one MI355X; "
              "DESIGN.md describes how it would shard)", file=sys.stderr)
        sys.exit(2)
    import torch

    if args.calibrate_copy:
        src = torch.ones(args.calibrate_copy << 17, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each_way": src.numel() * 8, "copies": 3}), flush=True)
        return

    from mpifield_amd import _lib, binning
    from mpifield_amd import kernels as K

    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream(dev)
    n, nodes, comps = args.markers, args.nodes, args.components
    geometry = (0.0, 0.0, 0.0, 1.0 / nodes, 1.0 / nodes, 1.0 / nodes, nodes, nodes, nodes, 1, 1, 1)
    log(f"generating {n} markers")
    rng = np.random.default_rng(2)
    pos_h = rng.random((3, n), dtype=np.float32)
    val_h = rng.normal(1.0, 0.1, (comps, n)).astype(np.float32)
    pos = torch.as_tensor(pos_h, device=dev)
    val = torch.as_tensor(val_h, device=dev)
    field = torch.zeros((nodes, nodes, nodes, comps), dtype=torch.float64, device=dev)
    result = torch.empty((comps, n), dtype=torch.float32, device=dev)
    prof = Profile(_lib.load())

    # ---- headline: bins kept between steps (the sort before timing) ----
    log("binning (before timing)")
    bins = binning.bin_markers(pos, val, (1, 2), geometry)
    torch.cuda.synchronize()

    def headline():
        binning.scatter_binned(bins, field)
        binning.gather_binned(bins, field, result)

    prof.reset(False)
    for _ in range(args.warmup):
        headline()
    torch.cuda.synchronize()
    prof.reset(True)
    ms_step, wall_step = time_steps(torch, headline, args.steps, 0, stream)
    sc_ms, sc_n = prof.query("scatter_tiled")
    ga_ms, ga_n = prof.query("gather_tiled")
    prof.reset(False)
    alg = algorithmic_bytes(n, nodes, comps)
    value = n / (ms_step * 1e-3)
    sc_step, ga_step = sc_ms / args.steps, ga_ms / args.steps
    kernels = {
        "scatter_tiled": {"ms_per_step": sc_step, "launches_per_step": sc_n / args.steps,
                          "algorithmic_bytes": alg["scatter"], "GBs": alg["scatter"] / (sc_step * 1e-3) / 1e9},
        "gather_tiled": {"ms_per_step": ga_step, "launches_per_step": ga_n / args.steps,
                         "algorithmic_bytes": alg["gather"], "GBs": alg["gather"] / (ga_step * 1e-3) / 1e9},
    }
    dominant = max(kernels, key=lambda k: kernels[k]["ms_per_step"])
    dk = kernels[dominant]
    achieved = dk["GBs"]
    launches = dk["launches_per_step"]
    traffic, traffic_source = measured_traffic(dominant, n, nodes, comps)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "markers/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic",
        "config": {
            "workload": f"3D quadratic B-spline scatter+gather, {nodes}^3 periodic nodes, {comps} components, "
            f"{n} float32 markers kept cell-sorted (binned before timing); float64 stencils, int64 fixed-point "
            "deposit, float64 field",
            "global_batch": n,
            "parallelism": "single-gpu",
        },
        "hbm_fraction_pair": alg["pair"] / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "wall_ms_per_step": wall_step,
        "kernels": kernels,
        "roofline": {
            "bound": "hbm",
            "kernel": dominant,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": traffic_source,
            "algorithmic_bytes_per_launch": dk["algorithmic_bytes"] / launches if launches else None,
            "launches_per_step": launches,
        },
    }
    if not args.only_headline:
        # ---- error against the float64 oracle on a sample of markers ----
        from oracle import oracle as orc

        check = torch.zeros_like(field)
        binning.scatter_binned(bins, check)
        sample = 1 << 16
        idx = np.random.default_rng(7).choice(n, sample, replace=False)
        sorted_pos = bins.positions[:, torch.as_tensor(idx, device=dev)].double().cpu().numpy()
        check_h = check.cpu().numpy()
        res_s = torch.empty_like(result)
        binning.gather_binned(bins, check, res_s)
        want_g = np.zeros((comps, sample))
        orc.gather("quadratic", sorted_pos, check_h, want_g, *geometry, threads=os.cpu_count() or 1)
        got_g = res_s[:, torch.as_tensor(idx, device=dev)].double().cpu().numpy()
        # the deposit against the oracle's float64 deposit of the same markers
        want_f = np.zeros_like(check_h)
        orc.scatter("quadratic", pos_h.astype(np.float64), val_h.astype(np.float64), want_f, *geometry, threads=1)
        line["error_vs_fp64_oracle"] = {
            "scatter_max_rel": float(np.abs(check_h - want_f).max() / np.abs(want_f).max()),
            "gather_max_rel": float(np.abs(got_g - want_g).max() / np.abs(want_g).max()),
        }
        del check
        log("shuffled step (sort inside) and the float64-atomic form")

        # ---- the sort inside the step: shuffled input, results back in input order ----
        def shuffled():
            b = binning.bin_markers(pos, val, (1, 2), geometry)
            binning.scatter_binned(b, field)
            binning.gather_binned(b, field, result)
            b.unsort(result)

        sh_ms, _ = time_steps(torch, shuffled, max(3, args.steps // 4), 1, stream)

        # ---- the plain float64-atomic one-thread-per-marker form, same markers: the direct
        # entry points (mpf_scatter in atomic mode, mpf_gather), not the binned route ----
        def direct_gather(p):
            ndim, lower, cell, num, code, _ = K._split(geometry)
            _lib.call("mpf_gather", 1, 0, 3, p.data_ptr(), p.stride(0), _lib.F32, n, field.data_ptr(), comps,
                      result.data_ptr(), result.stride(0), lower.ctypes.data, cell.ctypes.data, num.ctypes.data,
                      code.ctypes.data, None, K.stream_handle(dev))

        def atomic_sorted_direct():
            K.scatter(1, 0, bins.positions, bins.values, field, geometry, mode=_lib.MODE_ATOMIC)
            direct_gather(bins.positions)

        def atomic_input_direct():
            K.scatter(1, 0, pos, val, field, geometry, mode=_lib.MODE_ATOMIC)
            direct_gather(pos)

        at_s, _ = time_steps(torch, atomic_sorted_direct, max(3, args.steps // 4), 1, stream)
        at_u, _ = time_steps(torch, atomic_input_direct, 3, 1, stream)
        line["variants"] = {
            "binned_sort_inside_step_markers_per_s": n / (sh_ms * 1e-3),
            "binned_sort_inside_step_ms": sh_ms,
            "fp64_atomic_cell_sorted_markers_per_s": n / (at_s * 1e-3),
            "fp64_atomic_cell_sorted_ms": at_s,
            "fp64_atomic_input_order_markers_per_s": n / (at_u * 1e-3),
            "fp64_atomic_input_order_ms": at_u,
        }
    if not (args.no_cpu or args.only_headline):
        log("CPU oracle sample")
        line["cpu_baseline"] = cpu_baseline(pos_h, val_h, nodes, min(args.cpu_sample, n))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
