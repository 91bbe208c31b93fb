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
input order), with an incremental re-binning inside it (the binned markers moved by up to
half a cell, results kept in bins order) and the plain float64-atomic
one-thread-per-marker form on the same markers.

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
    p.add_argument("--config", type=int, default=2, choices=(2, 3, 4, 5, 6, 7, 8, 9, 10),
                   help="2: the headline (default); 3: 2D quadratic-edge path deposit, 1024^2, 2^25 paths; "
                   "4: 4096 batched 2D patches, ~2e8 markers; 5: differentiable 3D cubic gather of scatter, "
                   "128^3, 2^24 markers; 6: 2D Whitney path deposit, same paths as 3; 7: halo fold + fill, "
                   "256^3 x 3, halo 2; 8: 2D quadratic-edge field at 2^25 markers; 9: 3D quadratic faces (2-forms) "
                   "scatter+gather, 256^3 x 3, 2^25 markers; 10: 3D quadratic-edge path deposit, 128^3, 2^23 paths "
                   "(one JSON line each)")
    return p.parse_args()


def algorithmic_bytes(n: int, nodes: int, comps: int, ndim: int = 3, marker_bytes: int = 4) -> dict:
    """scatter: read positions + values once, read-modify-write the float64 field once;
    gather: read positions, write results, read the field once.  Binning traffic excluded."""
    g = nodes**ndim
    scatter = n * (marker_bytes * ndim + marker_bytes * comps) + 2 * g * comps * 8
    gather = n * (marker_bytes * ndim + marker_bytes * comps) + g * comps * 8
    return {"scatter": scatter, "gather": gather, "pair": scatter + gather}


TRAFFIC_SUMMARY = ROOT / "profiles" / "quadratic3d_256cubed_64M_traffic.json"
PATHS_VALU = ROOT / "profiles" / "paths_valu.json"
HEADLINE_VALU = ROOT / "profiles" / "quadratic3d_256cubed_64M_valu.json"


def paths_valu_roofline(kernel: str, paths: int, kernel_ms: float) -> dict | None:
    """The VALU roofline of a tiled path deposit: the lane operations per path that rocprofv3
    counted (committed summary) times the paths, over the kernel's HIP-event time."""
    if not PATHS_VALU.exists() or kernel_ms <= 0.0:
        return None
    summary = json.loads(PATHS_VALU.read_text())
    ops = summary["valu_lane_ops_per_path"].get(kernel)
    if ops is None:
        return None
    peak = summary["peak_lane_ops_per_s"] / 1e9
    achieved = ops * paths / (kernel_ms * 1e-3) / 1e9
    return {"bound": "valu", "kernel": kernel, "achieved": achieved, "peak": peak, "unit": "G lane-ops/s",
            "frac": achieved / peak, "lane_ops_per_path": ops, "kernel_ms": kernel_ms,
            "source": str(PATHS_VALU.relative_to(ROOT))}


def time_binned_paths(torch, fn, steps: int) -> tuple[float, float]:
    """(ms per call of fn by HIP events, ms per call of the library's spline_paths_tiled)."""
    from mpifield_amd import _lib

    prof = Profile(_lib.load())
    prof.reset(False)
    fn()
    torch.cuda.synchronize()
    prof.reset(True)
    ms = time_events(torch, fn, steps, 0)
    kms, count = prof.query("spline_paths_tiled")
    prof.reset(False)
    return ms, (kms / count if count else 0.0)


def measured_traffic(kernel: str, n: int, nodes: int, comps: int) -> tuple[float | None, str | None]:
    """HBM bytes per launch of `kernel` at the default workload, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (counters scaled by a plain copy of known size); None for
    other workloads or when the summary is absent."""
    if (n, nodes, comps) != (1 << 26, 256, 3) or not TRAFFIC_SUMMARY.exists():
        return None, None
    summary = json.loads(TRAFFIC_SUMMARY.read_text())
    for name, value in summary.get("traffic_bytes_per_launch", {}).items():
        if name.startswith(kernel) or (kernel == "gather_tiled" and name.startswith("gather_lanes")):
            return float(value), str(TRAFFIC_SUMMARY.relative_to(ROOT))
    return None, None


def log(msg: str) -> None:
    print(f"# {msg}", file=sys.stderr, flush=True)


def cpu_baseline(pos: np.ndarray, val: np.ndarray, nodes: int, sample: int, guard_lifted: bool = True) -> dict:
    """The oracle (C, float64, the reference's OpenMP policy) on `sample` markers; with
    guard_lifted, also the whole batch with the policy's private-copy bound lifted."""
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
    line = {
        "value": sample / (t2 - t0),
        "unit": "markers/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {sample} of the same 64M markers (input order) on the 256^3 x 3 grid: scatter {t1 - t0:.2f} s "
        f"(serial, the reference's MAX_BUFFER_VALUES policy) + gather {t2 - t1:.2f} s on {threads} threads",
    }
    if guard_lifted:
        # the same kernels on every marker with the private-copy bound lifted: each thread
        # deposits its chunk into its own 256^3 x 3 copy, the copies summed in thread order
        # (threads x 403 MB of host memory); a labelled variant, not the reference's policy
        del p, v, out
        n = pos.shape[1]
        p = pos.astype(np.float64)
        v = val.astype(np.float64)
        field = np.zeros((nodes, nodes, nodes, v.shape[0]))
        t0 = time.perf_counter()
        orc.scatter("quadratic", p, v, field, *geometry, threads=threads, guard=False)
        t1 = time.perf_counter()
        out = np.zeros_like(v)
        orc.gather("quadratic", p, field, out, *geometry, threads=threads)
        t2 = time.perf_counter()
        line["guard_lifted"] = {
            "value": n / (t2 - t0),
            "unit": "markers/s",
            "cores": threads,
            "sample": f"all {n} markers, the reference's MAX_BUFFER_VALUES guard lifted (not its policy): scatter "
            f"{t1 - t0:.2f} s into {threads} private copies + gather {t2 - t1:.2f} s on {threads} threads",
        }
    return line


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


def time_events(torch, fn, steps: int, warmup: int) -> float:
    """ms per call of fn by HIP events on the current stream."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(steps):
        fn()
    stop.record()
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / steps


def cpu_cores() -> int:
    """Host threads for the oracle's OpenMP loops: the box's share (16) at most."""
    return max(1, min(16, os.cpu_count() or 1))


def cpu_seconds(fn) -> float:
    t0 = time.perf_counter()
    fn()
    return time.perf_counter() - t0


def paths_2d(torch, n: int, nodes: int):
    """The 1024^2 path workload: start uniform in [0,1)^2, per-axis displacement
    clip(N(0, h/4), +-h/2), end = start + displacement (not wrapped), unit charges."""
    h = 1.0 / nodes
    gen = torch.Generator(device="cuda").manual_seed(3)
    start = torch.rand((2, n), generator=gen, device="cuda", dtype=torch.float32)
    d = (torch.randn((2, n), generator=gen, device="cuda") * (0.25 * h)).clamp(-0.5 * h, 0.5 * h)
    end = (start.double() + d.double()).float()
    q = torch.ones(n, device="cuda", dtype=torch.float32)
    return (0.0, 0.0, h, h, nodes, nodes, 1, 1), start, end, q


def config3(torch, args) -> dict:
    """2D charge-conserving path deposit (QUADRATIC_EDGES, spline degree 2) on a 1024^2 periodic
    grid: 2^25 float32 paths, start uniform, per-axis displacement clip(N(0, h/4), +-h/2),
    unit charges.  The value: paths whose starts are kept binned (the markers' positions at the
    start of a step, sorted before timing as in the headline); beside it the same deposit with
    the sort inside and the float64-atomic form.  The discrete continuity G^T j = rho(end) -
    rho(start) against the quadratic node charge, all on the GPU."""
    from mpifield_amd import _lib, binning
    from mpifield_amd import kernels as K

    nodes, n = 1024, 1 << 25
    g, start, end, q = paths_2d(torch, n, nodes)
    j = torch.zeros((nodes, nodes, 2), dtype=torch.float64, device="cuda")
    out = {}
    for label, mode in (("fixed", _lib.MODE_FIXED), ("atomic", _lib.MODE_ATOMIC)):
        ms = time_events(torch, lambda: K.deposit_paths(1, 2, start, end, q, j, g, mode=mode), args.steps, 1)
        out[label] = ms
    bins = binning.bin_markers(start, None, (_lib.KIND_SPLINE, 2), g)  # before timing
    end_b, q_b = end[:, bins.order].contiguous(), q[bins.order].contiguous()
    out["binned"], kernel_ms = time_binned_paths(torch, lambda: binning.deposit_paths_binned(bins, end_b, q_b, j, 1, 2),
                                                 args.steps)
    j.zero_()
    binning.deposit_paths_binned(bins, end_b, q_b, j, 1, 2)
    rho0 = torch.zeros((nodes, nodes, 1), dtype=torch.float64, device="cuda")
    rho1 = torch.zeros_like(rho0)
    K.scatter(2, 2, start, q[None], rho0, g, mode=_lib.MODE_FIXED)
    K.scatter(2, 2, end, q[None], rho1, g, mode=_lib.MODE_FIXED)
    residual = (K.divergence(j, (1, 1)) - (rho1 - rho0)[..., 0]).abs().max().item()
    alg = n * (16 + 4) + 2 * nodes * nodes * 2 * 8
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as orc

        s_h, e_h, q_h = start.double().cpu().numpy(), end.double().cpu().numpy(), q.double().cpu().numpy()
        edges = np.zeros((nodes, nodes, 2))
        sec = cpu_seconds(lambda: orc.deposit_spline(s_h, e_h, q_h, edges, *g, degree=2))
        cpu = {"value": n / sec, "unit": "paths/s", "cores": 1, "kind": "port",
               "sample": f"all {n} paths: {sec:.2f} s (the reference's spline deposit is serial)"}
    return {
        "metric": "paths/s, 2D QUADRATIC_EDGES path deposit 1024^2, 2^25 paths", "value": n / (out["binned"] * 1e-3),
        "unit": "paths/s", "n_gpus": 1, "steps": args.steps, "ms_per_step": out["binned"], "higher_is_better": True,
        "dtype": "int64", "data": "synthetic",
        "config": {"workload": "2D spline-edge (degree 2) charge-conserving path deposit, 1024^2 periodic, "
                   f"{n} float32 paths, displacements clip(N(0,h/4), h/2); starts kept binned (sorted before "
                   "timing), tiled fixed point"},
        "roofline": {"bound": "hbm", "achieved": alg / (out["binned"] * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": alg / (out["binned"] * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None},
        "roofline_valu": paths_valu_roofline("spline_paths_tiled<float, 2, 2, 256>", n, kernel_ms),
        "variants": {"sort_inside_ms": out["fixed"], "sort_inside_paths_per_s": n / (out["fixed"] * 1e-3),
                     "fp64_atomic_ms": out["atomic"], "fp64_atomic_paths_per_s": n / (out["atomic"] * 1e-3)},
        "continuity_max_abs_residual": residual, "total_charge": float(n), "cpu_baseline": cpu,
    }


def config10(torch, args) -> dict:
    """3D charge-conserving path deposit (QUADRATIC_EDGES, spline degree 2) on a 128^3 periodic
    grid: 2^23 float32 paths, start uniform, per-axis displacement clip(N(0, h/4), +-h/2),
    unit charges.  The value: starts kept binned (sorted before timing); beside it the same
    deposit with the sort inside and the float64-atomic form; the continuity residual on the GPU."""
    from mpifield_amd import _lib, binning
    from mpifield_amd import kernels as K

    nodes, n = 128, 1 << 23
    h = 1.0 / nodes
    g = (0.0, 0.0, 0.0, h, h, h, nodes, nodes, nodes, 1, 1, 1)
    gen = torch.Generator(device="cuda").manual_seed(10)
    start = torch.rand((3, n), generator=gen, device="cuda", dtype=torch.float32)
    d = (torch.randn((3, n), generator=gen, device="cuda") * (0.25 * h)).clamp(-0.5 * h, 0.5 * h)
    end = (start.double() + d.double()).float()
    q = torch.ones(n, device="cuda", dtype=torch.float32)
    j = torch.zeros((nodes, nodes, nodes, 3), dtype=torch.float64, device="cuda")
    out = {}
    for label, mode in (("fixed", _lib.MODE_FIXED), ("atomic", _lib.MODE_ATOMIC)):
        out[label] = time_events(torch, lambda: K.deposit_paths(1, 2, start, end, q, j, g, mode=mode), args.steps, 1)
    bins = binning.bin_markers(start, None, (_lib.KIND_SPLINE, 2), g)  # before timing
    end_b, q_b = end[:, bins.order].contiguous(), q[bins.order].contiguous()
    out["binned"], kernel_ms = time_binned_paths(torch, lambda: binning.deposit_paths_binned(bins, end_b, q_b, j, 1, 2),
                                                 args.steps)
    j.zero_()
    binning.deposit_paths_binned(bins, end_b, q_b, j, 1, 2)
    rho0 = torch.zeros((nodes, nodes, nodes, 1), dtype=torch.float64, device="cuda")
    rho1 = torch.zeros_like(rho0)
    K.scatter(2, 2, start, q[None], rho0, g, mode=_lib.MODE_FIXED)
    K.scatter(2, 2, end, q[None], rho1, g, mode=_lib.MODE_FIXED)
    residual = (K.divergence(j, (1, 1, 1)) - (rho1 - rho0)[..., 0]).abs().max().item()
    alg = n * (24 + 4) + 2 * nodes**3 * 3 * 8
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as orc

        sample = 1 << 20
        s_h, e_h = start[:, :sample].double().cpu().numpy(), end[:, :sample].double().cpu().numpy()
        q_h = q[:sample].double().cpu().numpy()
        edges = np.zeros((nodes, nodes, nodes, 3))
        sec = cpu_seconds(lambda: orc.deposit_spline(s_h, e_h, q_h, edges, *g, degree=2))
        cpu = {"value": sample / sec, "unit": "paths/s", "cores": 1, "kind": "port",
               "sample": f"first {sample} paths: {sec:.2f} s (the reference's spline deposit is serial)"}
    return {
        "metric": "paths/s, 3D QUADRATIC_EDGES path deposit 128^3, 2^23 paths", "value": n / (out["binned"] * 1e-3),
        "unit": "paths/s", "n_gpus": 1, "steps": args.steps, "ms_per_step": out["binned"], "higher_is_better": True,
        "dtype": "int64", "data": "synthetic",
        "config": {"workload": "3D spline-edge (degree 2) charge-conserving path deposit, 128^3 periodic, "
                   f"{n} float32 paths, displacements clip(N(0,h/4), h/2); starts kept binned (sorted before "
                   "timing), tiled fixed point"},
        "roofline": {"bound": "hbm", "achieved": alg / (out["binned"] * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": alg / (out["binned"] * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None},
        "roofline_valu": paths_valu_roofline("spline_paths_tiled<float, 3, 2, 512>", n, kernel_ms),
        "variants": {"sort_inside_ms": out["fixed"], "sort_inside_paths_per_s": n / (out["fixed"] * 1e-3),
                     "fp64_atomic_ms": out["atomic"]},
        "continuity_max_abs_residual": residual, "cpu_baseline": cpu,
    }


def config6(torch, args) -> dict:
    """The WHITNEY (degree-1 edge) path deposit of the same 1024^2 workload: the value with the
    starts kept binned (sorted before timing); beside it the sort inside the call and float64
    atomics; the continuity G^T j = rho(end) - rho(start) with the linear node charge."""
    from mpifield_amd import _lib, binning
    from mpifield_amd import kernels as K

    nodes, n = 1024, 1 << 25
    g, start, end, q = paths_2d(torch, n, nodes)
    j = torch.zeros((nodes, nodes, 2), dtype=torch.float64, device="cuda")
    out = {}
    for label, mode in (("fixed", _lib.MODE_FIXED), ("atomic", _lib.MODE_ATOMIC)):
        out[label] = time_events(torch, lambda: K.deposit_paths(0, 1, start, end, q, j, g, mode=mode),
                                 max(1, args.steps // 2), 1)
    bins = binning.bin_markers(start, None, (_lib.KIND_LINEAR, 1), g)  # before timing
    end_b, q_b = end[:, bins.order].contiguous(), q[bins.order].contiguous()
    out["binned"] = time_events(torch, lambda: binning.deposit_paths_binned(bins, end_b, q_b, j, 0, 1), args.steps, 1)
    j.zero_()
    binning.deposit_paths_binned(bins, end_b, q_b, j, 0, 1)
    rho0 = torch.zeros((nodes, nodes, 1), dtype=torch.float64, device="cuda")
    rho1 = torch.zeros_like(rho0)
    K.scatter(0, 1, start, q[None], rho0, g, mode=_lib.MODE_FIXED)
    K.scatter(0, 1, end, q[None], rho1, g, mode=_lib.MODE_FIXED)
    residual = (K.divergence(j, (1, 1)) - (rho1 - rho0)[..., 0]).abs().max().item()
    alg = n * (16 + 4) + 2 * nodes * nodes * 2 * 8
    best = out["binned"]
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as orc

        s_h, e_h, q_h = start.double().cpu().numpy(), end.double().cpu().numpy(), q.double().cpu().numpy()
        edges = np.zeros((nodes, nodes, 2))
        cores = cpu_cores()
        sec = cpu_seconds(lambda: orc.deposit_whitney(s_h, e_h, q_h, edges, *g, threads=cores))
        cpu = {"value": n / sec, "unit": "paths/s", "cores": cores, "kind": "port",
               "sample": f"all {n} paths: {sec:.2f} s (OpenMP with private copies, as the reference)"}
    return {
        "metric": "paths/s, 2D WHITNEY path deposit 1024^2, 2^25 paths", "value": n / (best * 1e-3),
        "unit": "paths/s", "n_gpus": 1, "ms_per_step": best, "higher_is_better": True, "dtype": "int64",
        "data": "synthetic",
        "config": {"workload": f"2D Whitney charge-conserving path deposit, 1024^2 periodic, {n} float32 paths, "
                   "displacements clip(N(0,h/4), h/2); starts kept binned (sorted before timing), tiled fixed point"},
        "roofline": {"bound": "hbm", "achieved": alg / (best * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": alg / (best * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None},
        "variants": {"sort_inside_ms": out["fixed"], "fp64_atomic_ms": out["atomic"]},
        "continuity_max_abs_residual": residual, "cpu_baseline": cpu,
    }


def config8(torch, args) -> dict:
    """The QUADRATIC_EDGES field at markers (the transpose of config 3): 2^25 float32 markers
    (the paths' start points) on the 1024^2 periodic grid, a random edge field."""
    from mpifield_amd import kernels as K

    nodes, n = 1024, 1 << 25
    g, start, _, _ = paths_2d(torch, n, nodes)
    gen = torch.Generator(device="cuda").manual_seed(8)
    edges = torch.randn((nodes, nodes, 2), generator=gen, device="cuda", dtype=torch.float64)
    result = torch.empty((2, n), dtype=torch.float32, device="cuda")
    ms = time_events(torch, lambda: K.gather_edges(1, 2, start, edges, result, g), args.steps, 2)
    alg = n * (8 + 8) + nodes * nodes * 2 * 8
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as orc

        sample = 1 << 22
        p_h = start[:, :sample].double().cpu().numpy()
        e_h = edges.cpu().numpy()
        r_h = np.zeros((2, sample))
        sec = cpu_seconds(lambda: orc.gather_edges(p_h, e_h, r_h, *g, degree=2))
        got = result[:, :sample].double().cpu().numpy()
        err = float(np.abs(got - r_h).max() / np.abs(r_h).max())
        cpu = {"value": sample / sec, "unit": "markers/s", "cores": 1, "kind": "port",
               "sample": f"first {sample} markers: {sec:.2f} s; max relative difference of the GPU result {err:.1e}"}
    return {
        "metric": "markers/s, 2D QUADRATIC_EDGES edge gather 1024^2, 2^25 markers", "value": n / (ms * 1e-3),
        "unit": "markers/s", "n_gpus": 1, "ms_per_step": ms, "higher_is_better": True, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"2D spline-edge (degree 2) field at {n} float32 markers, 1024^2 periodic, float64 edges; "
                   "one marker per work-item"},
        "roofline": {"bound": "hbm", "achieved": alg / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None},
        "cpu_baseline": cpu,
    }


def config9(torch, args) -> dict:
    """The 2-forms (faces) of degree 2 in 3D on a 256^3 periodic grid: 2^25 float32 markers
    uniform in [0, 1)^3, 3 values N(1, 0.1) each; one step = the fixed-point deposit of the
    face values + the gather of a face field.  Timed on the markers sorted by stencil cell
    (the order a particle code keeps; the sort runs once before timing) and in input order."""
    from mpifield_amd import binning
    from mpifield_amd import kernels as K

    nodes, n, C = 256, 1 << 25, 3
    h = 1.0 / nodes
    g = (0.0, 0.0, 0.0, h, h, h, nodes, nodes, nodes, 1, 1, 1)
    gen = torch.Generator(device="cuda").manual_seed(9)
    pos = torch.rand((3, n), generator=gen, device="cuda", dtype=torch.float32)
    val = (1.0 + 0.1 * torch.randn((C, n), generator=gen, device="cuda")).float()
    bins = binning.bin_markers(pos, val, (2, 2), g)
    spos, sval = bins.positions, bins.values
    field = torch.zeros((nodes, nodes, nodes, C), dtype=torch.float64, device="cuda")
    faces = torch.randn((nodes, nodes, nodes, C), generator=gen, device="cuda", dtype=torch.float64)
    result = torch.empty((C, n), dtype=torch.float32, device="cuda")
    from mpifield_amd import _lib

    out = {}
    prof = Profile(_lib.load())
    prof.reset(False)
    for _ in range(2):
        binning.scatter_forms_binned(bins, field, 2)
    prof.reset(True)  # the library's HIP events around the tiled launch (the finalize pass excluded)
    sc = time_events(torch, lambda: binning.scatter_forms_binned(bins, field, 2), args.steps, 0)
    tiled_ms, tiled_n = prof.query("scatter_forms_tiled")
    prof.reset(False)
    ga = time_events(torch, lambda: binning.gather_forms_binned(bins, faces, result, 2), args.steps, 2)
    with K.deposit_mode("fixed"):
        out["input"] = (time_events(torch, lambda: K.scatter_forms(2, 2, pos, val, field, g), max(1, args.steps // 4), 1),
                        time_events(torch, lambda: K.gather_forms(2, 2, pos, faces, result, g), args.steps, 2))
    with K.deposit_mode("atomic"):
        out["direct_atomic_sorted"] = time_events(torch, lambda: K.scatter_forms(2, 2, spos, sval, field, g),
                                                  max(1, args.steps // 4), 1)
    launch_ms = tiled_ms / max(tiled_n, 1)
    per_marker = 12 + 12  # float32 positions + values (scatter) or results (gather)
    alg_sc, alg_ga = n * per_marker + 2 * nodes**3 * C * 8, n * per_marker + nodes**3 * C * 8
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as orc

        sample, cores = 1 << 21, cpu_cores()
        p_h = spos[:, :sample].double().cpu().numpy()
        v_h = sval[:, :sample].double().cpu().numpy()
        f_h = np.zeros((nodes, nodes, nodes, C))
        r_h = np.zeros((C, sample))
        t_sc = cpu_seconds(lambda: orc.scatter_forms(p_h, v_h, f_h, *g, degree=2, form=2, threads=cores))
        t_ga = cpu_seconds(lambda: orc.gather_forms(p_h, f_h, r_h, *g, degree=2, form=2, threads=cores))
        cpu = {"value": sample / (t_sc + t_ga), "unit": "markers/s", "cores": cores, "kind": "port",
               "sample": f"first {sample} sorted markers: scatter {t_sc:.2f} s (serial: the grid exceeds the "
                         f"reference's MAX_BUFFER_VALUES) + gather {t_ga:.2f} s on {cores} threads"}
    return {
        "metric": "markers/s, 3D quadratic faces (2-forms) scatter+gather 256^3 x 3, 2^25 markers",
        "value": n / ((sc + ga) * 1e-3), "unit": "markers/s", "n_gpus": 1, "ms_per_step": sc + ga,
        "higher_is_better": True, "dtype": "int64", "data": "synthetic",
        "config": {"workload": f"3D 2-forms of degree 2, 256^3 x 3 periodic, {n} float32 markers sorted by stencil "
                   "cell: tiled fixed-point deposit + tiled gather (bins made once before timing, results in bins "
                   "order)"},
        "roofline": {"bound": "hbm", "achieved": alg_sc / (launch_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": alg_sc / (launch_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "scatter_forms_tiled", "launch_ms": launch_ms},
        "scatter_ms": sc, "gather_ms": ga, "gather_GBs": alg_ga / (ga * 1e-3) / 1e9,
        "input_order": {"scatter_ms": out["input"][0], "gather_ms": out["input"][1],
                        "note": "both bin inside the call (tiled; the gather unsorts its results)"},
        "direct_fp64_atomic_sorted_scatter_ms": out["direct_atomic_sorted"],
        "cpu_baseline": cpu,
    }


def config7(torch, args) -> dict:
    """Halo fold then fill of one block's storage: 256^3 interior nodes, halo 2 on every side,
    3 components, periodic (the exchange a single block does locally)."""
    from mpifield_amd import kernels as K

    inner, h, C = 256, 2, 3
    side = inner + 2 * h
    gen = torch.Generator(device="cuda").manual_seed(7)
    data = torch.randn((side, side, side, C), generator=gen, device="cuda", dtype=torch.float64)
    per = (True, True, True)
    fold = time_events(torch, lambda: K.halo("fold", data, h, per), args.steps, 2)
    fill = time_events(torch, lambda: K.halo("fill", data, h, per), args.steps, 2)
    slab = h * side * side * C * 8  # one halo slab of one axis side
    alg_fold, alg_fill = 3 * 2 * slab * 4, 3 * 2 * slab * 2  # per axis and side: fold r/w halo + r/w interior
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as orc

        host = data.cpu().numpy()
        reps = 20
        sec = cpu_seconds(lambda: [(orc.halo_fold(host, h, per), orc.halo_fill(host, h, per)) for _ in range(reps)])
        cpu = {"value": reps * (alg_fold + alg_fill) / sec / 1e9, "unit": "GB/s", "cores": 1, "kind": "port",
               "sample": f"the same storage, fold then fill {reps} times: {sec:.2f} s"}
    total = fold + fill
    return {
        "metric": "GB/s, halo fold + fill, 256^3 x 3 float64 storage, halo 2", "value": (alg_fold + alg_fill) / (total * 1e-3) / 1e9,
        "unit": "GB/s", "n_gpus": 1, "ms_per_step": total, "higher_is_better": True, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{side}^3 x {C} float64 storage ({inner}^3 interior, halo {h}), periodic, fold then fill"},
        "roofline": {"bound": "hbm", "achieved": (alg_fold + alg_fill) / (total * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": (alg_fold + alg_fill) / (total * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None},
        "fold_ms": fold, "fill_ms": fill, "cpu_baseline": cpu,
    }


def config4(torch, args) -> dict:
    """4096 2D patches of 16^2..128^2 nodes, spline degree 1..3 per patch, markers per patch
    log-uniform in [1e3, 1e6] rescaled to about 2e8 in total, float32 positions uniform in the
    patch, one component, bounded patches: one persistent deposit launch and one gather."""
    import mpifield_amd as m
    from mpifield_amd import patches as P

    rng = np.random.default_rng(4)
    count = 4096
    sides = rng.integers(16, 129, count)
    degrees = rng.integers(1, 4, count)
    raw = np.exp(rng.uniform(np.log(1e3), np.log(1e6), count))
    markers = np.clip(np.round(raw * 2e8 / raw.sum()), 1e3, 1e6).astype(np.int64)
    weightings = {1: m.LINEAR, 2: m.QUADRATIC, 3: m.CUBIC}
    batch = P.PatchBatch(ndim=2, ncomp=1)
    lows, cells = [], []
    for i in range(count):
        cell = 1.0 / float(sides[i] - 1)
        lower = (float(i % 64), float(i // 64))
        batch.add(lower, (cell, cell), (int(sides[i]), int(sides[i])), weightings[int(degrees[i])],
                          int(markers[i]))
        lows.append(lower)
        cells.append(cell)
    total = int(markers.sum())
    batch.build()
    # positions uniform in each patch, made on the GPU
    dev = torch.device("cuda")
    patch_of = torch.repeat_interleave(torch.arange(count, device=dev), torch.as_tensor(markers, device=dev))
    lo = torch.as_tensor(np.array(lows), device=dev, dtype=torch.float64)[patch_of].T
    span = torch.as_tensor((sides - 1) * np.array(cells), device=dev, dtype=torch.float64)[patch_of]
    gen = torch.Generator(device="cuda").manual_seed(4)
    pos = (lo + torch.rand((2, total), generator=gen, device=dev, dtype=torch.float64) * span).float()
    val = torch.ones((1, total), device=dev, dtype=torch.float32)
    fields = torch.zeros(batch.field_values, dtype=torch.float64, device=dev)
    result = torch.empty((1, total), dtype=torch.float32, device=dev)
    reps = max(1, args.steps // 4)
    # input order (markers uniform in each patch, unsorted)
    sc_in = time_events(torch, lambda: batch.scatter(pos, val, fields), reps, 1)
    ga_in = time_events(torch, lambda: batch.gather(pos, fields, result), reps, 1)
    # the headline: each patch's markers kept sorted by cell (sorted before timing, as the
    # 3D headline keeps its markers binned), the launch autotuned on them
    pos, val, _ = batch.sort(pos, val)
    tuned = P.autotune(batch, pos, val, fields)  # (chunk, block, workgroups) of the deposit launch
    sc = time_events(torch, lambda: batch.scatter(pos, val, fields), reps, 1)
    ga = time_events(torch, lambda: batch.gather(pos, fields, result), reps, 1)
    alg = total * 12 * 2 + int((3 * sides.astype(np.int64) ** 2 * 8).sum())
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as orc

        cores, names = cpu_cores(), {1: "linear", 2: "quadratic", 3: "spline"}
        done, sec, first = 0, 0.0, 0
        pos_h = pos.cpu().numpy()
        for i in range(count):
            m_i = int(markers[i])
            p_h = pos_h[:, first:first + m_i].astype(np.float64)
            v_h = np.ones((1, m_i))
            geo = (*lows[i], cells[i], cells[i], int(sides[i]), int(sides[i]), 0, 0)
            f_h = np.zeros((int(sides[i]), int(sides[i]), 1))
            r_h = np.zeros((1, m_i))
            kind, deg = names[int(degrees[i])], int(degrees[i])
            sec += cpu_seconds(lambda: (orc.scatter(kind, p_h, v_h, f_h, *geo, degree=deg, threads=cores),
                                        orc.gather(kind, p_h, f_h, r_h, *geo, degree=deg, threads=cores)))
            first += m_i
            done += m_i
        cpu = {"value": done / sec, "unit": "markers/s", "cores": cores, "kind": "port",
               "sample": f"all {count} patches ({done} markers) one by one, scatter then gather: {sec:.2f} s"}
    return {
        "metric": "markers/s, 4096 batched 2D patches scatter+gather, ~2e8 markers", "value": total / ((sc + ga) * 1e-3),
        "unit": "markers/s", "n_gpus": 1, "ms_per_step": sc + ga, "higher_is_better": True, "dtype": "int64",
        "data": "synthetic",
        "config": {"workload": f"4096 2D patches 16^2-128^2 nodes, spline degree 1-3, {total} float32 markers "
                   "(log-uniform per patch, rescaled to ~2e8), one component, one persistent launch each way; "
                   "each patch's markers kept sorted by cell (sorted before timing)"},
        "roofline": {"bound": "hbm", "achieved": alg / ((sc + ga) * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": alg / ((sc + ga) * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None},
        "scatter_ms": sc, "gather_ms": ga, "markers": total, "autotuned_chunk_block_workgroups": list(tuned),
        "variants": {"input_order_scatter_ms": sc_in, "input_order_gather_ms": ga_in,
                     "input_order_ms_per_step": sc_in + ga_in},
        "cpu_baseline": cpu,
    }


def config5(torch, args) -> dict:
    """The differentiable 3D cubic gather of scatter on a 128^3 periodic grid: 2^24 float32
    markers, field = sum of 8 Gaussians; one forward (scatter, gather) and backward (gradients
    with respect to positions, weights and field)."""
    import mpifield_amd as m
    from mpifield_amd import autograd as AG

    nodes, n = 128, 1 << 24
    cell = 1.0 / nodes
    grid = m.BlockGrid.from_geometry([0.0] * 3, [cell] * 3, [nodes] * 3, [True] * 3)
    gen = torch.Generator(device="cuda").manual_seed(5)
    pos = torch.rand((3, n), generator=gen, device="cuda", dtype=torch.float32).requires_grad_(True)
    w = torch.ones((1, n), device="cuda", dtype=torch.float32).requires_grad_(True)
    axis = (torch.arange(nodes, device="cuda", dtype=torch.float64) + 0.0) * cell
    X, Y, Z = torch.meshgrid(axis, axis, axis, indexing="ij")
    field = torch.zeros((nodes, nodes, nodes, 1), dtype=torch.float64, device="cuda")
    for _ in range(8):
        c = torch.rand(3, generator=gen, device="cuda", dtype=torch.float64)
        sigma = 0.05 + 0.15 * torch.rand((), generator=gen, device="cuda", dtype=torch.float64)
        amp = torch.randn((), generator=gen, device="cuda", dtype=torch.float64)
        field[..., 0] += amp * torch.exp(-((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) / (2 * sigma**2))
    field.requires_grad_(True)

    def step():
        # the positions binned once per step and shared by the scatter, the gather and both
        # backward passes (the binning is inside the timed step: positions change every step)
        pos.grad = w.grad = field.grad = None
        bins = AG.bin_positions(pos, m.CUBIC, grid)
        rho = AG.scatter(pos, w, m.CUBIC, grid, bins=bins)
        g = AG.gather(pos, rho + field, m.CUBIC, grid, bins=bins)
        g.double().sum().backward()

    ms = time_events(torch, step, max(1, args.steps // 4), 1)
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as orc

        sample, cores = n, cpu_cores()
        p_h = pos[:, :sample].detach().double().cpu().numpy()
        f_h = field.detach().cpu().numpy().copy()
        r_h = np.zeros((1, sample))
        geo = (0.0, 0.0, 0.0, cell, cell, cell, nodes, nodes, nodes, 1, 1, 1)
        sec = cpu_seconds(lambda: (orc.scatter("spline", p_h, np.ones((1, sample)), f_h, *geo, degree=3, threads=cores),
                                   orc.gather("spline", p_h, f_h, r_h, *geo, degree=3, threads=cores)))
        cpu = {"value": sample / sec, "unit": "markers/s", "cores": cores, "kind": "port",
               "sample": f"all {sample} markers, forward only (scatter then gather; the reference has no "
                         f"differentiation): {sec:.2f} s"}
    return {
        "metric": "markers/s, differentiable 3D cubic gather(scatter) 128^3, 2^24 markers, forward+backward",
        "value": n / (ms * 1e-3), "unit": "markers/s", "n_gpus": 1, "ms_per_step": ms, "higher_is_better": True,
        "dtype": "int64", "data": "synthetic",
        "config": {"workload": f"128^3 periodic, {n} float32 markers, CUBIC, field = 8 Gaussians; positions binned "
                   "once per step, scatter, gather, backward to positions, weights and field"},
        "cpu_baseline": cpu,
    }


def main() -> None:
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the path runs on one MI355X; "
              "the split over GPUs by spatial blocks with halo folds is not built)", file=sys.stderr)
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

    if args.config != 2:
        line = {3: config3, 4: config4, 5: config5, 6: config6, 7: config7, 8: config8, 9: config9,
                10: config10}[args.config](
            torch, args)
        print(json.dumps(line), flush=True)
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
    bins = binning.bin_markers(pos, val, (1, 2), geometry).prepare_items()  # bins and the tiles' work lists
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
    roofline_valu = None
    if (n, nodes, comps) == (1 << 26, 256, 3) and HEADLINE_VALU.exists():
        # the dominant kernel against the VALU issue rate as well: lane operations per marker
        # counted by rocprofv3 (committed summary) over its HIP-event time
        vs = json.loads(HEADLINE_VALU.read_text())
        ops = vs["valu_lane_ops_per_marker"][dominant]
        peak = vs["peak_lane_ops_per_s"] / 1e9
        ach = ops * n / (dk["ms_per_step"] * 1e-3) / 1e9
        roofline_valu = {"bound": "valu", "kernel": dominant, "achieved": ach, "peak": peak, "unit": "G lane-ops/s",
                         "frac": ach / peak, "lane_ops_per_marker": ops,
                         "source": str(HEADLINE_VALU.relative_to(ROOT))}
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
        "roofline_valu": roofline_valu,
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

        # ---- re-binning inside the step: the bins' sorted markers moved by up to half a cell
        # per axis since they were binned (a pusher's step), binned again incrementally
        # (mpf_rebin_markers), results kept in bins order ----
        log("re-binning moved markers (incremental) inside the step")
        gen = torch.Generator(device=dev).manual_seed(5)
        moved = torch.remainder(bins.positions + (torch.rand(bins.positions.shape, generator=gen, device=dev) - 0.5)
                                / nodes, 1.0)

        def rebinned():
            b = bins.rebin(moved)
            binning.scatter_binned(b, field)
            binning.gather_binned(b, field, result)

        rb_ms, _ = time_steps(torch, rebinned, max(3, args.steps // 4), 1, stream)
        rbo_ms, _ = time_steps(torch, lambda: bins.rebin(moved), max(3, args.steps // 4), 1, stream)
        full_ms, _ = time_steps(torch, lambda: binning.bin_markers(moved, bins.values, (1, 2), geometry),
                                max(3, args.steps // 4), 1, stream)
        # the deposit of the re-binned markers against that of a fresh binning of them
        fresh = binning.bin_markers(moved, bins.values, (1, 2), geometry)
        again = bins.rebin(moved)
        f1, f2 = torch.zeros_like(field), torch.zeros_like(field)
        binning.scatter_binned(again, f1)
        binning.scatter_binned(fresh, f2)
        rebin_equal = bool(torch.equal(f1, f2) and torch.equal(again.start, fresh.start))
        del fresh, again, f1, f2, moved

        # ---- the drop-in API (Field.scatter + Field.gather, the reference's field.py:1097-1180):
        # positions given each call (the library bins inside both calls), and the positions'
        # bins made once with Field.bin_markers and passed instead ----
        log("Field.scatter + Field.gather (drop-in route)")
        import mpifield_amd as m

        fld = m.Field(3, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), nodes, num_components=comps, periodic=True)

        def api_positions():
            fld.scatter(pos, val, m.QUADRATIC)
            fld.gather(pos, m.QUADRATIC)

        api_ms, _ = time_steps(torch, api_positions, max(3, args.steps // 4), 1, stream)
        fbins = fld.bin_markers(pos, m.QUADRATIC)

        def api_bins():
            fld.scatter(fbins, val, m.QUADRATIC)
            fld.gather(fbins, m.QUADRATIC)

        apib_ms, _ = time_steps(torch, api_bins, max(3, args.steps // 4), 1, stream)
        del fld, fbins

        # ---- the plain float64-atomic one-thread-per-marker form, same markers: the direct
        # entry points (mpf_scatter in atomic mode, mpf_gather with the in-call binning off) ----
        def atomic_sorted_direct():
            with K.direct_kernels():
                K.scatter(1, 0, bins.positions, bins.values, field, geometry, mode=_lib.MODE_ATOMIC)
                K.gather(1, 0, bins.positions, field, result, geometry)

        def atomic_input_direct():
            with K.direct_kernels():
                K.scatter(1, 0, pos, val, field, geometry, mode=_lib.MODE_ATOMIC)
                K.gather(1, 0, pos, field, result, geometry)

        at_s, _ = time_steps(torch, atomic_sorted_direct, max(3, args.steps // 4), 1, stream)
        at_u, _ = time_steps(torch, atomic_input_direct, 3, 1, stream)
        line["variants"] = {
            "binned_sort_inside_step_markers_per_s": n / (sh_ms * 1e-3),
            "binned_sort_inside_step_ms": sh_ms,
            "rebinned_inside_step_ms": rb_ms,
            "rebinned_inside_step_markers_per_s": n / (rb_ms * 1e-3),
            "rebin_only_ms": rbo_ms,
            "full_sort_of_moved_markers_ms": full_ms,
            "rebin_deposit_equals_fresh_binning": rebin_equal,
            "rebin_workload": "the binned markers moved by uniform [-0.5, 0.5] cell per axis (periodic wrap)",
            "field_api_positions_step_ms": api_ms,
            "field_api_bins_kept_step_ms": apib_ms,
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
