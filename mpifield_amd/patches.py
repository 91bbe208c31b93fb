"""Many refined-mesh patches deposited and gathered in one persistent launch each.

The reference runs one `Weighting.scatter_block` per patch, each on its own
``BlockGrid.from_geometry(lower, cell_size, num_nodes, periodic)`` (the caller loops over
patches). `PatchBatch` holds the whole batch -- every patch's geometry, weighting (linear,
quadratic or a spline degree 0..3) and its slices of one flat marker array and one flat
field buffer -- and runs it with `mpf_patches_scatter` / `mpf_patches_gather`:

    batch = PatchBatch(ndim=2, ncomp=1)
    for lower, cell, num, weighting, n in ...:
        batch.add(lower, cell, num, weighting, n)      # markers and field slices in order
    batch.build(positions.shape[1])
    batch.scatter(positions, values, fields)           # one persistent launch + finalize
    batch.gather(positions, fields, result)

`autotune` times the work-item size (markers per item) and the workgroup shape on the
batch itself and keeps the fastest per (ndim, degrees, ncomp); `capture` records the
deposit in a HIP graph for replay.
"""

from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from mpifield_amd import _lib
from mpifield_amd import kernels as K
from mpifield_amd.binning import kind_of

#: the tuned (chunk, block, workgroups) per (ndim, degrees, ncomp), filled by autotune()
TUNED: dict[tuple, tuple[int, int, int]] = {}
DEFAULT = (8192, 512, 0)


@dataclass
class PatchBatch:
    ndim: int
    ncomp: int = 1
    patches: list = field(default_factory=list)
    field_values: int = 0
    markers: int = 0
    plan: int | None = None
    config: tuple[int, int, int] = DEFAULT

    def add(self, lower, cell_size, num_nodes, weighting, num_markers: int, periodic=None) -> tuple[int, slice]:
        """Append a patch whose markers are the next ``num_markers`` of the batch; returns
        (field offset, marker slice)."""
        if len(num_nodes) != self.ndim:
            raise ValueError(f"a {self.ndim}D batch takes {self.ndim}D patches")
        kind, degree = kind_of(weighting)
        p = _lib.Patch()
        per = (False,) * self.ndim if periodic is None else tuple(bool(x) for x in periodic)
        for d in range(self.ndim):
            p.lower[d] = float(lower[d])
            p.cell_size[d] = float(cell_size[d])
            p.num_nodes[d] = int(num_nodes[d])
            p.code[d] = int(per[d])
        p.kind, p.degree = kind, degree
        p.first_marker, p.num_markers = self.markers, int(num_markers)
        p.field_offset = self.field_values
        self.patches.append(p)
        offset = self.field_values
        self.field_values += int(np.prod(num_nodes)) * self.ncomp
        first = self.markers
        self.markers += int(num_markers)
        return offset, slice(first, self.markers)

    def field_view(self, fields: torch.Tensor, index: int) -> torch.Tensor:
        """Patch ``index``'s field, ``num_nodes + (ncomp,)``, as a view of the flat buffer."""
        p = self.patches[index]
        shape = tuple(int(p.num_nodes[d]) for d in range(self.ndim)) + (self.ncomp,)
        size = int(np.prod(shape))
        return fields[p.field_offset : p.field_offset + size].view(shape)

    def key(self) -> tuple:
        degrees = tuple(sorted({(p.kind, p.degree) for p in self.patches}))
        return (self.ndim, degrees, self.ncomp)

    def build(self, chunk: int | None = None) -> None:
        """Upload the patch table and the work list (items of at most ``chunk`` markers)."""
        self.close()
        tuned = TUNED.get(self.key())
        if tuned is not None:
            self.config = tuned
        chunk = self.config[0] if chunk is None else chunk
        array = (_lib.Patch * max(1, len(self.patches)))(*self.patches)
        lib = _lib.load()
        plan = lib.mpf_patch_plan_create(self.ndim, len(self.patches), ctypes.addressof(array), chunk, self.ncomp)
        if not plan:
            raise ValueError(lib.mpf_last_error().decode())
        self.plan = plan
        self.config = (chunk, self.config[1], self.config[2])

    def close(self) -> None:
        if self.plan:
            _lib.load().mpf_patch_plan_destroy(self.plan)
            self.plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def scatter(self, positions: torch.Tensor, values: torch.Tensor, fields: torch.Tensor) -> None:
        """Add every patch's markers into its field (fields: the flat float64 buffer)."""
        if self.plan is None:
            self.build()
        pos, ps = K.marker_rows(positions, self.ndim)
        val, vs = K.marker_rows(values, self.ncomp, pos.dtype)
        fields = K._field(fields, "fields")
        if fields.numel() < self.field_values or pos.shape[1] < self.markers:
            raise ValueError("fields or markers smaller than the batch")
        _, block, workgroups = self.config
        with torch.cuda.device(fields.device):
            _lib.call("mpf_patches_scatter", self.plan, pos.data_ptr(), ps, val.data_ptr(), vs, K._dtype_code(pos),
                      self.ncomp, fields.data_ptr(), block, workgroups, None, K.stream_handle(fields.device))

    def gather(self, positions: torch.Tensor, fields: torch.Tensor, result: torch.Tensor) -> None:
        """Interpolate every patch's field at its markers into ``result`` (ncomp, N)."""
        if self.plan is None:
            self.build()
        pos, ps = K.marker_rows(positions, self.ndim)
        fields = K._field(fields, "fields")
        if result.dtype != pos.dtype or result.stride(-1) != 1:
            raise ValueError("result must have the positions' dtype and unit element stride")
        rs = result.stride(0) if result.dim() == 2 and result.shape[0] > 1 else result.shape[-1]
        with torch.cuda.device(fields.device):
            _lib.call("mpf_patches_gather", self.plan, pos.data_ptr(), ps, K._dtype_code(pos), fields.data_ptr(),
                      self.ncomp, result.data_ptr(), rs, 0, None, K.stream_handle(fields.device))

    def sort(self, positions: torch.Tensor, values: torch.Tensor | None = None):
        """The batch's markers ordered by cell within each patch, the order a particle code
        keeps between steps: returns (positions, values, order) with column i of the result
        column order[i] of the input.  Each patch's markers stay within its range, so the
        plan serves the sorted arrays as they are; the deposit gives the same integers in
        any order, and on sorted markers a work-item sums runs of one cell in registers."""
        pos, _ = K.marker_rows(positions, self.ndim)
        dev = pos.device
        counts = torch.as_tensor([int(p.num_markers) for p in self.patches], device=dev)
        patch_of = torch.repeat_interleave(torch.arange(len(self.patches), device=dev), counts)
        lower = torch.as_tensor([[p.lower[d] for d in range(self.ndim)] for p in self.patches],
                                dtype=torch.float64, device=dev)[patch_of]
        cell = torch.as_tensor([[p.cell_size[d] for d in range(self.ndim)] for p in self.patches],
                               dtype=torch.float64, device=dev)[patch_of]
        nodes = torch.as_tensor([[p.num_nodes[d] for d in range(self.ndim)] for p in self.patches],
                                dtype=torch.int64, device=dev)
        sizes = nodes.prod(dim=1)
        offsets = torch.cumsum(sizes, 0) - sizes
        num = nodes[patch_of]
        key = offsets[patch_of]
        stride = torch.ones(len(patch_of), dtype=torch.int64, device=dev)
        for d in reversed(range(self.ndim)):
            rel = (pos[d, : self.markers].double() - lower[:, d]) / cell[:, d]
            c = torch.nan_to_num(torch.floor(rel), nan=0.0).clamp(0, None).to(torch.int64)
            c = torch.minimum(c, num[:, d] - 1)
            key = key + c * stride
            stride = stride * num[:, d]
        order = torch.argsort(key, stable=True)
        out_pos = positions[:, : self.markers][:, order].contiguous()
        out_val = None if values is None else values[:, : self.markers][:, order].contiguous()
        return out_pos, out_val, order

    def capture(self, positions, values, fields) -> torch.cuda.CUDAGraph:
        """Record one deposit of the batch in a HIP graph; replay with ``graph.replay()``.

        The library's scratch belongs to the (device, stream) pair, so the deposit runs once
        on the capture stream first: the scratch then exists and the captured call allocates
        nothing.  ``fields`` gets that deposit as well as the captured one."""
        graph = torch.cuda.CUDAGraph()
        stream = torch.cuda.Stream()
        stream.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(stream):
            self.scatter(positions, values, fields)
            stream.synchronize()
            with torch.cuda.graph(graph, stream=stream):
                self.scatter(positions, values, fields)
        torch.cuda.current_stream().wait_stream(stream)
        return graph


def autotune(batch: PatchBatch, positions, values, fields, candidates=None, repeats: int = 3) -> tuple:
    """Time (chunk, block, workgroups) candidates on the batch; keep and return the fastest."""
    if candidates is None:
        candidates = [(c, b, w) for c in (2048, 8192, 32768) for b in (256, 512) for w in (512, 1024)]
    scratch = torch.zeros_like(fields)
    best, best_t = None, float("inf")
    for chunk, block, workgroups in candidates:
        batch.config = (chunk, block, workgroups)
        batch.build(chunk)
        batch.scatter(positions, values, scratch)  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(repeats):
            batch.scatter(positions, values, scratch)
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / repeats
        if t < best_t:
            best, best_t = (chunk, block, workgroups), t
    TUNED[batch.key()] = best
    batch.config = best
    batch.build(best[0])
    return best
