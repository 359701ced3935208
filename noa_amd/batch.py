"""Batched extraction and insertion: several calls in one persistent launch each, replayable from a HIP graph.

A call is described exactly as for ``noa_amd.xform.extract_central_slices_3d`` / ``insert_central_slices_3d``
(same arguments, same checks, same results); a batch of calls runs as one launch of the library's persistent kernels
(``noa_hip_extract_batch_*`` / ``noa_hip_insert_batch_*``). The launch grid does not depend on the batch, so the
two launches are captured once in a HIP graph (``torch.cuda.CUDAGraph``) and replayed for every later batch: a batch
is prepared on the host into pinned memory, copied to the slot's device image, and the graph is replayed, all on the
slot's stream. No device memory is allocated per batch.

noa has no counterpart: its only asynchrony is ``Stream::enqueue`` (src/noa/runtime/cuda/Stream.hpp:205-239).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import torch

from . import abi
from . import xform as nx
from ._lib import last_error, lib
from .xform import VolumeBlocks  # noqa: F401  (a prepared volume accepted by extract_call)


@dataclass
class Call:
    """One prepared call: the C descriptor and the objects that keep its host-side inputs alive."""
    args: abi.ProjectArgs
    keep: list = field(default_factory=list)


def extract_call(remap, volume, volume_shape, slices, slice_shape, fwd_rotation,
                 options: nx.ExtractCentralSlicesOptions | None = None) -> Call:
    """The descriptor of extract_central_slices_3d(remap, volume, None, volume_shape, slices, None, slice_shape,
    None, fwd_rotation, options). `volume` may be a VolumeBlocks (the volume's block copy, made once)."""
    opts = options or nx.ExtractCentralSlicesOptions()
    keep: list = []
    args = abi.default_args()
    interp = nx._interp_of(volume, None, opts.interp)
    slice_shape = nx._shape4(slice_shape, "slice_shape")
    nx._fill(args, remap=remap, inp=volume, inp_w=None, out=slices, out_w=None, in_shape=nx._shape4(volume_shape, "volume_shape"),
             out_shape=slice_shape, rotation=fwd_rotation, scaling=None, count=slice_shape[0], opts=opts, keep=keep,
             allow_value=False, interp=interp)
    args.fftfreq_sinc = float(opts.w_windowed_sinc.fftfreq_sinc)
    args.fftfreq_blackman = float(opts.w_windowed_sinc.fftfreq_blackman)
    return Call(args, keep)


def insert_call(remap, slices, slice_weights, slice_shape, volume, volume_weight, volume_shape, inv_rotation,
                options: nx.InsertCentralSlicesOptions | None = None) -> Call:
    """The descriptor of insert_central_slices_3d(remap, slices, slice_weights, slice_shape, volume, volume_weight,
    volume_shape, None, inv_rotation, options)."""
    opts = options or nx.InsertCentralSlicesOptions()
    keep: list = []
    args = abi.default_args()
    interp = nx._interp_of(slices, slice_weights, opts.interp)
    slice_shape = nx._shape4(slice_shape, "slice_shape")
    nx._fill(args, remap=remap, inp=slices, inp_w=slice_weights, out=volume, out_w=volume_weight, in_shape=slice_shape,
             out_shape=nx._shape4(volume_shape, "volume_shape"), rotation=inv_rotation, scaling=None,
             count=slice_shape[0], opts=opts, keep=keep, allow_value=True, interp=interp)
    args.fftfreq_sinc = float(opts.windowed_sinc.fftfreq_sinc)
    args.fftfreq_blackman = float(opts.windowed_sinc.fftfreq_blackman)
    return Call(args, keep)


def _prepare(fn_name: str, calls: list, host: torch.Tensor) -> int:
    arr = (abi.ProjectArgs * max(1, len(calls)))(*[c.args for c in calls])
    need = ctypes.c_int64(0)
    status = getattr(lib(), fn_name)(arr, len(calls), ctypes.c_void_p(host.data_ptr()), host.numel(),
                                     ctypes.byref(need))
    if status != abi.OK:
        raise nx.NoaError(f"{last_error()} (image of {need.value} bytes, capacity {host.numel()})")
    return int(need.value)


class BatchSlot:
    """Host (pinned) and device images of one batch, the stream it runs on, and its graph.

    capacity: bytes of each image (the insertion image holds the calls' parameters and every work item: about
    16 bytes per active sub-brick of 8x8x16 voxels, e.g. 80 kB per 256^3 call).
    """

    def __init__(self, device: torch.device | str = "cuda", capacity: int = 8 << 20, graph: bool = True):
        self.device = torch.device(device)
        self.capacity = int(capacity)
        self.host_extract = torch.empty(self.capacity, dtype=torch.uint8, pin_memory=True)
        self.host_insert = torch.empty(self.capacity, dtype=torch.uint8, pin_memory=True)
        self.dev_extract = torch.zeros(self.capacity, dtype=torch.uint8, device=self.device)
        self.dev_insert = torch.zeros(self.capacity, dtype=torch.uint8, device=self.device)
        self.stream = torch.cuda.Stream(self.device)          # default priority
        self.bytes_extract = 0
        self.bytes_insert = 0
        self.graph: torch.cuda.CUDAGraph | None = None
        if graph:
            self._capture()

    def _launch(self) -> None:
        s = ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
        if lib().noa_hip_insert_batch_launch(ctypes.c_void_p(self.dev_insert.data_ptr()), s) != abi.OK:
            raise nx.NoaError(last_error())
        if lib().noa_hip_extract_batch_launch(ctypes.c_void_p(self.dev_extract.data_ptr()), s) != abi.OK:
            raise nx.NoaError(last_error())

    def _capture(self) -> None:
        # Empty images (zero calls) while capturing; a capture records the launches, it does not run them.
        self.prepare([], [])
        with torch.cuda.stream(self.stream):
            self.upload()
        self.stream.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=self.stream):
            self._launch()
        self.graph = g

    def prepare(self, extract_calls: list, insert_calls: list) -> None:
        """Writes both images into pinned host memory (nothing is launched). The slot must be idle."""
        self.bytes_extract = _prepare("noa_hip_extract_batch_prepare", extract_calls, self.host_extract)
        self.bytes_insert = _prepare("noa_hip_insert_batch_prepare", insert_calls, self.host_insert)

    def upload(self) -> None:
        """Copies the images to the device on the current stream."""
        self.dev_extract[: self.bytes_extract].copy_(self.host_extract[: self.bytes_extract], non_blocking=True)
        self.dev_insert[: self.bytes_insert].copy_(self.host_insert[: self.bytes_insert], non_blocking=True)

    def run(self) -> torch.cuda.Event:
        """Uploads the prepared batch and runs it (graph replay, or direct launches) on the slot's stream; returns
        an event recorded after it."""
        with torch.cuda.stream(self.stream):
            self.upload()
            if self.graph is not None:
                self.graph.replay()
            else:
                self._launch()
            ev = torch.cuda.Event()
            ev.record(self.stream)
        return ev


def run_batch(extract_calls: list, insert_calls: list, slot: BatchSlot | None = None) -> torch.cuda.Event:
    """Runs one batch (after the work already on the current stream) and returns its completion event."""
    slot = slot or BatchSlot(graph=False)
    slot.stream.wait_stream(torch.cuda.current_stream(slot.device))
    slot.prepare(extract_calls, insert_calls)
    return slot.run()
