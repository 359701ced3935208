"""Continuous-batching request loop for central-slice projection (extraction) and backprojection (insertion).

Reference volumes stay resident in HBM. Requests arrive over time. Each request's output (extracted slices, or the
accumulator volume and weights of an insertion) is carved from one preallocated device pool: the pool is the HBM
budget, and a request is admitted when its output fits (a request larger than the whole pool runs alone, with its
own allocation, when nothing else is in flight). Every step, the admitted requests that are not running yet become
one batch: all their extraction calls run as one persistent launch and all their insertion calls as another
(noa_amd.batch), replayed from a HIP graph captured once per slot. Several slots (each its own stream, images and
graph) keep batches in flight while the next one is formed, so requests that arrive while others are on the GPU
are admitted as soon as memory and a slot free up: batching is continuous, not a queue drained in rounds. Nothing
is allocated on the device per request or per batch (outputs come from the pool, accumulators are cleared there).

Results do not depend on the mix: an extracted slice is computed independently of the other work of its launch,
and an insertion request accumulates into its own volume with a deterministic kernel. A request completes when the
event recorded after its batch has passed; its latency is completion minus arrival.
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field

import torch

from . import batch as nb
from . import xform as nx

ALIGN = 256


@dataclass
class Request:
    """One projection ("extract") or backprojection ("insert") request.

    rotations: (K, 3, 3) float32 on the device (or a Mat33Array); forward rotations for extraction, inverse
    rotations for insertion. slices, weights: insertion inputs, (K, 1, box, box//2+1) complex64 and float32 on the
    device ("h" layout).
    """
    rid: int
    kind: str
    box: int
    rotations: torch.Tensor | nx.Mat33Array
    volume: int = 0
    slices: torch.Tensor | None = None
    weights: torch.Tensor | None = None
    arrival: float = 0.0
    result: torch.Tensor | None = None
    result_weight: torch.Tensor | None = None
    admitted: float | None = None
    done: float | None = None
    ready: torch.cuda.Event | None = None       # recorded at submit() on the stream that produced the inputs
    offset: int | None = None                   # pool offset of the output (None: dedicated allocation)

    @property
    def count(self) -> int:
        r = self.rotations
        return int(r.count if isinstance(r, nx.Mat33Array) else r.shape[0])

    def output_bytes(self) -> int:
        wr = self.box // 2 + 1
        if self.kind == "extract":
            return _aligned(self.count * self.box * wr * 8)
        return _aligned(self.box * self.box * wr * 8) + _aligned(self.box * self.box * wr * 4)


def _aligned(n: int) -> int:
    return (n + ALIGN - 1) // ALIGN * ALIGN


class DevicePool:
    """First-fit allocator over one preallocated device buffer (offsets aligned to 256 bytes)."""

    def __init__(self, nbytes: int, device: torch.device):
        self.size = _aligned(int(nbytes))
        self.buf = torch.empty(self.size, dtype=torch.uint8, device=device)
        self.free = [(0, self.size)]

    def alloc(self, nbytes: int) -> int | None:
        for i, (off, size) in enumerate(self.free):
            if size >= nbytes:
                if size == nbytes:
                    self.free.pop(i)
                else:
                    self.free[i] = (off + nbytes, size - nbytes)
                return off
        return None

    def release(self, off: int, nbytes: int) -> None:
        self.free.append((off, nbytes))
        self.free.sort()
        merged = []
        for o, s in self.free:
            if merged and merged[-1][0] + merged[-1][1] == o:
                merged[-1] = (merged[-1][0], merged[-1][1] + s)
            else:
                merged.append((o, s))
        self.free = merged

    def view(self, off: int, shape, dtype: torch.dtype) -> torch.Tensor:
        nbytes = torch.Size(shape).numel() * torch.empty((), dtype=dtype).element_size()
        return self.buf[off:off + nbytes].view(dtype).view(shape)


@dataclass
class _Launch:
    requests: list
    event: torch.cuda.Event
    slot: nb.BatchSlot
    calls: list                                  # keeps the calls' host objects alive until the batch has run


@dataclass
class ServeStats:
    latencies_ms: list = field(default_factory=list)
    slices: int = 0
    launches: int = 0
    merged: int = 0
    batch_sizes: list = field(default_factory=list)

    def percentile(self, q: float) -> float:
        if not self.latencies_ms:
            return float("nan")
        xs = sorted(self.latencies_ms)
        k = min(len(xs) - 1, max(0, int(round(q / 100.0 * (len(xs) - 1)))))
        return xs[k]


class Server:
    """volumes: {box: [complex64 (1, box, box, box//2+1) "hc" volumes on the device]}.

    budget_bytes: size of the output pool. n_streams: batch slots (each with its own stream and graph).
    max_batch_slices: slices above which a batch takes no further request.
    volume_blocks: extract from block copies of the resident volumes (noa_amd.xform.VolumeBlocks, made once here:
    the gather reads fewer cache lines than from the row-major volumes, with the same results).
    keep_results: copy each finished request's outputs out of the pool (into req.result / req.result_weight) before
    the pool space is reused; otherwise the outputs are only valid inside `on_done`.
    """

    def __init__(self, volumes: dict, budget_bytes: int, n_streams: int = 3, device: torch.device | str = "cuda",
                 keep_results: bool = True, on_done=None, graphs: bool = True, max_batch_slices: int = 8192,
                 volume_blocks: bool = True):
        self.device = torch.device(device)
        self.volumes = volumes
        # What the extraction calls read: the volumes' block copies, or the volumes themselves.
        self.sources = ({box: [nx.VolumeBlocks(v) for v in vols] for box, vols in volumes.items()}
                        if volume_blocks else volumes)
        self.budget = int(budget_bytes)
        self.pool = DevicePool(self.budget, self.device)
        self.in_use = 0
        self.oversize = 0                           # admitted requests larger than the pool (run alone)
        self.slots = [nb.BatchSlot(self.device, graph=graphs) for _ in range(max(1, n_streams))]
        self.free_slots = deque(self.slots)
        self.pending: deque[Request] = deque()
        self.admitted: list[Request] = []
        self.running: list[_Launch] = []
        self.keep_results = keep_results
        self.max_batch_slices = int(max_batch_slices)
        self.on_done = on_done
        self.stats = ServeStats()
        self.t0 = time.perf_counter()

    def now(self) -> float:
        return time.perf_counter() - self.t0

    def submit(self, req: Request) -> None:
        """Queues a request. Its inputs may still be in flight on the caller's current stream: an event recorded
        there now orders the batch that reads them after their producer."""
        if not isinstance(req.rotations, nx.Mat33Array):
            req.rotations = nx.Mat33Array(req.rotations, device=self.device)
        req.ready = torch.cuda.Event()
        req.ready.record(torch.cuda.current_stream(self.device))
        self.pending.append(req)

    def idle(self) -> bool:
        return not self.pending and not self.admitted and not self.running

    # -- admission: FIFO while the output fits in the pool; a request larger than the pool runs alone.
    def _admit(self) -> None:
        while self.pending:
            req = self.pending[0]
            need = req.output_bytes()
            if self.oversize:
                break                               # a request larger than the pool is running alone
            if need > self.pool.size:
                if self.in_use > 0 or self.running or self.admitted:
                    break
                req.offset = None
                self.oversize += 1
            else:
                off = self.pool.alloc(need)
                if off is None:
                    break
                req.offset = off
            self.pending.popleft()
            self.in_use += need
            req.admitted = self.now()
            self.admitted.append(req)

    def _outputs(self, req: Request):
        wr = req.box // 2 + 1
        if req.kind == "extract":
            shape = (req.count, 1, req.box, wr)
            if req.offset is None:
                return torch.empty(shape, dtype=torch.complex64, device=self.device), None
            return self.pool.view(req.offset, shape, torch.complex64), None
        shape = (1, req.box, req.box, wr)
        vbytes = _aligned(req.box * req.box * wr * 8)
        if req.offset is None:
            return (torch.zeros(shape, dtype=torch.complex64, device=self.device),
                    torch.zeros(shape, dtype=torch.float32, device=self.device))
        return (self.pool.view(req.offset, shape, torch.complex64),
                self.pool.view(req.offset + vbytes, shape, torch.float32))

    # -- launches: every admitted request joins one batch on a free slot.
    def _launch(self) -> None:
        if not self.admitted or not self.free_slots:
            return
        slot = self.free_slots.popleft()
        # A batch takes admitted requests in order up to max_batch_slices (at least one request); the rest wait for
        # the next free slot, so that a burst becomes several pipelined batches rather than one long one.
        n, take = 0, 0
        for r in self.admitted:
            if take and n + r.count > self.max_batch_slices:
                break
            n += r.count
            take += 1
        reqs, self.admitted = self.admitted[:take], self.admitted[take:]
        ext, ins = [], []
        with torch.cuda.stream(slot.stream):
            for r in reqs:
                if r.ready is not None:
                    slot.stream.wait_event(r.ready)
                out, wout = self._outputs(r)
                r.result, r.result_weight = out, wout
                box = r.box
                if r.kind == "extract":
                    ext.append(nb.extract_call("hc2hc", self.sources[box][r.volume], (1, box, box, box), out,
                                               (r.count, 1, box, box), r.rotations))
                else:
                    out.zero_()
                    wout.zero_()
                    ins.append(nb.insert_call("h2h", r.slices, r.weights, (r.count, 1, box, box), out, wout,
                                              (1, box, box, box), r.rotations))
        slot.prepare(ext, ins)
        ev = slot.run()
        self.running.append(_Launch(reqs, ev, slot, ext + ins))
        self.stats.launches += 1
        self.stats.merged += len(reqs) - 1
        self.stats.batch_sizes.append(len(reqs))

    def _finish(self, r: Request) -> None:
        if self.on_done is not None:
            self.on_done(r)
        if r.offset is not None:
            if self.keep_results:
                # Copied before the pool space can be handed to a later batch on another stream.
                r.result = r.result.clone()
                if r.result_weight is not None:
                    r.result_weight = r.result_weight.clone()
                torch.cuda.current_stream(self.device).synchronize()
            else:
                r.result = r.result_weight = None
            self.pool.release(r.offset, r.output_bytes())
            r.offset = None
        else:
            self.oversize -= 1
        self.in_use -= r.output_bytes()

    def _poll(self) -> int:
        finished = 0
        still = []
        for launch in self.running:
            if launch.event.query():
                t = self.now()
                for r in launch.requests:
                    r.done = t
                    self.stats.latencies_ms.append((t - r.arrival) * 1e3)
                    self.stats.slices += r.count
                    self._finish(r)
                self.free_slots.append(launch.slot)
                finished += len(launch.requests)
            else:
                still.append(launch)
        self.running = still
        return finished

    def step(self) -> int:
        """Admit, launch, and retire what has finished; returns the number of requests completed."""
        self._admit()
        self._launch()
        return self._poll()


def run_open_loop(server: Server, requests: list, timeout_s: float = 120.0) -> ServeStats:
    """Submits each request at its arrival time (seconds from now) and steps the server until all are done."""
    order = sorted(requests, key=lambda r: r.arrival)
    server.t0 = time.perf_counter()
    i = 0
    done = 0
    while done < len(order):
        now = server.now()
        while i < len(order) and order[i].arrival <= now:
            server.submit(order[i])
            i += 1
        done += server.step()
        if server.now() > timeout_s:
            raise RuntimeError(f"serving did not finish within {timeout_s} s ({done}/{len(order)} requests done)")
        if i < len(order) and server.idle():
            # Nothing to do until the next arrival.
            wait = order[i].arrival - server.now()
            if wait > 0:
                time.sleep(min(wait, 1e-3))
    return server.stats


def run_alone(server: Server, req: Request) -> Request:
    """The same request through the same code with nothing else in flight (for mix-independence checks)."""
    server.submit(req)
    while not server.idle():
        server.step()
    torch.cuda.synchronize(server.device)
    return req
