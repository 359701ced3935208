"""Continuous-batching request loop for central-slice projection (extraction) and backprojection (insertion).

Reference volumes stay resident in HBM. Requests arrive over time; the loop admits them while an HBM budget for
their outputs allows, merges the admitted extraction requests that read the same volume into one launch, gives each
insertion request its own accumulator, and launches on several streams so that independent work overlaps on the
device. A request completes when the event recorded after its launch has passed; its latency is completion time
minus arrival time. Requests that arrive while others run are admitted as soon as budget frees up: batching is
continuous, not a queue drained in rounds.

Results do not depend on the mix: an extraction slice is computed independently of the other slices of its launch,
and an insertion request accumulates into its own volume with a deterministic kernel.
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field

import torch

from . import xform as nx


@dataclass
class Request:
    """One projection ("extract") or backprojection ("insert") request.

    rotations: (K, 3, 3) float32 on the device; forward rotations for extraction, inverse rotations for insertion.
    slices, weights: insertion inputs, (K, 1, box, box//2+1) complex64 and float32 on the device ("h" layout).
    """
    rid: int
    kind: str
    box: int
    rotations: torch.Tensor
    volume: int = 0
    slices: torch.Tensor | None = None
    weights: torch.Tensor | None = None
    arrival: float = 0.0
    result: torch.Tensor | None = None
    result_weight: torch.Tensor | None = None
    admitted: float | None = None
    done: float | None = None
    ready: torch.cuda.Event | None = None       # recorded at submit() on the stream that produced the inputs

    @property
    def count(self) -> int:
        return int(self.rotations.shape[0])

    def output_bytes(self) -> int:
        wr = self.box // 2 + 1
        if self.kind == "extract":
            return self.count * self.box * wr * 8
        return self.box * self.box * wr * (8 + 4)


@dataclass
class _Launch:
    requests: list
    event: torch.cuda.Event
    nbytes: int


@dataclass
class ServeStats:
    latencies_ms: list = field(default_factory=list)
    slices: int = 0
    launches: int = 0
    merged: int = 0

    def percentile(self, q: float) -> float:
        if not self.latencies_ms:
            return float("nan")
        xs = sorted(self.latencies_ms)
        k = min(len(xs) - 1, max(0, int(round(q / 100.0 * (len(xs) - 1)))))
        return xs[k]


class Server:
    """volumes: {box: [complex64 (1, box, box, box//2+1) "hc" volumes on the device]}."""

    def __init__(self, volumes: dict, budget_bytes: int, n_streams: int = 3, max_merge_slices: int = 4096,
                 device: torch.device | str = "cuda"):
        self.device = torch.device(device)
        self.volumes = volumes
        self.budget = int(budget_bytes)
        self.in_use = 0
        self.max_merge = max_merge_slices
        self.streams = [torch.cuda.Stream(self.device) for _ in range(n_streams)]   # default priority
        self.next_stream = 0
        self.pending: deque[Request] = deque()
        self.admitted: list[Request] = []
        self.running: list[_Launch] = []
        self.stats = ServeStats()
        self.t0 = time.perf_counter()

    def now(self) -> float:
        return time.perf_counter() - self.t0

    def submit(self, req: Request) -> None:
        """Queues a request. Its inputs may still be in flight on the caller's current stream: an event recorded
        there now orders every launch that reads them after their producer."""
        req.ready = torch.cuda.Event()
        req.ready.record(torch.cuda.current_stream(self.device))
        self.pending.append(req)

    @staticmethod
    def _inputs_on(stream: torch.cuda.Stream, reqs) -> None:
        """Makes `stream` wait for each request's inputs, and ties their memory to `stream` so that the caching
        allocator does not reuse it while the launch may still read it."""
        for r in reqs:
            if r.ready is not None:
                stream.wait_event(r.ready)
            for t in (r.rotations, r.slices, r.weights):
                if t is not None and t.is_cuda:
                    t.record_stream(stream)

    def idle(self) -> bool:
        return not self.pending and not self.admitted and not self.running

    # -- admission: FIFO while the budget allows; a request larger than the budget runs alone.
    def _admit(self) -> None:
        while self.pending:
            req = self.pending[0]
            need = req.output_bytes()
            if self.in_use + need > self.budget and (self.in_use > 0 or self.running):
                break
            self.pending.popleft()
            self.in_use += need
            req.admitted = self.now()
            self.admitted.append(req)

    def _stream(self) -> torch.cuda.Stream:
        s = self.streams[self.next_stream]
        self.next_stream = (self.next_stream + 1) % len(self.streams)
        return s

    # -- launches: extraction requests on the same volume are merged (up to max_merge slices).
    def _launch(self) -> None:
        if not self.admitted:
            return
        groups: dict = {}
        inserts = []
        for req in self.admitted:
            if req.kind == "extract":
                groups.setdefault((req.box, req.volume), []).append(req)
            else:
                inserts.append(req)
        self.admitted = []
        for (box, vid), reqs in groups.items():
            batch: list[Request] = []
            n = 0
            for req in reqs + [None]:
                if req is not None and (n + req.count <= self.max_merge or not batch):
                    batch.append(req)
                    n += req.count
                    continue
                if batch:
                    self._launch_extract(box, vid, batch, n)
                batch, n = ([req], req.count) if req is not None else ([], 0)
        for req in inserts:
            self._launch_insert(req)

    def _launch_extract(self, box: int, vid: int, reqs: list, n: int) -> None:
        wr = box // 2 + 1
        stream = self._stream()
        self._inputs_on(stream, reqs)
        with torch.cuda.stream(stream):
            rot = torch.cat([r.rotations for r in reqs]) if len(reqs) > 1 else reqs[0].rotations
            out = torch.empty((n, 1, box, wr), dtype=torch.complex64, device=self.device)
            nx.extract_central_slices_3d("hc2hc", self.volumes[box][vid], None, (1, box, box, box), out, None,
                                         (n, 1, box, box), None, rot)
            ev = torch.cuda.Event()
            ev.record(stream)
        off = 0
        for r in reqs:
            r.result = out[off:off + r.count]
            off += r.count
        self.running.append(_Launch(reqs, ev, sum(r.output_bytes() for r in reqs)))
        self.stats.launches += 1
        self.stats.merged += len(reqs) - 1

    def _launch_insert(self, req: Request) -> None:
        box, wr = req.box, req.box // 2 + 1
        stream = self._stream()
        self._inputs_on(stream, [req])
        with torch.cuda.stream(stream):
            vol = torch.zeros((1, box, box, wr), dtype=torch.complex64, device=self.device)
            wvol = torch.zeros((1, box, box, wr), dtype=torch.float32, device=self.device)
            nx.insert_central_slices_3d("h2h", req.slices, req.weights, (req.count, 1, box, box), vol, wvol,
                                        (1, box, box, box), None, req.rotations)
            ev = torch.cuda.Event()
            ev.record(stream)
        req.result, req.result_weight = vol, wvol
        self.running.append(_Launch([req], ev, req.output_bytes()))
        self.stats.launches += 1

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
                self.in_use -= launch.nbytes
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
