"""Launch autotuner: measures candidate launch parameters on the device and fills the library's tuning table.

noa fixes its launch shapes at compile time (``IwiseOptions``: block size and elements per thread,
src/noa/runtime/Iwise.hpp:14-60; the CUDA iwise kernels reshape every block to 32x16, runtime/cuda/Iwise.cuh:94-175).
Here the kernels take their tile shapes from a table keyed by (operator, box, layout, interpolation, volume type),
``noa_hip_set_tuning``; this module times the candidates of each key with HIP events on the launch stream and keeps
the fastest. Every candidate computes the same result (the parameters only change the work decomposition), so
tuning never changes an output bit.

Operators and their parameters (``values``):
  OP_EXTRACT        volume layout (1 row-major gather, 2 block-copy gather, 3 LDS tiles), crossing slices per tile item
  OP_INSERT         kernel (1 one-wave, 2 slab where allowed), expected slices per work item, x taps of a slab row
                    (0 auto: one load, 2: two loads), item order (0 by work, 1 by first slice, 2 XCD-aware, 3
                    pose-adaptive: segments cut from the passing slices counted on the device)
  OP_EXTRACT_BATCH  pixels per lane (1-4), workgroups of the persistent grid
  OP_INSERT_BATCH   workgroups of the persistent grid
"""
from __future__ import annotations

import ctypes
import itertools
import json
from dataclasses import asdict, dataclass
from pathlib import Path
from typing import Callable, Iterable

import torch

from . import abi
from ._lib import last_error, lib

OPS = {"extract": abi.OP_EXTRACT, "insert": abi.OP_INSERT, "extract_batch": abi.OP_EXTRACT_BATCH,
       "insert_batch": abi.OP_INSERT_BATCH}

# Candidate values per operator (cartesian products).
SPACE = {
    "extract": {"layout": [1, 2, 3], "pairs": [64, 128, 256]},
    "insert": {"kernel": [1, 2], "list": [256, 512, 1024], "paired": [0], "order": [0, 2, 3]},
    "extract_batch": {"ppt": [1, 2, 3, 4], "grid": [1024, 2048, 4096]},
    "insert_batch": {"grid": [3072, 6144, 12288]},
}


@dataclass
class Entry:
    """One row of the tuning table (noa_hip_tuning)."""
    op: str
    box: int = 0          # 0: any
    remap: int = -1       # -1: any
    interp: int = -1
    dtype: int = -1
    values: tuple = (0, 0, 0, 0)
    ms: float | None = None


_table: list[Entry] = []


def apply(entries: Iterable[Entry]) -> None:
    """Replaces the library's tuning table with `entries` (and keeps them as this module's table)."""
    global _table
    _table = list(entries)
    arr = (abi.Tuning * max(1, len(_table)))()
    for i, e in enumerate(_table):
        arr[i].op, arr[i].box, arr[i].remap = OPS[e.op], e.box, e.remap
        arr[i].interp, arr[i].dtype = e.interp, e.dtype
        for k in range(4):
            arr[i].values[k] = int(e.values[k]) if k < len(e.values) else 0
    if lib().noa_hip_set_tuning(arr, len(_table)) != abi.OK:
        raise RuntimeError(last_error())


def table() -> list[Entry]:
    return list(_table)


def lookup(op: str, box: int = 0, remap: int = -1, interp: int = -1, dtype: int = -1) -> tuple:
    """The values the library's table gives a key (zeros where nothing matches)."""
    vals = (ctypes.c_int32 * 4)()
    lib().noa_hip_get_tuning(OPS[op], box, remap, interp, dtype, vals)
    return tuple(vals)


def save(path: str | Path) -> None:
    Path(path).write_text(json.dumps([asdict(e) for e in _table], indent=1))


def load(path: str | Path) -> list[Entry]:
    entries = [Entry(**{**d, "values": tuple(d["values"])}) for d in json.loads(Path(path).read_text())]
    apply(entries)
    return entries


def time_ms(fn: Callable[[], None], reps: int = 5, device: torch.device | str = "cuda") -> float:
    """Median time of `fn` in ms, from HIP events on the current stream (after one untimed call)."""
    fn()
    torch.cuda.synchronize(device)
    times = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        b.synchronize()
        times.append(a.elapsed_time(b))
    times.sort()
    return times[len(times) // 2]


def tune(op: str, fn: Callable[[], None], key: dict | None = None, space: dict | None = None, reps: int = 5,
         log: Callable[[str], None] | None = None, candidates: list | None = None) -> Entry:
    """Times `fn` under every candidate of `op` for one key and adds the fastest to the table.

    key: box / remap / interp / dtype of the entry (omitted fields match anything). candidates: explicit value
    tuples (else the cartesian product of `space`, by default SPACE[op])."""
    key = key or {}
    if candidates is None:
        space = space or SPACE[op]
        candidates = list(itertools.product(*(space[n] for n in space)))
    kept = [e for e in _table if not (e.op == op and all(getattr(e, k) == v for k, v in key.items()))]
    best: Entry | None = None
    for combo in candidates:
        cand = Entry(op=op, **key, values=tuple(combo) + (0,) * (4 - len(combo)))
        apply(kept + [cand])
        ms = time_ms(fn, reps)
        if log:
            log(f"tune {op} {key} {tuple(combo)}: {ms:.4f} ms")
        if best is None or ms < best.ms:
            cand.ms = ms
            best = cand
    apply(kept + [best])
    return best
