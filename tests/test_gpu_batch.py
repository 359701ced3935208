"""Batched extraction and insertion (one persistent launch per batch, replayed from a HIP graph).

Bars: a batched extraction equals the same calls made one by one bit for bit (same per-pixel expressions); a batched
insertion equals the one-wave kernel run one call at a time with unsplit sub-bricks bit for bit (same per-voxel sums
in slice order), and the oracle within rel-L2 1e-6. Graph replays are checked one by one against eager results, with
eager work and synchronisations between the replays.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle as O
from helpers import random_rotations, random_spectrum, rel_l2

pytestmark = pytest.mark.gpu

DEV = "cuda"

EXTRACT_JOBS = [   # remap, box, slices, cutoff
    ("hc2hc", 64, 37, 0.5),
    ("h2h", 96, 5, 0.45),
    ("hc2h", 33, 300, 0.5),
    ("h2hc", 128, 1, 0.5),
    ("hc2hc", 64, 64, 0.3),
]
INSERT_JOBS = [    # remap, box, slices, weights
    ("h2h", 64, 40, True),
    ("hc2hc", 96, 25, True),
    ("h2h", 128, 9, False),
    ("hc2h", 33, 60, True),
    ("h2hc", 64, 1, True),
]


def nx():
    import noa_amd.xform as m
    return m


def nb():
    import noa_amd.batch as m
    return m


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def make_extract(rng, jobs):
    out = []
    for remap, n, b, cut in jobs:
        vol = dev(random_spectrum((1, n, n, n // 2 + 1), rng))
        rot = dev(random_rotations(max(b, 1), rng)[:b].reshape(b, 3, 3))
        out.append(dict(remap=remap, n=n, b=b, cut=cut, vol=vol, rot=rot))
    return out


def extract_eager(j):
    res = torch.full((j["b"], 1, j["n"], j["n"] // 2 + 1), float("nan"), dtype=torch.complex64, device=DEV)
    if j["b"]:
        nx().extract_central_slices_3d(j["remap"], j["vol"], None, (1, j["n"], j["n"], j["n"]), res, None,
                                       (j["b"], 1, j["n"], j["n"]), None, j["rot"],
                                       nx().ExtractCentralSlicesOptions(fftfreq_cutoff=j["cut"]))
    return res


def extract_calls(jobs, outs):
    return [nb().extract_call(j["remap"], j["vol"], (1, j["n"], j["n"], j["n"]), o, (j["b"], 1, j["n"], j["n"]),
                              j["rot"], nx().ExtractCentralSlicesOptions(fftfreq_cutoff=j["cut"]))
            for j, o in zip(jobs, outs)]


def test_batched_extraction_equals_single_calls(monkeypatch):
    print("case batch extract direct", flush=True)
    monkeypatch.setenv("NOA_HIP_EXTRACT_LAYOUT", "linear")
    rng = np.random.default_rng(2000)
    jobs = make_extract(rng, EXTRACT_JOBS)
    want = [extract_eager(j) for j in jobs]
    outs = [torch.full_like(w, float("nan")) for w in want]
    ev = nb().run_batch(extract_calls(jobs, outs), [])
    ev.synchronize()
    for j, w, o in zip(jobs, want, outs):
        np.testing.assert_array_equal(o.cpu().numpy().view(np.uint32), w.cpu().numpy().view(np.uint32))
    # One job against the oracle as well.
    j = jobs[0]
    ref = np.zeros((j["b"], 1, j["n"], j["n"] // 2 + 1), np.complex64)
    O.extract_central_slices_3d(j["remap"], j["vol"].cpu().numpy(), None, (1, j["n"], j["n"], j["n"]), ref, None,
                                (j["b"], 1, j["n"], j["n"]), None, j["rot"].cpu().numpy(), fftfreq_cutoff=j["cut"])
    np.testing.assert_array_equal(outs[0].cpu().numpy(), ref)


def make_insert(rng, jobs):
    out = []
    for remap, n, b, w in jobs:
        sl = dev(random_spectrum((b, 1, n, n // 2 + 1), rng))
        wt = dev(rng.uniform(0.1, 1.0, (b, 1, n, n // 2 + 1)).astype(np.float32)) if w else None
        inv = dev(np.transpose(random_rotations(b, rng), (0, 2, 1)))
        out.append(dict(remap=remap, n=n, b=b, sl=sl, w=wt, inv=inv))
    return out


def insert_eager(j, monkeypatch):
    # The one-wave kernel without split sub-bricks: the batched kernel's work decomposition.
    monkeypatch.setenv("NOA_HIP_INSERT_KERNEL", "wave")
    monkeypatch.setenv("NOA_HIP_INSERT_LIST", "1000000000")
    n = j["n"]
    vol = torch.zeros((1, n, n, n // 2 + 1), dtype=torch.complex64, device=DEV)
    wv = torch.zeros((1, n, n, n // 2 + 1), dtype=torch.float32, device=DEV)
    nx().insert_central_slices_3d(j["remap"], j["sl"], j["w"], (j["b"], 1, n, n), vol, wv, (1, n, n, n), None, j["inv"])
    monkeypatch.delenv("NOA_HIP_INSERT_KERNEL")
    monkeypatch.delenv("NOA_HIP_INSERT_LIST")
    return vol, wv


def insert_calls(jobs, outs):
    return [nb().insert_call(j["remap"], j["sl"], j["w"], (j["b"], 1, j["n"], j["n"]), v, w, (1, j["n"], j["n"], j["n"]),
                             j["inv"]) for j, (v, w) in zip(jobs, outs)]


def test_batched_insertion_equals_single_calls(monkeypatch):
    print("case batch insert direct", flush=True)
    rng = np.random.default_rng(2100)
    jobs = make_insert(rng, INSERT_JOBS)
    want = [insert_eager(j, monkeypatch) for j in jobs]
    outs = [(torch.zeros_like(v), torch.zeros_like(w)) for v, w in want]
    ev = nb().run_batch([], insert_calls(jobs, outs))
    ev.synchronize()
    for (wv, ww), (ov, ow) in zip(want, outs):
        np.testing.assert_array_equal(ov.cpu().numpy(), wv.cpu().numpy())
        np.testing.assert_array_equal(ow.cpu().numpy(), ww.cpu().numpy())
    j = jobs[0]
    n = j["n"]
    ref = np.zeros((1, n, n, n // 2 + 1), np.complex64)
    refw = np.zeros((1, n, n, n // 2 + 1), np.float32)
    O.insert_central_slices_3d(j["remap"], j["sl"].cpu().numpy(), j["w"].cpu().numpy(), (j["b"], 1, n, n), ref, refw,
                               (1, n, n, n), None, j["inv"].cpu().numpy(), n_threads=8)
    assert rel_l2(outs[0][0].cpu().numpy(), ref) < 1e-6
    assert rel_l2(outs[0][1].cpu().numpy(), refw) < 1e-6


def test_graph_replays_of_different_batches(monkeypatch):
    """One slot, its launches captured once; three different batches (extraction + insertion) replayed, each
    checked against eager calls, with eager work and synchronisations between replays."""
    print("case batch graph replays", flush=True)
    monkeypatch.setenv("NOA_HIP_EXTRACT_LAYOUT", "linear")
    slot = nb().BatchSlot(DEV, graph=True)
    assert slot.graph is not None
    for rep, seed in enumerate((2200, 2201, 2202)):
        rng = np.random.default_rng(seed)
        ejobs = make_extract(rng, EXTRACT_JOBS[rep:rep + 3])
        ijobs = make_insert(rng, INSERT_JOBS[rep:rep + 3])
        ewant = [extract_eager(j) for j in ejobs]
        iwant = [insert_eager(j, monkeypatch) for j in ijobs]
        monkeypatch.setenv("NOA_HIP_EXTRACT_LAYOUT", "linear")
        eouts = [torch.full_like(w, float("nan")) for w in ewant]
        iouts = [(torch.zeros_like(v), torch.zeros_like(w)) for v, w in iwant]
        torch.cuda.synchronize()
        slot.prepare(extract_calls(ejobs, eouts), insert_calls(ijobs, iouts))
        ev = slot.run()
        _ = torch.randn(1 << 16, device=DEV).sum().item()     # eager work between replays
        ev.synchronize()
        for w, o in zip(ewant, eouts):
            np.testing.assert_array_equal(o.cpu().numpy().view(np.uint32), w.cpu().numpy().view(np.uint32))
        for (wv, ww), (ov, ow) in zip(iwant, iouts):
            np.testing.assert_array_equal(ov.cpu().numpy(), wv.cpu().numpy())
            np.testing.assert_array_equal(ow.cpu().numpy(), ww.cpu().numpy())


def test_batch_refuses_unsupported_calls():
    print("case batch refusals", flush=True)
    rng = np.random.default_rng(2300)
    n, b = 32, 4
    vol = dev(random_spectrum((1, n, n, n // 2 + 1), rng))
    rot = dev(random_rotations(b, rng))
    out = torch.empty((b, 1, n, n // 2 + 1), dtype=torch.complex64, device=DEV)
    opts = nx().ExtractCentralSlicesOptions(interp=nx().Interp.CUBIC)
    call = nb().extract_call("hc2hc", vol, (1, n, n, n), out, (b, 1, n, n), rot, opts)
    with pytest.raises(nx().NoaError, match="LINEAR"):
        nb().run_batch([call], [])


@pytest.mark.parametrize("op,values", [("extract_batch", (1, 512)), ("extract_batch", (4, 4096)),
                                       ("insert_batch", (1024,))])
def test_tuning_changes_no_result(monkeypatch, op, values):
    """Batched launches under other tile shapes (pixels per lane, grid sizes) give the same bits."""
    print(f"case batch tuning {op} {values}", flush=True)
    from noa_amd import autotune
    monkeypatch.setenv("NOA_HIP_EXTRACT_LAYOUT", "linear")
    rng = np.random.default_rng(2400)
    ejobs = make_extract(rng, EXTRACT_JOBS[:3])
    ijobs = make_insert(rng, INSERT_JOBS[:2])
    ewant = [extract_eager(j) for j in ejobs]
    iwant = [insert_eager(j, monkeypatch) for j in ijobs]
    try:
        autotune.apply([autotune.Entry(op=op, values=tuple(values) + (0,) * (4 - len(values)))])
        assert autotune.lookup(op)[: len(values)] == tuple(values)
        eouts = [torch.full_like(w, float("nan")) for w in ewant]
        iouts = [(torch.zeros_like(v), torch.zeros_like(w)) for v, w in iwant]
        nb().run_batch(extract_calls(ejobs, eouts), insert_calls(ijobs, iouts)).synchronize()
    finally:
        autotune.apply([])
    for w, o in zip(ewant, eouts):
        np.testing.assert_array_equal(o.cpu().numpy().view(np.uint32), w.cpu().numpy().view(np.uint32))
    for (wv, ww), (ov, ow) in zip(iwant, iouts):
        np.testing.assert_array_equal(ov.cpu().numpy(), wv.cpu().numpy())


def test_batched_extraction_from_block_copies(monkeypatch):
    """Batched extraction from VolumeBlocks (the 4x2x4-block copy made once) gives the bits of the single calls on
    the plain volumes; after `update` with another volume, those of the other volume; a single call refuses it."""
    print("case batch extract from block copies", flush=True)
    monkeypatch.setenv("NOA_HIP_EXTRACT_LAYOUT", "linear")
    rng = np.random.default_rng(2500)
    jobs = make_extract(rng, EXTRACT_JOBS)
    want = [extract_eager(j) for j in jobs]
    blocks = [nx().VolumeBlocks(j["vol"]) for j in jobs]
    outs = [torch.full_like(w, float("nan")) for w in want]
    calls = [nb().extract_call(j["remap"], vb, (1, j["n"], j["n"], j["n"]), o, (j["b"], 1, j["n"], j["n"]), j["rot"],
                               nx().ExtractCentralSlicesOptions(fftfreq_cutoff=j["cut"]))
             for j, vb, o in zip(jobs, blocks, outs)]
    nb().run_batch(calls, []).synchronize()
    for w, o in zip(want, outs):
        np.testing.assert_array_equal(o.cpu().numpy().view(np.uint32), w.cpu().numpy().view(np.uint32))

    # A new volume copied into the same blocks.
    j = jobs[1]
    j2 = dict(j, vol=dev(random_spectrum(tuple(j["vol"].shape), rng)))
    blocks[1].update(j2["vol"])
    want2 = extract_eager(j2)
    out2 = torch.full_like(want2, float("nan"))
    call = nb().extract_call(j["remap"], blocks[1], (1, j["n"], j["n"], j["n"]), out2, (j["b"], 1, j["n"], j["n"]),
                             j["rot"], nx().ExtractCentralSlicesOptions(fftfreq_cutoff=j["cut"]))
    nb().run_batch([call], []).synchronize()
    np.testing.assert_array_equal(out2.cpu().numpy().view(np.uint32), want2.cpu().numpy().view(np.uint32))

    # Refusals, all raised on the host before any launch.
    with pytest.raises(nx().NoaError, match="batched extraction only"):
        nx().extract_central_slices_3d(j["remap"], blocks[1], None, (1, j["n"], j["n"], j["n"]), out2, None,
                                       (j["b"], 1, j["n"], j["n"]), None, j["rot"])
    with pytest.raises(nx().NoaError):
        nb().run_batch([nb().extract_call(j["remap"], blocks[1], (1, 64, 64, 64), out2, (j["b"], 1, j["n"], j["n"]),
                                          j["rot"])], [])
    with pytest.raises(nx().NoaError, match="expected shape"):
        blocks[1].update(jobs[0]["vol"])
