"""The continuous-batching request loop: every result is the same alone and in the mix, and merging happens."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _mix(device, n=24, boxes=(64, 128, 192), max_k=40, seed=5):
    from noa_amd import workloads
    from noa_amd.serve import Request
    rng = np.random.default_rng(seed)
    g = torch.Generator(device=device).manual_seed(seed)
    reqs = []
    for i in range(n):
        box = int(rng.choice(boxes))
        k = int(rng.integers(1, max_k + 1))
        poses = workloads.uniform_rotations(k, rng)
        arrival = 0.0004 * (i // 4)          # groups of four arriving together
        if i % 2 == 0:
            reqs.append(Request(rid=i, kind="extract", box=box, volume=int(rng.integers(0, 2)),
                                rotations=torch.from_numpy(poses).to(device), arrival=arrival))
        else:
            spec = torch.fft.rfft2(torch.randn((k, box, box), generator=g, device=device), norm="forward")
            w = torch.rand((k, box, box // 2 + 1), generator=g, device=device)
            inv = np.ascontiguousarray(np.transpose(poses, (0, 2, 1)))
            reqs.append(Request(rid=i, kind="insert", box=box, rotations=torch.from_numpy(inv).to(device),
                                slices=spec.reshape(k, 1, box, box // 2 + 1).contiguous(),
                                weights=w.reshape(k, 1, box, box // 2 + 1).contiguous(), arrival=arrival))
    return reqs


def _clone(r):
    from noa_amd.serve import Request
    return Request(rid=r.rid, kind=r.kind, box=r.box, rotations=r.rotations, volume=r.volume, slices=r.slices,
                   weights=r.weights, arrival=0.0)


def test_results_are_the_same_alone_and_in_the_mix():
    from noa_amd import workloads
    from noa_amd.serve import Server, run_alone, run_open_loop
    dev = torch.device("cuda")
    vols = {b: v[:2] for b, v in workloads.serving_volumes(dev, per_box=2).items() if b in (64, 128, 192)}
    reqs = _mix(dev)
    torch.cuda.synchronize()
    server = Server(vols, budget_bytes=256 << 20, n_streams=3)
    stats = run_open_loop(server, reqs, timeout_s=60)
    torch.cuda.synchronize()
    assert len(stats.latencies_ms) == len(reqs)
    assert all(r.done is not None and r.done >= r.arrival for r in reqs)
    assert stats.merged > 0, "no extraction requests were merged into a shared launch"
    for r in reqs:
        alone = run_alone(Server(vols, budget_bytes=256 << 20, n_streams=1), _clone(r))
        assert torch.equal(r.result, alone.result), f"request {r.rid} ({r.kind}) differs alone and in the mix"
        if r.kind == "insert":
            assert torch.equal(r.result_weight, alone.result_weight)


def test_inputs_produced_just_before_submit_are_waited_for():
    """Requests whose inputs are still being written on the caller's stream when they are submitted (no
    synchronisation in between) give the same results as the same requests run alone on settled inputs."""
    from noa_amd import workloads
    from noa_amd.serve import Server, run_alone
    dev = torch.device("cuda")
    vols = {b: v[:2] for b, v in workloads.serving_volumes(dev, per_box=2).items() if b in (64, 128, 192)}
    reqs = _mix(dev, n=12, seed=8)
    torch.cuda.synchronize()
    want = [run_alone(Server(vols, budget_bytes=256 << 20, n_streams=1), _clone(r)) for r in reqs]
    # Overwrite every input with a long chain of work on the current stream, ending in the same values, and submit
    # right away: a launch that does not wait would read the intermediate values.
    server = Server(vols, budget_bytes=256 << 20, n_streams=3)
    fresh = []
    for r in reqs:
        c = _clone(r)
        rot = c.rotations.clone()
        c.rotations = torch.full_like(rot, float("nan"))
        for _ in range(20):
            c.rotations.mul_(1.0)
        c.rotations.copy_(rot)
        if c.kind == "insert":
            sl, w = c.slices.clone(), c.weights.clone()
            c.slices = torch.zeros_like(sl)
            c.weights = torch.zeros_like(w)
            c.slices.copy_(sl)
            c.weights.copy_(w)
        server.submit(c)
        fresh.append(c)
    while not server.idle():
        server.step()
    torch.cuda.synchronize()
    for got, ref in zip(fresh, want):
        assert torch.equal(got.result, ref.result), f"request {got.rid} ({got.kind}) read unsettled inputs"


def test_budget_is_respected_and_large_requests_still_run():
    from noa_amd import workloads
    from noa_amd.serve import Server, run_open_loop
    dev = torch.device("cuda")
    vols = {b: v[:2] for b, v in workloads.serving_volumes(dev, per_box=2).items() if b in (64, 128, 192)}
    reqs = _mix(dev, n=10, seed=6)
    torch.cuda.synchronize()
    budget = 4 << 20          # smaller than one 192^3 accumulator: such requests run alone
    server = Server(vols, budget_bytes=budget, n_streams=2)
    peak = 0
    orig = server._launch

    def launch_and_watch():
        nonlocal peak
        orig()
        peak = max(peak, server.in_use)
    server._launch = launch_and_watch
    stats = run_open_loop(server, reqs, timeout_s=60)
    assert len(stats.latencies_ms) == len(reqs)
    largest = max(r.output_bytes() for r in reqs)
    assert peak <= max(budget, largest)


def test_full_serving_mix_alone_and_in_the_mix():
    """The serving workload's whole box range (64-384) and slice counts up to 512: requests of every box, extraction
    and insertion, including K = 512 at 384^3, give the same bits alone and in the mix."""
    from noa_amd import workloads
    from noa_amd.serve import Request, Server, run_alone, run_open_loop
    dev = torch.device("cuda")
    vols = workloads.serving_volumes(dev, per_box=1)
    rng = np.random.default_rng(21)
    g = torch.Generator(device=dev).manual_seed(21)
    reqs = []
    plan = [(b, kind) for b in workloads.SERVING_BOXES for kind in ("extract", "insert")]
    for i, (box, kind) in enumerate(plan):
        k = 512 if box == 384 else int(rng.integers(1, 513))
        poses = workloads.uniform_rotations(k, rng)
        if kind == "extract":
            reqs.append(Request(rid=i, kind="extract", box=box, volume=0, rotations=torch.from_numpy(poses).to(dev),
                                arrival=0.0))
        else:
            spec = torch.fft.rfft2(torch.randn((k, box, box), generator=g, device=dev), norm="forward")
            w = torch.rand((k, box, box // 2 + 1), generator=g, device=dev)
            inv = np.ascontiguousarray(np.transpose(poses, (0, 2, 1)))
            reqs.append(Request(rid=i, kind="insert", box=box, rotations=torch.from_numpy(inv).to(dev),
                                slices=spec.reshape(k, 1, box, box // 2 + 1).contiguous(),
                                weights=w.reshape(k, 1, box, box // 2 + 1).contiguous(), arrival=0.0))
    print(f"case full serving mix: {len(reqs)} requests, boxes {sorted({r.box for r in reqs})}", flush=True)
    torch.cuda.synchronize()
    server = Server(vols, budget_bytes=4 << 30, n_streams=3)
    stats = run_open_loop(server, reqs, timeout_s=120)
    torch.cuda.synchronize()
    assert len(stats.latencies_ms) == len(reqs)
    assert max(stats.batch_sizes) > 1
    for r in reqs:
        alone = run_alone(Server(vols, budget_bytes=4 << 30, n_streams=1), _clone(r))
        assert torch.equal(r.result, alone.result), f"request {r.rid} ({r.kind}, box {r.box}) differs"
        if r.kind == "insert":
            assert torch.equal(r.result_weight, alone.result_weight)


def test_block_copy_volumes_change_no_result():
    """Extraction requests served from the resident volumes' block copies (the default) and from the volumes
    themselves give the same bits."""
    from noa_amd import workloads
    from noa_amd.serve import Server, run_alone
    dev = torch.device("cuda")
    vols = {b: v[:2] for b, v in workloads.serving_volumes(dev, per_box=2).items() if b in (64, 128, 192)}
    reqs = [r for r in _mix(dev, n=12, seed=9) if r.kind == "extract"]
    print(f"case serving block copies: {len(reqs)} extraction requests", flush=True)
    torch.cuda.synchronize()
    for r in reqs:
        a = run_alone(Server(vols, budget_bytes=256 << 20, n_streams=1), _clone(r))
        b = run_alone(Server(vols, budget_bytes=256 << 20, n_streams=1, volume_blocks=False), _clone(r))
        assert torch.equal(a.result, b.result), f"request {r.rid} differs with and without block copies"
