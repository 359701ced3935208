"""The pose-adaptive insertion plan (item order 3): segments cut from the passing slices counted on the device, for
orientation sets far from uniform (a single-axis tilt series, one plane repeated). Bars: rel-L2 1e-6 against the
oracle and against the default plan (the segments regroup the per-voxel sums), deterministic run to run."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle as O
from helpers import random_spectrum, rel_l2

pytestmark = pytest.mark.gpu

DEV = "cuda"


def nx():
    import noa_amd.xform as m
    return m


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def poses(kind, b, rng):
    m = nx()
    if kind == "tilt":
        ang = np.stack([np.zeros(b), np.deg2rad(rng.uniform(-60, 60, b)), rng.uniform(0, 2 * np.pi, b)], -1)
        return m.euler2matrix(ang, axes="zyz", dtype=np.float32)
    return np.repeat(m.rotate_y(np.deg2rad(23.0)).astype(np.float32)[None], b, 0)


@pytest.mark.parametrize("kernel", ["slab", "wave"])
@pytest.mark.parametrize("kind", ["tilt", "plane"])
def test_pose_plan_matches_the_oracle_and_the_default_plan(monkeypatch, kernel, kind):
    print(f"case pose plan {kind} {kernel}", flush=True)
    rng = np.random.default_rng(500)
    n, b = 64, 600
    sl = random_spectrum((b, 1, n, n // 2 + 1), rng)
    w = rng.uniform(0.1, 1.0, (b, 1, n, n // 2 + 1)).astype(np.float32)
    inv = np.ascontiguousarray(np.transpose(poses(kind, b, rng), (0, 2, 1)))
    want_v = np.zeros((1, n, n, n // 2 + 1), np.complex64)
    want_w = np.zeros((1, n, n, n // 2 + 1), np.float32)
    O.insert_central_slices_3d("h2h", sl, w, (b, 1, n, n), want_v, want_w, (1, n, n, n), None, inv, n_threads=8)
    monkeypatch.setenv("NOA_HIP_INSERT_KERNEL", kernel)
    monkeypatch.setenv("NOA_HIP_INSERT_LIST", "64")
    outs = {}
    for order in ("3", "0", "3"):
        monkeypatch.setenv("NOA_HIP_INSERT_ORDER", order)
        gv = torch.zeros((1, n, n, n // 2 + 1), dtype=torch.complex64, device=DEV)
        gw = torch.zeros((1, n, n, n // 2 + 1), dtype=torch.float32, device=DEV)
        nx().insert_central_slices_3d("h2h", dev(sl), dev(w), (b, 1, n, n), gv, gw, (1, n, n, n), None, dev(inv))
        torch.cuda.synchronize()
        got = (gv.cpu().numpy(), gw.cpu().numpy())
        if order in outs:
            np.testing.assert_array_equal(got[0], outs[order][0])
            np.testing.assert_array_equal(got[1], outs[order][1])
        outs[order] = got
    assert rel_l2(outs["3"][0], want_v) < 1e-6
    assert rel_l2(outs["3"][1], want_w) < 1e-6
    assert rel_l2(outs["3"][0], outs["0"][0]) < 1e-6
    assert rel_l2(outs["3"][1], outs["0"][1]) < 1e-6


def test_pose_plan_with_many_slices_matches_the_default_plan(monkeypatch):
    """5,000 slices of a tilt series at 64^3: bins of 128 slices, against the default plan (rel-L2 1e-6; the oracle
    at this size is covered by the smaller cases)."""
    print("case pose plan 5000 slices", flush=True)
    rng = np.random.default_rng(501)
    n, b = 64, 5000
    sl = dev(random_spectrum((b, 1, n, n // 2 + 1), rng))
    w = dev(rng.uniform(0.1, 1.0, (b, 1, n, n // 2 + 1)).astype(np.float32))
    inv = dev(np.ascontiguousarray(np.transpose(poses("tilt", b, rng), (0, 2, 1))))
    outs = []
    for order in ("3", "0"):
        monkeypatch.setenv("NOA_HIP_INSERT_ORDER", order)
        gv = torch.zeros((1, n, n, n // 2 + 1), dtype=torch.complex64, device=DEV)
        gw = torch.zeros((1, n, n, n // 2 + 1), dtype=torch.float32, device=DEV)
        nx().insert_central_slices_3d("h2h", sl, w, (b, 1, n, n), gv, gw, (1, n, n, n), None, inv)
        torch.cuda.synchronize()
        outs.append((gv.cpu().numpy(), gw.cpu().numpy()))
    assert rel_l2(outs[0][0], outs[1][0]) < 1e-6
    assert rel_l2(outs[0][1], outs[1][1]) < 1e-6
