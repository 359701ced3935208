"""noa_hip_insert_central_slices_3d_noalloc: the insertion without device allocations (static scratch per GPU).

It runs the same plan and kernels as noa_hip_insert_central_slices_3d, so its output must equal that entry point's bit
for bit. Two calls in a row share the scratch; the second, smaller one must not see the first one's leftovers. A call
whose work exceeds the scratch is refused before anything is launched.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from helpers import random_rotations, random_spectrum

pytestmark = pytest.mark.gpu

DEV = "cuda"


def nx():
    import noa_amd.xform as m
    return m


def call(fn_name, remap, sl, w, inv, n_slices, vol, wvol, n):
    """One insertion through the named C entry point, with the arguments the Python mirror builds."""
    from noa_amd import abi
    m = nx()
    keep = []
    args = abi.default_args()
    opts = m.InsertCentralSlicesOptions()
    device = m._fill(args, remap=remap, inp=sl, inp_w=w, out=vol, out_w=wvol, in_shape=(n_slices, 1, n, n),
                     out_shape=(1, n, n, n), rotation=inv, scaling=None, count=n_slices, opts=opts,
                     keep=keep, allow_value=True, interp=m.Interp.LINEAR)
    args.fftfreq_sinc = float(opts.windowed_sinc.fftfreq_sinc)
    args.fftfreq_blackman = float(opts.windowed_sinc.fftfreq_blackman)
    m._call(fn_name, args, device)


def inputs(rng, n, b):
    sl = torch.from_numpy(random_spectrum((b, 1, n, n // 2 + 1), rng, np.complex64)).to(DEV)
    w = torch.from_numpy(rng.uniform(0.1, 1.0, (b, 1, n, n // 2 + 1)).astype(np.float32)).to(DEV)
    inv = torch.from_numpy(np.ascontiguousarray(np.transpose(random_rotations(b, rng), (0, 2, 1)))).to(DEV)
    return sl, w, inv


def both(remap, sl, w, inv, b, n):
    out = {}
    for fn in ("noa_hip_insert_central_slices_3d", "noa_hip_insert_central_slices_3d_noalloc"):
        vol = torch.zeros((1, n, n, n // 2 + 1), dtype=torch.complex64, device=DEV)
        wvol = torch.zeros((1, n, n, n // 2 + 1), dtype=torch.float32, device=DEV)
        call(fn, remap, sl, w, inv, b, vol, wvol, n)
        torch.cuda.synchronize()
        out[fn] = (vol.cpu().numpy(), wvol.cpu().numpy())
    return out["noa_hip_insert_central_slices_3d"], out["noa_hip_insert_central_slices_3d_noalloc"]


@pytest.mark.parametrize("remap,n,b1,b2,kernel", [
    ("h2h", 128, 1500, 200, "slab"),     # the first call has split sub-bricks (partial sums in the scratch)
    ("h2h", 128, 1500, 200, "wave"),
    ("hc2hc", 96, 700, 90, "wave"),      # size not a power of two
])
def test_noalloc_equals_the_allocating_entry_point_twice_in_a_row(monkeypatch, remap, n, b1, b2, kernel):
    print(f"case noalloc {remap} n={n} b={b1} then b={b2} kernel={kernel}", flush=True)
    monkeypatch.setenv("NOA_HIP_INSERT_KERNEL", kernel)
    # The allocation-free entry point does not bound the slices per work item and keeps the kernel's fixed work-item
    # size (its partial sums stay in the scratch); with the same bound off and the same size, the allocating entry
    # point builds the same plan: the same bits.
    monkeypatch.setenv("NOA_HIP_INSERT_SPAN", "0")
    monkeypatch.setenv("NOA_HIP_INSERT_LIST", "512" if kernel == "slab" else "256")
    rng = np.random.default_rng(900 + n)
    first = inputs(rng, n, b1)
    second = inputs(rng, n, b2)
    for (sl, w, inv), b in ((first, b1), (second, b2)):
        (ref_v, ref_w), (got_v, got_w) = both(remap, sl, w, inv, b, n)
        assert np.abs(ref_v).max() > 0 and np.abs(ref_w).max() > 0
        np.testing.assert_array_equal(got_v, ref_v)
        np.testing.assert_array_equal(got_w, ref_w)


def test_noalloc_matches_the_allocating_default_plan():
    """With its default plan (at most 1,024 slices per work item, work items sized to the call) the allocating
    entry point regroups the partial sums differently: the same result within rel-L2 1e-6."""
    print("case noalloc vs default plan h2h n=128 b=3000", flush=True)
    rng = np.random.default_rng(960)
    n, b = 128, 3000
    sl, w, inv = inputs(rng, n, b)
    (ref_v, ref_w), (got_v, got_w) = both("h2h", sl, w, inv, b, n)
    rel = lambda a, r: float(np.linalg.norm(a - r) / np.linalg.norm(r))
    assert rel(got_v, ref_v) < 1e-6 and rel(got_w, ref_w) < 1e-6


def test_noalloc_refuses_work_beyond_the_scratch():
    """256^3 with 8192 slices needs ~120 MB of partial sums: refused with the size, the volume untouched."""
    print("case noalloc refusal 256^3 x 8192", flush=True)
    m = nx()
    rng = np.random.default_rng(950)
    n, b = 256, 8192
    sl = torch.zeros((b, 1, n, n // 2 + 1), dtype=torch.complex64, device=DEV)
    inv = torch.from_numpy(np.ascontiguousarray(np.transpose(random_rotations(b, rng), (0, 2, 1)))).to(DEV)
    vol = torch.full((1, n, n, n // 2 + 1), 1 + 1j, dtype=torch.complex64, device=DEV)
    with pytest.raises(m.NoaError, match="bytes of scratch"):
        call("noa_hip_insert_central_slices_3d_noalloc", "h2h", sl, None, inv, b, vol, None, n)
    torch.cuda.synchronize()
    assert bool((vol == (1 + 1j)).all())
