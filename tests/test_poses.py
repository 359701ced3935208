"""Pins of the pose construction (noa_amd.xform.euler2matrix, rotate_*, quaternions) that every workload uses.

euler2matrix restates xform/core/Euler.cpp:218-274 and rotate_x/y/z xform/core/Transform.hpp:82-129. noa's matrices
act on (z, y, x) vectors; the reference's own test (tests/noa/xform/TestXformCoreEuler.cpp:40-56) compares them, after
reversing the axis order (`permute(Vec{2, 1, 0})`, zyx -> xyz), with matrices made in Python. scipy's Rotation is
an independent implementation of the same textbook compositions, so it pins the axis order, the composition order
(intrinsic r1*r2*r3, extrinsic r3*r2*r1), the handedness and the sign of every axis rotation.
"""
import itertools

import numpy as np
import pytest

from noa_amd import xform as nx

scipy_rot = pytest.importorskip("scipy.spatial.transform")

# The 12 proper Euler and Tait-Bryan sequences noa accepts (Euler.cpp:218-274 rejects repeated neighbours).
SEQUENCES = ["".join(p) for p in itertools.product("ZYX", repeat=3) if p[0] != p[1] and p[1] != p[2]]


def zyx_to_xyz(m):
    """TestXformCoreEuler.cpp:45-47: matrix.permute(Vec{2, 1, 0})."""
    return m[..., ::-1, ::-1]


def test_twelve_sequences():
    assert len(SEQUENCES) == 12


@pytest.mark.parametrize("axes", SEQUENCES)
@pytest.mark.parametrize("intrinsic", [True, False])
@pytest.mark.parametrize("right_handed", [True, False])
def test_euler2matrix_is_scipys_composition(axes, intrinsic, right_handed):
    rng = np.random.default_rng(11)
    angles = rng.uniform(-np.pi, np.pi, (32, 3))
    got = zyx_to_xyz(nx.euler2matrix(angles, axes=axes.lower(), intrinsic=intrinsic, right_handed=right_handed))
    seq = axes.upper() if intrinsic else axes.lower()        # scipy: upper case = intrinsic
    signed = angles if right_handed else -angles
    want = scipy_rot.Rotation.from_euler(seq, signed).as_matrix()
    np.testing.assert_allclose(got, want, atol=1e-12)


def test_euler2matrix_reference_asset_angles_f32():
    """The reference's asset test computes f32 matrices of (deg2rad of) its yaml angles for every sequence and
    compares them to 1e-6 (TestXformCoreEuler.cpp:17-56). Its yaml angles are not available; a fixed triplet of
    degrees stands in, with the same f32 evaluation and tolerance."""
    ang = np.deg2rad(np.array([[25.0, -44.0, 131.0]], np.float32)).astype(np.float32)
    for axes in SEQUENCES:
        for intrinsic in (True, False):
            for rh in (True, False):
                got = zyx_to_xyz(nx.euler2matrix(ang, axes=axes, intrinsic=intrinsic, right_handed=rh,
                                                 dtype=np.float32))
                assert got.dtype == np.float32
                seq = axes if intrinsic else axes.lower()
                want = scipy_rot.Rotation.from_euler(seq, ang.astype(np.float64) * (1 if rh else -1)).as_matrix()
                np.testing.assert_allclose(got, want, atol=1e-6)


@pytest.mark.parametrize("axes", SEQUENCES)
def test_euler_round_trip_through_an_independent_inverse(axes):
    """euler2matrix -> angles (scipy's inverse of the xyz matrix) -> euler2matrix returns the same matrix
    (TestXformCoreEuler.cpp:59-78 does this with noa's matrix2euler)."""
    angles = np.deg2rad(np.array([10.0, 20.0, 30.0]))
    for intrinsic in (True, False):
        m = nx.euler2matrix(angles, axes=axes, intrinsic=intrinsic)
        seq = axes if intrinsic else axes.lower()
        back = scipy_rot.Rotation.from_matrix(zyx_to_xyz(m)).as_euler(seq)
        np.testing.assert_allclose(nx.euler2matrix(back, axes=axes, intrinsic=intrinsic), m, atol=1e-12)


def test_rotate_z_90_maps_x_to_y():
    """TestXformCoreQuaternion.cpp:20-33: rotate_z(90 deg) * (z=0, y=0, x=1) == (0, 1, 0), for the matrix and for
    the quaternion made from it."""
    m = nx.rotate_z(np.deg2rad(90.0))
    v = np.array([0.0, 0.0, 1.0])
    np.testing.assert_allclose(m @ v, [0.0, 1.0, 0.0], atol=1e-12)
    q = nx.quaternion_from_matrix(m)
    np.testing.assert_allclose(nx.quaternion_to_matrix(q) @ v, [0.0, 1.0, 0.0], atol=1e-12)


def test_rotate_y_and_x_single_axis():
    """Right-handed rotations about y (z -> x) and x (y -> z) by 90 deg, in zyx component order."""
    ry = nx.rotate_y(np.deg2rad(90.0))
    rx = nx.rotate_x(np.deg2rad(90.0))
    np.testing.assert_allclose(ry @ np.array([1.0, 0, 0]), [0, 0, 1.0], atol=1e-12)
    np.testing.assert_allclose(rx @ np.array([0, 1.0, 0]), [1.0, 0, 0], atol=1e-12)


def test_quaternion_rotation_equals_the_matrix():
    """TestXformCoreQuaternion.cpp:35-58: for random Euler angles, matrix * v == quaternion's matrix * v."""
    rng = np.random.default_rng(3)
    m = nx.euler2matrix(rng.uniform(-1, 1, (1000, 3)))
    q = nx.quaternion_from_matrix(m)
    v = rng.uniform(-1, 1, (1000, 3))
    np.testing.assert_allclose(np.einsum("bij,bj->bi", nx.quaternion_to_matrix(q), v),
                               np.einsum("bij,bj->bi", m, v), atol=1e-12)


def test_rotations_are_proper_and_orthonormal():
    rng = np.random.default_rng(5)
    m = nx.euler2matrix(rng.uniform(-np.pi, np.pi, (200, 3)), axes="zyz")
    np.testing.assert_allclose(m @ np.transpose(m, (0, 2, 1)), np.broadcast_to(np.eye(3), m.shape), atol=1e-12)
    np.testing.assert_allclose(np.linalg.det(m), 1.0, atol=1e-12)


def test_workload_poses_are_uniform_on_so3():
    """The workloads' poses (uniform ZYZ with theta = arccos U(-1, 1)) rotate a fixed axis uniformly over the
    sphere: each component's mean is ~0 and its second moment ~1/3."""
    from noa_amd.workloads import uniform_rotations
    m = uniform_rotations(20000, np.random.default_rng(9), dtype=np.float64)
    n = m[:, 0, :]                       # image of the z axis under R^T, i.e. the slice normals
    assert np.all(np.abs(n.mean(0)) < 0.02)
    np.testing.assert_allclose((n ** 2).mean(0), 1 / 3, atol=0.01)
