"""eSRRF 3D on the host: the oracle against the reference's own C kernels, closed forms of the running statistics,
and the operator's argument checks (which run before any GPU work)."""
import numpy as np
import pytest

from oracle import ref32

CASES = [  # (Mxy, Mz, radius, sensitivity, mode, iw)
    (2, 2, 1.5, 1, "average", True),
    (1, 1, 1.5, 2, "average", True),
    (1, 2, 1.2, 1, "std", False),
    (2, 1, 2.0, 3, "std", True),
    (3, 2, 1.5, 1, "average", True),
]


def _stack(shape, seed=0):
    rng = np.random.default_rng(seed)
    return (rng.poisson(50, size=shape) + rng.random(shape) * 20).astype(np.float32)


@pytest.fixture(scope="module")
def oracle():
    ref32.build()
    return ref32


@pytest.mark.parametrize("case", CASES)
def test_esrrf3d_restatement_bit_exact_vs_reference(oracle, case):
    if not oracle.reference_available():
        pytest.skip("oracle/_ref/libnpx_ref.so is not built (reference sources absent)")
    Mxy, Mz, R, S, mode, iw = case
    x = _stack((3, 10, 18, 20), seed=Mxy * 10 + Mz)
    mine = oracle.esrrf3d(x, Mxy, Mz, R, 2.8, 4.0, S, mode, iw)
    oracle.use_reference_3d(True)
    oracle.use_reference_kernels(True)
    try:
        ref = oracle.esrrf3d(x, Mxy, Mz, R, 2.8, 4.0, S, mode, iw)
    finally:
        oracle.use_reference_3d(False)
        oracle.use_reference_kernels(False)
    assert np.array_equal(mine, ref, equal_nan=True)


def test_esrrf3d_closed_forms(oracle):
    # a constant volume has zero gradients: no sample with G.d < 0, so every voxel is 0
    assert not oracle.esrrf3d(np.full((2, 8, 12, 12), 5.0, np.float32), 2, 2).any()
    # identical frames: the running mean is the one-frame value exactly and the std is exactly 0
    v = _stack((1, 8, 14, 14), seed=3)
    rep = np.repeat(v, 4, axis=0)
    one = oracle.esrrf3d(v, 2, 2, 1.5)
    assert np.array_equal(oracle.esrrf3d(rep, 2, 2, 1.5), one) and one.any()
    assert not oracle.esrrf3d(rep, 2, 2, 1.5, mode="std").any()
    # std of two different frames, from their one-frame volumes: sqrt(((a - m)^2 + (b - m)^2) / 2) = |a - b| / 2
    w = _stack((1, 8, 14, 14), seed=4)
    a, b = oracle.esrrf3d(v, 2, 1, 1.5), oracle.esrrf3d(w, 2, 1, 1.5)
    sd = oracle.esrrf3d(np.concatenate([v, w]), 2, 1, 1.5, mode="std")
    assert np.allclose(sd, np.abs(a - b) / 2, rtol=1e-5, atol=1e-5 * np.abs(a).max())


def test_esrrf3d_operator_checks(capsys):
    from nanopyx_amd.core.transform import eSRRF3D_ST
    op = eSRRF3D_ST(verbose=False)
    x = np.zeros((1, 10, 20, 20), np.float32)
    assert op.run(x, radius=15) is None                     # radius above half the rows
    assert op.run(x, radius=6) is None                      # 2 * radius above half the columns
    assert op.run(x, PSF_ratio=15) is None                  # radius_z = 1.5 * 15 / 4 above half the slices
    assert op.run(np.zeros((1, 1, 4, 4, 4), np.float32)) is None
    assert "Warning" in capsys.readouterr().out
