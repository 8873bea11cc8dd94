"""Pins of the FP64 convolution reference (oracle/unet_ref.py) by identities
that do not reuse its code paths: adjointness of the data gradient, the
weight gradient as the derivative of <conv(x, w), dy> (linear in w), and a
direct per-voxel sum at random points."""
import numpy as np

from oracle import unet_ref as R


def test_dgrad_is_adjoint_and_wgrad_is_weight_derivative():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(2, 4, 5, 6, 3))
    w = rng.normal(size=(4, 27 * 3))
    dy = rng.normal(size=(2, 4, 5, 6, 4))
    y = R.conv3d(x, w)
    assert np.isclose(np.sum(y * dy), np.sum(x * R.conv3d_dgrad(dy, w, 3)))
    assert np.isclose(np.sum(y * dy), np.sum(w * R.conv3d_wgrad(x, dy)))


def test_conv_matches_direct_sum_at_points():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(1, 5, 4, 6, 2))
    w = rng.normal(size=(3, 27 * 2))
    b = rng.normal(size=3)
    y = R.conv3d(x, w, b)
    W = w.reshape(3, 3, 3, 3, 2)     # cout, dz, dy, dx, cin
    for (z, yy, xx) in [(0, 0, 0), (2, 1, 3), (4, 3, 5), (1, 2, 0)]:
        s = b.copy()
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    zz, y2, x2 = z + dz, yy + dy, xx + dx
                    if 0 <= zz < 5 and 0 <= y2 < 4 and 0 <= x2 < 6:
                        s += W[:, dz + 1, dy + 1, dx + 1, :] @ x[0, zz, y2, x2]
        assert np.allclose(y[0, z, yy, xx], s)
