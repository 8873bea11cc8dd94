"""TEST INFRASTRUCTURE -- FP64 NumPy reference of the learned preconditioner's
convolution arithmetic (the checker of the BF16 matrix-core kernels; there is
no reference-repository counterpart).  Channels-last tensors x[n][d][h][w][c],
weights w[cout][27*cin] with taps (dz, dy, dx) row-major and cin fastest,
"same" zero padding -- the layout of flub_conv3d_bf16.
"""
import numpy as np

TAPS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def _shift(x, dz, dy, dx):
    """out[.., z, y, x, :] = x[.., z+dz, y+dy, x+dx, :] with zero outside."""
    out = np.zeros_like(x)
    n, D, H, W, C = x.shape
    zs, ze = max(0, -dz), min(D, D - dz)
    ys, ye = max(0, -dy), min(H, H - dy)
    xs, xe = max(0, -dx), min(W, W - dx)
    out[:, zs:ze, ys:ye, xs:xe] = x[:, zs + dz:ze + dz, ys + dy:ye + dy, xs + dx:xe + dx]
    return out


def conv3d(x, w, b=None, relu=False):
    x = np.asarray(x, np.float64)
    cout = w.shape[0]
    cin = x.shape[-1]
    w = np.asarray(w, np.float64).reshape(cout, 27, cin)
    y = np.zeros(x.shape[:-1] + (cout,))
    for t, (dz, dy, dx) in enumerate(TAPS):
        y += _shift(x, dz, dy, dx) @ w[:, t, :].T
    if b is not None:
        y += np.asarray(b, np.float64)
    return np.maximum(y, 0.0) if relu else y


def conv3d_wgrad(x, dy):
    """dL/dw[cout][27*cin] for y = conv3d(x, w): sum over voxels of dy (x) shifted x."""
    x = np.asarray(x, np.float64)
    dy = np.asarray(dy, np.float64)
    cin, cout = x.shape[-1], dy.shape[-1]
    dw = np.zeros((cout, 27, cin))
    for t, (dz, dyy, dx) in enumerate(TAPS):
        xs = _shift(x, dz, dyy, dx).reshape(-1, cin)
        dw[:, t, :] = dy.reshape(-1, cout).T @ xs
    return dw.reshape(cout, 27 * cin)


def conv3d_dgrad(dy, w, cin):
    """dL/dx for y = conv3d(x, w) (adjoint: transposed taps and channels)."""
    dy = np.asarray(dy, np.float64)
    cout = dy.shape[-1]
    w = np.asarray(w, np.float64).reshape(cout, 27, cin)
    dx = np.zeros(dy.shape[:-1] + (cin,))
    for t, (dz, dyy, dxx) in enumerate(TAPS):
        dx += _shift(dy @ w[:, t, :], -dz, -dyy, -dxx)
    return dx
