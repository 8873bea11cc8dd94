"""Pin of the oracle's V-cycle (oracle/oracle_body.h
orc_mgpcg_apply_preconditioner, restating multigrid/mgpcg.py:296-349) and of
its variant without a reference counterpart (restricted residuals times a
coarse-correction scale, a cap on the sweeps per level) against an
independent vectorised numpy V-cycle in FP64:

* red-black half-sweeps as whole-colour array updates: phase p updates the
  cells with (i + j + k) odd for p = 0 and even for p = 1 (mgpcg.py:72-86
  maps k = 2k' + 1 onto the rows with (i + j) & 1 == phase), z = 0 where
  divd = 0;
* divd counted from the tags on FLUID cells, 0 elsewhere (_compute_divd and
  neighbor_sum_tag, mgpcg.py:51-58,145-160), the coarse tags by restrict_tag
  (mgpcg.py:162-174);
* restriction as a reshape-sum of r + ndt(z) over the 2x2x2 children, zero
  under a non-FLUID parent; piecewise-constant prolongation by repetition.

The two agree to FP64 rounding on the reference cycle and on the variants; a
scale applied on one level only, or a cap applied to level 0, would not.
"""
import numpy as np
import pytest

from oracle import oracle as O


def _ndt(x, tag):
    out = np.zeros_like(x)
    for ax in range(3):
        n = x.shape[ax]
        for s in (1, -1):
            idx = np.clip(np.arange(n) + s, 0, n - 1)
            xn, tn = np.take(x, idx, axis=ax), np.take(tag, idx, axis=ax)
            out += (np.where(tn == 0, 0.0, xn) - x) * (tn < 2)
    out[tag != 1] = 0.0
    return out


def _divd(tag):
    d = np.zeros(tag.shape)
    for ax in range(3):
        n = tag.shape[ax]
        shp = [1, 1, 1]
        shp[ax] = n
        for s in (1, -1):
            idx = np.arange(n) + s
            inside = ((idx >= 0) & (idx < n)).reshape(shp)
            d += (np.take(tag, np.clip(idx, 0, n - 1), axis=ax) < 2) & inside
    d[tag != 1] = 0.0
    return d


def _restrict_tag(t):
    a = t.reshape(t.shape[0] // 2, 2, t.shape[1] // 2, 2, t.shape[2] // 2, 2)
    return np.where((a == 0).any(axis=(1, 3, 5)), 0, np.where((a == 1).any(axis=(1, 3, 5)), 1, 2))


def numpy_vcycle(r0, tag, levels, pre_post, bottom, cs=1.0, cap=0):
    tags = [tag]
    for _ in range(levels - 1):
        tags.append(_restrict_tag(tags[-1]))
    divd = [_divd(t) for t in tags]
    odd = [(np.indices(t.shape).sum(0) & 1) == 1 for t in tags]

    def sweep(z, r, l, phase):
        cells = odd[l] if phase == 0 else ~odd[l]
        d = divd[l]
        new = np.where(d != 0, z + (r + _ndt(z, tags[l])) / np.where(d != 0, d, 1.0), 0.0)
        z[cells] = new[cells]

    def count(l):
        s = pre_post << l
        return min(s, cap) if cap > 0 else s

    z, r = [None] * levels, [None] * levels
    r[0] = r0.astype(np.float64)
    for l in range(levels - 1):
        z[l] = np.zeros_like(r[l])
        for _ in range(count(l)):
            sweep(z[l], r[l], l, 0)
            sweep(z[l], r[l], l, 1)
        res = r[l] + _ndt(z[l], tags[l])
        c = res.reshape(res.shape[0] // 2, 2, res.shape[1] // 2, 2, res.shape[2] // 2, 2).sum(axis=(1, 3, 5))
        r[l + 1] = np.where(tags[l + 1] == 1, c, 0.0) * cs
    lb = levels - 1
    z[lb] = np.zeros_like(r[lb])
    for _ in range(bottom // 2):
        sweep(z[lb], r[lb], lb, 0)
        sweep(z[lb], r[lb], lb, 1)
    for _ in range(bottom // 2):
        sweep(z[lb], r[lb], lb, 1)
        sweep(z[lb], r[lb], lb, 0)
    for l in reversed(range(levels - 1)):
        z[l] = z[l] + np.repeat(np.repeat(np.repeat(z[l + 1], 2, 0), 2, 1), 2, 2)
        for _ in range(count(l)):
            sweep(z[l], r[l], l, 1)
            sweep(z[l], r[l], l, 0)
    return z[0]


def _problem(shape, seed):
    rng = np.random.default_rng(seed)
    tag = rng.choice(3, size=shape, p=[0.08, 0.86, 0.06]).astype(np.int32)
    r = rng.normal(0, 1, shape)
    r[tag != 1] = 0.0
    return tag, r


@pytest.mark.parametrize("shape,levels,cs,cap", [((32, 32, 32), 4, 1.0, 0), ((32, 32, 32), 4, 0.5, 3),
                                                 ((32, 16, 32), 4, 0.6, 2), ((64, 32, 32), 5, 0.5, 3)])
def test_oracle_vcycle_matches_numpy(shape, levels, cs, cap):
    tag, r = _problem(shape, 5)
    o = O.MGPCG(shape, levels, real=np.float64, pre_and_post_smoothing=2, bottom_smoothing=6)
    o.set_tags(tag)
    o.set_cycle(cs, cap)
    o.put("r", r, 0)
    o.apply_preconditioner("phases")
    got = o.get("z", 0)
    want = numpy_vcycle(r, tag, levels, 2, 6, cs, cap)
    assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want)
    if (cs, cap) != (1.0, 0):   # the variant is a different cycle
        ref = numpy_vcycle(r, tag, levels, 2, 6)
        assert np.linalg.norm(got - ref) > 1e-3 * np.linalg.norm(ref)
