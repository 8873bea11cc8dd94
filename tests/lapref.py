"""Independent constructions used to pin the oracle (scipy/NumPy, no oracle code).

assemble(tag): the pressure matrix of the tagged MAC grid built from its graph
definition -- unknowns are the FLUID cells; every face between two non-SOLID
in-domain cells is an edge; A_II counts the edges at I (AIR neighbours count,
Dirichlet p = 0), A_IJ = -1 between adjacent FLUID cells.  Domain walls and
SOLID neighbours contribute nothing (Neumann).
"""
import numpy as np
import scipy.sparse as sp


def assemble(tag):
    tag = np.asarray(tag)
    fl = tag == 1
    ids = -np.ones(tag.shape, np.int64)
    ids[fl] = np.arange(int(fl.sum()))
    nf = int(fl.sum())
    rows, cols = [], []
    diag = np.zeros(nf)
    for a in range(3):
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[a] = slice(0, -1)
        hi[a] = slice(1, None)
        t0, t1 = tag[tuple(lo)], tag[tuple(hi)]
        i0, i1 = ids[tuple(lo)], ids[tuple(hi)]
        ff = (t0 == 1) & (t1 == 1)
        rows += [i0[ff], i1[ff]]
        cols += [i1[ff], i0[ff]]
        np.add.at(diag, i0[ff], 1)
        np.add.at(diag, i1[ff], 1)
        np.add.at(diag, i0[(t0 == 1) & (t1 == 0)], 1)
        np.add.at(diag, i1[(t0 == 0) & (t1 == 1)], 1)
    r = np.concatenate(rows) if rows else np.zeros(0, np.int64)
    c = np.concatenate(cols) if cols else np.zeros(0, np.int64)
    A = sp.csr_matrix((-np.ones(len(r)), (r, c)), shape=(nf, nf)) + sp.diags(diag)
    return A.tocsr(), ids


def random_tags(shape, seed, p=(0.2, 0.65, 0.15)):
    return np.random.default_rng(seed).choice(3, size=shape, p=list(p)).astype(np.int32)


def face_gradient(p):
    """G p on the three face grids: interior faces p_r - p_l, wall faces 0."""
    nx, ny, nz = p.shape
    gu = np.zeros((nx + 1, ny, nz))
    gv = np.zeros((nx, ny + 1, nz))
    gw = np.zeros((nx, ny, nz + 1))
    gu[1:-1] = p[1:] - p[:-1]
    gv[:, 1:-1] = p[:, 1:] - p[:, :-1]
    gw[:, :, 1:-1] = p[:, :, 1:] - p[:, :, :-1]
    return gu, gv, gw
