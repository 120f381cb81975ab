"""vhier layout helpers mirroring dg_vmethods.jl and schemes.jl.

The vector layout is the wire format of the drop-in: multi-levels in Julia
CartesianIndices order (dimension 1 fastest) with the sparse cutoff sum(l) <= n on
0-based levels, then cells, then modes, dimension 1 fastest.  Tuples are 1-based like
the reference's CartesianIndex values.
"""
from __future__ import annotations

import itertools

import numpy as np

from . import _lib as L


def get_size(D, k, n, scheme="sparse"):
    """dg_vmethods.jl:35-45 get_size(Val(D), k, n, scheme)."""
    if scheme not in L.SCHEME:
        raise ValueError(f"unknown scheme {scheme!r}")
    return int(L.check(L.lib().gsg_get_size(D, k, n, L.SCHEME[scheme])))


def _levels(D, n, scheme):
    """CartesianIndices((n+1,)*D) order with dimension 1 fastest; 1-based levels."""
    for rev in itertools.product(range(1, n + 2), repeat=D):
        lev = rev[::-1]
        if scheme == "sparse" and sum(lev) > n + D:
            continue
        yield lev


def V2Dref(D, k, n, scheme="sparse"):
    """dg_vmethods.jl:123-142 V2Dref: list of ((level), (cell), (mode)) tuples."""
    N = get_size(D, k, n, scheme)
    lev = np.zeros((N, D), np.int32)
    cel = np.zeros((N, D), np.int32)
    mod = np.zeros((N, D), np.int32)
    L.check(L.lib().gsg_v2dref(D, k, n, L.SCHEME[scheme], lev.ctypes.data_as(L.ip32), cel.ctypes.data_as(L.ip32),
                               mod.ctypes.data_as(L.ip32), N))
    return [(tuple(a), tuple(b), tuple(c)) for a, b, c in zip(lev.tolist(), cel.tolist(), mod.tolist())]


def V2Dref_arrays(D, k, n, scheme="sparse"):
    """V2Dref as three N x D int32 arrays (levels, cells, modes), 1-based."""
    N = get_size(D, k, n, scheme)
    lev = np.zeros((N, D), np.int32)
    cel = np.zeros((N, D), np.int32)
    mod = np.zeros((N, D), np.int32)
    L.check(L.lib().gsg_v2dref(D, k, n, L.SCHEME[scheme], lev.ctypes.data_as(L.ip32), cel.ctypes.data_as(L.ip32),
                               mod.ctypes.data_as(L.ip32), N))
    return lev, cel, mod


def D2Vref(D, k, n, scheme="sparse"):
    """dg_vmethods.jl:102-121 D2Vref: dict (level, cell, mode) -> 0-based index."""
    return {t: j for j, t in enumerate(V2Dref(D, k, n, scheme))}


def V2D(D, k, n, vect, scheme="sparse"):
    """dg_vmethods.jl:155-179 V2D: dict level -> array of shape (cells..., modes...)
    indexed [c_1-1, ..., c_D-1, m_1-1, ..., m_D-1]."""
    vect = np.asarray(vect)
    if vect.shape != (get_size(D, k, n, scheme),):
        raise ValueError("vector length does not match get_size")
    out = {}
    j = 0
    for lev in _levels(D, n, scheme):
        cells = tuple(1 << max(0, l - 2) for l in lev)
        size = int(np.prod(cells)) * k ** D
        blk = vect[j:j + size]
        # memory order: modes fastest (dim 1 fastest), then cells (dim 1 fastest)
        arr = blk.reshape(cells[::-1] + (k,) * D)  # (c_D..c_1, m_D..m_1)
        arr = np.transpose(arr, tuple(range(D - 1, -1, -1)) + tuple(range(2 * D - 1, D - 1, -1)))
        out[lev] = np.array(arr)
        j += size
    return out


def D2V(D, k, n, coeffs, scheme="sparse"):
    """dg_vmethods.jl:127-152 D2V: inverse of V2D."""
    parts = []
    for lev in _levels(D, n, scheme):
        arr = np.asarray(coeffs[lev])
        arr = np.transpose(arr, tuple(range(D - 1, -1, -1)) + tuple(range(2 * D - 1, D - 1, -1)))
        parts.append(np.ascontiguousarray(arr).ravel())
    return np.concatenate(parts) if parts else np.zeros(0)
