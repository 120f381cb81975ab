"""FP64 host projection of initial data onto the sparse tensor basis.

Mirrors vcoeffs_DG (dg_vmethods.jl:149-179), tensor_construct
(tensor_construct.jl:19-63), the exact coefficients of basic_function_exact_coeffs.jl
and cos_coeffs of examples/traveling_wave.jl:18-48.  The reference integrates with
adaptive hcubature (rtol 1e-8, dg_methods.jl:10-12); here each smooth piece of every
basis function is integrated with tensor Gauss-Legendre rules, which is exact to
rounding for polynomial data and converges spectrally for smooth data.
"""
from __future__ import annotations

import numpy as np

from . import layout
from .operators import dg_basis


def _legendre_norm(j, t):
    """Orthonormal Legendre polynomial of degree j on [-1, 1]."""
    c = np.zeros(j + 1)
    c[j] = 1.0
    return np.polynomial.legendre.legval(t, c) * np.sqrt((2 * j + 1) / 2.0)


def eval_v(k, level, cell, mode, x, _cache={}):
    """v(k, level, cell, mode, x) (dg_methods.jl:27-36), vectorised in x; cell and
    mode 1-based."""
    x = np.asarray(x, dtype=np.float64)
    if level == 0:
        return _legendre_norm(mode - 1, 2 * x - 1) * np.sqrt(2.0)
    if k not in _cache:
        _cache[k] = dg_basis(k)
    row = _cache[k][mode - 1]
    y = (1 << level) * x - (2 * cell - 1)
    s = np.zeros_like(y)
    sg = np.where(np.signbit(y), -1.0, 1.0)
    for i in range(k - 1, -1, -1):
        s = s * y + row[i] + sg * row[i + k]
    return np.where(np.abs(y) > 1.0, 0.0, s) * np.sqrt(float(1 << level))


def _pieces(level, cell):
    nc = 1 << max(0, level - 1)
    a, b = (cell - 1) / nc, cell / nc
    return [(a, 0.5 * (a + b)), (0.5 * (a + b), b)]


def vcoeffs_1d(k, n, f, panels=4, q=12):
    """vcoeffs_DG(1, k, n, f) for a vectorised f on [0, 1]."""
    t, w = np.polynomial.legendre.leggauss(q)
    out = np.zeros(k << n)
    j = 0
    for l in range(n + 1):
        for c in range(1, (1 << max(0, l - 1)) + 1):
            xs, ws = [], []
            for a, b in _pieces(l, c):
                for p in range(panels):
                    pa, pb = a + (b - a) * p / panels, a + (b - a) * (p + 1) / panels
                    xs.append(0.5 * (pa + pb) + 0.5 * (pb - pa) * t)
                    ws.append(0.5 * (pb - pa) * w)
            x = np.concatenate(xs)
            wt = np.concatenate(ws)
            fx = np.asarray(f(x), dtype=np.float64) * wt
            for m in range(1, k + 1):
                out[j] = np.dot(fx, eval_v(k, l, c, m, x))
                j += 1
    return out


def vcoeffs_DG(D, k, n, f, scheme="sparse", q=None, device=None):
    """dg_vmethods.jl:149-179 vcoeffs_DG(D, k, n, f; scheme): f takes an array of
    shape (D, npts) and returns npts values.  With `device` (a torch device), f takes
    and returns torch tensors on that device and the projection runs there
    (vcoeffs_DG_device)."""
    if device is not None:
        return vcoeffs_DG_device(D, k, n, f, scheme=scheme, q=q, device=device)
    q = q or (k + 4)
    t, w = np.polynomial.legendre.leggauss(q)
    out = np.zeros(layout.get_size(D, k, n, scheme))
    j = 0
    for lev in layout._levels(D, n, scheme):
        lv = [l - 1 for l in lev]
        nc = [1 << max(0, l - 1) for l in lv]
        # per dimension: nodes (nc*2q), weights, basis values [cell, mode, node]
        X, W, V = [], [], []
        for d in range(D):
            xs, ws = [], []
            for c in range(1, nc[d] + 1):
                for a, b in _pieces(lv[d], c):
                    xs.append(0.5 * (a + b) + 0.5 * (b - a) * t)
                    ws.append(0.5 * (b - a) * w)
            x = np.concatenate(xs)
            X.append(x)
            W.append(np.concatenate(ws))
            P = 2 * q
            vals = np.zeros((nc[d], k, P))
            for c in range(1, nc[d] + 1):
                xc = x[(c - 1) * P:c * P]
                for m in range(1, k + 1):
                    vals[c - 1, m - 1] = eval_v(k, lv[d], c, m, xc)
            V.append(vals)
        grids = np.meshgrid(*X, indexing="ij")
        fx = np.asarray(f(np.stack([g.ravel() for g in grids])), dtype=np.float64).reshape(grids[0].shape)
        for d in range(D):
            shape = [1] * D
            shape[d] = -1
            fx = fx * W[d].reshape(shape)
        # contract dimension by dimension: axis d (nc_d*P nodes) -> (nc_d*k coefficients)
        arr = fx
        P = 2 * q
        for d in range(D):
            sh = arr.shape
            arr = arr.reshape(sh[:d] + (nc[d], P) + sh[d + 1:])
            arr = np.moveaxis(arr, (d, d + 1), (0, 1))
            rest = arr.shape[2:]
            arr = np.einsum("cp...,cmp->cm...", arr, V[d]).reshape((nc[d] * k,) + rest)
            arr = np.moveaxis(arr, 0, d)
        arr = arr.reshape(sum(([nc[d], k] for d in range(D)), []))
        # memory order: modes (dim 1 fastest) inside cells (dim 1 fastest)
        perm = [2 * d for d in range(D - 1, -1, -1)] + [2 * d + 1 for d in range(D - 1, -1, -1)]
        blk = np.transpose(arr, perm).ravel()
        out[j:j + blk.size] = blk
        j += blk.size
    return out


def coeffs_DG(D, k, n, f, scheme="sparse", q=None, device=None):
    """dg_methods.jl:113-143 coeffs_DG(D, k, n, f; scheme): the Galerkin coefficients in
    the reference's Dict form, level (1-based tuple) -> array of shape (cells..., modes...)
    indexed [c_1-1, ..., c_D-1, m_1-1, ..., m_D-1] (the layout V2D gives), computed by the
    same projection as vcoeffs_DG (on `device` when given)."""
    v = vcoeffs_DG(D, k, n, f, scheme=scheme, q=q, device=device)
    if device is not None:
        v = v.cpu().numpy()
    return layout.V2D(D, k, n, v, scheme)


def vcoeffs_DG_device(D, k, n, f, scheme="sparse", q=None, device=None, max_points=1 << 24, broadcast=False):
    """vcoeffs_DG(D, k, n, f; scheme) (dg_vmethods.jl:149-179) on the device, FP64.

    The rule is the host vcoeffs_DG's: q Gauss-Legendre nodes (default k + 4) on each
    half of every 1-D cell, tensorised per cell.  f receives a float64 tensor X of shape
    (D, npts) on the device (X[d] the d-th coordinates) and returns npts values; it is
    called on a range of cells at a time, at most max_points nodes (or one cell).  The
    node coordinates and the contraction into coefficients are HIP kernels
    (gsg_proj_nodes, gsg_proj_contract).  Returns a float64 tensor of length
    get_size(D, k, n) on the device.

    broadcast=True: f receives a list of D tensors, X[d] of shape (cells, 1, ..., 2q,
    ..., 1) with the 2q node coordinates of direction d on axis D - d (axes: cell, then
    directions D, ..., 1), and returns values broadcastable to (cells, 2q, ..., 2q).
    For f built from elementwise operations this gives the same values at the same
    nodes, bit for bit, while the library writes only D 2q coordinates per cell
    (gsg_proj_nodes_1d) and f reads far less."""
    import torch
    q = q or (k + 4)
    dev = torch.device(device if device is not None else "cuda")
    if dev.type != "cuda":
        raise ValueError(f"vcoeffs_DG_device runs on a GPU, not {dev}")
    if dev.index is None:
        dev = torch.device("cuda", torch.cuda.current_device())
    with torch.cuda.device(dev):  # the handle's tables live on the device of the output
        return _project(D, k, n, f, scheme, q, dev, max_points, broadcast)


def _project(D, k, n, f, scheme, q, dev, max_points, broadcast=False):
    import ctypes as C
    import torch
    from . import _lib as L
    h = L.vp()
    L.check(L.lib().gsg_proj_create(D, k, n, L.SCHEME[scheme], q, C.byref(h)))
    lib = L.lib()
    try:
        N, ncells, ppc = lib.gsg_proj_len(h), lib.gsg_proj_cells(h), lib.gsg_proj_points_per_cell(h)
        chunk = max(1, min(ncells, int(max_points) // ppc))
        out = torch.empty(N, dtype=torch.float64, device=dev)
        P = 2 * q
        if broadcast:
            X = torch.empty(D * chunk * P, dtype=torch.float64, device=dev)
        else:
            X = torch.empty((D, chunk * ppc), dtype=torch.float64, device=dev)
        stream = C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        for c0 in range(0, ncells, chunk):
            nc = min(chunk, ncells - c0)
            npts = nc * ppc
            if broadcast:
                L.check(lib.gsg_proj_nodes_1d(h, c0, nc, C.c_void_p(X.data_ptr()), D * nc * P, stream))
                X1 = X[:D * nc * P].view(D, nc, P)
                Xs = []
                for d in range(D):
                    shape = [nc] + [1] * D
                    shape[D - d] = P
                    Xs.append(X1[d].reshape(shape))
                fv = f(Xs)
                if not torch.is_tensor(fv):
                    raise ValueError(f"f must return a tensor on {dev}")
                try:
                    fv = fv.to(device=dev, dtype=torch.float64).expand([nc] + [P] * D)
                except RuntimeError as e:
                    raise ValueError(f"f's result does not broadcast to (cells, {P}, ..., {P}): {e}") from None
            else:
                L.check(lib.gsg_proj_nodes(h, c0, nc, C.c_void_p(X.data_ptr()), X.shape[1], stream))
                fv = f(X[:, :npts])
                if not torch.is_tensor(fv) or fv.numel() != npts:
                    raise ValueError(f"f must return a tensor of {npts} values on {dev}")
            fv = fv.to(device=dev, dtype=torch.float64).reshape(-1).contiguous()
            L.check(lib.gsg_proj_contract(h, c0, nc, C.c_void_p(fv.data_ptr()), npts, C.c_void_p(out.data_ptr()), N,
                                          stream))
        return out
    finally:
        lib.gsg_proj_destroy(h)


def tensor_construct_device(D, k, n, vcoeff_array, scheme="sparse", device=None):
    """tensor_construct (tensor_construct.jl:57-63) on the device (gsg_proj_tensor):
    returns a float64 tensor of length get_size(D, k, n) on the device."""
    import ctypes as C
    import torch
    from . import _lib as L
    dev = torch.device(device if device is not None else "cuda")
    if dev.type != "cuda":
        raise ValueError(f"tensor_construct_device runs on a GPU, not {dev}")
    if dev.index is None:
        dev = torch.device("cuda", torch.cuda.current_device())
    if len(vcoeff_array) != D:
        raise ValueError(f"need {D} 1-D vectors")
    vecs = [torch.as_tensor(v, dtype=torch.float64).to(dev).contiguous() for v in vcoeff_array]
    for v in vecs:
        if v.numel() != (k << n):
            raise ValueError(f"a 1-D vector has {v.numel()} entries, not k 2^n = {k << n}")
    with torch.cuda.device(dev):
        lib = L.lib()
        h = L.vp()
        L.check(lib.gsg_proj_create(D, k, n, L.SCHEME[scheme], 1, C.byref(h)))
        try:
            N = lib.gsg_proj_len(h)
            out = torch.empty(N, dtype=torch.float64, device=dev)
            ptrs = (L.vp * D)(*[v.data_ptr() for v in vecs])
            stream = C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
            L.check(lib.gsg_proj_tensor(h, ptrs, C.c_void_p(out.data_ptr()), N, stream))
            torch.cuda.current_stream(dev).synchronize()  # vecs may be freed on return
            return out
        finally:
            lib.gsg_proj_destroy(h)


def tensor_construct(D, k, n, vcoeff_array, scheme="sparse", device=None):
    """tensor_construct.jl:57-63: coefficients of prod_d f_d(x_d) from the 1-D
    vhier vectors of the f_d.  With `device`, the products run there
    (tensor_construct_device) and a device tensor is returned."""
    if device is not None:
        return tensor_construct_device(D, k, n, vcoeff_array, scheme, device)
    lev, cel, mod = layout.V2Dref_arrays(D, k, n, scheme)
    out = np.ones(lev.shape[0])
    for d in range(D):
        l = lev[:, d] - 1
        base = np.where(l == 0, 0, 1 << np.maximum(l - 1, 0))
        idx = k * (base + cel[:, d] - 1) + mod[:, d] - 1
        out = out * np.asarray(vcoeff_array[d], dtype=np.float64)[idx]
    return out


def cos_coeffs(k, n, m, scheme="sparse", phase=0.0, A=1.0, device=None):
    """examples/traveling_wave.jl:18-48: A cos(2 pi m . x + phase).  With `device`, the
    D-dimensional products run there and a device tensor is returned (the 1-D
    projections stay on the host)."""
    D = len(m)
    wn = 2 * np.pi * np.asarray(m, dtype=np.float64)
    sines = [vcoeffs_1d(k, n, lambda x, i=i: np.sin(wn[i] * x + (phase if i == 0 else 0.0))) for i in range(D)]
    coss = [vcoeffs_1d(k, n, lambda x, i=i: np.cos(wn[i] * x + (phase if i == 0 else 0.0))) for i in range(D)]
    ans = None
    for code in range(1 << D):
        sc = [(code >> i) & 1 for i in range(D)]
        ns = sum(sc)
        if ns % 2:
            continue
        sign = 1.0 if ns % 4 == 0 else -1.0
        term = sign * tensor_construct(D, k, n, [sines[i] if sc[i] else coss[i] for i in range(D)], scheme, device)
        ans = term if ans is None else ans + term
    return A * ans


def sin_coeffs(k, n, m, scheme="sparse", phase=0.0, A=1.0, device=None):
    """examples/traveling_wave.jl:53-56."""
    return cos_coeffs(k, n, m, scheme=scheme, phase=phase - np.pi / 2, A=A, device=device)


_IDX_CACHE = {}
_PROJ_CACHE = {}


def _index_1d(D, k, n, scheme):
    """N x D array of 1-D vhier indices of every DoF (cached per grid)."""
    key = (D, k, n, scheme)
    if key not in _IDX_CACHE:
        lev, cel, mod = layout.V2Dref_arrays(D, k, n, scheme)
        l0 = lev.astype(np.int64) - 1
        base = np.where(l0 == 0, 0, np.left_shift(1, np.maximum(l0 - 1, 0)))
        _IDX_CACHE[key] = k * (base + cel - 1) + mod - 1
    return _IDX_CACHE[key]


def _exp_proj(k, n, m):
    """1-D projection of exp(2 pi i m x) (cached)."""
    key = (k, n, int(m))
    if key not in _PROJ_CACHE:
        w = 2 * np.pi * m
        _PROJ_CACHE[key] = vcoeffs_1d(k, n, lambda x: np.cos(w * x)) + 1j * vcoeffs_1d(k, n, lambda x: np.sin(w * x))
    return _PROJ_CACHE[key]


def fourier_coeffs(k, n, modes, amps, phases, scheme="sparse"):
    """sum_j amps[j] cos(2 pi modes[j] . x + phases[j]), by the same separable expansion
    as cos_coeffs (the real part of a product of 1-D exponentials, the phase carried by
    dimension 1)."""
    D = len(modes[0])
    idx = _index_1d(D, k, n, scheme)
    out = np.zeros(idx.shape[0])
    for m, a, p in zip(modes, amps, phases):
        prod = np.exp(1j * float(p)) * _exp_proj(k, n, m[0])[idx[:, 0]]
        for d in range(1, D):
            prod = prod * _exp_proj(k, n, m[d])[idx[:, d]]
        out += float(a) * prod.real
    return out


def get_one_modal(D, k, n):
    """basic_function_exact_coeffs.jl:5-8."""
    v = np.zeros(layout.get_size(D, k, n))
    v[0] = 1.0
    return v


def get_xi_modal(D, i, k, n):
    """basic_function_exact_coeffs.jl:11-17: coefficients of 2 x_i - 1 (the
    reference's 'velocity' coordinate on [-1, 1])."""
    N1 = k << n
    one = np.zeros(N1)
    one[0] = 1.0
    xi = np.zeros(N1)
    if k > 1:
        xi[1] = 1.0 / np.sqrt(3.0)
    return tensor_construct(D, k, n, [xi if d == i - 1 else one for d in range(D)])
