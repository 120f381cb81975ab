"""Time evolution mirroring pdes.jl.

The reference hands `(t,x)->RHS*x` to ODE.jl's adaptive ode45/ode78
(pdes.jl:62-68,113-119,206-213).  This build integrates with fixed-step classical RK4
on the device (the operator application and the stage update are one fused kernel
per stage); `order` therefore accepts "rk4" only and raises ValueError otherwise,
as the reference raises ArgumentError(:order) for unknown orders.  Results come back
as (tout, yout) like ODE.jl.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from . import operators as ops


def _steps(time0, time1, dt, nsteps):
    if nsteps is None:
        if dt is None:
            raise ValueError("give dt or nsteps")
        nsteps = int(round((time1 - time0) / dt))
    dt = (time1 - time0) / nsteps if nsteps > 0 else 0.0
    return dt, int(nsteps)


def _evolve(op, y0, time0, time1, order, dt, nsteps, save_every):
    if order != "rk4":
        raise ValueError(f"order {order!r}: this build integrates with fixed-step RK4 (order='rk4')")
    dt, nsteps = _steps(time0, time1, dt, nsteps)
    y0 = L.f64(y0)
    if save_every and save_every > 0:
        y, snaps = op.rk4(y0, dt, nsteps, save_every=save_every)
        tout = [time0] + [time0 + (i + 1) * save_every * dt for i in range(len(snaps))]
        yout = [y0.copy()] + [s.copy() for s in snaps]
        if nsteps % save_every:
            tout.append(time1)
            yout.append(y)
        return tout, yout
    y = op.rk4(y0, dt, nsteps)
    return [time0, time1], [y0.copy(), y]


def _coeffs(D, k, n, f, scheme):
    """Coefficient vector of f: f itself when it is a vector, vcoeffs_DG(D, k, n, f)
    when it is a function (of an array of shape (D, npts))."""
    if callable(f):
        from .ics import vcoeffs_DG
        return vcoeffs_DG(D, k, n, f, scheme=scheme)
    return L.f64(f)


def wave_evolve(D, k, n, f0coeffs, v0coeffs, time0, time1, order="rk4", scheme="sparse", dt=None, nsteps=None,
                save_every=0, precision="fp64"):
    """pdes.jl:54-70 wave_evolve(D, k, n, f0coeffs, v0coeffs, time0, time1), and the
    overload of pdes.jl:74-83 that takes the initial position and velocity profiles as
    functions (here of an array of shape (D, npts)) and projects them with vcoeffs_DG."""
    f0coeffs = _coeffs(D, k, n, f0coeffs, scheme)
    v0coeffs = _coeffs(D, k, n, v0coeffs, scheme)
    op = ops.wave_matrix(D, k, n, scheme=scheme, precision=precision)
    try:
        y0 = np.concatenate([L.f64(f0coeffs), L.f64(v0coeffs)])
        if y0.size != op.len:
            raise ValueError("coefficient vectors do not match get_size(D, k, n)")
        return _evolve(op, y0, time0, time1, order, dt, nsteps, save_every)
    finally:
        op.close()


def wave_evolve_1D(k, max_level, f0coeffs, v0coeffs, time0, time1, order="rk4", dt=None, nsteps=None,
                   save_every=0, precision="fp64"):
    """pdes.jl:89-121 in the hierarchical basis (equal to wave_evolve with D = 1)."""
    return wave_evolve(1, k, max_level, f0coeffs, v0coeffs, time0, time1, order=order, dt=dt, nsteps=nsteps,
                       save_every=save_every, precision=precision)


def transport_evolve(D, k, n, u0coeffs, a, time0, time1, order="rk4", scheme="sparse", dt=None, nsteps=None,
                     save_every=0, precision="fp64"):
    """u_t + a . grad u = 0: the composition -sum_d a_d Ds[d] u of pdes.jl:179,186,191.
    u0coeffs may also be a function, projected with vcoeffs_DG as wave_evolve does."""
    u0coeffs = _coeffs(D, k, n, u0coeffs, scheme)
    op = ops.transport_matrix(D, k, n, a, scheme=scheme, precision=precision)
    try:
        return _evolve(op, u0coeffs, time0, time1, order, dt, nsteps, save_every)
    finally:
        op.close()


def norm_squared(coeffs):
    """pdes.jl:230-237."""
    c = np.asarray(coeffs)
    return float(np.dot(c, c))


def energy_func(D, k, n, soln, scheme="sparse"):
    """pdes.jl:258-273: E = sum_d |D_d u|^2 + |v|^2 for each saved state."""
    times, states = soln
    Ds = ops.grad_matrix(D, k, n, scheme=scheme)
    try:
        energies = []
        for y in states:
            y = L.f64(y)
            N = y.size // 2
            energies.append(sum(norm_squared(Dd @ y[:N]) for Dd in Ds) + norm_squared(y[N:]))
        return list(times), energies
    finally:
        for Dd in Ds:
            Dd.close()


def energy_func_1D(k, level, soln):
    """pdes.jl:239-255 in the hierarchical basis."""
    return energy_func(1, k, level, soln)


SCHEDULE = {"auto": 0, "streams": 1, "fused": 2}


class Batch:
    """Many independent transport/wave solves advanced together: every RK stage of
    every solve is covered by one launch per (D, k) class."""

    def __init__(self, precision="fp64"):
        self.precision = precision
        self._h = L.vp()
        L.check(L.lib().gsg_batch_create(L.PREC[precision], C.byref(self._h)))
        self.lens = []

    def add(self, kind, D, k, n, y0, dt, nsteps, a=None, scheme="sparse"):
        desc = L.make_desc(D, k, n, kind, scheme, self.precision, coef=a)
        y = L.f64(y0)
        slot = C.c_int32()
        L.check(L.lib().gsg_batch_add(self._h, C.byref(desc), L.dptr(y), y.size, float(dt), int(nsteps),
                                      C.byref(slot)))
        self.lens.append(y.size)
        return slot.value

    def run(self):
        L.check(L.lib().gsg_batch_run(self._h))

    def step(self, steps):
        L.check(L.lib().gsg_batch_step(self._h, int(steps)))

    def get(self, slot):
        y = np.zeros(self.lens[slot])
        L.check(L.lib().gsg_batch_get(self._h, int(slot), L.dptr(y), y.size))
        return y

    def schedule(self, which="auto"):
        """Launch schedule of the RK stages (FP32): 'streams' (each (D, k) class its own
        launches, the classes on four streams), 'fused' (one launch per stage for every
        class the fused kernel serves) or 'auto' (the tuned choice)."""
        L.check(L.lib().gsg_batch_schedule(self._h, SCHEDULE[which]))

    def profile(self, enable=True):
        """Time every stage launch with HIP events (see kernel_stats)."""
        L.check(L.lib().gsg_batch_profile(self._h, 1 if enable else 0))

    def kernel_stats(self):
        """Per (D, k) kernel class: launches, summed milliseconds and algorithmic bytes."""
        cnt = C.c_int32()
        L.check(L.lib().gsg_batch_kernel_stats(self._h, None, 0, C.byref(cnt)))
        arr = (L.KernelStat * max(cnt.value, 1))()
        L.check(L.lib().gsg_batch_kernel_stats(self._h, C.cast(arr, C.c_void_p), cnt.value, C.byref(cnt)))
        return [{f: getattr(arr[i], f) for f, _ in L.KernelStat._fields_} for i in range(cnt.value)]

    @property
    def dofs(self):
        return int(L.lib().gsg_batch_dofs(self._h))

    def close(self):
        if self._h:
            L.lib().gsg_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
