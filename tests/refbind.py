"""Calls into oracle/_ref/libarmref.so: the reference's own CERNsource
routines (kernel.f, link_list.f, density.f, internal_force.f, art_visc.f,
art_heat.f, av_vel.f, direct_find.f) compiled unmodified by oracle/Makefile.

The as-shipped options.inc has dim = 3, so positions and velocities are
passed as x(3, n) with z = 0.  Their large local arrays need a big stack,
so every call runs on a helper thread with a 512 MiB stack.
"""
import ctypes as C
import os
import threading

import numpy as np

from oracle import orc

_lib = None


def available():
    return os.path.exists(orc.REF_PATH)


def lib():
    global _lib
    if _lib is None:
        _lib = C.CDLL(orc.REF_PATH)
    return _lib


def run_big_stack(fn, *a):
    out = {}

    def body():
        try:
            out["r"] = fn(*a)
        except BaseException as e:  # re-raised on the caller thread
            out["e"] = e

    old = threading.stack_size(1 << 29)
    try:
        t = threading.Thread(target=body)
        t.start()
        t.join()
    finally:
        threading.stack_size(old)
    if "e" in out:
        raise out["e"]
    return out["r"]


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _i(v):
    return C.byref(C.c_int(v))


def vec3(a, b, c=None):
    X = np.zeros((len(a), 3))
    X[:, 0] = a
    X[:, 1] = b
    if c is not None:
        X[:, 2] = c
    return X


def hvec(h, n):
    """Per-particle smoothing lengths from a scalar or an array."""
    return np.ascontiguousarray(np.broadcast_to(np.asarray(h, np.float64), (n,)))


_lib90 = None


def available90():
    return os.path.exists(orc.REF90_PATH)


def lib90():
    """The other tree's routines (FORTRANsource/external_force.f90, kernel.f90)."""
    global _lib90
    if _lib90 is None:
        _lib90 = C.CDLL(orc.REF90_PATH)
    return _lib90


def kernel90(r, h, skf=0):
    """W of FORTRANsource/kernel.f90 (dim = 3) at distance r along x."""
    w = C.c_double(0.0)
    dx = np.array([r, 0.0, 0.0])
    dw = np.zeros(3)
    lib90().kernel_(C.byref(C.c_double(r)), _p(dx), C.byref(C.c_double(h)), C.byref(w), _p(dw),
                    _i(skf))
    return w.value


def contact_displacement(np_, nv, x, y, vx, vy, h, pi, pj, dx0=None, dy0=None):
    """contact_displacement of FORTRANsource/external_force.f90:191-244 on
    the pair list (0-based pi, pj); d starts at (dx0, dy0) or zero."""
    def f():
        n = len(x)
        X = vec3(x, y)
        V = vec3(vx, vy)
        H = hvec(h, n)
        D = vec3(np.zeros(n) if dx0 is None else dx0, np.zeros(n) if dy0 is None else dy0)
        P1 = np.ascontiguousarray(pi + 1, np.int32)
        P2 = np.ascontiguousarray(pj + 1, np.int32)
        lib90().contact_displacement_(_i(np_), _i(nv), _p(X), _p(V), _p(H), _p(D),
                                      _i(len(pi)), _p(P1), _p(P2))
        return D[:, 0].copy(), D[:, 1].copy()
    return run_big_stack(f)


def smooth_vector(np_, x, y, h, mass, rho, vx, vy, pi, pj, skf=0):
    """smooth_vector of FORTRANsource/external_force.f90:247-339."""
    def f():
        n = len(x)
        X = vec3(x, y)
        H = hvec(h, n)
        M = np.full(n, mass)
        R = np.ascontiguousarray(rho, np.float64)
        Vv = vec3(vx, vy)
        P1 = np.ascontiguousarray(pi + 1, np.int32)
        P2 = np.ascontiguousarray(pj + 1, np.int32)
        lib90().smooth_vector_(_i(np_), _p(X), _p(H), _p(M), _p(R), _p(Vv), _i(len(pi)),
                               _p(P1), _p(P2), _i(skf))
        return Vv[:, 0].copy(), Vv[:, 1].copy()
    return run_big_stack(f)


def link_list(x, y, h, z=None):
    def f():
        n = len(x)
        X = vec3(x, y, z)
        H = hvec(h, n)
        cap = n * 64 + 64
        pi = np.zeros(cap, np.int32)
        pj = np.zeros(cap, np.int32)
        niac = C.c_int(0)
        lib().link_list_(_i(n), _p(X), _p(H), C.byref(niac), _p(pi), _p(pj), _i(0))
        return pi[:niac.value] - 1, pj[:niac.value] - 1
    return run_big_stack(f)


def direct_find(x, y, h, z=None):
    def f():
        n = len(x)
        X = vec3(x, y, z)
        H = np.full(n, h)
        cap = n * 64 + 64
        pi = np.zeros(cap, np.int32)
        pj = np.zeros(cap, np.int32)
        niac = C.c_int(0)
        lib().direct_find_(_i(n), _p(X), _p(H), C.byref(niac), _p(pi), _p(pj), _i(0))
        return pi[:niac.value] - 1, pj[:niac.value] - 1
    return run_big_stack(f)


def pair_routines(x, y, vx, vy, rho, p, u, c, h, mass, pi, pj, z=None, vz=None):
    """Runs sum_density, con_density, int_force (pa_sph 0 and 1),
    art_visc, art_heat and av_vel with every particle counted as real.
    With z and vz the third components are real (the build's dim = 3) and
    the z parts of the vector outputs are returned too (*_az, xs_z)."""
    def f():
        n = len(x)
        X = vec3(x, y, z)
        V = vec3(vx, vy, vz)
        H = hvec(h, n)
        M = np.full(n, mass)
        P1 = np.ascontiguousarray(pi + 1, np.int32)
        P2 = np.ascontiguousarray(pj + 1, np.int32)
        niac = len(pi)
        out = {}
        rs = np.zeros(n)
        lib().sum_density_(_i(n), _p(X), _p(H), _p(M), _p(rs), _i(niac), _p(P1), _p(P2), _i(0))
        out["sum_rho"] = rs
        dr = np.zeros(n)
        lib().con_density_(_i(n), _p(X), _p(V), _p(H), _p(M), _p(dr), _i(niac), _p(P1),
                           _p(P2), _i(0))
        out["drho"] = dr
        R = np.ascontiguousarray(rho, np.float64)
        Pp = np.ascontiguousarray(p, np.float64)
        U = np.ascontiguousarray(u, np.float64)
        Cc = np.ascontiguousarray(c, np.float64)
        mat = np.ones(n, np.int32)
        eta = np.zeros(n)
        dt = C.c_double(1e-4)
        visc = C.c_int(0)
        for pa in (0, 1):
            dv = np.zeros((n, 3))
            du = np.zeros(n)
            tds = np.zeros(n)
            lib().int_force_(_i(n), _p(X), _p(V), _p(mat), _p(H), _p(M), _p(R), _p(Pp), _p(eta),
                             _p(dv), _p(du), _p(tds), _i(niac), _p(P1), _p(P2), C.byref(dt),
                             _i(0), _i(pa), C.byref(visc))
            out[f"if{pa}_ax"], out[f"if{pa}_ay"], out[f"if{pa}_du"] = dv[:, 0].copy(), dv[:, 1].copy(), du
            if z is not None:
                out[f"if{pa}_az"] = dv[:, 2].copy()
        dv = np.zeros((n, 3))
        du = np.zeros(n)
        lib().art_visc_(_i(n), _p(X), _p(V), _p(H), _p(M), _p(R), _p(Cc), _p(dv), _p(du),
                        _i(niac), _p(P1), _p(P2), _i(0))
        out["av_ax"], out["av_ay"], out["av_du"] = dv[:, 0].copy(), dv[:, 1].copy(), du
        if z is not None:
            out["av_az"] = dv[:, 2].copy()
        du = np.zeros(n)
        lib().art_heat_(_i(n), _p(X), _p(V), _p(H), _p(M), _p(R), _p(U), _p(Cc), _p(du),
                        _i(niac), _p(P1), _p(P2), _i(0))
        out["ah_du"] = du
        av = np.zeros((n, 3))
        lib().av_vel_(_i(n), _p(X), _p(V), _p(H), _p(M), _p(R), _p(av), _i(niac), _p(P1),
                      _p(P2), _i(0))
        out["xs_x"], out["xs_y"] = av[:, 0].copy(), av[:, 1].copy()
        if z is not None:
            out["xs_z"] = av[:, 2].copy()
        return out
    return run_big_stack(f)


def int_force_visc(x, y, vx, vy, rho, p, h, mass, eta, pi, pj, pa_sph):
    """internal_force.f with visc = .true. and a uniform eta, every particle
    counted (nt = n): returns dvdt x, dvdt y and dudt."""
    def f():
        n = len(x)
        X = vec3(x, y)
        V = vec3(vx, vy)
        H = np.full(n, h)
        M = np.full(n, mass)
        P1 = np.ascontiguousarray(pi + 1, np.int32)
        P2 = np.ascontiguousarray(pj + 1, np.int32)
        R = np.ascontiguousarray(rho, np.float64)
        Pp = np.ascontiguousarray(p, np.float64)
        mat = np.ones(n, np.int32)
        E = np.full(n, eta)
        dt = C.c_double(1e-4)
        visc = C.c_int(1)
        dv = np.zeros((n, 3))
        du = np.zeros(n)
        tds = np.zeros(n)
        lib().int_force_(_i(n), _p(X), _p(V), _p(mat), _p(H), _p(M), _p(R), _p(Pp), _p(E),
                         _p(dv), _p(du), _p(tds), _i(len(pi)), _p(P1), _p(P2), C.byref(dt),
                         _i(0), _i(pa_sph), C.byref(visc))
        return dv[:, 0].copy(), dv[:, 1].copy(), du, tds
    return run_big_stack(f)


def eos_poly(rho, u, rho0, a1, a2, a3, b0, b1, t1, t2, pmin):
    """eos.f eos_poly per particle."""
    def f():
        p = np.zeros(len(rho))
        for k in range(len(rho)):
            r = C.c_double(rho[k])
            e = C.c_double(u[k])
            out = C.c_double(0.0)
            args = [C.c_double(v) for v in (rho0, a1, a2, a3, b0, b1, t1, t2, pmin)]
            lib().eos_poly_(C.byref(r), C.byref(e), C.byref(out), *[C.byref(a) for a in args])
            p[k] = out.value
        return p
    return run_big_stack(f)


def h_upgrade(x, y, vx, vy, h, mass, rho, pi, pj, dt, sle):
    """hsml.f h_upgrade over all particles (nt = n), cubic kernel (skf 0);
    returns the new smoothing lengths.  The build has dim = 3: the update
    is dh/dt = (h/3) div v, or h = 2 (m/rho)^(1/3)."""
    def f():
        n = len(x)
        X = vec3(x, y)
        V = vec3(vx, vy)
        H = np.array(hvec(h, n))
        M = np.full(n, mass)
        R = np.ascontiguousarray(rho, np.float64)
        P1 = np.ascontiguousarray(pi + 1, np.int32)
        P2 = np.ascontiguousarray(pj + 1, np.int32)
        d = C.c_double(dt)
        lib().h_upgrade_(_i(n), _p(X), _p(V), _p(H), _p(M), _p(R), _i(len(pi)), _p(P1), _p(P2),
                         C.byref(d), _i(sle), _i(0))
        return H
    return run_big_stack(f)


def ghost_particles(x, y, vx, vy, h, mass, rho, p, np_, nv, pi, pj):
    """ghost.f ghost_particles over the given pair list (0-based, i < j);
    returns (ng, index_ghost (0-based ghost index per real particle, -1 for
    none), x, y, vx, vy, h, mass, rho, p of the n + ng particles)."""
    def f():
        n = np_ + nv
        cap = n + np_
        X = np.zeros((cap, 3))
        X[:n, 0], X[:n, 1] = x, y
        V = np.zeros((cap, 3))
        V[:n, 0], V[:n, 1] = vx, vy
        H = np.zeros(cap)
        H[:n] = hvec(h, n)
        M = np.zeros(cap)
        M[:n] = mass
        R = np.zeros(cap)
        R[:n] = rho
        Pp = np.zeros(cap)
        Pp[:n] = p
        P1 = np.ascontiguousarray(pi + 1, np.int32)
        P2 = np.ascontiguousarray(pj + 1, np.int32)
        ng = C.c_int(0)
        idx = np.zeros(max(np_, 1), np.int32)
        lib().ghost_particles_(_i(np_), _i(nv), C.byref(ng), _p(X), _p(V), _p(H), _p(M), _p(R),
                               _p(Pp), _i(len(pi)), _p(P1), _p(P2), _p(idx))
        k = n + ng.value
        return dict(ng=ng.value, index=idx[:np_].astype(np.int64) - 1, x=X[:k, 0].copy(),
                    y=X[:k, 1].copy(), vx=V[:k, 0].copy(), vy=V[:k, 1].copy(), h=H[:k].copy(),
                    mass=M[:k].copy(), rho=R[:k].copy(), p=Pp[:k].copy())
    return run_big_stack(f)


def fluka_heat(x, y, rho, origin, cell, nbin, data, z=None, tfluka=0):
    """fluka_heat.f fluka_heat (as shipped, dim = 3; Cartesian binning
    tfluka = 0, or 1 / 7 for its R-phi-z branch) for every particle, at
    z = 0 unless z is given; dedt starts at zero, as single_step.f:92-96
    leaves it (the 3D branch does not touch particles outside the
    binning)."""
    def f():
        n = len(x)
        X = vec3(x, y, z)
        R = np.ascontiguousarray(rho, np.float64)
        out = np.zeros(n)
        O = np.ascontiguousarray(origin, np.float64)
        D = np.ascontiguousarray(cell, np.float64)
        N = np.ascontiguousarray(nbin, np.int32)
        M = np.ascontiguousarray(np.asarray(data, np.float64).reshape(-1))
        lib().fluka_heat_(_i(n), _p(X), _p(R), _p(out), _p(O), _p(D), _p(N), _i(tfluka), _p(M))
        return out
    return run_big_stack(f)


def read_binning(kind, dirpath, translation, sf, ef, npart, dt, count):
    """read_binning_txt / read_binning_bin of fluka_heat.f on the file
    fluka_binning.FORM / fluka_binning in dirpath: (origin, cell, n, type,
    data[:count])."""
    def f():
        tr = np.ascontiguousarray(translation, np.float64)
        O = np.zeros(3)
        D = np.zeros(3)
        N = np.zeros(3, np.int32)
        T = C.c_int(-1)
        data = np.zeros(count)
        d = (dirpath.rstrip("/") + "/").encode()
        fn = lib().read_binning_txt_ if kind == "txt" else lib().read_binning_bin_
        dbl = lambda v: C.byref(C.c_double(v))  # noqa: E731
        fn(_p(tr), dbl(sf), dbl(ef), dbl(npart), dbl(dt), _p(O), _p(D), _p(N), C.byref(T),
           _p(data), C.c_char_p(d), C.c_size_t(len(d)))
        return O, D, N, T.value, data
    return run_big_stack(f)
