"""ctypes binding of the CPU oracle (oracle/_build/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the CPU
baseline leg of bench.py, never by the armando_amd package.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liborc.so")
REF_PATH = os.path.join(HERE, "_ref", "libarmref.so")
REF90_PATH = os.path.join(HERE, "_ref", "libarmref90.so")

K_WENDLAND2D, K_CUBIC2D, K_CUBIC3D, K_WENDLAND3D = 0, 1, 2, 3
ART_VISC, ART_HEAT, AV_VEL, DBC, GRAVITY, VISC, HEAT_SOURCE = 1, 2, 4, 8, 16, 32, 64
CONDUCTION, PERIODIC_X, GHOSTS, CONTACT = 128, 256, 512, 1024
EOS_TAIT, EOS_LINEAR = 0, 1


class Params(C.Structure):
    _fields_ = [
        ("kernel", C.c_int), ("flags", C.c_int), ("pa_sph", C.c_int),
        ("density_method", C.c_int),
        ("h", C.c_double), ("mass", C.c_double),
        ("rho0", C.c_double), ("c0", C.c_double), ("pmin", C.c_double),
        ("av_alpha", C.c_double), ("av_beta", C.c_double), ("av_eta", C.c_double),
        ("ah_g1", C.c_double), ("ah_g2", C.c_double), ("ah_eta", C.c_double),
        ("xsph_eps", C.c_double), ("gx", C.c_double), ("gy", C.c_double),
        ("dt", C.c_double),
        ("ox", C.c_double), ("oy", C.c_double), ("cell", C.c_double),
        ("nx", C.c_int), ("ny", C.c_int),
        ("visc_eta", C.c_double), ("eos", C.c_int), ("eos_p0", C.c_double),
        ("hs_q0", C.c_double), ("hs_x", C.c_double), ("hs_y", C.c_double),
        ("hs_sigma", C.c_double),
        ("dim", C.c_int), ("oz", C.c_double), ("gz", C.c_double), ("nz", C.c_int),
        ("cond_k", C.c_double), ("cp", C.c_double), ("sle", C.c_int),
        ("dep", C.POINTER(C.c_double)), ("dep_n", C.c_int * 3), ("dep_o", C.c_double * 3),
        ("dep_d", C.c_double * 3), ("dep_tend", C.c_double), ("dep_on", C.c_int),
        ("contact_hdr", C.c_double), ("dep_type", C.c_int),
    ]


class State(C.Structure):
    _fields_ = [("n", C.c_int), ("np", C.c_int)] + [
        (name, C.POINTER(C.c_double)) for name in (
            "x", "y", "vx", "vy", "rho", "u", "p", "c", "drho", "du", "ax", "ay",
            "avx", "avy", "divv", "rho_min", "u_min", "vx_min", "vy_min", "h")
    ] + [("nb", C.c_int), ("cap", C.c_int), ("have_prev", C.c_int),
         ("xp", C.POINTER(C.c_double)), ("yp", C.POINTER(C.c_double)),
         ("gpar", C.POINTER(C.c_int))]


STATE3_FIELDS = ("x", "y", "z", "vx", "vy", "vz", "rho", "u", "p", "c", "drho", "du", "ax",
                 "ay", "az", "avx", "avy", "avz", "divv", "rho_min", "u_min", "vx_min",
                 "vy_min", "vz_min")


class State3(C.Structure):
    _fields_ = [("n", C.c_int), ("np", C.c_int)] + [
        (name, C.POINTER(C.c_double)) for name in STATE3_FIELDS]


STATE_FIELDS = ("x", "y", "vx", "vy", "rho", "u", "p", "c", "drho", "du", "ax", "ay",
                "avx", "avy", "divv", "rho_min", "u_min", "vx_min", "vy_min", "h")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_PP = C.POINTER(Params)

_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
    L = C.CDLL(LIB_PATH)
    sig = {
        "orc_kernel": (C.c_double, [_PP, C.c_double, C.POINTER(C.c_double)]),
        "orc_link_list": (C.c_int, [_PP, C.c_int, _dp, _dp, C.c_int, C.POINTER(C.c_int), _ip, _ip]),
        "orc_direct_find": (C.c_int, [_PP, C.c_int, _dp, _dp, C.c_int, C.POINTER(C.c_int), _ip, _ip]),
        "orc_keys": (C.c_int, [_PP, C.c_int, _dp, _dp, _ip]),
        "orc_sort_cells": (None, [C.c_int, _ip, C.c_int, _ip, _ip]),
        "orc_sum_density": (None, [_PP, C.c_int, C.c_int, _ip, _ip, _dp, _dp, _dp]),
        "orc_con_density": (None, [_PP, C.c_int, C.c_int, _ip, _ip, _dp, _dp, _dp, _dp, _dp]),
        "orc_int_force": (None, [_PP, C.c_int, C.c_int, _ip, _ip, _dp, _dp, _dp, _dp, _dp, _dp,
                                 _dp, _dp, _dp]),
        "orc_art_visc": (None, [_PP, C.c_int, _ip, _ip, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                _dp]),
        "orc_art_heat": (None, [_PP, C.c_int, C.c_int, _ip, _ip, _dp, _dp, _dp, _dp, _dp, _dp,
                                _dp, _dp, _dp]),
        "orc_av_vel": (None, [_PP, C.c_int, _ip, _ip, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
        "orc_eos_tait": (None, [_PP, C.c_int, _dp, _dp, _dp]),
        "orc_eos": (None, [_PP, C.c_int, _dp, _dp, _dp]),
        "orc_fluka_heat": (None, [_PP, C.c_int, _dp, _dp, _dp, _dp]),
        "orc3_fluka_heat": (None, [_PP, C.c_int, _dp, _dp, _dp, _dp, _dp]),
        "orc_state_new": (C.POINTER(State), [C.c_int, C.c_int]),
        "orc_state_new_ghosts": (C.POINTER(State), [C.c_int, C.c_int]),
        "orc_ghost_particles": (C.c_int, [_PP, C.POINTER(State)]),
        "orc_state_free": (None, [C.POINTER(State)]),
        "orc_single_step": (C.c_int, [_PP, C.POINTER(State), C.c_int, _ip, _ip]),
        "orc_single_step_gather": (C.c_int, [_PP, C.POINTER(State)]),
        "orc_predictor": (None, [_PP, C.POINTER(State)]),
        "orc_corrector": (None, [_PP, C.POINTER(State), C.c_int]),
        "orc_run": (C.c_int, [_PP, C.POINTER(State), C.c_int, C.c_int, C.c_int]),
        "orc_conservation": (None, [_PP, C.POINTER(State), _dp]),
        "orc_adjoint": (C.c_int, [_PP, C.POINTER(State), C.c_int, _ip, C.POINTER(C.c_double),
                                  C.POINTER(C.c_double), _dp, _dp]),
        "orc_num_threads": (C.c_int, []),
        "orc_set_threads": (None, [C.c_int]),
        "orc_split_f32": (None, [_PP, C.c_int, _dp, _dp, _ip, _ip, _fp, _fp]),
        "orc_pairs_f32": (C.c_int, [_PP, C.c_int, _ip, _ip, _fp, _fp, C.c_int,
                                    C.POINTER(C.c_int), _ip, _ip]),
        "orc_pairs_f64": (C.c_int, [_PP, C.c_int, _dp, _dp, C.c_int, C.POINTER(C.c_int), _ip,
                                    _ip]),
        "orc_clist_f32": (None, [C.c_int, _fp, _fp, C.c_float, C.c_float, C.c_float, C.c_int,
                                 C.c_int, C.c_float, C.c_int, _ip, _ip]),
        "orc3_state_new": (C.POINTER(State3), [C.c_int, C.c_int]),
        "orc3_state_free": (None, [C.POINTER(State3)]),
        "orc3_keys": (C.c_int, [_PP, C.c_int, _dp, _dp, _dp, _ip]),
        "orc3_pairs_f64": (C.c_int, [_PP, C.c_int, _dp, _dp, _dp, C.c_int, C.POINTER(C.c_int),
                                     _ip, _ip]),
        "orc3_split_f32": (None, [_PP, C.c_int, _dp, _dp, _dp, _ip, _ip, _ip, _fp, _fp, _fp]),
        "orc3_pairs_f32": (C.c_int, [_PP, C.c_int, _ip, _ip, _ip, _fp, _fp, _fp, C.c_int,
                                     C.POINTER(C.c_int), _ip, _ip]),
        "orc3_routines": (C.c_int, [_PP, C.c_int] + [_dp] * 10 + [_dp] * 9),
        "orc3_single_step_gather": (C.c_int, [_PP, C.POINTER(State3)]),
        "orc3_predictor": (None, [_PP, C.POINTER(State3)]),
        "orc3_corrector": (None, [_PP, C.POINTER(State3), C.c_int]),
        "orc3_run": (C.c_int, [_PP, C.POINTER(State3), C.c_int, C.c_int]),
        "orc_contact_displacement": (None, [_PP, C.c_int, C.c_int, _dp, _dp, _dp, _dp, C.c_int,
                                            _ip, _ip, _dp, _dp]),
        "orc_smooth_vector": (None, [_PP, C.c_int, _dp, _dp, _dp, C.c_int, _ip, _ip, _dp, _dp]),
        "orc_kat_interaction_f32": (None, [C.c_int, C.c_int, _ip, _fp, _fp, _fp, _fp, _fp, _fp,
                                           C.c_float, C.c_float, C.c_float, _fp, _fp, _fp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def params(**kw):
    """Params with every field defaulted to zero, then overridden."""
    P = Params()
    for k, v in kw.items():
        setattr(P, k, v)
    return P


def params_from(case):
    """Oracle Params from a case dictionary (see armando_amd.cases)."""
    keys = [f for f, _ in Params._fields_]
    return params(**{k: case[k] for k in keys if k in case})


def set_deposition(P, origin, cell, n, data, t_end=float("inf"), on=1, typ=0):
    """Attach a deposition map (W/m^3, shape (n[2], n[1], n[0]) or flat with
    the first axis fastest) to P; P keeps a reference to the array.  typ:
    FLUKA's binning type, 0 Cartesian (x, y, z), 1 or 7 R-phi-z (r, phi in
    degrees, z)."""
    arr = np.ascontiguousarray(np.asarray(data, dtype=np.float64).reshape(-1))
    assert arr.size == int(n[0]) * int(n[1]) * int(n[2])
    P._dep_keep = arr
    P.dep = arr.ctypes.data_as(C.POINTER(C.c_double))
    for k in range(3):
        P.dep_n[k] = int(n[k])
        P.dep_o[k] = float(origin[k])
        P.dep_d[k] = float(cell[k])
    P.dep_tend = float(t_end)
    P.dep_on = int(on)
    if int(typ) % 10 not in (0, 1, 7):
        raise ValueError(f"binning type {typ}")
    P.dep_type = 0 if int(typ) % 10 == 0 else 1
    return P


def fluka_heat(P, x, y, rho, np_real=None):
    """dedt of fluka_heat.f (2D, Cartesian binning) for particles 0 .. np-1."""
    x, y, rho = _d(x), _d(y), _d(rho)
    n = len(x) if np_real is None else np_real
    out = np.zeros(n)
    lib().orc_fluka_heat(C.byref(P), n, x, y, rho, out)
    return out


def fluka_heat3(P, x, y, z, rho):
    """dedt of fluka_heat.f (3D, Cartesian binning) for every particle."""
    x, y, z, rho = _d(x), _d(y), _d(z), _d(rho)
    out = np.zeros(len(x))
    lib().orc3_fluka_heat(C.byref(P), len(x), x, y, z, rho, out)
    return out


# ------------------------------------------------------------------ pairs
def link_list(P, x, y, cap=None):
    n = len(x)
    cap = cap or max(64, n * 16)
    pi = np.zeros(cap, np.int32)
    pj = np.zeros(cap, np.int32)
    niac = C.c_int(0)
    rc = lib().orc_link_list(C.byref(P), n, np.ascontiguousarray(x, np.float64),
                             np.ascontiguousarray(y, np.float64), cap, C.byref(niac), pi, pj)
    if rc:
        raise RuntimeError(f"orc_link_list rc={rc}")
    return pi[:niac.value].copy(), pj[:niac.value].copy()


def direct_find(P, x, y, cap=None):
    n = len(x)
    cap = cap or max(64, n * 16)
    pi = np.zeros(cap, np.int32)
    pj = np.zeros(cap, np.int32)
    niac = C.c_int(0)
    rc = lib().orc_direct_find(C.byref(P), n, np.ascontiguousarray(x, np.float64),
                               np.ascontiguousarray(y, np.float64), cap, C.byref(niac), pi, pj)
    if rc:
        raise RuntimeError(f"orc_direct_find rc={rc}")
    return pi[:niac.value].copy(), pj[:niac.value].copy()


def pairs_f64(P, x, y, cap=None):
    n = len(x)
    cap = cap or max(64, n * 16)
    pi = np.zeros(cap, np.int32)
    pj = np.zeros(cap, np.int32)
    niac = C.c_int(0)
    rc = lib().orc_pairs_f64(C.byref(P), n, np.ascontiguousarray(x, np.float64),
                             np.ascontiguousarray(y, np.float64), cap, C.byref(niac), pi, pj)
    if rc:
        raise RuntimeError(f"orc_pairs_f64 rc={rc}")
    return pi[:niac.value].copy(), pj[:niac.value].copy()


def split_f32(P, x, y):
    n = len(x)
    cx = np.zeros(n, np.int32)
    cy = np.zeros(n, np.int32)
    ox = np.zeros(n, np.float32)
    oy = np.zeros(n, np.float32)
    lib().orc_split_f32(C.byref(P), n, np.ascontiguousarray(x, np.float64),
                        np.ascontiguousarray(y, np.float64), cx, cy, ox, oy)
    return cx, cy, ox, oy


def pairs_f32(P, cx, cy, offx, offy, cap=None):
    n = len(cx)
    cap = cap or max(64, n * 16)
    pi = np.zeros(cap, np.int32)
    pj = np.zeros(cap, np.int32)
    niac = C.c_int(0)
    rc = lib().orc_pairs_f32(C.byref(P), n, np.ascontiguousarray(cx, np.int32),
                             np.ascontiguousarray(cy, np.int32),
                             np.ascontiguousarray(offx, np.float32),
                             np.ascontiguousarray(offy, np.float32), cap, C.byref(niac), pi, pj)
    if rc:
        raise RuntimeError(f"orc_pairs_f32 rc={rc}")
    return pi[:niac.value].copy(), pj[:niac.value].copy()


def keys(P, x, y):
    n = len(x)
    k = np.zeros(n, np.int32)
    bad = lib().orc_keys(C.byref(P), n, np.ascontiguousarray(x, np.float64),
                         np.ascontiguousarray(y, np.float64), k)
    return k, bad


def sort_cells(key, ncell):
    n = len(key)
    sid = np.zeros(n, np.int32)
    start = np.zeros(ncell + 1, np.int32)
    lib().orc_sort_cells(n, np.ascontiguousarray(key, np.int32), ncell, sid, start)
    return sid, start


# ------------------------------------------------------------------ state
class OracleState:
    """Owns an orc_state; fields are exposed as numpy views."""

    def __init__(self, n, np_fluid):
        self.ptr = lib().orc_state_new(n, np_fluid)
        self.n = n
        self.np = np_fluid

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().orc_state_free(self.ptr)
            self.ptr = None

    def field(self, name):
        p = getattr(self.ptr.contents, name)
        return np.ctypeslib.as_array(p, shape=(self.n,))

    def set(self, **kw):
        for k, v in kw.items():
            self.field(k)[:] = v

    def get(self, *names):
        return {k: self.field(k).copy() for k in names}


class GhostState(OracleState):
    """An orc_state of np real and nv virtual particles with room for np
    ghosts; fields are views of the live particles (n = nb + ghosts)."""

    def __init__(self, np_fluid, nv):
        self.ptr = lib().orc_state_new_ghosts(np_fluid, nv)
        self.np = np_fluid
        self.nb = np_fluid + nv

    @property
    def n(self):
        return self.ptr.contents.n

    @n.setter
    def n(self, value):
        pass

    @property
    def ng(self):
        return self.ptr.contents.n - self.nb

    def set_prev(self, x, y):
        """Positions of the last neighbour search (real particles)."""
        c = self.ptr.contents
        np.ctypeslib.as_array(c.xp, shape=(self.np,))[:] = x[:self.np]
        np.ctypeslib.as_array(c.yp, shape=(self.np,))[:] = y[:self.np]
        c.have_prev = 1

    def parents(self):
        return np.ctypeslib.as_array(self.ptr.contents.gpar, shape=(max(self.ng, 1),))[
            :self.ng].copy()


def ghost_particles(P, st):
    ng = lib().orc_ghost_particles(C.byref(P), st.ptr)
    if ng < 0:
        raise RuntimeError("orc_ghost_particles: capacity exceeded")
    return ng


def run(P, st, its0, nsteps, gather=True):
    rc = lib().orc_run(C.byref(P), st.ptr, its0, nsteps, 1 if gather else 0)
    if rc:
        raise RuntimeError(f"orc_run rc={rc}")


def single_step_gather(P, st):
    rc = lib().orc_single_step_gather(C.byref(P), st.ptr)
    if rc:
        raise RuntimeError(f"orc_single_step_gather rc={rc}")


def conservation(P, st):
    out = np.zeros(5, np.float64)
    lib().orc_conservation(C.byref(P), st.ptr, out)
    return out


def predictor(P, st):
    """The half step of time_integration.f:136-154 on the stored rates."""
    lib().orc_predictor(C.byref(P), st.ptr)


def corrector(P, st, its):
    """The update of time_integration.f:186-217 for step its."""
    lib().orc_corrector(C.byref(P), st.ptr, its)


def adjoint(P, st, nsteps, objmask):
    """J and dJ/d(alpha, vx0, vy0) by the oracle's discrete adjoint."""
    n = st.n
    J = C.c_double(0)
    da = C.c_double(0)
    gx = np.zeros(n)
    gy = np.zeros(n)
    rc = lib().orc_adjoint(C.byref(P), st.ptr, nsteps, np.ascontiguousarray(objmask, np.int32),
                           C.byref(J), C.byref(da), gx, gy)
    if rc:
        raise RuntimeError(f"orc_adjoint rc={rc}")
    return J.value, da.value, gx, gy


# ------------------------------------------------------------------ 3D
def _d(a):
    return np.ascontiguousarray(a, np.float64)


def keys3(P, x, y, z):
    k = np.zeros(len(x), np.int32)
    bad = lib().orc3_keys(C.byref(P), len(x), _d(x), _d(y), _d(z), k)
    return k, bad


def pairs3_f64(P, x, y, z, cap=None):
    n = len(x)
    cap = cap or max(64, n * 64)
    pi = np.zeros(cap, np.int32)
    pj = np.zeros(cap, np.int32)
    niac = C.c_int(0)
    rc = lib().orc3_pairs_f64(C.byref(P), n, _d(x), _d(y), _d(z), cap, C.byref(niac), pi, pj)
    if rc:
        raise RuntimeError(f"orc3_pairs_f64 rc={rc}")
    return pi[:niac.value].copy(), pj[:niac.value].copy()


def split3_f32(P, x, y, z):
    n = len(x)
    cells = [np.zeros(n, np.int32) for _ in range(3)]
    offs = [np.zeros(n, np.float32) for _ in range(3)]
    lib().orc3_split_f32(C.byref(P), n, _d(x), _d(y), _d(z), *cells, *offs)
    return cells, offs


def pairs3_f32(P, cells, offs, cap=None):
    n = len(cells[0])
    cap = cap or max(64, n * 64)
    pi = np.zeros(cap, np.int32)
    pj = np.zeros(cap, np.int32)
    niac = C.c_int(0)
    rc = lib().orc3_pairs_f32(C.byref(P), n, *[np.ascontiguousarray(c, np.int32) for c in cells],
                              *[np.ascontiguousarray(o, np.float32) for o in offs], cap,
                              C.byref(niac), pi, pj)
    if rc:
        raise RuntimeError(f"orc3_pairs_f32 rc={rc}")
    return pi[:niac.value].copy(), pj[:niac.value].copy()


def routines3(P, x, y, z, vx, vy, vz, rho, p, u, c):
    """Per-routine outputs of one 3D step in gather form: a dict of
    rho_sum, drho, if_a (n, 3), if_du, av_a (n, 3), av_du, ah_du, divv,
    xs (n, 3)."""
    n = len(x)
    out = {k: np.zeros(n) for k in ("rho_sum", "drho", "if_du", "av_du", "ah_du", "divv")}
    for k in ("if_a", "av_a", "xs"):
        out[k] = np.zeros((n, 3))
    rc = lib().orc3_routines(C.byref(P), n, *[_d(a) for a in (x, y, z, vx, vy, vz, rho, p, u, c)],
                             out["rho_sum"], out["drho"], out["if_a"], out["if_du"], out["av_a"],
                             out["av_du"], out["ah_du"], out["divv"], out["xs"])
    if rc:
        raise RuntimeError(f"orc3_routines rc={rc}")
    return out


class OracleState3:
    """Owns an orc3_state; fields are exposed as numpy views."""

    def __init__(self, n, np_fluid):
        self.ptr = lib().orc3_state_new(n, np_fluid)
        self.n = n
        self.np = np_fluid

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().orc3_state_free(self.ptr)
            self.ptr = None

    def field(self, name):
        p = getattr(self.ptr.contents, name)
        return np.ctypeslib.as_array(p, shape=(self.n,))

    def set(self, **kw):
        for k, v in kw.items():
            self.field(k)[:] = v

    def get(self, *names):
        return {k: self.field(k).copy() for k in names}


def run3(P, st, its0, nsteps):
    rc = lib().orc3_run(C.byref(P), st.ptr, its0, nsteps)
    if rc:
        raise RuntimeError(f"orc3_run rc={rc}")


def single_step3(P, st):
    rc = lib().orc3_single_step_gather(C.byref(P), st.ptr)
    if rc:
        raise RuntimeError(f"orc3_single_step_gather rc={rc}")
