"""armando_amd -- MI355X (gfx950) WCSPH time step of the Armando code.

The product is the C-ABI shared library `armando_amd/libarmando.so` built
from `armando_amd/csrc` (declared in `include/armando.h`).  This module is a
thin ctypes mirror of that interface whose methods carry the reference's
routine names (link_list, density, eos, internal_force, av_vel, single_step,
time_integration) so that scripts and tests read like the reference loop.

There is no CPU fallback: if the library is missing or no HIP device is
present, the calls raise.
"""
import ctypes as C
import os

import numpy as np

from . import cases  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libarmando.so")

ARM_OK = 0
ERRORS = {-1: "ARM_ERR_ARG", -2: "ARM_ERR_HIP", -3: "ARM_ERR_CAPACITY", -4: "ARM_ERR_DOMAIN",
          -5: "ARM_ERR_UNSUPPORTED", -6: "ARM_ERR_STATE"}
ART_VISC, ART_HEAT, AV_VEL, DBC, GRAVITY, VISC, HEAT_SOURCE = 1, 2, 4, 8, 16, 32, 64
CONDUCTION, PERIODIC_X, GHOSTS, CONTACT = 128, 256, 512, 1024
EOS_TAIT, EOS_LINEAR = 0, 1
FP64, FP32 = 0, 1
PASSES = ("keys", "sort", "reorder", "density", "eos", "force", "integrate")

EXPORTS = ("arm_create", "arm_destroy", "arm_last_error", "arm_domain_violations", "arm_upload",
           "arm_download", "arm_download_rates", "arm_link_list", "arm_get_pairs",
           "arm_get_cells", "arm_density", "arm_eos", "arm_internal_force", "arm_av_vel",
           "arm_single_step", "arm_predictor", "arm_corrector", "arm_time_integration",
           "arm_conservation", "arm_pass_times", "arm_reset_pass_times", "arm_set_timing",
           "arm_set_gather_path", "arm_adjoint_rollout", "arm_rollout_forward",
           "arm_rollout_backward", "arm_update_params", "arm_set_objective",
           "arm_rollout_param_grads", "arm_conservation3", "arm_upload_h", "arm_download_h",
           "arm_ghost_particles", "arm_ghost_count", "arm_download_ghosts",
           "arm_set_deposition", "arm_set_deposition_binning", "arm_link_list_static")
ARM_STATIC_MAX_PARTICLES = 1 << 20  # arm_link_list_static limits (include/armando.h)
ARM_STATIC_MAX_CELLS = 1 << 20


class ArmParams(C.Structure):
    _fields_ = [
        ("precision", C.c_int), ("flags", C.c_int), ("pa_sph", C.c_int),
        ("density_method", C.c_int),
        ("h", C.c_double), ("mass", C.c_double), ("rho0", C.c_double), ("c0", C.c_double),
        ("pmin", C.c_double), ("av_alpha", C.c_double), ("av_beta", C.c_double),
        ("av_eta", C.c_double), ("ah_g1", C.c_double), ("ah_g2", C.c_double),
        ("ah_eta", C.c_double), ("xsph_eps", C.c_double), ("gx", C.c_double),
        ("gy", C.c_double), ("dt", C.c_double), ("ox", C.c_double), ("oy", C.c_double),
        ("cell", C.c_double), ("nx", C.c_int), ("ny", C.c_int),
        ("visc_eta", C.c_double), ("eos", C.c_int), ("eos_p0", C.c_double),
        ("hs_q0", C.c_double), ("hs_x", C.c_double), ("hs_y", C.c_double),
        ("hs_sigma", C.c_double),
        ("dim", C.c_int), ("oz", C.c_double), ("gz", C.c_double), ("nz", C.c_int),
        ("cond_k", C.c_double), ("cp", C.c_double), ("sle", C.c_int),
        ("contact_hdr", C.c_double),
    ]


# fields a case may leave out, with their defaults (no physical viscosity,
# Tait EOS)
PARAM_DEFAULTS = {"visc_eta": 0.0, "eos": 0, "eos_p0": 0.0, "hs_q0": 0.0, "hs_x": 0.0,
                  "hs_y": 0.0, "hs_sigma": 0.0, "dim": 2, "oz": 0.0, "gz": 0.0, "nz": 1,
                  "cond_k": 0.0, "cp": 0.0, "sle": 0, "contact_hdr": 0.0}
OBJ_PRESSURE, OBJ_ENERGY = 0, 1


class ArmandoError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_lib = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def load_library(path=LIB_PATH):
    """Load the C-ABI library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} not built: run __graft_entry__.build()")
    L = C.CDLL(path)
    vp = C.c_void_p
    dpn = C.POINTER(C.c_double)
    sig = {
        "arm_create": (C.c_int, [C.POINTER(ArmParams), C.c_int, C.POINTER(vp)]),
        "arm_destroy": (None, [vp]),
        "arm_last_error": (C.c_char_p, [vp]),
        "arm_domain_violations": (C.c_int, [vp]),
        "arm_upload": (C.c_int, [vp, C.c_int, C.c_int, dpn, dpn, dpn, dpn]),
        "arm_download": (C.c_int, [vp, dpn, dpn, dpn, dpn, dpn]),
        "arm_download_rates": (C.c_int, [vp, dpn, dpn, dpn, dpn, dpn]),
        "arm_link_list": (C.c_int, [vp, C.POINTER(C.c_int)]),
        "arm_get_pairs": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int), _ip, _ip]),
        "arm_get_cells": (C.c_int, [vp, _ip, _ip, _ip]),
        "arm_density": (C.c_int, [vp, C.c_int]),
        "arm_eos": (C.c_int, [vp]),
        "arm_internal_force": (C.c_int, [vp, C.c_int]),
        "arm_av_vel": (C.c_int, [vp]),
        "arm_single_step": (C.c_int, [vp]),
        "arm_predictor": (C.c_int, [vp]),
        "arm_corrector": (C.c_int, [vp, C.c_int]),
        "arm_time_integration": (C.c_int, [vp, C.c_int, C.c_int]),
        "arm_conservation": (C.c_int, [vp, _dp]),
        "arm_conservation3": (C.c_int, [vp, _dp]),
        "arm_upload_h": (C.c_int, [vp, _dp]),
        "arm_download_h": (C.c_int, [vp, _dp]),
        "arm_pass_times": (C.c_int, [vp, C.c_int, dpn, C.POINTER(C.c_int)]),
        "arm_reset_pass_times": (C.c_int, [vp]),
        "arm_set_timing": (C.c_int, [vp, C.c_int]),
        "arm_set_gather_path": (C.c_int, [vp, C.c_int]),
        "arm_adjoint_rollout": (C.c_int, [vp, C.c_int, C.c_int, _ip, C.c_int, dpn, dpn, dpn,
                                          dpn, dpn]),
        "arm_rollout_forward": (C.c_int, [vp, C.c_int, C.c_int, _ip, C.c_int, dpn, dpn]),
        "arm_rollout_backward": (C.c_int, [vp, C.c_double, dpn, dpn, dpn]),
        "arm_update_params": (C.c_int, [vp, C.POINTER(ArmParams)]),
        "arm_set_objective": (C.c_int, [vp, C.c_int]),
        "arm_rollout_param_grads": (C.c_int, [vp, dpn, C.c_int]),
        "arm_set_deposition": (C.c_int, [vp, dpn, dpn, C.c_void_p, dpn, C.c_double]),
        "arm_set_deposition_binning": (C.c_int, [vp, C.c_int, dpn, dpn, C.c_void_p, dpn,
                                                 C.c_double]),
        # device pointers: x, key_sorted, id_sorted, cell_start, outside, stream
        "arm_link_list_static": (C.c_int, [C.POINTER(ArmParams), C.c_int] + [vp] * 6),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def params_from(case):
    P = ArmParams()
    for k, _ in ArmParams._fields_:
        setattr(P, k, case[k] if k in case else PARAM_DEFAULTS[k])
    return P


class Simulation:
    """A device context holding one particle set (one per host thread)."""

    def __init__(self, case, capacity=None):
        self.L = load_library()
        self.case = case
        self.n = int(case["n"])
        self.np = int(case["np"])
        self.params = params_from(case)
        self.dim = 3 if self.params.dim == 3 else 2
        h = C.c_void_p()
        if capacity is None:
            # mirror ghosts: room for the live set and the next one
            capacity = self.n + (2 * self.np if self.params.flags & GHOSTS else 0)
        rc = self.L.arm_create(C.byref(self.params), int(capacity), C.byref(h))
        self.ctx = h
        self._check(rc)
        X = np.ascontiguousarray(np.stack([case[k] for k in self._axes()], 1), np.float64)
        V = np.ascontiguousarray(np.stack([case["v" + k] for k in self._axes()], 1), np.float64)
        self._check(self.L.arm_upload(self.ctx, self.np, self.n - self.np, _ptr(X), _ptr(V),
                                      _ptr(np.ascontiguousarray(case["rho"], np.float64)),
                                      _ptr(np.ascontiguousarray(case["u"], np.float64))))

    def _axes(self):
        return ("x", "y", "z")[:self.dim]

    def _check(self, rc):
        if rc != ARM_OK:
            msg = self.L.arm_last_error(self.ctx).decode() if self.ctx else ""
            raise ArmandoError(rc, msg)

    def close(self):
        if getattr(self, "ctx", None):
            self.L.arm_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- the reference's routines ------------------------------------------
    def link_list(self, count=False):
        niac = C.c_int(0)
        self._check(self.L.arm_link_list(self.ctx, C.byref(niac) if count else None))
        return niac.value if count else None

    def density(self, method=0):
        self._check(self.L.arm_density(self.ctx, method))

    def eos(self):
        self._check(self.L.arm_eos(self.ctx))

    def internal_force(self, flags):
        self._check(self.L.arm_internal_force(self.ctx, flags))

    def av_vel(self):
        self._check(self.L.arm_av_vel(self.ctx))

    def single_step(self):
        self._check(self.L.arm_single_step(self.ctx))

    def predictor(self):
        self._check(self.L.arm_predictor(self.ctx))

    def corrector(self, its):
        self._check(self.L.arm_corrector(self.ctx, its))

    def time_integration(self, its0, nsteps):
        self._check(self.L.arm_time_integration(self.ctx, its0, nsteps))

    # ---- data ---------------------------------------------------------------
    def state(self):
        n, d = self.n, self.dim
        X = np.zeros((n, d))
        V = np.zeros((n, d))
        rho = np.zeros(n)
        p = np.zeros(n)
        u = np.zeros(n)
        self._check(self.L.arm_download(self.ctx, _ptr(X), _ptr(V), _ptr(rho), _ptr(p), _ptr(u)))
        out = dict(rho=rho, p=p, u=u)
        for a, k in enumerate(self._axes()):
            out[k] = X[:, a].copy()
            out["v" + k] = V[:, a].copy()
        return out

    # ---- mirror ghosts (ghost.f) -------------------------------------------
    def ghost_particles(self):
        ng = C.c_int(0)
        self._check(self.L.arm_ghost_particles(self.ctx, C.byref(ng)))
        return ng.value

    def ghost_count(self):
        ng = C.c_int(0)
        self._check(self.L.arm_ghost_count(self.ctx, C.byref(ng)))
        return ng.value

    def ghosts(self):
        """The ghosts of the current state in number order."""
        ng = self.ghost_count()
        X = np.zeros((max(ng, 1), 2))
        V = np.zeros((max(ng, 1), 2))
        rho, p, u = np.zeros(max(ng, 1)), np.zeros(max(ng, 1)), np.zeros(max(ng, 1))
        par = np.zeros(max(ng, 1), np.int32)
        self._check(self.L.arm_download_ghosts(self.ctx, _ptr(X), _ptr(V), _ptr(rho), _ptr(p),
                                               _ptr(u), _ptr(par)))
        return dict(x=X[:ng, 0].copy(), y=X[:ng, 1].copy(), vx=V[:ng, 0].copy(),
                    vy=V[:ng, 1].copy(), rho=rho[:ng], p=p[:ng], u=u[:ng], parent=par[:ng])

    def rates(self):
        n, d = self.n, self.dim
        dr = np.zeros(n)
        A = np.zeros((n, d))
        du = np.zeros(n)
        W = np.zeros((n, d))
        dv = np.zeros(n)
        self._check(self.L.arm_download_rates(self.ctx, _ptr(dr), _ptr(A), _ptr(du), _ptr(W),
                                              _ptr(dv)))
        out = dict(drho=dr, du=du, divv=dv)
        for a, k in enumerate(self._axes()):
            out["a" + k] = A[:, a].copy()
            out["av" + k] = W[:, a].copy()
        return out

    def pairs(self, cap=None):
        niac = C.c_int(0)
        cap = cap or self.n * (16 if self.dim == 2 else 64) + 64
        pi = np.zeros(cap, np.int32)
        pj = np.zeros(cap, np.int32)
        self._check(self.L.arm_get_pairs(self.ctx, cap, C.byref(niac), pi, pj))
        return pi[:niac.value].copy(), pj[:niac.value].copy()

    def cells(self):
        ncell = int(self.case["nx"]) * int(self.case["ny"]) * \
            (int(self.case["nz"]) if self.dim == 3 else 1)
        live = self.n + (self.ghost_count() if self.params.flags & GHOSTS else 0)
        k = np.zeros(live, np.int32)
        s = np.zeros(live, np.int32)
        cs = np.zeros(ncell + 1, np.int32)
        self._check(self.L.arm_get_cells(self.ctx, k, s, cs))
        return k, s, cs

    def conservation(self):
        """mass, momentum x, momentum y, kinetic energy, internal energy."""
        out = np.zeros(5)
        self._check(self.L.arm_conservation(self.ctx, out))
        return out

    def set_h(self, h):
        """Per-particle smoothing lengths (sle != 0), original order."""
        self._check(self.L.arm_upload_h(self.ctx, np.ascontiguousarray(h, np.float64)))

    def h(self):
        """Per-particle smoothing lengths (sle != 0), original order."""
        out = np.zeros(self.n)
        self._check(self.L.arm_download_h(self.ctx, out))
        return out

    def conservation3(self):
        """mass, momentum x, y, z, kinetic energy, internal energy."""
        out = np.zeros(6)
        self._check(self.L.arm_conservation3(self.ctx, out))
        return out

    def adjoint_rollout(self, nsteps, ckpt_every, obj_ids):
        """J, dJ/dalpha, dJ/dv0 (n, dim) and (forward, backward) device ms."""
        ids = np.ascontiguousarray(obj_ids, np.int32)
        J = C.c_double(0)
        da = C.c_double(0)
        g = np.zeros((self.n, self.dim))
        tf = C.c_double(0)
        tb = C.c_double(0)
        self._check(self.L.arm_adjoint_rollout(self.ctx, nsteps, ckpt_every, ids, len(ids),
                                               C.byref(J), C.byref(da), _ptr(g), C.byref(tf),
                                               C.byref(tb)))
        return J.value, da.value, g, (tf.value, tb.value)

    def rollout_forward(self, nsteps, ckpt_every, obj_ids):
        """J and the forward device ms; keeps the checkpoints for
        rollout_backward."""
        ids = np.ascontiguousarray(obj_ids, np.int32)
        J = C.c_double(0)
        tf = C.c_double(0)
        self._check(self.L.arm_rollout_forward(self.ctx, nsteps, ckpt_every, ids, len(ids),
                                               C.byref(J), C.byref(tf)))
        return J.value, tf.value

    def rollout_backward(self, seed=1.0):
        """seed * (dJ/dalpha, dJ/dv0 (n, dim)) and the backward device ms."""
        da = C.c_double(0)
        g = np.zeros((self.n, self.dim))
        tb = C.c_double(0)
        self._check(self.L.arm_rollout_backward(self.ctx, seed, C.byref(da), _ptr(g),
                                                C.byref(tb)))
        return da.value, g, tb.value

    def set_deposition(self, binning=None, t_end=float("inf")):
        """The energy-deposition heat source of fluka_heat.f: `binning` a
        dict with origin (3), cell (3), n (3), data (n0 n1 n2 values in W/m^3,
        first axis fastest) and optionally type (FLUKA's tfluka: 0 Cartesian
        x, y, z; 1 or 7 R-phi-z, r, phi in degrees, z), as
        `armando_amd.io.read_fluka_*` return; it acts in the steps with
        its dt <= t_end.  None switches the source off."""
        if binning is None:
            self._check(self.L.arm_set_deposition(self.ctx, None, None, None, None, 0.0))
            return
        o = np.ascontiguousarray(binning["origin"], np.float64)
        d = np.ascontiguousarray(binning["cell"], np.float64)
        nb = np.ascontiguousarray(binning["n"], np.int32)
        data = np.ascontiguousarray(np.asarray(binning["data"], np.float64).reshape(-1))
        if o.size != 3 or d.size != 3 or nb.size != 3 or data.size != int(np.prod(nb.astype(np.int64))):
            raise ValueError("binning needs origin, cell and n of 3 entries and n0*n1*n2 values")
        self._check(self.L.arm_set_deposition_binning(self.ctx, int(binning.get("type", 0)),
                                                       _ptr(o), _ptr(d),
                                                       nb.ctypes.data_as(C.c_void_p), _ptr(data),
                                                       float(t_end)))

    def set_objective(self, kind):
        """OBJ_PRESSURE (J = sum dt sum p^2) or OBJ_ENERGY (u^2) for the next rollout."""
        self._check(self.L.arm_set_objective(self.ctx, int(kind)))

    def param_grads(self):
        """(dJ/d av_alpha, dJ/d hs_q0, dJ/d visc_eta) of the last rollout backward."""
        out = np.zeros(3)
        self._check(self.L.arm_rollout_param_grads(self.ctx, _ptr(out), 3))
        return float(out[0]), float(out[1]), float(out[2])

    def set_gather_path(self, tiled=True):
        self._check(self.L.arm_set_gather_path(self.ctx, 1 if tiled else 0))

    def set_timing(self, on=True):
        self._check(self.L.arm_set_timing(self.ctx, 1 if on else 0))

    def reset_pass_times(self):
        self._check(self.L.arm_reset_pass_times(self.ctx))

    def pass_times(self):
        ms = (C.c_double * len(PASSES))()
        cnt = (C.c_int * len(PASSES))()
        self._check(self.L.arm_pass_times(self.ctx, len(PASSES), ms, cnt))
        return {PASSES[k]: (ms[k], cnt[k]) for k in range(len(PASSES))}
