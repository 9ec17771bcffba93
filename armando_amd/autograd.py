"""torch.autograd over the adjoint rollout.

`RolloutLoss(case, obj_ids, nsteps, ckpt_every)(v0, alpha)` returns the
objective J (a 0-d float64 tensor) of `nsteps` of time_integration from the
case's initial state with initial velocities v0 (n x 2) and artificial
viscosity alpha.  Its backward is the hand-written adjoint of
libarmando.so (checkpointed every `ckpt_every` steps), wrapped by the C++
torch::autograd::Function in csrc/autograd.cpp.
"""
import importlib.machinery
import importlib.util
import os
import sysconfig

import numpy as np
import torch

from . import PARAM_DEFAULTS, ArmParams, load_library

HERE = os.path.dirname(os.path.abspath(__file__))
_ext = None


def extension():
    global _ext
    if _ext is None:
        load_library()
        path = os.path.join(HERE, "armando_autograd" + sysconfig.get_config_var("EXT_SUFFIX"))
        if not os.path.exists(path):
            raise RuntimeError(f"{path} not built: run __graft_entry__.build()")
        loader = importlib.machinery.ExtensionFileLoader("armando_autograd", path)
        spec = importlib.util.spec_from_loader("armando_autograd", loader)
        mod = importlib.util.module_from_spec(spec)
        loader.exec_module(mod)
        _ext = mod
    return _ext


class RolloutLoss:
    """J(v0, alpha, q0, eta) = sum over nsteps of dt * sum_{obj} f, f = p^2
    (objective OBJ_PRESSURE) or u^2 (OBJ_ENERGY), as a torch function whose
    backward is the hand-written adjoint (checkpoint every ckpt_every steps).
    v0: (n, 2) initial velocities; alpha: artificial viscosity; q0: heat
    source amplitude (used when the case has the heat source); eta: dynamic
    viscosity (used when the case has the physical viscosity)."""

    def __init__(self, case, obj_ids, nsteps, ckpt_every, objective=0):
        ext = extension()
        params = {k: case[k] if k in case else PARAM_DEFAULTS[k] for k, _ in ArmParams._fields_}
        x = torch.from_numpy(np.stack([case["x"], case["y"]], 1).astype(np.float64))
        rho = torch.from_numpy(np.asarray(case["rho"], np.float64))
        u = torch.from_numpy(np.asarray(case["u"], np.float64))
        obj = torch.from_numpy(np.asarray(obj_ids, np.int32))
        self.n = int(case["n"])
        self.q0 = float(params["hs_q0"])
        self.eta = float(params["visc_eta"])
        self.handle = ext.make_rollout(params, x, rho, u, int(case["np"]), int(case["nv"]), obj,
                                       int(nsteps), int(ckpt_every), int(objective))

    def __call__(self, v0, alpha, q0=None, eta=None):
        if q0 is None:
            q0 = torch.tensor(self.q0, dtype=torch.float64)
        if eta is None:
            eta = torch.tensor(self.eta, dtype=torch.float64)
        return extension().rollout_loss(self.handle, v0, alpha, q0, eta)

    def close(self):
        if getattr(self, "handle", 0):
            extension().free_rollout(self.handle)
            self.handle = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
