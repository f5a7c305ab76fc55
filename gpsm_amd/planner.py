"""Learned join-order planner: weights on disk, installation into a context.

The network (32 -> 128 -> 128 -> 1, ReLU, BF16 weights, FP32 accumulation)
runs on the device through the C ABI (gpsm_mlp_*); this module only stores
and loads its FP32 master parameters.
"""
import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import check, ptr

WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "planner_weights.npy")
NORM = os.path.join(os.path.dirname(os.path.abspath(__file__)), "planner_norm.npy")


class MLP:
    def __init__(self, ctx, seed=0):
        h = C.c_void_p()
        check(_lib.lib().gpsm_mlp_create(ctx.h, seed, C.byref(h)), "gpsm_mlp_create")
        self.h = h
        self.ctx = ctx

    def free(self):
        if self.h:
            _lib.lib().gpsm_mlp_free(self.h)
            self.h = None

    def get(self):
        p = np.zeros(_lib.N_PARAMS, dtype=np.float32)
        check(_lib.lib().gpsm_mlp_get_params(self.h, ptr(p)), "gpsm_mlp_get_params")
        return p

    def set(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.shape == (_lib.N_PARAMS,)
        check(_lib.lib().gpsm_mlp_set_params(self.h, ptr(p)), "gpsm_mlp_set_params")

    def set_norm(self, mu, inv_sd):
        mu = np.ascontiguousarray(mu, dtype=np.float32)
        inv_sd = np.ascontiguousarray(inv_sd, dtype=np.float32)
        check(_lib.lib().gpsm_mlp_set_input_norm(self.h, ptr(mu), ptr(inv_sd)), "gpsm_mlp_set_input_norm")

    def get_norm(self):
        mu = np.zeros(_lib.FEAT, dtype=np.float32)
        isd = np.zeros(_lib.FEAT, dtype=np.float32)
        check(_lib.lib().gpsm_mlp_get_input_norm(self.h, ptr(mu), ptr(isd)), "gpsm_mlp_get_input_norm")
        return mu, isd

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros(x.shape[0], dtype=np.float32)
        check(_lib.lib().gpsm_mlp_forward(self.h, ptr(x), x.shape[0], ptr(y)), "gpsm_mlp_forward")
        return y

    def grad(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(_lib.N_PARAMS, dtype=np.float32)
        loss = C.c_float(0)
        check(_lib.lib().gpsm_mlp_grad(self.h, ptr(x), ptr(t), x.shape[0], C.byref(loss), ptr(g)), "gpsm_mlp_grad")
        return loss.value, g

    def train_step(self, x, t, lr):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = C.c_float(0)
        check(_lib.lib().gpsm_mlp_train_step(self.h, ptr(x), ptr(t), x.shape[0], lr, C.byref(loss)),
              "gpsm_mlp_train_step")
        return loss.value


def install(ctx, path=WEIGHTS, norm_path=NORM):
    """Loads trained weights (if present) into a device MLP and makes it the
    context's planner.  Returns the MLP, or None without weights."""
    if not os.path.exists(path) or not os.path.exists(norm_path):
        return None
    m = MLP(ctx)
    m.set(np.load(path, allow_pickle=False))
    norm = np.load(norm_path, allow_pickle=False)
    m.set_norm(norm[:_lib.FEAT], norm[_lib.FEAT:])
    check(_lib.lib().gpsm_set_planner(ctx.h, m.h), "gpsm_set_planner")
    if ctx.planner is not None:  # replaced: release the previous network
        ctx.planner.free()
    ctx.planner = m
    return m
