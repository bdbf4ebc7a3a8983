"""Host side of the learned pair screen (device work goes through the C ABI).

A pair screen predicts, from cheap features, the MI the plain pass would
give a pair, so that pairs predicted below a cutoff can be left out of the
full pass (the reference's A-MIE / A-MIST take such a reduced pair list from
a user neighbour file, Centre.cpp:459-593).  The network and its features
are described in ``include/centre_hip.h``; this module initialises the
parameters, turns MI into the regression target and back, and selects the
pairs to keep.
"""
from __future__ import annotations

import numpy as np

from . import hip

F, H = 8, 64
N_PARAMS = H * F + H + H * H + H + H + 1 + F + F
OFF_MU = N_PARAMS - 2 * F
OFF_SD = N_PARAMS - F
MI_FLOOR = 1e-5          # MI (nats) below which the target saturates


def target(mi: np.ndarray) -> np.ndarray:
    """Regression target: log MI, floored."""
    return np.log(np.maximum(np.asarray(mi, np.float64), MI_FLOOR)).astype(np.float32)


def init_params(feat: np.ndarray, seed: int) -> np.ndarray:
    """Glorot-uniform weights, zero biases, input standardisation from feat."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS, np.float32)

    def glorot(n_out, n_in):
        a = np.sqrt(6.0 / (n_in + n_out))
        return rng.uniform(-a, a, size=(n_out, n_in)).astype(np.float32)

    o = 0
    p[o:o + H * F] = glorot(H, F).ravel(); o += H * F + H
    p[o:o + H * H] = glorot(H, H).ravel(); o += H * H + H
    p[o:o + H] = glorot(1, H).ravel(); o += H + 1
    feat = np.asarray(feat, np.float64)
    mu = feat.mean(0)
    sd = feat.std(0)
    sd[sd < 1e-6] = 1.0
    p[OFF_MU:OFF_SD] = mu
    p[OFF_SD:] = sd
    return p


class Screen:
    """MLP parameters resident on the device of ``ctx``."""

    def __init__(self, ctx: "hip.Context", params: np.ndarray):
        if hip.lib().centre_screen_n_params() != N_PARAMS:
            raise RuntimeError("parameter layout differs between the library and this module")
        self.ctx = ctx
        self.d_params = ctx.to_device(np.asarray(params, np.float32))

    def params(self) -> np.ndarray:
        return self.d_params.numpy()

    def train(self, d_feat: "hip.DevArray", d_target: "hip.DevArray", n: int, epochs: int = 4,
              batch: int = 4096, lr: float = 2e-3, seed: int = 1) -> np.ndarray:
        return self.ctx.mlp_train(self.d_params.ptr, d_feat.ptr, d_target.ptr, n, epochs, batch,
                                  lr, seed)

    def predict(self, d_feat: "hip.DevArray", n: int) -> "hip.DevArray":
        """Predicted log(MI) for the first n feature rows."""
        out = self.ctx.alloc(n, np.float32)
        self.ctx.mlp_forward(self.d_params.ptr, d_feat.ptr, n, out.ptr)
        return out


def features(ctx: "hip.Context", d_codes: "hip.DevArray", n_frame: int, d_types, d_ext,
             nbins: int, n_sub: int = 2000) -> "hip.DevArray":
    """[P, 8] features of all pairs of the device code matrix [D, n_frame]."""
    D = d_codes.shape[0]
    P = D * (D - 1) // 2
    d_feat = ctx.alloc((P, F), np.float32)
    n_sub = min(n_sub, n_frame) // 4 * 4
    ctx.screen_features(d_codes.ptr, d_codes.shape[1], D, n_frame, d_types.ptr, d_ext.ptr, nbins,
                        n_sub, d_feat.ptr)
    return d_feat


def keep_mask(pred_log_mi: np.ndarray, cutoff: float) -> np.ndarray:
    """Pairs whose predicted MI reaches the cutoff (nats); compared in the
    predictions' FP32."""
    return np.asarray(pred_log_mi, np.float32) >= np.float32(np.log(cutoff))
