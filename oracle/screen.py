"""CPU restatement of the learned pair screen: features and the MLP.

TEST INFRASTRUCTURE ONLY (see the package docstring).  The screen has no
reference counterpart; this is the build's own model written plainly in
float64 numpy, pinned by its own tests: closed-form feature values
(correlation of a DOF with itself or its mirror, MI of a DOF with itself
equals its entropy) and a central finite-difference check of the analytic
gradient.

Features of pair (i < j), as the device computes them:
  S1_i, S1_j (ML over all frames, one step, Jacobian on), circular and linear
  correlation of the codes on an evenly strided subsample, ML MI on the
  subsample with the bins merged to nbc = min(nb, 8) (code * nbc // nb),
  type_i / 3, type_j / 3, same-type flag.
"""
from __future__ import annotations

import numpy as np

from . import oracle as _o

F, H = 8, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + H * F
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H
OFF_MU = OFF_B3 + 1
OFF_SD = OFF_MU + F
N_PARAMS = OFF_SD + F


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[OFF_W1:OFF_B1].reshape(H, F), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(H, H),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3], p[OFF_B3], p[OFF_MU:OFF_SD], p[OFF_SD:N_PARAMS])


def forward(p, feat):
    """Predicted log(MI) per sample (float64)."""
    W1, b1, W2, b2, w3, b3, mu, sd = unpack(p)
    x = (np.asarray(feat, np.float64) - mu) / sd
    h1 = np.maximum(x @ W1.T + b1, 0.0)
    h2 = np.maximum(h1 @ W2.T + b2, 0.0)
    return h2 @ w3 + b3


def loss_grad(p, feat, target):
    """(sum 0.5 (y - t)^2, gradient w.r.t. the parameter vector); the mu and
    sd entries get zero gradient, as the device leaves them untouched."""
    W1, b1, W2, b2, w3, b3, mu, sd = unpack(p)
    x = (np.asarray(feat, np.float64) - mu) / sd
    z1 = x @ W1.T + b1
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0.0)
    y = h2 @ w3 + b3
    dy = y - np.asarray(target, np.float64)
    g = np.zeros(N_PARAMS)
    g[OFF_W3:OFF_B3] = dy @ h2
    g[OFF_B3] = dy.sum()
    d2 = np.outer(dy, w3) * (z2 > 0)
    g[OFF_W2:OFF_B2] = (d2.T @ h1).ravel()
    g[OFF_B2:OFF_W3] = d2.sum(0)
    d1 = (d2 @ W2) * (z1 > 0)
    g[OFF_W1:OFF_B1] = (d1.T @ x).ravel()
    g[OFF_B1:OFF_W2] = d1.sum(0)
    return 0.5 * float(dy @ dy), g


def features(codes, types, ext, nb, n_sub):
    """[P, 8] features for all pairs i < j in triangle order."""
    codes = np.asarray(codes, np.uint8)
    D, Fr = codes.shape
    s1 = _o.dof_entropies(codes, types, ext, 1, schemes=(nb,))[:, 0, 0, 0]
    step = Fr // n_sub
    sub = np.ascontiguousarray(codes[:, ::step][:, :n_sub])
    iu, ju = np.triu_indices(D, 1)
    pairs = np.stack([iu, ju], 1).astype(np.uint32)
    nbc = min(nb, 8)
    merged = (sub.astype(np.int64) * nbc // nb).astype(np.uint8)
    unit = np.tile([0.0, float(nbc), 0.0, 0.0], (D, 1))
    s2 = _o.pair_entropies(merged, types, unit, pairs, 1, schemes=(nbc,),
                           jacobian=False)[:, 0, 0, 0]
    s1s = _o.dof_entropies(merged, types, unit, 1, schemes=(nbc,), jacobian=False)[:, 0, 0, 0]
    mi = s1s[iu] + s1s[ju] - s2
    c = sub.astype(np.float64)
    xc = c - c.mean(1, keepdims=True)
    sdl = np.sqrt((xc * xc).sum(1))
    X = np.divide(xc, sdl[:, None], out=np.zeros_like(xc), where=sdl[:, None] > 0)
    th = (c + 0.5) * (2 * np.pi / nb)
    mu = np.arctan2(np.sin(th).sum(1), np.cos(th).sum(1))
    ys = np.sin(th - mu[:, None])
    sdc = np.sqrt((ys * ys).sum(1))
    Y = np.divide(ys, sdc[:, None], out=np.zeros_like(ys), where=sdc[:, None] > 0)
    lin = (X @ X.T)[iu, ju]
    circ = (Y @ Y.T)[iu, ju]
    t = np.asarray(types, np.float64)
    return np.stack([s1[iu], s1[ju], circ, lin, mi, t[iu] / 3, t[ju] / 3,
                     (t[iu] == t[ju]).astype(np.float64)], 1)
