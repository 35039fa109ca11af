"""CPU oracle of the surrogate MLP ensemble (TEST INFRASTRUCTURE ONLY).

The network is the project's own (the reference has no learned part):
din -> H -> ... -> H -> 1 with exact GELU, x -> 0.5 x (1 + erf(x / sqrt 2)),
on the mean squared error.  This module restates its forward pass and its
gradients in float64 numpy in two modes:

  * exact:  no rounding; its gradients are checked against central finite
            differences of its own forward pass;
  * device: the rounding points of the device kernels — inputs, weight copies,
            hidden activations and back-propagated pre-activation gradients
            rounded to bfloat16 (round to nearest even), pre-activations and
            outputs rounded to float32 — so that device gradients can be
            compared with it closely.

Only the tests and the bench's checks import it.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf


def bf16(a: np.ndarray) -> np.ndarray:
    """Round float values to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


class MLPOracle:
    def __init__(self, din: int, hidden: int, layers: int):
        self.din, self.H, self.L = din, hidden, layers

    def count(self) -> int:
        H = self.H
        c = H * self.din + H
        for _ in range(1, self.L):
            c += H * H + H
        return c + H + 1

    def unpack(self, p: np.ndarray):
        p = np.asarray(p, dtype=np.float64)
        H, q = self.H, 0
        Ws, bs = [], []
        for l in range(self.L):
            kin = self.din if l == 0 else H
            Ws.append(p[q:q + H * kin].reshape(H, kin))
            q += H * kin
            bs.append(p[q:q + H])
            q += H
        wo = p[q:q + H]
        q += H
        bo = p[q]
        return Ws, bs, wo, bo

    def pack(self, Ws, bs, wo, bo) -> np.ndarray:
        parts = []
        for W, b in zip(Ws, bs):
            parts += [np.ravel(W), np.ravel(b)]
        parts += [np.ravel(wo), np.array([bo])]
        return np.concatenate(parts)

    def forward(self, p, x, device=True):
        Ws, bs, wo, bo = self.unpack(p)
        rb = bf16 if device else (lambda a: np.asarray(a, dtype=np.float64))
        r32 = f32 if device else (lambda a: np.asarray(a, dtype=np.float64))
        a = rb(x)
        Zs, As = [], [a]
        for W, b in zip(Ws, bs):
            z = r32(a @ rb(W).T + b)
            a = rb(gelu(z))
            Zs.append(z)
            As.append(a)
        y = r32(a @ wo + bo)
        return y, (Zs, As)

    def gradients(self, p, x, t, device=True):
        """(flat gradient of mean((y - t)^2), loss) for one batch."""
        Ws, bs, wo, bo = self.unpack(p)
        rb = bf16 if device else (lambda a: np.asarray(a, dtype=np.float64))
        y, (Zs, As) = self.forward(p, x, device)
        B = len(t)
        r = y - np.asarray(t, dtype=np.float64)
        loss = float(np.sum(r * r) / B)
        dy = 2.0 * r / B
        gwo = As[-1].T @ dy
        gbo = float(np.sum(dy))
        dz = rb(dy[:, None] * wo[None, :] * gelu_grad(Zs[-1]))
        gW = [None] * self.L
        gb = [None] * self.L
        for l in range(self.L - 1, -1, -1):
            gW[l] = dz.T @ As[l]
            gb[l] = dz.sum(axis=0)
            if l > 0:
                dz = rb((dz @ rb(Ws[l])) * gelu_grad(Zs[l - 1]))
        return self.pack(gW, gb, gwo, gbo), loss
