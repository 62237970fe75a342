"""CPU restatement of the learned-guess network (test infrastructure only).

No reference counterpart: the reference starts each elliptic solve from the
previous solution (odeSolver.cpp:652-720).  This module restates, in float64
numpy, the network the engine trains and evaluates on the GPU
(qdeqcycle_fdm_amd/csrc/nn.hip, guess.hip), so that the GPU forward and
backward passes can be checked against it:

  x [B, nk]  -> Conv1d(1 -> C, width T, stride S, zero padding; output
                position p reads nodes S*p - (T-S)/2 + t) + ReLU -> H0 [B, P0*C]
                (feature p*C + c)
             -> W1 [H, F0], ReLU -> W2 [H, H], ReLU -> W3 [H, H], ReLU
             -> W4 [nout, H] -> y [B, nout]      (no biases)
  loss = mean over B*nout of (y - t)^2

With ``bf16=True`` the values are rounded to bfloat16 (round to nearest
even) at the points where the GPU stores them in BF16: the four weight
matrices, H0..H3, dloss/dy and the masked hidden gradients dZ1..dZ3; the
encoder and dZ0 stay unrounded (FP32 on the GPU).  With ``bf16=False`` the
pass is plain float64 (used for the finite-difference check of the
gradient formulas).
"""
from dataclasses import dataclass

import numpy as np


def bf16_round(a) -> np.ndarray:
    """Rounds to the nearest bfloat16 (ties to even) via float32 bits; returns
    float64.  NaN and infinities pass through."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    rounded = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    out = rounded.astype(np.uint32).view(np.float32)
    out = np.where(np.isfinite(f), out, f)
    return out.astype(np.float64)


@dataclass
class Dims:
    nk: int
    factor: int = 4
    channels: int = 8
    width: int = 8
    hidden: int = 512

    @property
    def nc(self):
        return (self.nk - 1) // self.factor + 1

    @property
    def nout(self):
        return self.nc * self.nc

    @property
    def P0(self):
        return (self.nk - 1) // self.factor

    @property
    def F0(self):
        return self.P0 * self.channels

    def shapes(self):
        return [(self.channels, self.width), (self.hidden, self.F0),
                (self.hidden, self.hidden), (self.hidden, self.hidden),
                (self.nout, self.hidden)]

    def num_params(self):
        return sum(r * c for r, c in self.shapes())


def unpack(dims: Dims, flat):
    out, o = [], 0
    for r, c in dims.shapes():
        out.append(np.asarray(flat[o:o + r * c], dtype=np.float64).reshape(r, c))
        o += r * c
    return out


def pack(tensors) -> np.ndarray:
    return np.concatenate([np.asarray(t, dtype=np.float64).ravel() for t in tensors])


def conv_index(dims: Dims):
    """(P0, T) node index of each encoder tap (-1 = zero padding)."""
    S, T = dims.factor, dims.width
    p = np.arange(dims.P0)[:, None]
    t = np.arange(T)[None, :]
    q = S * p - (T - S) // 2 + t
    return np.where((q >= 0) & (q < dims.nk), q, -1)


def _conv(dims, x, Wc):
    q = conv_index(dims)
    xp = np.concatenate([x, np.zeros((x.shape[0], 1))], axis=1)  # index -1 -> 0
    patches = xp[:, q]                        # [B, P0, T]
    z = np.einsum("bpt,ct->bpc", patches, Wc)  # [B, P0, C]
    return z.reshape(x.shape[0], -1), patches


def forward(dims: Dims, flat, x, bf16=True, cache=False):
    r = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    Wc, W1, W2, W3, W4 = unpack(dims, flat)
    Wb = [r(W) for W in (W1, W2, W3, W4)]
    x = np.asarray(x, dtype=np.float64)
    z0, patches = _conv(dims, x, Wc)
    H = [r(np.maximum(z0, 0.0))]
    for l in range(3):
        H.append(r(np.maximum(H[-1] @ Wb[l].T, 0.0)))
    y = H[3] @ Wb[3].T
    if cache:
        return y, dict(H=H, Wb=Wb, patches=patches)
    return y


def forward_backward(dims: Dims, flat, x, t, bf16=True):
    """Returns (loss, flat gradient)."""
    r = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    B = x.shape[0]
    y, c = forward(dims, flat, x, bf16=bf16, cache=True)
    H, Wb, patches = c["H"], c["Wb"], c["patches"]
    d = y - np.asarray(t, dtype=np.float64)
    loss = float(np.mean(d * d))
    dY = r(2.0 * d / (B * dims.nout))
    g4 = dY.T @ H[3]
    dH = dY @ Wb[3]
    grads = [None, None, None, None, g4]
    for l in (2, 1, 0):
        dZ = r(dH * (H[l + 1] > 0))
        grads[l + 1] = dZ.T @ H[l]
        dH = dZ @ Wb[l]
    dZ0 = dH * (H[0] > 0)                               # [B, P0*C]
    dZ0 = dZ0.reshape(B, dims.P0, dims.channels)
    grads[0] = np.einsum("bpc,bpt->ct", dZ0, patches)
    return loss, pack(grads)


def init_params(dims: Dims, seed=0) -> np.ndarray:
    """He-uniform weights (an independent draw; the engine seeds its own)."""
    rng = np.random.default_rng(seed)
    out = []
    for i, (rws, cols) in enumerate(dims.shapes()):
        bound = np.sqrt((3.0 if i == 4 else 6.0) / cols)
        out.append(rng.uniform(-bound, bound, size=(rws, cols)))
    return pack(out).astype(np.float32)
