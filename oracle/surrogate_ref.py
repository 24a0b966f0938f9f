"""CPU oracle of the surrogate network (test infrastructure only: imported by tests/, smoke()
and bench.py's baseline leg, never by the product path).

A plain numpy restatement, in any float dtype, of the network the library trains on the GPU
(modlefork_amd/csrc/sg_model.cpp; the network has no counterpart in the reference). Parameters
use the library's canonical order and shapes: weights [cout][taps][cin] then biases [cout], per
layer, as listed by param_shapes(). Pinned by tests/test_surrogate_oracle.py against torch
autograd on the CPU and against central finite differences in float64.

    x   [B][Lin][in_ch]
    h1 = relu(conv1d(x)),  h(l+1) = h(l) + relu(conv1d(h(l)))   (kernel K, layer l dilated by
                                                                 enc_dilation[l], zero padding)
    hc = h[:, out_offset : out_offset + Lout]
    P = hc Wp^T + bp,  Q = hc Wq^T
    z0[b, d, i] = relu(P[b, i] + Q[b, i + d] + E[d])  if i + d < Lout else 0
    z(j+1) = z(j) + relu(conv3x3(z(j)))                                 on the (d, i) grid
    out[b, d, i] = conv3x3(z_last) (1 channel)
    loss = mean over b and valid pixels (i + d < Lout) of (out - y)^2
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class SgConfig:
    in_len: int = 600
    out_offset: int = 100
    out_len: int = 400
    ndiag: int = 100
    in_ch: int = 6
    enc_ch: int = 128
    enc_layers: int = 8
    enc_k: int = 5
    head_ch: int = 64
    head_layers: int = 4
    enc_dilation: tuple = (1, 1, 2, 4, 8, 16, 32, 1)

    def dil(self, l: int) -> int:
        return self.enc_dilation[l] if l < len(self.enc_dilation) and self.enc_dilation[l] else 1


def param_shapes(c: SgConfig):
    C, H, K = c.enc_ch, c.head_ch, c.enc_k
    s = []
    for l in range(c.enc_layers):
        s.append((f"enc{l}.w", (C, K, c.in_ch if l == 0 else C)))
        s.append((f"enc{l}.b", (C,)))
    s += [("proj_p.w", (H, 1, C)), ("proj_p.b", (H,)), ("proj_q.w", (H, 1, C)),
          ("emb", (c.ndiag, H))]
    for j in range(c.head_layers - 1):
        s.append((f"head{j}.w", (H, 9, H)))
        s.append((f"head{j}.b", (H,)))
    s += [("out.w", (1, 9, H)), ("out.b", (1,))]
    return s


def num_params(c: SgConfig) -> int:
    return sum(int(np.prod(sh)) for _, sh in param_shapes(c))


def unflatten(c: SgConfig, flat: np.ndarray) -> dict:
    out, k = {}, 0
    for name, sh in param_shapes(c):
        n = int(np.prod(sh))
        out[name] = flat[k:k + n].reshape(sh)
        k += n
    assert k == len(flat)
    return out


def flatten(c: SgConfig, d: dict) -> np.ndarray:
    return np.concatenate([d[name].reshape(-1) for name, _ in param_shapes(c)])


def valid_mask(c: SgConfig) -> np.ndarray:
    d = np.arange(c.ndiag)[:, None]
    i = np.arange(c.out_len)[None, :]
    return (i + d) < c.out_len


# ---- convolutions (zero padding, centred taps; 1D taps dilated by dil) ----
def conv1d(x, w, b, dil=1):
    """x [B, L, Cin], w [Cout, K, Cin] -> [B, L, Cout]; tap t reads l + (t - K // 2) * dil"""
    K = w.shape[1]
    L = x.shape[1]
    pad = (K // 2) * dil
    xp = np.pad(x, ((0, 0), (pad, pad), (0, 0)))
    y = np.zeros(x.shape[:2] + (w.shape[0],), dtype=x.dtype)
    for t in range(K):
        y += xp[:, t * dil:t * dil + L, :] @ w[:, t, :].T
    return y + b


def conv1d_backward(x, w, g, dil=1):
    K = w.shape[1]
    L = x.shape[1]
    pad = (K // 2) * dil
    xp = np.pad(x, ((0, 0), (pad, pad), (0, 0)))
    dw = np.zeros_like(w)
    dxp = np.zeros_like(xp)
    for t in range(K):
        dw[:, t, :] = np.einsum("blo,bli->oi", g, xp[:, t * dil:t * dil + L, :])
        dxp[:, t * dil:t * dil + L, :] += g @ w[:, t, :]
    return dxp[:, pad:pad + L, :], dw, g.sum(axis=(0, 1))


def conv2d3(x, w, b):
    """x [B, D, I, Cin], w [Cout, 9, Cin] (taps 3 * td + ti, offset (td - 1, ti - 1))"""
    D, I = x.shape[1], x.shape[2]
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    y = np.zeros(x.shape[:3] + (w.shape[0],), dtype=x.dtype)
    for t in range(9):
        td, ti = divmod(t, 3)
        y += xp[:, td:td + D, ti:ti + I, :] @ w[:, t, :].T
    return y + b


def conv2d3_backward(x, w, g):
    D, I = x.shape[1], x.shape[2]
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    dw = np.zeros_like(w)
    dxp = np.zeros_like(xp)
    for t in range(9):
        td, ti = divmod(t, 3)
        dw[:, t, :] = np.einsum("bdio,bdic->oc", g, xp[:, td:td + D, ti:ti + I, :])
        dxp[:, td:td + D, ti:ti + I, :] += g @ w[:, t, :]
    return dxp[:, 1:1 + D, 1:1 + I, :], dw, g.sum(axis=(0, 1, 2))


def relu(v):
    return np.maximum(v, 0)


# ---- network ----
def forward(c: SgConfig, p: dict, x: np.ndarray):
    cache = {"x": x}
    h = relu(conv1d(x, p["enc0.w"], p["enc0.b"], c.dil(0)))
    cache["h1"] = h
    hs, rs = [None, h], [None]
    for l in range(1, c.enc_layers):
        r = relu(conv1d(h, p[f"enc{l}.w"], p[f"enc{l}.b"], c.dil(l)))
        h = h + r
        hs.append(h)
        rs.append(r)
    cache["hs"], cache["rs"] = hs, rs
    hc = h[:, c.out_offset:c.out_offset + c.out_len, :]
    cache["hc"] = hc
    P = hc @ p["proj_p.w"][:, 0, :].T + p["proj_p.b"]
    Q = hc @ p["proj_q.w"][:, 0, :].T
    B, L, H, ND = x.shape[0], c.out_len, c.head_ch, c.ndiag
    z = np.zeros((B, ND, L, H), dtype=x.dtype)
    for d in range(ND):
        z[:, d, :L - d, :] = relu(P[:, :L - d, :] + Q[:, d:, :] + p["emb"][d])
    cache["z0"] = z
    zs, r2s = [z], []
    for j in range(c.head_layers - 1):
        r = relu(conv2d3(z, p[f"head{j}.w"], p[f"head{j}.b"]))
        z = z + r
        zs.append(z)
        r2s.append(r)
    cache["zs"], cache["r2s"] = zs, r2s
    out = conv2d3(z, p["out.w"], p["out.b"])[..., 0]
    return out, cache


def loss_and_grads(c: SgConfig, p: dict, x: np.ndarray, y: np.ndarray):
    out, cache = forward(c, p, x)
    m = valid_mask(c)[None]
    nvalid = m.sum() * x.shape[0]
    r = np.where(m, out - y, 0)
    loss = float((r * r).sum() / nvalid)
    gout = (2.0 / nvalid) * r
    g = {}
    zs, r2s = cache["zs"], cache["r2s"]
    gz, g["out.w"], g["out.b"] = conv2d3_backward(zs[-1], p["out.w"], gout[..., None])
    for j in range(c.head_layers - 2, -1, -1):
        gm = gz * (r2s[j] > 0)
        dx, g[f"head{j}.w"], g[f"head{j}.b"] = conv2d3_backward(zs[j], p[f"head{j}.w"], gm)
        gz = gz + dx
    z0 = cache["z0"]
    gs = gz * (z0 > 0)
    B, L, H, ND = x.shape[0], c.out_len, c.head_ch, c.ndiag
    dP = np.zeros((B, L, H), dtype=x.dtype)
    dQ = np.zeros((B, L, H), dtype=x.dtype)
    g["emb"] = np.zeros((ND, H), dtype=x.dtype)
    for d in range(ND):
        dP[:, :L - d, :] += gs[:, d, :L - d, :]
        dQ[:, d:, :] += gs[:, d, :L - d, :]
        g["emb"][d] = gs[:, d, :L - d, :].sum(axis=(0, 1))
    hc = cache["hc"]
    g["proj_p.w"] = np.einsum("blo,bli->oi", dP, hc)[:, None, :]
    g["proj_p.b"] = dP.sum(axis=(0, 1))
    g["proj_q.w"] = np.einsum("blo,bli->oi", dQ, hc)[:, None, :]
    dhc = dP @ p["proj_p.w"][:, 0, :] + dQ @ p["proj_q.w"][:, 0, :]
    gh = np.zeros_like(cache["hs"][-1])
    gh[:, c.out_offset:c.out_offset + c.out_len, :] = dhc
    hs, rs = cache["hs"], cache["rs"]
    for l in range(c.enc_layers - 1, 0, -1):
        gm = gh * (rs[l] > 0)
        dx, g[f"enc{l}.w"], g[f"enc{l}.b"] = conv1d_backward(hs[l], p[f"enc{l}.w"], gm, c.dil(l))
        gh = gh + dx
    gm = gh * (cache["h1"] > 0)
    _, g["enc0.w"], g["enc0.b"] = conv1d_backward(x, p["enc0.w"], gm, c.dil(0))
    return loss, flatten(c, g), out


def adamw(p, m, v, g, step, lr, wd, b1=0.9, b2=0.999, eps=1e-8):
    """One AdamW step (decoupled weight decay, bias-corrected moments); returns new (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * (mh / (np.sqrt(vh) + eps) + wd * p), m, v
