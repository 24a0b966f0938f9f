"""Pins the surrogate network's CPU oracle (oracle/surrogate_ref.py).

The network has no counterpart in the reference, so the oracle is checked against an
independent implementation of the same mathematics (torch autograd on the CPU, fp32) and
against central finite differences of its own loss in float64: a dropped term, a transposed
weight or a shifted tap fails one of them."""
import numpy as np
import pytest

from oracle import surrogate_ref as SR

TINY = SR.SgConfig(in_len=24, out_offset=4, out_len=16, ndiag=6, in_ch=5, enc_ch=8, enc_layers=3,
                   enc_k=5, head_ch=6, head_layers=3, enc_dilation=(1, 2, 3))


def random_problem(c, seed, dtype=np.float64, batch=2):
    rng = np.random.default_rng(seed)
    flat = (rng.standard_normal(SR.num_params(c)) * 0.3).astype(dtype)
    x = rng.standard_normal((batch, c.in_len, c.in_ch)).astype(dtype)
    y = rng.standard_normal((batch, c.ndiag, c.out_len)).astype(dtype)
    return flat, x, y


def torch_loss(c, flat, x, y):
    torch = pytest.importorskip("torch")
    F = torch.nn.functional
    p = {k: torch.tensor(v, requires_grad=True) for k, v in SR.unflatten(c, flat.copy()).items()}
    xt = torch.tensor(x)

    def c1(h, w, b, dil):  # [B, L, Cin], w [Cout, K, Cin]
        return F.conv1d(h.transpose(1, 2), w.permute(0, 2, 1), b, padding=(w.shape[1] // 2) * dil,
                        dilation=dil).transpose(1, 2)

    def c2(z, w, b):  # [B, D, I, Cin], w [Cout, 9, Cin]
        wk = w.reshape(w.shape[0], 3, 3, w.shape[2]).permute(0, 3, 1, 2)
        return F.conv2d(z.permute(0, 3, 1, 2), wk, b, padding=1).permute(0, 2, 3, 1)

    h = torch.relu(c1(xt, p["enc0.w"], p["enc0.b"], c.dil(0)))
    for l in range(1, c.enc_layers):
        h = h + torch.relu(c1(h, p[f"enc{l}.w"], p[f"enc{l}.b"], c.dil(l)))
    hc = h[:, c.out_offset:c.out_offset + c.out_len]
    P = hc @ p["proj_p.w"][:, 0, :].T + p["proj_p.b"]
    Q = hc @ p["proj_q.w"][:, 0, :].T
    L = c.out_len
    rows = []
    for d in range(c.ndiag):
        v = torch.relu(P[:, :L - d] + Q[:, d:] + p["emb"][d])
        rows.append(torch.cat([v, torch.zeros(x.shape[0], d, c.head_ch, dtype=v.dtype)], dim=1))
    z = torch.stack(rows, dim=1)
    for j in range(c.head_layers - 1):
        z = z + torch.relu(c2(z, p[f"head{j}.w"], p[f"head{j}.b"]))
    out = c2(z, p["out.w"], p["out.b"])[..., 0]
    m = torch.tensor(SR.valid_mask(c))[None]
    r = torch.where(m, out - torch.tensor(y), torch.zeros_like(out))
    loss = (r * r).sum() / (m.sum() * x.shape[0])
    loss.backward()
    g = np.concatenate([p[k].grad.numpy().reshape(-1) for k, _ in SR.param_shapes(c)])
    return float(loss.detach()), g, out.detach().numpy()


def test_param_count_of_default_network():
    c = SR.SgConfig()
    # 8 x 128-ch k=5 encoder, two 1x1 projections to 64, 100 x 64 embedding, 3 x (64-ch 3x3),
    # 1-channel 3x3 output
    expect = (128 * 5 * 6 + 128) + 7 * (128 * 5 * 128 + 128) + (64 * 128 + 64) + 64 * 128 + \
        100 * 64 + 3 * (64 * 9 * 64 + 64) + (9 * 64 + 1)
    assert SR.num_params(c) == expect


@pytest.mark.parametrize("seed", [0, 1])
def test_oracle_matches_torch_autograd(seed):
    flat, x, y = random_problem(TINY, seed, np.float64)
    loss, g, out = SR.loss_and_grads(TINY, SR.unflatten(TINY, flat), x, y)
    tl, tg, tout = torch_loss(TINY, flat, x, y)
    assert np.allclose(out, tout, rtol=1e-10, atol=1e-10)
    assert abs(loss - tl) < 1e-10 * max(1.0, abs(tl))
    assert np.allclose(g, tg, rtol=1e-8, atol=1e-10)


def test_oracle_gradient_matches_finite_differences():
    flat, x, y = random_problem(TINY, 7, np.float64)
    _, g, _ = SR.loss_and_grads(TINY, SR.unflatten(TINY, flat), x, y)
    rng = np.random.default_rng(3)
    # a few entries of every tensor
    offs, k = [], 0
    for _, sh in SR.param_shapes(TINY):
        n = int(np.prod(sh))
        offs += list(k + rng.choice(n, size=min(3, n), replace=False))
        k += n
    eps = 1e-6
    for j in offs:
        fp, fm = flat.copy(), flat.copy()
        fp[j] += eps
        fm[j] -= eps
        lp = SR.loss_and_grads(TINY, SR.unflatten(TINY, fp), x, y)[0]
        lm = SR.loss_and_grads(TINY, SR.unflatten(TINY, fm), x, y)[0]
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[j]) < 1e-6 + 1e-4 * abs(fd), (j, fd, g[j])


def test_adamw_first_step_is_sign_times_lr():
    """With zero moments, the first bias-corrected step moves each weight by lr * g / (|g| + eps)
    plus the decay term."""
    p = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -0.01, 0.0])
    q, m, v = SR.adamw(p, np.zeros(3), np.zeros(3), g, 1, lr=0.1, wd=0.01)
    expect = p - 0.1 * (g / (np.abs(g) + 1e-8) + 0.01 * p)
    assert np.allclose(q, expect, rtol=1e-12)
