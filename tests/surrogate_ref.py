"""A plain PyTorch fp32 restatement of the surrogate network (run on the CPU)
for the GPU kernels' numerics tests.  Weights are rounded to BF16 and
activations are rounded to BF16 where the device stores them; sums are
FP32.  The weight vector has the documented order of
mms_surrogate_weight_count (include/mms_gpu.h)."""
import math

import numpy as np
import torch

C = 64


def _bf(t):
    return t.to(torch.bfloat16).to(torch.float32)


def gelu(x):
    return 0.5 * x * (1.0 + torch.tanh(0.7978845608 * (x + 0.044715 * x * x * x)))


def unpack(w):
    w = torch.from_numpy(np.asarray(w, np.float32))
    o = 0

    def take(n, shape):
        nonlocal o
        t = w[o:o + n].reshape(shape)
        o += n
        return t
    p = {}
    p["lift"] = take(C * 27, (C, 3, 3, 3))
    p["blift"] = take(C, (C,))
    for k in ("1", "2"):
        p["w" + k] = take(C * C * 9, (C, C, 3, 3))
        p["b" + k] = take(C, (C,))
    p["mode"] = take(32 * 32 * C * C, (32 * 32, C, C))
    p["w3"] = take(C * C * 9, (C, C, 3, 3))
    p["b3"] = take(C, (C,))
    p["wp"] = take(C, (C,))
    p["bp"] = take(1, (1,))
    assert o == w.numel()
    return p


def basis(n):
    x = np.arange(n, dtype=np.float64)
    rows = []
    for k in range(32):
        f = (k + 1) // 2
        a = 2 * math.pi * f * x / n
        rows.append(np.ones(n) if k == 0 else (np.cos(a) if k % 2 else np.sin(a)))
    return _bf(torch.from_numpy(np.sqrt(2.0 / n) * np.stack(rows)).to(torch.float32))


def planes(rho, zf, rscale):
    """Input planes of every layer of films of zf layers: [Z, 3, y, x] (own
    charge, layers at distance 1, layers at distance >= 2), times rscale."""
    Z = rho.shape[0]
    out = torch.zeros((Z, 3) + tuple(rho.shape[1:]), dtype=torch.float32)
    for z in range(Z):
        f0 = (z // zf) * zf
        for k in range(zf):
            d = abs(k - (z - f0))
            out[z, min(d, 2)] += rho[f0 + k]
    return out * np.float32(rscale)


def forward_t(p, rho, zf, rscale, oscale):
    """Differentiable forward pass: p the unpacked parameters (tensors), rho
    [Z, y, x] float32 -> potential [Z, y, x]."""
    Z, yd, xd = rho.shape
    x = _bf(planes(rho, zf, rscale))
    conv = torch.nn.functional.conv2d
    h0 = _bf(gelu(conv(x, _bf(p["lift"]), p["blift"], padding=1)))
    h1 = _bf(gelu(conv(h0, _bf(p["w1"]), p["b1"], padding=1)))
    Bx, By = basis(xd), basis(yd)                                         # [32, n]
    hl = h1.permute(0, 2, 3, 1)                                           # [z, y, x, c]
    X1 = torch.einsum("kx,zyxc->zykc", Bx, hl)                             # FP32
    H = torch.einsum("ly,zykc->zlkc", By, _bf(X1))                         # [z, ky, kx, c]
    Wm = _bf(p["mode"]).reshape(32, 32, C, C)                              # [ky, kx, co, ci]
    G = torch.einsum("lkoi,zlki->zlko", Wm, H)
    X2 = torch.einsum("ly,zlko->zyko", By, _bf(G))
    S = torch.einsum("kx,zyko->zyxo", Bx, _bf(X2)).permute(0, 3, 1, 2)  # [z, c, y, x], FP32 (added in the conv's sums)
    h2 = _bf(gelu(conv(h1, _bf(p["w2"]), p["b2"], padding=1) + S))
    h3 = _bf(gelu(conv(h2, _bf(p["w3"]), p["b3"], padding=1)))
    return (torch.einsum("zcyx,c->zyx", h3, p["wp"]) + p["bp"]) * np.float32(oscale)


def forward(weights, rho, mat, dims, rscale, oscale):
    """rho [zd*yd*xd] float32 of one film -> potential [zd*yd*xd] (mat is
    not an input of the network)."""
    xd, yd, zd = dims
    p = unpack(weights)
    r = torch.from_numpy(np.asarray(rho, np.float32).reshape(zd, yd, xd))
    with torch.no_grad():
        out = forward_t(p, r, zd, rscale, oscale)
    return out.reshape(-1).numpy()


BLOCKS = ("lift", "blift", "w1", "b1", "w2", "b2", "mode", "w3", "b3", "wp", "bp")


def loss_and_grad(weights, rho, phi, dims, nfilm, rscale, oscale):
    """Loss mean_f ||phi_s - phi||^2 / ||phi||^2 over nfilm films of dims
    (rho, phi [nfilm * n]) and its gradient in the API weight order."""
    xd, yd, zd = dims
    p = {k: v.clone().requires_grad_(True) for k, v in unpack(weights).items()}
    r = torch.from_numpy(np.asarray(rho, np.float32).reshape(nfilm * zd, yd, xd))
    t = torch.from_numpy(np.asarray(phi, np.float32).reshape(nfilm, -1))
    out = forward_t(p, r, zd, rscale, oscale).reshape(nfilm, -1)
    inv = torch.from_numpy(np.array([1.0 / np.sum(np.asarray(phi, np.float64).reshape(nfilm, -1)[f] ** 2)
                                     for f in range(nfilm)], np.float32))
    loss = (((out - t) ** 2).sum(1) * inv).mean()
    loss.backward()
    g = np.concatenate([p[k].grad.reshape(-1).numpy() for k in BLOCKS]).astype(np.float32)
    return float(loss.detach()), g


def block_slices():
    """(name, slice) of each block of the API weight vector."""
    sizes = {"lift": C * 27, "blift": C, "w1": C * C * 9, "b1": C, "w2": C * C * 9, "b2": C,
             "mode": 32 * 32 * C * C, "w3": C * C * 9, "b3": C, "wp": C, "bp": 1}
    o, out = 0, []
    for k in BLOCKS:
        out.append((k, slice(o, o + sizes[k])))
        o += sizes[k]
    return out
