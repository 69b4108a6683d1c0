"""Float64 CPU restatement of the surrogate network with torch autograd.

Used by the GPU tests to check the library's forward pass and gradients.
The depth transforms use torch.fft.rfft / irfft (not the library's DFT
matrices), the activation torch's tanh-form GELU and the complex mode
mixing a complex einsum, so it shares no arithmetic with the kernels.
"""
import torch
import torch.nn.functional as Fn


def bf16_round(t):
    return t.to(torch.bfloat16).to(torch.float64)


def reference_params(fno):
    """The parameters as the GPU path uses them: products see bf16 weights,
    bias epilogues see the fp32 biases."""
    out = {}
    for n, t in fno.params.items():
        t = t.detach().cpu().to(torch.float64)
        out[n] = t if n.startswith(("b", "c")) else bf16_round(t)
    return out


def forward(p, x, cfg):
    """x float64 [S, Z, F] -> prediction [S*Z]."""
    S, Z, F = x.shape
    h = x @ p["W0"] + p["b0"]
    for l in range(1, cfg.layers + 1):
        X = torch.fft.rfft(h, dim=1)[:, :cfg.modes, :]                 # [S, Km, C]
        W = torch.complex(p[f"Wr{l}"], p[f"Wi{l}"])                    # [Km, C, C]
        M = torch.einsum("skc,kcd->skd", X, W)
        full = torch.zeros(S, Z // 2 + 1, h.shape[2], dtype=M.dtype)
        full[:, :cfg.modes] = M
        Y = torch.fft.irfft(full, n=Z, dim=1)
        h = Fn.gelu(h @ p[f"Wp{l}"] + p[f"bp{l}"] + Y, approximate="tanh")
    q = Fn.gelu(h @ p["P1"] + p["c1"], approximate="tanh")
    return (q @ p["P2"] + p["c2"]).reshape(-1)


def loss_and_grads(p, x, target, mask, cfg):
    p = {n: t.clone().requires_grad_(True) for n, t in p.items()}
    pred = forward(p, x, cfg)
    loss = torch.sum(mask * (pred - target) ** 2) / torch.sum(mask)
    loss.backward()
    return float(loss), pred.detach(), {n: t.grad for n, t in p.items()}
