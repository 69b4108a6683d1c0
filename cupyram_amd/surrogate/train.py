"""Training of the surrogate on PE-generated TL fields, and its validation.

    python -m cupyram_amd.surrogate.train --domain wedge --n-train 50000 \
        --members 4 --epochs 8 --out cupyram_amd/surrogate/weights

generates the training fields with the complex64 PE march (this library),
trains each ensemble member with the library's forward/backward/Adam kernels
and reports, on held-out PE solves, the TL error of the network beside two
baselines that use no network: the nearest training field in normalised
query space, and the mean training field.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import torch

from ..cupyram import CuPyRAM
from . import dataset as D
from .encode import ENCODING, Domain, Queries, denormalise_tl, encode, global_features, normalise_tl
from .fno import FNO, FNOConfig, infer

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "weights")


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


def pe_fields(dom: Domain, q: Queries, precision="c64", chunk=10000, log=_log):
    """TL (dB) of every query on the domain grid (cell means or point samples
    per the domain's target), float32 [n, nr, nzp] on the GPU, and the march
    time (s) and cell-steps."""
    n = len(q)
    out = torch.empty(n, dom.nr, dom.nzp, dtype=torch.float32, device="cuda")
    t_march, cells = 0.0, 0.0
    starts = list(range(0, n, chunk))

    def prepare(a):
        # host set-up of one chunk (keyword arguments, grids, packed tables,
        # Pade sets); ctypes releases the GIL while the march runs, so the next
        # chunk is prepared on a second thread meanwhile
        idx = np.arange(a, min(n, a + chunk))
        m = CuPyRAM(**D.march_kwargs(dom, q.take(idx), precision=precision))
        m.setup()
        return idx, m

    with ThreadPoolExecutor(max_workers=1) as pool:
        nxt = pool.submit(prepare, starts[0]) if starts else None
        for i, a in enumerate(starts):
            idx, m = nxt.result()
            if i + 1 < len(starts):
                nxt = pool.submit(prepare, starts[i + 1])
            t0 = time.perf_counter()
            res = m.run()
            t_march += time.perf_counter() - t0
            r = m._rays
            cells += float(np.sum(r["nz"].astype(np.float64) * r["nsteps"]))
            key = "TL Cells" if dom.target == "cells" else "TL Grid"
            tlg = np.asarray(res[key]).reshape(len(idx), dom.nzp, dom.nr)
            # [depth][range] per field to the network's [range][depth], converted on the GPU
            out[a:a + len(idx)] = torch.from_numpy(tlg).cuda().transpose(1, 2).float()
            if log:
                log(f"  PE {precision} {dom.name}: {a + len(idx)}/{n} fields ({t_march:.1f} s)")
    return out, t_march, cells


def _lr(step, total, lr0, warm=100):
    if step < warm:
        return lr0 * (step + 1) / warm
    return 1e-2 * lr0 + 0.5 * (lr0 - 1e-2 * lr0) * (1 + math.cos(math.pi * (step - warm) / max(1, total - warm)))


def train_member(dom, q, tl, cfg=FNOConfig(), epochs=4, batch=64, lr=1e-3, seed=0, log=_log, init=None):
    """Train one network on queries q with TL fields tl ([n, nr, nzp] GPU),
    from random weights or from the member file `init` (further training)."""
    if init is not None:
        fno, dom_init = load_member(init)
        if dom_init.name != dom.name:
            raise ValueError(f"{init} was trained on domain {dom_init.name}, not {dom.name}")
    else:
        fno = FNO(cfg, seed=seed)
    target, mask = normalise_tl(dom, tl)
    n = len(q)
    steps_per_epoch = n // batch
    total = epochs * steps_per_epoch
    rng = np.random.default_rng(seed + 1000)
    step, t0 = 0, time.perf_counter()
    hist = []
    for ep in range(epochs):
        perm = rng.permutation(n)
        acc, cnt = 0.0, 0.0
        loss_dev = torch.zeros(2, device="cuda")
        for b in range(steps_per_epoch):
            idx = np.sort(perm[b * batch:(b + 1) * batch])
            x = encode(dom, q.take(idx), cfg.features)
            ti = torch.as_tensor(idx, device="cuda")
            pred = fno.forward(x, train=True)
            loss, dpred = fno.loss_grad(pred, target[ti].reshape(-1), mask[ti].reshape(-1))
            loss_dev += loss
            fno.backward(dpred)
            fno.step(_lr(step, total, lr), wd=1e-4)
            step += 1
            if log and step % 500 == 0:
                log(f"  member {seed} step {step}/{total} ({time.perf_counter() - t0:.1f} s)")
        torch.cuda.synchronize()
        ls = loss_dev.cpu().numpy()
        rms_db = math.sqrt(ls[0] / max(ls[1], 1)) * dom.tl_std
        hist.append(rms_db)
        if log:
            log(f"  member {seed} epoch {ep + 1}/{epochs}: train RMS {rms_db:.2f} dB "
                f"({step} steps, {time.perf_counter() - t0:.1f} s)")
    return fno, {"train_rms_db": hist, "steps": step, "seconds": time.perf_counter() - t0}


def predict(dom, fno, q, batch=128, cfg=FNOConfig()):
    """Surrogate TL (dB) [n, nr, nzp] on the GPU."""
    n = len(q)
    out = torch.empty(n, dom.nr, dom.nzp, dtype=torch.float32, device="cuda")
    for a in range(0, n, batch):
        idx = np.arange(a, min(n, a + batch))
        y = infer(fno, encode(dom, q.take(idx), cfg.features))
        out[a:a + len(idx)] = denormalise_tl(dom, y.view(len(idx), dom.nr, dom.nzp))
    return out


def clipped(dom, tl):
    return torch.nan_to_num(tl, nan=dom.tl_hi).clamp(dom.tl_lo, dom.tl_hi)


def field_rms(dom, a, b):
    """Per-field RMS difference (dB) of two TL sets, both clipped to the
    domain's TL range."""
    d = clipped(dom, a) - clipped(dom, b)
    return torch.sqrt(torch.mean(d * d, dim=(1, 2)))


def query_matrix(dom, q):
    """Normalised query vectors (the global part of the encoding)."""
    return global_features(dom, q)


def nearest_neighbour(dom, q_train, tl_train, q_val, chunk=1024):
    a = query_matrix(dom, q_train)
    out = torch.empty(len(q_val), dom.nr, dom.nzp, device="cuda")
    b = query_matrix(dom, q_val)
    for s in range(0, len(q_val), chunk):
        d = torch.cdist(b[s:s + chunk], a)
        out[s:s + chunk] = tl_train[d.argmin(dim=1)]
    return out


def summary(x):
    x = x.detach().cpu().numpy()
    return {"rms_db": float(np.sqrt(np.mean(x ** 2))), "median_field_rms_db": float(np.median(x)),
            "p90_field_rms_db": float(np.percentile(x, 90)), "frac_fields_under_3db": float(np.mean(x < 3.0))}


def by_band(dom, q, err, nbands=6):
    """Per-field RMS error summarised over log-spaced frequency bands."""
    e = err.detach().cpu().numpy()
    edges = np.exp(np.linspace(np.log(dom.freq_lo), np.log(dom.freq_hi), nbands + 1))
    out = []
    for lo, hi in zip(edges[:-1], edges[1:]):
        m = (q.freq >= lo) & (q.freq <= hi)
        if m.any():
            out.append({"freq_lo": float(lo), "freq_hi": float(hi), "fields": int(m.sum()),
                        "rms_db": float(np.sqrt(np.mean(e[m] ** 2))), "median_db": float(np.median(e[m])),
                        "frac_under_3db": float(np.mean(e[m] < 3.0))})
    return out


def save_member(fno, dom, path):
    from safetensors.torch import save_file
    d = {}
    for n, t in fno.params.items():
        t = t.detach().cpu()
        d[n] = t.contiguous() if n.startswith(("b", "c")) else t.to(torch.bfloat16).contiguous()
    meta = {"domain": dom.name, "config": json.dumps(list(fno.cfg.__dict__.values())),
            "encoding": ENCODING, "target": dom.target,
            "format": "FNO params; product weights bf16 (as used), biases f32"}
    save_file(d, path, metadata=meta)


def load_member(path, device="cuda"):
    from safetensors import safe_open
    from safetensors.torch import load_file
    with safe_open(path, "pt") as f:
        meta = f.metadata()
    if meta.get("encoding") != ENCODING:
        raise ValueError(f"{path}: weights trained on input encoding {meta.get('encoding')!r}, "
                         f"this library encodes {ENCODING!r}; retrain")
    cfg = FNOConfig(*json.loads(meta["config"]))
    fno = FNO(cfg, seed=0, device=device)
    fno.load_state_dict({n: t.float() for n, t in load_file(path).items()})
    return fno, D.DOMAINS[meta["domain"]]


def load_ensemble(domain, directory=WEIGHTS):
    files = sorted(f for f in os.listdir(directory) if f.startswith(domain + "_m") and f.endswith(".safetensors"))
    if not files:
        raise FileNotFoundError(f"no trained {domain} members under {directory}")
    return [load_member(os.path.join(directory, f))[0] for f in files]


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--domain", choices=sorted(D.DOMAINS), default="wedge")
    ap.add_argument("--n-train", type=int, default=50000)
    ap.add_argument("--n-val", type=int, default=2000)
    ap.add_argument("--members", type=int, default=4)
    ap.add_argument("--resume", default=None,
                    help="directory of trained members to train further (same member files as --out writes)")
    ap.add_argument("--first-member", type=int, default=0,
                    help="index of the first member trained here (members of one ensemble may be trained "
                         "in separate runs on the same data)")
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=100)
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--report", default=None)
    ap.add_argument("--eval-only", action="store_true",
                    help="evaluate the trained members found under --out on fresh validation fields")
    args = ap.parse_args(argv)
    if args.eval_only:
        return evaluate(args)
    dom = D.DOMAINS[args.domain]
    gen = D.deep_queries if dom.name == "deep_munk" else D.wedge_queries
    t0 = time.perf_counter()
    q_tr = gen(args.n_train, args.seed)
    q_va = gen(args.n_val, args.seed + 1)
    tl_tr, t_tr, c_tr = pe_fields(dom, q_tr)
    tl_va, t_va, c_va = pe_fields(dom, q_va)
    _log(f"PE data: {args.n_train} + {args.n_val} fields, {c_tr + c_va:.3g} cell-steps, "
         f"{t_tr + t_va:.1f} s in the march")
    rep = {"domain": dom.name, "n_train": args.n_train, "n_val": args.n_val, "pe_seconds": t_tr + t_va,
           "pe_cell_steps": c_tr + c_va, "members": []}
    os.makedirs(args.out, exist_ok=True)
    preds = []
    for m in range(args.first_member, args.first_member + args.members):
        init = os.path.join(args.resume, f"{dom.name}_m{m}.safetensors") if args.resume else None
        fno, hist = train_member(dom, q_tr, tl_tr, epochs=args.epochs, batch=args.batch, lr=args.lr,
                                 seed=args.seed + 10 + m, init=init)
        save_member(fno, dom, os.path.join(args.out, f"{dom.name}_m{m}.safetensors"))
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        p = predict(dom, fno, q_va)
        torch.cuda.synchronize()
        t_inf = time.perf_counter() - t1
        preds.append(p)
        err = summary(field_rms(dom, p, tl_va))
        _log(f"member {m}: validation {err}, inference {args.n_val / t_inf:.0f} fields/s")
        rep["members"].append({**hist, "val": err, "inference_fields_per_s": args.n_val / t_inf})
        fno.release()
        del fno
    ens = torch.stack(preds).mean(0)
    spread = torch.sqrt(torch.stack(preds).var(0, unbiased=False).mean(dim=(1, 2)))
    ens_err = field_rms(dom, ens, tl_va)
    rep["ensemble_val"] = summary(ens_err)
    rep["ensemble_val_by_band"] = by_band(dom, q_va, ens_err)
    rep["ensemble_spread_median_db"] = float(spread.median())
    rep["nearest_neighbour_val"] = summary(field_rms(dom, nearest_neighbour(dom, q_tr, tl_tr, q_va), tl_va))
    mean_field = tl_tr.mean(0, keepdim=True).expand_as(tl_va)
    rep["mean_field_val"] = summary(field_rms(dom, mean_field, tl_va))
    rep["seconds"] = time.perf_counter() - t0
    _log(json.dumps(rep, indent=1))
    if args.report:
        with open(args.report, "w") as fh:
            json.dump(rep, fh, indent=1)


def evaluate(args):
    dom = D.DOMAINS[args.domain]
    gen = D.deep_queries if dom.name == "deep_munk" else D.wedge_queries
    members = load_ensemble(dom.name, args.out)
    q_va = gen(args.n_val, args.seed + 1)
    tl_va, t_va, c_va = pe_fields(dom, q_va)
    preds = [predict(dom, m, q_va) for m in members]
    ens = torch.stack(preds).mean(0)
    err = field_rms(dom, ens, tl_va)
    spread = torch.sqrt(torch.stack(preds).var(0, unbiased=False).mean(dim=(1, 2)))
    rep = {"domain": dom.name, "n_val": args.n_val, "members": len(members), "ensemble_val": summary(err),
           "ensemble_val_by_band": by_band(dom, q_va, err),
           "spread_vs_error_corr": float(np.corrcoef(spread.cpu().numpy(), err.cpu().numpy())[0, 1])}
    _log(json.dumps(rep, indent=1))
    if args.report:
        with open(args.report, "w") as fh:
            json.dump(rep, fh, indent=1)
    return 0


if __name__ == "__main__":
    main()
