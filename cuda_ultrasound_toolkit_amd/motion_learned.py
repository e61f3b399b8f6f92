"""Learned 3-D motion estimation: data, training and evaluation of the
displacement-proposal network.

Data: envelope-volume pairs beamformed by the plain pipeline from row-column
acquisitions (128 tx x 128 ch x 4096 samples -> 128^3) before and after a
rigid scatterer shift U[-6, 6] voxels per axis, with 2 % amplitude
decorrelation and independent noise; every 16^3 block at stride 8 is labelled
by the plain full +-8 ZNCC search.

Training runs on the GPU through the library's own kernels
(cut_pnet_gather / cut_pnet_loss_grad / cut_pnet_adam); blocks whose
full-search peak is below the fallback threshold are left out, since the
learned search hands them to the full search anyway.  Augmentation mirrors
axes (negating the label component) and offsets the source block by up to 2
voxels (label minus offset).

Usage (GPU):
    python -m cuda_ultrasound_toolkit_amd.motion_learned --pairs 160 --steps 24000 --out weights.safetensors
"""
from __future__ import annotations

import argparse
import json
import math
import os
import time
from dataclasses import dataclass, field

import numpy as np

from . import api, motion_data

WEIGHTS = os.path.join(os.path.dirname(__file__), "weights", "pnet_rca128.safetensors")
INPUT_MODE = 1  # the network reads the 2x2x2-pooled 32^3 search regions (cut_pnet_set_input_mode)
LAYERS = [(2, 16), (16, 32), (32, 32), (32, 32)]
GRID_STRIDE = 8


def init_params(seed: int) -> np.ndarray:
    """He-normal conv weights, zero biases, small FC."""
    rng = np.random.default_rng(seed)
    p = np.zeros(api.pnet_param_count(), np.float32)
    o = 0
    for ci, co in LAYERS:
        p[o:o + co * ci * 27] = rng.standard_normal(co * ci * 27) * math.sqrt(2.0 / (ci * 27))
        o += co * ci * 27 + co
    p[o:o + 768] = rng.standard_normal(768) * math.sqrt(1.0 / 256) * 0.5
    return p


def save_params(path: str, params: np.ndarray, meta: dict) -> None:
    from safetensors.numpy import save_file
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    save_file({"params": np.ascontiguousarray(params, np.float32)}, path,
              metadata={"meta": json.dumps(meta, sort_keys=True)})


def load_params(path: str = WEIGHTS) -> tuple[np.ndarray, dict]:
    from safetensors import safe_open
    with safe_open(path, framework="numpy") as f:
        p = f.get_tensor("params")
        meta = json.loads((f.metadata() or {}).get("meta", "{}"))
    if p.shape != (api.pnet_param_count(),):
        raise ValueError(f"{path}: {p.shape[0]} parameters, the network has {api.pnet_param_count()}")
    return p.astype(np.float32), meta


@dataclass
class PairSet:
    """Device-resident envelope pairs with their full-search labels."""
    shape: tuple
    vols: list = field(default_factory=list)        # [(tpl, src)] torch tensors
    shifts: list = field(default_factory=list)      # true shifts (voxels)
    disp: list = field(default_factory=list)        # full-search displacement, int [nb][3] (numpy)
    peak: list = field(default_factory=list)        # full-search peak, [nb]
    full_ms: list = field(default_factory=list)

    @property
    def nb(self) -> int:
        nz, ny, nx = self.shape
        s = GRID_STRIDE
        return ((nx - 16) // s + 1) * ((ny - 16) // s + 1) * ((nz - 16) // s + 1)

    def block_origins(self) -> np.ndarray:
        nz, ny, nx = self.shape
        gx, gy = (nx - 16) // GRID_STRIDE + 1, (ny - 16) // GRID_STRIDE + 1
        b = np.arange(self.nb)
        return np.stack([b % gx, (b // gx) % gy, b // (gx * gy)], 1) * GRID_STRIDE


def _timed(fn, stream):
    import torch
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1)


def make_pairs(n: int, seed: int, dev, log=print, limit: float = 6.0, **rca_kw) -> PairSet:
    import torch
    rng = np.random.default_rng(seed)
    shifts = motion_data.random_shifts(rng, n, limit)
    stream = torch.cuda.current_stream()
    out = None
    for k in range(n):
        t0 = time.time()
        tpl, src, _ = motion_data.envelope_pair(seed * 100003 + k, shifts[k], dev, **rca_kw)
        if out is None:
            out = PairSet(tuple(tpl.shape))
        nb = out.nb
        disp = torch.zeros((nb, 3), dtype=torch.int32, device=dev)
        peak = torch.zeros(nb, dtype=torch.float32, device=dev)
        ms = _timed(lambda: api.ncc3d_search(tpl, src, GRID_STRIDE, disp, peak), stream)
        out.vols.append((tpl, src))
        out.shifts.append(shifts[k])
        out.disp.append(disp.cpu().numpy())
        out.peak.append(peak.cpu().numpy())
        out.full_ms.append(ms)
        if log and (k % 8 == 0 or k == n - 1):
            log(f"pair {k + 1}/{n} shift {np.round(shifts[k], 2)} peak>=0.6 {np.mean(out.peak[-1] >= 0.6):.3f} "
                f"({time.time() - t0:.2f} s)")
    return out


def train(pairs: PairSet, steps: int = 6000, batch: int = 256, lr: float = 1e-3, seed: int = 0,
          threshold: float = 0.6, max_offset: int = 2, wd: float = 1e-4, params=None, log=print,
          eval_fn=None, eval_every: int = 1000) -> np.ndarray:
    import torch
    dev = pairs.vols[0][0].device
    rng = np.random.default_rng(seed)
    nb = pairs.nb
    origins = pairs.block_origins()
    nz, ny, nx = pairs.shape
    # candidate (pair, block) with a confident label
    cand = [(k, b) for k in range(len(pairs.vols)) for b in np.nonzero(pairs.peak[k] >= threshold)[0]]
    cand = np.array(cand, np.int64)
    labels = np.stack([pairs.disp[k] for k in range(len(pairs.vols))])  # [pairs][nb][3]
    if log:
        log(f"training on {len(cand)} confident blocks of {len(pairs.vols) * nb}")
    vol_ptrs = []
    for tpl, src in pairs.vols:
        vol_ptrs += [tpl.data_ptr(), src.data_ptr()]
    d_vols = torch.tensor(vol_ptrs, dtype=torch.int64, device=dev)
    p = torch.from_numpy(init_params(seed) if params is None else params.astype(np.float32)).to(dev)
    g = torch.zeros_like(p)
    m = torch.zeros_like(p)
    v = torch.zeros_like(p)
    x = torch.zeros((batch, 2, 16, 16, 16), dtype=torch.float32, device=dev)
    loss = torch.zeros(1, dtype=torch.float32, device=dev)
    hi = np.array([nx - 16, ny - 16, nz - 16])
    run = 0.0
    t0 = time.time()
    for step in range(1, steps + 1):
        pick = cand[rng.integers(0, len(cand), batch)]
        k, b = pick[:, 0], pick[:, 1]
        d = labels[k, b].astype(np.int64)
        off = rng.integers(-max_offset, max_offset + 1, (batch, 3))
        o = origins[b]
        off = np.clip(o + off, 0, hi) - o                 # source block inside the volume
        off = np.where(np.abs(d - off) <= 8, off, 0)      # label stays within the search range
        flips = rng.integers(0, 8, batch)
        t = (d - off).astype(np.float32)
        for a in range(3):
            t[:, a] = np.where(flips & (1 << a), -t[:, a], t[:, a])
        desc = np.stack([k, b, off[:, 0], off[:, 1], off[:, 2], flips], 1).astype(np.int32)
        assert api.pnet_gather(d_vols, pairs.shape, GRID_STRIDE, torch.from_numpy(desc).to(dev), x)
        lr_t = lr * 0.5 * (1 + math.cos(math.pi * (step - 1) / steps))
        assert api.pnet_loss_grad(p, x, torch.from_numpy(t).to(dev), g, loss)
        assert api.pnet_adam(p, g, m, v, lr=lr_t, step=step, wd=wd)
        if step % 50 == 0:
            run = 0.9 * run + 0.1 * loss.item() if run else loss.item()
        if log and (step % 250 == 0 or step == steps):
            log(f"step {step}/{steps} loss {run:.4f} lr {lr_t:.2e} ({time.time() - t0:.1f} s)")
        if eval_fn is not None and (step % eval_every == 0 or step == steps):
            eval_fn(p.cpu().numpy(), step)
    return p.cpu().numpy()


def evaluate(params: np.ndarray, pairs: PairSet, radius: int = 2, threshold: float = 0.6, boundary_check=True,
             repeats: int = 3) -> dict:
    """Learned search against the full search on every block of every pair."""
    import torch
    dev = pairs.vols[0][0].device
    stream = torch.cuda.current_stream()
    dp = torch.from_numpy(params.astype(np.float32)).to(dev)
    nb = pairs.nb
    agree = fall = within = inner = conf = 0
    t_learn = t_full = 0.0
    for k, (tpl, src) in enumerate(pairs.vols):
        disp = torch.zeros((nb, 3), dtype=torch.int32, device=dev)
        peak = torch.zeros(nb, dtype=torch.float32, device=dev)
        path = torch.zeros(nb, dtype=torch.int32, device=dev)
        pred = torch.zeros((nb, 3), dtype=torch.float32, device=dev)
        ok, nf = api.motion3d_learned(dp, tpl, src, GRID_STRIDE, disp, peak, radius=radius, threshold=threshold,
                                      boundary_check=boundary_check, d_path=path, d_pred=pred, count_fallbacks=True)
        assert ok
        ms = [_timed(lambda: api.motion3d_learned(dp, tpl, src, GRID_STRIDE, disp, peak, radius=radius,
                                                  threshold=threshold, boundary_check=boundary_check), stream)
              for _ in range(repeats)]
        d2 = torch.zeros_like(disp)
        p2 = torch.zeros_like(peak)
        fms = [_timed(lambda: api.ncc3d_search(tpl, src, GRID_STRIDE, d2, p2), stream) for _ in range(repeats)]
        t_learn += min(ms)
        t_full += min(fms)
        got = disp.cpu().numpy()
        agree += int(np.all(got == pairs.disp[k], axis=1).sum())
        fall += nf
        pr = np.rint(pred.cpu().numpy()).astype(np.int64)
        c = pairs.peak[k] >= threshold
        conf += int(c.sum())
        err = np.abs(pr - pairs.disp[k]).max(axis=1)
        within += int((err <= radius)[c].sum())
        inner += int((err < radius)[c].sum())
    n = nb * len(pairs.vols)
    return {
        "blocks": n,
        "agreement": agree / n,
        "fallback_rate": fall / n,
        "proposal_within_radius": within / max(conf, 1),
        "proposal_inside_window": inner / max(conf, 1),
        "confident_fraction": conf / n,
        "learned_ms_per_pair": t_learn / len(pairs.vols),
        "full_ms_per_pair": t_full / len(pairs.vols),
        "speedup": t_full / t_learn if t_learn > 0 else float("nan"),
        "learned_blocks_per_s": n / (t_learn * 1e-3),
        "full_blocks_per_s": n / (t_full * 1e-3),
    }


def main(argv=None) -> int:
    import torch
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--pairs", type=int, default=160)
    ap.add_argument("--eval-pairs", type=int, default=8)
    ap.add_argument("--steps", type=int, default=24000)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--report", default=None, help="JSON file for the final evaluation")
    a = ap.parse_args(argv)
    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    api.set_stream(stream.cuda_stream)
    assert api.set_das_options(cubic=False)
    assert api.pnet_set_input_mode(INPUT_MODE)

    def log(msg):
        print(msg, flush=True)

    log("generating training pairs")
    tr = make_pairs(a.pairs, a.seed, dev, log)
    log("generating evaluation pairs")
    ev = make_pairs(a.eval_pairs, a.seed + 1000, dev, log)
    zero = np.zeros(api.pnet_param_count(), np.float32)
    base = evaluate(zero, ev)
    log(f"without the network (proposal 0): {json.dumps(base)}")

    def eval_fn(p, step):
        r = evaluate(p, ev, repeats=1)
        log(f"  eval @ {step}: agreement {r['agreement']:.4f} fallback {r['fallback_rate']:.3f} "
            f"within-radius {r['proposal_within_radius']:.3f} inside {r['proposal_inside_window']:.3f} "
            f"speedup {r['speedup']:.2f}")

    params = train(tr, steps=a.steps, batch=a.batch, lr=a.lr, seed=a.seed, log=log, eval_fn=eval_fn,
                   eval_every=max(1, a.steps // 4))
    sweep = {}
    for thr in (0.6, 0.65, 0.7, 0.75, 0.8):
        sweep[f"{thr:.2f}"] = evaluate(params, ev, threshold=thr)
        log(f"trained, threshold {thr:.2f}: {json.dumps(sweep[f'{thr:.2f}'])}")
    meta = {"pairs": a.pairs, "steps": a.steps, "batch": a.batch, "lr": a.lr, "seed": a.seed,
            "input_mode": INPUT_MODE, "eval_pairs": a.eval_pairs, "eval_by_threshold": sweep,
            "baseline_zero_proposal": base}
    save_params(a.out, params, meta)
    if a.report:
        with open(a.report, "w") as f:
            json.dump(meta, f, indent=1)
    log(f"saved {a.out}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
