#!/usr/bin/env python3
"""Train the learned far field on the GPU and save its parameters.

    python cloud-renderer_amd/train_farfield.py [--k 32] [--scenes 512] [--rows 16777216]

Training rows (features, far-field targets) come from the plain march over
seeded clumped-sphere scenes at 256^3 (seeds 1000 + i; the evaluation scenes
use seeds 0..15), cones started at fragment-like points.  Rows are cut to
--rows (2^24 by default).  The network is trained with the library's own
forward/backward kernels and Adam (cosine learning-rate decay), validated on
rows of held-out scenes, and evaluated on a held-out frame: PSNR of the
learned trace and of the trace truncated after k samples, both against the
plain trace, and the fallback rate.  Parameters and the summary are written
to cloud-renderer_amd/weights/farfield_k<k>.npz.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import sys
import time
from dataclasses import replace
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent


def _load_pkg():
    name = "cloud_renderer_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, PKG_DIR / "__init__.py",
                                                  submodule_search_locations=[str(PKG_DIR)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg: str) -> None:
    print(msg, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--k", type=int, default=32)
    ap.add_argument("--scenes", type=int, default=512)
    ap.add_argument("--per-scene", type=int, default=40000)
    ap.add_argument("--rows", type=int, default=1 << 24)
    ap.add_argument("--epochs", type=int, default=16)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--lr-final", type=float, default=5e-5)
    ap.add_argument("--err-bound", type=float, default=2e-3)
    ap.add_argument("--eval-w", type=int, default=1920)
    ap.add_argument("--eval-h", type=int, default=1080)
    ap.add_argument("--out", type=str, default=None)
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args()

    import torch

    if not torch.cuda.is_available():
        log("train_farfield.py needs a GPU")
        return 2
    cv = _load_pkg()
    from cloud_renderer_amd import farfield as ff

    ctx = cv.Context(0)
    t0 = time.perf_counter()
    scenes = ff.training_scenes(cv, args.scenes, seed0=1000)
    X, T = ff.generate_segments(cv, scenes, args.per_scene, args.k, seed=args.seed, context=ctx, log=log)
    torch.cuda.synchronize()
    n_all = int(X.shape[0])
    if n_all > args.rows:
        X, T = X[:args.rows].contiguous(), T[:args.rows].contiguous()
    log(f"[train] {n_all} rows from {args.scenes} scenes, keeping {int(X.shape[0])}, "
        f"{time.perf_counter() - t0:.1f} s")
    val_scenes = ff.training_scenes(cv, 8, seed0=5000)
    Xv, Tv = ff.generate_segments(cv, val_scenes, 8192, args.k, seed=args.seed + 1, context=ctx)

    mlp = cv.FarFieldMLP(ctx, cv.FarFieldMLP.init_params(args.seed))
    _, v0 = mlp.gradient(Xv, Tv)
    t1 = time.perf_counter()
    losses = ff.train(mlp, X, T, args.epochs, batch=args.batch, lr=args.lr, lr_final=args.lr_final,
                      seed=args.seed, log=log)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t1
    _, v1 = mlp.gradient(Xv, Tv)
    log(f"[train] validation loss {v0:.3e} -> {v1:.3e}; training {t_train:.1f} s")

    # held-out frame: plain256 seed 0, camera 0
    ev = cv.scenes.scaled(cv.scenes.plain256(0), args.eval_w, args.eval_h)
    ev = replace(ev, light_w=1920, light_h=1080)
    bounds = [args.err_bound, 5e-3, 1e-2, 2e-2, 5e-2, 1e30]
    res = ff.evaluate(cv, mlp, ev, args.k, bounds, context=ctx)
    plain, st_p = res["plain"]
    trunc, st_t = res["truncated"]
    sweep = []
    for b in bounds:
        img, st = res[b]
        rows = max(1, int(st["mlp_rows"]))
        sweep.append({"err_bound": b, "psnr_vs_plain": round(ff.psnr(img, plain), 2),
                      "maxabs_vs_plain": float(np.abs(img - plain).max()),
                      "fallback_rate": round(int(st["mlp_fallbacks"]) / rows, 4), "mlp_rows": int(st["mlp_rows"]),
                      "samples": int(st["samples"]), "ms_cone": round(st["ms_cone"], 4)})
        log(f"[train] bound {b:g}: " + json.dumps(sweep[-1]))
    summary = {
        "k": args.k, "rows": int(X.shape[0]), "rows_generated": n_all, "scenes": args.scenes,
        "epochs": args.epochs, "batch": args.batch, "lr": args.lr, "lr_final": args.lr_final,
        "train_loss": [round(x, 6) for x in losses], "val_loss_init": v0, "val_loss": v1,
        "train_seconds": round(t_train, 1),
        "eval": {"scene": f"plain256 seed 0 camera 0 at {args.eval_w}x{args.eval_h}",
                 "plain": {"samples": int(st_p["samples"]), "ms_cone": round(st_p["ms_cone"], 4)},
                 "truncated": {"psnr_vs_plain": round(ff.psnr(trunc, plain), 2), "samples": int(st_t["samples"]),
                               "ms_cone": round(st_t["ms_cone"], 4)},
                 "learned": sweep},
    }
    log("[train] " + json.dumps(summary))
    out = Path(args.out) if args.out else ff.weights_path(args.k)
    out.parent.mkdir(parents=True, exist_ok=True)
    np.savez_compressed(out, params=mlp.get_params(), summary=np.array(json.dumps(summary)))
    log(f"[train] wrote {out}")
    mlp.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
