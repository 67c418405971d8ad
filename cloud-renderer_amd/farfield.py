"""Training driver of the learned far field (host side).

Training rows come from the plain march on the GPU (cr_volume_segments): for
each seeded scene the volume is voxelized, cones are started at points
distributed like camera fragments (uniform over billboard discs facing an
orbit camera, lifted onto the sphere: pos = c + du S + dv U + B r sc, the
fragment position of res/conetrace_frag.glsl:185-186), and every cone still
running after k samples yields (features, far-field targets).  The network is
then trained with the library's own forward/backward kernels and Adam
(cr_mlp_train_step).  Everything runs on the GPU; torch only holds buffers
and shuffles row indices.
"""
from __future__ import annotations

import math
import time
from dataclasses import replace
from pathlib import Path
from typing import Callable, List, Optional

import numpy as np

WEIGHTS_DIR = Path(__file__).resolve().parent / "weights"


def weights_path(k: int = 32) -> Path:
    return WEIGHTS_DIR / f"farfield_k{k}.npz"


def fragment_starts(scene, n: int, rng: np.random.Generator, radius: float = 3.0) -> np.ndarray:
    """n cone start points (texture space, float32 (n, 3)) distributed like the
    camera fragments of `scene` seen from random azimuths of an orbit camera."""
    xyz = scene.xyz.astype(np.float64) + np.asarray(scene.position, np.float64)
    r = scene.scales.astype(np.float64) * float(scene.fluffiness)
    w = r * r
    j = rng.choice(len(r), size=n, p=w / w.sum())
    az = rng.uniform(0.0, 2.0 * math.pi, size=n)
    eye = np.stack([radius * np.sin(az), np.zeros(n), radius * np.cos(az)], axis=1)
    back = eye / np.linalg.norm(eye, axis=1, keepdims=True)          # view back axis (toward the eye)
    up = np.array([0.0, 1.0, 0.0])
    side = np.cross(up, back)
    side /= np.linalg.norm(side, axis=1, keepdims=True)
    upv = np.cross(back, side)
    rho = np.sqrt(rng.uniform(0.0, 1.0, size=n)) * math.sqrt(1.0 - 0.01 ** 2)
    phi = rng.uniform(0.0, 2.0 * math.pi, size=n)
    du, dv = rho * np.cos(phi), rho * np.sin(phi)
    sc = np.sqrt(np.maximum(0.0, 1.0 - rho * rho))
    rr = r[j][:, None]
    pos = xyz[j] + rr * (du[:, None] * side + dv[:, None] * upv + sc[:, None] * back)
    lo = np.asarray(scene.position, np.float64) + np.asarray(scene.bounds_lo, np.float64)
    hi = np.asarray(scene.position, np.float64) + np.asarray(scene.bounds_hi, np.float64)
    return ((pos - lo) / (hi - lo)).astype(np.float32)


def generate_segments(cv, scenes: List, per_scene: int, k: int, seed: int = 0, context=None,
                      log: Optional[Callable[[str], None]] = None):
    """Run the plain march over `scenes` and keep the rows still running after
    k samples.  Returns (X (m, 17), T (m, 5)) float32 CUDA tensors."""
    import torch

    rng = np.random.Generator(np.random.PCG64(seed))
    xs, ts = [], []
    vol = None
    shape = None
    t0 = time.perf_counter()
    for i, sc in enumerate(scenes):
        key = (sc.N, sc.levels, sc.fmt, sc.mode, sc.bounds_lo, sc.bounds_hi, sc.position)
        if vol is None or key != shape:
            if vol is not None:
                vol.close()
            vol = cv.CloudVolume(sc.N, (sc.bounds_lo, sc.bounds_hi), sc.position, sc.levels, sc.fmt, sc.mode,
                                 context=context)
            shape = key
        vol.fluffiness = sc.fluffiness
        vol.set_billboards(sc.xyz, sc.scales)
        if sc.noise is not None:
            vol.set_noise(sc.noise.octaves, sc.noise.seed, sc.noise.base_freq, sc.noise.lacunarity, sc.noise.gain)
        else:
            vol.set_noise(0)
        _, sun, _, vox, tr = _objects(cv, sc, vol)
        vox.voxelize(vol, sun)
        p0 = torch.from_numpy(fragment_starts(sc, per_scene, rng)).cuda()
        f, t, v = tr.segments(vol, sun, p0, k)
        keep = v.bool()
        xs.append(f[keep])
        ts.append(t[keep])
        if log and (i + 1) % 64 == 0:
            torch.cuda.synchronize()
            log(f"[farfield] segments: {i + 1}/{len(scenes)} scenes, "
                f"{sum(int(x.shape[0]) for x in xs)} rows, {time.perf_counter() - t0:.1f} s")
    if vol is not None:
        vol.close()
    return torch.cat(xs), torch.cat(ts)


def _objects(cv, sc, vol):
    sun = cv.Sun(position=tuple(float(x) for x in sc.sun_pos), direction=tuple(float(x) for x in sc.sun_dir),
                 directional=True)
    vox = cv.VoxelizeShader(sc.light_w, sc.light_h)
    c = sc.cone
    tr = cv.ConeTraceShader(vctSteps=c.steps, vctConeAngle=c.cone_angle, vctConeInitialHeight=c.initial_height,
                            vctLodOffset=c.lod_offset, vctDownScaling=c.down_scaling, coneMode=c.cone_mode,
                            alphaStop=c.alpha_stop, sigma=c.sigma, composite=c.composite,
                            quantizeRGBA8=bool(c.quantize_rgba8), compensated=bool(c.compensated),
                            clearColor=c.clear_color)
    return vol, sun, None, vox, tr


def train(mlp, X, T, epochs: int, batch: int = 16384, lr: float = 2e-3, lr_final: float = 1e-4, seed: int = 0,
          log: Optional[Callable[[str], None]] = None) -> List[float]:
    """Adam over shuffled mini-batches with a cosine learning-rate decay from
    lr to lr_final.  Returns the mean loss of every epoch."""
    import torch

    n = int(X.shape[0])
    steps_per_epoch = max(1, n // batch)
    total = epochs * steps_per_epoch
    g = torch.Generator(device=X.device)
    g.manual_seed(seed)
    losses = []
    step = 0
    t0 = time.perf_counter()
    for ep in range(epochs):
        perm = torch.randperm(n, device=X.device, generator=g)
        acc, cnt = 0.0, 0
        for b in range(steps_per_epoch):
            idx = perm[b * batch:(b + 1) * batch]
            xb = X.index_select(0, idx).contiguous()
            tb = T.index_select(0, idx).contiguous()
            frac = step / max(1, total - 1)
            cur = lr_final + 0.5 * (lr - lr_final) * (1.0 + math.cos(math.pi * frac))
            mlp.train_step(xb, tb, cur)
            step += 1
            if b % 64 == 0 or b == steps_per_epoch - 1:
                acc += mlp.last_loss()
                cnt += 1
        losses.append(acc / cnt)
        if log:
            log(f"[farfield] epoch {ep + 1}/{epochs}: loss {losses[-1]:.3e}, {time.perf_counter() - t0:.1f} s")
    return losses


def training_scenes(cv, count: int, seed0: int = 1000, W: int = 480, H: int = 270):
    """Seeded plain256 scenes (clumped spheres, hemisphere sun) for training;
    the light map is rendered at W x H (it only decides the lit shell)."""
    out = []
    for i in range(count):
        s = cv.scenes.plain256(seed0 + i)
        out.append(replace(s, light_w=1920, light_h=1080, W=W, H=H))
    return out


def evaluate(cv, mlp, scene, k: int, bounds, cam_index: int = 0, context=None, repeats: int = 5):
    """Render `scene` plainly, truncated after k samples, and learned at every
    error bound in `bounds`.  Returns a dict with the images (CPU numpy), the
    cone-stage time of each (median of `repeats` event-timed calls) and the
    learned passes' counters."""
    import torch

    vol, sun, cam, vox, tr = cv.scene_objects(scene, cam_index, context=context)
    ctx = vol.ctx
    W, H = scene.W, scene.H

    def run(fn):
        img = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
        ctx.set_profiling(1)
        fn(img)
        st = ctx.stats()
        times = []
        for _ in range(repeats):
            fn(img)
            times.append(ctx.stats()["ms_cone"])
        ctx.set_profiling(0)
        torch.cuda.synchronize()
        st["ms_cone"] = float(np.median(times))
        return img.cpu().numpy(), st

    try:
        vox.voxelize(vol, sun)
        out = {}
        out["plain"] = run(lambda img: tr.coneTrace(vol, sun, cam, img, W, H))
        steps = tr.vctSteps
        tr.vctSteps = min(k, steps)
        out["truncated"] = run(lambda img: tr.coneTrace(vol, sun, cam, img, W, H))
        tr.vctSteps = steps
        for b in bounds:
            out[b] = run(lambda img, b=b: tr.coneTraceLearned(vol, sun, cam, mlp, img, W, H, k=k, err_bound=b))
        return out
    finally:
        vol.close()


def psnr(a: np.ndarray, b: np.ndarray) -> float:
    mse = float(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2))
    return float("inf") if mse == 0 else 10.0 * math.log10(1.0 / mse)
