"""Batched scenes (one hipGraph, one cone-trace launch over all scenes' tiles).

The per-scene path (cr_voxelize + cr_cone_trace) is pinned to the oracle by
the parity tests; here the batch must reproduce it bit for bit:
  * under several launch shapes (tile 8x8 / 16x4 / 4x16, 1-4 waves per
    workgroup, one workgroup per tile or a small persistent grid);
  * on every replay of the captured graph, with eager per-scene work
    between the replays;
  * after a scene's billboards change (the graph is captured again);
  * after the autotuner has picked a launch shape.
Profiling counters of an eager batch run equal the sums of the scenes'.
"""
from dataclasses import replace

import numpy as np
import pytest

from gpu_common import log

pytestmark = pytest.mark.gpu


def _scenes(cv, n=6, seed=3):
    out = []
    for sc in cv.scenes.batched64(seed, count=n):
        side = min(sc.W, 384)
        out.append(replace(sc, W=side, H=side, light_w=side, light_h=side))
    return out


def _single(cv, ctx, scene):
    import torch
    vol, sun, cam, vox, tr = cv.scene_objects(scene, 0, context=ctx)
    try:
        vox.voxelize(vol, sun)
        img = torch.zeros((scene.H, scene.W, 4), dtype=torch.float32, device="cuda")
        tr.coneTrace(vol, sun, cam, img, scene.W, scene.H)
        torch.cuda.synchronize()
        return img.cpu().numpy()
    finally:
        vol.close()


def _batch(cv, ctx, scenes):
    import torch
    entries, objs = [], []
    for sc in scenes:
        vol, sun, cam, vox, tr = cv.scene_objects(sc, 0, context=ctx)
        img = torch.full((sc.H, sc.W, 4), -1.0, dtype=torch.float32, device="cuda")
        entries.append((vol, sun, cam, tr, img, vox))
        objs.append(vol)
    return cv.Batch(entries, context=ctx), entries


def _check(entries, want, label):
    for i, (vol, sun, cam, tr, img, vox) in enumerate(entries):
        got = img.cpu().numpy()
        assert np.array_equal(got.view(np.uint32), want[i].view(np.uint32)), f"{label}: scene {i} differs"


CONFIGS = [(8, 1, 0), (16, 1, 0), (4, 4, 0), (8, 2, 64), (16, 4, 16)]


def test_batch_matches_single_scenes(cv, gpu_ctx):
    import torch
    scenes = _scenes(cv)
    log(f"[batch] {len(scenes)} scenes: " + ", ".join(f"N={s.N} S={s.S} {s.W}x{s.H}" for s in scenes))
    want = [_single(cv, gpu_ctx, sc) for sc in scenes]
    b, entries = _batch(cv, gpu_ctx, scenes)
    try:
        for cfg in CONFIGS:
            b.set_config(*cfg)
            for rep in range(3):
                for _, _, _, _, img, _ in entries:
                    img.fill_(-1.0)
                b.run()
                torch.cuda.synchronize()
                _check(entries, want, f"config {cfg} replay {rep}")
                # eager per-scene work between replays
                _single(cv, gpu_ctx, scenes[rep % len(scenes)])
            log(f"[batch] config {cfg}: 3 replays equal the per-scene images")
    finally:
        b.close()


def test_batch_recaptures_after_billboard_change(cv, gpu_ctx):
    import torch
    scenes = _scenes(cv, n=3, seed=4)
    b, entries = _batch(cv, gpu_ctx, scenes)
    try:
        b.run()
        torch.cuda.synchronize()
        rng = np.random.default_rng(1)
        moved = replace(scenes[1], xyz=(scenes[1].xyz + rng.normal(0, 0.05, scenes[1].xyz.shape)).astype(np.float32))
        vol = entries[1][0]
        vol.set_billboards(moved.xyz, moved.scales)
        b.run()
        torch.cuda.synchronize()
        want = [_single(cv, gpu_ctx, sc) for sc in (scenes[0], moved, scenes[2])]
        _check(entries, want, "after set_billboards")
        log("[batch] recaptured after a billboard change")
    finally:
        b.close()


def test_batch_autotune_and_counters(cv, gpu_ctx):
    import torch
    scenes = _scenes(cv, n=4, seed=5)
    want = [_single(cv, gpu_ctx, sc) for sc in scenes]
    samples = 0
    for sc in scenes:
        vol, sun, cam, vox, tr = cv.scene_objects(sc, 0, context=gpu_ctx)
        img = torch.zeros((sc.H, sc.W, 4), dtype=torch.float32, device="cuda")
        vox.voxelize(vol, sun)
        gpu_ctx.set_profiling(1)
        tr.coneTrace(vol, sun, cam, img, sc.W, sc.H)
        samples += gpu_ctx.stats()["samples"]
        gpu_ctx.set_profiling(0)
        vol.close()
    b, entries = _batch(cv, gpu_ctx, scenes)
    try:
        best, ms = b.autotune(repeats=2)
        log(f"[batch] autotuned: {best}, {ms:.3f} ms per run")
        assert best["tile_w"] in (4, 8, 16) and best["waves_per_group"] in (1, 2, 4) and ms > 0
        for _, _, _, _, img, _ in entries:
            img.fill_(-1.0)
        b.run()
        torch.cuda.synchronize()
        _check(entries, want, "autotuned")
        gpu_ctx.set_profiling(1)
        b.run()
        st = gpu_ctx.stats()
        gpu_ctx.set_profiling(0)
        assert st["samples"] == samples
        _check(entries, want, "eager profiled run")
    finally:
        b.close()


def test_batch_after_single_traces_on_its_volumes(cv, gpu_ctx):
    """Single voxelize + trace calls on a batch's volumes between runs, one
    with a larger image (its camera buffers grow and are reallocated): the
    next run captures again where needed and still reproduces every scene."""
    import torch
    scenes = _scenes(cv, n=3, seed=6)
    want = [_single(cv, gpu_ctx, sc) for sc in scenes]
    b, entries = _batch(cv, gpu_ctx, scenes)
    try:
        b.run()
        torch.cuda.synchronize()
        _check(entries, want, "first run")
        vol, sun, cam, tr, _, vox = entries[0]
        sc = scenes[0]
        for k in range(3):
            # two traces, so both camera buffer pairs are used; the second
            # one of round 1 is larger than anything the batch traced
            for t in range(2):
                w, h = (2 * sc.W, 2 * sc.H) if (k == 1 and t == 1) else (sc.W, sc.H)
                big = torch.zeros((h, w, 4), dtype=torch.float32, device="cuda")
                vox.voxelize(vol, sun)
                tr.coneTrace(vol, sun, cam, big, w, h)
            torch.cuda.synchronize()
            for _, _, _, _, img, _ in entries:
                img.fill_(-1.0)
            b.run()
            torch.cuda.synchronize()
            _check(entries, want, f"run after single trace {k}")
        log("[batch] runs after single traces on a batch volume equal the per-scene images")
    finally:
        b.close()


def test_batched64_full_size(cv, gpu_ctx):
    """The batched workload at its native sizes: the 64 scenes of
    batched64(0) (volume sides 64..256, 256..8192 spheres, images up to
    1024^2) rendered by one batch graph equal the per-scene path bit for bit,
    and four of them -- the largest (N = 256, 1024^2, 5986 spheres), the one
    with most spheres, a mid-size one and the smallest -- equal the oracle's
    FP32 pass on every 8th image row bit for bit and its FP64 replay within
    1e-3."""
    import torch
    import oracle_lib as O
    from gpu_common import assert_bit_equal, psnr

    scenes = cv.scenes.batched64(0)
    log(f"[batch64] {len(scenes)} scenes, {sum(s.W * s.H for s in scenes)} pixels")
    want = [_single(cv, gpu_ctx, sc) for sc in scenes]
    log("[batch64] per-scene images done")
    b, entries = _batch(cv, gpu_ctx, scenes)
    try:
        b.run()
        torch.cuda.synchronize()
        _check(entries, want, "batched64 full")
        b.run()
        torch.cuda.synchronize()
        _check(entries, want, "batched64 full, second replay")
        log("[batch64] two graph runs equal the per-scene images")
    finally:
        b.close()
        for e in entries:
            e[0].close()
    size = [s.N ** 3 * s.W * s.H for s in scenes]
    picks = [int(np.argmax(size)), int(np.argmax([s.S for s in scenes])), int(np.argsort(size)[len(size) // 2]),
             int(np.argmin(size))]
    step = 8
    for i in dict.fromkeys(picks):
        sc = scenes[i]
        log(f"[batch64] scene {i}: N={sc.N} S={sc.S} {sc.W}x{sc.H} against the oracle, every {step}th row")
        vr = O.voxelize(sc)
        nat, f64, _, _ = O.build_levels(sc, vr.shell, vr.fill)
        cam = sc.cameras[0]
        order = O.sort_boards(sc, cam.eye)
        i32, i64, _ = O.cone_trace(sc, nat, f64, order, cam, row_step=step)
        rows = np.arange(0, sc.H, step)
        assert_bit_equal(f"scene {i} rows (FP32)", want[i][rows], i32[rows])
        err = float(np.abs(want[i][rows].astype(np.float64) - i64[rows]).max())
        log(f"[batch64] scene {i}: rows bit-exact, max-abs vs FP64 {err:.2e}, "
            f"PSNR {psnr(want[i][rows], i64[rows]):.1f} dB")
        assert err <= 1e-3
