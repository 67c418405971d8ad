"""Shared helpers for the GPU parity tests: run a scene through the HIP
library and through the oracle, and compare."""
from __future__ import annotations

import sys
from dataclasses import dataclass

import numpy as np

import oracle_lib as O


def log(msg: str) -> None:
    print(msg, flush=True)
    sys.stdout.flush()


@dataclass
class GpuRun:
    state: np.ndarray
    count: int
    hash: int
    posmap: np.ndarray
    levels: list
    order: np.ndarray
    image: np.ndarray
    stats: dict


def run_gpu(cv, scene, cam_index=0, ctx=None, read_levels=True, rows=None, pixel_format=0):
    import torch
    vol, sun, cam, vox, tr = cv.scene_objects(scene, cam_index, context=ctx)
    try:
        vox.voxelize(vol, sun)
        if pixel_format == 0:
            img = torch.zeros((scene.H, scene.W, 4), dtype=torch.float32, device="cuda")
        else:
            img = torch.zeros((scene.H, scene.W, 4), dtype=torch.uint8, device="cuda")
        tr.coneTrace(vol, sun, cam, img, scene.W, scene.H, pixel_format=pixel_format)
        torch.cuda.synchronize()
        state = vol.read_occupancy()
        cnt, h = vol.occupancy()
        pm = vol.read_position_map(scene.light_w, scene.light_h)
        levels = [vol.read_level(k) for k in range(scene.levels)] if read_levels else []
        order = vol.read_draw_order()
        image = img.cpu().numpy()
        stats = vol.ctx.stats()
    finally:
        vol.close()
    return GpuRun(state, cnt, h, pm, levels, order, image, stats)


@dataclass
class OracleRun:
    vox: object
    count: int
    hash: int
    levels: list
    levels64: list
    order: np.ndarray
    img32: np.ndarray
    img64: np.ndarray
    stats: object


def run_oracle(scene, cam_index=0, row0=0, row1=None, want_f64=True):
    vr = O.voxelize(scene)
    cnt, h = O.occupancy(scene.N, vr.shell, vr.fill)
    nat, f64, nl, fl = O.build_levels(scene, vr.shell, vr.fill, want_f64=want_f64)
    cam = scene.cameras[cam_index]
    order = O.sort_boards(scene, cam.eye)
    i32, i64, st = O.cone_trace(scene, nat, f64, order, cam, row0, row1, want_f64=want_f64)
    return OracleRun(vr, cnt, h, nl, fl, order, i32, i64, st)


def bits(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a)
    if a.dtype == np.float32:
        return a.view(np.uint32)
    return a


def assert_bit_equal(name: str, got: np.ndarray, want: np.ndarray) -> None:
    assert got.shape == want.shape, f"{name}: shape {got.shape} != {want.shape}"
    gb, wb = bits(got), bits(want)
    if not np.array_equal(gb, wb):
        diff = np.argwhere(gb != wb)
        first = tuple(diff[0])
        raise AssertionError(f"{name}: {len(diff)} of {gb.size} elements differ; first at {first}: "
                             f"got {got[first]!r} want {want[first]!r}")


def psnr(a: np.ndarray, b: np.ndarray) -> float:
    mse = float(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2))
    return float("inf") if mse == 0 else 10.0 * np.log10(1.0 / mse)
