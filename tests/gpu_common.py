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
    active: int = 0


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
        active = vol.active_voxels()
        image = img.cpu().numpy()
        stats = vol.ctx.stats()
    finally:
        vol.close()
    return GpuRun(state, cnt, h, pm, levels, order, image, stats, active)


@dataclass
class OracleRun:
    vox: object
    count: int
    hash: int
    levels: list
    levels64: list
    order: np.ndarray
    img32: np.ndarray
    img64: np.ndarray          # FP64 replay over the exact (unrounded) mip means
    stats: object
    img64_stored: np.ndarray = None  # R8 volumes: FP64 replay over the stored R8 levels (byte / 255)
    active: int = 0            # activeVoxels: level-0 texels with r != 0 (the lit shell)


def run_oracle(scene, cam_index=0, row0=0, row1=None, want_f64=True):
    vr = O.voxelize(scene)
    cnt, h = O.occupancy(scene.N, vr.shell, vr.fill)
    nat, f64, nl, fl = O.build_levels(scene, vr.shell, vr.fill, want_f64=want_f64)
    cam = scene.cameras[cam_index]
    order = O.sort_boards(scene, cam.eye)
    i32, i64, st = O.cone_trace(scene, nat, f64, order, cam, row0, row1, want_f64=want_f64)
    i64s = None
    if want_f64 and scene.fmt == 0:
        # the same replay over the R8 levels as stored, so that the FP64
        # comparison judges the arithmetic and not the R8 rounding of the mips
        _, i64s, _ = O.cone_trace(scene, nat, nat.astype(np.float64) / 255.0, order, cam, row0, row1)
    active = int(np.count_nonzero(vr.shell))
    return OracleRun(vr, cnt, h, nl, fl, order, i32, i64, st, i64s, active)


def r8_mip_bound(scene) -> float:
    """Largest change the R8 rounding of the mip levels (each texel rounded to
    the nearest 1/255, level 0 exact) can make to a pixel when the fragment
    alpha does not depend on the cone (noise march on: alpha is the march's
    opacity).  A sample that reads a level above 0 is a convex combination of
    texels each off by <= 0.5/255, so the weighted cone sum moves by at most
    0.5/255 * sum of i / (steps * downScaling) over those steps; the noise
    colour col <= minNoiseColor + 2 * noiseColorScale scales it and the blends
    are convex combinations of fragment colours."""
    h, r, lvl, fr = O.step_table(O.cone_of(scene.cone), scene.levels, scene.N)
    c = scene.cone
    w = sum((i + 1) / (c.steps * c.down_scaling) for i in range(c.steps) if lvl[i] > 0 or fr[i] > 0)
    nm = scene.noise_march
    col = nm.min_color + 2.0 * nm.color_scale
    return 0.5 / 255.0 * w * col


def glm_perspective(fovy, aspect, znear, zfar):
    """glm::perspective (right-handed, clip z in [-1, 1]) in float32, column-major."""
    f32 = np.float32
    t = f32(np.tan(f32(fovy) / f32(2)))
    P = np.zeros(16, np.float32)
    P[0] = f32(1) / (f32(aspect) * t)
    P[5] = f32(1) / t
    P[10] = -(f32(zfar) + f32(znear)) / (f32(zfar) - f32(znear))
    P[11] = -1.0
    P[14] = -(f32(2) * f32(zfar) * f32(znear)) / (f32(zfar) - f32(znear))
    return P


def glm_look_at(eye, target, up):
    """glm::lookAt (right-handed) in float32, column-major."""
    e, c, u = (np.asarray(v, np.float32) for v in (eye, target, up))
    f = c - e
    f = f / np.float32(np.sqrt(np.dot(f, f)))
    sv = np.cross(f, u).astype(np.float32)
    sv = sv / np.float32(np.sqrt(np.dot(sv, sv)))
    uv = np.cross(sv, f).astype(np.float32)
    M = np.zeros(16, np.float32)
    M[0], M[4], M[8] = sv
    M[1], M[5], M[9] = uv
    M[2], M[6], M[10] = -f
    M[12], M[13], M[14], M[15] = -np.dot(sv, e), -np.dot(uv, e), np.dot(f, e), 1.0
    return M


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
