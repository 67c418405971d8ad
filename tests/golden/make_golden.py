"""Regenerate the golden fixtures of the small workload from the CPU oracle.

    python tests/golden/make_golden.py

Each fixture (.npz) holds, for one seeded scene: its billboards (xyz,
scales), the level-0 occupancy state (packed bits of shell and fill), its
count and hash, activeVoxels (level-0 texels with r != 0), SHA-256 digests of
every mip level in native format, the draw order, the FP32 image and the
FP64-replay image (stored as FP32; for R8 volumes the replay over the stored
R8 levels).  The small workload's fixtures are rendered at 64 x 64; the
reference application's default frame (32^3 R8 volume at (25, 0, 0), 200
billboards, point sun, noise march and sun sprite on, RGBA8 framebuffer) at
320 x 180, which keeps its integer aspect of 1 (320 / 180).
"""
import hashlib
import sys
from dataclasses import replace
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))

import conftest  # noqa: E402,F401  (loads the package)
import oracle_lib as O  # noqa: E402

import cloud_renderer_amd as cv  # noqa: E402


def fixture_scenes():
    base = replace(cv.scenes.small32(0), W=64, H=64)
    return {
        "small32_64": base,
        "small32_ref_64": cv.scenes.reference_mode(base),
        "small32_fp16_64": replace(base, fmt=2, noise=cv.scenes.Noise(seed=3)),
        "reference_default_320": cv.scenes.reference_default(0, W=320, H=180),
    }


def fixture_scene(name, fixture=None):
    """The scene of a fixture, with the billboards it stores (when given)."""
    scene = fixture_scenes()[name]
    if fixture is not None and "xyz" in fixture:
        scene = replace(scene, xyz=fixture["xyz"], scales=fixture["scales"])
    return scene


def compute(scene):
    vr = O.voxelize(scene)
    cnt, h = O.occupancy(scene.N, vr.shell, vr.fill)
    nat, f64, nl, fl = O.build_levels(scene, vr.shell, vr.fill)
    cam = scene.cameras[0]
    order = O.sort_boards(scene, cam.eye)
    if scene.fmt == 0:
        f64 = nat.astype(np.float64) / 255.0    # FP64 replay over the stored R8 levels
    i32, i64, st = O.cone_trace(scene, nat, f64, order, cam)
    digests = np.array([hashlib.sha256(np.ascontiguousarray(x).tobytes()).hexdigest() for x in nl])
    return dict(xyz=np.ascontiguousarray(scene.xyz, np.float32), scales=np.ascontiguousarray(scene.scales, np.float32),
                shell_bits=np.packbits(vr.shell.reshape(-1)), fill_bits=np.packbits(vr.fill.reshape(-1)),
                count=np.uint64(cnt), hash=np.uint64(h), active=np.uint64(np.count_nonzero(vr.shell)),
                level_sha256=digests, order=order, image32=i32, image64=i64.astype(np.float32))


def main():
    for name, scene in fixture_scenes().items():
        d = compute(scene)
        np.savez_compressed(HERE / f"{name}.npz", **d)
        print(name, int(d["count"]), hex(int(d["hash"])))


if __name__ == "__main__":
    main()
