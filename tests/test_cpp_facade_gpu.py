"""The C++ facade (include/cloudvox.hpp) drives the reference application's
default frame through libcloudvox.so the way src/main.cpp:69-124 drives the
GL passes (tests/cpp/reference_frame.cpp), and the frame equals the committed
golden fixture of the oracle bit for bit: activeVoxels, the occupancy count
and hash, and every byte of the RGBA8 image."""
import importlib.util
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REPO = Path(__file__).resolve().parent.parent


def _build_mod():
    spec = importlib.util.spec_from_file_location("cr_build", REPO / "cloud-renderer_amd" / "build.py")
    b = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(b)
    return b


def test_cpp_facade_reference_frame(cv, gpu_ctx, tmp_path):
    from gpu_common import log
    b = _build_mod()
    exe = b.CPP_FRAME_BIN
    if not exe.exists():
        exe = tmp_path / "reference_frame"
        p = subprocess.run(b.cpp_frame_command(exe), capture_output=True, text=True)
        assert p.returncode == 0, p.stderr
    with np.load(REPO / "tests" / "golden" / "reference_default_320.npz", allow_pickle=False) as z:
        want = {k: z[k] for k in z.files}
    xyz, scales = want["xyz"], want["scales"]
    bb = tmp_path / "billboards.bin"
    with open(bb, "wb") as f:
        f.write(np.uint32(scales.shape[0]).tobytes())
        f.write(np.ascontiguousarray(xyz, np.float32).tobytes())
        f.write(np.ascontiguousarray(scales, np.float32).tobytes())
    out = tmp_path / "frame.bin"
    W, H = 320, 180
    log(f"[cpp facade] reference default frame {W}x{H} through cloudvox.hpp")
    sys.stdout.flush()
    p = subprocess.run([str(exe), str(bb), str(W), str(H), "0", str(out)], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    raw = out.read_bytes()
    active, count, h = np.frombuffer(raw[:24], np.uint64)
    image = np.frombuffer(raw[24:], np.uint8).reshape(H, W, 4)
    assert int(active) == int(want["active"])
    assert (int(count), int(h)) == (int(want["count"]), int(want["hash"]))
    want8 = np.rint(want["image32"].astype(np.float64) * 255.0).astype(np.uint8)
    assert np.array_equal(want8.astype(np.float32) / np.float32(255), want["image32"])   # fixture is RGBA8 content
    diff = np.argwhere(image != want8)
    assert diff.size == 0, f"{len(diff)} bytes differ, first at {tuple(diff[0])}"
    log(f"[cpp facade] ok: activeVoxels {int(active)}, image bytes equal")
