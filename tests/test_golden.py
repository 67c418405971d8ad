"""Committed golden fixtures (tests/golden/*.npz, made by make_golden.py):
the oracle reproduces them bit for bit on the CPU, and the HIP path reproduces
them bit for bit on the GPU (without running the oracle there)."""
import hashlib
import sys
from pathlib import Path

import numpy as np
import pytest

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE / "golden"))

from make_golden import compute, fixture_scene  # noqa: E402

NAMES = ["small32_64", "small32_ref_64", "small32_fp16_64", "reference_default_320"]


def load(name):
    with np.load(HERE / "golden" / f"{name}.npz", allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_golden(name):
    want = load(name)
    got = compute(fixture_scene(name, want))
    for k in ("shell_bits", "fill_bits", "count", "hash", "active", "level_sha256", "order"):
        assert np.array_equal(got[k], want[k]), k
    assert np.array_equal(got["image32"].view(np.uint32), want["image32"].view(np.uint32))
    assert np.abs(got["image64"] - want["image64"]).max() == 0


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_reproduces_golden(cv, gpu_ctx, name):
    from gpu_common import log, run_gpu
    want = load(name)
    scene = fixture_scene(name, want)
    log(f"[golden] {name} on the GPU")
    g = run_gpu(cv, scene, ctx=gpu_ctx)
    N = scene.N
    shell = np.unpackbits(want["shell_bits"])[: N ** 3].reshape(N, N, N)
    fill = np.unpackbits(want["fill_bits"])[: N ** 3].reshape(N, N, N)
    state = np.where(shell != 0, 2, np.where(fill != 0, 1, 0)).astype(np.uint8)
    assert np.array_equal(g.state, state)
    assert (g.count, g.hash) == (int(want["count"]), int(want["hash"]))
    assert g.active == int(want["active"])
    digests = [hashlib.sha256(np.ascontiguousarray(x).tobytes()).hexdigest() for x in g.levels]
    assert digests == list(want["level_sha256"])
    assert np.array_equal(g.order, want["order"])
    assert np.array_equal(g.image.view(np.uint32), want["image32"].view(np.uint32))
    if not (scene.cone.quantize_rgba8 and scene.cone.composite == 0):   # RGBA8 rounding is in no FP64 replay
        assert np.abs(g.image - want["image64"]).max() <= (2e-3 if scene.fmt == 2 else 1e-3)
