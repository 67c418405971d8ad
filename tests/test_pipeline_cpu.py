"""Whole-frame properties of the oracle and the scene generator (CPU only)."""
import math
from dataclasses import replace

import numpy as np

import oracle_lib as O


def _frame(scene, **cone):
    sc = replace(scene, cone=replace(scene.cone, **cone)) if cone else scene
    vr = O.voxelize(sc)
    nat, f64, _, _ = O.build_levels(sc, vr.shell, vr.fill)
    cam = sc.cameras[0]
    order = O.sort_boards(sc, cam.eye)
    return O.cone_trace(sc, nat, f64, order, cam)


def test_back_to_front_blend_equals_front_to_back_under(cv):
    # a17 (GL SRC_ALPHA blend, far to near) and N3 (front-to-back under, near
    # to far) agree on RGB in exact arithmetic when nothing terminates early
    # (alpha_stop also ends the cone march, so both runs use the same value)
    sc = replace(cv.scenes.small32(1), W=96, H=96)
    ftb, _, _ = _frame(sc, composite=1, alpha_stop=2.0)
    gl, _, _ = _frame(sc, composite=0, quantize_rgba8=0, clear_color=(0.0, 0.0, 0.0, 0.0), alpha_stop=2.0)
    assert np.abs(ftb[..., :3] - gl[..., :3]).max() < 2e-6
    assert ftb[..., 3].max() > 0.5


def test_fp32_within_tolerance_of_fp64_replay(cv):
    for sc in (cv.scenes.small32(0), cv.scenes.small32(4)):
        i32, i64, st = _frame(sc)
        assert st.fragments > 1000
        assert np.abs(i32 - i64).max() <= 1e-3


def test_early_out_matches_closed_form_saturation(cv):
    # with alpha_stop the composite never exceeds 1 and stops within one fragment of it
    sc = cv.scenes.small32(2)
    i32, _, _ = _frame(sc, cone_mode=1)
    a = i32[..., 3]
    assert a.max() <= 1.0 + 1e-6
    assert np.all((a == 0) | (a > 0.01))


def test_reference_mode_weighted_sum_range(cv):
    sc = cv.scenes.reference_mode(cv.scenes.small32(0))
    i32, i64, st = _frame(sc)
    # GL blend over the clear colour: untouched pixels keep the RGBA8-quantised clear colour
    clear = np.floor(np.clip(np.array(sc.cone.clear_color, np.float32), 0, 1) * np.float32(255) + np.float32(0.5))
    corner = i32[0, 0]
    assert np.array_equal(corner, (clear / np.float32(255)).astype(np.float32))
    assert st.samples == st.fragments * sc.cone.steps      # no exit test in the reference cone


def test_scene_generators_follow_the_workload_definitions(cv):
    s = cv.scenes
    a = s.small32(3)
    assert a.S == 64 and a.N == 32 and (a.W, a.H) == (256, 256)
    assert np.all(np.abs(a.xyz) <= 0.6) and np.all((a.scales >= 0.08) & (a.scales <= 0.2))
    assert np.allclose(a.sun_dir, np.array([1, 1, 0.3]) / np.linalg.norm([1, 1, 0.3]), atol=1e-6)
    b = s.plain256(7)
    assert b.S == 4096 and b.N == 256 and (b.W, b.H) == (1920, 1080) and len(b.cameras) == 16
    assert np.all((b.scales >= 0.01) & (b.scales <= 0.2))
    assert abs(np.median(b.scales) - 0.05) < 0.005
    assert b.sun_dir[1] >= 0 and abs(np.linalg.norm(b.sun_dir) - 1) < 1e-6
    for c in b.cameras:
        assert abs(math.hypot(c.eye[0], c.eye[2]) - 3.0) < 1e-6 and c.eye[1] == 0
    assert b.cone.steps == 256 and abs(b.cone.cone_angle - math.radians(20)) < 1e-7
    c2 = s.plain256(7)
    assert np.array_equal(b.xyz, c2.xyz) and np.array_equal(b.scales, c2.scales)      # seeded
    f = s.fp16_512(0)
    assert f.S == 16384 and f.N == 512 and f.fmt == 2 and f.noise.octaves == 4
    sc = s.batched64(0)
    assert len(sc) == 64
    assert {x.N for x in sc} <= {64, 96, 128, 192, 256} and {x.W for x in sc} <= {256, 512, 768, 1024}
    assert all(256 <= x.S <= 8192 for x in sc)
