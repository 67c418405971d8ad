"""Closed-form pins of the oracle's camera pass (the cone-trace fragment
setup of res/conetrace_frag.glsl:136-138, 176-191 and the camera of
src/Camera.cpp:59-60 as ConeTraceShader uploads it,
src/Shaders/ConeTraceShader.cpp:64-69).

Each check compares the oracle with geometry computed here in float64 from the
scene's definition, so a reversed billboard normal, a reversed light
direction, a transposed view matrix or a wrong projection scale fails:
  * cone start points lie on the camera-facing hemisphere
    z = c_z + sqrt(r^2 - d^2) at the pixel centres (lookAt + perspective, an
    explicit aspect, and the same camera handed over as P and V matrices);
  * a point light on +x lights a one-sided slab through the weighted cone and
    the same light on -x leaves every pixel exactly 0;
  * the noise march's view ray is the unit vector from the target to the eye
    (V's third row), the projection scales are tan(fovy / 2) * aspect and
    tan(fovy / 2) (the reference's fovy = 45 taken as radians), and the depth
    range recovered from a glm::perspective matrix is the one it was built from.
"""
import math

import numpy as np
import pytest

import oracle_lib as O
from gpu_common import glm_look_at, glm_perspective

F = np.float32


def _affine_z_levels(N, L):
    """RGBA32F levels whose r is the texture coordinate z of the texel centre
    ((z + 1/2) / n at level k of side n): box means of an affine field at
    texel centres are the field itself, so every level and every trilinear
    sample inside the volume returns t_z exactly (up to rounding)."""
    parts = []
    for k in range(L):
        n = max(1, N >> k)
        z = (np.arange(n, dtype=np.float64) + 0.5) / n
        r = np.broadcast_to(z[:, None, None], (n, n, n))
        t = np.zeros((n, n, n, 4), np.float32)
        t[..., 0] = r
        t[..., 3] = 1.0
        parts.append(t.reshape(-1))
    nat = np.concatenate(parts).astype(np.float32)
    return nat, nat.astype(np.float64)


def _scene(cv, cam, W, H, cone, xyz=((0.0, 0.0, 0.0),), r=(0.5,), sun_pos=(10.0, 0.0, 0.0), directional=1,
           sun_dir=(1.0, 0.0, 0.0)):
    return cv.scenes.Scene(name="pin", N=32, levels=6, fmt=1, mode=1, xyz=np.asarray(xyz, F).reshape(-1, 3),
                           scales=np.asarray(r, F), sun_dir=np.asarray(sun_dir, F), W=W, H=H, light_w=8,
                           light_h=8, cameras=[cam], cone=cone, directional=directional, sun_position=sun_pos)


def _trace(scene, nat, f64):
    order = np.arange(scene.S, dtype=np.int32)
    i32, i64, _ = O.cone_trace(scene, nat, f64, order, scene.cameras[0])
    return i32, i64


@pytest.mark.parametrize("form", ["lookat", "aspect", "matrices"])
def test_oracle_camera_hemisphere(cv, form):
    """One sphere of radius 0.5 at the origin, camera on +z looking at it, a
    one-step weighted cone along +x through a volume whose r is t_z: the
    image's alpha is t_z of the cone start point (the x offset of the one
    step leaves z unchanged), which must be the camera-facing hemisphere
    point above the pixel centre's hit with the billboard plane z = 0."""
    sc = cv.scenes
    W, H, r, D = 72, 48, 0.5, 3.0
    fovy = math.radians(50.0)
    aspect = {"lookat": 0.0, "aspect": 1.0, "matrices": 0.8}[form]
    cam = sc.Camera(eye=(0.0, 0.0, D), target=(0.0, 0.0, 0.0), fovy=fovy, aspect=aspect)
    if form == "matrices":
        cam = sc.Camera(eye=(0.0, 0.0, D), P=glm_perspective(fovy, aspect, 0.01, 2500.0),
                        V=glm_look_at((0.0, 0.0, D), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0)))
    cone = sc.Cone(steps=1, cone_angle=0.9, initial_height=0.1, cone_mode=0, composite=1)
    scene = _scene(cv, cam, W, H, cone, r=(r,))
    nat, f64 = _affine_z_levels(32, 6)
    i32, i64 = _trace(scene, nat, f64)
    # closed form, float64
    ax = aspect if aspect > 0 else W / H
    t = math.tan(float(F(fovy)) / 2)
    xs = ((2 * (np.arange(W) + 0.5)) / W - 1) * t * ax * D
    ys = ((2 * (np.arange(H) + 0.5)) / H - 1) * t * D
    X, Y = np.meshgrid(xs, ys)
    q = np.sqrt(X * X + Y * Y) / r
    sc64 = np.sqrt(np.maximum(0.0, 1.0 - q * q))
    covered = sc64 >= 0.01
    margin = np.abs(sc64 - 0.01) < 1e-4
    got = i32[..., 3] > 0
    assert np.array_equal(got[~margin], covered[~margin])
    assert covered.sum() > 200
    zpos = r * sc64                              # + sqrt(r^2 - d^2): the side facing the camera
    want = (zpos + 1.0) / 2.0                    # t_z in a volume over [-1, 1]
    inner = covered & ~margin
    assert np.abs(i64[..., 3][inner] - want[inner]).max() < 1e-6
    assert np.abs(i32[..., 3][inner] - want[inner]).max() < 2e-6
    # the far side (a flipped normal) is off by 2 sqrt(r^2 - d^2) / 2 at the centre
    assert np.abs(i32[..., 3][inner] - (1.0 - zpos[inner]) / 2.0).max() > 0.2


def _slab_levels(N, L, x0):
    """Level 0: r = 1 where the texel's x index is >= x0, else 0; levels above
    0 are zero (so no coarse sample blurs the slab across)."""
    parts = []
    for k in range(L):
        n = max(1, N >> k)
        t = np.zeros((n, n, n, 4), np.float32)
        if k == 0:
            t[:, :, x0:, 0] = 1.0
        parts.append(t.reshape(-1))
    nat = np.concatenate(parts).astype(np.float32)
    return nat, nat.astype(np.float64)


def test_oracle_point_light_direction(cv):
    """The weighted cone toward a point light (dir = lightPos - pos,
    res/conetrace_frag.glsl:190): with the volume occupied only at x >= 0.5
    and a narrow cone, a light at +x gives a positive sum on the sphere's
    right half and a light at -x gives exactly 0 everywhere."""
    sc = cv.scenes
    W = H = 48
    cam = sc.Camera(eye=(0.0, 0.0, 3.0), target=(0.0, 0.0, 0.0), fovy=math.radians(30.0))
    cone = sc.Cone(steps=64, cone_angle=0.2, initial_height=0.1, cone_mode=0, composite=0,
                   clear_color=(0.0, 0.0, 0.0, 0.0))
    nat, f64 = _slab_levels(32, 6, 24)           # x >= 0.5 in [-1, 1]
    imgs = {}
    for side in (1.0, -1.0):
        scene = _scene(cv, cam, W, H, cone, r=(0.3,), sun_pos=(10.0 * side, 0.0, 0.0), directional=0)
        imgs[side], _ = _trace(scene, nat, f64)
    right = imgs[1.0][..., 3]
    left = imgs[-1.0][..., 3]
    assert np.count_nonzero(right > 0) > 50
    assert np.all(left == 0)
    # pixels on the sphere's right half (x > 0.1 world) all see the slab
    t = math.tan(float(F(math.radians(30.0))) / 2)
    xs = ((2 * (np.arange(W) + 0.5)) / W - 1) * t * 3.0
    ys = ((2 * (np.arange(H) + 0.5)) / H - 1) * t * 3.0
    X, Y = np.meshgrid(xs, ys)
    rightish = (X > 0.1) & (np.sqrt(X * X + Y * Y) < 0.29)
    assert rightish.sum() > 20 and np.all(right[rightish] > 0)


def test_oracle_camera_setup_closed_forms(cv):
    sc = cv.scenes
    eye, tgt = np.array([1.3, -0.7, 2.1]), np.array([0.2, 0.4, -0.3])
    cam = sc.Camera(eye=tuple(eye), target=tuple(tgt), fovy=0.9)
    s = O.camera_setup(cam, 200, 120)
    back = (eye - tgt) / np.linalg.norm(eye - tgt)
    # noise view ray = normalize(V[0][2], V[1][2], V[2][2]) = unit vector target -> eye
    assert np.abs(s["ray"] - back).max() < 2e-7
    assert np.abs(s["B"] - back).max() < 2e-7
    assert abs(float(np.dot(s["S"], back))) < 1e-6 and abs(float(np.dot(s["U"], back))) < 1e-6
    assert abs(float(s["S"][1])) < 1e-7            # the right axis is horizontal (up = +y)
    assert np.allclose(np.cross(s["S"], s["U"]), back, atol=1e-6)
    assert abs(s["ky"] - math.tan(float(F(0.9)) / 2)) < 1e-7
    assert abs(s["kx"] - math.tan(float(F(0.9)) / 2) * 200 / 120) < 1e-6
    # the reference: perspective(45.f, (float)(1280 / 720), .01f, 2500.f), fovy taken as radians
    ref = sc.reference_camera(1280, 720)
    r = O.camera_setup(ref, 1280, 720)
    t45 = math.tan(22.5)
    assert ref.aspect == 1.0 and abs(r["kx"] - t45) < 1e-6 and abs(r["ky"] - t45) < 1e-6
    assert abs(t45 - 0.5573) < 1e-3 and abs(2 * math.atan(t45) - 1.0178) < 1e-4   # effective 58.3 degrees
    assert np.abs(r["ray"] - np.array([-1.0, 0.0, 0.0])).max() < 1e-6            # looking down +x
    # the same camera as matrices: equal view, recovered depth range
    P = glm_perspective(45.0, 1.0, 0.01, 2500.0)
    V = glm_look_at(ref.eye, ref.target, ref.up)
    m = O.camera_setup(sc.Camera(eye=ref.eye, P=P, V=V), 1280, 720)
    assert np.array_equal(m["V"], r["V"])
    assert abs(m["kx"] - t45) < 1e-6 and abs(m["ky"] - t45) < 1e-6
    assert abs(m["znear"] - 0.01) < 1e-6 and abs(m["zfar"] - 2500.0) / 2500.0 < 1e-2
    # matrices that are not a symmetric perspective projection are refused
    Q = P.copy()
    Q[11] = 0.0
    assert O.camera_setup(sc.Camera(eye=ref.eye, P=Q, V=V), 1280, 720) is None
    Q = P.copy()
    Q[8] = 0.1                                   # off-centre projection
    assert O.camera_setup(sc.Camera(eye=ref.eye, P=Q, V=V), 1280, 720) is None


def test_oracle_noise_march_step_limits(cv):
    """min = max = 1 would divide by iSteps - 1 = 0 (res/conetrace_frag.glsl:154-158): refused."""
    nm = cv.scenes.NoiseMarch
    assert O.noise_march_valid(nm())
    assert O.noise_march_valid(nm(min_steps=1, max_steps=2))
    assert not O.noise_march_valid(nm(min_steps=1, max_steps=1))
    assert not O.noise_march_valid(nm(min_steps=3, max_steps=2))
