"""GPU parity: the HIP path against the CPU oracle on seeded scenes.

Bars (stated per check):
  * occupancy (lit shell / chord fill state per voxel), its count and hash,
    activeVoxels (level-0 texels with r != 0), the pass-1 position map, every
    mip level and the draw order: bit-exact;
  * the FP32 image: bit-exact against the oracle's FP32 pass (same op order);
  * the FP32 image against the oracle's FP64 replay: max-abs <= 1e-3 on
    every channel (FP32 storage), <= 2e-3 for RGBA16F storage, whose
    rounding the FP64 replay does not share (PSNR is reported as well);
  * R8 volumes (reference semantics): the FP64 replay runs over the R8
    levels as stored (byte / 255), so 1e-3 judges the arithmetic; the replay
    over the unrounded mip means is reported beside it and, where the
    fragment alpha does not depend on the cone (noise march on), bounded by
    what the R8 rounding of the mips can change (r8_mip_bound);
  * GL blending into an RGBA8 framebuffer rounds after every blend, which no
    FP64 replay shares: those images are bit-exact against the FP32 pass and
    their FP64 accuracy is judged on a float-blend twin case.
Every reference-semantics case listed as informative keeps its weighted
cone sums below 1 (no pixel saturates), so a wrong sample, weight or light
direction changes the image.
"""
from dataclasses import replace

import numpy as np
import pytest

import oracle_lib as O
from gpu_common import assert_bit_equal, glm_look_at, glm_perspective, log, psnr, r8_mip_bound, run_gpu, run_oracle

pytestmark = pytest.mark.gpu

TOL_F32 = 1e-3
TOL_F16 = 2e-3


def _scenes(cv):
    sc = cv.scenes
    base = sc.small32(0)
    cases = {
        "small32": base,
        "small32_seed3": sc.small32(3),
        "small32_ftb": replace(base, cone=replace(base.cone, cone_mode=1)),
        "small32_ref": sc.reference_mode(base),
        "small32_ref_rgba32f": replace(sc.reference_mode(base), fmt=1),
        "small32_ref_float_blend": replace(sc.reference_mode(base),
                                           cone=replace(sc.reference_mode(base).cone, quantize_rgba8=0)),
        "small32_fp16_noise": replace(base, fmt=2, noise=sc.Noise(seed=7), cone=replace(base.cone, compensated=1)),
        # cells of 32/3, 32/7.5, 32/18.75 voxels: boundaries cut the 2x2x2 blocks
        "small32_fp32_noise_odd": replace(base, fmt=1, noise=sc.Noise(octaves=3, seed=9, base_freq=3.0, lacunarity=2.5)),
        # cells of 32 and 16 voxels: every lane of a wave in one cell (lattice
        # hash on scalar registers), then cells of 8 (one cell per lane)
        "small32_fp16_noise_wide": replace(base, fmt=2, noise=sc.Noise(octaves=3, seed=12, base_freq=1.0),
                                           cone=replace(base.cone, compensated=1)),
        # soft-particle noise march in front of the cone (reference defaults;
        # then wind offsets, a fifth octave and 1..12 steps on the rgba path)
        "small32_ref_noise_march": replace(sc.reference_mode(base), noise_march=sc.NoiseMarch(seed=3)),
        "small32_ftb_noise_march": replace(base, cone=replace(base.cone, cone_mode=1),
                                           noise_march=sc.NoiseMarch(seed=9, octaves=5, min_steps=1, max_steps=12,
                                                                     offsets=(0.013, -0.02, 0.005), adjust_size=2.0,
                                                                     step_size=0.02)),
        # sun sprite before the clouds: sun behind the volume, in view; the
        # default radii, then a wide ring over float blending with the march
        "small32_ref_sun": replace(sc.reference_mode(base), sun_dir=_unit3((0.15, 0.25, -1.0)),
                                   sun_sprite=sc.SunSprite()),
        "small32_blend_sun_march": replace(base, sun_dir=_unit3((-0.2, 0.1, -1.0)),
                                           cone=replace(base.cone, composite=0, quantize_rgba8=0),
                                           sun_sprite=sc.SunSprite(inner_color=(0.9, 0.6, 0.3),
                                                                   outer_color=(0.1, 0.2, 1.0), inner_radius=0.7,
                                                                   outer_radius=3.5),
                                           noise_march=sc.NoiseMarch(seed=4)),
        "small32_comp_blend": replace(base, cone=replace(base.cone, composite=0)),
        "small32_rect": replace(base, W=200, H=120, light_w=130, light_h=210),
        "npot24": replace(base, N=24, levels=5),
        "npot96_plain": replace(sc.scaled(sc.plain256(1), 320, 180), N=96, levels=7),
        # spheres whose chord slabs overflow the raster's LDS bitmap at 128^3:
        # pass 1 cuts their tiles into depth slices (light_slices K = 2..4)
        "sliced128": _sliced(sc, base),
        # light along a volume axis: chords advance in one coordinate only
        # (pass 1's brick words change along one axis, the slab boxes are thin)
        "small32_sun_x": replace(base, sun_dir=_unit3((1.0, 0.0, 0.0))),
        "small32_sun_negz": replace(base, sun_dir=_unit3((0.0, 0.0, -1.0))),
        # the reference's lit-shell semantics stored as RGBA16F
        "small32_ref_fp16": replace(sc.reference_mode(base), fmt=2),
        # edge cases: no billboards, one, all outside the volume, an image
        # and light map of a few pixels (tiles cut by every edge)
        "empty": replace(base, xyz=np.zeros((0, 3), np.float32), scales=np.zeros(0, np.float32)),
        "single": replace(base, xyz=base.xyz[:1].copy(), scales=base.scales[:1].copy()),
        "outside": replace(base, xyz=(base.xyz + np.asarray([1.9, 0.0, 0.0], np.float32)).astype(np.float32)),
        "tiny_img": replace(base, W=7, H=5, light_w=9, light_h=3),
        "ties": tie_scene(sc),
        "ties_ref": sc.reference_mode(tie_scene(sc)),
        # reference semantics in a range that carries information: with
        # vctDownScaling 16 no weighted sum reaches 1 (float blend over a
        # zero clear; then the RGBA8 framebuffer over the reference clear)
        "small32_ref_ds16": _ref_ds(sc, base, 16.0, quantize=0, clear=(0.0, 0.0, 0.0, 0.0)),
        "small32_ref_ds16_q8": _ref_ds(sc, base, 16.0, quantize=1),
        # off-origin volume with bounds of its own, fluffiness != 1, camera
        # and point sun placed around it
        "small32_ref_offset_fluffy": _offset_fluffy(sc, base),
        # the reference application's default frame (src/main.cpp defaults)
        "refdef": sc.reference_default(0),
        "refdef_float": sc.reference_default(0, quantize_rgba8=0),
        "refdef_fluffy": sc.reference_default(1, fluffiness=0.6, run_time=2.5, quantize_rgba8=0),
        "refdef_nonoise_ds8": _refdef_nonoise(sc),
        # the same frame with P and V handed over as matrices
        "refdef_matrices": _with_matrices(sc.reference_default(2, quantize_rgba8=0)),
    }
    return cases


def _sliced(sc, base):
    rng = np.random.default_rng(11)
    xyz = rng.uniform(-0.35, 0.35, size=(6, 3)).astype(np.float32)
    r = rng.uniform(0.35, 0.6, size=6).astype(np.float32)
    return replace(base, name="sliced128", N=128, levels=8, xyz=xyz, scales=r, sun_dir=_unit3((1.0, 0.8, 0.5)),
                   W=256, H=256, light_w=256, light_h=256)


def _ref_ds(sc, base, ds, quantize, clear=None):
    r = sc.reference_mode(base)
    cone = replace(r.cone, down_scaling=ds, quantize_rgba8=quantize)
    if clear is not None:
        cone = replace(cone, clear_color=clear)
    return replace(r, cone=cone)


def _offset_fluffy(sc, base):
    r = _ref_ds(sc, base, 16.0, quantize=0)
    pos = (10.0, -3.0, 2.0)
    cam = sc.Camera(eye=(10.6, -2.7, 5.5), target=pos)
    return replace(r, position=pos, bounds_lo=(-1.2, -0.9, -1.1), bounds_hi=(1.0, 1.3, 0.9), fluffiness=0.8,
                   cameras=[cam], sun_position=(13.0, 6.0, -1.0))


def _refdef_nonoise(sc):
    r = sc.reference_default(0, noise_march=False, quantize_rgba8=0)
    return replace(r, cone=replace(r.cone, down_scaling=8.0))


def _with_matrices(scene):
    c = scene.cameras[0]
    P = glm_perspective(c.fovy, c.aspect or scene.W / scene.H, c.znear, c.zfar)
    V = glm_look_at(c.eye, c.target, c.up)
    return replace(scene, cameras=[replace(c, P=P, V=V)])


def _unit3(v):
    v = np.asarray(v, np.float64)
    return (v / np.linalg.norm(v)).astype(np.float32)


def tie_scene(sc):
    """64 billboards in groups of four related by 90-degree turns about the
    camera axis (plus exact duplicates): their FP32 camera distances are equal
    bit for bit, which exercises the selection-sort tie replay."""
    base = sc.small32(11)
    rng = np.random.default_rng(11)
    pts, rad = [], []
    for g in range(14):
        x, y, z = rng.uniform(-0.5, 0.5, 3).astype(np.float32)
        r = np.float32(rng.uniform(0.08, 0.2))
        for (a, b) in ((x, y), (-y, x), (-x, -y), (y, -x)):
            pts.append((a, b, z))
            rad.append(r)
    for k in range(8):  # exact duplicates
        pts.append(pts[k * 3])
        rad.append(rad[k * 3])
    perm = rng.permutation(len(pts))
    xyz = np.asarray(pts, np.float32)[perm]
    r = np.asarray(rad, np.float32)[perm]
    return replace(base, xyz=xyz, scales=r)


CASES = ["small32", "small32_seed3", "small32_ftb", "small32_ref", "small32_ref_rgba32f", "small32_ref_float_blend",
         "small32_fp16_noise", "small32_fp32_noise_odd", "small32_fp16_noise_wide", "small32_ref_noise_march",
         "small32_ftb_noise_march", "small32_ref_sun", "small32_blend_sun_march", "small32_comp_blend",
         "small32_rect", "npot24", "npot96_plain", "sliced128", "small32_sun_x", "small32_sun_negz",
         "small32_ref_fp16", "empty", "single", "outside", "tiny_img", "ties", "ties_ref", "small32_ref_ds16",
         "small32_ref_ds16_q8",
         "small32_ref_offset_fluffy", "refdef", "refdef_float", "refdef_fluffy", "refdef_nonoise_ds8",
         "refdef_matrices"]
# reference-semantics cases whose pixels must stay below saturation
INFORMATIVE = {"small32_ref_ds16", "small32_ref_ds16_q8", "small32_ref_offset_fluffy", "refdef", "refdef_float",
               "refdef_fluffy", "refdef_nonoise_ds8", "refdef_matrices"}


@pytest.mark.parametrize("case", CASES)
def test_parity(cv, gpu_ctx, case):
    scene = _scenes(cv)[case]
    log(f"[parity] case={case} N={scene.N} S={scene.S} img={scene.W}x{scene.H} fmt={scene.fmt} mode={scene.mode}")
    g = run_gpu(cv, scene, ctx=gpu_ctx)
    o = run_oracle(scene)
    # occupancy: bit-exact
    assert_bit_equal("occupancy state", g.state, o.vox.state())
    assert (g.count, g.hash) == (o.count, o.hash), "occupancy count/hash"
    assert_bit_equal("position map", g.posmap, o.vox.posmap)
    for k, (gl, ol) in enumerate(zip(g.levels, o.levels)):
        assert_bit_equal(f"level {k}", gl, ol)
    assert_bit_equal("draw order", g.order, o.order)
    assert g.active == o.active, "activeVoxels"
    # image: FP32 bit-exact against the oracle's FP32 pass
    assert_bit_equal("image fp32", g.image, o.img32)
    check_fp64(case, scene, g, o)


def check_fp64(case, scene, g, o):
    """The FP32 image against the oracle's FP64 replay (bars in the module docstring)."""
    img = g.image.astype(np.float64)
    err = float(np.abs(img - o.img64).max())
    quantized = scene.cone.quantize_rgba8 and scene.cone.composite == 0
    msg = (f"[parity] case={case} max_abs_vs_fp64={err:.3e} psnr={psnr(g.image, o.img64):.1f} dB "
           f"fragments={g.stats['fragments']} samples={g.stats['samples']}")
    if scene.fmt == 0:
        err_s = float(np.abs(img - o.img64_stored).max())
        msg += f" max_abs_vs_fp64_stored_r8={err_s:.3e}"
        if scene.noise_march is not None and not quantized:
            bound = r8_mip_bound(scene)
            msg += f" r8_mip_bound={bound:.3e}"
            assert err <= bound + TOL_F32, f"max-abs {err} vs the unrounded FP64 replay exceeds {bound} + 1e-3"
        if not quantized:
            assert err_s <= TOL_F32, f"max-abs {err_s} vs the FP64 replay over the stored R8 levels exceeds 1e-3"
    elif not quantized:
        tol = TOL_F16 if scene.fmt == 2 else TOL_F32
        assert err <= tol, f"max-abs {err} vs FP64 replay exceeds {tol}"
    if case in INFORMATIVE:
        clear = np.asarray(scene.cone.clear_color, np.float32)
        if quantized:  # the RGBA8 framebuffer holds the clear colour rounded to 8 bits
            clear = (np.floor(np.clip(clear, 0, 1) * np.float32(255) + np.float32(0.5)) / np.float32(255)).astype(
                np.float32)
        covered = np.any(g.image != clear, axis=-1)
        sat = int(np.count_nonzero((g.image[..., 0] >= 1.0) & covered))
        msg += f" saturated_px={sat}/{int(np.count_nonzero(covered))}"
        assert np.count_nonzero(covered) > 100, "the case covers too few pixels to carry information"
        assert sat <= 0.05 * np.count_nonzero(covered), "more than 5% of the covered pixels saturate"
    log(msg)


def test_determinism(cv, gpu_ctx):
    scene = cv.scenes.small32(5)
    log("[determinism] small32 seed 5, two runs")
    a = run_gpu(cv, scene, ctx=gpu_ctx, read_levels=True)
    b = run_gpu(cv, scene, ctx=gpu_ctx, read_levels=True)
    assert_bit_equal("state", a.state, b.state)
    assert_bit_equal("image", a.image, b.image)
    for k in range(len(a.levels)):
        assert_bit_equal(f"level {k}", a.levels[k], b.levels[k])


def test_rgba8_output(cv, gpu_ctx):
    scene = cv.scenes.small32(2)
    log("[rgba8] small32 seed 2 RGBA8 output")
    g8 = run_gpu(cv, scene, ctx=gpu_ctx, read_levels=False, pixel_format=1)
    g32 = run_gpu(cv, scene, ctx=gpu_ctx, read_levels=False, pixel_format=0)
    want = np.floor(np.clip(g32.image, 0, 1) * np.float32(255) + np.float32(0.5)).astype(np.uint8)
    assert_bit_equal("rgba8", g8.image, want)


def test_volume_side_limit(cv, gpu_ctx):
    # texel indices inside a level are 32-bit: sides above 1624 are refused
    # before anything is allocated
    log("[limits] cr_volume_create side 1625 and 0")
    for dim in (1625, 0):
        with pytest.raises(cv.CloudVoxError) as e:
            cv.CloudVolume(dim, mips=1, context=gpu_ctx)
        assert e.value.code == cv._abi.CR_EINVAL and "dim must be in [1, 1624]" in str(e.value)


def test_pipelined_frames_match_serial(cv, gpu_ctx):
    """Frames queued back to back on one stream (the voxelization of frame
    i + 1 overlaps the cone trace of frame i on the volume's own streams) give
    the images of frames run one at a time.  The sun moves every frame, so a
    pass that reads or writes the wrong frame's bitmaps or levels changes the
    picture."""
    import math
    import torch

    scene = cv.scenes.plain256(0)
    F = 4
    log(f"[pipeline] plain256 seed 0, {F} frames with a moving sun")
    vol, sun0, _, vox, tr = cv.scene_objects(scene, 0, context=gpu_ctx)
    try:
        suns, cams = [], []
        for i in range(F):
            a = 0.7 * i
            p = sun0.position
            pos = (p[0] * math.cos(a) - p[2] * math.sin(a), p[1], p[0] * math.sin(a) + p[2] * math.cos(a))
            n = math.sqrt(sum(x * x for x in pos))
            suns.append(cv.Sun(position=pos, direction=tuple(x / n for x in pos), directional=sun0.directional))
            c = scene.cameras[(3 * i) % len(scene.cameras)]
            cams.append(cv.Camera(eye=c.eye, target=c.target, up=c.up, fovy=c.fovy, znear=c.znear, zfar=c.zfar))
        shape = (scene.H, scene.W, 4)
        serial = []
        for i in range(F):
            img = torch.zeros(shape, dtype=torch.float32, device="cuda")
            vox.voxelize(vol, suns[i])
            tr.coneTrace(vol, suns[i], cams[i], img, scene.W, scene.H)
            torch.cuda.synchronize()
            serial.append(img.cpu().numpy())
        # the volume state after the last serial frame
        want_state = vol.read_occupancy()
        want_count = vol.occupancy()
        want_levels = [vol.read_level(k) for k in range(scene.levels)]
        want_pm = vol.read_position_map(scene.light_w, scene.light_h)
        want_order = vol.read_draw_order()
        stream = torch.cuda.current_stream()
        imgs = [torch.zeros(shape, dtype=torch.float32, device="cuda") for _ in range(F)]
        for rep in range(2):
            for i in range(F):
                vox.voxelize(vol, suns[i], stream=stream)
                tr.coneTrace(vol, suns[i], cams[i], imgs[i], scene.W, scene.H, stream=stream)
            torch.cuda.synchronize()
            for i in range(F):
                assert_bit_equal(f"pipelined frame {i} (pass {rep})", imgs[i].cpu().numpy(), serial[i])
            # readbacks see the last queued frame
            assert_bit_equal(f"occupancy state (pass {rep})", vol.read_occupancy(), want_state)
            assert vol.occupancy() == want_count
            for k in range(scene.levels):
                assert_bit_equal(f"level {k} (pass {rep})", vol.read_level(k), want_levels[k])
            assert_bit_equal(f"position map (pass {rep})", vol.read_position_map(scene.light_w, scene.light_h),
                             want_pm)
            assert np.array_equal(vol.read_draw_order(), want_order)
        # a counter run (serial, on the caller's stream) between queued frames
        for img in imgs:
            img.zero_()
        for i in range(F):
            if i == 2:
                gpu_ctx.set_profiling(1)
            try:
                vox.voxelize(vol, suns[i], stream=stream)
                tr.coneTrace(vol, suns[i], cams[i], imgs[i], scene.W, scene.H, stream=stream)
            finally:
                if i == 2:
                    gpu_ctx.set_profiling(0)
        torch.cuda.synchronize()
        for i in range(F):
            assert_bit_equal(f"frame {i} around a counter run", imgs[i].cpu().numpy(), serial[i])
        assert len({serial[i].tobytes() for i in range(F)}) == F, "the frames differ"
    finally:
        vol.close()


def test_pipelined_frames_with_growing_buffers(cv, gpu_ctx):
    """Queued frames whose light map and image sizes grow from frame to frame
    (the depth keys, the work list and the camera buffers are reallocated
    while earlier frames may still be queued) give the images of frames run
    one at a time."""
    import torch

    scene = cv.scenes.small32(4)
    sizes = [(48, 40, 64, 48), (96, 80, 128, 96), (64, 64, 200, 150), (160, 120, 256, 200)]
    log(f"[pipeline] small32 seed 4, {len(sizes)} frames with growing light maps and images")

    def run(queued):
        vol, sun, cam, _, tr = cv.scene_objects(scene, 0, context=gpu_ctx)
        try:
            out = []
            for (lw, lh, w, h) in sizes:
                vox = cv.VoxelizeShader(lw, lh)
                img = torch.zeros((h, w, 4), dtype=torch.float32, device="cuda")
                vox.voxelize(vol, sun)
                tr.coneTrace(vol, sun, cam, img, w, h)
                if not queued:
                    torch.cuda.synchronize()
                out.append(img)
            torch.cuda.synchronize()
            return [o.cpu().numpy() for o in out]
        finally:
            vol.close()

    want = run(False)
    got = run(True)
    for i in range(len(sizes)):
        assert_bit_equal(f"frame {i} ({sizes[i]})", got[i], want[i])


def test_noise_march_map_and_limits(cv, gpu_ctx):
    """The runtime's noise map equals the oracle's byte for byte; the learned
    trace and the batch refuse a volume with the march on; switching it off
    gives the plain image back."""
    sc = cv.scenes
    base = sc.small32(6)
    log("[noise march] map bytes seed 21, refusals, switch off")
    scene = replace(base, noise_march=sc.NoiseMarch(seed=21))
    vol, sun, cam, vox, tr = cv.scene_objects(scene, 0, context=gpu_ctx)
    try:
        vol.set_noise_march(tr.noise_params())
        assert_bit_equal("noise map", vol.read_noise_map(), O.noise_map(21))
        vox.voxelize(vol, sun)
        import torch
        img = torch.zeros((scene.H, scene.W, 4), dtype=torch.float32, device="cuda")
        with pytest.raises(cv.CloudVoxError) as e:
            tr.coneTraceLearned(vol, sun, cam, cv.FarFieldMLP(context=gpu_ctx), img, scene.W, scene.H, k=8)
        assert e.value.code == cv._abi.CR_EUNSUPPORTED
        tr.coneTrace(vol, sun, cam, img, scene.W, scene.H)
        torch.cuda.synchronize()
        with_nm = img.cpu().numpy()
        tr.doNoiseSample = False
        tr.coneTrace(vol, sun, cam, img, scene.W, scene.H)
        torch.cuda.synchronize()
        plain = img.cpu().numpy()
    finally:
        vol.close()
    want = run_gpu(cv, base, ctx=gpu_ctx, read_levels=False).image
    assert_bit_equal("noise march off", plain, want)
    assert not np.array_equal(with_nm, plain)


def test_runtime_refusals_camera_and_noise_steps(cv, gpu_ctx):
    """The runtime refuses what the oracle refuses: min = max = 1 noise steps
    (iSteps - 1 = 0 divides), a projection that is not a symmetric
    perspective, and a degenerate lookAt; fovy = 45 (radians, as the
    reference passes it) is accepted."""
    import torch
    sc = cv.scenes
    scene = replace(sc.reference_default(0, W=64, H=36), noise_march=sc.NoiseMarch(min_steps=1, max_steps=1))
    log("[refusals] noise steps 1..1, bad P, degenerate lookAt")
    vol, sun, cam, vox, tr = cv.scene_objects(scene, 0, context=gpu_ctx)
    try:
        vox.voxelize(vol, sun)
        img = torch.zeros((scene.H, scene.W, 4), dtype=torch.float32, device="cuda")
        with pytest.raises(cv.CloudVoxError) as e:
            tr.coneTrace(vol, sun, cam, img, scene.W, scene.H)
        assert e.value.code == cv._abi.CR_EINVAL and "max_steps >= 2" in str(e.value)
        tr.doNoiseSample = False
        tr.coneTrace(vol, sun, cam, img, scene.W, scene.H)      # fovy 45, aspect 1: accepted
        P = glm_perspective(45.0, 1.0, 0.01, 2500.0)
        V = glm_look_at(cam.eye, cam.target, cam.up)
        Q = P.copy()
        Q[8] = 0.25
        bad = cv.Camera(eye=cam.eye, P=Q, V=V)
        with pytest.raises(cv.CloudVoxError) as e:
            tr.coneTrace(vol, sun, bad, img, scene.W, scene.H)
        assert e.value.code == cv._abi.CR_EUNSUPPORTED
        with pytest.raises(cv.CloudVoxError) as e:
            tr.coneTrace(vol, sun, cv.Camera(eye=(0, 0, 0), target=(0, 0, 0)), img, scene.W, scene.H)
        assert e.value.code == cv._abi.CR_EINVAL
        torch.cuda.synchronize()
    finally:
        vol.close()


@pytest.mark.parametrize("case", ["small32", "small32_fp16_noise", "npot96_plain", "small32_comp_blend", "plain256",
                                  "fp16_512"])
def test_table_addressing_matches_direct(cv, gpu_ctx, case):
    """The front-to-back cone trace reads its corner offsets from per-level
    LDS tables (4-wave workgroups, 8 at 512^3); a context created with
    CR_CONTEXT_DIRECT_ADDRESSING computes them per sample.  Same images bit
    for bit, at the small sizes and at the 256^3 / 512^3 workloads."""
    sc = cv.scenes
    scenes = {"plain256": sc.plain256(0), "fp16_512": sc.fp16_512(0)}
    scene = scenes[case] if case in scenes else _scenes(cv)[case]
    log(f"[tables] {case}: N={scene.N} {scene.W}x{scene.H} fmt={scene.fmt}, table vs direct addressing")
    direct = cv.Context(0, flags=cv._abi.CR_CONTEXT_DIRECT_ADDRESSING)
    try:
        a = run_gpu(cv, scene, ctx=gpu_ctx, read_levels=False)
        b = run_gpu(cv, scene, ctx=direct, read_levels=False)
    finally:
        direct.close()
    assert_bit_equal("image (tables vs direct)", a.image, b.image)


def test_reference_draw_order_frames(cv, gpu_ctx):
    """The reference keeps its billboard arrays sorted in place
    (src/CloudVolume.cpp:65-82) after the frame's draw call was set up, so
    frame f voxelizes and draws in the order sorted at frame f - 1 and each
    selection sort starts from the previous order.  With CR_ORDER_UPLOAD and
    CloudVolume.sortBoards after each trace, three frames of the tie scene
    under a moving camera equal the oracle run the same way (sort from the
    current arrays, permute, upload): images, occupancy and every
    permutation bit for bit, and the orders differ from fresh sorts."""
    import torch
    sc = cv.scenes
    scene = sc.reference_mode(tie_scene(sc))
    scene = replace(scene, cone=replace(scene.cone, down_scaling=16.0, quantize_rgba8=0))
    eyes = [(0.0, 0.0, 3.0), (0.4, 0.3, 2.9), (-0.5, 0.1, 2.8)]
    log("[draw order] tie scene, 3 frames, upload order + in-place sort")
    vol, sun, _, vox, tr = cv.scene_objects(scene, 0, context=gpu_ctx)
    xyz, scales = scene.xyz.copy(), scene.scales.copy()
    try:
        vol.set_draw_order(cv._abi.CR_ORDER_UPLOAD)
        for f, eye in enumerate(eyes):
            cam = cv.Camera(eye=eye)
            vox.voxelize(vol, sun)
            img = torch.zeros((scene.H, scene.W, 4), dtype=torch.float32, device="cuda")
            tr.coneTrace(vol, sun, cam, img, scene.W, scene.H)
            torch.cuda.synchronize()
            got_img, got_state = img.cpu().numpy(), vol.read_occupancy()
            perm = vol.sortBoards(eye)
            vol.update()
            # the oracle, run the same way on the current arrays
            cur = replace(scene, xyz=xyz, scales=scales, cameras=[sc.Camera(eye=eye)])
            vr = O.voxelize(cur)
            nat, f64, _, _ = O.build_levels(cur, vr.shell, vr.fill)
            ident = np.arange(cur.S, dtype=np.int32)
            i32, _, _ = O.cone_trace(cur, nat, f64, ident, cur.cameras[0])
            want_perm = O.sort_boards(cur, eye)
            assert_bit_equal(f"frame {f} occupancy", got_state, vr.state())
            assert_bit_equal(f"frame {f} image", got_img, i32)
            assert np.array_equal(perm, want_perm), f"frame {f} permutation"
            if f > 0:
                fresh = O.sort_boards(replace(cur, xyz=scene.xyz, scales=scene.scales), eye)
                log(f"[draw order] frame {f}: permuted arrays vs a fresh sort from upload order: "
                    f"{int(np.count_nonzero(scene.xyz[fresh][:, 0] != xyz[want_perm][:, 0]))} positions differ")
            xyz, scales = xyz[want_perm], scales[want_perm]
            assert np.array_equal(vol.billboards.positions, xyz)
    finally:
        vol.close()
