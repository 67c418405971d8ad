"""Seeded synthetic scenes for the benchmark workloads.

The reference builds its billboards with rand() seeded from time(0)
(src/Util.hpp:19-45, src/main.cpp:70, src/CloudVolume.cpp:120-133), so runs
are not reproducible.  These generators replace it with numpy's PCG64 and the
distributions of the five benchmark workloads:

  small32     32^3 RGBA32F, 64 uniform spheres, 256x256, 30 deg cone, 64 steps
  plain256    256^3 RGBA32F, 4096 clumped spheres, 1920x1080, 20 deg, 256 steps,
              camera on a radius-3 orbit at 16 seeded azimuths
  fp16_512    512^3 RGBA16F with 4-octave value-noise density, 16384 clumped spheres
  batched64   64 scenes of uneven volume / image sizes
  learned256  plain256 scenes for the learned far field

The sun is directional; its position for the light view (Sun::update,
src/Sun.hpp:26-43) is centre + direction * SUN_DISTANCE.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field, replace
from typing import List, Optional

import numpy as np

SUN_DISTANCE = 10.0
DEFAULT_SIGMA = 24.0  # extinction per unit texture length of the density cone mode


@dataclass
class Camera:
    eye: tuple
    target: tuple = (0.0, 0.0, 0.0)
    up: tuple = (0.0, 1.0, 0.0)
    fovy: float = math.radians(60.0)
    znear: float = 0.01
    zfar: float = 2500.0   # src/Camera.cpp:59
    aspect: float = 0.0    # 0: W / H
    P: Optional[np.ndarray] = None  # column-major matrices replacing the above (eye stays the sort point)
    V: Optional[np.ndarray] = None


@dataclass
class Cone:
    steps: int = 16                 # src/Shaders/ConeTraceShader.hpp:28-32 defaults
    cone_angle: float = 0.9
    initial_height: float = 0.1
    lod_offset: float = 0.0
    down_scaling: float = 1.0
    cone_mode: int = 2
    alpha_stop: float = 0.99
    sigma: float = DEFAULT_SIGMA
    composite: int = 1
    quantize_rgba8: int = 0
    compensated: int = 0
    clear_color: tuple = (0.2, 0.3, 0.5, 1.0)  # src/main.cpp:112


@dataclass
class Noise:
    octaves: int = 4
    seed: int = 0
    base_freq: float = 4.0
    lacunarity: float = 2.0
    gain: float = 0.5


@dataclass(frozen=True)
class NoiseMarch:
    """Soft-particle noise march (ConeTraceShader defaults, src/Shaders/ConeTraceShader.hpp:14-25)."""
    seed: int = 0
    step_size: float = 0.01
    opacity: float = 4.0
    octaves: int = 4
    freq_step: float = 3.0
    pers_step: float = 0.5
    adjust_size: float = 40.0
    min_steps: int = 2
    max_steps: int = 8
    min_color: float = 0.2
    color_scale: float = 0.45
    offsets: tuple = (0.0, 0.0, 0.0)


@dataclass(frozen=True)
class SunSprite:
    """Sun sprite drawn before the clouds (src/main.cpp:40-43 defaults)."""
    inner_color: tuple = (1.0, 1.0, 1.0)
    outer_color: tuple = (1.0, 1.0, 0.0)
    inner_radius: float = 1.0
    outer_radius: float = 2.0


@dataclass
class Scene:
    name: str
    N: int
    levels: int
    fmt: int                 # 0 R8, 1 RGBA32F, 2 RGBA16F
    mode: int                # 0 reference, 1 rgba
    xyz: np.ndarray          # (S, 3) float32, relative to position
    scales: np.ndarray       # (S,) float32
    sun_dir: np.ndarray      # (3,) float32 unit
    W: int
    H: int
    light_w: int
    light_h: int
    cameras: List[Camera]
    cone: Cone
    bounds_lo: tuple = (-1.0, -1.0, -1.0)
    bounds_hi: tuple = (1.0, 1.0, 1.0)
    position: tuple = (0.0, 0.0, 0.0)
    fluffiness: float = 1.0
    noise: Optional[Noise] = None
    noise_march: Optional[NoiseMarch] = None
    sun_sprite: Optional[SunSprite] = None
    directional: int = 1
    seed: int = 0
    sun_position: Optional[tuple] = None   # Sun::position when given (else centre + sun_dir * SUN_DISTANCE)
    meta: dict = field(default_factory=dict)

    @property
    def S(self) -> int:
        return int(self.xyz.shape[0])

    @property
    def sun_pos(self) -> np.ndarray:
        if self.sun_position is not None:
            return np.asarray(self.sun_position, np.float32)
        return (np.asarray(self.position, np.float32) + self.sun_dir * np.float32(SUN_DISTANCE)).astype(np.float32)


def _unit(v) -> np.ndarray:
    v = np.asarray(v, np.float64)
    return (v / np.linalg.norm(v)).astype(np.float32)


def full_levels(N: int) -> int:
    return int(math.floor(math.log2(N))) + 1


def clumped_spheres(rng: np.random.Generator, S: int):
    """8 Gaussian clumps (sigma 0.25, centres U[-0.5,0.5]^3), radii lognormal(ln 0.05, 0.4) in [0.01, 0.2]."""
    clumps = rng.uniform(-0.5, 0.5, size=(8, 3))
    which = rng.integers(0, 8, size=S)
    xyz = clumps[which] + rng.normal(0.0, 0.25, size=(S, 3))
    r = np.clip(np.exp(rng.normal(math.log(0.05), 0.4, size=S)), 0.01, 0.2)
    return xyz.astype(np.float32), r.astype(np.float32)


def hemisphere_sun(rng: np.random.Generator) -> np.ndarray:
    v = rng.normal(size=3)
    v[1] = abs(v[1])
    return _unit(v)


def orbit_cameras(rng: np.random.Generator, n: int, radius: float = 3.0) -> List[Camera]:
    az = rng.uniform(0.0, 2.0 * math.pi, size=n)
    return [Camera(eye=(float(radius * math.sin(a)), 0.0, float(radius * math.cos(a)))) for a in az]


def small32(seed: int = 0) -> Scene:
    rng = np.random.Generator(np.random.PCG64(seed))
    xyz = rng.uniform(-0.6, 0.6, size=(64, 3)).astype(np.float32)
    r = rng.uniform(0.08, 0.2, size=64).astype(np.float32)
    return Scene(name="small32", N=32, levels=6, fmt=1, mode=1, xyz=xyz, scales=r,
                 sun_dir=_unit((1.0, 1.0, 0.3)), W=256, H=256, light_w=256, light_h=256,
                 cameras=[Camera(eye=(0.0, 0.0, 3.0))],
                 cone=Cone(steps=64, cone_angle=math.radians(30.0)), seed=seed)


def plain256(seed: int = 0, W: int = 1920, H: int = 1080) -> Scene:
    rng = np.random.Generator(np.random.PCG64(seed))
    xyz, r = clumped_spheres(rng, 4096)
    sun = hemisphere_sun(rng)
    cams = orbit_cameras(rng, 16)
    return Scene(name="plain256", N=256, levels=9, fmt=1, mode=1, xyz=xyz, scales=r, sun_dir=sun,
                 W=W, H=H, light_w=W, light_h=H, cameras=cams,
                 cone=Cone(steps=256, cone_angle=math.radians(20.0)), seed=seed)


def fp16_512(seed: int = 0, W: int = 1920, H: int = 1080) -> Scene:
    rng = np.random.Generator(np.random.PCG64(seed))
    xyz, r = clumped_spheres(rng, 16384)
    sun = hemisphere_sun(rng)
    cams = orbit_cameras(rng, 16)
    return Scene(name="fp16_512", N=512, levels=10, fmt=2, mode=1, xyz=xyz, scales=r, sun_dir=sun,
                 W=W, H=H, light_w=W, light_h=H, cameras=cams,
                 cone=Cone(steps=256, cone_angle=math.radians(20.0), compensated=1),
                 noise=Noise(seed=seed), seed=seed)


def batched64(seed: int = 0, count: int = 64) -> List[Scene]:
    rng = np.random.Generator(np.random.PCG64(seed))
    scenes = []
    for i in range(count):
        N = int(rng.choice([64, 96, 128, 192, 256]))
        S = int(round(math.exp(rng.uniform(math.log(256), math.log(8192)))))
        side = int(rng.choice([256, 512, 768, 1024]))
        xyz, r = clumped_spheres(rng, S)
        sun = hemisphere_sun(rng)
        cam = orbit_cameras(rng, 1)
        scenes.append(Scene(name=f"batched64[{i}]", N=N, levels=full_levels(N), fmt=1, mode=1, xyz=xyz,
                            scales=r, sun_dir=sun, W=side, H=side, light_w=side, light_h=side, cameras=cam,
                            cone=Cone(steps=256, cone_angle=math.radians(20.0)), seed=seed * 1000 + i))
    return scenes


def learned256(seed: int = 0, W: int = 1920, H: int = 1080) -> Scene:
    s = plain256(seed, W, H)
    return replace(s, name="learned256")


def reference_mode(scene: Scene) -> Scene:
    """The shipped shaders' semantics on the same geometry: R8 volume holding
    the lit shell, weighted R-sum cone (res/conetrace_frag.glsl:64-79), point
    light at the sun position, GL back-to-front blending over the clear colour."""
    cone = replace(scene.cone, cone_mode=0, composite=0, quantize_rgba8=1)
    return replace(scene, name=scene.name + "/ref", fmt=0, mode=0, directional=0, cone=cone, noise=None)


def reference_camera(W: int, H: int) -> Camera:
    """The reference's start-up camera (src/Camera.cpp:12-61 with phi = theta
    = 0 and position = 0): lookAt = normalize((cos 0 cos 0, sin 0,
    cos 0 cos(PI/2.f - 0))) = (1, 0, -4.371139e-08), V = lookAt(0, that,
    +y), P = perspective(45.f, (float)(W / H), .01f, 2500.f) with glm taking
    45 as radians and W / H an integer division."""
    lz = float(np.float32(math.cos(float(np.float32(3.14159265359) / np.float32(2.0)))))
    return Camera(eye=(0.0, 0.0, 0.0), target=(1.0, 0.0, lz), up=(0.0, 1.0, 0.0), fovy=45.0, znear=0.01,
                  zfar=2500.0, aspect=float(W // H))


def reference_default(seed: int = 0, W: int = 1280, H: int = 720, noise_march: bool = True, run_time: float = 0.0,
                      fluffiness: float = 1.0, sun_sprite: bool = True, quantize_rgba8: int = 1) -> Scene:
    """The reference application's default frame (src/main.cpp:25-46, 81-82,
    112-123): a 32^3 GL_R8 volume with 4 mips at position (25, 0, 0) with
    bounds (-5, 5) on every axis, 200 billboards with offsets uniform in
    [-2.5, 2.5]^3 and scales uniform in [1, 2.5] (regenerateBillboards,
    src/CloudVolume.cpp:120-133, seeded here in place of rand()), the point
    sun at (5, 20, -5) with its sprite, the cone defaults vctSteps 16,
    vctConeAngle 0.9, vctConeInitialHeight 0.1, vctLodOffset 0,
    vctDownScaling 1, the noise march on with its defaults and octave offsets
    windVel * runTime = (0.01 * run_time, 0, 0)
    (src/Shaders/ConeTraceShader.hpp:14-36, ConeTraceShader.cpp:54-61), the
    GL blend into an RGBA8 framebuffer cleared to (0.2, 0.3, 0.5, 1), a
    light map the size of the window (src/Shaders/VoxelizeShader.cpp:20-22)
    and the start-up camera (reference_camera)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    xyz = rng.uniform(-2.5, 2.5, size=(200, 3)).astype(np.float32)
    r = rng.uniform(1.0, 2.5, size=200).astype(np.float32)
    pos = (25.0, 0.0, 0.0)
    sun = (5.0, 20.0, -5.0)
    sd = _unit(np.asarray(sun, np.float64) - np.asarray(pos, np.float64))
    nm = NoiseMarch(seed=seed, offsets=(float(np.float32(0.01) * np.float32(run_time)), 0.0, 0.0)) if noise_march else None
    return Scene(name="reference_default", N=32, levels=4, fmt=0, mode=0, xyz=xyz, scales=r, sun_dir=sd, W=W, H=H,
                 light_w=W, light_h=H, cameras=[reference_camera(W, H)],
                 cone=Cone(steps=16, cone_angle=0.9, initial_height=0.1, lod_offset=0.0, down_scaling=1.0,
                           cone_mode=0, composite=0, quantize_rgba8=quantize_rgba8),
                 bounds_lo=(-5.0, -5.0, -5.0), bounds_hi=(5.0, 5.0, 5.0), position=pos, fluffiness=fluffiness,
                 noise_march=nm, sun_sprite=SunSprite() if sun_sprite else None, directional=0, seed=seed,
                 sun_position=sun)


def scaled(scene: Scene, W: int, H: int) -> Scene:
    """Same scene rendered (and light-mapped) at another resolution."""
    return replace(scene, W=W, H=H, light_w=W, light_h=H)


CONFIGS = {
    "small32": small32,
    "plain256": plain256,
    "fp16_512": fp16_512,
    "learned256": learned256,
    "reference_default": reference_default,
}
