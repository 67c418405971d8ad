"""cloud-renderer_amd: MI355X (gfx950) billboard-cloud voxelize / mip / cone-trace path.

Python host layer over the C ABI in include/cloudvox.h (libcloudvox.so).  The
classes mirror the reference's host objects so that code written against
Cloud-Renderer's render passes reads the same:

  CloudVolume      src/CloudVolume.hpp:12-55   (position, x/y/zBounds, dimension,
                                                levels, voxelSize, billboards, fluffiness)
  Sun              src/Sun.hpp:12-19           (position; plus a directional flag)
  Camera           src/Camera.hpp:25-45        (eye/target/up, fovy, near/far)
  VoxelizeShader   src/Shaders/VoxelizeShader.hpp:18   voxelize(volume)
  ConeTraceShader  src/Shaders/ConeTraceShader.hpp:12  coneTrace(volume), with the
                   vct* parameters of src/Shaders/ConeTraceShader.hpp:28-32

All work runs on the GPU through the HIP library; there is no CPU fallback.
Output images are device buffers (a torch CUDA tensor or a raw device pointer).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _abi
from ._abi import (CR_COMPOSITE_FTB, CR_COMPOSITE_GL_BLEND, CR_CONE_FTB, CR_CONE_FTB_DENSITY,  # noqa: F401
                   CR_CONE_WEIGHTED, CR_FORMAT_R8, CR_FORMAT_RGBA16F, CR_FORMAT_RGBA32F, CR_MODE_REF,
                   CR_MODE_RGBA, CR_PIXEL_RGBA8, CR_PIXEL_RGBA32F, CloudVoxError, cr_camera, cr_cone_params,
                   cr_light, cr_noise_march, cr_stats, cr_sun_sprite)

__all__ = ["Context", "CloudVolume", "Sun", "Camera", "VoxelizeShader", "ConeTraceShader", "FarFieldMLP", "Batch",
           "CloudVoxError", "lib", "scenes"]


def lib():
    return _abi.lib()


def _check(rc: int, ctx_handle=None) -> None:
    if rc != 0:
        msg = lib().cr_last_error(ctx_handle)
        raise CloudVoxError(rc, msg.decode() if msg else "")


def _stream_handle(stream):
    if stream is None:
        return None
    if hasattr(stream, "cuda_stream"):
        return C.c_void_p(stream.cuda_stream)
    return C.c_void_p(int(stream))


def _f3(v):
    return _abi.F3(*[float(x) for x in v])


class Context:
    """One HIP device (cr_context)."""

    def __init__(self, device: int = 0, profiling: bool = False, flags: int = 0):
        h = C.c_void_p()
        _check(lib().cr_context_create(int(device), int(flags), C.byref(h)))
        self.handle = h
        self.device = device
        if profiling:
            self.set_profiling(True)

    def set_profiling(self, level) -> None:
        """0 off, 1 per-call counters and stage times (synchronises), 2 stage times summed without syncs."""
        _check(lib().cr_context_set_profiling(self.handle, int(level)), self.handle)

    def stats(self) -> dict:
        s = cr_stats()
        _check(lib().cr_get_stats(self.handle, C.byref(s)), self.handle)
        return {name: getattr(s, name) for name, _ in cr_stats._fields_}

    def close(self) -> None:
        if self.handle:
            lib().cr_context_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx: Optional[Context] = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


@dataclass
class Billboards:
    """CloudVolume::Billboards (src/CloudVolume.hpp:14-23)."""
    positions: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), np.float32))
    scales: np.ndarray = field(default_factory=lambda: np.zeros((0,), np.float32))

    @property
    def count(self) -> int:
        return int(self.positions.shape[0])


class CloudVolume:
    """CloudVolume(dim, bounds, position, mips) (src/CloudVolume.cpp:7-55) on the GPU.

    bounds is (lo, hi) applied to x, y and z like the reference's glm::vec2, or a
    pair of 3-vectors.  fmt: CR_FORMAT_*; mode: CR_MODE_REF (shipped shaders) or
    CR_MODE_RGBA (chord fill + lit shell).
    """

    def __init__(self, dim: int, bounds=(-1.0, 1.0), position=(0.0, 0.0, 0.0), mips: int = 4,
                 fmt: int = CR_FORMAT_RGBA32F, mode: int = CR_MODE_RGBA, context: Optional[Context] = None):
        self.ctx = context or default_context()
        b = np.asarray(bounds, np.float32)
        if b.shape == (2,):
            lo, hi = np.full(3, b[0], np.float32), np.full(3, b[1], np.float32)
        else:
            lo, hi = b[0].astype(np.float32), b[1].astype(np.float32)
        self.position = np.asarray(position, np.float32)
        self.xBounds = (float(lo[0]), float(hi[0]))
        self.yBounds = (float(lo[1]), float(hi[1]))
        self.zBounds = (float(lo[2]), float(hi[2]))
        self.dimension = int(dim)
        self.levels = int(mips)
        self.format = int(fmt)
        self.mode = int(mode)
        rng = hi - lo
        self.range = rng
        self.billboards = Billboards()
        self.fluffiness = 1.0
        h = C.c_void_p()
        lo3, hi3, p3 = _f3(lo), _f3(hi), _f3(self.position)
        _check(lib().cr_volume_create(self.ctx.handle, self.dimension, lo3, hi3, p3, self.levels, self.format,
                                      self.mode, C.byref(h)), self.ctx.handle)
        self.handle = h
        self.voxelSize = rng / np.float32(dim)

    # -- billboards -------------------------------------------------------
    def set_billboards(self, positions, scales) -> None:
        self.billboards = Billboards(np.ascontiguousarray(positions, np.float32).reshape(-1, 3),
                                     np.ascontiguousarray(scales, np.float32).reshape(-1))
        self.update()

    def update(self) -> None:
        """CloudVolume::update -> uploadBillboards (src/CloudVolume.cpp:84-93, 139-164)."""
        bb = self.billboards
        if bb.positions.shape[0] != bb.scales.shape[0]:
            raise ValueError("positions and scales differ in length")
        xyz = np.ascontiguousarray(bb.positions, np.float32)
        sc = np.ascontiguousarray(bb.scales, np.float32)
        _check(lib().cr_volume_set_billboards(self.handle, _abi.fptr(xyz), _abi.fptr(sc), bb.count,
                                              float(self.fluffiness)), self.ctx.handle)

    def set_noise(self, octaves=4, seed=0, base_freq=4.0, lacunarity=2.0, gain=0.5) -> None:
        _check(lib().cr_volume_set_noise(self.handle, int(octaves), int(seed) & 0xffffffff, float(base_freq),
                                         float(lacunarity), float(gain)), self.ctx.handle)

    def set_noise_march(self, nm: Optional[cr_noise_march]) -> None:
        """Switch the soft-particle noise march of coneTrace on (nm.enabled) or off (None)."""
        _check(lib().cr_volume_set_noise_march(self.handle, C.byref(nm) if nm is not None else None),
               self.ctx.handle)

    def set_sun_sprite(self, inner_color=(1.0, 1.0, 1.0), outer_color=(1.0, 1.0, 0.0), inner_radius=1.0,
                       outer_radius=2.0, enabled=True) -> None:
        """SunShader::render before the clouds (Sun::innerColor ... defaults, src/main.cpp:40-43);
        enabled=False switches it off."""
        sp = cr_sun_sprite(1 if enabled else 0, (C.c_float * 3)(*map(float, inner_color)),
                           (C.c_float * 3)(*map(float, outer_color)), float(inner_radius), float(outer_radius))
        _check(lib().cr_volume_set_sun_sprite(self.handle, C.byref(sp)), self.ctx.handle)

    def read_noise_map(self) -> np.ndarray:
        """The noise map bytes, (32, 32, 32, 4) int8 indexed [z, y, x, (r, g, b, a)]."""
        out = np.empty((32, 32, 32, 4), np.int8)
        _check(lib().cr_volume_read_noise_map(self.handle, out.ctypes.data_as(C.c_void_p), out.nbytes),
               self.ctx.handle)
        return out

    # -- readback ----------------------------------------------------------
    def occupancy(self, stream=None):
        cnt, h = C.c_uint64(), C.c_uint64()
        _check(lib().cr_volume_occupancy(self.handle, C.byref(cnt), C.byref(h), _stream_handle(stream)),
               self.ctx.handle)
        return int(cnt.value), int(h.value)

    def active_voxels(self, stream=None) -> int:
        """VoxelShader's activeVoxels (src/Shaders/VoxelShader.cpp:110-115): level-0 texels with r != 0."""
        cnt = C.c_uint64()
        _check(lib().cr_volume_active_voxels(self.handle, C.byref(cnt), _stream_handle(stream)), self.ctx.handle)
        return int(cnt.value)

    def level_dim(self, k: int) -> int:
        return max(1, self.dimension >> k)

    def read_level(self, k: int, stream=None) -> np.ndarray:
        n = self.level_dim(k)
        if self.format == CR_FORMAT_R8:
            out = np.empty((n, n, n), np.uint8)
        elif self.format == CR_FORMAT_RGBA32F:
            out = np.empty((n, n, n, 4), np.float32)
        else:
            out = np.empty((n, n, n, 4), np.uint16)
        _check(lib().cr_volume_read_level(self.handle, k, out.ctypes.data_as(C.c_void_p), out.nbytes,
                                          _stream_handle(stream)), self.ctx.handle)
        return out  # indexed [z, y, x, (c)]

    def read_occupancy(self, stream=None) -> np.ndarray:
        n = self.dimension
        out = np.empty((n, n, n), np.uint8)
        _check(lib().cr_volume_read_occupancy(self.handle, out.ctypes.data_as(C.c_void_p), out.nbytes,
                                              _stream_handle(stream)), self.ctx.handle)
        return out

    def read_position_map(self, w: int, h: int, stream=None) -> np.ndarray:
        out = np.empty((h, w, 4), np.float32)
        _check(lib().cr_volume_read_position_map(self.handle, out.ctypes.data_as(C.c_void_p), out.nbytes,
                                                 _stream_handle(stream)), self.ctx.handle)
        return out

    def read_draw_order(self, stream=None) -> np.ndarray:
        out = np.empty(max(self.billboards.count, 1), np.int32)
        _check(lib().cr_volume_read_draw_order(self.handle, out.ctypes.data_as(C.c_void_p), out.shape[0],
                                               _stream_handle(stream)), self.ctx.handle)
        return out[:self.billboards.count]

    def set_draw_order(self, mode: int) -> None:
        """_abi.CR_ORDER_SORTED (cone traces sort first) or CR_ORDER_UPLOAD (draw in upload order)."""
        _check(lib().cr_volume_set_draw_order(self.handle, int(mode)), self.ctx.handle)

    def sortBoards(self, point, stream=None) -> np.ndarray:
        """CloudVolume::sortBoards(point) (src/CloudVolume.cpp:65-82): permutes
        billboards.positions / scales in place, far to near from point, as the
        selection sort leaves them (equal distances included), and returns the
        permutation (new index -> old index).  The next update() uploads them."""
        n = self.billboards.count
        order = np.empty(max(n, 1), np.int32)
        _check(lib().cr_volume_sort_boards(self.handle, _f3(point), order.ctypes.data_as(C.c_void_p), order.shape[0],
                                           _stream_handle(stream)), self.ctx.handle)
        order = order[:n]
        self.billboards = Billboards(np.ascontiguousarray(self.billboards.positions[order]),
                                     np.ascontiguousarray(self.billboards.scales[order]))
        return order

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().cr_volume_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class Sun:
    """Sun (src/Sun.hpp:12-19).  position drives the light view; the cone goes
    toward `position` (point light, reference) or along `direction`."""
    position: tuple = (5.0, 20.0, -5.0)   # src/main.cpp:37
    direction: tuple = (0.0, 1.0, 0.0)
    directional: bool = False

    def to_c(self) -> cr_light:
        return cr_light(_f3(self.position), _f3(self.direction), 1 if self.directional else 0)


@dataclass
class Camera:
    """Camera (src/Camera.hpp:25-45, src/Camera.cpp:59-60): V = lookAt(eye,
    target, up), P = perspective(fovy, aspect, znear, zfar) with glm's
    arithmetic (fovy in radians as glm takes it, so the reference's 45 is
    accepted as is; aspect 0 = width / height).  P and V (column-major 4x4,
    as ConeTraceShader uploads them) replace everything but eye, the sort
    point, when both are given."""
    eye: tuple = (0.0, 0.0, 3.0)
    target: tuple = (0.0, 0.0, 0.0)
    up: tuple = (0.0, 1.0, 0.0)
    fovy: float = math.radians(60.0)
    znear: float = 0.01
    zfar: float = 2500.0
    aspect: float = 0.0
    P: Optional[np.ndarray] = None
    V: Optional[np.ndarray] = None

    def to_c(self) -> cr_camera:
        use = self.P is not None and self.V is not None
        Pm = np.asarray(self.P, np.float32).reshape(16) if use else np.zeros(16, np.float32)
        Vm = np.asarray(self.V, np.float32).reshape(16) if use else np.zeros(16, np.float32)
        return cr_camera(_f3(self.eye), _f3(self.target), _f3(self.up), float(self.fovy), float(self.znear),
                         float(self.zfar), float(self.aspect), 1 if use else 0,
                         (C.c_float * 16)(*[float(x) for x in Pm]), (C.c_float * 16)(*[float(x) for x in Vm]))


class VoxelizeShader:
    """VoxelizeShader::voxelize (src/Shaders/VoxelizeShader.cpp:18-31).  The
    position map is light_w x light_h (the reference uses the window size)."""

    def __init__(self, light_w: int = 1280, light_h: int = 720):
        self.width = light_w
        self.height = light_h

    def voxelize(self, volume: CloudVolume, sun: Sun, stream=None) -> None:
        light = sun.to_c()
        _check(lib().cr_voxelize(volume.handle, C.byref(light), int(self.width), int(self.height),
                                 _stream_handle(stream)), volume.ctx.handle)


class ConeTraceShader:
    """ConeTraceShader::coneTrace (src/Shaders/ConeTraceShader.cpp:15-82)."""

    def __init__(self, **kw):
        # src/Shaders/ConeTraceShader.hpp:28-32
        self.vctSteps = 16
        self.vctConeAngle = 0.9
        self.vctConeInitialHeight = 0.1
        self.vctLodOffset = 0.0
        self.vctDownScaling = 1.0
        self.coneMode = CR_CONE_WEIGHTED
        self.alphaStop = 0.99
        self.sigma = 24.0
        self.composite = CR_COMPOSITE_GL_BLEND
        self.quantizeRGBA8 = True
        self.compensated = False
        self.clearColor = (0.2, 0.3, 0.5, 1.0)   # src/main.cpp:112
        # noise march, src/Shaders/ConeTraceShader.hpp:14-25 (the reference
        # starts with doNoiseSample = true; here it starts off, the cone
        # output alone being the default look)
        self.doNoiseSample = False
        self.noiseSeed = 0            # replaces rand() in initNoiseMap
        self.stepSize = 0.01
        self.noiseOpacity = 4.0
        self.numOctaves = 4
        self.freqStep = 3.0
        self.persStep = 0.5
        self.adjustSize = 40.0
        self.minNoiseSteps = 2
        self.maxNoiseSteps = 8
        self.minNoiseColor = 0.2
        self.noiseColorScale = 0.45
        self.windVel = (0.01, 0.0, 0.0)
        self.runTime = 0.0            # Window::runTime: octaveOffsets = windVel * runTime
        for k, v in kw.items():
            if not hasattr(self, k):
                raise AttributeError(k)
            setattr(self, k, v)

    def params(self) -> cr_cone_params:
        return cr_cone_params(int(self.vctSteps), float(self.vctConeAngle), float(self.vctConeInitialHeight),
                              float(self.vctLodOffset), float(self.vctDownScaling), int(self.coneMode),
                              float(self.alphaStop), float(self.sigma), int(self.composite),
                              1 if self.quantizeRGBA8 else 0, 1 if self.compensated else 0,
                              _abi.F4(*[float(c) for c in self.clearColor]))

    def noise_params(self) -> Optional[cr_noise_march]:
        if not self.doNoiseSample:
            return None
        off = (C.c_float * 3)(*[float(np.float32(w) * np.float64(self.runTime)) for w in self.windVel])
        return cr_noise_march(1, int(self.noiseSeed) & 0xffffffff, float(self.stepSize), float(self.noiseOpacity),
                              int(self.numOctaves), float(self.freqStep), float(self.persStep),
                              float(self.adjustSize), int(self.minNoiseSteps), int(self.maxNoiseSteps),
                              float(self.minNoiseColor), float(self.noiseColorScale), off)

    @staticmethod
    def _image_args(image, width, pixel_format, pitch, stream):
        if hasattr(image, "data_ptr"):
            ptr = image.data_ptr()
            if pitch is None:
                pitch = image.stride(0) * image.element_size()
            if stream is None:
                import torch
                stream = torch.cuda.current_stream(image.device)
        else:
            ptr = int(image)
            if pitch is None:
                pitch = width * (16 if pixel_format == CR_PIXEL_RGBA32F else 4)
        return ptr, pitch, stream

    def coneTrace(self, volume: CloudVolume, sun: Sun, camera: Camera, image, width: int, height: int,
                  pixel_format: int = CR_PIXEL_RGBA32F, pitch: Optional[int] = None, stream=None) -> None:
        """Render into `image`: a torch CUDA tensor (H, W, 4) or a raw device pointer (int)."""
        ptr, pitch, stream = self._image_args(image, width, pixel_format, pitch, stream)
        light = sun.to_c()
        cam = camera.to_c()
        p = self.params()
        volume.set_noise_march(self.noise_params())
        _check(lib().cr_cone_trace(volume.handle, C.byref(light), C.byref(cam), C.byref(p), C.c_void_p(ptr),
                                   int(width), int(height), C.c_size_t(pitch), int(pixel_format),
                                   _stream_handle(stream)), volume.ctx.handle)

    def coneTraceLearned(self, volume: CloudVolume, sun: Sun, camera: Camera, mlp: "FarFieldMLP", image,
                         width: int, height: int, k: int = 32, err_bound: float = 2e-3,
                         pixel_format: int = CR_PIXEL_RGBA32F, pitch: Optional[int] = None, stream=None) -> None:
        """coneTrace with the far field after k samples predicted by `mlp`;
        cones whose predicted error times remaining transmittance exceeds
        err_bound fall back to the plain march (err_bound < 0: all do)."""
        ptr, pitch, stream = self._image_args(image, width, pixel_format, pitch, stream)
        light = sun.to_c()
        cam = camera.to_c()
        p = self.params()
        volume.set_noise_march(self.noise_params())
        _check(lib().cr_cone_trace_learned(volume.handle, C.byref(light), C.byref(cam), C.byref(p), mlp.handle,
                                           int(k), float(err_bound), C.c_void_p(ptr), int(width), int(height),
                                           C.c_size_t(pitch), int(pixel_format), _stream_handle(stream)),
               volume.ctx.handle)

    def segments(self, volume: CloudVolume, sun: Sun, p0, k: int = 32, stream=None):
        """Training rows of the far-field MLP from the plain march: p0 is a
        (n, 3) float32 CUDA tensor of cone start points in texture space.
        Returns (features (n, 17), targets (n, 5), valid (n,) int32) tensors."""
        import torch
        n = int(p0.shape[0])
        dev = p0.device
        feat = torch.empty((n, _abi.CR_MLP_INPUTS), dtype=torch.float32, device=dev)
        tgt = torch.empty((n, _abi.CR_MLP_OUTPUTS), dtype=torch.float32, device=dev)
        valid = torch.empty((n,), dtype=torch.int32, device=dev)
        if stream is None:
            stream = torch.cuda.current_stream(dev)
        light = sun.to_c()
        p = self.params()
        _check(lib().cr_volume_segments(volume.handle, C.byref(light), C.byref(p), int(k),
                                        C.c_void_p(p0.contiguous().data_ptr()), n, C.c_void_p(feat.data_ptr()),
                                        C.c_void_p(tgt.data_ptr()), C.c_void_p(valid.data_ptr()),
                                        _stream_handle(stream)), volume.ctx.handle)
        return feat, tgt, valid


class FarFieldMLP:
    """The learned far field (cr_mlp): 17 -> 64 x 4 (GELU) -> 5, BF16 weights on
    the matrix cores, FP32 master parameters trained with Adam by the
    library's own forward/backward kernels."""

    PARAMS = _abi.CR_MLP_PARAMS

    def __init__(self, context: Optional[Context] = None, params: Optional[np.ndarray] = None):
        self.ctx = context or default_context()
        h = C.c_void_p()
        _check(lib().cr_mlp_create(self.ctx.handle, C.byref(h)), self.ctx.handle)
        self.handle = h
        if params is not None:
            self.set_params(params)

    @staticmethod
    def init_params(seed: int = 0) -> np.ndarray:
        """He-style initialisation (normal, std sqrt(2 / fan_in)), zero biases."""
        rng = np.random.Generator(np.random.PCG64(seed))
        dims = [(17, 64), (64, 64), (64, 64), (64, 64), (64, 5)]
        out = []
        for fin, fout in dims:
            out.append(rng.normal(0.0, math.sqrt(2.0 / fin), size=(fout, fin)).astype(np.float32).ravel())
            out.append(np.zeros(fout, np.float32))
        p = np.concatenate(out)
        assert p.size == FarFieldMLP.PARAMS
        return p

    def set_params(self, params: np.ndarray) -> None:
        a = np.ascontiguousarray(params, np.float32).ravel()
        _check(lib().cr_mlp_set_params(self.handle, _abi.fptr(a), int(a.size)), self.ctx.handle)

    def get_params(self) -> np.ndarray:
        a = np.empty(self.PARAMS, np.float32)
        _check(lib().cr_mlp_get_params(self.handle, _abi.fptr(a), int(a.size)), self.ctx.handle)
        return a

    def forward(self, x, stream=None):
        import torch
        x = x.contiguous()
        n = int(x.shape[0])
        y = torch.empty((n, _abi.CR_MLP_OUTPUTS), dtype=torch.float32, device=x.device)
        if stream is None:
            stream = torch.cuda.current_stream(x.device)
        _check(lib().cr_mlp_forward(self.handle, C.c_void_p(x.data_ptr()), n, C.c_void_p(y.data_ptr()),
                                    _stream_handle(stream)), self.ctx.handle)
        return y

    def gradient(self, x, t, stream=None):
        import torch
        x, t = x.contiguous(), t.contiguous()
        g = np.empty(self.PARAMS, np.float32)
        loss = C.c_float()
        if stream is None:
            stream = torch.cuda.current_stream(x.device)
        _check(lib().cr_mlp_gradient(self.handle, C.c_void_p(x.data_ptr()), C.c_void_p(t.data_ptr()),
                                     int(x.shape[0]), _abi.fptr(g), C.byref(loss), _stream_handle(stream)),
               self.ctx.handle)
        return g, float(loss.value)

    def train_step(self, x, t, lr: float, stream=None) -> None:
        import torch
        if stream is None:
            stream = torch.cuda.current_stream(x.device)
        _check(lib().cr_mlp_train_step(self.handle, C.c_void_p(x.data_ptr()), C.c_void_p(t.data_ptr()),
                                       int(x.shape[0]), float(lr), _stream_handle(stream)), self.ctx.handle)

    def train_step_static(self, x, t, lr: float, stream=None) -> None:
        """train_step without allocation (cr_mlp_train_step_static): at most
        CR_MLP_STATIC_ROWS rows; calls overlapping on one GPU race."""
        import torch
        if stream is None:
            stream = torch.cuda.current_stream(x.device)
        _check(lib().cr_mlp_train_step_static(self.handle, C.c_void_p(x.data_ptr()), C.c_void_p(t.data_ptr()),
                                              int(x.shape[0]), float(lr), _stream_handle(stream)), self.ctx.handle)

    def last_loss(self, stream=None) -> float:
        import torch
        loss = C.c_float()
        if stream is None:
            stream = torch.cuda.current_stream()
        _check(lib().cr_mlp_last_loss(self.handle, C.byref(loss), _stream_handle(stream)), self.ctx.handle)
        return float(loss.value)

    def save(self, path) -> None:
        np.savez_compressed(path, params=self.get_params())

    @staticmethod
    def load_params(path) -> np.ndarray:
        with np.load(path, allow_pickle=False) as z:
            return np.asarray(z["params"], np.float32)

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().cr_mlp_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many scenes rendered by one hipGraph (cr_batch): every scene's voxelize,
    draw order and binning, then one cone-trace launch over all scenes' tiles.

    entries: list of (volume, sun, camera, tracer, image, voxelizer) with
    images as torch CUDA tensors (H, W, 4) float32."""

    def __init__(self, entries, context: Optional[Context] = None):
        self.ctx = context or entries[0][0].ctx
        items = (_abi.cr_batch_item * len(entries))()
        self._keep = []
        for i, (vol, sun, cam, tr, img, vox) in enumerate(entries):
            it = items[i]
            it.volume = vol.handle
            it.light = sun.to_c()
            it.light_w, it.light_h = int(vox.width), int(vox.height)
            it.camera = cam.to_c()
            it.params = tr.params()
            it.d_image = img.data_ptr()
            it.height, it.width = int(img.shape[0]), int(img.shape[1])
            it.pitch_bytes = img.stride(0) * img.element_size()
            it.pixel_format = CR_PIXEL_RGBA32F if img.element_size() == 4 else CR_PIXEL_RGBA8
            self._keep.append((vol, img))
        h = C.c_void_p()
        _check(lib().cr_batch_create(self.ctx.handle, items, len(entries), C.byref(h)), self.ctx.handle)
        self.handle = h

    def config(self) -> dict:
        c = _abi.cr_batch_config()
        _check(lib().cr_batch_get_config(self.handle, C.byref(c)), self.ctx.handle)
        return {"tile_w": c.tile_w, "waves_per_group": c.waves_per_group, "groups": c.groups}

    def set_config(self, tile_w: int = 8, waves_per_group: int = 1, groups: int = 0) -> None:
        c = _abi.cr_batch_config(int(tile_w), int(waves_per_group), int(groups))
        _check(lib().cr_batch_set_config(self.handle, C.byref(c)), self.ctx.handle)

    def run(self, stream=None) -> None:
        if stream is None:
            import torch
            stream = torch.cuda.current_stream()
        _check(lib().cr_batch_run(self.handle, _stream_handle(stream)), self.ctx.handle)

    def autotune(self, repeats: int = 3, stream=None):
        if stream is None:
            import torch
            stream = torch.cuda.current_stream()
        best = _abi.cr_batch_config()
        ms = C.c_float()
        _check(lib().cr_batch_autotune(self.handle, int(repeats), _stream_handle(stream), C.byref(best),
                                       C.byref(ms)), self.ctx.handle)
        return {"tile_w": best.tile_w, "waves_per_group": best.waves_per_group, "groups": best.groups}, float(ms.value)

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().cr_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


from . import scenes  # noqa: E402


def scene_objects(scene, camera_index: int = 0, context: Optional[Context] = None):
    """Build (volume, sun, camera, voxelizer, tracer) for a scenes.Scene."""
    vol = CloudVolume(scene.N, (scene.bounds_lo, scene.bounds_hi), scene.position, scene.levels, scene.fmt,
                      scene.mode, context=context)
    vol.fluffiness = scene.fluffiness
    vol.set_billboards(scene.xyz, scene.scales)
    if scene.noise is not None:
        n = scene.noise
        vol.set_noise(n.octaves, n.seed, n.base_freq, n.lacunarity, n.gain)
    sun = Sun(position=tuple(float(x) for x in scene.sun_pos), direction=tuple(float(x) for x in scene.sun_dir),
              directional=bool(scene.directional))
    c = scene.cameras[camera_index]
    cam = Camera(eye=c.eye, target=c.target, up=c.up, fovy=c.fovy, znear=c.znear, zfar=c.zfar,
                 aspect=getattr(c, "aspect", 0.0), P=getattr(c, "P", None), V=getattr(c, "V", None))
    vox = VoxelizeShader(scene.light_w, scene.light_h)
    k = scene.cone
    tr = ConeTraceShader(vctSteps=k.steps, vctConeAngle=k.cone_angle, vctConeInitialHeight=k.initial_height,
                         vctLodOffset=k.lod_offset, vctDownScaling=k.down_scaling, coneMode=k.cone_mode,
                         alphaStop=k.alpha_stop, sigma=k.sigma, composite=k.composite,
                         quantizeRGBA8=bool(k.quantize_rgba8), compensated=bool(k.compensated),
                         clearColor=k.clear_color)
    if scene.sun_sprite is not None:
        q = scene.sun_sprite
        vol.set_sun_sprite(q.inner_color, q.outer_color, q.inner_radius, q.outer_radius)
    if scene.noise_march is not None:
        m = scene.noise_march
        tr.doNoiseSample = True
        tr.noiseSeed, tr.stepSize, tr.noiseOpacity, tr.numOctaves = m.seed, m.step_size, m.opacity, m.octaves
        tr.freqStep, tr.persStep, tr.adjustSize = m.freq_step, m.pers_step, m.adjust_size
        tr.minNoiseSteps, tr.maxNoiseSteps, tr.minNoiseColor, tr.noiseColorScale = (m.min_steps, m.max_steps,
                                                                                    m.min_color, m.color_scale)
        # offsets given directly: windVel = offsets, runTime = 1
        tr.windVel, tr.runTime = tuple(m.offsets), 1.0
    return vol, sun, cam, vox, tr
