"""ctypes wrapper of the CPU oracle (oracle/cv_oracle.cpp) for the tests.

Test infrastructure only: the product library never imports this module.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent.parent
ORACLE_LIB = REPO / "oracle" / "libcv_oracle.so"

F3 = C.c_float * 3
F4 = C.c_float * 4


class cvo_volume(C.Structure):
    _fields_ = [("dim", C.c_int), ("levels", C.c_int), ("format", C.c_int), ("mode", C.c_int),
                ("bounds_lo", F3), ("bounds_hi", F3), ("position", F3), ("noise_octaves", C.c_int),
                ("noise_seed", C.c_uint32), ("noise_base_freq", C.c_float), ("noise_lacunarity", C.c_float),
                ("noise_gain", C.c_float)]


class cvo_light(C.Structure):
    _fields_ = [("position", F3), ("direction", F3), ("directional", C.c_int)]


class cvo_camera(C.Structure):
    _fields_ = [("eye", F3), ("target", F3), ("up", F3), ("fovy", C.c_float), ("znear", C.c_float),
                ("zfar", C.c_float), ("aspect", C.c_float), ("use_matrices", C.c_int), ("P", C.c_float * 16),
                ("V", C.c_float * 16)]


class cvo_noise_march(C.Structure):
    _fields_ = [("enabled", C.c_int), ("seed", C.c_uint32), ("step_size", C.c_float), ("opacity", C.c_float),
                ("octaves", C.c_int), ("freq_step", C.c_float), ("pers_step", C.c_float),
                ("adjust_size", C.c_float), ("min_steps", C.c_int), ("max_steps", C.c_int),
                ("min_color", C.c_float), ("color_scale", C.c_float), ("offsets", C.c_float * 3)]


def noise_march_of(m) -> cvo_noise_march:
    return cvo_noise_march(1, int(m.seed) & 0xffffffff, m.step_size, m.opacity, int(m.octaves), m.freq_step,
                           m.pers_step, m.adjust_size, int(m.min_steps), int(m.max_steps), m.min_color,
                           m.color_scale, (C.c_float * 3)(*[float(o) for o in m.offsets]))


class cvo_sun_sprite(C.Structure):
    _fields_ = [("enabled", C.c_int), ("inner_color", C.c_float * 3), ("outer_color", C.c_float * 3),
                ("inner_radius", C.c_float), ("outer_radius", C.c_float)]


class cvo_cone(C.Structure):
    _fields_ = [("steps", C.c_int), ("cone_angle", C.c_float), ("initial_height", C.c_float),
                ("lod_offset", C.c_float), ("down_scaling", C.c_float), ("cone_mode", C.c_int),
                ("alpha_stop", C.c_float), ("sigma", C.c_float), ("composite", C.c_int),
                ("quantize_rgba8", C.c_int), ("compensated", C.c_int), ("clear_color", F4)]


class cvo_lightview(C.Structure):
    _fields_ = [("V", C.c_float * 16), ("near_plane", F3), ("clip", C.c_float), ("ortho", F4),
                ("znear", C.c_float), ("zfar", C.c_float)]


class cvo_stats(C.Structure):
    _fields_ = [("fragments", C.c_int64), ("samples", C.c_int64), ("two_level", C.c_int64),
                ("light_fragments", C.c_int64), ("chord_stores", C.c_int64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not ORACLE_LIB.exists():
            raise ImportError(f"{ORACLE_LIB} missing: run `make -C oracle`")
        L = C.CDLL(str(ORACLE_LIB))
        vp = C.c_void_p
        P = C.POINTER
        L.cvo_sun_update.argtypes = [P(cvo_volume), P(C.c_float), P(cvo_lightview)]
        L.cvo_voxelize.argtypes = [P(cvo_volume), C.c_int, vp, vp, P(cvo_light), C.c_int, C.c_int, vp, vp, vp, vp,
                                   P(cvo_stats)]
        L.cvo_voxelize.restype = C.c_int
        L.cvo_occupancy.argtypes = [C.c_int, vp, vp, P(C.c_uint64), P(C.c_uint64)]
        L.cvo_build_levels.argtypes = [P(cvo_volume), vp, vp, vp, vp]
        L.cvo_build_levels.restype = C.c_int
        L.cvo_sort_boards.argtypes = [C.c_int, vp, P(C.c_float), P(C.c_float), vp]
        L.cvo_board_distances.argtypes = [C.c_int, vp, P(C.c_float), P(C.c_float), vp]
        L.cvo_f2h_array.argtypes = [vp, vp, C.c_int64]
        L.cvo_h2f_array.argtypes = [vp, vp, C.c_int64]
        L.cvo_exp2_neg_array.argtypes = [vp, vp, C.c_int64]
        L.cvo_noise_voxels.argtypes = [P(cvo_volume), vp, C.c_int64, vp, vp]
        L.cvo_voxel_index.argtypes = [P(cvo_volume), vp, C.c_int64, vp]
        L.cvo_cone_trace.argtypes = [P(cvo_volume), vp, vp, C.c_int, vp, vp, vp, P(cvo_camera), P(cvo_light),
                                     P(cvo_cone), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp,
                                     P(cvo_stats)]
        L.cvo_cone_trace.restype = C.c_int
        L.cvo_cone_trace_nm.argtypes = L.cvo_cone_trace.argtypes + [P(cvo_noise_march)]
        L.cvo_cone_trace_nm.restype = C.c_int
        L.cvo_cone_trace_ex.argtypes = L.cvo_cone_trace_nm.argtypes + [P(cvo_sun_sprite)]
        L.cvo_cone_trace_ex.restype = C.c_int
        L.cvo_noise_map.argtypes = [C.c_uint32, vp]
        L.cvo_noise_sample.argtypes = [C.c_uint32, vp, C.c_int64, vp, vp]
        L.cvo_noise_fragment.argtypes = [P(cvo_noise_march), vp, vp]
        L.cvo_noise_march_valid.argtypes = [P(cvo_noise_march)]
        L.cvo_noise_march_valid.restype = C.c_int
        L.cvo_camera_setup.argtypes = [P(cvo_camera), C.c_int, C.c_int, vp]
        L.cvo_camera_setup.restype = C.c_int
        L.cvo_step_table.argtypes = [P(cvo_cone), C.c_int, C.c_int, vp, vp, vp, vp]
        L.cvo_sample.argtypes = [P(cvo_volume), vp, vp, C.c_float, vp, C.c_int64, vp, vp]
        L.cvo_trace_one.argtypes = [P(cvo_volume), vp, vp, P(cvo_cone), P(C.c_float), P(C.c_float), vp, vp,
                                    P(C.c_int32)]
        L.cvo_num_threads.restype = C.c_int
        L.cvo_set_num_threads.argtypes = [C.c_int]
        L.cvo_segments.argtypes = [P(cvo_volume), vp, P(cvo_cone), P(cvo_light), C.c_int, vp, C.c_int64, vp, vp, vp]
        L.cvo_segments.restype = C.c_int
        L.cvo_mlp_param_count.restype = C.c_int
        L.cvo_mlp_forward.argtypes = [vp, vp, C.c_int64, C.c_int, vp]
        L.cvo_mlp_forward.restype = C.c_int
        L.cvo_mlp_grad.argtypes = [vp, vp, vp, C.c_int64, C.c_int, vp, vp, P(C.c_double)]
        L.cvo_mlp_grad.restype = C.c_int
        _lib = L
    return _lib


def ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def f3(v):
    return F3(*[float(x) for x in v])


def make_volume(N, levels, fmt=1, mode=1, lo=(-1, -1, -1), hi=(1, 1, 1), pos=(0, 0, 0), noise=None) -> cvo_volume:
    v = cvo_volume()
    v.dim, v.levels, v.format, v.mode = int(N), int(levels), int(fmt), int(mode)
    v.bounds_lo, v.bounds_hi, v.position = f3(lo), f3(hi), f3(pos)
    if noise is not None:
        v.noise_octaves, v.noise_seed = int(noise.octaves), int(noise.seed) & 0xffffffff
        v.noise_base_freq, v.noise_lacunarity, v.noise_gain = noise.base_freq, noise.lacunarity, noise.gain
    return v


def volume_of(scene) -> cvo_volume:
    return make_volume(scene.N, scene.levels, scene.fmt, scene.mode, scene.bounds_lo, scene.bounds_hi,
                       scene.position, scene.noise)


def light_of(scene) -> cvo_light:
    L = cvo_light()
    L.position = f3(scene.sun_pos)
    L.direction = f3(scene.sun_dir)
    L.directional = int(scene.directional)
    return L


def camera_of(c) -> cvo_camera:
    k = cvo_camera()
    k.eye, k.target, k.up = f3(c.eye), f3(c.target), f3(c.up)
    k.fovy, k.znear, k.zfar = float(c.fovy), float(c.znear), float(c.zfar)
    k.aspect = float(getattr(c, "aspect", 0.0))
    P, V = getattr(c, "P", None), getattr(c, "V", None)
    if P is not None and V is not None:
        k.use_matrices = 1
        k.P = (C.c_float * 16)(*[float(x) for x in np.asarray(P, np.float32).reshape(16)])
        k.V = (C.c_float * 16)(*[float(x) for x in np.asarray(V, np.float32).reshape(16)])
    return k


def cone_of(c) -> cvo_cone:
    k = cvo_cone()
    k.steps, k.cone_angle, k.initial_height = int(c.steps), float(c.cone_angle), float(c.initial_height)
    k.lod_offset, k.down_scaling, k.cone_mode = float(c.lod_offset), float(c.down_scaling), int(c.cone_mode)
    k.alpha_stop, k.sigma, k.composite = float(c.alpha_stop), float(c.sigma), int(c.composite)
    k.quantize_rgba8, k.compensated = int(c.quantize_rgba8), int(c.compensated)
    k.clear_color = F4(*[float(x) for x in c.clear_color])
    return k


def radii(scene) -> np.ndarray:
    r = scene.scales.astype(np.float32)
    if scene.fluffiness != 1.0:
        r = (r * np.float32(scene.fluffiness)).astype(np.float32)
    return np.ascontiguousarray(r)


@dataclass
class VoxResult:
    shell: np.ndarray   # [z, y, x] uint8
    fill: np.ndarray
    posmap: np.ndarray  # [h, w, 4]
    winner: np.ndarray  # [h, w]
    stats: cvo_stats

    def state(self) -> np.ndarray:
        st = np.where(self.shell != 0, 2, np.where(self.fill != 0, 1, 0)).astype(np.uint8)
        return st


def voxelize(scene) -> VoxResult:
    N = scene.N
    vol = volume_of(scene)
    shell = np.zeros((N, N, N), np.uint8)
    fill = np.zeros((N, N, N), np.uint8)
    pm = np.zeros((scene.light_h, scene.light_w, 4), np.float32)
    win = np.zeros((scene.light_h, scene.light_w), np.int32)
    st = cvo_stats()
    xyz = np.ascontiguousarray(scene.xyz, np.float32)
    r = radii(scene)
    L = light_of(scene)
    rc = lib().cvo_voxelize(C.byref(vol), scene.S, ptr(xyz), ptr(r), C.byref(L), scene.light_w, scene.light_h,
                            ptr(shell), ptr(fill), ptr(pm), ptr(win), C.byref(st))
    assert rc == 0
    if scene.mode == 0:
        fill[:] = 0
    return VoxResult(shell, fill, pm, win, st)


def occupancy(N, shell, fill):
    c, h = C.c_uint64(), C.c_uint64()
    lib().cvo_occupancy(N, ptr(shell), ptr(fill), C.byref(c), C.byref(h))
    return int(c.value), int(h.value)


def level_dims(N, L):
    return [max(1, N >> k) for k in range(L)]


def build_levels(scene, shell, fill, want_f64=True):
    """Returns (native_flat, f64_flat, native_levels list, f64_levels list)."""
    N, L, fmt = scene.N, scene.levels, scene.fmt
    C_ = 1 if fmt == 0 else 4
    dims = level_dims(N, L)
    total = sum(d ** 3 for d in dims) * C_
    dt = {0: np.uint8, 1: np.float32, 2: np.uint16}[fmt]
    nat = np.zeros(total, dt)
    f64 = np.zeros(total, np.float64) if want_f64 else None
    vol = volume_of(scene)
    rc = lib().cvo_build_levels(C.byref(vol), ptr(shell), ptr(fill if scene.mode == 1 else None), ptr(nat), ptr(f64))
    assert rc == 0
    nl, fl, off = [], [], 0
    for d in dims:
        n = d ** 3 * C_
        shp = (d, d, d) if C_ == 1 else (d, d, d, 4)
        nl.append(nat[off:off + n].reshape(shp))
        if f64 is not None:
            fl.append(f64[off:off + n].reshape(shp))
        off += n
    return nat, f64, nl, fl


def sort_boards(scene, eye):
    order = np.zeros(scene.S, np.int32)
    xyz = np.ascontiguousarray(scene.xyz, np.float32)
    lib().cvo_sort_boards(scene.S, ptr(xyz), f3(scene.position), f3(eye), ptr(order))
    return order


def cone_trace(scene, nat, f64, order, cam, row0=0, row1=None, want_f64=True, row_step=1):
    W, H = scene.W, scene.H
    row1 = H if row1 is None else row1
    vol = volume_of(scene)
    img32 = np.zeros((H, W, 4), np.float32)
    img64 = np.zeros((H, W, 4), np.float64) if (want_f64 and f64 is not None) else None
    st = cvo_stats()
    xyz = np.ascontiguousarray(scene.xyz, np.float32)
    r = radii(scene)
    k = camera_of(cam)
    L = light_of(scene)
    cp = cone_of(scene.cone)
    args = (C.byref(vol), ptr(nat), ptr(f64 if img64 is not None else None), scene.S, ptr(xyz), ptr(r), ptr(order),
            C.byref(k), C.byref(L), C.byref(cp), W, H, int(row0), int(row1), int(row_step), ptr(img32), ptr(img64),
            C.byref(st))
    nm = getattr(scene, "noise_march", None)
    sun = getattr(scene, "sun_sprite", None)
    if nm is not None or sun is not None:
        m = noise_march_of(nm) if nm is not None else None
        q = (cvo_sun_sprite(1, (C.c_float * 3)(*sun.inner_color), (C.c_float * 3)(*sun.outer_color),
                            sun.inner_radius, sun.outer_radius) if sun is not None else None)
        rc = lib().cvo_cone_trace_ex(*args, C.byref(m) if m is not None else None,
                                     C.byref(q) if q is not None else None)
    else:
        rc = lib().cvo_cone_trace(*args)
    assert rc == 0
    return img32, img64, st


def camera_setup(cam, W: int, H: int):
    """(V[16], kx, ky, znear, zfar, S, U, B, noise ray) of the oracle's camera setup, or None."""
    out = np.zeros(32, np.float32)
    if lib().cvo_camera_setup(C.byref(camera_of(cam)), int(W), int(H), ptr(out)) != 0:
        return None
    return dict(V=out[:16], kx=out[16], ky=out[17], znear=out[18], zfar=out[19], S=out[20:23], U=out[23:26],
                B=out[26:29], ray=out[29:32])


def noise_march_valid(m) -> bool:
    return bool(lib().cvo_noise_march_valid(C.byref(noise_march_of(m))))


def noise_map(seed: int) -> np.ndarray:
    """The noise-march map bytes, (32, 32, 32, 4) int8 indexed [z, y, x, (r, g, b, a)]."""
    out = np.zeros((32, 32, 32, 4), np.int8)
    lib().cvo_noise_map(int(seed) & 0xffffffff, ptr(out))
    return out


def noise_sample(seed: int, pts: np.ndarray):
    pts = np.ascontiguousarray(pts, np.float32)
    n = pts.shape[0]
    o32 = np.zeros((n, 4), np.float32)
    o64 = np.zeros((n, 4), np.float64)
    lib().cvo_noise_sample(int(seed) & 0xffffffff, ptr(pts), n, ptr(o32), ptr(o64))
    return o32, o64


def noise_fragment(m, fp, c, r, tu, tv, ray):
    """One fragment of the noise branch (FP32): (col, alpha, steps, kept)."""
    nm = noise_march_of(m)
    inp = np.array(list(fp) + list(c) + [r, tu, tv] + list(ray), np.float32)
    out = np.zeros(4, np.float32)
    lib().cvo_noise_fragment(C.byref(nm), ptr(inp), ptr(out))
    return float(out[0]), float(out[1]), int(out[2]), bool(out[3])


def f2h(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, np.float32)
    out = np.zeros(a.shape, np.uint16)
    lib().cvo_f2h_array(ptr(a), ptr(out), a.size)
    return out


def h2f(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, np.uint16)
    out = np.zeros(a.shape, np.float32)
    lib().cvo_h2f_array(ptr(a), ptr(out), a.size)
    return out


def exp2_neg(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, np.float32)
    out = np.zeros(a.shape, np.float32)
    lib().cvo_exp2_neg_array(ptr(a), ptr(out), a.size)
    return out


def voxel_index(vol: cvo_volume, pts: np.ndarray) -> np.ndarray:
    pts = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
    out = np.zeros(pts.shape[0], np.int64)
    lib().cvo_voxel_index(C.byref(vol), ptr(pts), pts.shape[0], ptr(out))
    return out


def sun_update(vol: cvo_volume, sun_pos) -> cvo_lightview:
    lv = cvo_lightview()
    lib().cvo_sun_update(C.byref(vol), f3(sun_pos), C.byref(lv))
    return lv


def step_table(cone: cvo_cone, levels: int, dim: int):
    n = cone.steps
    h = np.zeros(n, np.float32)
    r = np.zeros(n, np.float32)
    lvl = np.zeros(n, np.int32)
    fr = np.zeros(n, np.float32)
    lib().cvo_step_table(C.byref(cone), levels, dim, ptr(h), ptr(r), ptr(lvl), ptr(fr))
    return h, r, lvl, fr


def sample(vol: cvo_volume, nat, f64, lam: float, pts):
    pts = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
    ch = 1 if vol.format == 0 else 4
    o32 = np.zeros((pts.shape[0], ch), np.float32)
    o64 = np.zeros((pts.shape[0], ch), np.float64) if f64 is not None else None
    lib().cvo_sample(C.byref(vol), ptr(nat), ptr(f64), float(lam), ptr(pts), pts.shape[0], ptr(o32), ptr(o64))
    return o32, o64


def trace_one(vol: cvo_volume, nat, f64, cone: cvo_cone, p0, d):
    o32 = np.zeros(4, np.float32)
    o64 = np.zeros(4, np.float64)
    n = C.c_int32()
    lib().cvo_trace_one(C.byref(vol), ptr(nat), ptr(f64), C.byref(cone), f3(p0), f3(d), ptr(o32),
                        ptr(o64 if f64 is not None else None), C.byref(n))
    return o32, (o64 if f64 is not None else None), int(n.value)


def noise_voxels(vol: cvo_volume, xyz_int: np.ndarray):
    xyz_int = np.ascontiguousarray(xyz_int, np.int32).reshape(-1, 3)
    o32 = np.zeros(xyz_int.shape[0], np.float32)
    o64 = np.zeros(xyz_int.shape[0], np.float64)
    lib().cvo_noise_voxels(C.byref(vol), ptr(xyz_int), xyz_int.shape[0], ptr(o32), ptr(o64))
    return o32, o64


# ---- learned far field ------------------------------------------------------

def segments(vol: cvo_volume, nat, cone: cvo_cone, light: cvo_light, k: int, p0):
    """Features (n, 17), targets (n, 5), valid (n,) of cones started at p0 (n, 3)."""
    p0 = np.ascontiguousarray(p0, np.float32).reshape(-1, 3)
    n = p0.shape[0]
    feat = np.zeros((n, 17), np.float32)
    tgt = np.zeros((n, 5), np.float32)
    valid = np.zeros(n, np.int32)
    rc = lib().cvo_segments(C.byref(vol), ptr(nat), C.byref(cone), C.byref(light), int(k), ptr(p0), n, ptr(feat),
                            ptr(tgt), ptr(valid))
    if rc != 0:
        raise ValueError("cvo_segments rejected its arguments")
    return feat, tgt, valid


def mlp_forward(params, x, emulate=1):
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32).reshape(-1, 17)
    y = np.zeros((x.shape[0], 5), np.float32)
    lib().cvo_mlp_forward(ptr(params), ptr(x), x.shape[0], int(emulate), ptr(y))
    return y


def mlp_grad(params, x, t, emulate=1, e_fixed=None):
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32).reshape(-1, 17)
    t = np.ascontiguousarray(t, np.float32).reshape(-1, 5)
    ef = None if e_fixed is None else np.ascontiguousarray(e_fixed, np.float32)
    g = np.zeros(lib().cvo_mlp_param_count(), np.float64)
    loss = C.c_double()
    lib().cvo_mlp_grad(ptr(params), ptr(x), ptr(t), x.shape[0], int(emulate), ptr(ef), ptr(g), C.byref(loss))
    return g, float(loss.value)
