"""Pins of the CPU oracle (oracle/cv_oracle.cpp) that do not restate its code.

The reference has no tests, fixtures or golden vectors, so each oracle function
is checked against what the mathematics fixes: closed forms, brute force on
tiny inputs, invariants and textbook routines (numpy's IEEE binary16,
math.exp2, a literal selection sort).  Each check is chosen so that a plausible
mistake (dropped term, wrong sign or index, transposed operand) fails.
"""
import math

import numpy as np
import pytest

import oracle_lib as O

F = np.float32


def vol(N=32, L=6, fmt=1, mode=1, lo=(-1, -1, -1), hi=(1, 1, 1), pos=(0, 0, 0), noise=None):
    return O.make_volume(N, L, fmt, mode, lo, hi, pos, noise)


# ---------------------------------------------------------------------------
# index map (res/first_voxelize.glsl:23-37)
# ---------------------------------------------------------------------------
def test_oracle_index_map():
    v = vol(N=32)
    vs = 2.0 / 32
    pts = np.array([
        [-1, -1, -1],                       # lo -> voxel 0
        [1, 0, 0],                          # x = hi -> N, dropped
        [-1.01, -1, -1],                    # lerp -0.16 truncates to 0 (kept)
        [-1 - 1.2 * vs, -1, -1],            # lerp -1.2 -> dropped
        [-1 + 5 * vs + 1e-4, -1 + 7.5 * vs, -1 + 2.9 * vs],   # -> (5, 7, 2)
        [0.999, 0.999, 0.999],              # -> (31, 31, 31)
    ], dtype=F)
    got = O.voxel_index(v, pts)
    N = 32
    assert got[0] == 0
    assert got[1] == -1
    assert got[2] == 0
    assert got[3] == -1
    assert got[4] == 5 + N * (7 + N * 2)
    assert got[5] == 31 + N * (31 + N * 31)
    # world bounds are position + bounds (src/Shaders/VoxelizeShader.cpp:124-126)
    v2 = vol(N=32, pos=(10, 0, 0))
    assert O.voxel_index(v2, np.array([[9.0 + 5 * vs + 1e-4, -1, -1]], F))[0] == 5
    assert O.voxel_index(v2, np.array([[0.0, 0.0, 0.0]], F))[0] == -1


# ---------------------------------------------------------------------------
# binary16 conversions
# ---------------------------------------------------------------------------
def test_oracle_half_conversion():
    rng = np.random.default_rng(0)
    x = np.concatenate([
        rng.standard_normal(20000).astype(F) * F(100),
        rng.uniform(-1, 1, 20000).astype(F) * F(1e-4),          # subnormal halves
        np.array([0.0, -0.0, 65504, 65519.9, 65520, 1e6, -1e6, np.inf, -np.inf, 2 ** -24, 2 ** -25,
                  3 * 2 ** -26, 2 ** -14, 1 + 2 ** -11, 1 + 3 * 2 ** -11, 0.1, 1 / 3], F),
    ])
    # exact ties: halfway between two halves (round to even)
    h = np.arange(0x0001, 0x7bff, 37, dtype=np.uint16).view(np.float16).astype(F)
    nxt = (np.arange(0x0001, 0x7bff, 37, dtype=np.uint16) + 1).view(np.float16).astype(F)
    ties = ((h.astype(np.float64) + nxt.astype(np.float64)) / 2).astype(F)
    x = np.concatenate([x, ties])
    with np.errstate(over="ignore"):
        want = x.astype(np.float16).view(np.uint16)
    got = O.f2h(x)
    assert np.array_equal(got, want)
    allh = np.arange(65536, dtype=np.uint32).astype(np.uint16)
    back = O.h2f(allh)
    ref = allh.view(np.float16).astype(F)
    nan = np.isnan(ref)
    assert np.array_equal(np.isnan(back), nan)
    assert np.array_equal(back[~nan].view(np.uint32), ref[~nan].view(np.uint32))


def test_oracle_exp2():
    x = np.concatenate([np.linspace(-130, 0, 20001), -np.logspace(-8, 2, 2000)]).astype(F)
    got = O.exp2_neg(x).astype(np.float64)
    want = np.exp2(x.astype(np.float64))
    want[x <= -126] = 0.0
    ok = want > 0
    rel = np.abs(got[ok] - want[ok]) / want[ok]
    assert rel.max() < 4e-7
    assert np.all(got[~ok] == 0)
    assert O.exp2_neg(np.array([0.0], F))[0] == 1.0


# ---------------------------------------------------------------------------
# light view (src/Sun.hpp:26-43)
# ---------------------------------------------------------------------------
def _mat(lv):
    return np.array(lv.V[:], np.float64).reshape(4, 4).T  # column-major -> row-major


@pytest.mark.parametrize("sun", [(10, 0, 0), (3, 4, -2), (0, 10, 0), (0.5, -7, 0.3)])
def test_oracle_sun_update(sun):
    v = vol(N=32)
    lv = O.sun_update(v, sun)
    M = _mat(lv)
    R = M[:3, :3]
    assert np.allclose(R @ R.T, np.eye(3), atol=1e-6)
    assert abs(np.linalg.det(R) - 1) < 1e-6
    d = -np.asarray(sun, np.float64) / np.linalg.norm(sun)      # lookDir toward the volume
    look_pos = -d * math.sqrt(3)
    # the eye maps to the origin and the volume centre lies straight ahead at sqrt(3)
    assert np.allclose(M @ np.append(look_pos, 1), [0, 0, 0, 1], atol=1e-5)
    assert np.allclose(M @ np.array([0, 0, 0, 1]), [0, 0, -math.sqrt(3), 1], atol=1e-5)
    assert np.allclose(np.array(lv.near_plane[:]), look_pos + d * 0.01, atol=1e-6)
    assert abs(lv.clip - (2 * math.sqrt(3) - 0.01)) < 1e-5
    assert list(lv.ortho) == [-2, 2, -2, 2]
    assert abs(lv.zfar - (0.01 + lv.clip)) < 1e-6


# ---------------------------------------------------------------------------
# pass 1: analytic hemisphere under an axis-aligned sun
# ---------------------------------------------------------------------------
class _Scene:
    def __init__(self, xyz, r, sun_pos, W=64, H=48, N=32, mode=1, fmt=1):
        self.xyz = np.asarray(xyz, F).reshape(-1, 3)
        self.scales = np.asarray(r, F).reshape(-1)
        self.fluffiness = 1.0
        self.N, self.levels, self.fmt, self.mode = N, 6, fmt, mode
        self.bounds_lo, self.bounds_hi, self.position = (-1, -1, -1), (1, 1, 1), (0, 0, 0)
        self.noise = None
        self.light_w, self.light_h = W, H
        self._sun = np.asarray(sun_pos, F)
        self.sun_dir = self._sun / np.linalg.norm(self._sun)
        self.directional = 1

    @property
    def S(self):
        return self.xyz.shape[0]

    @property
    def sun_pos(self):
        return self._sun


def test_oracle_light_hemisphere():
    r = 0.5
    sc = _Scene([[0.1, -0.2, 0.0]], [r], (0, 0, 10), W=96, H=80)
    res = O.voxelize(sc)
    W, H = sc.light_w, sc.light_h
    # light looks down -z with view x = world x, view y = world y (lookAt with up = +y)
    px = (np.arange(W) + 0.5) * 2 / W - 1
    py = (np.arange(H) + 0.5) * 2 / H - 1
    X, Y = np.meshgrid(px * 2, py * 2)        # ortho window [-2, 2]^2
    dx, dy = X - 0.1, Y + 0.2
    q = np.sqrt(dx * dx + dy * dy) / r
    sc64 = np.sqrt(np.maximum(0, 1 - q * q))
    covered = sc64 >= 0.01
    margin = np.abs(sc64 - 0.01) < 1e-5
    got = res.winner >= 0
    assert np.array_equal(got[~margin], covered[~margin])
    assert got.sum() > 0.9 * math.pi * (r / 4 * W) * (r / 4 * H)
    # the position map holds the hemisphere point facing the sun: z = +sqrt(r^2 - d^2)
    pm = res.posmap[got]
    zt = np.sqrt(np.maximum(0, r * r - dx[got] ** 2 - dy[got] ** 2))
    assert np.allclose(pm[:, 0], X[got], atol=2e-6)
    assert np.allclose(pm[:, 1], Y[got], atol=2e-6)
    assert np.allclose(pm[:, 2], zt, atol=2e-5)
    assert np.all(res.posmap[got][:, 3] == 1) and np.all(res.posmap[~got][:, 3] == 0)


def test_oracle_depth_ties():
    # identical spheres: the first drawn wins every tie (GL_LESS)
    sc = _Scene([[0, 0, 0], [0, 0, 0]], [0.4, 0.4], (0, 0, 10))
    res = O.voxelize(sc)
    assert set(np.unique(res.winner)) == {-1, 0}
    # the sphere nearer to the sun wins where both cover, in either draw order
    for order in ([0, 1], [1, 0]):
        xyz = np.array([[0, 0, 0.3], [0, 0, -0.3]], F)[order]
        sc = _Scene(xyz, [0.3, 0.3], (0, 0, 10))
        res = O.voxelize(sc)
        near = order.index(0)
        assert set(np.unique(res.winner)) == {-1, near}
    # concentric: the larger sphere's cap is nearer the sun everywhere
    sc = _Scene([[0, 0, 0], [0, 0, 0]], [0.2, 0.4], (0, 0, 10))
    res = O.voxelize(sc)
    assert set(np.unique(res.winner)) == {-1, 1}


# ---------------------------------------------------------------------------
# pass 2 and chord fill: brute force over the fragments
# ---------------------------------------------------------------------------
def _brute_index(p, N=32):
    """calculateVoxelIndex in float32 for world bounds [-1, 1]^3; -1 when dropped."""
    p = p.astype(F)
    lerp = F(N) * ((p - F(-1)) / F(2))
    ok = np.all((lerp > -1) & (lerp < N), axis=-1)
    ijk = np.trunc(lerp).astype(np.int64)
    lin = ijk[..., 0] + N * (ijk[..., 1] + N * ijk[..., 2])
    return np.where(ok, lin, -1)


def test_oracle_splat_bruteforce():
    rng = np.random.default_rng(3)
    xyz = rng.uniform(-0.5, 0.5, (6, 3))
    sc = _Scene(xyz, rng.uniform(0.1, 0.25, 6), (3, 7, 2), W=80, H=64, mode=0)
    res = O.voxelize(sc)
    step = F(2.0 / 32)
    k = F(1.0) / np.sqrt(F(3.0))
    wp = res.posmap[res.winner >= 0][:, :3].astype(F)
    want = set()
    offs = [(0, 0, 0)] + [(sx, sy, sz) for sx in (1, -1) for sy in (1, -1) for sz in (1, -1)]
    for o in offs:
        d = np.array(o, F) * (step * k)
        idx = _brute_index(wp + d)
        want |= set(idx[idx >= 0].tolist())
    got = set(np.flatnonzero(res.shell.reshape(-1)).tolist())
    assert got == want
    assert res.fill.sum() == 0  # reference mode: no chord fill


def test_oracle_chord_bruteforce():
    rng = np.random.default_rng(5)
    xyz = rng.uniform(-0.4, 0.4, (3, 3))
    r = rng.uniform(0.15, 0.3, 3)
    sun = (1.0, 2.0, 0.5)
    sc = _Scene(xyz, r, sun, W=48, H=40, N=16, mode=1)
    res = O.voxelize(sc)
    v = vol(N=16)
    lv = O.sun_update(v, sc.sun_pos)
    V = np.array(lv.V[:], F).reshape(4, 4)            # V[col][row]
    S_ax = np.array([V[0, 0], V[1, 0], V[2, 0]], F)
    U_ax = np.array([V[0, 1], V[1, 1], V[2, 1]], F)
    B = np.array([V[0, 2], V[1, 2], V[2, 2]], F)
    B = (B * (F(1) / np.sqrt(np.sum(B * B, dtype=F)))).astype(F)
    W, H = sc.light_w, sc.light_h
    want = set()
    step = F(2.0 / 16)
    for j in range(3):
        c = xyz[j].astype(F)
        s = F(r[j])
        cv = np.array([((V[0, a] * c[0] + V[1, a] * c[1]) + V[2, a] * c[2]) + V[3, a] for a in range(3)], F)
        for py_ in range(H):
            for px_ in range(W):
                xn = (F(2) * (F(px_) + F(0.5))) / F(W) - F(1)
                yn = (F(2) * (F(py_) + F(0.5))) / F(H) - F(1)
                du = (xn * F(2) + F(0)) - cv[0]
                dv = (yn * F(2) + F(0)) - cv[1]
                q = np.sqrt(du * du + dv * dv) / s
                scn = np.sqrt(max(F(0), F(1) - q * q))
                if scn < F(0.01):
                    continue
                fp = (c + du * S_ax) + dv * U_ax
                dist = s * scn
                start = fp - B * dist
                i = F(0)
                while i < F(2) * dist:
                    idx = _brute_index((start + B * i)[None, :], N=16)[0]
                    if idx >= 0:
                        want.add(int(idx))
                    i = F(i + step)
    got = set(np.flatnonzero(res.fill.reshape(-1)).tolist())
    assert got == want
    assert len(got) > 50


# ---------------------------------------------------------------------------
# mip chain (glGenerateMipmap, src/Shaders/VoxelizeShader.cpp:105)
# ---------------------------------------------------------------------------
class _VS:
    def __init__(self, N, L, fmt=1, mode=1, noise=None):
        self.N, self.levels, self.fmt, self.mode = N, L, fmt, mode
        self.bounds_lo, self.bounds_hi, self.position = (-1, -1, -1), (1, 1, 1), (0, 0, 0)
        self.noise = noise


def test_oracle_mips_constant():
    N = 16
    shell = np.ones((N, N, N), np.uint8)
    _, _, nl, fl = O.build_levels(_VS(N, 5), shell, np.zeros_like(shell))
    for lvl in nl:
        assert np.all(lvl == 1.0)


def test_oracle_mips_single_voxel():
    N = 16
    shell = np.zeros((N, N, N), np.uint8)
    shell[5, 6, 7] = 1          # z, y, x
    _, _, nl, _ = O.build_levels(_VS(N, 5), shell, np.zeros_like(shell))
    assert nl[1][2, 3, 3, 0] == 1 / 8 and np.count_nonzero(nl[1][..., 0]) == 1
    assert nl[2][1, 1, 1, 0] == 1 / 64 and np.count_nonzero(nl[2][..., 0]) == 1
    assert nl[4][0, 0, 0, 0] == 1 / 4096


def test_oracle_mips_mean_and_modes():
    rng = np.random.default_rng(7)
    N = 32
    shell = (rng.random((N, N, N)) < 0.2).astype(np.uint8)
    fill = (rng.random((N, N, N)) < 0.3).astype(np.uint8)
    _, f64, nl, fl = O.build_levels(_VS(N, 6), shell, fill)
    a0 = nl[0][..., 3].mean()
    for lvl in nl:
        assert abs(lvl[..., 3].mean() - a0) < 1e-6     # mean preserved (power-of-two N)
    # rgba semantics: shell (1,1,1,1), fill only (0,0,0,1)
    st = np.where(shell != 0, 2, np.where(fill != 0, 1, 0))
    assert np.all(nl[0][st == 2] == 1)
    assert np.all(nl[0][st == 1] == [0, 0, 0, 1])
    assert np.all(nl[0][st == 0] == 0)
    # level k is the 2x2x2 box mean of level k-1
    for k in range(1, 6):
        prev = fl[k - 1]
        n = prev.shape[0] // 2
        box = prev.reshape(n, 2, n, 2, n, 2, 4).mean(axis=(1, 3, 5))
        assert np.allclose(fl[k], box, atol=1e-12)
        assert np.allclose(nl[k], box, atol=1e-6)


def test_oracle_mips_npot_and_r8():
    rng = np.random.default_rng(9)
    N = 6       # 6 -> 3 -> 1: the last step takes all three children per axis
    shell = (rng.random((N, N, N)) < 0.5).astype(np.uint8)
    _, f64, nl, fl = O.build_levels(_VS(N, 3), shell, np.zeros_like(shell))
    l1 = shell.reshape(3, 2, 3, 2, 3, 2).mean(axis=(1, 3, 5))
    assert np.allclose(fl[1][..., 0], l1)
    assert abs(fl[2][0, 0, 0, 0] - l1.mean()) < 1e-12
    # R8 (the reference's GL_R8): 255 on the shell, integer means rounded half up
    nat, _, nl8, _ = O.build_levels(_VS(8, 4, fmt=0, mode=0), np.pad(shell, 1)[:8, :8, :8], np.zeros((8, 8, 8), np.uint8))
    s8 = np.pad(shell, 1)[:8, :8, :8].astype(np.int64) * 255
    assert np.array_equal(nl8[0], s8)
    sums = s8.reshape(4, 2, 4, 2, 4, 2).sum(axis=(1, 3, 5))
    assert np.array_equal(nl8[1], (sums + 4) // 8)


def test_oracle_fp16_levels_round_once_per_level():
    rng = np.random.default_rng(11)
    N = 16
    shell = (rng.random((N, N, N)) < 0.3).astype(np.uint8)
    fill = (rng.random((N, N, N)) < 0.3).astype(np.uint8)
    n32 = O.build_levels(_VS(N, 5, fmt=1, mode=1, noise=_noise()), shell, fill)[2]
    n16 = O.build_levels(_VS(N, 5, fmt=2, mode=1, noise=_noise()), shell, fill)[2]
    # level 0 is the FP32 value rounded to nearest binary16
    assert np.array_equal(n16[0], n32[0].astype(np.float16).view(np.uint16))
    # level k = round16(box mean of the rounded level k-1)
    for k in range(1, 5):
        prev = n16[k - 1].view(np.float16).astype(np.float32)
        n = prev.shape[0] // 2
        box = prev.reshape(n, 2, n, 2, n, 2, 4).astype(np.float64).sum(axis=(1, 3, 5)) / 8
        assert np.array_equal(n16[k], box.astype(np.float32).astype(np.float16).view(np.uint16))


def _noise(octaves=4, seed=5, base=4.0):
    class N_:
        pass
    n = N_()
    n.octaves, n.seed, n.base_freq, n.lacunarity, n.gain = octaves, seed, base, 2.0, 0.5
    return n


def _fmix32(h):
    h = np.uint32(h)
    h ^= h >> np.uint32(16)
    h = np.uint32((int(h) * 0x85EBCA6B) & 0xFFFFFFFF)
    h ^= h >> np.uint32(13)
    h = np.uint32((int(h) * 0xC2B2AE35) & 0xFFFFFFFF)
    h ^= h >> np.uint32(16)
    return int(h)


def test_oracle_value_noise_lattice():
    # one octave with the lattice on voxel centres: (x + 0.5) / 8 * 16 = 2x + 1
    v = vol(N=8, noise=_noise(octaves=1, seed=123, base=16.0))
    pts = np.array([[x, y, z] for x in range(8) for y in range(0, 8, 3) for z in (0, 7)], np.int32)
    n32, n64 = O.noise_voxels(v, pts)
    for (x, y, z), a, b in zip(pts.tolist(), n32, n64):
        h = 123
        h ^= ((2 * x + 1) * 0x8DA6B343) & 0xFFFFFFFF
        h ^= ((2 * y + 1) * 0xD8163841) & 0xFFFFFFFF
        h ^= ((2 * z + 1) * 0xCB1AB31F) & 0xFFFFFFFF
        want = (_fmix32(h) >> 8) / 16777216.0
        assert a == np.float32(want) and b == want
    # 4 octaves: values in [0, 1), FP32 agrees with FP64
    v4 = vol(N=32, noise=_noise())
    pts = np.random.default_rng(0).integers(0, 32, (500, 3)).astype(np.int32)
    n32, n64 = O.noise_voxels(v4, pts)
    assert np.all((n64 >= 0) & (n64 < 1)) and np.abs(n32 - n64).max() < 1e-6
    assert 0.3 < n64.mean() < 0.7


# ---------------------------------------------------------------------------
# sampler (textureLod, GL 4.4 8.14)
# ---------------------------------------------------------------------------
def _levels_from(l0):
    """linear RGBA32F levels with exact box means, for the sampler tests"""
    lv = [l0]
    while lv[-1].shape[0] > 1:
        p = lv[-1]
        n = p.shape[0] // 2
        lv.append(p.reshape(n, 2, n, 2, n, 2, 4).mean(axis=(1, 3, 5)))
    return lv


def test_oracle_sampler_centres_affine_clamp():
    N = 16
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    l0 = np.stack([0.25 + 0.01 * x, 0.5 - 0.02 * y, 0.1 + 0.03 * z, 0.3 + 0.01 * x + 0.02 * y], -1)
    lv = _levels_from(l0)
    v = vol(N=N, L=len(lv))
    nat = np.concatenate([a.reshape(-1) for a in lv]).astype(F)
    f64 = np.concatenate([a.reshape(-1) for a in lv]).astype(np.float64)
    # texel centres return the texel
    idx = np.array([[0, 0, 0], [3, 7, 11], [15, 15, 15], [8, 2, 5]])
    pts = (idx + 0.5) / N
    o32, o64 = O.sample(v, nat, f64, 0.0, pts)
    assert np.allclose(o64, l0[idx[:, 2], idx[:, 1], idx[:, 0]], atol=1e-12)
    # exact on an affine field inside the volume (level 0)
    rng = np.random.default_rng(1)
    p = rng.uniform(0.5 / N, 1 - 0.5 / N, (200, 3)).astype(F)  # the sampler takes FP32 coordinates
    o32, o64 = O.sample(v, nat, f64, -3.0, p)
    u = p.astype(np.float64) * N - 0.5
    want = np.stack([0.25 + 0.01 * u[:, 0], 0.5 - 0.02 * u[:, 1], 0.1 + 0.03 * u[:, 2],
                     0.3 + 0.01 * u[:, 0] + 0.02 * u[:, 1]], -1)
    assert np.abs(o64 - want).max() < 1e-12 and np.abs(o32 - want).max() < 1e-5
    # clamp to edge: beyond the border the edge texel is returned
    o32, o64 = O.sample(v, nat, f64, 0.0, np.array([[-0.3, 0.5 / N, 1.7]]))
    assert np.allclose(o64[0], l0[N - 1, 0, 0], atol=1e-12)
    # lambda >= L-1: top level only; lambda = 1.5: halfway between levels 1 and 2
    o32, o64 = O.sample(v, nat, f64, 9.0, np.array([[0.3, 0.6, 0.1]]))
    assert np.allclose(o64[0], lv[-1][0, 0, 0], atol=1e-12)
    c = np.array([[0.5, 0.5, 0.5]])
    a, _ = O.sample(v, nat, f64, 1.0, c)
    b, _ = O.sample(v, nat, f64, 2.0, c)
    m, _ = O.sample(v, nat, f64, 1.5, c)
    assert np.allclose(m, (a + b) / 2, atol=1e-6)


# ---------------------------------------------------------------------------
# cone (res/conetrace_frag.glsl:64-79) and the front-to-back extension
# ---------------------------------------------------------------------------
def _cone(steps=16, angle=0.9, h0=0.1, mode=0, alpha_stop=0.99, sigma=24.0, ds=1.0, comp=0):
    k = O.cvo_cone()
    k.steps, k.cone_angle, k.initial_height, k.lod_offset, k.down_scaling = steps, angle, h0, 0.0, ds
    k.cone_mode, k.alpha_stop, k.sigma, k.composite, k.quantize_rgba8, k.compensated = mode, alpha_stop, sigma, 1, 0, comp
    k.clear_color = O.F4(0, 0, 0, 0)
    return k


def _const_levels(N, L, rgba):
    dims = [max(1, N >> k) for k in range(L)]
    nat = np.concatenate([np.tile(np.asarray(rgba, F), d ** 3) for d in dims]).astype(F)
    return nat, nat.astype(np.float64)


def test_oracle_cone_constant_weighted():
    N, L = 16, 5
    nat, f64 = _const_levels(N, L, (0.6, 0.6, 0.6, 0.6))
    v = vol(N=N, L=L)
    for steps, ds in ((16, 1.0), (5, 2.0), (64, 0.5)):
        o32, o64, n = O.trace_one(v, nat, f64, _cone(steps=steps, ds=ds), (0.5, 0.5, 0.5), (0.01, 0.0, 0.0))
        assert n == steps
        v06 = float(F(0.6))                                           # the stored texel value
        assert abs(o64[0] - v06 * (steps + 1) / (2 * ds)) < 1e-9     # sum of v * i / (steps * ds)
        assert abs(o32[0] - o64[0]) < 1e-4


def test_oracle_cone_ftb_constant_alpha():
    N, L = 16, 5
    for alpha in (0.3, 0.05):
        nat, f64 = _const_levels(N, L, (0.2, 0.4, 0.8, alpha))
        v = vol(N=N, L=L)
        o32, o64, n = O.trace_one(v, nat, f64, _cone(steps=200, mode=1, h0=0.001, angle=1e-4), (0.5, 0.5, 0.5),
                                  (1e-6, 0, 0))
        a32 = float(F(alpha))                                         # stored texel values
        rgb = np.array([F(0.2), F(0.4), F(0.8)], np.float64)
        n_expect = math.ceil(math.log(0.01) / math.log(1 - a32))
        assert n == n_expect
        A = 1 - (1 - a32) ** n
        assert abs(o64[3] - A) < 1e-12
        assert np.allclose(o64[:3], rgb * A, atol=1e-12)


def test_oracle_cone_density_closed_form():
    N, L = 16, 5
    nat, f64 = _const_levels(N, L, (1.0, 1.0, 1.0, 1.0))
    v = vol(N=N, L=L)
    cone = _cone(steps=20, mode=2, sigma=3.0, alpha_stop=2.0, angle=0.5)
    h, r, lvl, fr = O.step_table(cone, L, N)
    # stays inside: start at the centre, tiny direction
    o32, o64, n = O.trace_one(v, nat, f64, cone, (0.5, 0.5, 0.5), (1e-7, 0, 0))
    assert n == 20
    tan_half = math.tan(float(F(0.5)) / 2)      # parameters arrive as FP32
    hh, tau = float(F(0.1)), 0.0
    for i in range(20):
        rr = hh * tan_half
        tau += 3.0 * rr / N
        hh += rr
    assert abs(o64[3] - (1 - math.exp(-tau))) < 1e-9
    assert abs(o32[3] - o64[3]) < 1e-5


def test_oracle_cone_empty_and_exit():
    N, L = 16, 5
    nat, f64 = _const_levels(N, L, (0, 0, 0, 0))
    v = vol(N=N, L=L)
    o32, o64, n = O.trace_one(v, nat, f64, _cone(mode=1, steps=50), (0.5, 0.5, 0.5), (0.1, 0, 0))
    assert np.all(o64 == 0)
    # front-to-back stops before the first sample outside [0,1]^3
    cone = _cone(mode=1, steps=50, h0=1.0, angle=0.5)
    h, _, _, _ = O.step_table(cone, L, N)
    o32, o64, n = O.trace_one(v, nat, f64, cone, (0.5, 0.5, 0.5), (1.0 / N, 0, 0))
    inside = [0.5 + hh / N <= 1.0 for hh in h]
    assert n == inside.index(False)


def test_oracle_step_table():
    cone = _cone(steps=40, angle=math.radians(20), h0=0.1)
    h, r, lvl, fr = O.step_table(cone, 9, 256)
    t = math.tan(math.radians(10))
    for i in range(40):
        assert abs(h[i] - 0.1 * (1 + t) ** i) <= 1e-5 * h[i]
        lam = math.log2(max(1.0, 2 * r[i]))
        if lam <= 0:
            assert lvl[i] == 0 and fr[i] == 0
        elif lam >= 8:
            assert lvl[i] == 8 and fr[i] == 0
        else:
            assert lvl[i] == math.floor(lam) and abs(fr[i] - (lam - math.floor(lam))) < 1e-5


# ---------------------------------------------------------------------------
# sortBoards (src/CloudVolume.cpp:65-82)
# ---------------------------------------------------------------------------
def _literal_selection(d):
    d = list(d)
    idx = list(range(len(d)))
    for i in range(len(d)):
        m = i
        for j in range(i + 1, len(d)):
            if d[m] < d[j]:
                m = j
        d[i], d[m] = d[m], d[i]
        idx[i], idx[m] = idx[m], idx[i]
    return idx


class _SortScene:
    def __init__(self, xyz):
        self.xyz = np.asarray(xyz, F)
        self.position = (0, 0, 0)

    @property
    def S(self):
        return self.xyz.shape[0]


def test_oracle_sort_distinct_and_ties():
    rng = np.random.default_rng(4)
    xyz = rng.uniform(-1, 1, (300, 3)).astype(F)
    eye = (0.3, 0.1, 3.0)
    order = O.sort_boards(_SortScene(xyz), eye)
    d = np.sqrt(np.sum((np.asarray(eye, np.float64) - xyz.astype(np.float64)) ** 2, axis=1))
    assert np.array_equal(order, np.argsort(-d, kind="stable"))
    # ties: four-fold rotations about the camera axis give equal FP32 distances
    base = rng.uniform(-0.5, 0.5, (12, 3)).astype(F)
    pts = []
    for x, y, z in base:
        pts += [(x, y, z), (-y, x, z), (-x, -y, z), (y, -x, z)]
    pts = np.array(pts, F)[rng.permutation(48)]
    order = O.sort_boards(_SortScene(pts), (0, 0, 3))
    d32 = np.zeros(48, F)
    O.lib().cvo_board_distances(48, O.ptr(np.ascontiguousarray(pts)), O.f3((0, 0, 0)), O.f3((0, 0, 3)), O.ptr(d32))
    assert len(np.unique(d32)) < 48
    assert list(order) == _literal_selection(d32.tolist())


# ---------------------------------------------------------------------------
# soft-particle noise march (res/conetrace_frag.glsl:103-174) and its map
# (src/Shaders/ConeTraceShader.cpp:100-159)
# ---------------------------------------------------------------------------
def _hash_u(seed, x, y, z):
    h = seed
    h ^= (x * 0x8DA6B343) & 0xFFFFFFFF
    h ^= (y * 0xD8163841) & 0xFFFFFFFF
    h ^= (z * 0xCB1AB31F) & 0xFFFFFFFF
    return (_fmix32(h) >> 8) / 16777216.0


def test_oracle_noise_map_alpha():
    m = O.noise_map(7)
    a = m[..., 3].astype(np.int64)
    # (char)genRandom(-128, 128): truncation toward zero of u * 256 - 128
    for (x, y, z) in [(0, 0, 0), (31, 0, 5), (3, 17, 29), (31, 31, 31)]:
        v = np.float32(np.float32(_hash_u(7, x, y, z)) * np.float32(256.0)) + np.float32(-128.0)
        assert a[z, y, x] == int(np.trunc(v))
    # uniform bytes over [-128, 127]: every value from -127 to 126 occurs, mean ~ -0.5
    vals = set(np.unique(a).tolist())
    assert set(range(-127, 127)) <= vals and vals <= set(range(-128, 128))
    assert abs(a.mean() + 0.5) < 2.0
    assert not np.array_equal(m, O.noise_map(8))


def test_oracle_noise_map_normals():
    m = O.noise_map(3)
    d = m[..., 3].astype(np.float64) / 128.0                       # [z, y, x]
    def nb(ax, s):
        return np.roll(d, -s, axis=ax)
    # as written: d(+1) - d(-1) / heightAdjust (the division binds to the second term)
    g = np.stack([nb(2, 1) - nb(2, -1) / 0.5, nb(1, 1) - nb(1, -1) / 0.5, nb(0, 1) - nb(0, -1) / 0.5], -1)
    central = np.stack([nb(2, 1) - nb(2, -1), nb(1, 1) - nb(1, -1), nb(0, 1) - nb(0, -1)], -1)
    n = m[..., :3].astype(np.float64) / 128.0
    ok = ~np.any(m[..., :3] == -128, axis=-1)   # +-1.0 components wrap to -128
    assert ok.mean() > 0.99
    ln = np.linalg.norm(n, axis=-1)
    assert np.all(np.abs(ln[ok] - 1.0) < 0.03)
    cos_g = np.sum(n * g, -1) / (ln * np.linalg.norm(g, axis=-1))
    cos_c = np.sum(n * central, -1) / (ln * np.linalg.norm(central, axis=-1))
    assert cos_g[ok].min() > 0.99
    assert np.mean(cos_c[ok] < 0.99) > 0.5   # the central difference points elsewhere
    # alpha bytes are untouched by the normal pass
    assert np.array_equal(m[..., 3], O.noise_map(3)[..., 3])


def _gl_repeat_trilinear(m, p):
    """GL_LINEAR + GL_REPEAT, weight-product form, FP64 (GL 4.4 eq. 8.10)."""
    t = np.maximum(m.astype(np.float64) / 127.0, -1.0)             # [z, y, x, c]
    u = np.asarray(p, np.float64) * 32.0 - 0.5
    i0 = np.floor(u).astype(np.int64)
    a = u - i0
    out = np.zeros(4)
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                w = (a[0] if dx else 1 - a[0]) * (a[1] if dy else 1 - a[1]) * (a[2] if dz else 1 - a[2])
                out += w * t[(i0[2] + dz) % 32, (i0[1] + dy) % 32, (i0[0] + dx) % 32]
    return out


def test_oracle_noise_sample():
    seed = 11
    m = O.noise_map(seed)
    rng = np.random.default_rng(1)
    pts = np.concatenate([rng.uniform(-3, 3, (300, 3)), rng.uniform(0, 1, (100, 3)),
                          [[0.5 / 32, 0.5 / 32, 0.5 / 32], [31.5 / 32, 1.5 / 32, -0.5 / 32], [0.0, 1.0, 2.0],
                           [-1e-9, 1e-9, 33.0]]]).astype(np.float32)
    o32, o64 = O.noise_sample(seed, pts)
    for p, a, b in zip(pts, o32, o64):
        want = _gl_repeat_trilinear(m, p.astype(np.float64))
        assert np.abs(b - want).max() < 1e-12
        assert np.abs(a - want).max() < 1e-5   # FP32 weights: u = 32 s - 0.5 holds ~2^-17 of fraction at |s| = 3
    # texel centres return the texel; wrap-around joins texel 31 to texel 0
    c = O.noise_sample(seed, np.array([[3.5 / 32, 7.5 / 32, 30.5 / 32]], np.float32))[0][0]
    assert np.array_equal(c, np.maximum(m[30, 7, 3].astype(np.float32) / np.float32(127), -1))
    edge = O.noise_sample(seed, np.array([[0.0, 0.5 / 32, 0.5 / 32]], np.float32))[1][0]
    t = np.maximum(m.astype(np.float64) / 127.0, -1.0)
    assert np.abs(edge - 0.5 * (t[0, 0, 31] + t[0, 0, 0])).max() < 1e-12


def _noise_fragment_f64(m, P, fp, c, r, tu, tv, ray):
    """res/conetrace_frag.glsl:136-172 restated in FP64 from the GLSL text."""
    fp, c, ray = (np.asarray(v, np.float64) for v in (fp, c, ray))
    delta = fp - c
    A = ray @ ray
    B = 2 * (delta @ ray)
    Cc = delta @ delta - r * r
    disc = B * B - 4.0 * A * Cc
    if disc < 0.01:
        return None
    sq = math.sqrt(disc)
    tnear, tfar = (-B - sq) / (2 * A), (-B + sq) / (2 * A)
    wn, wf = fp + ray * tnear, fp + ray * tfar
    unit = (wn - c) / r
    ltn, ltf = wn / P.adjust_size, wf / P.adjust_size
    it = np.linalg.norm(ltf - ltn) / P.step_size
    it = min(it, P.max_steps - P.min_steps) + P.min_steps
    cur = ltn.copy()
    dt = (ltf - ltn) / (it - 1)
    ro = rl = 0.0
    i = 0
    while i < it:
        nv = np.zeros(4)
        freq = pers = 1.0
        for o in range(P.octaves):
            off = P.offsets[o] if o < 3 else 0.0
            nv += pers * _gl_repeat_trilinear(m, (cur + off) * freq)
            freq *= P.freq_step
            pers *= P.pers_step
        nv[3] = abs(nv[3])
        nxyz = nv[:3] + unit / np.linalg.norm(unit)
        ro += nv[3] * (1 - unit @ unit)
        rl += min(max(nxyz[1] * 0.5 + 0.5, 0.0), 1.0)
        cur += dt
        unit += dt
        i += 1
    col = P.min_color + P.color_scale * rl / (it - 1)
    alpha = 1 - math.hypot((tu + 1) / 2 - 0.5, (tv + 1) / 2 - 0.5) * 2
    return col, min(max(ro * P.opacity / (it - 1), 0.0), 1.0) * alpha, i


def _march(**kw):
    from types import SimpleNamespace
    base = dict(seed=5, step_size=0.01, opacity=4.0, octaves=4, freq_step=3.0, pers_step=0.5, adjust_size=40.0,
                min_steps=2, max_steps=8, min_color=0.2, color_scale=0.45, offsets=(0.0, 0.0, 0.0))
    base.update(kw)
    return SimpleNamespace(**base)


@pytest.mark.parametrize("kw", [{}, dict(offsets=(0.013, -0.02, 0.005), octaves=5, min_steps=1, max_steps=12,
                                         adjust_size=0.7, step_size=0.05)])
def test_oracle_noise_fragment(kw):
    P = _march(**kw)
    m = O.noise_map(P.seed)
    rng = np.random.default_rng(4)
    kept = dropped = 0
    for _ in range(40):
        c = rng.uniform(-0.6, 0.6, 3).astype(np.float32)
        r = np.float32(rng.uniform(0.05, 0.3))
        ray = rng.normal(size=3)
        ray = (ray / np.linalg.norm(ray)).astype(np.float32)
        # a point of the billboard plane through c, perpendicular to the ray
        e1 = np.cross(ray, [0.0, 0.0, 1.0] if abs(ray[2]) < 0.9 else [1.0, 0.0, 0.0])
        e1 /= np.linalg.norm(e1)
        e2 = np.cross(ray, e1)
        tu, tv = rng.uniform(-1, 1, 2).astype(np.float32)
        fp = (c + r * tu * e1 + r * tv * e2).astype(np.float32)
        col, alpha, steps, ok = O.noise_fragment(P, fp, c, float(r), float(tu), float(tv), ray)
        want = _noise_fragment_f64(m, P, fp, c, float(r), float(tu), float(tv), ray)
        if want is None or not ok:
            # the discard decision (disc < 0.01) agrees away from its boundary
            assert want is None and not ok or abs(4 * (float(r) ** 2 - float(tu * tu + tv * tv) * float(r) ** 2)
                                                   - 0.01) < 1e-4
            dropped += 1
            continue
        kept += 1
        assert steps == want[2]
        assert abs(col - want[0]) < 1e-4 and abs(alpha - want[1]) < 1e-4, (col, alpha, want)
        assert 0.0 <= alpha <= 1.0 and col >= P.min_color
    assert kept >= 10 and dropped >= 3


# ---------------------------------------------------------------------------
# sun sprite (res/sun_frag.glsl:14-28, src/Shaders/SunShader.cpp:7-42)
# ---------------------------------------------------------------------------
def test_oracle_sun_sprite(cv):
    """Empty scene, float GL blend: the image is the sprite over the clear
    colour, restated in FP64 from the GLSL text (glm::lookAt + perspective,
    pixel ray hit with the sprite's view-aligned plane)."""
    from dataclasses import replace
    from gpu_common import run_oracle
    sc = cv.scenes
    sun_dir = np.array([0.3, -0.2, -1.0]) / np.linalg.norm([0.3, -0.2, -1.0])
    sp = sc.SunSprite(inner_color=(0.9, 0.5, 0.1), outer_color=(0.2, 0.4, 1.0), inner_radius=0.8, outer_radius=2.5)
    base = sc.small32(0)
    scene = replace(sc.reference_mode(base), xyz=np.zeros((0, 3), np.float32), scales=np.zeros(0, np.float32),
                    sun_dir=sun_dir.astype(np.float32), sun_sprite=sp, W=96, H=64, light_w=32, light_h=32,
                    cone=replace(sc.reference_mode(base).cone, quantize_rgba8=0))
    o = run_oracle(scene)
    cam = scene.cameras[0]
    eye, tgt, up = (np.asarray(v, np.float64) for v in (cam.eye, cam.target, cam.up))
    f = (tgt - eye) / np.linalg.norm(tgt - eye)
    s = np.cross(f, up); s /= np.linalg.norm(s)
    u = np.cross(s, f)
    p = np.asarray(scene.sun_pos, np.float64)
    sv = np.array([s @ (p - eye), u @ (p - eye), -(f @ (p - eye))])     # view space
    th = math.tan(cam.fovy / 2)
    clear = np.asarray(scene.cone.clear_color, np.float64)
    want = np.tile(clear, (scene.H, scene.W, 1))
    inside = boundary = 0
    for py in range(scene.H):
        for px in range(scene.W):
            rx = ((2 * (px + 0.5)) / scene.W - 1) * th * scene.W / scene.H
            ry = ((2 * (py + 0.5)) / scene.H - 1) * th
            dz = -sv[2]
            du, dv = rx * dz - sv[0], ry * dz - sv[1]
            if max(abs(du), abs(dv)) > sp.outer_radius:
                continue
            dist = math.hypot(du, dv)
            if dist < sp.inner_radius:
                q, a = np.asarray(sp.inner_color), 1.0
            else:
                k = (dist - sp.inner_radius) / (sp.outer_radius - sp.inner_radius)
                if abs(k - 0.99) < 1e-4 or abs(dist - sp.inner_radius) < 1e-5:
                    boundary += 1
                    want[py, px] = o.img64[py, px]   # FP32 decision at the edge
                    continue
                if k > 0.99:
                    continue
                q, a = np.asarray(sp.outer_color) * k + np.asarray(sp.inner_color) * (1 - k), 1 - k
            inside += 1
            want[py, px, :3] = q * a + clear[:3] * (1 - a)
            want[py, px, 3] = a * a + clear[3] * (1 - a)
    assert inside > 200
    assert np.abs(o.img64 - want).max() < 1e-6
    assert np.abs(o.img32 - want).max() < 1e-5
    # the pixel nearest the sprite centre is the inner colour
    cy, cx = np.unravel_index(np.argmax(np.all(np.isclose(want[..., :3], sp.inner_color), -1)), want.shape[:2])
    assert np.allclose(o.img32[cy, cx, :3], sp.inner_color)
