"""Learned far field, CPU side: the oracle's MLP and segment generator.

Pins:
  * MLP backward (pure FP64 network) against central finite differences of
    the loss, every parameter block, with the error target held constant;
  * the bf16-emulating network stays within bf16 rounding of the FP64 one;
  * segment targets against two independent plain marches of the oracle
    (truncated after k samples, and full): bit-exact;
  * segment features: position, directions, and the LOD samples against the
    oracle sampler at the documented heights; bit-exact;
  * closed forms on a constant volume: t_rgb = light * t_a, 1 - a_k equals
    the product of the per-step transmittances.
"""
import math

import numpy as np
import pytest

import oracle_lib as O

NP = 13957


def _params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    dims = [(17, 64), (64, 64), (64, 64), (64, 64), (64, 5)]
    out = []
    for fin, fout in dims:
        out.append((rng.normal(0, math.sqrt(2.0 / fin), (fout, fin)) * scale).astype(np.float32).ravel())
        out.append((rng.normal(0, 0.1, fout) * scale).astype(np.float32))
    p = np.concatenate(out)
    assert p.size == NP
    return p


def _rows(seed, n):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n, 17)).astype(np.float32)
    t = rng.uniform(0, 1, (n, 5)).astype(np.float32)
    return x, t


def test_param_count():
    assert O.lib().cvo_mlp_param_count() == NP


def test_mlp_grad_finite_differences():
    p = _params(1, 0.7)
    x, t = _rows(2, 12)
    y = O.mlp_forward(p, x, emulate=0)
    e = np.max(np.abs(np.clip(y[:, :4], 0, 1) - t[:, :4]), axis=1).astype(np.float32)
    g, loss = O.mlp_grad(p, x, t, emulate=0, e_fixed=e)
    # blocks: W0 b0 W1 b1 W2 b2 W3 b3 W4 b4
    offs = [0, 1088, 1152, 5248, 5312, 9408, 9472, 13568, 13632, 13952, 13957]
    rng = np.random.default_rng(3)
    checked = 0
    for b in range(10):
        idx = rng.choice(np.arange(offs[b], offs[b + 1]), size=min(6, offs[b + 1] - offs[b]), replace=False)
        for i in idx:
            h = np.float32(1e-2 * max(1.0, abs(float(p[i]))))
            pp, pm = p.copy(), p.copy()
            pp[i] = p[i] + h
            pm[i] = p[i] - h
            _, lp = O.mlp_grad(pp, x, t, emulate=0, e_fixed=e)
            _, lm = O.mlp_grad(pm, x, t, emulate=0, e_fixed=e)
            fd = (lp - lm) / (float(pp[i]) - float(pm[i]))
            assert abs(fd - g[i]) <= 1e-4 * max(1.0, abs(g[i])) + 1e-7, (b, i, fd, g[i])
            checked += 1
    assert checked >= 50
    assert loss > 0


def test_mlp_loss_closed_form():
    # zero network: y = 0, e = max_c t_c; loss = mean(w^2 sum_c t_c^2 + e^2)
    p = np.zeros(NP, np.float32)
    x, t = _rows(4, 9)
    g, loss = O.mlp_grad(p, x, t, emulate=0)
    e = t[:, :4].max(axis=1).astype(np.float64)
    w = t[:, 4].astype(np.float64)
    want = np.mean(w ** 2 * (t[:, :4].astype(np.float64) ** 2).sum(axis=1) + e ** 2)
    assert abs(loss - want) <= 1e-12 * max(1.0, want)
    # output bias gradient = mean of dL/dy
    gb = np.concatenate([np.mean(-2 * w[:, None] ** 2 * t[:, :4], axis=0), [np.mean(-2 * e)]])
    assert np.allclose(g[13952:13957], gb, rtol=1e-12, atol=1e-15)


def test_mlp_emulation_is_bf16_close():
    p = _params(5)
    x, _ = _rows(6, 256)
    y0 = O.mlp_forward(p, x, emulate=0)
    y1 = O.mlp_forward(p, x, emulate=1)
    scale = np.abs(y0).max()
    d = np.abs(y0 - y1).max()
    assert 0 < d <= 0.05 * scale


# ---- segments ----------------------------------------------------------

def _const_volume(N, L, rgba):
    lv = []
    for k in range(L):
        n = max(1, N >> k)
        lv.append(np.broadcast_to(np.asarray(rgba, np.float32), (n, n, n, 4)).ravel())
    return np.ascontiguousarray(np.concatenate(lv))


def _cone(steps=64, angle=20.0, mode=2, sigma=24.0, comp=0):
    c = O.cvo_cone()
    c.steps, c.cone_angle, c.initial_height = steps, math.radians(angle), 0.1
    c.lod_offset, c.down_scaling, c.cone_mode = 0.0, 1.0, mode
    c.alpha_stop, c.sigma, c.composite, c.compensated = 0.99, sigma, 1, comp
    return c


def _light(d):
    L = O.cvo_light()
    d = np.asarray(d, np.float64)
    d = d / np.linalg.norm(d)
    L.direction = O.f3(d)
    L.position = O.f3(d * 10)
    L.directional = 1
    return L


def test_segments_constant_volume_closed_form():
    N, L, k = 64, 7, 12
    light_v = 0.6
    rho = 0.05
    nat = _const_volume(N, L, (light_v, light_v, light_v, rho))
    vol = O.make_volume(N, L)
    cone = _cone(steps=64, sigma=8.0)
    lt = _light((0.3, 1.0, 0.2))
    rng = np.random.default_rng(7)
    p0 = rng.uniform(0.1, 0.4, (40, 3)).astype(np.float32)
    feat, tgt, valid = O.segments(vol, nat, cone, lt, k, p0)
    assert valid.sum() > 30
    hh, rr, lvl, frac = O.step_table(cone, L, N)
    kk = rr.astype(np.float64) * (np.float32(cone.sigma) * np.float32(1.4426950408889634) / np.float32(N))
    alpha = 1 - np.exp2(-kk * np.float64(np.float32(rho)))
    trans_k = np.prod(1 - alpha[:k])
    for i in np.nonzero(valid)[0]:
        f, t = feat[i], tgt[i]
        # density / light features where sampled, zero beyond the volume
        for j in range(4):
            d, l = f[9 + 2 * j], f[10 + 2 * j]
            assert (d, l) == (0, 0) or (abs(d - np.float32(rho)) < 1e-6 and abs(l - np.float32(light_v)) < 1e-6)
        assert abs(t[4] - trans_k) <= 2e-6
        assert abs(t[0] - light_v * t[3]) <= 2e-5 and t[0] == t[1] == t[2]
        assert 0 <= t[3] <= 1


def test_segments_exit_before_k_is_invalid():
    N, L = 64, 7
    nat = _const_volume(N, L, (1, 1, 1, 0.01))
    vol = O.make_volume(N, L)
    cone = _cone(steps=64)
    lt = _light((0, 1, 0))
    p0 = np.array([[0.5, 0.999, 0.5], [0.5, 0.2, 0.5]], np.float32)
    _, _, valid = O.segments(vol, nat, cone, lt, 20, p0)
    assert valid.tolist() == [0, 1]


def _scene_volume():
    import conftest  # noqa: F401
    import cloud_renderer_amd as cv
    from dataclasses import replace
    sc = cv.scenes.small32(1)
    sc = replace(sc, N=64, levels=7, cone=replace(sc.cone, steps=96, cone_angle=math.radians(20.0)))
    vr = O.voxelize(sc)
    nat, _, _, _ = O.build_levels(sc, vr.shell, vr.fill, want_f64=False)
    return cv, sc, nat


@pytest.mark.parametrize("k", [24, 33])
def test_segments_match_plain_marches(k):
    cv, sc, nat = _scene_volume()
    from dataclasses import replace
    import farfield_starts
    vol = O.volume_of(sc)
    cone = O.cone_of(sc.cone)
    lt = O.light_of(sc)
    p0 = farfield_starts.starts(cv, sc, 300, 11)
    feat, tgt, valid = O.segments(vol, nat, cone, lt, k, p0)
    assert valid.sum() > 50
    d = (sc.sun_dir / np.linalg.norm(sc.sun_dir.astype(np.float32))).astype(np.float32)
    # the oracle's directional cone direction: normalize(L) / N in FP32
    inv = np.float32(1.0) / np.sqrt(np.float32(np.dot(d.astype(np.float32), d.astype(np.float32))))
    ld = (sc.sun_dir.astype(np.float32) * inv).astype(np.float32)
    dn = (ld / np.float32(sc.N)).astype(np.float32)
    cone_k = O.cone_of(replace(sc.cone, steps=k))
    for i in range(len(p0)):
        c_full, _, n_full = O.trace_one(vol, nat, None, cone, p0[i], dn)
        c_k, _, n_k = O.trace_one(vol, nat, None, cone_k, p0[i], dn)
        alive = n_k == k and c_k[3] < np.float32(sc.cone.alpha_stop) and k < sc.cone.steps
        assert bool(valid[i]) == alive
        if not alive:
            assert not tgt[i].any() and not feat[i].any()
            continue
        om = np.float32(1.0) - c_k[3]
        want = np.array([(c_full[0] - c_k[0]) / om, (c_full[1] - c_k[1]) / om, (c_full[2] - c_k[2]) / om,
                         (c_full[3] - c_k[3]) / om, om], np.float32)
        assert np.array_equal(tgt[i].view(np.uint32), want.view(np.uint32)), (i, tgt[i], want)
        assert np.allclose(feat[i, 3:6], ld, rtol=0, atol=1e-7)
        assert np.array_equal(feat[i, 6:9], ld)


def test_segment_features_match_sampler():
    cv, sc, nat = _scene_volume()
    import farfield_starts
    k = 16
    vol = O.volume_of(sc)
    cone = O.cone_of(sc.cone)
    lt = O.light_of(sc)
    p0 = farfield_starts.starts(cv, sc, 200, 12)
    feat, tgt, valid = O.segments(vol, nat, cone, lt, k, p0)
    hh, _, _, _ = O.step_table(cone, sc.levels, sc.N)
    tanh = np.tan(np.float32(cone.cone_angle) / np.float32(2))
    L = sc.levels
    ld = feat[np.nonzero(valid)[0][0], 6:9]
    dn = (ld / np.float32(sc.N)).astype(np.float32)
    for i in np.nonzero(valid)[0][:60]:
        hk = np.float64(hh[k - 1])
        xk = (hk * dn.astype(np.float64) + p0[i].astype(np.float64)).astype(np.float32)
        assert np.allclose(feat[i, :3], 2 * xk - 1, rtol=0, atol=4e-7)
        for j in range(4):
            lod = 3 + j
            hj = max(np.float32(hh[k - 1]), np.float32(np.float32(1 << (lod - 1)) / np.float32(tanh)))
            y = (np.float64(hj) * dn.astype(np.float64) + p0[i].astype(np.float64)).astype(np.float32)
            if np.all((y >= 0) & (y <= 1)):
                o32, _ = O.sample(vol, nat, None, float(min(lod, L - 1)), y[None])
                assert abs(feat[i, 9 + 2 * j] - o32[0, 3]) <= 1e-6
                assert abs(feat[i, 10 + 2 * j] - o32[0, 0]) <= 1e-6
            else:
                assert feat[i, 9 + 2 * j] == 0 and feat[i, 10 + 2 * j] == 0
