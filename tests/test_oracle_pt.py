"""CPU tests that pin the path-tracer oracle: PCG32 against its published
demo output, the FP32 restatement against the independent FP64 brute-force
estimator, and closed forms for the sky and sun terms. No GPU."""
import numpy as np
import pytest

from test_oracle_scene import w2o_of


def tiny_pt_scene(scenes, W=48, H=32):
    """24 instances of four small heightfield patches, camera in front."""
    sc = scenes.instanced_scene(mesh_res=(4, 6, 8, 10), n_inst=24, seed=3, span=1.5)
    cam = scenes.look_at((0.0, 1.5, -4.0), (0.0, 0.0, 0.0), W, H, 60.0)
    return sc, cam


def rel_mse(a, ref):
    return float(((a - ref) ** 2 / (ref ** 2 + 1e-2)).mean())


def test_pcg32_known_answer(oracle):
    # pcg32-demo: pcg32_srandom_r(&rng, 42u, 54u), first six outputs
    want = [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]
    assert oracle.pcg32(42, 54, 6).tolist() == want


@pytest.mark.parametrize("spp,depth,rr", [(1, 8, 3), (4, 8, 3), (2, 3, 1)])
def test_fp32_path_tracer_matches_fp64(oracle, scenes, spp, depth, rr):
    sc, cam = tiny_pt_scene(scenes)
    osc = oracle.OracleScene(sc["meshes"], sc["o2w"], sc["blas_of"], w2o_of(sc["o2w"]))
    p = oracle.path_params(48, 32, spp=spp, max_depth=depth, rr_start=rr, seed=7)
    rgb, aov, cnt = oracle.path_trace_2l(osc, sc["albedo"], scenes.cam_array(cam), p)
    wt, wi, _ = scenes.instance_world_tris(sc)
    ref = oracle.path_trace_f64(wt, wi, sc["albedo"], scenes.cam_array(cam), p)
    assert (aov[6] > 0).mean() > 0.5                      # most camera rays hit geometry
    assert cnt[1] > 0.3 * cnt[0]                          # paths continue
    assert rel_mse(rgb, ref) < 1e-6
    # counts: one camera ray per pixel and sample, populations never grow
    assert cnt[0] == 48 * 32 * spp
    ext, sh = cnt[:depth], cnt[depth:]
    assert np.all(ext[1:] <= ext[:-1]) and np.all(sh <= ext)


def test_sky_closed_form(oracle, scenes):
    """Nothing in view: every pixel is the sky gradient of its camera ray."""
    sc = scenes.instanced_scene(mesh_res=(3,), n_inst=2, seed=1, span=1.0)
    sc["o2w"][:, 3] = 0.0
    sc["o2w"][:, 7] = 0.0
    sc["o2w"][:, 11] = -50.0          # far behind the camera
    cam = scenes.look_at((0.0, 0.0, 0.0), (0.2, 0.3, 1.0), 40, 30, 70.0)
    osc = oracle.OracleScene(sc["meshes"], sc["o2w"], sc["blas_of"], w2o_of(sc["o2w"]))
    p = oracle.path_params(40, 30, spp=1, seed=3)
    rgb, aov, cnt = oracle.path_trace_2l(osc, sc["albedo"], scenes.cam_array(cam), p)
    assert cnt[1] == 0 and np.all(aov == 0)
    # camera directions (jitter from the same PCG32 stream)
    W, H = 40, 30
    c = scenes.cam_array(cam).astype(np.float64)
    e = np.zeros((H, W))
    for i in range(W * H):
        py, px = divmod(i, W)
        j = oracle.pcg32((3 << 32) ^ 0, i, 2) >> 8
        jx, jy = j.astype(np.float64) * 2.0 ** -24
        sx, sy = (px + jx) / W * 2 - 1, (py + jy) / H * 2 - 1
        d = c[3:6] + sx * c[12] * c[6:9] + sy * c[13] * c[9:12]
        e[py, px] = max(d[1] / np.linalg.norm(d), 0.0)
    hz, zn = np.array([1.0, 1.0, 1.0]), np.array([0.35, 0.55, 1.0])
    want = hz[:, None, None] + (zn - hz)[:, None, None] * e[None]
    assert np.abs(rgb - want).max() < 1e-5


def test_sun_closed_form(oracle, scenes):
    """One large quad facing the sun, depth 1: albedo * E * cos / pi."""
    quad = np.array([[-50, 0, -50, 50, 0, -50, 50, 0, 50], [-50, 0, -50, 50, 0, 50, -50, 0, 50]], np.float32)
    eye = np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0], np.float32)[None]
    alb = np.array([[0.25, 0.5, 0.75]], np.float32)
    osc = oracle.OracleScene([quad], eye, [0], eye)
    cam = scenes.look_at((0.0, 3.0, 0.0), (0.0, 0.0, 0.5), 16, 16, 60.0)
    sun = np.array([0.35, 0.85, -0.40])
    p = oracle.path_params(16, 16, spp=1, max_depth=1, seed=5, sun_dir=tuple(sun))
    rgb, aov, cnt = oracle.path_trace_2l(osc, alb, scenes.cam_array(cam), p)
    assert np.all(aov[6] > 0)
    cos = (sun / np.linalg.norm(sun))[1]
    want = alb[0] * np.array([3.0, 2.85, 2.6]) * cos / np.pi
    assert np.allclose(rgb.reshape(3, -1), want[:, None], rtol=1e-6)
    assert np.allclose(aov[3:6].reshape(3, -1), [[0], [1], [0]])
