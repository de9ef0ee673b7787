"""Static collider set on the CPU: the 16-bit PNG reader, the oracle's range-image loader against
closed forms, and the oracle's static contact force against a known answer."""
import os

import numpy as np
import pytest

import oracle
from diploma_thesis_amd import capi, scenes, staticscene

HERE = os.path.dirname(os.path.abspath(__file__))
DEPTH = os.path.join(HERE, "golden", "RangeImages", "Box", "MyPicDepth.png")


def test_png_round_trip_all_filters(tmp_path):
    rng = np.random.default_rng(0)
    img = rng.integers(0, 65536, size=(23, 17), dtype=np.uint16)
    img[5:9] = 1234                              # flat rows (filters see zero differences)
    filters = [k % 5 for k in range(23)]
    path = str(tmp_path / "t.png")
    staticscene.write_png16(path, img, filters)
    assert np.array_equal(staticscene.read_png16(path), img)


def test_depth_fixture_decodes():
    img = staticscene.read_png16(DEPTH)
    assert img.shape == (480, 640) and img.dtype == np.uint16
    assert 0.05 < (img == 0).mean() < 0.5 and 200 < int(img.max()) < 10000


def gravity_rotation():
    g = np.array([0.478957, -8.010560, 4.166928])
    g = (g / np.linalg.norm(g)).astype(np.float32).astype(np.float64)
    return np.array([[1.0, -g[0], 0.0], [0.0, -g[1], g[2]], [0.0, -g[2], -g[1]]])   # rows = (x, y, z) of G


def test_range_image_closed_form():
    H, W = 24, 32
    depth = np.full((H, W), 2500, np.uint16)       # d = 0.5 everywhere
    depth[3, 4] = 0                                # no depth -> pushed to z = -1000 before the transform
    cam = np.eye(4, dtype=np.float32)
    cam[0, 1] = 0.25                               # column 0, row 1 (column-major storage)
    pos, norm, radii, vao = oracle.range_image(depth, cam, 0.01)
    last = (W - 1) * (H - 1)
    assert np.isnan(pos[last + 1:]).all() and np.isfinite(pos[:last + 1]).all()
    assert np.all(radii[:last + 1] == np.float32(0.01))
    # closed form in float64: t = transpose(cam) (X, Y, d), u = G t, p = (u_x, -u_y, -u_z)
    M = cam.reshape(4, 4).T.astype(np.float64)     # M[r, c] = element (row r, col c)
    G = gravity_rotation()
    ys, xs = np.divmod(np.arange(last + 1), W)
    d = np.where(depth.reshape(-1)[:last + 1] == 0, -1000.0, depth.reshape(-1)[:last + 1] / 5000.0)
    X = d * (xs - 318.6) / 517.3
    Y = d * (ys - 255.3) / 516.5
    t = M.T[:3, :3] @ np.stack([X, Y, d])          # transpose(cam) applied to (X, Y, d, 0)
    u = G @ t
    want = np.stack([u[0], -u[1], -u[2]], 1)
    assert np.allclose(pos[:last + 1, :3], want, rtol=2e-6, atol=2e-6 * np.abs(want).max())
    assert np.allclose(vao[:last + 1, 3], (xs / W).astype(np.float32)) and np.allclose(vao[:last + 1, 4], ys / H)
    # interior normals of the flat part: unit length, orthogonal to the image's x and y directions
    y, x = 10, 12
    n = norm[y * W + x, :3].astype(np.float64)
    ex = pos[y * W + x + 1, :3] - pos[y * W + x - 1, :3]
    ey = pos[(y + 1) * W + x, :3] - pos[(y - 1) * W + x, :3]
    assert abs(np.linalg.norm(n) - 1) < 1e-6 and abs(n @ ex) < 1e-6 and abs(n @ ey) < 1e-6
    border = np.zeros((H, W), bool)
    border[0], border[-1], border[:, 0], border[:, -1] = True, True, True, True
    assert np.isnan(norm.reshape(H, W, 4)[border]).all()


def test_static_force_known_answer():
    sc = scenes.mix_scene(1024, 3)
    st = oracle.State(sc, np.float64)
    st.force[:] = 0
    st.torque[:] = 0
    r = float(sc.params.particleRadius)
    k = float(sc.params.spring)
    i = 0                                          # a body particle
    assert sc.rb[i] >= 0
    p = st.pos[i, :3].copy()
    gap = 1.3 * r                                  # centre-to-point distance < 2r
    sp = np.array([[p[0], p[1] - gap, p[2], 0.0]], np.float32)
    sn = np.array([[0.0, 1.0, 0.0, 0.0]], np.float32)
    st.vel[i, :3] = 0.0
    st.rel[i, :3] = (0.1, -0.2, 0.3)
    # cell bounds of the single static point
    h, _ = oracle.calc_hash(sp, sc.params)
    cs = np.full(sc.params.numCells, 0xFFFFFFFF, np.uint32)
    ce = np.zeros(sc.params.numCells, np.uint32)
    cs[h[0]], ce[h[0]] = 0, 1
    st.dem_static(cs, ce, sp, sn)
    dist = np.linalg.norm(sp[0, :3].astype(np.float64) - p)
    want = np.array([0.0, k * (2 * r - dist), 0.0]) + float(sc.params.attraction) * (sp[0, :3] - p)
    assert np.allclose(st.force[i, :3], want, rtol=1e-12, atol=1e-9)
    assert np.allclose(st.torque[i, :3], np.cross(st.rel[i, :3], st.force[i, :3])) and st.torque[i, :3].any()
    # every other particle is out of reach: nothing added
    far = np.linalg.norm(st.pos[:, :3] - sp[0, :3], axis=1) >= 2 * r
    assert not st.force[far, :3].any()
