"""The dense-pile generator (arsim_generate_packed_pile) against what it promises: particle counts,
ellipsoids voxelised at spacing r with 33 to ~2,000 particles, no two bodies closer than the lattice
distance, grains on a face-centred cubic lattice with 12 neighbours in contact in the bulk and none
closer than the surface gap to a body, and each body's inverse inertia the inverse of its point-mass
sum.  Counted with numpy / scipy on the host."""
import numpy as np
import pytest
from scipy.spatial import cKDTree

from diploma_thesis_amd import capi, scenes


@pytest.fixture(scope="module")
def pile():
    return scenes.config2(seed=2, scale=0.02)


def test_pile_counts_and_body_sizes(pile):
    sc = pile
    assert sc.n == int(round(4194304 * 0.02)) and sc.b == int(round(16384 * 0.02))
    sizes = np.bincount(sc.rb[sc.rb >= 0])
    assert sizes.min() >= 33 and sizes.max() <= 2145
    assert np.array_equal(sizes, sc.ppo)
    # heavy tail: the median body is far below the mean of the extremes
    assert np.median(sizes) < 0.2 * sizes.max()
    # bodies first, contiguous, in body order; grains after them
    nb = int(sizes.sum())
    assert np.all(sc.rb[:nb] >= 0) and np.all(sc.rb[nb:] == -1) and np.all(np.diff(sc.rb[:nb]) >= 0)


def test_ellipsoid_voxels_at_spacing_r(pile):
    """Body-frame offsets lie on a lattice of spacing r (up to the centring shift of each body)."""
    sc = pile
    r = sc.params.particleRadius
    for b in range(0, sc.b, max(1, sc.b // 20)):
        st, n = sc.pstart[b], sc.ppo[b]
        f = sc.body_frame[st:st + n, :3].astype(np.float64)
        d = (f - f[0]) / r
        assert np.max(np.abs(d - np.round(d))) < 1e-4


def test_no_body_pair_or_grain_too_close(pile):
    sc = pile
    r = sc.params.particleRadius
    pos = sc.pos[:, :3].astype(np.float64)
    tree = cKDTree(pos)
    pairs = tree.query_pairs(2 * r, output_type="ndarray")
    a, b = sc.rb[pairs[:, 0]], sc.rb[pairs[:, 1]]
    d = np.linalg.norm(pos[pairs[:, 0]] - pos[pairs[:, 1]], axis=1)
    between_bodies = (a >= 0) & (b >= 0) & (a != b)
    assert not between_bodies.any() or d[between_bodies].min() >= 2 * r * 0.98 - 1e-4
    grain_body = (a < 0) != (b < 0)
    assert d[grain_body].min() >= 2 * r * 0.9 - 1e-4
    grain_grain = (a < 0) & (b < 0)
    assert d[grain_grain].min() >= 2 * r * 0.98 - 1e-4


def test_bulk_grains_touch_twelve_neighbours(pile):
    sc = pile
    r = sc.params.particleRadius
    g = sc.rb < 0
    gp = sc.pos[g, :3].astype(np.float64)
    tree = cKDTree(gp)
    counts = np.array([len(x) - 1 for x in tree.query_ball_point(gp, 2 * r * 0.99)])
    assert counts.max() == 12
    # a grain with no body particle within 4r (a skipped site lies within 0.9 * 2r of one) and away from the walls and the pile top is in the bulk
    allpos = sc.pos[:, :3].astype(np.float64)
    body_tree = cKDTree(allpos[sc.rb >= 0])
    near_body = np.array([len(x) > 0 for x in body_tree.query_ball_point(gp, 2 * r * 2.0)])
    edge = sc.cfg.boxMax[0]
    top = gp[:, 1].max()
    inner = (~near_body & (gp[:, 0] > 2) & (gp[:, 0] < edge - 2) & (gp[:, 2] > 2) & (gp[:, 2] < edge - 2)
             & (gp[:, 1] > 2) & (gp[:, 1] < top - 2))
    assert inner.sum() > 100 and np.all(counts[inner] == 12)


def test_inverse_inertia_of_point_mass_sum(pile):
    sc = pile
    I = sc.rb_inertia.reshape(-1, 3, 3)
    for b in range(0, sc.b, max(1, sc.b // 10)):
        st, n = sc.pstart[b], sc.ppo[b]
        f = sc.body_frame[st:st + n, :3].astype(np.float64)
        J = np.eye(3) * np.sum(f * f) - f.T @ f
        # stored column-major: the transpose of the row-major reading (J is symmetric)
        np.testing.assert_allclose(I[b].T, np.linalg.inv(J), rtol=1e-5, atol=1e-9)
        assert sc.rb_mass[b] == n


def test_pile_generator_refuses_impossible_specs():
    sp = capi.ArSimPileSpec(seed=1, numEllipsoids=10, targetParticles=100, edge=40.0, radius=0.5, latticeGap=0.98,
                            surfaceGap=0.9, gridLog2=7, method=capi.METHOD_GRID)
    import ctypes as C
    out = C.c_void_p()
    # ten bodies of at least 33 particles do not fit a 100-particle target
    assert capi.lib.arsim_generate_packed_pile(C.byref(sp), C.byref(out)) == capi.ARSIM_ERR_INVALID_ARG
    sp.targetParticles, sp.edge = 5000, 6.0
    assert capi.lib.arsim_generate_packed_pile(C.byref(sp), C.byref(out)) == capi.ARSIM_ERR_INVALID_ARG
    sp.edge, sp.latticeGap = 40.0, 1.5
    assert capi.lib.arsim_generate_packed_pile(C.byref(sp), C.byref(out)) == capi.ARSIM_ERR_INVALID_ARG
