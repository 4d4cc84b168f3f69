"""Host-side behaviour of magtense.magstatics (no GPU): the Tiles setters,
easy-axis conventions, grid_config and refine_prism, as written in
python/src/magtense/magstatics.py:9-741."""

import math

import numpy as np
import pytest

from magtense import magstatics as ms

MU0 = 4 * math.pi * 1e-7


def test_defaults_and_layout():
    t = ms.Tiles(3, mag_angle=[0.0, 0.0])
    assert t.n == 3
    assert t.size.shape == (3, 3) and t.size.dtype == np.float64 and t.size.flags.f_contiguous
    assert t.vertices.shape == (3, 3, 4)
    assert t.tile_type.dtype == np.int32 and np.all(t.tile_type == 1)
    assert np.all(t.mu_r_ea == 1.0) and np.all(t.mu_r_oa == 1.0)
    assert np.all(t.use_sym == 0) and np.all(t.sym_op == 1.0)
    assert np.all(t.magnet_type == 1) and np.all(t.incl_it == 1)


def test_easy_axis_sets_magnetisation():
    t = ms.Tiles(2, tile_type=2, M_rem=1.2 / MU0, easy_axis=[1, 1, 0])
    u = np.around(np.array([1, 1, 0]) / math.sqrt(2), 9)
    assert np.array_equal(t.u_ea[0], u) and np.array_equal(t.u_ea[1], u)
    assert np.allclose(t.M, (1.2 / MU0) * u)
    # u_oa1 = u_ea x (1,0,0), u_oa2 = u_ea x u_oa1 (rounded to 9 decimals)
    assert np.allclose(t.u_oa1[0], np.around(np.cross(u, [1, 0, 0]), 9))
    assert np.allclose(t.u_oa2[0], np.around(np.cross(u, t.u_oa1[0]), 9))
    # along x the helper axis is y
    t2 = ms.Tiles(1, M_rem=1.0, easy_axis=[2, 0, 0])
    assert np.array_equal(t2.u_ea[0], [1, 0, 0])
    assert np.allclose(t2.u_oa1[0], np.cross([1, 0, 0], [0, 1, 0]))
    # one tile through the tuple form
    t.u_ea = ([0, 0, 3], 1)
    assert np.array_equal(t.u_ea[1], [0, 0, 1]) and np.allclose(t.M[1], [0, 0, 1.2 / MU0])


def test_spherical_easy_axis_and_seed():
    t = ms.Tiles(1, M_rem=2.0, mag_angle=[math.pi / 2, math.pi / 2])
    assert np.allclose(t.u_ea[0], [0, 1, 0], atol=1e-9)
    assert np.allclose(t.M[0], [0, 2.0, 0], atol=1e-8)
    a = ms.Tiles(4, M_rem=1.0)
    b = ms.Tiles(4, M_rem=1.0)
    assert np.array_equal(a.u_ea, b.u_ea)  # default_rng(42)
    rand = np.random.default_rng(42).random(size=(4, 2))
    th, ph = math.pi * rand[2, 0], 2 * math.pi * rand[2, 1]
    assert np.allclose(a.u_ea[2], [math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)], atol=1e-9)
    with pytest.raises(ValueError):
        a._set_ea_i(1.0, 0)  # both spherical angles are required


def test_setters_broadcast_tuple_and_rows():
    t = ms.Tiles(3, mag_angle=[0.0, 0.0])
    t.offset = [1.0, 2.0, 3.0]
    assert np.array_equal(t.offset, np.tile([1.0, 2.0, 3.0], (3, 1)))
    t.offset = ([9.0, 9.0, 9.0], 1)
    assert np.array_equal(t.offset[1], [9.0, 9.0, 9.0])
    t.size = [[1, 1, 1], [2, 2, 2], [3, 3, 3]]
    assert np.array_equal(t.size[2], [3, 3, 3])
    t.M_rem = 5.0
    assert np.array_equal(t.M_rem, [5.0] * 3)
    t.tile_type = 2
    assert t.tile_type.dtype == np.int32 and np.all(t.tile_type == 2)
    t.rot = [0.1, 0.2, 0.3]  # a flat vector sets tile 0 only
    assert np.array_equal(t.rot[0], [0.1, 0.2, 0.3]) and np.all(t.rot[1:] == 0)
    t[2].offset = [4.0, 5.0, 6.0]
    assert np.array_equal(t.offset[2], [4.0, 5.0, 6.0])
    assert np.array_equal(t[2].offset, [4.0, 5.0, 6.0])
    with pytest.raises(IndexError):
        t[3]


def test_grid_config_and_refine():
    tiles, pts = ms.grid_config([4, 4, 1], [0.04, 0.04, 0.01], n_tiles=5, n_pts=[3, 2, 1])
    assert tiles.n == 5 and pts.shape == (6, 3) and pts.flags.f_contiguous
    assert np.all(tiles.tile_type == 2)
    assert np.allclose(tiles.size, [0.01, 0.01, 0.01])
    assert np.allclose(tiles.M_rem, 1.2 / MU0)
    cells = np.round(tiles.offset / 0.01 - 0.5).astype(int)
    assert len({tuple(c) for c in cells}) == 5
    assert np.all((cells >= 0) & (cells < [4, 4, 1]))
    t = ms.Tiles(1, size=[0.6, 0.2, 0.4], offset=[1.0, 2.0, 3.0], tile_type=2, M_rem=1.0, easy_axis=[0, 0, 1],
                 rot=[0.0, 0.0, math.pi / 2])
    t.refine_prism(0, [3, 2, 2])
    assert t.n == 12
    assert np.allclose(t.size, [0.2, 0.1, 0.2])
    assert np.allclose(t.offset.mean(axis=0), [1.0, 2.0, 3.0])
    assert np.allclose(t.M, [0, 0, 1.0])
    # sub-prism centres are rotated with the tile: x-extent of the cloud is the original y edge
    span = t.offset.max(axis=0) - t.offset.min(axis=0)
    assert np.allclose(span, [0.1, 0.4, 0.2])


def test_non_prism_rejected_before_gpu():
    t = ms.Tiles(2, tile_type=[2, 6], size=[0.1, 0.1, 0.1], M_rem=1.0, easy_axis=[1, 0, 0])
    with pytest.raises(NotImplementedError):
        ms.get_demag_tensor(t, np.zeros((1, 3)))
    with pytest.raises(NotImplementedError):
        ms.run_simulation(t, np.zeros((1, 3)))
