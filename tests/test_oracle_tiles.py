"""Pins of the oracle's prism-tile restatement (getNFromTiles / getHFromTiles
for prisms, DemagFieldGetSolution.f90:685-948) that do not retype its formula:

* rotation matrices: orthogonality, Rinv = R^T, and the closed form of a
  rotation about z;
* a prism rotated by 90 degrees about z gives the field of the unrotated prism
  with swapped edges (rotation applied consistently to position and M);
* an eight-fold symmetric tile equals eight explicit tiles at the mirrored
  offsets with mirrored (and signed) magnetisation;
* the stored-tensor path (useStoredN) reproduces the direct field;
* the COMSOL 5.4 line profiles shipped with the reference
  (documentation/examples_FEM_validation/Validation_prism, tile of
  python/examples/magnetostatics/validation.py:9-20) at FEM accuracy, which
  pins the rotation sign and order.
"""

import math
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
MU0 = 4 * math.pi * 1e-7


def test_rotation_matrices_closed_form(oracle):
    rng = np.random.default_rng(0)
    for _ in range(20):
        ang = rng.uniform(-math.pi, math.pi, 3)
        R, Ri = oracle.rotation_matrices(ang)
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-15)
        assert np.allclose(Ri, R.T, atol=1e-15)
        assert abs(np.linalg.det(R) - 1.0) < 1e-14
    t = 0.3
    R, _ = oracle.rotation_matrices([0.0, 0.0, t])
    # rotMat = Rz(-t): the position is rotated into the tile frame
    assert np.allclose(R, [[math.cos(t), math.sin(t), 0], [-math.sin(t), math.cos(t), 0], [0, 0, 1]], atol=1e-16)
    R, _ = oracle.rotation_matrices([0.0, 0.0, 0.0])
    assert np.array_equal(R, np.eye(3))


def test_quarter_turn_equals_swapped_edges(oracle):
    rng = np.random.default_rng(1)
    abc = np.array([0.6, 0.2, 0.3])
    off = np.array([0.1, -0.2, 0.05])
    M = np.array([3e5, -1e5, 7e5])
    pts = off + rng.uniform(-1.0, 1.0, size=(200, 3))
    H_rot = oracle.tiles_field([abc], [off], [M], pts, rot=[[0.0, 0.0, math.pi / 2]])
    H_swp = oracle.tiles_field([abc[[1, 0, 2]]], [off], [M], pts)
    assert np.abs(H_rot - H_swp).max() <= 1e-10 * np.abs(H_swp).max()
    # a non-quarter turn is different
    H_other = oracle.tiles_field([abc], [off], [M], pts, rot=[[0.0, 0.0, math.pi / 3]])
    assert np.abs(H_other - H_swp).max() > 1e-3 * np.abs(H_swp).max()


@pytest.mark.parametrize("ops", [(1.0, 1.0, 1.0), (1.0, -1.0, 1.0), (-1.0, 1.0, -1.0)])
def test_symmetric_tile_equals_eight_mirrored_tiles(oracle, ops):
    rng = np.random.default_rng(2)
    abc = np.array([0.2, 0.3, 0.1])
    off = np.array([0.4, 0.5, 0.3])
    M = np.array([2e5, 5e5, -3e5])
    pts = rng.uniform(-1.5, 1.5, size=(150, 3))
    H_sym = oracle.tiles_field([abc], [off], [M], pts, use_sym=[1], sym_ops=[ops])
    o1, o2, o3 = ops  # symmetryOps about the xy, xz and yz planes
    mirrors = [(1, 1, 1, 1.0), (-1, 1, 1, o3), (1, -1, 1, o2), (1, 1, -1, o1),
               (-1, -1, 1, o2 * o3), (-1, 1, -1, o1 * o3), (1, -1, -1, o1 * o2), (-1, -1, -1, o1 * o2 * o3)]
    offs, Ms = [], []
    for sx, sy, sz, c in mirrors:
        S = np.array([sx, sy, sz], dtype=float)
        offs.append(S * off)
        Ms.append(c * S * M)
    H_8 = oracle.tiles_field([abc] * 8, offs, Ms, pts)
    assert np.abs(H_sym - H_8).max() <= 1e-10 * np.abs(H_8).max()


def test_stored_tensor_reproduces_field(oracle):
    rng = np.random.default_rng(3)
    nt, npt = 7, 40
    size = rng.uniform(0.1, 0.5, (nt, 3))
    off = rng.uniform(-1, 1, (nt, 3))
    rot = rng.uniform(-math.pi, math.pi, (nt, 3))
    M = rng.normal(size=(nt, 3)) * 1e5
    pts = rng.uniform(-2, 2, (npt, 3))
    use_sym = np.array([0, 1, 0, 0, 1, 0, 0])
    N = oracle.tiles_tensor(size, off, pts, rot=rot, use_sym=use_sym)
    assert N.shape == (nt, npt, 3, 3) and N.flags.f_contiguous
    H1 = oracle.tiles_field(size, off, M, pts, rot=rot, use_sym=use_sym)
    H2 = oracle.tiles_field(size, off, M, pts, rot=rot, use_sym=use_sym, N=N)
    assert np.array_equal(H1, H2)
    # the non-symmetric tensors are symmetric 3x3 matrices
    plain = use_sym == 0
    assert np.array_equal(N[plain], np.swapaxes(N[plain], 2, 3))
    with pytest.raises(ValueError):
        oracle.tiles_tensor(size, off, pts, tile_type=[2, 2, 3, 2, 2, 2, 2])


def comsol_case():
    ea = np.array([0.35355339, 0.61237244, 0.70710678])
    ea = np.around(ea / np.linalg.norm(ea), decimals=9)
    return dict(size=[0.6, 0.1, 0.3], offset=[0.5, 0.4, 0.1], rot=[math.pi / 2, -math.pi / 3, math.pi / 4],
                M=(1.2 / MU0) * ea)


def comsol_lines():
    d = np.load(os.path.join(HERE, "golden", "comsol_prism_normH.npz"))
    case = comsol_case()
    off = np.asarray(case["offset"])
    for k, ax in enumerate("xyz"):
        line = d[ax]
        pts = np.tile(off, (line.shape[0], 1))
        pts[:, k] = line[:, 0]
        yield ax, pts, line[:, 1]


def comsol_check(field_fn, oracle):
    """|H| mu0 along the three lines vs COMSOL; points exactly on the magnet
    surface (where H jumps) are excluded."""
    case = comsol_case()
    R, _ = oracle.rotation_matrices(case["rot"])
    half = np.asarray(case["size"]) / 2
    worst = {}
    for ax, pts, ref in comsol_lines():
        H = field_fn(case, pts)
        hn = np.linalg.norm(H, axis=1) * MU0
        local = (R @ (pts - np.asarray(case["offset"])).T).T
        dist = np.min(np.abs(np.abs(local) - half), axis=1)
        err = np.abs(hn - ref) / np.abs(ref).max()
        off_surface = dist > 1e-9
        far = dist > 0.01
        assert off_surface.sum() >= ref.size - 2
        assert err[off_surface].max() < 6e-3, (ax, err[off_surface].max())
        assert err[far].max() < 1e-3, (ax, err[far].max())
        worst[ax] = (float(err[off_surface].max()), float(err[far].max()))
    return worst


def test_comsol_prism_lines(oracle):
    def fn(case, pts):
        return oracle.tiles_field([case["size"]], [case["offset"]], [case["M"]], pts, rot=[case["rot"]])

    worst = comsol_check(fn, oracle)
    print("COMSOL prism: max |H| error / peak (off-surface, >1 cm from surface)", worst)
