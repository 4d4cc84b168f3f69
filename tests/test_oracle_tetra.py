"""Oracle checks for the tetrahedron tensor (tileType 5:
getN_tensor_tetrahedron, TileNComponents.f90:944-961, over getN_Triangle and
its face integrals, TileTriangle.f90:26-286).

* Superposition: the six Kuhn tetrahedra of a box (vertices 0, e_s1,
  e_s1 + e_s2, a for every axis order s) fill it, so their tensors sum to the
  box's prism tensor (pinned independently: closed forms, quadrature) at every
  point off their faces, inside and outside.
* Inside a tetrahedron the trace is -1 (MagTense sign), outside 0; the result
  does not depend on the order of the vertices; N is symmetric to rounding
  (the entries are sums of O(1) face terms, so rounding is absolute, ~1e-15);
  far away it tends to the point-dipole tensor; degenerate faces are refused
  (the reference stops).
"""

import itertools

import numpy as np
import pytest


def kuhn(a, origin=np.zeros(3)):
    tets = []
    for perm in itertools.permutations(range(3)):
        v = [origin.copy()]
        cur = origin.copy()
        for ax in perm:
            cur = cur.copy()
            cur[ax] += a[ax]
            v.append(cur)
        tets.append(np.array(v))
    return tets


def _off_faces(u, gap=0.04):
    return all(abs(u[i] - u[j]) > gap for i in range(3) for j in range(i + 1, 3)) and np.all(
        np.abs(u) > gap) and np.all(np.abs(u - 1) > gap)


def test_kuhn_tetrahedra_sum_to_the_prism(oracle):
    rng = np.random.default_rng(3)
    a = np.array([4e-9, 3e-9, 2e-9])
    tets = kuhn(a)
    comp = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    n_in = n_out = 0
    while n_in + n_out < 60:
        u = rng.uniform(0, 1, 3) if (n_in + n_out) % 2 else rng.uniform(-1.5, 2.5, 3)
        if not _off_faces(u):
            continue
        r = u * a
        S = sum(oracle.tetra_n(t, r) for t in tets)
        k = oracle.prism_k6(a, r - a / 2)
        for c, (i, j) in enumerate(comp):
            # the face terms are O(1) and cancel: absolute bar
            assert abs(S[i, j] - k[c]) <= 1e-12, (u, i, j, S[i, j], k[c])
        inside = np.all((u > 0) & (u < 1))
        n_in += inside
        n_out += not inside
        assert abs(np.trace(S) - (-1.0 if inside else 0.0)) < 1e-10
    assert n_in >= 5 and n_out >= 5


def test_tetrahedron_invariants(oracle):
    rng = np.random.default_rng(9)
    for _ in range(10):
        v = rng.uniform(-5e-9, 5e-9, (4, 3))
        vol = abs(np.linalg.det(v[1:] - v[0])) / 6
        if vol < 5e-27:
            continue
        c = v.mean(axis=0)
        bary = rng.dirichlet(np.ones(4))
        inside = bary @ v
        far = c + np.array([60e-9, -40e-9, 30e-9])
        for r, tr in ((c, -1.0), (inside, -1.0), (far, 0.0)):
            N = oracle.tetra_n(v, r)
            # entries are sums of O(1) face terms: rounding is absolute, ~1e-15
            assert abs(np.trace(N) - tr) < 1e-12
            assert np.abs(N - N.T).max() <= 1e-13
            for perm in ((1, 0, 2, 3), (3, 2, 1, 0), (2, 3, 0, 1)):
                Np = oracle.tetra_n(v[list(perm)], r)
                assert np.abs(Np - N).max() <= 1e-13
        # far field: the point-dipole tensor V/(4 pi r^3) (3 rr - I), MagTense sign
        d = far - c
        rn = np.linalg.norm(d)
        dip = vol / (4 * np.pi * rn ** 3) * (3 * np.outer(d, d) / rn ** 2 - np.eye(3))
        Nf = oracle.tetra_n(v, far)
        assert np.abs(Nf - dip).max() < 0.05 * np.abs(dip).max()


def test_degenerate_face_refused(oracle):
    v = np.array([[0, 0, 0], [1e-9, 0, 0], [2e-9, 0, 0], [0, 1e-9, 1e-9]], dtype=float)
    with pytest.raises(ValueError):
        oracle.tetra_n(v, np.array([5e-9, 5e-9, 5e-9]))


def box_mesh(nb=(3, 2, 1), a=(5e-9, 4e-9, 3e-9)):
    """Boxes on a lattice, each split into its six Kuhn tetrahedra (a conforming
    mesh): nodes (lattice points), 1-based elements (n, 4), centroids, and the
    box of every element."""
    a = np.asarray(a)
    idx = {}
    nodes = []

    def node(p):
        key = tuple(int(v) for v in p)
        if key not in idx:
            idx[key] = len(nodes) + 1
            nodes.append(np.asarray(p, dtype=float) * a)
        return idx[key]

    elems, box = [], []
    for bz in range(nb[2]):
        for by in range(nb[1]):
            for bx in range(nb[0]):
                o = np.array([bx, by, bz])
                for perm in itertools.permutations(range(3)):
                    cur = o.copy()
                    vs = [node(cur)]
                    for ax in perm:
                        cur = cur.copy()
                        cur[ax] += 1
                        vs.append(node(cur))
                    elems.append(vs)
                    box.append((bx, by, bz))
    nodes = np.array(nodes)
    elems = np.array(elems)
    cent = nodes[elems - 1].mean(axis=1)
    return nodes, elems, cent, np.array(box), a


def test_tetra_grid_tensor_sums_to_box_prisms(oracle):
    """Tetrahedral micromagnetic grid (LandauLifshitzEquationSolver.f90:988-1041):
    the demag field of the six elements of one box, all magnetised alike,
    equals that box's prism field at every element centroid (none lies on a
    face plane of the triangulation)."""
    nodes, elems, cent, box, a = box_mesh()
    n = len(elems)
    exch = (np.ones(n, dtype=np.int32), np.ones(n, dtype=np.int32), np.array([1]), np.array([0.0]))
    q = oracle.Problem([n, 1, 1], [1, 1, 1], grid_pts=cent, exch=exch, grid_nodes=nodes, grid_elements=elems,
                       demag_fp64=True, Ms=1.0)
    b0 = (1, 1, 0)
    sel = np.all(box == b0, axis=1)
    centre = (np.array(b0) + 0.5) * a
    K = oracle.tensor_points(centre[None, :], a[None, :], cent)  # (6, n, 1): the box prism at every centroid
    comp = [[0, 1, 2], [1, 3, 4], [2, 4, 5]]
    for b in range(3):
        m = np.zeros((3, n))
        m[b, sel] = 1.0
        _, H = q.dmdt(0.0, m.reshape(-1))
        Hd = H[2].reshape(3, n)
        for c in range(3):
            ref = -K[comp[c][b], :, 0]
            assert np.abs(Hd[c] - ref).max() <= 1e-12, (b, c, np.abs(Hd[c] - ref).max())


def test_tetra_tiles_match_tetra_tensor(oracle):
    """Tetrahedron tiles through the tile routine (offset, rotation, eight-fold
    symmetry) against the element tensor: N = T(R (p - offset)), and the
    symmetric tile equals its eight mirrored evaluations."""
    rng = np.random.default_rng(4)
    v = np.array([[0, 0, 0], [4e-9, 0, 0], [0, 3e-9, 0], [0, 0, 2e-9]], dtype=float)
    verts = np.transpose(v)[None, :, :]  # (1, 3, 4)
    off = np.array([[1e-9, -2e-9, 0.5e-9]])
    ang = np.array([[0.3, -0.2, 0.7]])
    pts = rng.uniform(-8e-9, 8e-9, (20, 3))
    N = oracle.tiles_tensor(np.zeros((1, 3)), off, pts, rot=ang, tile_type=[5], vertices=verts)
    R, _ = oracle.rotation_matrices(ang[0])
    for p in range(len(pts)):
        ref = oracle.tetra_n(v, R @ (pts[p] - off[0]))
        assert np.abs(N[0, p] - ref).max() <= 1e-13  # O(1) face terms; numpy rounds R (p - off) differently
    ops = np.array([[1.0, -1.0, 1.0]])
    Ns = oracle.tiles_tensor(np.zeros((1, 3)), off, pts, rot=ang, tile_type=[5], vertices=verts, use_sym=[1],
                             sym_ops=ops)
    mir = [(1, 1, 1), (-1, 1, 1), (1, -1, 1), (1, 1, -1), (-1, -1, 1), (-1, 1, -1), (1, -1, -1), (-1, -1, -1)]
    o0, o1, o2 = ops[0]
    cs = [1.0, o2, o1, o0, o1 * o2, o0 * o2, o0 * o1, o2 * (o1 * o0)]
    for p in range(len(pts)):
        acc = np.zeros((3, 3))
        for j, s in enumerate(mir):
            s = np.array(s, dtype=float)
            acc += s[:, None] * oracle.tetra_n(v, R @ (s * pts[p] - off[0])) * cs[j]
        assert np.abs(Ns[0, p] - acc).max() <= 1e-13
