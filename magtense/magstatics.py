"""Magnetostatic prism tiles on the MI355X: ``Tiles``, ``grid_config``,
``get_demag_tensor`` and ``get_H_field`` with the interface of
``magtense.magstatics`` (python/src/magtense/magstatics.py:9-741,865-950).

The tensor and field evaluations run in ``libmagtense_hip.so``
(``mt_tiles_demag_tensor`` / ``mt_tiles_h_field``, the prism part of
getNFromTiles / getHFromTiles, FortranToPythonIO.f90:25-223).  Rectangular
prisms (``tile_type == 2``) and tetrahedra (``tile_type == 5``, ``vertices``
(3 x 4 per tile) in the tile's frame, the offset subtracted from the point as
for every tile) are supported; any other tile type raises
``NotImplementedError``.  The iterative soft-magnet solver
(``run_simulation`` / ``iterate_magnetization`` of the reference) is not part
of this package.

``Tiles`` keeps the reference's array layout (Fortran-ordered float64 / int32
arrays, one row per tile) and its setter conventions: a ``(value, index)``
tuple sets one tile, a flat 3-vector (or scalar) is broadcast to every tile,
and a sequence of ``n`` rows replaces the whole array.
"""

import ctypes as C
from typing import List, Optional, Tuple, Union

import numpy as np

from magtense_amd import _lib

MU0 = 4 * np.pi * 1e-7
PRISM = 2
TETRAHEDRON = 5


class Tile:
    """View of tile ``index`` of a ``Tiles`` object (magstatics.py:9-41):
    reading an attribute returns that tile's row, assigning writes it."""

    def __init__(self, parent, index):
        self._parent = parent
        self._index = index

    def __getattr__(self, name):
        attr = getattr(self._parent, name)
        return attr[self._index] if isinstance(attr, np.ndarray) else attr

    def __setattr__(self, name, value):
        if name in ("_parent", "_index"):
            super().__setattr__(name, value)
            return
        attr = getattr(self._parent, name)
        if isinstance(attr, np.ndarray):
            attr[self._index] = value


def _zeros3(n):
    return np.zeros((n, 3), dtype=np.float64, order="F")


class Tiles:
    """Input of the MagTense tile routines (the Fortran derived type MagTile).

    Attributes (one row per tile): center_pos, dev_center, size (a, b, c),
    vertices (3x4), M, u_ea, u_oa1, u_oa2, mu_r_ea, mu_r_oa, M_rem, tile_type
    (2 = prism), offset, rot (rotation angles about x, y, z), color,
    magnet_type, stfcn_index, incl_it, use_sym, sym_op, M_rel.
    """

    # per-attribute (shape suffix, dtype, fill value) of the backing arrays
    _LAYOUT = {
        "center_pos": ((3,), np.float64, 0.0), "dev_center": ((3,), np.float64, 0.0),
        "size": ((3,), np.float64, 0.0), "vertices": ((3, 4), np.float64, 0.0),
        "M": ((3,), np.float64, 0.0), "u_ea": ((3,), np.float64, 0.0),
        "u_oa1": ((3,), np.float64, 0.0), "u_oa2": ((3,), np.float64, 0.0),
        "mu_r_ea": ((), np.float64, 1.0), "mu_r_oa": ((), np.float64, 1.0),
        "M_rem": ((), np.float64, 0.0), "tile_type": ((), np.int32, 1),
        "offset": ((3,), np.float64, 0.0), "rot": ((3,), np.float64, 0.0),
        "color": ((3,), np.float64, 0.0), "magnet_type": ((), np.int32, 1),
        "stfcn_index": ((), np.int32, 1), "incl_it": ((), np.int32, 1),
        "use_sym": ((), np.int32, 0), "sym_op": ((3,), np.float64, 1.0),
        "M_rel": ((), np.float64, 0.0),
    }

    def __init__(self, n: int, center_pos: Optional[List] = None, dev_center: Optional[List] = None,
                 size: Optional[List] = None, vertices: Optional[List] = None,
                 mu_r_ea: Union[int, List, None] = None, mu_r_oa: Union[int, List, None] = None,
                 tile_type: Union[int, List, None] = None, offset: Optional[List] = None,
                 rot: Optional[List] = None, M_rem: Union[int, List, None] = None,
                 easy_axis: Optional[List] = None, color: Optional[List] = None,
                 magnet_type: Optional[List] = None, mag_angle: Optional[List] = None) -> None:
        for name, (shape, dtype, fill) in self._LAYOUT.items():
            setattr(self, "_" + name, np.full((n, *shape), fill, dtype=dtype, order="F"))
        self._n = n
        # same assignment order as the reference constructor (magstatics.py:116-141):
        # tile_type before rot (spheroid axes), M_rem before the easy axis (M = M_rem u_ea)
        for name, val in (("center_pos", center_pos), ("dev_center", dev_center), ("size", size),
                          ("vertices", vertices), ("mu_r_ea", mu_r_ea), ("mu_r_oa", mu_r_oa),
                          ("tile_type", tile_type), ("offset", offset), ("rot", rot), ("M_rem", M_rem)):
            if val is not None:
                setattr(self, name, val)
        if easy_axis is not None:
            self.u_ea = easy_axis
        else:
            self.set_easy_axis(mag_angle)
        if color is not None:
            self.color = color
        if magnet_type is not None:
            self.magnet_type = magnet_type

    def __str__(self):
        return "".join(f"Tile_{i} with coordinates {self.offset[i]}.\n" for i in range(self.n))

    def __getitem__(self, index):
        if index < 0 or index >= self.n:
            raise IndexError("Index out of bounds")
        return Tile(self, index)

    @property
    def n(self):
        return self._n

    @n.setter
    def n(self, val):
        self._n = val

    # ---- setter conventions ---------------------------------------------------
    def _set_rows(self, name, val):
        """3-vector attributes: (row, i) tuple, one broadcast row, or n rows."""
        arr = getattr(self, "_" + name)
        if isinstance(val, Tuple):
            arr[val[1]] = np.asarray(val[0])
        elif isinstance(val[0], (int, float)):
            arr[:] = np.asarray(val)
        elif len(val) == self.n:
            setattr(self, "_" + name, np.asarray(val))

    def _set_scalars(self, name, val, dtype=None):
        """Per-tile scalars: (value, i) tuple, one broadcast value, or n values."""
        if isinstance(val, Tuple):
            getattr(self, "_" + name)[val[1]] = val[0]
        elif isinstance(val, (int, float)):
            setattr(self, "_" + name, np.asarray([val] * self.n, dtype=dtype))
        elif len(val) == self.n:
            setattr(self, "_" + name, np.asarray(val, dtype=dtype))

    center_pos = property(lambda self: self._center_pos, lambda self, v: self._set_rows("center_pos", v))
    dev_center = property(lambda self: self._dev_center, lambda self, v: self._set_rows("dev_center", v))
    size = property(lambda self: self._size, lambda self, v: self._set_rows("size", v))
    offset = property(lambda self: self._offset, lambda self, v: self._set_rows("offset", v))
    color = property(lambda self: self._color, lambda self, v: self._set_rows("color", v))
    mu_r_ea = property(lambda self: self._mu_r_ea, lambda self, v: self._set_scalars("mu_r_ea", v))
    mu_r_oa = property(lambda self: self._mu_r_oa, lambda self, v: self._set_scalars("mu_r_oa", v))
    M_rem = property(lambda self: self._M_rem, lambda self, v: self._set_scalars("M_rem", v))
    magnet_type = property(lambda self: self._magnet_type, lambda self, v: self._set_scalars("magnet_type", v))
    stfcn_index = property(lambda self: self._stfcn_index, lambda self, v: self._set_scalars("stfcn_index", v))
    incl_it = property(lambda self: self._incl_it, lambda self, v: self._set_scalars("incl_it", v))
    use_sym = property(lambda self: self._use_sym)
    sym_op = property(lambda self: self._sym_op)
    u_oa1 = property(lambda self: self._u_oa1)
    u_oa2 = property(lambda self: self._u_oa2)

    @property
    def vertices(self):
        return self._vertices

    @vertices.setter
    def vertices(self, val):
        if isinstance(val, Tuple):
            v = np.asarray(val[0])
            if v.shape not in ((3, 4), (4, 3)):
                raise ValueError("Four 3-D vertices have to be defined!")
            self._vertices[val[1]] = v if v.shape == (3, 4) else v.T
            return
        v = np.asarray(val)
        if v.shape == (3, 4):
            self._vertices[:] = v
        elif v.shape == (4, 3):
            self._vertices[:] = v.T
        else:
            assert v[0].shape == (3, 4)
            self._vertices = v

    @property
    def tile_type(self):
        return self._tile_type

    @tile_type.setter
    def tile_type(self, val):
        if isinstance(val, Tuple):
            self._tile_type[val[1]] = val[0]
        elif isinstance(val, (int, float)):
            self._tile_type = np.asarray([val] * self.n, dtype=np.int32)
        elif len(val) == self.n:
            self._tile_type = np.asarray(val, dtype=np.int32)

    @property
    def rot(self):
        return self._rot

    @rot.setter
    def rot(self, val):
        # spheroids (tile_type 7) take a rotation axis, stored as Euler angles
        def conv(v, i):
            return _euler_to_rot_axis(v) if self.tile_type[i] == 7 else np.asarray(v)

        if isinstance(val, Tuple):
            self._rot[val[1]] = conv(val[0], val[1])
        elif isinstance(val[0], (List, np.ndarray)):
            for i in range(self.n):
                self._rot[i] = conv(val[i], i)
        else:
            self._rot[0] = conv(val, 0)

    @property
    def M(self):
        return self._M

    @M.setter
    def M(self, val):
        if isinstance(val, Tuple):
            self._M[val[1]] = val[0]
        elif isinstance(val[0], (int, float)):
            assert len(val) == 3
            self._M = np.asarray([val] * self.n)
        elif len(val) == self.n:
            assert len(val[0]) == 3
            self._M = np.asarray(val)

    @property
    def M_rel(self):
        return self._M_rel

    @M_rel.setter
    def M_rel(self, val):
        if isinstance(val, Tuple):
            self._M_rel[val[1]] = val[0]
        elif isinstance(val, (int, float)):
            self._M_rel = [val] * self.n
        elif len(val) == self.n:
            self._M_rel = val

    @property
    def u_ea(self):
        return self._u_ea

    @u_ea.setter
    def u_ea(self, val):
        # unit easy axis rounded to 9 decimals, M = M_rem u_ea (magstatics.py:313-331)
        if isinstance(val, Tuple):
            i = val[1]
            self._u_ea[i] = np.around(val[0] / np.linalg.norm(val[0]), decimals=9)
            self.M = (self.M_rem[i] * self.u_ea[i], i)
            return
        if isinstance(val[0], (int, float)):
            val = [val] * self.n
        for i, ea in enumerate(val):
            self._u_ea[i] = np.around(ea / np.linalg.norm(ea), decimals=9)
            self.M = (self.M_rem[i] * self.u_ea[i], i)
            w = np.array([1, 0, 0]) if (ea[1] != 0 or ea[2] != 0) else np.array([0, 1, 0])
            self._u_oa1[i] = np.around(np.cross(self.u_ea[i], w), decimals=9)
            self._u_oa2[i] = np.around(np.cross(self.u_ea[i], self.u_oa1[i]), decimals=9)

    def set_easy_axis(self, val: Optional[List] = None, idx: Optional[int] = None, seed: int = 42) -> None:
        """Easy axis from spherical angles (polar in [0, pi], azimuth in [0, 2 pi]);
        random angles from ``default_rng(seed)`` when ``val`` is None."""
        rand = np.random.default_rng(seed).random(size=(self.n, 2)) if val is None else None
        span = [np.pi, 2 * np.pi]
        if idx is not None:
            self._set_ea_i(span * rand[idx] if val is None else val, idx)
            return
        for i in range(self.n):
            if val is None:
                self._set_ea_i(span * rand[i], i)
            elif isinstance(val[0], (int, float)):
                self._set_ea_i(val, i)
            else:
                self._set_ea_i(val[i], i)

    def _set_ea_i(self, val, i):
        if isinstance(val, (int, float)):
            raise ValueError("Both spherical angles have to be set!")
        theta, phi = val
        self.u_ea = ([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)], i)
        self._u_oa1[i] = [np.sin(theta) * np.sin(phi), np.sin(theta) * (-np.cos(phi)), 0]
        self._u_oa2[i] = [0.5 * np.sin(2 * theta) * np.cos(phi), 0.5 * np.sin(2 * theta) * np.sin(phi),
                          -1 * np.sin(theta) ** 2]

    def _add_tiles(self, n: int) -> None:
        for name, (shape, dtype, fill) in self._LAYOUT.items():
            old = getattr(self, "_" + name)
            new = np.full((n, *shape), fill, dtype=dtype, order="F")
            setattr(self, "_" + name, np.append(old, new, axis=0))
        self._n += n

    def refine_prism(self, idx: Union[int, List], mat: List):
        """Split prism(s) ``idx`` into mat[0] x mat[1] x mat[2] sub-prisms
        (magstatics.py:571-616); tile ``idx`` becomes the last sub-prism."""
        for i in ([idx] if isinstance(idx, (int, float)) else idx):
            self._refine_prism_i(i, mat)

    def _refine_prism_i(self, i, mat):
        assert self.tile_type[i] == PRISM
        old_n = self.n
        self._add_tiles(int(np.prod(mat)) - 1)
        R = get_rotmat(self.rot[i])
        step = self.size[i] / mat
        c0 = -self.size[i] / 2 + step / 2
        local = [c0 + [step[0] * j, step[1] * k, step[2] * m]
                 for j in range(mat[0]) for k in range(mat[1]) for m in range(mat[2])]
        centres = (np.dot(R, np.array(local).T)).T + self.offset[i]
        copied = ("M", "M_rel", "color", "magnet_type", "mu_r_ea", "mu_r_oa", "rot", "tile_type", "u_ea", "u_oa1",
                  "u_oa2", "incl_it")
        for q in range(len(local)):
            dst = q + old_n
            if dst == self.n:  # the last sub-prism reuses the original slot
                self._offset[i] = centres[q]
                self._size[i] = self.size[i] / mat
                continue
            self._size[dst] = self.size[i] / mat
            for name in copied:
                arr = getattr(self, "_" + name)
                arr[dst] = arr[i]
            self._offset[dst] = centres[q]


def _euler_to_rot_axis(euler):
    """Rotation axis of a spheroid -> the Euler angles MagTense stores
    (magstatics.py:619-646)."""
    Ry = np.array([[np.cos(np.pi / 2), 0, np.sin(np.pi / 2)], [0, 1, 0],
                   [-np.sin(np.pi / 2), 0, np.cos(np.pi / 2)]])
    ax = np.dot(Ry, np.asarray(euler).T)
    rot_x = -np.arctan2(ax[1], ax[0])
    rot_y = np.arccos(ax[2] / np.sqrt(ax[0] ** 2 + ax[1] ** 2 + ax[2] ** 2)) - np.pi / 2
    return np.array([rot_x, rot_y, 0])


def get_rotmat(rot):
    """Rx(rot[0]) Ry(rot[1]) Rz(rot[2]) (magstatics.py:649-669)."""
    c, s = np.cos, np.sin
    rx = np.array([[1, 0, 0], [0, c(rot[0]), -s(rot[0])], [0, s(rot[0]), c(rot[0])]])
    ry = np.array([[c(rot[1]), 0, s(rot[1])], [0, 1, 0], [-s(rot[1]), 0, c(rot[1])]])
    rz = np.array([[c(rot[2]), -s(rot[2]), 0], [s(rot[2]), c(rot[2]), 0], [0, 0, 1]])
    return rx @ ry @ rz


def grid_config(spots: Union[List, np.ndarray], area: Union[List, np.ndarray], filled_pos: Optional[List] = None,
                n_pts: List = [20, 20, 1], mode: str = "uniform", n_tiles: Optional[int] = None,
                mag_angles: Optional[List] = None, B_rem: float = 1.2, seed: int = 42) -> Tuple[Tiles, np.ndarray]:
    """Prism magnets on a ``spots`` lattice filling ``area`` plus evaluation
    points (magstatics.py:672-741)."""
    spots = np.asarray(spots)
    area = np.asarray(area)
    rng = np.random.default_rng(seed)
    if mode == "uniform":
        seg = area / np.asarray(n_pts)
        pts = [[(i + 0.5) * seg[0], (j + 0.5) * seg[1], (k + 0.5) * seg[2]]
               for i in range(n_pts[0]) for j in range(n_pts[1]) for k in range(n_pts[2])]
    elif mode == "center":
        seg = (area[0] / 10) / n_pts[0]
        seg_angle = np.pi / n_pts[1]
        seg_layer = area[2] / n_pts[2]
        pts = [[i * seg, j * seg_angle, k * seg_layer]
               for i in range(n_pts[0]) for j in range(n_pts[1]) for k in range(n_pts[2])]
    else:
        raise NotImplementedError()

    if filled_pos is None:
        total = int(np.prod(spots))
        if n_tiles is None:
            n_tiles = 1 + rng.integers(total)
        elif n_tiles < 1 or n_tiles > total:
            raise ValueError("n_tiles out of range!")
        lin = rng.choice(total, size=n_tiles, replace=False)
        plane = spots[0] * spots[1]
        filled_pos = np.stack([lin % spots[0], (lin % plane) // spots[0], lin // plane], axis=1).astype(int).tolist()
    elif len(filled_pos) == 0:
        raise ValueError("filled_pos is empty!")

    tiles = Tiles(n=len(filled_pos), size=area / spots, tile_type=PRISM, M_rem=B_rem / MU0, color=[1, 0, 0],
                  mag_angle=mag_angles, offset=[(np.asarray(pos) + 0.5) * (area / spots) for pos in filled_pos])
    return tiles, np.asarray(pts, dtype=np.float64, order="F")


# ---- GPU evaluation -------------------------------------------------------------

def _tiles_struct(tiles: Tiles, with_m: bool):
    """Fortran-ordered copies of the tile arrays and the mt_tiles struct over them."""
    n = tiles.n
    types = np.asarray(tiles.tile_type).reshape(n)
    bad = np.nonzero((types != PRISM) & (types != TETRAHEDRON))[0]
    if bad.size:
        raise NotImplementedError(
            f"tile {int(bad[0])} has tile_type {int(tiles.tile_type[bad[0]])}: rectangular prisms (2) and "
            "tetrahedra (5) are supported by the MI355X backend")
    keep = {
        "size": np.asfortranarray(np.asarray(tiles.size, dtype=np.float64).reshape(n, 3)),
        "offset": np.asfortranarray(np.asarray(tiles.offset, dtype=np.float64).reshape(n, 3)),
        "rot": np.asfortranarray(np.asarray(tiles.rot, dtype=np.float64).reshape(n, 3)),
        "M": np.asfortranarray(np.asarray(tiles.M, dtype=np.float64).reshape(n, 3)),
        "tile_type": np.ascontiguousarray(np.asarray(tiles.tile_type, dtype=np.int32).reshape(n)),
        "use_sym": np.ascontiguousarray(np.asarray(tiles.use_sym, dtype=np.int32).reshape(n)),
        "sym_ops": np.asfortranarray(np.asarray(tiles.sym_op, dtype=np.float64).reshape(n, 3)),
        "vertices": np.asfortranarray(np.asarray(tiles.vertices, dtype=np.float64).reshape(n, 3, 4)),
    }
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    st = _lib.MtTiles()
    st.n_tiles = n
    for k in ("size", "offset", "rot", "sym_ops"):
        setattr(st, k, keep[k].ctypes.data_as(dp))
    st.M = keep["M"].ctypes.data_as(dp) if with_m else None
    st.tile_type = keep["tile_type"].ctypes.data_as(ip)
    st.use_sym = keep["use_sym"].ctypes.data_as(ip)
    st.vertices = keep["vertices"].ctypes.data_as(dp)
    return st, keep


def _points(pts):
    P = np.asfortranarray(np.asarray(pts, dtype=np.float64).reshape(-1, 3))
    if P.shape[0] == 0:
        raise ValueError("no evaluation points")
    return P


def get_demag_tensor(tiles: Tiles, pts: np.ndarray) -> np.ndarray:
    """Demagnetisation tensor of every tile at every point:
    N(n_tiles, n_pts, 3, 3), float64, Fortran order, MagTense sign, each tile's
    tensor in its rotated frame (getNFromTiles, FortranToPythonIO.f90:25-93)."""
    st, keep = _tiles_struct(tiles, with_m=False)
    P = _points(pts)
    N = np.zeros((tiles.n, P.shape[0], 3, 3), dtype=np.float64, order="F")
    L = _lib.lib()
    _lib.check(L.mt_tiles_demag_tensor(C.byref(st), P.shape[0], P.ctypes.data, N.ctypes.data))
    return N


def get_H_field(tiles: Tiles, pts: np.ndarray, demag_tensor: Optional[np.ndarray] = None) -> np.ndarray:
    """Field H(n_pts, 3) [A/m] of all tiles at the points; a precomputed
    ``demag_tensor`` (from ``get_demag_tensor``) skips the tensor evaluation
    (getHFromTiles, FortranToPythonIO.f90:106-223)."""
    st, keep = _tiles_struct(tiles, with_m=True)
    P = _points(pts)
    Nin = None
    if demag_tensor is not None:
        Nin = np.asfortranarray(np.asarray(demag_tensor, dtype=np.float64))
        if Nin.shape != (tiles.n, P.shape[0], 3, 3):
            raise ValueError(f"demag_tensor has shape {Nin.shape}, expected {(tiles.n, P.shape[0], 3, 3)}")
    H = np.zeros((P.shape[0], 3), dtype=np.float64, order="F")
    L = _lib.lib()
    _lib.check(L.mt_tiles_h_field(C.byref(st), P.shape[0], P.ctypes.data,
                                  None if Nin is None else Nin.ctypes.data, H.ctypes.data))
    return H


def run_simulation(*args, **kwargs):
    """The reference's iterative magnetostatic solver (state functions, soft
    magnets) is not provided by the MI355X backend."""
    raise NotImplementedError("magstatics.run_simulation (iterative magnetisation) is not supported; "
                              "use get_H_field for hard prism magnets")


def iterate_magnetization(*args, **kwargs):
    raise NotImplementedError("magstatics.iterate_magnetization is not supported by the MI355X backend")
