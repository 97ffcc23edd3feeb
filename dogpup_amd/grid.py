"""Voxel-grid basis, mesh-to-mesh interpolation and display slicing.

Host side of ``mesh2grid`` (``DOGPUP/dMesh.m:286-380``), ``mesh2mesh``
(``:383-411``) and ``mesh_slice``/``mesh_slice3`` (``:414-517``).  The
geometry runs in libdogpup_hip.so:
  - ``Locator``: binned point location in the tetrahedra
    (dogpup_locator_*), replacing triangulation/pointLocation;
  - grid -> mesh weights in the grid's Kuhn split with the reference's
    in-mesh filter, nearest-voxel fallback and row normalisation
    (dogpup_grid_to_mesh), replacing delaunay/pointLocation/nearestNeighbor;
  - every interpolation matrix is stored as 4 (index, weight) pairs per row
    and applied by dogpup_ell4_apply, to vectors or to whole fluence fields in
    the solver layout.

Grid points are ordered as ndgrid's [X(:), Y(:), Z(:)]: voxel
v = ix + nx*(iy + ny*iz).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib

DEV = "cuda"


def _f64(a):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=DEV)


def _i32(a):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.int32), device=DEV)


class Locator:
    """Point location in one tetrahedral mesh (dogpup_locator_create)."""

    def __init__(self, node, elem):
        node = node if torch.is_tensor(node) else _f64(node)
        elem = elem if torch.is_tensor(elem) else _i32(elem)
        if node.dtype != torch.float64 or elem.dtype != torch.int32:
            raise TypeError("node must be float64 and elem int32")
        self.nn, self.ne = int(node.shape[0]), int(elem.shape[0])
        h = C.c_void_p()
        _lib.call("dogpup_locator_create", C.byref(h), _lib.ptr(node.contiguous()), self.nn,
                  _lib.ptr(elem.contiguous()), self.ne, _lib.stream_ptr())
        self.h = h

    def __del__(self):
        try:
            if getattr(self, "h", None):
                _lib.lib().dogpup_locator_destroy(self.h)
        except Exception:
            pass

    def stats(self):
        out = np.zeros(4, np.int64)
        _lib.call("dogpup_locator_stats", self.h, _lib.ptr(out))
        return {"bins": tuple(int(v) for v in out[:3]), "pairs": int(out[3])}

    def find(self, pts, tol=1e-9):
        """(eid, w, nodes) device tensors: containing element (-1 outside),
        barycentric weights (NaN outside) and the element's node ids (-1)."""
        p = pts if torch.is_tensor(pts) else _f64(np.atleast_2d(pts))
        p = p.to(DEV, torch.float64).contiguous()
        n = int(p.shape[0])
        eid = torch.empty(n, dtype=torch.int32, device=DEV)
        w = torch.empty((n, 4), dtype=torch.float64, device=DEV)
        nodes = torch.empty((n, 4), dtype=torch.int32, device=DEV)
        _lib.call("dogpup_locator_find", self.h, _lib.ptr(p), n, float(tol), _lib.ptr(eid), _lib.ptr(w),
                  _lib.ptr(nodes), _lib.stream_ptr())
        return eid, w, nodes


@dataclass
class Ell4:
    """Sparse rows x cols matrix with at most 4 entries per row (index -1 =
    no entry), applied on the device."""

    idx: torch.Tensor      # rows x 4 int32
    w: torch.Tensor        # rows x 4 float64
    rows: int
    cols: int

    def apply(self, x, width=None, ngroups=1):
        """y = M x.  x: cols (x width) float64, or a [ngroups][cols][width]
        block (e.g. a complex128 solver-layout field viewed as real pairs)."""
        xt = x if torch.is_tensor(x) else _f64(x)
        xt = xt.to(DEV, torch.float64).contiguous()
        if width is None:
            width = 1 if xt.dim() == 1 else int(xt.shape[-1])
        if xt.numel() != ngroups * self.cols * width:
            raise ValueError(f"input has {xt.numel()} values, expected {ngroups} x {self.cols} x {width}")
        out = torch.empty(ngroups * self.rows * width, dtype=torch.float64, device=DEV)
        _lib.call("dogpup_ell4_apply", _lib.ptr(xt), _lib.ptr(out), _lib.ptr(self.idx), _lib.ptr(self.w),
                  self.rows, self.cols, int(width), int(ngroups), _lib.stream_ptr())
        if xt.dim() == 1 and width == 1 and ngroups == 1:
            return out
        return out.view(ngroups, self.rows, width) if ngroups > 1 else out.view(self.rows, width)

    def rows_subset(self, rows) -> "Ell4":
        r = torch.as_tensor(rows, device=DEV, dtype=torch.long)
        return Ell4(self.idx.index_select(0, r).contiguous(), self.w.index_select(0, r).contiguous(),
                    int(r.numel()), self.cols)

    def to_scipy(self):
        import scipy.sparse as sp
        idx = self.idx.cpu().numpy()
        w = self.w.cpu().numpy()
        keep = idx >= 0
        r = np.repeat(np.arange(self.rows), 4).reshape(self.rows, 4)[keep]
        return sp.csr_matrix((w[keep], (r, idx[keep])), shape=(self.rows, self.cols))


def uniform_axis(v, name):
    v = np.asarray(v, float).reshape(-1)
    if v.size < 2:
        raise ValueError(f"grid axis {name} needs at least 2 points")
    d = np.diff(v)
    if not np.all(d > 0) or np.max(np.abs(d - d[0])) > 1e-9 * max(abs(d[0]), 1.0):
        raise ValueError(f"grid axis {name} must be uniformly spaced and increasing")
    return float(v[0]), float(d[0]), int(v.size)


def ndgrid_points(x, y, z):
    """[X(:), Y(:), Z(:)] of ndgrid(x, y, z), x fastest (dMesh.m:308-309)."""
    X, Y, Z = np.meshgrid(np.asarray(x, float), np.asarray(y, float), np.asarray(z, float), indexing="ij")
    return np.stack([X.ravel(order="F"), Y.ravel(order="F"), Z.ravel(order="F")], axis=1)


@dataclass
class GridBasis:
    """What mesh2grid attaches to the mesh (dMesh.m:294-300) plus the
    device-side matrices: m2g (NV x NN), g2m (NN x NV), the in-mesh voxel
    list and m2g restricted to it."""

    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    grid: np.ndarray              # NV x 3
    gridSize: tuple
    dxyz: float
    gridinmesh: np.ndarray        # NV bool
    m2g: Ell4
    g2m: Ell4
    in_idx: np.ndarray            # in-mesh voxel ids, ascending
    m2g_in: Ell4                  # m2g rows of the in-mesh voxels
    n_fallback: int = 0

    @property
    def nv(self):
        return self.grid.shape[0]

    @property
    def nv_in(self):
        return int(self.in_idx.size)

    def to_grid(self, f):
        """m2g * f for a nodal vector (NN) -> NV."""
        return self.m2g.apply(f)

    def to_mesh(self, g):
        """g2m * g for a voxel vector (NV) -> NN."""
        return self.g2m.apply(g)

    def field_to_voxels(self, field):
        """m2g applied to every column and frequency of a FluenceField: a
        FluenceField over the in-mesh voxels (rows = in_idx)."""
        from .mesh import FluenceField
        ng = field.data.shape[0]
        x = torch.view_as_real(field.data).reshape(ng, field.nn, 2 * field.nsp)
        y = self.m2g_in.apply(x, width=2 * field.nsp, ngroups=ng)
        data = torch.view_as_complex(y.view(ng, self.nv_in, field.nsp, 2))
        return FluenceField(data, self.nv_in, field.nf, field.ncols, field.nsp, field.iters, field.relres,
                            field.status, dict(field.stats))


def mesh2grid(node, elem, x, y, z, tol=1e-9, locator: Locator | None = None) -> GridBasis:
    """Voxel basis of a mesh (dMesh.m:286-380): m2g from the containing
    tetrahedron of each grid point, gridinmesh, g2m from each node's tetrahedron
    of the grid's Kuhn split (see dogpup_grid_to_mesh)."""
    o = []
    for v, name in ((x, "x"), (y, "y"), (z, "z")):
        o.append(uniform_axis(v, name))
    origin = np.array([a[0] for a in o])
    spacing = np.array([a[1] for a in o])
    dims = np.array([a[2] for a in o], np.int32)
    node_np = node.cpu().numpy() if torch.is_tensor(node) else np.asarray(node, float)
    loc = locator or Locator(node, elem)
    grid = ndgrid_points(x, y, z)
    nv = grid.shape[0]
    eid, w, nodes = loc.find(_f64(grid), tol)
    inmesh_d = (eid >= 0)
    w = torch.where(inmesh_d[:, None], w, torch.zeros_like(w))
    m2g = Ell4(nodes, w.contiguous(), nv, loc.nn)
    inm_u8 = inmesh_d.to(torch.uint8).contiguous()
    nn = node_np.shape[0]
    gi = torch.empty((nn, 4), dtype=torch.int32, device=DEV)
    gw = torch.empty((nn, 4), dtype=torch.float64, device=DEV)
    counts = np.zeros(2, np.int32)
    node_d = node if torch.is_tensor(node) else _f64(node_np)
    _lib.call("dogpup_grid_to_mesh", _lib.ptr(node_d.contiguous()), nn, _lib.ptr(origin), _lib.ptr(spacing),
              _lib.ptr(dims), _lib.ptr(inm_u8), _lib.ptr(gi), _lib.ptr(gw), _lib.ptr(counts), _lib.stream_ptr())
    g2m = Ell4(gi, gw, nn, nv)
    inmesh = inmesh_d.cpu().numpy()
    in_idx = np.flatnonzero(inmesh)
    return GridBasis(np.asarray(x, float), np.asarray(y, float), np.asarray(z, float), grid,
                     (int(dims[0]), int(dims[1]), int(dims[2])), abs(float(spacing[0])), inmesh, m2g, g2m, in_idx,
                     m2g.rows_subset(in_idx), int(counts[0]))


def mesh2mesh(node, val, old_node, old_elem, tol=1e-9, locator: Locator | None = None):
    """Interpolate ``val`` (old NN x C, or old NN) from the old mesh to
    ``node`` (dMesh.m:383-411).  Points outside the old mesh take the value
    of the nearest located point (the reference's intent; its own fill
    indexes coordinates instead of node ids)."""
    loc = locator or Locator(old_node, old_elem)
    node_np = node.cpu().numpy() if torch.is_tensor(node) else np.asarray(node, float)
    eid, w, nodes = loc.find(_f64(node_np), tol)
    ok = eid >= 0
    w = torch.where(ok[:, None], w, torch.zeros_like(w))
    v = val if torch.is_tensor(val) else _f64(val)
    v = v.to(DEV, torch.float64)
    squeeze = v.dim() == 1
    v2 = v.reshape(loc.nn, -1).contiguous()
    out = Ell4(nodes, w.contiguous(), node_np.shape[0], loc.nn).apply(v2, width=v2.shape[1])
    out = out.reshape(node_np.shape[0], -1)
    okh = ok.cpu().numpy()
    if not okh.all():
        if not okh.any():
            raise ValueError("no point of the new mesh lies in the old mesh")
        good = np.flatnonzero(okh)
        bad = np.flatnonzero(~okh)
        from scipy.spatial import cKDTree
        _, k = cKDTree(node_np[good]).query(node_np[bad])
        out[torch.as_tensor(bad, device=DEV)] = out[torch.as_tensor(good[k], device=DEV)]
    return out[:, 0] if squeeze else out


def _slice_points(node, plane, res=256):
    if not isinstance(plane, str):
        raise TypeError("Char input required for slicing plane")
    key = plane[:2].lower()
    if key not in ("x=", "y=", "z="):
        raise ValueError("Format for slicing plane is similar to 'x=30'")
    lo, hi = float(np.min(node)), float(np.max(node))
    x = np.linspace(lo, hi, res)
    c = float(plane.split("=", 1)[1])
    if key == "x=":
        X, Y, Z = np.meshgrid([c], x, x)
    elif key == "y=":
        X, Y, Z = np.meshgrid(x, [c], x)
    else:
        X, Y, Z = np.meshgrid(x, x, [c])
    return np.stack([X.ravel(order="F"), Y.ravel(order="F"), Z.ravel(order="F")], axis=1)


def mesh_slice(node, elem, plane, locator: Locator | None = None):
    """(sliceMat, points) of mesh_slice (dMesh.m:414-470): a 256 x 256 plane
    of the mesh's bounding cube; sliceMat (as Ell4) interpolates nodal values
    to it.  Points outside the mesh get NaN, as the reference's NaN entries
    produce (returned as the ``outside`` mask)."""
    node_np = node.cpu().numpy() if torch.is_tensor(node) else np.asarray(node, float)
    pts = _slice_points(node_np, plane)
    loc = locator or Locator(node, elem)
    eid, w, nodes = loc.find(_f64(pts))
    ok = eid >= 0
    M = Ell4(nodes, torch.where(ok[:, None], w, torch.zeros_like(w)).contiguous(), pts.shape[0], loc.nn)
    return M, pts, (~ok).cpu().numpy()


def mesh_slice3(node, elem, locator: Locator | None = None):
    """(volMat, points) of mesh_slice3 (dMesh.m:476-517): a 64^3 grid over
    the bounding cube padded by one cell; rows outside the mesh are zero."""
    node_np = node.cpu().numpy() if torch.is_tensor(node) else np.asarray(node, float)
    lo, hi = float(np.min(node_np)), float(np.max(node_np))
    x = np.linspace(lo, hi, 62)
    dx = x[1] - x[0]
    x = np.concatenate([[x[0] - dx], x, [x[-1] + dx]])
    X, Y, Z = np.meshgrid(x, x, x)
    pts = np.stack([X.ravel(order="F"), Y.ravel(order="F"), Z.ravel(order="F")], axis=1)
    loc = locator or Locator(node, elem)
    eid, w, nodes = loc.find(_f64(pts))
    ok = eid >= 0
    M = Ell4(nodes, torch.where(ok[:, None], w, torch.zeros_like(w)).contiguous(), pts.shape[0], loc.nn)
    return M, pts
