"""Mesh and optode files, and iso2mesh-style box meshes.

The reference's users mesh with iso2mesh (``meshabox``, ``sortmesh``,
``tutorials/tut_2_fluence_and_sensitivity.m:5-7``) and keep meshes and optodes
in MATLAB .mat files (``tutorials/tut_3_weighted_reconstruction.m:4``).
MATLAB stores class objects (dMesh, dOptode) in .mat files as opaque MCOS
data that no reader outside MATLAB decodes, so the portable form is the
objects' fields as plain arrays or structs (in MATLAB: ``s = struct(mesh)``,
or saving node/elem/... directly).  This module reads and writes that form
as MAT v5/v7 (``scipy.io``; v7.3 files are HDF5 and need h5py):

  mesh   : node (NN x 3 mm), elem (NE x 4, or NE x 5 with iso2mesh's region
           label; 1-based), nr, mua, musp (scalar or NN) -- the dMesh
           constructor's arguments (dMesh.m:48-77);
  optode : s_positions, d_positions (N x 3), link (NM x 2, 1-based),
           s_avgPow, s_tpsf (or tpsf), tAxis, Nf -- the dOptode
           constructor's arguments (dOptode.m:34-74).

Fields are found at the top level of the file or inside any struct
variable; ``mesh_var``/``optode_var`` pick a struct by name.
"""
from __future__ import annotations

import math

import numpy as np

MESH_KEYS = ("node", "elem")
OPTODE_KEYS = ("s_positions", "d_positions", "link")


def _plain(v):
    """scipy.io mat_struct / 1x1 object arrays -> dict / ndarray."""
    from scipy.io.matlab import mat_struct
    if isinstance(v, mat_struct):
        return {k: _plain(getattr(v, k)) for k in v._fieldnames}
    if isinstance(v, np.ndarray) and v.dtype == object and v.size == 1:
        return _plain(v.item())
    if isinstance(v, np.ndarray) and v.dtype.names is not None and v.size == 1:
        return {k: _plain(v[k].item() if v[k].size == 1 else v[k]) for k in v.dtype.names}
    return v


def _find(tree: dict, keys, name=None):
    if name is not None:
        if name not in tree or not isinstance(tree[name], dict):
            raise KeyError(f"no struct variable {name!r} in the file")
        return tree[name]
    if all(k in tree for k in keys):
        return tree
    for v in tree.values():
        if isinstance(v, dict) and all(k in v for k in keys):
            return v
    return None


def read_mat(path, mesh_var=None, optode_var=None) -> dict:
    """Plain arrays of a mesh (and optode, if present) from a .mat file, with
    indices converted to 0-based: {'node', 'elem', 'nr', 'mua', 'musp',
    'region' (or None), 'optode': {...} or None}."""
    import scipy.io
    try:
        raw = scipy.io.loadmat(path, squeeze_me=True, struct_as_record=False)
    except NotImplementedError as e:      # MAT v7.3 (HDF5)
        raise ValueError(f"{path}: MAT v7.3 (HDF5) files need h5py; save with -v7 instead") from e
    tree = {k: _plain(v) for k, v in raw.items() if not k.startswith("__")}
    ms = _find(tree, MESH_KEYS, mesh_var)
    if ms is None:
        raise KeyError(f"{path}: no node/elem arrays found (top level or inside a struct)")
    node = np.asarray(ms["node"], float).reshape(-1, 3)
    elem = np.asarray(ms["elem"])
    elem = elem.reshape(-1, elem.shape[-1] if elem.ndim == 2 else 4)
    region = elem[:, 4].astype(np.int64) if elem.shape[1] >= 5 else None
    elem = elem[:, :4].astype(np.int64) - 1
    if elem.min() < 0 or elem.max() >= node.shape[0]:
        raise ValueError(f"{path}: elem indices out of range for {node.shape[0]} nodes (expected 1-based)")
    nn = node.shape[0]
    out = {"node": node, "elem": elem, "region": region,
           "nr": float(ms.get("nr", 1.4)),
           "mua": np.broadcast_to(np.asarray(ms.get("mua", 0.01), float).reshape(-1), (nn,)).copy()
           if np.size(ms.get("mua", 0.01)) in (1, nn) else None,
           "musp": np.broadcast_to(np.asarray(ms.get("musp", 1.0), float).reshape(-1), (nn,)).copy()
           if np.size(ms.get("musp", 1.0)) in (1, nn) else None}
    if out["mua"] is None or out["musp"] is None:
        raise ValueError(f"{path}: mua/musp must be scalars or one value per node")
    os_ = _find(tree, OPTODE_KEYS, optode_var)
    if os_ is None and isinstance(ms.get("optode"), dict):
        os_ = ms["optode"]
    if os_ is not None:
        t = np.asarray(os_["tAxis"], float).reshape(-1)
        tpsf = os_.get("s_tpsf", os_.get("tpsf"))
        out["optode"] = {
            "s_positions": np.asarray(os_["s_positions"], float).reshape(-1, 3),
            "d_positions": np.asarray(os_["d_positions"], float).reshape(-1, 3),
            "link": np.asarray(os_["link"], np.int64).reshape(-1, 2) - 1,
            "avgPow": float(os_.get("s_avgPow", os_.get("avgPow", 1.0))),
            "tpsf": np.asarray(tpsf, float).reshape(-1),
            "tAxis": t,
            "Nf": int(os_["Nf"]) if "Nf" in os_ else int(np.size(os_["fAxis"])),
        }
    else:
        out["optode"] = None
    return out


def write_mat(path, node, elem, nr, mua, musp, optode: dict | None = None, region=None):
    """Write the plain form read_mat reads (1-based indices, MAT v5 structs
    'mesh' and 'optode')."""
    import scipy.io
    elem = np.asarray(elem, np.int64) + 1
    if region is not None:
        elem = np.concatenate([elem, np.asarray(region, np.int64).reshape(-1, 1)], axis=1)
    mesh = {"node": np.asarray(node, float), "elem": elem.astype(np.float64), "nr": float(nr),
            "mua": np.asarray(mua, float).reshape(-1, 1), "musp": np.asarray(musp, float).reshape(-1, 1)}
    data = {"mesh": mesh}
    if optode is not None:
        data["optode"] = {"s_positions": np.asarray(optode["s_positions"], float),
                          "d_positions": np.asarray(optode["d_positions"], float),
                          "link": (np.asarray(optode["link"], np.int64) + 1).astype(np.float64),
                          "s_avgPow": float(optode["avgPow"]), "s_tpsf": np.asarray(optode["tpsf"], float),
                          "tAxis": np.asarray(optode["tAxis"], float), "Nf": float(optode["Nf"])}
    scipy.io.savemat(path, data, do_compression=True, oned_as="row")


def load_mat(path, mesh_var=None, optode_var=None, reference_faces=True):
    """A DMesh (with its DOptode added and snapped, if the file holds one)
    from a .mat file."""
    from .mesh import DMesh, DOptode
    d = read_mat(path, mesh_var, optode_var)
    D = DMesh(d["node"], d["elem"], d["nr"], np.stack([d["mua"], d["musp"]], 1), reference_faces)
    o = d["optode"]
    if o is not None:
        D.add_optode(DOptode(o["s_positions"], o["d_positions"], o["link"], o["avgPow"], o["tpsf"], o["tAxis"],
                             o["Nf"]))
    return D


def save_mat(path, mesh, raw_positions=None):
    """Save a DMesh (and its optode) in the plain form.  The optode's stored
    positions are the snapped ones, as in the reference; pass
    ``raw_positions=(s_pos, d_pos)`` to store others."""
    opt = None
    if mesh.optode is not None:
        o = mesh.optode
        s_pos, d_pos = raw_positions if raw_positions is not None else (o.s_positions, o.d_positions)
        opt = {"s_positions": s_pos, "d_positions": d_pos, "link": o.link, "avgPow": o.s_avgPow,
               "tpsf": o.s_tpsf, "tAxis": o.tAxis, "Nf": o.Nf}
    write_mat(path, mesh.node, mesh.elem, mesh.nr, mesh.mua, mesh.musp, opt)


def meshabox(p0, p1, maxvol):
    """Box mesh between corners p0 and p1 whose tetrahedra have volume at most
    ``maxvol`` mm^3, in the call form of iso2mesh's meshabox (used at
    tutorials/tut_1_meshing.m:5): returns (node, face, elem), elem 0-based
    NE x 4.  iso2mesh runs an external Delaunay mesher (TetGen) for an
    unstructured mesh; this builds the structured Kuhn (6-tetrahedra-per-cube)
    lattice instead, with per-axis spacing L/n <= (6 maxvol)^(1/3) so every
    tetrahedron (volume hx*hy*hz/6) satisfies the bound."""
    from .meshgen import kuhn_lattice, orient_positive
    from .topology import true_faces
    p0 = np.asarray(p0, float)
    p1 = np.asarray(p1, float)
    L = p1 - p0
    if np.any(L <= 0) or maxvol <= 0:
        raise ValueError("need p1 > p0 on every axis and maxvol > 0")
    h = (6.0 * maxvol) ** (1.0 / 3.0)
    n = np.maximum(1, np.ceil(L / h - 1e-9)).astype(int)
    node, elem = kuhn_lattice(n[0] + 1, n[1] + 1, n[2] + 1, 1.0)
    node = p0 + node * (L / n)
    elem = orient_positive(node, elem).astype(np.int64)
    face, _ = true_faces(node, elem)
    return node, face, elem


def sortmesh(node, elem):
    """Renumber nodes along a Morton curve and sort elements by their
    smallest node (the purpose of iso2mesh's sortmesh, tut_2:7: locality of
    the node and element order).  Returns (node, elem, order) with
    node_new = node[order]."""
    from .topology import morton_order
    node = np.asarray(node, float)
    elem = np.asarray(elem, np.int64)
    order = morton_order(node, elem)
    inv = np.empty_like(order)
    inv[order] = np.arange(order.size, dtype=order.dtype)
    e = inv[elem]
    e = e[np.lexsort((e.max(axis=1), e.min(axis=1)))]
    return node[order], e, order


def max_tet_volume(node, elem):
    x = np.asarray(node)[np.asarray(elem)]
    v = np.abs(np.einsum("ij,ij->i", np.cross(x[:, 1] - x[:, 0], x[:, 2] - x[:, 0]), x[:, 3] - x[:, 0])) / 6.0
    return float(v.max()) if v.size else math.nan
