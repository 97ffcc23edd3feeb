"""Mesh/optode .mat files (plain-struct form) and the iso2mesh-style box
mesher and reordering."""
import numpy as np
import pytest
import scipy.io

from dogpup_amd import meshgen, meshio, topology


def _vol(node, elem):
    x = node[elem]
    return np.einsum("ij,ij->i", np.cross(x[:, 1] - x[:, 0], x[:, 2] - x[:, 0]), x[:, 3] - x[:, 0]) / 6.0


def test_mat_round_trip_with_region_and_optode(tmp_path):
    m = meshgen.slab((10.0, 8.0, 6.0), 2.0, 0.01, 1.0, 1.4, seed=3)
    region = np.arange(m.elem.shape[0]) % 3 + 1
    opt = {"s_positions": np.array([[2.0, 2.0, 6.0], [8.0, 6.0, 6.0]]), "d_positions": np.array([[5.0, 4.0, 6.0]]),
           "link": np.array([[0, 0], [1, 0]]), "avgPow": 1e3, "tpsf": np.linspace(0, 1, 11),
           "tAxis": np.arange(11) * 1e-10, "Nf": 4}
    p = tmp_path / "m.mat"
    mua = m.mua.copy()
    mua[:5] = 0.02
    meshio.write_mat(p, m.node, m.elem, 1.33, mua, 0.9 * np.ones(m.node.shape[0]), opt, region)
    raw = scipy.io.loadmat(p)
    assert raw["mesh"]["elem"][0, 0].min() == 1 and raw["mesh"]["elem"][0, 0].shape[1] == 5   # 1-based + label
    d = meshio.read_mat(p)
    assert np.array_equal(d["elem"], m.elem) and np.allclose(d["node"], m.node)
    assert np.array_equal(d["region"], region)
    assert d["nr"] == 1.33 and np.allclose(d["mua"], mua) and np.allclose(d["musp"], 0.9)
    o = d["optode"]
    assert np.array_equal(o["link"], opt["link"]) and o["Nf"] == 4 and o["avgPow"] == 1e3
    assert np.allclose(o["s_positions"], opt["s_positions"]) and np.allclose(o["tpsf"], opt["tpsf"])


def test_read_top_level_arrays_and_named_struct(tmp_path):
    m = meshgen.slab((6.0, 6.0, 4.0), 2.0, 0.01, 1.0, 1.4, seed=1)
    p = tmp_path / "top.mat"
    scipy.io.savemat(p, {"node": m.node, "elem": m.elem + 1, "mua": 0.02, "musp": 1.1, "nr": 1.4})
    d = meshio.read_mat(p)
    assert np.array_equal(d["elem"], m.elem) and np.allclose(d["mua"], 0.02) and d["optode"] is None
    q = tmp_path / "named.mat"
    scipy.io.savemat(q, {"a": {"node": m.node, "elem": m.elem + 1}, "b": {"node": m.node[:4], "elem": [[1, 2, 3, 4]]}})
    assert meshio.read_mat(q, mesh_var="b")["node"].shape == (4, 3)
    bad = tmp_path / "bad.mat"
    scipy.io.savemat(bad, {"node": m.node, "elem": m.elem})                      # 0-based: rejected
    with pytest.raises(ValueError, match="1-based"):
        meshio.read_mat(bad)


def test_meshabox_bounds_volume_and_faces():
    node, face, elem = meshio.meshabox([0, 0, 0], [60, 30, 60], 1.2)
    v = _vol(node, elem)
    assert (v > 0).all() and v.max() <= 1.2 + 1e-12
    assert np.isclose(v.sum(), 60 * 30 * 60)
    assert np.allclose(node.min(0), 0) and np.allclose(node.max(0), [60, 30, 60])
    _, area = topology.true_faces(node, elem)
    assert np.isclose(area.sum(), 2 * (60 * 30 + 60 * 60 + 30 * 60))
    assert face.shape[1] == 3


def test_sortmesh_is_a_renumbering():
    m = meshgen.slab((8.0, 6.0, 4.0), 1.0, 0.01, 1.0, 1.4, seed=2)
    node, elem, order = meshio.sortmesh(m.node, m.elem)
    assert np.array_equal(np.sort(order), np.arange(m.node.shape[0]))
    assert np.allclose(node, m.node[order])
    a = {tuple(sorted(t)) for t in order[elem]}
    b = {tuple(sorted(t)) for t in m.elem}
    assert a == b
    assert np.isclose(np.abs(_vol(node, elem)).sum(), 8 * 6 * 4)
