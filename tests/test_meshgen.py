"""Synthetic mesh generators: Kuhn cylinders and the uneven 48-mesh batch."""
import numpy as np

from dogpup_amd import meshgen, topology


def _volumes(m):
    x = m.node[m.elem]
    return np.einsum("ij,ij->i", np.cross(x[:, 1] - x[:, 0], x[:, 2] - x[:, 0]), x[:, 3] - x[:, 0]) / 6.0


def test_cylinder_is_a_valid_staircase_cylinder():
    r, H, h = 10.0, 8.0, 1.0
    m = meshgen.cylinder(r, H, h)
    v = _volumes(m)
    assert (v > 0).all()                                     # oriented, none degenerate
    assert np.isclose(v, h ** 3 / 6.0).all()                 # whole Kuhn tetrahedra
    assert np.unique(m.elem).size == m.node.shape[0]         # every node used
    c = r
    rad = np.hypot(m.node[:, 0] - c, m.node[:, 1] - c)
    assert rad.max() <= r + 1e-9
    # the staircase volume is within the lattice's resolution of pi r^2 H
    assert abs(v.sum() - np.pi * r * r * H) / (np.pi * r * r * H) < 0.12
    # flat top face: the boundary holds the top disk of lattice points
    face, area = topology.true_faces(m.node, m.elem)
    top = np.isclose(m.node[face, 2], H).all(axis=1)
    assert abs(area[top].sum() - np.pi * r * r) / (np.pi * r * r) < 0.12


def test_uneven_batch_sizes_and_determinism():
    a = meshgen.uneven_batch(6, seed=3)
    b = meshgen.uneven_batch(6, seed=3)
    for x, y in zip(a, b):
        assert np.array_equal(x["mesh"].node, y["mesh"].node) and x["nf"] == y["nf"]
        assert np.array_equal(x["src"], y["src"])
    for x in a:
        m = x["mesh"]
        assert 8 <= len(x["src"]) <= 32 and 8 <= x["nf"] <= 64
        assert 0.005 <= m.mua[0] <= 0.02 and 0.5 <= m.musp[0] <= 1.5
        assert m.h in (1.0, 2.0, 3.0)
        assert 3000 <= m.node.shape[0] <= 260000
        # sources lie on the top face, inside the mesh's footprint
        assert np.allclose(x["src"][:, 2], m.extent[2])
        assert (x["src"][:, :2] >= 0).all() and (x["src"][:, :2] <= np.array(m.extent[:2])).all()


def test_batch_default_width_minimises_padding():
    from types import SimpleNamespace as NS
    from dogpup_amd.batch import MeshBatch
    from dogpup_amd.mesh import group_width
    big = [NS(nn=10000)]
    assert MeshBatch.default_width(big, [32], [8], "mixed") == 32          # no padding at 32: widest wins
    assert MeshBatch.default_width(big, [24], [8], "mixed") == 8           # 24 = 3 x 8
    assert MeshBatch.default_width(big, [12], [8], "mixed") == 16          # 16 caps it (group_width)
    two = [NS(nn=1000), NS(nn=50000)]
    assert MeshBatch.default_width(two, [9, 31], [4, 4], "mixed") == 16    # 16+32 columns pad less than 32+32
    assert MeshBatch.default_width(two, [16, 32], [4, 4], "mixed") == 16   # 8 and 16 tie: the wider group
    assert group_width(20, "fp64") == 16 and group_width(20, "mixed") == 32 and group_width(5) == 8
