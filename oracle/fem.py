"""Oracle: mesh preparation and P1 FEM assembly (test infrastructure only).

Restates ``DOGPUP/dMesh.m:945-1030`` (prepare_mesh) and ``dMesh.m:1033-1235``
(gen_fwdmat, TetQuadDat, TriQuadDat) in FP64 NumPy/SciPy, keeping the
reference's quadrature loops and its boundary-face selection.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp

C_VAC = 2.99792458e11  # mm/s, dMesh.m:42


@dataclass
class OMesh:
    node: np.ndarray
    elem: np.ndarray     # 0-based int64
    nr: float
    mua: np.ndarray
    musp: np.ndarray
    kappa: np.ndarray
    c: float = 0.0
    vol: np.ndarray = None
    face: np.ndarray = None
    area: np.ndarray = None
    bnd: np.ndarray = None
    R: np.ndarray = None
    gradScale: np.ndarray = None
    K: sp.csr_matrix = None   # upper triangle (triu), dMesh.m:1189
    M: sp.csr_matrix = None


def make_mesh(node, elem, nr, mua, musp, true_faces=False):
    """dMesh constructor (dMesh.m:48-77): kappa = 1/(3(mua+musp)), then
    prepare_mesh and gen_fwdmat."""
    node = np.asarray(node, float)
    nn = node.shape[0]
    mua = np.broadcast_to(np.asarray(mua, float), (nn,)).copy()
    musp = np.broadcast_to(np.asarray(musp, float), (nn,)).copy()
    m = OMesh(node=node, elem=np.asarray(elem, np.int64), nr=float(nr), mua=mua, musp=musp,
              kappa=1.0 / (3.0 * (mua + musp)))
    prepare_mesh(m)
    m.K, m.M = gen_fwdmat(m, true_faces=true_faces)
    return m


def update_properties(m: OMesh, mua, musp, true_faces=False):
    """dMesh.m:80-88."""
    m.mua = np.asarray(mua, float).copy()
    m.musp = np.asarray(musp, float).copy()
    m.kappa = 1.0 / (3.0 * (m.mua + m.musp))
    m.K, m.M = gen_fwdmat(m, true_faces=true_faces)
    return m


def robin_factor(nr):
    """R = 1/(2A), dMesh.m:1008-1011."""
    R0 = (nr - 1.0) ** 2 / (nr + 1.0) ** 2
    thetaC = np.arcsin(1.0 / nr)
    A = (2.0 / (1.0 - R0) - 1.0 + abs(np.cos(thetaC)) ** 3) / (1.0 - abs(np.cos(thetaC)) ** 2)
    return 1.0 / (2.0 * A)


def prepare_mesh(m: OMesh):
    """dMesh.m:945-1030."""
    node, elem = m.node, m.elem
    m.c = C_VAC / m.nr                                                    # :949
    a = node[elem[:, 1]] - node[elem[:, 0]]
    b = node[elem[:, 2]] - node[elem[:, 0]]
    c = node[elem[:, 3]] - node[elem[:, 0]]
    m.vol = np.abs(np.einsum("ij,ij->i", np.cross(a, b), c) / 6.0)       # :952-955
    faces = np.concatenate([elem[:, [0, 1, 2]], elem[:, [0, 1, 3]],
                            elem[:, [0, 2, 3]], elem[:, [1, 2, 3]]])       # :964-967
    key = np.sort(faces, axis=1)
    _, ix, jx, cnt = np.unique(key, axis=0, return_index=True, return_inverse=True,
                               return_counts=True)                         # :969
    once = cnt == 1                                                        # :971-972
    face = faces[ix[once]].copy()                                          # :973
    bnd = np.zeros(node.shape[0], np.int64)
    bnd[np.unique(face.ravel())] = 1                                       # :978-980
    com = node[face].mean(axis=1)                                          # :983-986
    centre = 0.5 * (node.max(axis=0) + node.min(axis=0))                   # :990-992
    vdir = centre[None, :] - com
    vdir /= np.linalg.norm(vdir, axis=1, keepdims=True)
    fa = node[face[:, 1]] - node[face[:, 0]]
    fb = node[face[:, 2]] - node[face[:, 0]]
    signA = np.sign(np.einsum("ij,ij->i", np.cross(fa, fb), vdir))        # :997
    flip = signA > 0
    face[flip, 1], face[flip, 2] = face[flip, 2].copy(), face[flip, 1].copy()  # :998
    fa = node[face[:, 1]] - node[face[:, 0]]
    fb = node[face[:, 2]] - node[face[:, 0]]
    m.area = 0.5 * np.linalg.norm(np.cross(fa, fb), axis=1)              # :1001-1002
    m.face, m.bnd = face, bnd
    m.R = robin_factor(m.nr) * bnd                                         # :1008-1011
    B = np.stack([a, b, c], axis=2)                                        # columns = edges
    BT = np.transpose(B, (0, 2, 1))
    invBT = np.linalg.inv(BT)                                              # :1021
    m.gradScale = invBT                                                    # NE x 3 x 3 pages
    return m


def tet_quad():
    """Keast 5-point rule, dMesh.m:1197-1215."""
    xa = np.array([1 / 4, 1 / 2, 1 / 6, 1 / 6, 1 / 6])
    ya = np.array([1 / 4, 1 / 6, 1 / 6, 1 / 6, 1 / 2])
    za = np.array([1 / 4, 1 / 6, 1 / 6, 1 / 2, 1 / 6])
    wt = np.array([-0.8, 0.45, 0.45, 0.45, 0.45])
    return xa, ya, za, wt


def tri_quad():
    """3-point degree-2 triangle rule, dMesh.m:1218-1235."""
    return np.array([2 / 3, 1 / 6, 1 / 6]), np.array([1 / 6, 2 / 3, 1 / 6]), np.array([1 / 3, 1 / 3, 1 / 3])


def boundary_faces_ref(m: OMesh):
    """The Robin-term face list of dMesh.m:1129-1140: tets whose 4 nodes are all
    boundary contribute all 4 faces, tets with exactly 3 boundary nodes
    contribute the face of those 3 (sorted node order)."""
    temp = np.sort(m.elem, axis=1)
    flag = m.bnd[temp] > 0
    mask = flag.sum(axis=1)
    t4 = temp[mask == 4]
    bface = np.concatenate([t4[:, [0, 1, 2]], t4[:, [0, 1, 3]], t4[:, [0, 2, 3]], t4[:, [1, 2, 3]]])
    t3 = temp[mask == 3]
    f3 = t3[flag[mask == 3]].reshape(-1, 3)   # nonzeros in row order (row-major)
    return np.concatenate([bface, f3])


def gen_fwdmat(m: OMesh, true_faces=False):
    """dMesh.m:1033-1194.  Returns triu(K), triu(M) as CSR.

    ``true_faces`` swaps the reference's face selection for the true boundary
    faces (mesh.face); the default follows the reference.
    """
    nn = m.node.shape[0]
    xa, ya, za, w3 = tet_quad()
    N3 = np.stack([1 - xa - ya - za, xa, ya, za])           # 4 x 5, :1056
    N3g = np.array([[-1, -1, -1], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    vol = m.vol
    kap = (N3.T @ m.kappa[m.elem].T) * w3[:, None]           # 5 x NE, :1078-1082
    mua = (N3.T @ m.mua[m.elem].T) * w3[:, None]
    ksum = kap.sum(axis=0)
    B = m.gradScale                                         # NE x 3 x 3 = inv(B')
    grads = np.einsum("eij,aj->eai", B, N3g)                # NE x 4 x 3
    EK = np.zeros((m.elem.shape[0], 4, 4))
    EM = np.zeros_like(EK)
    for i in range(4):                                      # :1099-1121
        for j in range(4):
            pre = vol * np.einsum("ek,ek->e", grads[:, i], grads[:, j])
            stiff = pre * ksum
            absrp = vol * (mua * (N3[i] * N3[j])[:, None]).sum(axis=0)
            EK[:, i, j] = stiff + absrp
            EM[:, i, j] = vol * (w3 * N3[i] * N3[j]).sum()
    I = np.repeat(m.elem[:, :, None], 4, axis=2)
    J = np.transpose(I, (0, 2, 1))
    K = sp.coo_matrix((EK.ravel(), (I.ravel(), J.ravel())), shape=(nn, nn)).tocsr()
    M = sp.coo_matrix((EM.ravel(), (I.ravel(), J.ravel())), shape=(nn, nn)).tocsr()
    # surface (Robin) term, :1129-1181
    bface = m.face if true_faces else boundary_faces_ref(m)
    fa = m.node[bface[:, 1]] - m.node[bface[:, 0]]
    fb = m.node[bface[:, 2]] - m.node[bface[:, 0]]
    area = 0.5 * np.linalg.norm(np.cross(fa, fb), axis=1)
    x2, y2, w2 = tri_quad()
    N2 = np.stack([1 - x2 - y2, x2, y2])                    # 3 x 3
    Rq = (N2.T @ m.R[bface].T) * w2[:, None]                # 3 x NF
    EB = np.zeros((bface.shape[0], 3, 3))
    for i in range(3):
        for j in range(3):
            EB[:, i, j] = area * (Rq * (N2[i] * N2[j])[:, None]).sum(axis=0)
    Ib = np.repeat(bface[:, :, None], 3, axis=2)
    Jb = np.transpose(Ib, (0, 2, 1))
    K = K + sp.coo_matrix((EB.ravel(), (Ib.ravel(), Jb.ravel())), shape=(nn, nn)).tocsr()
    return sp.triu(K).tocsr(), sp.triu(M).tocsr()


def full_system(m: OMesh, df: float):
    """A = K + i*df/c*M made full symmetric (fluGPU, dMesh.m:1241-1242).

    The imaginary part is for unit frequency; frequency k scales it by k."""
    A = (m.K + 1j * (df / m.c) * m.M).tocsr()
    A = (A + sp.triu(A, 1).T).tocsr()
    A.sort_indices()
    return A


def system_at(A: sp.csr_matrix, k: int) -> sp.csr_matrix:
    """A_k = Re(A) + i*k*Im(A) (smv_CUDA.cu:43-48, FSAIP_CUDA.cu:132-133)."""
    return (A.real + 1j * k * A.imag).tocsr()
