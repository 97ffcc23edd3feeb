"""The MATLAB gateways of dogpup_amd/mex/ (smv_mex, FSAImv_mex, FSAIP_gen_mex,
replacing smv_mexgateway.cpp:8-100, FSAImv_mexgateway.cpp:8-100 and
FSAIP_mexgateway.cpp:8-85), compiled against a project-written mex.h
stand-in (tests/fakemex/) and driven from Python.

CPU: each gateway linked against a recording stand-in of the mex-level
entry points, checking what the gateway passes on: argument-count errors
with the reference's identifiers, the block's dimensions (a 2-D x is one
frequency), nnz from numel(cPtr) whatever its orientation, real arrays
widened to complex, argument errors, the library's error message, and
empty inputs.
GPU: the same gateway sources linked against libdogpup_hip.so, against the
oracle's smv / fsaimv / fsaip_gen."""
import numpy as np
import pytest
import scipy.sparse as sp

from conftest import rel


@pytest.fixture(scope="module")
def mex():
    from fakemex import build as fb
    if not fb.up_to_date():                     # the package build makes them; a stale tree is rebuilt here
        fb.build()
    from fakemex import fm
    return fm.Mex()


@pytest.fixture()
def rec(mex):
    from fakemex import fm
    r, last = fm.recorder()
    r.rec_reset()
    r.rec_fail(0, b"")
    yield r, last
    r.rec_fail(0, b"")


def _csr(n, seed=0, density=0.3):
    M = sp.random(n, n, density=density, random_state=seed) + sp.eye(n)
    M = sp.csr_matrix(M)
    M.sort_indices()
    rng = np.random.default_rng(seed)
    v = rng.normal(size=M.nnz) + 1j * rng.normal(size=M.nnz)
    return M.indptr.astype(np.int32), M.indices.astype(np.int32), v


def _cx(rng, *shape):
    return rng.normal(size=shape) + 1j * rng.normal(size=shape)


REF_COUNT_ERRORS = [
    ("smv_mex", 5, "CUDA:fwdSubCSRmex:nrhs", "5 inputs required.", "CUDA:fwdSubCSRmex:nlhs", "1 outputs required."),
    ("FSAImv_mex", 4, "CUDA:FSAImv:nrhs", "4 inputs required.", "CUDA:FSAImv:nlhs", "1 outputs required."),
    ("FSAIP_gen_mex", 6, "CUDA:FSAIP_mex:nrhs", "6 inputs required.", "CUDA:FSAIP_mex:nlhs", "1 output required."),
]


@pytest.mark.parametrize("name,nin,id_in,msg_in,id_out,msg_out", REF_COUNT_ERRORS)
def test_argument_count_errors(mex, rec, name, nin, id_in, msg_in, id_out, msg_out):
    from fakemex.fm import MexError
    args = [np.zeros(2)] * (nin - 1)
    with pytest.raises(MexError) as e:
        mex.call(name, *args)
    assert (e.value.ident, e.value.msg) == (id_in, msg_in)
    with pytest.raises(MexError) as e:
        mex.call(name, *([np.zeros(2)] * (nin + 1)))
    assert (e.value.ident, e.value.msg) == (id_in, msg_in)
    with pytest.raises(MexError) as e:
        mex.call(name, *([np.zeros(2)] * nin), nlhs=0)
    assert (e.value.ident, e.value.msg) == (id_out, msg_out)
    with pytest.raises(MexError) as e:
        mex.call(name, *([np.zeros(2)] * nin), nlhs=2)
    assert e.value.ident == id_out
    assert rec[1]()[3] == 0                               # the library was never called


def test_smv_passes_block_and_matrix(mex, rec):
    from fakemex.fm import checksum
    rng = np.random.default_rng(1)
    rp, cp, v = _csr(9, 1)
    for shape in [(9, 4, 3), (9, 4)]:                   # 3-D, and 2-D = one frequency
        x = _cx(rng, *shape)
        nf = shape[2] if len(shape) > 2 else 1
        f = np.arange(nf, dtype=np.float64) * 1.5
        y = mex.call("smv_mex", x, f, v, rp, cp)
        fn, ints, sums, calls = rec[1]()
        assert (fn, ints, calls) == (1, [9, 4, nf, cp.size], 1 if len(shape) > 2 else 2)
        assert np.allclose(sums[:5], [checksum(x), checksum(f), checksum(v), checksum(rp), checksum(cp)])
        # y has x's shape, complex, filled in column-major storage order
        assert y.shape == x.shape and y.dtype == np.complex128
        flat = y.ravel(order="F")
        assert np.array_equal(flat, np.arange(0, 2 * x.size, 2) + 1j * np.arange(1, 2 * x.size, 2))


def test_smv_real_arrays_widen_to_complex(mex, rec):
    from fakemex.fm import checksum
    rng = np.random.default_rng(2)
    rp, cp, v = _csr(7, 2)
    x = rng.normal(size=(7, 3, 2))
    mex.call("smv_mex", x, np.array([0.0, 1.0]), v.real.copy(), rp, cp)
    sums = rec[1]()[2]
    assert np.isclose(sums[0], checksum(x.astype(np.complex128)))
    assert np.isclose(sums[2], checksum(v.real.astype(np.complex128)))


def test_fsaimv_nnz_from_numel_of_cptr(mex, rec):
    from fakemex.fm import checksum
    rng = np.random.default_rng(3)
    rp, cp, _ = _csr(8, 3)
    nnz, nf = cp.size, 3
    G = _cx(rng, nnz, nf)
    x = _cx(rng, 8, 5, nf)
    for c in (cp, cp[None, :]):                         # column or row vector: nnz = numel(cPtrG)
        y = mex.call("FSAImv_mex", x, G, rp, c)
        fn, ints, sums, _ = rec[1]()
        assert (fn, ints) == (2, [8, 5, nf, nnz])
        assert np.allclose(sums[:4], [checksum(x), checksum(G), checksum(rp), checksum(cp)])
        assert y.shape == x.shape
    # one frequency: 2-D x and an nnz x 1 G
    mex.call("FSAImv_mex", x[:, :, 0].copy(), G[:, :1].copy(), rp, cp)
    assert rec[1]()[1] == [8, 5, 1, nnz]


def test_fsaip_gen_dims_and_scalar(mex, rec):
    from fakemex.fm import checksum
    rA, cA, vA = _csr(10, 4)
    A = sp.csr_matrix((vA, cA, rA), shape=(10, 10))
    T = sp.tril(A).tocsr()
    T.sort_indices()
    rT, cT = T.indptr.astype(np.int32), T.indices.astype(np.int32)
    rG, cG = rT.copy(), cT.copy()
    for nfs in (4, 4.0):                               # the reference passes int32; any numeric scalar works
        out = mex.call("FSAIP_gen_mex", T.data, rT, cT, rG, cG[None, :], nfs)
        fn, ints, sums, _ = rec[1]()
        assert (fn, ints) == (3, [10, T.nnz, cG.size, 4])
        assert np.allclose(sums[:5], [checksum(T.data), checksum(rT), checksum(cT), checksum(rG), checksum(cG)])
        assert out.shape == (4, cG.size) and out.dtype == np.complex128


def test_argument_errors(mex, rec):
    from fakemex.fm import MexError
    rng = np.random.default_rng(5)
    rp, cp, v = _csr(6, 5)
    x = _cx(rng, 6, 2, 3)
    f = np.arange(3.0)
    bad = [
        ("smv_mex", (x, f, v, rp.astype(np.float64), cp), "rPtrA must be a real int32 array."),
        ("smv_mex", (x, f, v[:-1].copy(), rp, cp), f"A has {cp.size - 1} values for {cp.size} column indices."),
        ("smv_mex", (x, f, v, rp[:-1].copy(), cp), "rPtrA must have size(x,1)+1 = 7 elements."),
        ("smv_mex", (x, f[:2].copy(), v, rp, cp), "f must have at least size(x,3) = 3 elements."),
        ("smv_mex", (x, f + 0j, v, rp, cp), "f must be a real double array."),
        ("smv_mex", (x.astype(np.complex128).reshape(6, 2, 3, 1, 1)[..., None].repeat(2, 5), f, v, rp, cp),
         "x must have at most 3 dimensions (nodes x columns x frequencies)."),
        ("FSAImv_mex", (x, _cx(rng, cp.size, 2), rp, cp), f"G must be numel(cPtrG) x size(x,3) = {cp.size} x 3."),
        ("FSAIP_gen_mex", (v, rp, cp, rp, cp, 0), "nfreqs must be a positive integer (got 0)."),
        ("FSAIP_gen_mex", (v, rp, cp, rp, cp, 2.5), "nfreqs must be a positive integer (got 2.5)."),
        ("FSAIP_gen_mex", (v, rp, cp, rp, cp, np.array([2, 3], np.int32)), "nfreqs must be a scalar."),
        ("FSAIP_gen_mex", (v, rp[:-1].copy(), cp, rp, cp, 2), "rPtrA must have numel(rPtrG) = 7 elements."),
    ]
    for name, args, msg in bad:
        with pytest.raises(MexError) as e:
            mex.call(name, *args)
        assert e.value.ident == f"DOGPUP:{name}:arg" and e.value.msg == msg, (name, e.value)
    assert rec[1]()[3] == 0


def test_library_error_becomes_matlab_error(mex, rec):
    from fakemex.fm import MexError
    r, last = rec
    rng = np.random.default_rng(6)
    rp, cp, v = _csr(5, 6)
    r.rec_fail(-1, b"smv_mex: rPtrA decreases at row 2")
    with pytest.raises(MexError) as e:
        mex.call("smv_mex", _cx(rng, 5, 1, 1), np.zeros(1), v, rp, cp)
    assert (e.value.ident, e.value.msg) == ("DOGPUP:smv_mex:lib", "smv_mex: rPtrA decreases at row 2")
    assert last()[3] == 1


def test_empty_inputs_return_zeros_without_a_call(mex, rec):
    rp, cp, v = _csr(4, 7)
    y = mex.call("smv_mex", np.zeros((4, 0, 2), np.complex128), np.arange(2.0), v, rp, cp)
    assert y.shape == (4, 0, 2)
    y = mex.call("FSAImv_mex", np.ones((4, 2, 1), np.complex128), np.zeros((0, 1), np.complex128),
                 np.zeros(5, np.int32), np.zeros(0, np.int32))
    assert y.shape == (4, 2) and not y.any()
    out = mex.call("FSAIP_gen_mex", v, rp, cp, np.zeros(5, np.int32), np.zeros(0, np.int32), 3)
    assert out.shape == (3, 0)
    assert rec[1]()[3] == 0


# ---------------------------------------------------------------- GPU
@pytest.fixture(scope="module")
def system():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from conftest import small_slab
    from oracle import fem, transforms as otr
    m = small_slab(seed=4)
    om = fem.make_mesh(m.node, m.elem, m.nr, m.mua, m.musp)
    _, df = otr.freq_axis(np.arange(0, 501) * 20e-12, 3)
    A = fem.full_system(om, df).tocsr()
    A.sort_indices()
    return A


@pytest.mark.gpu
def test_gateways_on_gpu_match_oracle(mex, system):
    from oracle import fsai as ofsai, solver as osol
    A = system
    n, ns, nf = A.shape[0], 5, 3
    rng = np.random.default_rng(8)
    X = _cx(rng, n, ns, nf)
    f = np.arange(nf, dtype=np.float64)                  # dMesh.m:1259: f = 0:freqs-1
    rp, cp = A.indptr.astype(np.int32), A.indices.astype(np.int32)
    print("[case] smv_mex 3-D and 2-D", flush=True)
    y = mex.call("smv_mex", X, f[None, :], A.data, rp, cp, variant="lib")
    assert rel(y, osol.smv(A, f, X)) < 1e-14
    y1 = mex.call("smv_mex", X[:, :, 1].copy(), f[1:2].copy(), A.data, rp, cp, variant="lib")
    assert rel(y1, osol.smv(A, f[1:2], X[:, :, 1:2])[:, :, 0]) < 1e-14
    print("[case] FSAIP_gen_mex", flush=True)
    T = sp.tril(A).tocsr()
    T.sort_indices()
    rG, cG, valG = ofsai.fsaip_gen(A, nf)                 # per-row partial-pivot GE, nnzG x NF
    out = mex.call("FSAIP_gen_mex", T.data, T.indptr.astype(np.int32), T.indices.astype(np.int32),
                   rG.astype(np.int32), cG.astype(np.int32), np.int32(nf).item(), variant="lib")
    assert out.shape == (nf, cG.size)
    assert rel(out.T, valG) < 1e-12
    print("[case] FSAImv_mex", flush=True)
    G = np.asfortranarray(out.T)                         # dMesh.m:1384 transposes to nnzG x NF
    y = mex.call("FSAImv_mex", X, G, rG.astype(np.int32), cG.astype(np.int32)[None, :], variant="lib")
    assert rel(y, osol.fsaimv(rG, cG, valG, X)) < 1e-12
    print("[case] malformed CSR rejected on the host", flush=True)
    from fakemex.fm import MexError
    cbad = cp.copy()
    cbad[3] = n + 5
    with pytest.raises(MexError) as e:
        mex.call("smv_mex", X, f, A.data, rp, cbad, variant="lib")
    assert e.value.ident == "DOGPUP:smv_mex:lib" and "outside [0," in e.value.msg
