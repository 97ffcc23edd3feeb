"""The allocation-free fc2td entry point (dogpup_fc2td_field_noalloc) gives
the same FP32 time-domain field as dogpup_fc2td_field, bit for bit, on two
calls in a row with the second one smaller (fewer columns, frequencies and
bins), so the static scratch's leftover table from the first call is not
used by the second."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _call(name, F, out, f, t, nf, nt, ncol):
    from dogpup_amd import _lib
    _lib.call(name, _lib.ptr(F), _lib.ptr(out), _lib.ptr(f), _lib.ptr(t), nf, nt, ncol, _lib.stream_ptr())


@pytest.mark.parametrize("first,second", [((64, 300, 40_000), (9, 70, 1_000)),
                                          ((100, 200, 20_000), (64, 256, 12_345))])
def test_fc2td_noalloc_matches_fc2td_field(first, second):
    assert torch.cuda.is_available()
    rng = np.random.default_rng(5)
    for nf, nt, ncol in (first, second):
        print(f"[case] fc2td noalloc nf={nf} nt={nt} ncol={ncol}", flush=True)
        F = torch.as_tensor(rng.standard_normal((nf, ncol, 2)), device="cuda")
        f = torch.as_tensor(2 * np.pi * 50e6 * np.arange(nf), device="cuda")
        t = torch.as_tensor(np.arange(nt) * 20e-12, device="cuda")
        ref = torch.full((nt, ncol), float("nan"), dtype=torch.float32, device="cuda")
        got = torch.full((nt, ncol), float("nan"), dtype=torch.float32, device="cuda")
        _call("dogpup_fc2td_field", F, ref, f, t, nf, nt, ncol)
        _call("dogpup_fc2td_field_noalloc", F, got, f, t, nf, nt, ncol)
        torch.cuda.synchronize()
        assert torch.isfinite(ref).all()
        assert torch.equal(got, ref)
        # and both are the FP64 DFT of the block (fc2td.m:24-31) to FP32 rounding
        w = np.where(np.arange(nf) == 0, 0.5, 1.0)
        Fc = F[..., 0].cpu().numpy() + 1j * F[..., 1].cpu().numpy()
        E = np.exp(1j * np.outer(t.cpu().numpy(), f.cpu().numpy())) * w[None, :]
        y = 2 * (E @ Fc).real
        err = np.abs(got.cpu().numpy() - y).max() / np.abs(y).max()
        assert err < 1e-6, err


def test_fc2td_noalloc_rejects_bad_sizes():
    from dogpup_amd import _lib
    F = torch.zeros((2, 16, 2), dtype=torch.float64, device="cuda")
    out = torch.zeros((4, 16), dtype=torch.float32, device="cuda")
    f = torch.zeros(129, dtype=torch.float64, device="cuda")
    t = torch.zeros(4, dtype=torch.float64, device="cuda")
    with pytest.raises(_lib.DogpupError) as e:
        _call("dogpup_fc2td_field_noalloc", F, out, f, t, 129, 4, 16)
    assert e.value.code == _lib.ERR_ARG
