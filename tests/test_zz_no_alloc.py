"""The allocation-free multi-dot tacs_hip_mdot_noalloc against tacs_hip_mdot
(same kernels, scratch from the caller) through the C ABI: identical dots on
two calls in a row, the second smaller, in FP64 and FP32, plus the size
limit.  Runs alone or after the other tests (the static scratch keeps the
previous call's partials, which no call may depend on)."""
import numpy as np
import pytest

from tacs_amd._lib import lib

pytestmark = pytest.mark.gpu


def _mdot_pair(torch, L, prec, n, k, accumulate, seed):
    dt = torch.float64 if prec == 64 else torch.float32
    g = torch.Generator(device="cpu").manual_seed(seed)
    ldw = n + 17  # leading dimension beyond n: the kernels must honour it
    W = torch.randn(k * ldw, generator=g, dtype=torch.float64).to(dt).cuda()
    q = torch.randn(n, generator=g, dtype=torch.float64).to(dt).cuda()
    h0 = torch.randn(k, generator=g, dtype=torch.float64).cuda()
    h_ref, h_na = h0.clone(), h0.clone()
    scratch = torch.zeros(L.tacs_hip_red_scratch_bytes() // 8, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    assert L.tacs_hip_mdot(prec, n, k, W.data_ptr(), ldw, q.data_ptr(), scratch.data_ptr(),
                           h_ref.data_ptr(), accumulate, None) == 0
    assert L.tacs_hip_mdot_noalloc(prec, n, k, W.data_ptr(), ldw, q.data_ptr(), h_na.data_ptr(),
                                   accumulate, None) == 0
    torch.cuda.synchronize()
    # float64 dots of the stored values, as a plain reference of the op
    Wd = W.double().reshape(k, ldw)[:, :n]
    expect = Wd @ q.double() + (h0 if accumulate else 0.0)
    return h_ref.cpu().numpy(), h_na.cpu().numpy(), expect.cpu().numpy()


@pytest.mark.parametrize("prec", [64, 32])
def test_mdot_noalloc_matches_mdot(prec):
    import torch

    L = lib()
    # two calls in a row, the second with a smaller input (fewer blocks and
    # vectors: stale partials of the first call sit beyond what it reads)
    for n, k, acc, seed in [(3_000_001, 32, 0, 1), (40_000, 7, 1, 2), (1, 1, 0, 3)]:
        h_ref, h_na, expect = _mdot_pair(torch, L, prec, n, k, acc, seed)
        assert np.array_equal(h_ref, h_na), (n, k)
        tol = 1e-12 if prec == 64 else 1e-12  # FP64 accumulation of stored values
        assert np.abs(h_na - expect).max() <= tol * max(1.0, np.abs(expect).max()) * np.sqrt(n)


def test_mdot_noalloc_rejects_sizes_beyond_its_scratch():
    import torch

    L = lib()
    n = 1000
    W = torch.zeros(33 * n, dtype=torch.float64, device="cuda")
    q = torch.zeros(n, dtype=torch.float64, device="cuda")
    h = torch.full((33,), 7.0, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    assert L.tacs_hip_mdot_noalloc(64, n, 33, W.data_ptr(), n, q.data_ptr(), h.data_ptr(), 0,
                                   None) != 0
    assert b"32" in L.tacs_last_error()
    assert L.tacs_hip_mdot_noalloc(64, n, 0, W.data_ptr(), n, q.data_ptr(), h.data_ptr(), 0,
                                   None) != 0
    assert L.tacs_hip_mdot_noalloc(16, n, 4, W.data_ptr(), n, q.data_ptr(), h.data_ptr(), 0,
                                   None) != 0
    torch.cuda.synchronize()
    # nothing was launched: h is untouched
    assert np.all(h.cpu().numpy() == 7.0)
