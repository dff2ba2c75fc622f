"""pt_alias_build_device_noalloc: the alias-table build with its working
buffers in the library's own per-GPU device array instead of allocations.
Its tables must equal pt_alias_build_device's bit for bit, also on a second
call with fewer weights (the array keeps the previous call's contents)."""
import ctypes as C

import numpy as np
import pytest

import pathtracingtutorial_amd as pt
from pathtracingtutorial_amd import _capi

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

MAX_WEIGHTS = 1 << 20  # PT_ALIAS_NOALLOC_MAX_WEIGHTS in include/pt_amd.h


def _build(fn, w: torch.Tensor):
    n = w.numel()
    pmf = torch.full((n,), -1.0, dtype=torch.float32, device="cuda")
    alias = torch.full((n, 2), -1.0, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    _capi.check(getattr(pt.lib(), fn)(C.c_void_p(w.data_ptr()), n, C.c_void_p(pmf.data_ptr()),
                                      C.c_void_p(alias.data_ptr())))
    return pmf.view(torch.int32).cpu().numpy(), alias.view(torch.int32).cpu().numpy()


def _same_as_allocating(w: np.ndarray):
    wd = torch.from_numpy(w).cuda()
    got = _build("pt_alias_build_device_noalloc", wd)
    ref = _build("pt_alias_build_device", wd)
    assert np.array_equal(got[0], ref[0]), "pmf differs"
    assert np.array_equal(got[1], ref[1]), "alias table differs"
    return got


def test_noalloc_alias_equals_allocating_build_on_two_calls():
    rng = np.random.default_rng(5)
    big = (rng.pareto(1.2, 300000) + 1e-6).astype(np.float32)
    small = rng.random(4097).astype(np.float32)
    small[rng.random(4097) < 0.6] = 0.0
    print("noalloc alias: 300000 heavy-tail weights, then 4097 sparse weights", flush=True)
    _same_as_allocating(big)
    pmf, table = _same_as_allocating(small)
    # the second table is a table of the second weights, not a leftover
    exact = small.astype(np.float64) / small.astype(np.float64).sum()
    assert np.allclose(pmf.view(np.float32), exact, rtol=1e-6, atol=1e-12)
    assert np.all(table[:, 1] < small.size)


def test_noalloc_alias_at_the_maximum_size():
    rng = np.random.default_rng(6)
    w = rng.random(MAX_WEIGHTS).astype(np.float32)
    print(f"noalloc alias: {MAX_WEIGHTS} weights (the maximum)", flush=True)
    _same_as_allocating(w)


def test_noalloc_alias_rejects_too_many_and_all_zero():
    wd = torch.zeros(16, dtype=torch.float32, device="cuda")
    pmf = torch.empty(16, device="cuda")
    alias = torch.empty((16, 2), device="cuda")
    args = (C.c_void_p(wd.data_ptr()), None, C.c_void_p(pmf.data_ptr()), C.c_void_p(alias.data_ptr()))
    # the size is checked on the host before anything runs on the GPU
    with pytest.raises(pt.PtError, match="exceed the maximum"):
        _capi.check(pt.lib().pt_alias_build_device_noalloc(args[0], MAX_WEIGHTS + 1, args[2], args[3]))
    with pytest.raises(pt.PtError):
        _capi.check(pt.lib().pt_alias_build_device_noalloc(args[0], 16, args[2], args[3]))
    # still usable after the errors
    _same_as_allocating(np.array([1.0, 3.0], np.float32))
