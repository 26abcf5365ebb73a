"""The allocation-free conservation sums (mrb_reduce_conservation_noalloc)
against mrb_reduce_conservation, through the C ABI: bit-identical sums on
two calls in a row, the second on a smaller level, each written to a
caller-provided device array without a synchronisation."""
import ctypes

import numpy as np
import pytest

from marbles_amd import configs

pytestmark = pytest.mark.gpu


def _noalloc_sums(lib, lvl, torch):
    out = torch.full((5,), float("nan"), dtype=torch.float64, device="cuda")
    rc = lib.mrb_reduce_conservation_noalloc(lvl.h, ctypes.c_void_p(out.data_ptr()), None)
    assert rc == 0, lib.mrb_last_error()
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _host_sums(lib, lvl):
    out = (ctypes.c_double * 5)()
    assert lib.mrb_reduce_conservation(lvl.h, out, None) == 0, lib.mrb_last_error()
    return np.array(out[:])


def test_noalloc_sums_equal_the_allocating_entry_point(engine):
    import torch

    from marbles_amd import _lib

    lib = _lib.load()
    big = configs.build_level(configs.channel_sphere((48, 24, 24), radius=5.0, center=(16.0, 12.0, 12.0)))
    big.advance(5)
    a = _noalloc_sums(lib, big, torch)
    b = _host_sums(lib, big)
    assert np.array_equal(a, b), (a, b)
    # a smaller level next: fewer partial sums than the first call left in
    # the scratch array; its sums must not see the first call's
    small = configs.build_level(configs.taylor_green(16, tau=1.0))
    small.advance(3)
    c = _noalloc_sums(lib, small, torch)
    d = _host_sums(lib, small)
    assert np.array_equal(c, d), (c, d)
    assert not np.array_equal(c, a)
    # the shifted FP32 storage goes through the same kernels
    f32 = configs.build_level(configs.taylor_green(16, tau=1.0), storage="f32s")
    f32.advance(2)
    assert np.array_equal(_noalloc_sums(lib, f32, torch), _host_sums(lib, f32))
    for lvl in (big, small, f32):
        lvl.close()
