"""nnk_direct_apply_device_noalloc (partial sums in the library's static
device array) against nnk_direct_apply_device (partial sums in a
stream-ordered allocation): the same kernels and summation order, so the
results must be bit-identical, on two calls in a row, the second smaller
(stale partials of the first call must not leak into the second).  The
first call is also checked against the oracle's FP64 direct sum."""
import ctypes
from importlib import import_module

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    assert _t.cuda.is_available()
    return _t


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _both(nnk, torch, rng, ns, nt, eps, mu):
    L = import_module("stokes-nnkifmm_amd._lib")
    s = rng.random((ns, 3))
    q = rng.standard_normal((ns, 3))
    t = rng.random((nt, 3))
    ds, dq, dt = (torch.tensor(a, device="cuda") for a in (s, q, t))
    ua = torch.full((nt, 3), np.nan, dtype=torch.float64, device="cuda")
    ub = torch.full((nt, 3), np.nan, dtype=torch.float64, device="cuda")
    print(f"[no_alloc] ns={ns} nt={nt} eps={eps} mu={mu}", flush=True)
    L.check(nnk.lib().nnk_direct_apply_device(_vp(ds), ns, _vp(dq), _vp(dt), nt, eps, mu, _vp(ua), None))
    L.check(nnk.lib().nnk_direct_apply_device_noalloc(_vp(ds), ns, _vp(dq), _vp(dt), nt, eps, mu, _vp(ub), None))
    torch.cuda.synchronize()
    return s, q, t, ua.cpu().numpy(), ub.cpu().numpy()


def test_noalloc_direct_matches_allocating_entry_point(nnk, torch, oracle):
    rng = np.random.default_rng(11)
    # 4,096 targets over 300,000 sources: 128 source parts (12.6 MB of partials)
    s, q, t, ua, ub = _both(nnk, torch, rng, 300_000, 4096, 1e-3, 1.0)
    assert np.isfinite(ua).all()
    np.testing.assert_array_equal(ub, ua)
    idx = np.sort(rng.choice(4096, 48, replace=False))
    ref = oracle.direct(s, q, t[idx], 1e-3, 1.0)
    assert np.linalg.norm(ua[idx] - ref) / np.linalg.norm(ref) < 1e-12
    # smaller second call: 700 targets over 20,000 sources, 19 parts
    s, q, t, ua, ub = _both(nnk, torch, rng, 20_000, 700, 5e-3, 1.3)
    assert np.isfinite(ua).all()
    np.testing.assert_array_equal(ub, ua)
    ref = oracle.direct(s, q, t, 5e-3, 1.3)
    assert np.linalg.norm(ua - ref) / np.linalg.norm(ref) < 1e-13


def test_noalloc_direct_rejects_bad_arguments(nnk, torch):
    L = import_module("stokes-nnkifmm_amd._lib")
    d = torch.zeros((10, 3), dtype=torch.float64, device="cuda")
    rc = nnk.lib().nnk_direct_apply_device_noalloc(_vp(d), 10, _vp(d), _vp(d), 10, 0.0, 1.0, _vp(d), None)
    assert rc == 1 and b"eps" in nnk.lib().nnk_last_error()
    rc = nnk.lib().nnk_direct_apply_device_noalloc(None, 10, _vp(d), _vp(d), 10, 1e-3, 1.0, _vp(d), None)
    assert rc == 1
    L.check(nnk.lib().nnk_direct_apply_device_noalloc(_vp(d), 10, _vp(d), _vp(d), 0, 1e-3, 1.0, _vp(d), None))
