"""Device numerics against the library routines bit for bit on the GPU (tests/native/device_numerics.hip,
built by __graft_entry__.build()):

* sqrt_rn (phx_numerics.hpp), the flux square root: the compiler's v_rsq_f64 Newton sequence without
  its range scaling and special-value select, which are identities for 2^-767 <= x <= DBL_MAX, and the
  library sqrt for every other x.  Compared with the library sqrt (and that with numpy's IEEE sqrt)
  on a million random bit patterns of every exponent and on the edge classes: the threshold 2^-767
  and its neighbours, subnormals, +-0, +-inf, NaN, negative values and DBL_MAX.
* limiter_min4<false>, the Venkatakrishnan limiter's four-face minimum: fmin where no value is NaN,
  against the np.minimum chain (limiters/base.py:47-187) on limiter values with +0, 3/4 and NaN.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
LIB = Path(__file__).resolve().parent / "native" / "libdevice_numerics.so"


def _lib():
    import torch

    if not LIB.exists():
        pytest.fail(f"missing {LIB.name}: run __graft_entry__.build()")
    torch.zeros(1, device="cuda")
    lib = C.CDLL(str(LIB))
    for f in (lib.dn_sqrt, lib.dn_min4):
        f.restype = C.c_int
        f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
    return lib


def _run(fn, inp, n):
    import torch

    d_in = torch.from_numpy(np.ascontiguousarray(inp)).cuda()
    a = torch.empty(n, dtype=torch.float64, device="cuda")
    b = torch.empty(n, dtype=torch.float64, device="cuda")
    rc = fn(d_in.data_ptr(), a.data_ptr(), b.data_ptr(), n, torch.cuda.current_stream().cuda_stream)
    assert rc == 0, f"launch failed: {rc}"
    torch.cuda.synchronize()
    return a.cpu().numpy(), b.cpu().numpy()


def _same(x, y):
    return (x.view(np.uint64) == y.view(np.uint64)) | (np.isnan(x) & np.isnan(y))


def test_fast_square_root_equals_the_library_square_root():
    lib = _lib()
    rng = np.random.default_rng(5)
    rand = rng.integers(0, 2**64, 1_000_000, dtype=np.uint64).view(np.float64)
    t = 2.0**-767
    near = t * (1.0 + np.arange(-64, 65) * np.finfo(float).eps)
    edges = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, -1.0, -t, t, np.nextafter(t, 0.0), np.nextafter(t, 1.0),
                      np.finfo(float).max, np.nextafter(np.finfo(float).max, 0.0), np.finfo(float).tiny, 5e-324,
                      1e-310, 0.25, 0.5, 1.0, 2.0, 4.0, 1.4, 1e300, 1e-300])
    phys = rng.uniform(1e-3, 1e3, 100_000) * rng.uniform(0.1, 10.0, 100_000)
    x = np.concatenate([rand, near, edges, phys])
    fast, ref = _run(lib.dn_sqrt, x, len(x))
    bad = ~_same(fast, ref)
    assert not bad.any(), f"{bad.sum()} differ, e.g. x = {x[bad][:5]!r}"
    with np.errstate(invalid="ignore"):
        assert _same(ref, np.sqrt(x)).all()
    inrange = (x >= t) & (x <= np.finfo(float).max)
    assert inrange.sum() > 400_000 and (~inrange).sum() > 400_000   # both paths exercised


def test_venkatakrishnan_four_face_minimum_on_the_device():
    lib = _lib()
    rng = np.random.default_rng(6)
    n = 500_000
    s = rng.standard_normal(n) * np.exp(rng.uniform(-30, 30, n))
    s[: n // 20] = 0.0
    s[n // 20: n // 10] = -2.0
    s[n // 10: n // 10 + 1000] = 1e200                     # s*s overflows: NaN
    with np.errstate(over="ignore", invalid="ignore"):
        phi = (s * s + 2.0 * s) / (s * s + s + 2.0)
    phi[rng.random(n) < 0.1] = 0.75                        # flat faces
    assert np.isnan(phi).any() and (phi == 0.0).any() and not np.signbit(phi[phi == 0.0]).any()
    faces = phi[rng.integers(0, n, (n, 4))]
    fast, chain = _run(lib.dn_min4, faces.ravel(), n)
    assert _same(fast, chain).all()
    ref = np.minimum(np.minimum(np.minimum(faces[:, 0], faces[:, 1]), faces[:, 2]), faces[:, 3])
    assert _same(fast, ref).all() and np.isnan(ref).any()
