"""sp_dual_stats_noalloc, the entry point that allocates nothing: its partials
and totals live in a static device array instead of the backend's buffers.

It runs sp_dual_stats's own kernels with the same grid, so on every call its
result is bitwise sp_dual_stats's (dual_stats: crates/saddlepoint-cpu/src/
lib.rs:217-224). Checked on consecutive calls, a later one smaller than the
one before (a one-workgroup reduction that writes its totals directly, after a
1024-workgroup one that left partials behind), and larger again.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import netlib
from saddlepoint_amd import _lib, generate
from saddlepoint_amd._lib import DualStats
from saddlepoint_amd.backend import ScaledProblem

pytestmark = pytest.mark.gpu


def noalloc():
    """The entry point through the C ABI (its ctypes signature in _lib)."""
    return _lib.lib().sp_dual_stats_noalloc


def case(gpu, p, seed):
    pg = gpu.load(ScaledProblem.new(p))
    rng = np.random.default_rng(seed)
    y = gpu.upload(rng.normal(size=p.nrows))
    aty = gpu.upload(rng.normal(size=p.ncols) * 10.0)
    return pg, y, aty


def test_noalloc_dual_stats_is_bitwise_dual_stats(gpu):
    f = noalloc()
    big = generate.config1(scale=0.25)     # m + n = 325k: 159 workgroups, finished by sp_red_finish
    huge = generate.config1(scale=1.0)     # m + n = 1.3M: the 1024-workgroup cap
    small = netlib("afiro")                # m + n = 59: one workgroup, totals written directly
    cases = [("config1@0.25", big, 1), ("afiro", small, 2), ("config1", huge, 3), ("afiro", small, 4)]
    for name, p, seed in cases:
        pg, y, aty = case(gpu, p, seed)
        ref = gpu.dual_stats(pg, y, aty)
        got = DualStats()
        rc = f(gpu.h, pg.h, y.h, aty.h, C.byref(got))
        assert rc == 0, _lib.lib().sp_last_error(gpu.h).decode()
        print(f"{name}: m={p.nrows} n={p.ncols}: residual_sq {got.residual_sq!r} objective {got.objective!r}",
              flush=True)
        assert got.residual_sq == ref.residual_sq and got.objective == ref.objective, name
        assert np.isfinite(got.objective) and got.residual_sq > 0.0
        pg.free()


def test_noalloc_dual_stats_errors(gpu):
    f = noalloc()
    pg, y, aty = case(gpu, netlib("afiro"), 5)
    out = DualStats()
    assert f(gpu.h, pg.h, aty.h, y.h, C.byref(out)) == _lib.SP_ERR_SHAPE
    assert f(gpu.h, pg.h, y.h, aty.h, None) == _lib.SP_ERR_INVALID
    assert f(gpu.h, None, y.h, aty.h, C.byref(out)) == _lib.SP_ERR_INVALID
    assert f(gpu.h, pg.h, y.h, aty.h, C.byref(out)) == 0
    pg.free()
