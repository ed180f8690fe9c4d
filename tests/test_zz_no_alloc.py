"""ufx_cc_residual_norm_noalloc, the residual norm that allocates nothing
(its partial sums and result live in a static device array), through the
C-ABI against ufx_cc_residual_norm, bit for bit: on two contexts in a row,
the second smaller than the first (so the array still holds the first call's
partial sums beyond the second's), each for every variable and all four
together."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _norm(fn, s, var):
    from unsflow2d_amd import cc
    v = C.c_double(float("nan"))
    rc = fn(s._h, var, C.byref(v))
    assert rc == 0, cc.lib.ufx_last_error().decode()
    return v.value


def test_noalloc_norm_equals_allocating_norm():
    from unsflow2d_amd import cc, meshgen
    lib = cc.lib
    for nx, ny, seed, prec in ((300, 150, 3, "fp64"), (20, 8, 5, "mixed")):
        m = cc.Mesh.from_raw(meshgen.bump_channel(nx, ny, seed, h=0.06))
        s = cc.Solver(m, cc.Config(mach=0.6, precision=prec))
        print(f"case: noalloc norm, bump {nx}x{ny} seed {seed} {prec}, "
              f"{s.counts['noc']} cells", flush=True)
        s.keep_residual(True)
        s.iterate(3, hist=False)
        for var in (-1, 0, 1, 2, 3):
            a = _norm(lib.ufx_cc_residual_norm_noalloc, s, var)
            b = _norm(lib.ufx_cc_residual_norm, s, var)
            print(f"  var {var:2d}: {a!r} {b!r}", flush=True)
            assert np.float64(a).view(np.int64) == np.float64(b).view(np.int64)
            assert np.isfinite(a) and a > 0.0
        # a wrong variable is refused before any work
        v = C.c_double()
        assert lib.ufx_cc_residual_norm_noalloc(s._h, 7, C.byref(v)) == 1
        s.close()
        m.close()
