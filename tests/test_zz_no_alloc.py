"""sfh_gp_noalloc, the allocation-free form of sfh_gp (dense GP fit and
predictions with the scratch in a fixed device array): two calls in a row on
device inputs, the second smaller, each against sfh_gp on the same inputs and
against the oracle; sizes beyond the fixed scratch are refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(ctx, X, y, sigma, ell, eps, Xp):
    from statfem_amd import hip
    n, dim = X.shape
    np_ = Xp.shape[0]
    Xd, yd, Xpd = ctx.to_device(X), ctx.to_device(y), ctx.to_device(Xp)
    ld, md, vd = ctx.empty((1,)), ctx.empty((np_,)), ctx.empty((np_,))
    sd = ctx.empty((1,), dtype=np.int32)
    rc = hip.lib().sfh_gp_noalloc(ctx.h, n, dim, Xd.ptr, yd.ptr, sigma, ell, eps, np_, Xpd.ptr, ld.ptr, md.ptr,
                                  vd.ptr, sd.ptr)
    ctx.sync()
    return rc, ld.download()[0], md.download(), vd.download(), int(sd.download()[0])


@pytest.mark.parametrize("n,np_", [(96, 1500), (17, 40)])
def test_gp_noalloc_matches_sfh_gp(ctx, orc, n, np_):
    from statfem_amd import hip
    rng = np.random.default_rng(n)
    for rep, (nn, pp) in enumerate(((n, np_), (n // 2 + 1, np_ // 3 + 1))):   # second call smaller
        X = rng.uniform(size=(nn, 2))
        y = np.cos(4 * X[:, 0]) + X[:, 1] ** 2
        Xp = rng.uniform(size=(pp, 2))
        rc, ll, m, v, st = _noalloc(ctx, X, y, 0.9, 0.3, 1e-3, Xp)
        assert rc == 0 and st == 0, (rc, st)
        ll_ref, m_ref, v_ref = hip.gaussian_process(ctx, X, y, 0.9, 0.3, 1e-3, Xp)
        assert ll == ll_ref
        assert np.array_equal(m, m_ref) and np.array_equal(v, v_ref)
        lo, mo, vo = orc.gp(X, y, 0.9, 0.3, 1e-3, Xp)
        assert abs(ll - lo) <= 1e-10 * abs(lo)
        assert np.abs(m - mo).max() <= 1e-10 * np.abs(mo).max()
        assert np.abs(v - vo).max() <= 1e-10


def test_gp_noalloc_refuses_sizes_beyond_its_scratch(ctx):
    from statfem_amd import hip
    X = np.zeros((257, 1))
    d = ctx.to_device(X)
    y = ctx.to_device(np.zeros(257))
    out = ctx.empty((1,))
    st = ctx.empty((1,), dtype=np.int32)
    rc = hip.lib().sfh_gp_noalloc(ctx.h, 257, 1, d.ptr, y.ptr, 1.0, 0.3, 1e-3, 0, None, out.ptr, None, None, st.ptr)
    assert rc == -1
    assert b"at most 256" in hip.lib().sfh_last_error()
    rc = hip.lib().sfh_gp_noalloc(ctx.h, 10, 1, d.ptr, y.ptr, 1.0, 0.3, 1e-3, 2049, d.ptr, out.ptr, out.ptr,
                                  out.ptr, st.ptr)
    assert rc == -1
