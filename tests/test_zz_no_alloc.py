"""oct_topfrac_metric_noalloc (the sharpness metric with the library's own device scratch, for
callers that cannot allocate) against oct_topfrac_metric with a caller workspace, through the C ABI.

Both run the same kernels, and the metric is exact and deterministic, so the results must be
bit-identical.  Two calls in a row, the second on a smaller input, check that nothing of the first
call's scratch contents leaks into the second (the histograms and selection states are cleared by
the metric's first kernel).  Sizes above the 4 MiB scratch are rejected before any launch.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def xp():
    from oct_recon_af_v1_amd.backend import get_xp
    return get_xp("hip")[0]


def _mags(xp, shape, seed):
    torch = xp.torch
    g = torch.Generator(device=xp.device)
    g.manual_seed(seed)
    m = torch.rand(shape, generator=g, device=xp.device, dtype=torch.float32) * 3.0
    m[..., 100:140] *= 40.0                         # a bright band, as a reflector gives
    m[0, :5, :] = 0.0                               # ties at zero
    return m.contiguous()


def _with_workspace(xp, mag, zmin, frac):
    from oct_recon_af_v1_amd import _lib
    torch = xp.torch
    L = _lib.lib()
    B, nx, nz = mag.shape
    nb = L.oct_topfrac_metric_workspace_bytes(B, nx, nz, zmin)
    ws = torch.empty((nb,), dtype=torch.uint8, device=xp.device)
    out = torch.empty((B,), dtype=torch.float64, device=xp.device)
    _lib.check(L.oct_topfrac_metric(mag.data_ptr(), B, nx, nz, zmin, frac, out.data_ptr(), ws.data_ptr(), nb,
                                    xp.stream()), "oct_topfrac_metric")
    xp.synchronize()
    return out.cpu().numpy()


def _no_alloc(xp, mag, zmin, frac):
    from oct_recon_af_v1_amd import _lib
    torch = xp.torch
    B, nx, nz = mag.shape
    out = torch.full((B,), float("nan"), dtype=torch.float64, device=xp.device)
    _lib.check(_lib.lib().oct_topfrac_metric_noalloc(mag.data_ptr(), B, nx, nz, zmin, frac, out.data_ptr(),
                                                     xp.stream()), "oct_topfrac_metric_noalloc")
    xp.synchronize()
    return out.cpu().numpy()


def test_no_alloc_metric_equals_workspace_metric(xp):
    from oracle import ref_path as R
    cases = [((24, 512, 1024), 20, 0.01, 1), ((3, 200, 300), 7, 0.05, 2)]    # the second call is smaller
    for shape, zmin, frac, seed in cases:
        print(f"case: no-alloc topfrac metric, mags {shape}, zmin {zmin}, frac {frac}", flush=True)
        mag = _mags(xp, shape, seed)
        ref = _with_workspace(xp, mag, zmin, frac)
        got = _no_alloc(xp, mag, zmin, frac)
        assert np.array_equal(got, ref), (shape, got, ref)
        # and the value is the metric: log mean of the k largest |A|^2 (float64 oracle)
        o = R.top_percentile_metric(mag.cpu().numpy(), zmin, frac)
        np.testing.assert_allclose(got, o, rtol=0, atol=1e-12)


def test_no_alloc_metric_rejects_oversize(xp):
    from oct_recon_af_v1_amd import _lib
    L = _lib.lib()
    torch = xp.torch
    B, nx, nz = 1000, 16, 64                       # needs about 8 MiB of scratch
    assert L.oct_topfrac_metric_workspace_bytes(B, nx, nz, 0) > (4 << 20)
    mag = torch.ones((B, nx, nz), dtype=torch.float32, device=xp.device)
    out = torch.zeros((B,), dtype=torch.float64, device=xp.device)
    rc = L.oct_topfrac_metric_noalloc(mag.data_ptr(), B, nx, nz, 0, 0.01, out.data_ptr(), xp.stream())
    assert rc == -1
    assert b"no-allocation" in L.oct_last_error()
    xp.synchronize()
    assert float(out.abs().sum().item()) == 0.0     # nothing was launched
