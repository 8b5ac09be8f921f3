"""The allocation-free FC entry point (tvbx_corr_noalloc: device samples in,
device correlation out, every intermediate in the library's fixed device
scratch array) against tvbx_corr (host arrays, device buffers allocated
per call), through the C ABI: bit-identical results on two calls in a row,
the second smaller, and a clear error for sizes beyond the scratch."""

import ctypes

import numpy as np
import pytest

from tvb_hpc import hip

pytestmark = pytest.mark.gpu


def _samples(n, N, P, seed):
    rng = np.random.default_rng(seed)
    common = rng.standard_normal((n, 1, P))
    x = 0.6 * common + rng.standard_normal((n, N, P))
    return (x + 3.0).astype(np.float32)


def test_corr_noalloc_matches_corr(ctx):
    import torch
    # second call smaller; P = 100 goes through the pitched copy
    for n, N, P, seed in ((300, 84, 100, 1), (50, 30, 64, 2)):
        s = _samples(n, N, P, seed)
        want = ctx.corr(s)
        d_s = torch.from_numpy(s).cuda()
        d_c = torch.full((P, N, N), float('nan'), dtype=torch.float32,
                         device='cuda')
        torch.cuda.synchronize()
        ctx.corr_noalloc(d_s.data_ptr(), n, N, P, d_c.data_ptr())
        ctx.sync()
        got = d_c.cpu().numpy()
        print('n %d N %d P %d: max |diff| %.3e' % (n, N, P,
                                                  np.abs(got - want).max()))
        np.testing.assert_array_equal(got, want)


def test_corr_noalloc_rejects_sizes_beyond_scratch(ctx):
    import torch
    # N = 998, P = 1024: 4 GB of FP64 co-moments, more than the scratch
    # holds; the host rejects the call before anything is launched
    buf_in = torch.zeros(16, dtype=torch.float32, device='cuda')
    buf_out = torch.zeros(16, dtype=torch.float32, device='cuda')
    with pytest.raises(hip.TvbxError, match='scratch'):
        ctx.corr_noalloc(buf_in.data_ptr(), 10, 998, 1024, buf_out.data_ptr())
    # the context still works
    s = _samples(20, 8, 64, 3)
    assert np.isfinite(ctx.corr(s)).all()
