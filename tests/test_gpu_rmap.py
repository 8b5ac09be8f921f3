"""Region mapping (surface vertices <-> regions) through libtvbx's C ABI:
the reference's known answers (tvb_hpc/tests.py:453-473) through the
reference-shaped kernel objects, and bit-exact agreement with the oracle's
ascending-order float32 sums on random data."""

import numpy as np
import pytest

import oracle as O
from tvb_hpc.hip import HipTarget, RegionMap
from tvb_hpc.network import RMapFromAvg, RMapToAvg

pytestmark = pytest.mark.gpu


def test_rmap_to_avg_reference_case(ctx):
    knl = RMapToAvg().kernel(HipTarget())
    i = np.r_[:16].reshape((-1, 1))
    rmap = i // 4
    node = i.astype('f')
    roi = np.zeros((4, 1), 'f')
    knl(nroi=4, nvar=1, nnode=16, rmap=rmap, node=node, roi=roi)
    np.testing.assert_allclose(roi[:, 0], node.reshape((4, 4)).sum(axis=1))


def test_rmap_from_avg_reference_case(ctx):
    knl = RMapFromAvg().kernel(HipTarget())
    i = np.r_[:16].reshape((-1, 1))
    rmap = i // 4
    node = np.zeros((16, 1), 'f')
    roi = np.r_[:4].reshape((4, 1)).astype('f')
    knl(nroi=4, nvar=1, nnode=16, rmap=rmap, node=node, roi=roi)
    np.testing.assert_allclose(rmap, node)


@pytest.mark.parametrize('nnode,nroi,nvar,P', [
    (1000, 68, 1, 1), (5000, 76, 2, 3), (20000, 84, 1, 64),
    (4096, 10, 3, 130)])
def test_rmap_matches_oracle_bit_exact(ctx, nnode, nroi, nvar, P):
    print('rmap case', nnode, nroi, nvar, P, flush=True)
    rng = np.random.default_rng(nnode + P)
    # the last region has no vertices; regions of very different sizes
    rmap = np.minimum(rng.geometric(0.05, size=nnode) - 1, nroi - 2)
    rmap = rng.permutation(rmap).astype(np.int32)
    node = rng.standard_normal((nnode, nvar, P)).astype(np.float32)
    m = RegionMap(rmap, nroi)
    roi = m.to_avg(node)
    ref = O.rmap_to_avg(rmap, node, nroi)
    assert np.array_equal(roi, ref)
    assert not roi[nroi - 1].any()
    back = m.from_avg(ref)
    assert np.array_equal(back, O.rmap_from_avg(rmap, ref))


def test_rmap_device_tensors(ctx):
    torch = pytest.importorskip('torch')
    rng = np.random.default_rng(11)
    nnode, nroi, nvar, P = 30000, 80, 2, 256
    rmap = rng.integers(0, nroi, size=nnode).astype(np.int32)
    node = rng.standard_normal((nnode, nvar, P)).astype(np.float32)
    m = RegionMap(rmap, nroi)
    t = torch.from_numpy(node).cuda()
    roi_t = m.to_avg(t)
    assert roi_t.is_cuda
    ref = O.rmap_to_avg(rmap, node, nroi)
    assert np.array_equal(roi_t.cpu().numpy(), ref)
    out = torch.full((nnode, nvar, P), 7.0, device='cuda')
    m.from_avg(roi_t, out)
    assert np.array_equal(out.cpu().numpy(), O.rmap_from_avg(rmap, ref))
    # a 2-D (rows, nvar) array is n_inst = 1
    assert np.array_equal(m.to_avg(node[:, :, 0]), ref[:, :, 0])
