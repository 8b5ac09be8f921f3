"""The FC-fit loss on the device (tvbx_fc_loss: FP64 co-moments, the loss
and its derivative; the samples' gradient as batched float32 products
between two tvbx_transpose calls) against the same loss written in float64
torch (autograd.fc_loss, whose gradient torch's autograd derives on its
own), for a shared and a per-instance target, an instance count that is not
a multiple of 64, and the gradient workload's size (2000 samples, N = 84,
P = 4096)."""

import pytest
import torch

from tvb_hpc import autograd as AG

pytestmark = pytest.mark.gpu


def _case(M, N, P, per_instance, seed):
    g = torch.Generator(device='cpu').manual_seed(seed)
    # a shared slow signal with node-dependent gains gives structured FC
    base = torch.randn(M, 1, P, generator=g)
    gain = torch.rand(1, N, 1, generator=g)
    x = (torch.randn(M, N, P, generator=g) + 1.5 * gain * base).float()
    if per_instance:
        t = torch.randn(P, N, N, generator=g, dtype=torch.float64)
        tgt = torch.tanh(t + t.transpose(1, 2))
    else:
        t = torch.randn(N, N, generator=g, dtype=torch.float64)
        tgt = torch.tanh(t + t.T)
    gout = torch.rand(P, generator=g, dtype=torch.float64) + 0.5
    return x.cuda(), tgt.cuda(), gout.cuda()


@pytest.mark.parametrize('M,N,P,per_instance', [
    (300, 20, 128, False),
    (300, 20, 100, True),     # P not a multiple of 64: staged samples
    (64, 3, 64, False),       # the smallest N with two pairs
    (2000, 84, 4096, False),  # the gradient workload's samples
])
def test_fc_loss_device_matches_float64_torch(M, N, P, per_instance):
    x, tgt, gout = _case(M, N, P, per_instance, seed=M + N + P)
    xr = x.double().requires_grad_(True)
    lr = AG.fc_loss(xr, tgt)
    (lr * gout).sum().backward()
    xd = x.clone().requires_grad_(True)
    ld = AG.fc_loss_device(xd, tgt)
    (ld.double() * gout).sum().backward()
    torch.cuda.synchronize()
    err_l = (ld.double() - lr.detach()).abs().max().item()
    gr = xr.grad
    err_g = ((xd.grad.double() - gr).abs().max() / gr.abs().max()).item()
    print('M %d N %d P %d per_instance %s: loss err %.2e, grad err %.2e '
          '(of max |grad|)' % (M, N, P, per_instance, err_l, err_g))
    # the loss is FP64 on both sides, rounded to float32 on the device
    assert err_l <= 1e-6
    # float32 products of the gradient against float64 autograd
    assert err_g <= 1e-4


def test_fc_loss_device_refuses_host_samples():
    with pytest.raises(ValueError):
        AG.fc_loss_device(torch.zeros(4, 3, 2), torch.zeros(3, 3))


def test_transpose_with_centring():
    from tvb_hpc.hip import Context
    ctx = Context.default()
    g = torch.Generator(device='cpu').manual_seed(3)
    for rows, cols, sub_rows in ((130, 70, 13), (64, 64, 1), (5, 300, 5)):
        a = torch.randn(rows, cols, generator=g).cuda()
        sub = torch.randn(sub_rows, cols, generator=g).cuda()
        out = torch.empty(cols, rows, device='cuda')
        torch.cuda.synchronize()
        ctx.transpose(a.data_ptr(), rows, cols, out.data_ptr(),
                      sub.data_ptr(), sub_rows)
        ctx.sync()
        idx = torch.arange(rows, device='cuda') % sub_rows
        ref = (a - sub[idx]).T
        assert torch.equal(out, ref)
        ctx.transpose(a.data_ptr(), rows, cols, out.data_ptr())
        ctx.sync()
        assert torch.equal(out, a.T)
