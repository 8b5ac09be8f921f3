"""
PyTorch autograd wrapper of the device integrator's hand-written adjoint.

``integrate(sim, params, x0, n_steps, window)`` runs the simulation forward
(keeping a checkpoint every ``window`` steps) and returns its monitor samples
(n_samples, N, P) as a tensor; backward replays each window from its
checkpoint and runs the reverse sweep of the delayed Euler-Maruyama
recursion (csrc/adjoint_template.hip), returning dL/dparam for the tensors in
``params`` (coupling parameters (P,), node-and-instance model parameters
(N, P)).  Any differentiable torch loss of the samples can sit on top, e.g.
``fc_loss`` below (1 - Pearson correlation of simulated and target FC, the
Bayesian-inversion use the reference names in README.md:3-4).
``logp(sim, params, traj)`` is the Euler-Maruyama log-likelihood of a given
trajectory (scheme.py:42-48) with its parameter gradients.

PyTorch is used only for the autograd plumbing and the loss; the forward and
reverse sweeps are the HIP kernels behind libtvbx.
"""

import numpy as np
import torch


class _Integrate(torch.autograd.Function):

    @staticmethod
    def forward(ctx, sim, keys, x0, n_steps, window, seed, *values):
        params = {k: v.detach().double().cpu().numpy()
                  for k, v in zip(keys, values)}
        sim.set_params(params)
        sim.set_state(x0)
        sim.step = 0
        sim.init_history('zeros')
        sim.set_noise(seed)
        # one backward pass per forward: it may replay inside the checkpoints
        sim.set_checkpoint_reuse(False)
        sim.checkpoint_run(n_steps, window)
        ctx.sim, ctx.keys = sim, keys
        ctx.device = values[0].device
        ctx.dtypes = [v.dtype for v in values]
        if ctx.device.type == 'cuda':
            # samples stay on the device: one strided device-to-device copy
            n, first = sim.pending_samples()
            out = torch.empty((n, sim.N, sim.P), dtype=torch.float32,
                              device=ctx.device)
            sim.samples_into(out.data_ptr())
            ctx.first = first
            return out
        samples, first = sim.samples()
        ctx.first = first
        return torch.from_numpy(samples).to(ctx.device)

    @staticmethod
    def backward(ctx, d_samples):
        if d_samples.is_cuda:
            # dL/dsamples stays on the device; torch's stream must have
            # finished writing it before the simulation's stream reads it
            ds = d_samples.detach().float().contiguous()
            torch.cuda.current_stream(ds.device).synchronize()
            g = ctx.sim.adjoint(ds.data_ptr(), ctx.first,
                                n_samples=ds.shape[0])
        else:
            g = ctx.sim.adjoint(d_samples.detach().float().cpu().numpy(),
                                ctx.first)
        grads = [torch.from_numpy(np.asarray(g[k])).to(ctx.device, dt)
                 for k, dt in zip(ctx.keys, ctx.dtypes)]
        return (None, None, None, None, None, None) + tuple(grads)


def integrate(sim, params, x0, n_steps, window, seed=None):
    """Monitor samples (n_steps // tavg_period, N, P) of ``sim`` from state
    x0 (S, N, P), differentiable w.r.t. the tensors in ``params``
    ({'c:a': (P,), 'm:eta': (N, P), ...}; keys as in ``Simulation``).
    ``window``: steps between checkpoints, or 0 to record every step in
    device memory (no forward replay in the backward pass)."""
    keys = list(params)
    return _Integrate.apply(sim, keys, x0, int(n_steps), int(window), seed,
                            *[params[k] for k in keys])


class _Logp(torch.autograd.Function):

    @staticmethod
    def forward(ctx, sim, keys, traj, window, *values):
        params = {k: v.detach().double().cpu().numpy()
                  for k, v in zip(keys, values)}
        sim.set_params(params)
        sim.step = 0
        sim.set_state(traj[0])
        sim.init_history('zeros')
        lp, g = sim.logp(traj, window=window, grad=True)
        ctx.grads = g
        ctx.keys = keys
        ctx.device = values[0].device
        ctx.dtypes = [v.dtype for v in values]
        return torch.from_numpy(lp).to(ctx.device)

    @staticmethod
    def backward(ctx, d_logp):
        # logp[p] depends on instance p's parameters only
        go = d_logp.detach().double().cpu().numpy()
        grads = []
        for k, dt in zip(ctx.keys, ctx.dtypes):
            g = np.asarray(ctx.grads[k], np.float64) * go
            grads.append(torch.from_numpy(g).to(ctx.device, dt))
        return (None, None, None, None) + tuple(grads)


def logp(sim, params, traj, window=0):
    """Euler-Maruyama log-likelihood (P,) of the trajectory traj
    (n + 1, S, N, P) under ``sim`` (scheme.py:42-48, teacher forcing, history
    zeros before traj[0]), differentiable w.r.t. the tensors in ``params``
    (keys and shapes as in ``integrate``; ``sim`` created with
    adjoint=True)."""
    keys = list(params)
    tr = np.ascontiguousarray(traj, np.float32)
    return _Logp.apply(sim, keys, tr, int(window),
                       *[params[k] for k in keys])


def fc(samples):
    """Pearson FC per instance: samples (M, N, P) -> (P, N, N)."""
    x = samples - samples.mean(dim=0, keepdim=True)
    xp = x.permute(2, 0, 1).contiguous()              # (P, M, N)
    cov = torch.bmm(xp.transpose(1, 2).contiguous(), xp)
    d = torch.sqrt(torch.diagonal(cov, dim1=1, dim2=2))
    return cov / (d[:, :, None] * d[:, None, :])


def fc_loss(samples, fc_target):
    """Per-instance 1 - corr(upper triangle of FC(samples), of fc_target);
    samples (M, N, P), fc_target (N, N) or (P, N, N) -> (P,)."""
    f = fc(samples.double())
    N = f.shape[-1]
    iu = torch.triu_indices(N, N, 1)
    a = f[:, iu[0], iu[1]]
    b = torch.as_tensor(fc_target, dtype=a.dtype, device=a.device)
    b = b[..., iu[0], iu[1]]
    if b.ndim == 1:
        b = b.expand_as(a)
    a = a - a.mean(dim=1, keepdim=True)
    b = b - b.mean(dim=1, keepdim=True)
    r = (a * b).sum(dim=1) / torch.sqrt((a * a).sum(dim=1) * (b * b).sum(dim=1))
    return 1.0 - r


class _FcLossDevice(torch.autograd.Function):
    """fc_loss of device samples on the HIP kernels: the FP64 co-moments
    and the loss with its derivative in libtvbx (tvbx_fc_loss), the samples'
    gradient as one batched float32 product per instance,
    dL_p/dx_p = gout_p (x_p - mean_p) S_p, between two transposes
    (tvbx_transpose) that take the instance axis outermost and back."""

    @staticmethod
    def forward(ctx, samples, fc_target, hctx):
        x = samples.detach().float().contiguous()
        M, N, P = x.shape
        dev = x.device
        tgt = torch.as_tensor(fc_target, device=dev).float()
        per = tgt.ndim == 3
        # the kernel reads a per-instance target instance-innermost
        tgt = (tgt.permute(1, 2, 0) if per else tgt).contiguous()
        loss = torch.empty(P, dtype=torch.float32, device=dev)
        s = torch.empty((N, N, P), dtype=torch.float32, device=dev)
        mean = torch.empty((N, P), dtype=torch.float32, device=dev)
        # torch's stream has written the inputs before the context's reads
        torch.cuda.current_stream(dev).synchronize()
        hctx.fc_loss(x.data_ptr(), M, N, P, tgt.data_ptr(), per,
                     loss.data_ptr(), s.data_ptr(), mean.data_ptr())
        hctx.sync()
        ctx.save_for_backward(x, s, mean)
        ctx.hctx = hctx
        ctx.dtype = samples.dtype
        return loss.to(samples.dtype)

    @staticmethod
    def backward(ctx, gout):
        x, s, mean = ctx.saved_tensors
        M, N, P = x.shape
        hctx = ctx.hctx
        dev = x.device
        xt = torch.empty((P, M * N), dtype=torch.float32, device=dev)
        sp = (s.permute(2, 0, 1) *
              gout.detach().float()[:, None, None]).contiguous()
        torch.cuda.current_stream(dev).synchronize()
        hctx.transpose(x.data_ptr(), M * N, P, xt.data_ptr(),
                       mean.data_ptr(), N)
        hctx.sync()
        y = torch.bmm(xt.view(P, M, N), sp)
        dx = torch.empty((M, N, P), dtype=torch.float32, device=dev)
        torch.cuda.current_stream(dev).synchronize()
        hctx.transpose(y.data_ptr(), P, M * N, dx.data_ptr())
        hctx.sync()
        return dx.to(ctx.dtype), None, None


def fc_loss_device(samples, fc_target, context=None):
    """fc_loss (above) of samples (M, N, P) on the device, FP64 co-moments
    and loss in libtvbx, float32 gradient; fc_target (N, N) or (P, N, N).
    Differentiable with respect to ``samples``."""
    if not samples.is_cuda:
        raise ValueError('fc_loss_device needs device samples; fc_loss '
                         'takes host tensors')
    from .hip import Context
    hctx = context or Context.default(samples.device.index or 0)
    return _FcLossDevice.apply(samples, fc_target, hctx)

