"""BF16 MFMA convolution kernels and the surrogate network against torch CPU FP32.

The oracle is a library routine: torch's CPU conv3d / linear in FP32 on the same
bf16-rounded inputs and weights (and, for the backward, the same bf16-rounded output
gradient that the kernels multiply), so the only differences are FP32 summation order.
"""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

pytestmark = pytest.mark.gpu


def bf(t):
    return t.to(torch.bfloat16).float()


def reference_conv(x, w, b, stride, pad, activation):
    """CPU FP32: x [n,d,h,w,c] and w [co,k,k,k,ci] already bf16-valued."""
    pre = F.conv3d(x.permute(0, 4, 1, 2, 3), w.permute(0, 4, 1, 2, 3), b, stride=stride, padding=pad).permute(0, 2, 3, 4, 1)
    return pre, (F.gelu(pre) if activation else pre)


def rel(a, b):
    return float((a-b).norm()/b.norm())


@pytest.mark.parametrize('shape,ci,co,ks,stride,pad', [((2, 16, 12, 8), 1, 32, 3, 2, 1), ((3, 8, 8, 4), 32, 64, 3, 2, 1),
                                                       ((2, 5, 6, 3), 16, 40, 3, 1, 1), ((37, 1, 1, 1), 256, 32, 1, 1, 0)])
def test_conv_forward_and_gradients_match_cpu(gpu, shape, ci, co, ks, stride, pad):
    from torchfdtd_amd.surrogate import conv3d
    gen = torch.Generator().manual_seed(ci*co+ks)
    x = bf(torch.randn((*shape, ci), generator=gen))
    w = bf(torch.randn((co, ks, ks, ks, ci), generator=gen)/np.sqrt(ks**3*ci))
    b = torch.randn(co, generator=gen)
    for activation in (True, False):
        xg = x.cuda().requires_grad_(True)
        wg, bg = w.cuda().requires_grad_(True), b.cuda().requires_grad_(True)
        y = conv3d(xg, wg, bg, stride, pad, activation)
        xr, wr = x.clone().requires_grad_(True), w.clone().requires_grad_(True)
        pre, yr = reference_conv(xr, wr, b, stride, pad, activation)
        # Forward: the activation is stored in bf16.
        assert rel(y.cpu(), bf(yr.detach())) < 4e-3
        np.testing.assert_allclose(y.detach().cpu().numpy(), bf(yr.detach()).numpy(), rtol=1e-2, atol=1e-2)
        g = torch.randn(y.shape, generator=gen)
        y.backward(g.cuda())
        dpre = g*(torch.autograd.functional.vjp(F.gelu, pre.detach(), torch.ones_like(pre))[1] if activation else 1)
        dx, dw = torch.autograd.grad(pre, (xr, wr), bf(dpre))
        # GELU' is evaluated on each side before the bf16 rounding of dpre: a few roundings may differ by one ulp.
        tol = 1e-3 if activation else 1e-5
        assert rel(wg.grad.cpu(), dw) < tol
        assert rel(bg.grad.cpu(), dpre.sum((0, 1, 2, 3))) < 1e-5
        assert rel(xg.grad.cpu(), dx) < tol


def test_surrogate_network_gradients_and_training(gpu):
    from torchfdtd_amd.surrogate import SurrogateNet
    net = SurrogateNet(seed=3)
    assert 1.1e6 < net.parameter_count() < 1.3e6
    cpu = SurrogateNet(seed=3)
    net.cuda()
    gen = torch.Generator().manual_seed(1)
    rho = (torch.rand((6, 64, 64, 8), generator=gen) > .5).float()
    target = torch.rand((6, 32), generator=gen)
    loss = F.mse_loss(net(rho.cuda()), target.cuda())
    loss.backward()
    # CPU FP32 network on the same bf16-rounded weights (activations unrounded): bf16 rounding bounds the gap.
    with torch.no_grad():
        for p in cpu.parameters():
            p.copy_(bf(p))
    for p, q in zip(net.parameters(), cpu.parameters()):
        with torch.no_grad():
            p.copy_(q)
    x = rho.unsqueeze(-1)
    for w, b in zip(cpu.weights, cpu.biases):
        x = F.gelu(F.conv3d(x.permute(0, 4, 1, 2, 3), w.permute(0, 4, 1, 2, 3), b, stride=2, padding=1).permute(0, 2, 3, 4, 1))
    x = x.mean((1, 2, 3))
    x = F.gelu(x@cpu.head[0].T+cpu.head[1])
    ref_loss = F.mse_loss(x@cpu.head[2].T+cpu.head[3], target)
    ref_loss.backward()
    net.zero_grad()
    out = net(rho.cuda())
    F.mse_loss(out, target.cuda()).backward()
    assert abs(F.mse_loss(out, target.cuda()).item()-ref_loss.item()) < 2e-2*ref_loss.item()
    for p, q in zip(net.parameters(), cpu.parameters()):
        assert rel(p.grad.cpu(), q.grad) < 5e-2
    # A few AdamW steps on a fixed batch reduce the loss.
    opt = torch.optim.AdamW(net.parameters(), lr=3e-4)
    losses = []
    for _ in range(30):
        opt.zero_grad()
        value = F.mse_loss(net(rho.cuda()), target.cuda())
        value.backward()
        opt.step()
        losses.append(value.item())
    assert losses[-1] < .5*losses[0]


def test_slab_flux_cohorts_match_single_runs(gpu):
    from torchfdtd_amd.surrogate import slab_flux
    from torchfdtd_amd.workloads import slab_densities
    rho = slab_densities(3, 11)
    together = slab_flux(rho, steps=120, frequencies=8, cohort=3)
    alone = np.concatenate([slab_flux(rho[i:i+1], steps=120, frequencies=8, cohort=1) for i in range(3)])
    assert together.shape == (3, 8) and np.isfinite(together).all() and np.abs(together).max() > 0
    assert np.array_equal(together, alone)
    assert np.array_equal(slab_flux(rho, steps=120, frequencies=8, cohort=3), together)   # reused cohort, reset state


def test_trained_surrogate_screens_better_than_unscreened_picks(gpu):
    """A small version of the design study: solver data -> training -> ranking of held-out designs.

    The judge is the FDTD transmission the batched solver computed for every design; the method
    without the surrogate picks designs blindly.  The surrogate's top picks must beat the blind
    average figure of merit and rank the held-out designs with a positive correlation."""
    from torchfdtd_amd.surrogate import SurrogateNet, band_figure_of_merit, predict, slab_flux, slab_wavelengths, train
    from torchfdtd_amd.workloads import slab_densities
    n, held, steps, frequencies = 512, 96, 600, 8
    rho = slab_densities(n, 5)
    reference = slab_flux(np.zeros((1, *rho.shape[1:]), np.float32), steps=steps, frequencies=frequencies)[0]
    t = slab_flux(rho, steps=steps, frequencies=frequencies)/reference
    wavelengths = slab_wavelengths(frequencies)
    fom = band_figure_of_merit(torch.as_tensor(t), wavelengths, (1.50, 1.60)).numpy()
    x = torch.as_tensor(rho, device='cuda')
    y = torch.as_tensor(t, device='cuda', dtype=torch.float32)
    scale = y[held:].std()
    net = SurrogateNet(outputs=frequencies, seed=2).cuda()
    history = train(net, x[held:], (y[held:]-1)/scale, epochs=40, batch=64, lr=1e-3)
    assert history[-1] < history[0]
    predicted = predict(net, x[:held])*scale+1
    score = band_figure_of_merit(predicted.cpu(), wavelengths, (1.50, 1.60)).numpy()
    top = np.argsort(-score)[:12]
    assert fom[:held][top].mean() > fom[:held].mean()
    ranks = lambda v: np.argsort(np.argsort(v))
    assert np.corrcoef(ranks(score), ranks(fom[:held]))[0, 1] > .2
