"""Learned surrogate for inverse-design screening: a 3D conv encoder on BF16 MFMA kernels.

A density rho (64 x 64 x 8 design cells) maps to the transmission at 32 wavelengths through
four stride-2 3x3x3 convolutions (1 -> 32 -> 64 -> 128 -> 256 channels, GELU), a mean over
the remaining positions and an MLP head (256 -> 256 GELU -> 32): about 1.24 M parameters.
Every convolution and linear layer runs on the library's implicit-GEMM kernels
(``tfdtd_conv_forward`` / ``tfdtd_conv_dgrad`` / ``tfdtd_conv_wgrad``, bf16 operands, FP32
accumulation); weights are kept in FP32 (the optimiser's master copy) and rounded to bf16
per call.  The reference has no neural surrogate: this is the learned screen of the
inverse-design loop, with the FDTD solve as the arbiter (``screen``).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch
from torch import nn

from . import _native as N


class ConvShape(ctypes.Structure):
    _fields_ = [('n', ctypes.c_int32), ('di', ctypes.c_int32), ('hi', ctypes.c_int32), ('wi', ctypes.c_int32),
                ('ci', ctypes.c_int32), ('co', ctypes.c_int32), ('ks', ctypes.c_int32), ('stride', ctypes.c_int32),
                ('pad', ctypes.c_int32), ('kp', ctypes.c_int32)]


def padded_k(ks, ci):
    return (ks**3*ci+31)//32*32


def conv_shape(x_shape, ci, co, ks, stride, pad):
    n, d, h, w, c = x_shape
    if c != ci:
        raise ValueError(f'input has {c} channels, the layer expects {ci}')
    s = ConvShape(n, d, h, w, ci, co, ks, stride, pad, padded_k(ks, ci))
    out = (ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32())
    N.check(N.library().tfdtd_conv_output(ctypes.byref(s), *(ctypes.byref(v) for v in out)), 'tfdtd_conv_output')
    return s, tuple(v.value for v in out)


def forward_weight(weight):
    """FP32 [co][ks][ks][ks][ci] -> bf16 [co][kp] with zero padding."""
    co = weight.shape[0]
    k = weight[0].numel()
    w16 = torch.zeros((co, padded_k(weight.shape[1], weight.shape[-1])), device=weight.device, dtype=torch.bfloat16)
    w16[:, :k] = weight.reshape(co, k).to(torch.bfloat16)
    return w16


class _Conv(torch.autograd.Function):
    """y = act(conv(x) + b), x and y FP32 channels-last; bf16 operands on MFMA inside."""

    @staticmethod
    def forward(ctx, x, weight, bias, stride, pad, activation):
        lib = N.library()
        co, ks, ci = weight.shape[0], weight.shape[1], weight.shape[-1]
        shape, (do, ho, wo) = conv_shape(tuple(x.shape), ci, co, ks, stride, pad)
        x16 = x.detach().to(torch.bfloat16).contiguous()
        w16 = forward_weight(weight.detach())
        m = x.shape[0]*do*ho*wo
        pre = torch.empty((m, co), device=x.device, dtype=torch.float32)
        act = torch.empty((m, co), device=x.device, dtype=torch.bfloat16)
        b = bias.detach().float().contiguous()
        N.check(lib.tfdtd_conv_forward(ctypes.byref(shape), x16.data_ptr(), w16.data_ptr(), ctypes.c_void_p(b.data_ptr()),
                                       int(activation), ctypes.c_void_p(pre.data_ptr()), act.data_ptr(), N.stream_handle()),
                'tfdtd_conv_forward')
        ctx.save_for_backward(x16, w16, pre)
        ctx.config = (shape, ks, ci, co, int(activation), tuple(x.shape))
        return act.float().view(x.shape[0], do, ho, wo, co)

    @staticmethod
    def backward(ctx, grad):
        lib = N.library()
        x16, w16, pre = ctx.saved_tensors
        shape, ks, ci, co, activation, x_shape = ctx.config
        stream = N.stream_handle()
        m = pre.shape[0]
        dpre = torch.empty_like(pre)
        dpre16 = torch.empty(pre.shape, device=pre.device, dtype=torch.bfloat16)
        dy = grad.contiguous().view(m, co).float()
        N.check(lib.tfdtd_gelu_backward(ctypes.c_void_p(dy.data_ptr()), ctypes.c_void_p(pre.data_ptr()), m*co, activation,
                                        ctypes.c_void_p(dpre.data_ptr()), dpre16.data_ptr(), stream), 'tfdtd_gelu_backward')
        grad_bias = dpre.sum(0)
        dy_t = dpre16.t().contiguous()
        splits = max(1, min(64, m//4096))
        partial = torch.empty((splits, co, shape.kp), device=pre.device, dtype=torch.float32)
        dw = torch.empty((co, shape.kp), device=pre.device, dtype=torch.float32)
        N.check(lib.tfdtd_conv_wgrad(ctypes.byref(shape), dy_t.data_ptr(), x16.data_ptr(), splits,
                                     ctypes.c_void_p(partial.data_ptr()), ctypes.c_void_p(dw.data_ptr()), stream), 'tfdtd_conv_wgrad')
        taps = ks**3
        grad_weight = dw[:, :taps*ci].reshape(co, ks, ks, ks, ci)
        grad_x = None
        if ctx.needs_input_grad[0]:
            wt = w16[:, :taps*ci].reshape(co, taps, ci).permute(2, 1, 0).contiguous()
            dx = torch.empty((x_shape[0]*x_shape[1]*x_shape[2]*x_shape[3], ci), device=pre.device, dtype=torch.float32)
            N.check(lib.tfdtd_conv_dgrad(ctypes.byref(shape), dpre16.data_ptr(), wt.data_ptr(), ctypes.c_void_p(dx.data_ptr()),
                                         stream), 'tfdtd_conv_dgrad')
            grad_x = dx.view(x_shape)
        return grad_x, grad_weight, grad_bias, None, None, None


def conv3d(x, weight, bias, stride=1, pad=0, activation=True):
    """Channels-last FP32 x [n,d,h,w,ci], weight [co,ks,ks,ks,ci] -> [n,do,ho,wo,co] (bf16 MFMA inside)."""
    return _Conv.apply(x, weight, bias, stride, pad, activation)


def linear(x, weight, bias, activation=True):
    """[n, ci] -> [n, co] through the same kernels (a 1x1x1 convolution)."""
    y = _Conv.apply(x.reshape(x.shape[0], 1, 1, 1, x.shape[1]), weight.reshape(weight.shape[0], 1, 1, 1, weight.shape[1]),
                    bias, 1, 0, activation)
    return y.reshape(x.shape[0], weight.shape[0])


class SurrogateNet(nn.Module):
    """rho [n, 64, 64, 8] -> transmission [n, 32]."""

    def __init__(self, channels=(32, 64, 128, 256), hidden=256, outputs=32, seed=0):
        super().__init__()
        gen = torch.Generator().manual_seed(seed)
        self.weights, self.biases = nn.ParameterList(), nn.ParameterList()
        ci = 1
        for co in channels:
            bound = 1/math.sqrt(27*ci)
            self.weights.append(nn.Parameter((torch.rand((co, 3, 3, 3, ci), generator=gen)*2-1)*bound))
            self.biases.append(nn.Parameter((torch.rand(co, generator=gen)*2-1)*bound))
            ci = co
        self.head = nn.ParameterList()
        for a, b in ((ci, hidden), (hidden, outputs)):
            bound = 1/math.sqrt(a)
            self.head.append(nn.Parameter((torch.rand((b, a), generator=gen)*2-1)*bound))
            self.head.append(nn.Parameter((torch.rand(b, generator=gen)*2-1)*bound))

    def forward(self, rho):
        x = rho.unsqueeze(-1).float()
        for w, b in zip(self.weights, self.biases):
            x = conv3d(x, w, b, stride=2, pad=1, activation=True)
        x = x.mean(dim=(1, 2, 3))
        x = linear(x, self.head[0], self.head[1], activation=True)
        return linear(x, self.head[2], self.head[3], activation=False)

    def parameter_count(self):
        return sum(p.numel() for p in self.parameters())

    def figure_of_merit(self, transmission):
        """The design objective evaluated on predicted spectra (set by the caller, e.g. a band filter)."""
        if self.objective is None:
            raise ValueError('set SurrogateNet.objective to a spectrum -> figure of merit function first')
        return self.objective(transmission)

    objective = None


# ------------------------------------------------------------------ data from the batched solver and the screen

def _plane_flux(plane):
    """Signed flux through a z-normal plane per frequency, on the device (field_monitors.py:167-183)."""
    plane.flush()
    comps = plane.components
    v = plane.value
    ex, ey = v[..., comps.index('Ex')], v[..., comps.index('Ey')]
    hx, hy = v[..., comps.index('Hx')], v[..., comps.index('Hy')]
    density = .5*(ex*hy.conj()-ey*hx.conj()).real
    weights = torch.as_tensor(plane.plan['weights'], device=v.device, dtype=density.dtype)
    return density@weights


_COHORTS = {}


def slab_flux(densities, *, steps=2000, frequencies=32, cohort=64, launch='split', width=50, progress=None):
    """FDTD flux spectra (count, frequencies) of design-slab densities, in batched cohorts.

    One cohort of `cohort` grids (``batch.Cohort``, its step loop captured in HIP graphs of `width`
    steps) is built once per size and reused: each group of designs only resets the permittivities
    and the state.  Only the plane accumulators are reduced, on the device, to flux spectra."""
    from .batch import Cohort, Launch
    from .workloads import design_slab, slab_epsilon
    project = design_slab(steps, frequencies)
    monitor = project.monitors[0]
    monitor.record_fields = ('Ex', 'Ey', 'Hx', 'Hy')
    monitor.record_poynting = ('z',)
    out = []
    with torch.no_grad():
        for start in range(0, len(densities), cohort):
            members = densities[start:start+cohort]
            key = (len(members), steps, frequencies, launch, width)
            if key not in _COHORTS:
                while _COHORTS:
                    _COHORTS.popitem()
                torch.cuda.empty_cache()
                group = Cohort([project]*len(members), Launch(launch) if isinstance(launch, str) else launch,
                               epsilons=[slab_epsilon(project, rho) for rho in members])
                _COHORTS[key] = (group, group.graphs(min(width, steps)))
            group, graphs = _COHORTS[key]
            group.reset([slab_epsilon(project, rho) for rho in members])
            done = 0
            while done < steps:
                advance = graphs.width if steps-done >= graphs.width else 1
                graphs.replay(advance)
                done += advance
            out.append(torch.stack([_plane_flux(run.planes[0]) for run in group.runs]).cpu())
            if progress:
                progress(start+len(members))
    return torch.cat(out).numpy()


def slab_transmission(densities, reference_flux, **kwargs):
    return slab_flux(densities, **kwargs)/reference_flux


def train(net, densities, targets, *, epochs=20, batch=128, lr=3e-4, weight_decay=1e-2, seed=0, log=None, cosine=False,
          mirror=False):
    """AdamW on mean-squared error; densities (n, 64, 64, 8), targets (n, outputs), both on the device.
    `cosine`: decay the learning rate to zero over the epochs.  `mirror`: flip each batch along x
    and/or y at random -- the centred slab under a normally incident Ex plane source transmits the
    same flux for a design and its x or y mirror image."""
    gen = torch.Generator(device='cpu').manual_seed(seed)
    opt = torch.optim.AdamW(net.parameters(), lr=lr, weight_decay=weight_decay)
    schedule = torch.optim.lr_scheduler.CosineAnnealingLR(opt, epochs) if cosine else None
    n = densities.shape[0]
    history = []
    for epoch in range(epochs):
        order = torch.randperm(n, generator=gen)
        total = 0.
        for i in range(0, n, batch):
            idx = order[i:i+batch].to(densities.device)
            x = densities[idx]
            if mirror:
                flips = torch.randint(0, 4, (1,), generator=gen).item()
                if flips & 1:
                    x = x.flip(1)
                if flips & 2:
                    x = x.flip(2)
            opt.zero_grad(set_to_none=True)
            loss = torch.nn.functional.mse_loss(net(x), targets[idx])
            loss.backward()
            opt.step()
            total += loss.item()*len(idx)
        if schedule:
            schedule.step()
        history.append(total/n)
        if log:
            log(epoch, history[-1])
    return history


@torch.no_grad()
def predict(net, densities, batch=512):
    return torch.cat([net(densities[i:i+batch]) for i in range(0, densities.shape[0], batch)])


def tilt_figure_of_merit(transmission, wavelengths_um, edge=1.55):
    """Mean transmission below `edge` minus mean above: an edge-filter figure of merit per design."""
    t = torch.as_tensor(transmission)
    w = torch.as_tensor(wavelengths_um, dtype=t.dtype)
    return t[..., w < edge].mean(-1)-t[..., w >= edge].mean(-1)


def match_figure_of_merit(transmission, target):
    """Minus the RMS difference from a target spectrum: 0 is a perfect match."""
    t = torch.as_tensor(transmission)
    return -((t-torch.as_tensor(target, dtype=t.dtype))**2).mean(-1).sqrt()


def band_figure_of_merit(transmission, wavelengths_um, band=(1.50, 1.60)):
    """Mean transmission inside `band` minus mean outside: a band-pass figure of merit per design."""
    t = torch.as_tensor(transmission)
    w = torch.as_tensor(wavelengths_um, dtype=t.dtype)
    inside = (w >= band[0]) & (w <= band[1])
    return t[..., inside].mean(-1)-t[..., ~inside].mean(-1)


def propose(rng, count, best=None, flip=.08):
    """Candidate densities: fresh smoothed-threshold patterns (fill fractions drawn from SLAB_FILL), and
    (with a best design) mutations of it that re-threshold the best pattern plus filtered noise."""
    from .workloads import SLAB_DESIGN, SLAB_FILL, binary_design, conic_filter
    out = np.empty((count, *SLAB_DESIGN), dtype=np.float32)
    for i in range(count):
        if best is None or i % 2 == 0:
            out[i] = binary_design(rng, SLAB_DESIGN[:2], fill=rng.uniform(*SLAB_FILL))[:, :, None]
        else:
            base = best[:, :, 0].astype(np.float64)
            noisy = conic_filter(np.clip(base+flip*rng.normal(size=base.shape)*8, 0, 1), 2.5)
            out[i] = (noisy > .5)[:, :, None]
    return out


class Surrogate:
    """The network with its target normalisation: spectra are learned as (t - mean) / std per frequency."""

    def __init__(self, net, mean, std):
        self.net, self.mean, self.std = net, mean, std

    def fit(self, densities, spectra, **kwargs):
        return train(self.net, densities, (spectra-self.mean)/self.std, **kwargs)

    def __call__(self, densities):
        return predict(self.net, densities)*self.std+self.mean


def screen(surrogate, solve, fom, *, iterations=6, candidates=4096, verify=32, seed=0, replay=None, refine_epochs=0,
           log=None):
    """Surrogate-screened design loop: each iteration proposes `candidates` designs, ranks them by the
    surrogate's figure of merit and solves only the top `verify` with FDTD (`solve`: densities ->
    spectra [n, k] on the device; `fom`: spectra -> figure of merit).  With `refine_epochs`, the
    surrogate is refitted after each iteration on every design solved so far plus `replay` (a
    (densities, spectra) sample of its training set).  Returns the FDTD history [(solves, best FOM)],
    the surrogate's FOM error on every solved design and the best design."""
    rng = np.random.default_rng(seed)
    device = surrogate.mean.device
    best, best_fom, solves, history, errors = None, -np.inf, 0, [], []
    seen_x, seen_y = [], []
    for it in range(iterations):
        pool = propose(rng, candidates, best)
        x = torch.as_tensor(pool, device=device)
        predicted = fom(surrogate(x)).cpu().numpy()
        if best is None:
            top = np.argsort(-predicted)[:verify]
        else:
            # Half the solves go to the best-predicted mutations of the best design (odd pool entries),
            # half to the best-predicted fresh designs, so the local search always advances.
            half = verify//2
            picks = [parity+2*np.argsort(-predicted[parity::2])[:n] for parity, n in ((1, half), (0, verify-half))]
            top = np.concatenate(picks)
        spectra = solve(pool[top])
        judged = fom(spectra).cpu().numpy()
        solves += verify
        errors.extend((judged-predicted[top]).tolist())
        k = int(np.argmax(judged))
        if judged[k] > best_fom:
            best, best_fom = pool[top][k], float(judged[k])
        history.append((solves, best_fom))
        if log:
            log(f'screen: iteration {it}: {solves} solves, best {best_fom:.4f}, '
                f'surrogate error {np.sqrt(np.mean((judged-predicted[top])**2)):.4f}')
        if refine_epochs and it+1 < iterations:
            seen_x.append(x[torch.as_tensor(top, device=device)])
            seen_y.append(spectra)
            xs, ys = torch.cat(seen_x), torch.cat(seen_y)
            if replay is not None:
                xs, ys = torch.cat([xs, replay[0]]), torch.cat([ys, replay[1]])
            surrogate.fit(xs, ys, epochs=refine_epochs, batch=64, lr=3e-4, weight_decay=1e-4, seed=seed+it)
    return history, np.asarray(errors), best


def unscreened(solve, fom, *, iterations=24, per_iteration=32, seed=0):
    """The same loop without the surrogate: every proposed design is solved with FDTD."""
    rng = np.random.default_rng(seed)
    best, best_fom, solves, history = None, -np.inf, 0, []
    for _ in range(iterations):
        pool = propose(rng, per_iteration, best)
        judged = fom(solve(pool)).cpu().numpy()
        solves += per_iteration
        k = int(np.argmax(judged))
        if judged[k] > best_fom:
            best, best_fom = pool[k], float(judged[k])
        history.append((solves, best_fom))
    return history, best


# ------------------------------------------------------------------ the design study (data, training, screen)

SLAB_WAVELENGTHS_UM = None


def slab_wavelengths(frequencies=32):
    """Wavelengths (um) of the design-slab plane's frequency samples, in sample order."""
    from .monitors import frequency_samples
    from .workloads import design_slab
    return 299792458e6/np.asarray(frequency_samples(design_slab(frequencies=frequencies).monitors[0].spectrum))


def generate(start, count, *, seed=2026, steps=2000, frequencies=32, cohort=64, log=print):
    """Transmission spectra of designs start .. start+count of the seeded design sequence."""
    from .workloads import slab_densities
    densities = slab_densities(start+count, seed)[start:]
    reference = slab_flux(np.zeros((1, *densities.shape[1:]), np.float32), steps=steps, frequencies=frequencies)[0]
    flux = slab_flux(densities, steps=steps, frequencies=frequencies, cohort=cohort,
                     progress=lambda done: log(f'generate: {done}/{count} designs'))
    return flux/reference, reference


def study_objectives(data, held, wavelengths, band=(1.50, 1.60)):
    """The design campaigns of the study: (name, spectra -> FOM).  Band-pass (in-band minus out-of-band
    mean transmission), edge tilt (short minus long half), and matches to the spectra of four held-out
    designs (reachable targets that the surrogate never saw)."""
    flat = np.asarray(data[:held])
    level = flat.mean(1)
    picks = [int(np.argmin(np.abs(level-np.quantile(level, q)))) for q in (.25, .5, .75, .9)]
    out = [('bandpass', lambda t: band_figure_of_merit(t.cpu(), wavelengths, band)),
           ('tilt', lambda t: tilt_figure_of_merit(t.cpu(), wavelengths))]
    for k in picks:
        target = torch.as_tensor(flat[k], dtype=torch.float32)
        out.append((f'match-{k}', lambda t, target=target: match_figure_of_merit(t.cpu().float(), target)))
    return out


def design_study(data, *, seed=2026, epochs=150, batch=128, lr=1e-3, candidates=4096, verify=32, iterations=6,
                 baseline_iterations=48, refine_epochs=4, steps=2000, frequencies=32, band=(1.50, 1.60), objectives=None,
                 log=print):
    """Train the surrogate on `data` (transmission [n, 32] of the first n seeded designs), then run one
    screened and one unscreened design campaign per objective with the FDTD solve as the arbiter.
    The surrogate is trained once; each campaign starts from that network."""
    import copy
    import time
    from .workloads import slab_densities
    device = torch.device('cuda')
    n = data.shape[0]
    densities = torch.as_tensor(slab_densities(n, seed), device=device)
    targets = torch.as_tensor(data, device=device, dtype=torch.float32)
    held = max(1, n//10)
    mean, std = targets[held:].mean(0), targets[held:].std(0)
    surrogate = Surrogate(SurrogateNet(seed=seed).to(device), mean, std)
    started = time.perf_counter()
    history = surrogate.fit(densities[held:], targets[held:], epochs=epochs, batch=batch, lr=lr, weight_decay=1e-3,
                            seed=seed, cosine=True, mirror=True, log=lambda e, v: log(f'train: epoch {e} mse {v:.5f}'))
    train_seconds = time.perf_counter()-started
    predicted = surrogate(densities[:held])
    error = float(((predicted-targets[:held])**2).mean().sqrt())
    wavelengths = slab_wavelengths(frequencies)
    reference = slab_flux(np.zeros((1, *densities.shape[1:]), np.float32), steps=steps, frequencies=frequencies)[0]

    def solve(pool):
        return torch.as_tensor(slab_flux(pool, steps=steps, frequencies=frequencies)/reference, device=device,
                               dtype=torch.float32)
    gen = torch.Generator(device='cpu').manual_seed(seed)
    sample = torch.randperm(n-held, generator=gen)[:1024].to(device)+held
    campaigns = {}
    for name, fom in objectives or study_objectives(data, held, wavelengths, band):
        truth, guess = fom(targets[:held]).numpy(), fom(predicted).numpy()
        r2 = float(1-np.mean((guess-truth)**2)/np.var(truth))
        log(f'campaign {name}: held-out FOM R^2 {r2:.3f}')
        local = Surrogate(copy.deepcopy(surrogate.net), mean, std)
        screened, errors, _ = screen(local, solve, fom, iterations=iterations, candidates=candidates, verify=verify,
                                     seed=seed+1, replay=(densities[sample], targets[sample]), refine_epochs=refine_epochs,
                                     log=log)
        plain, _ = unscreened(solve, fom, iterations=baseline_iterations, per_iteration=verify, seed=seed+1)
        target = screened[-1][1]
        reached = next((solves for solves, value in plain if value >= target), None)
        checkpoints = [(solves, value, next((s for s, v in plain if v >= value), None)) for solves, value in screened]
        campaigns[name] = dict(checkpoints=checkpoints,heldout_fom_r2=r2, heldout_fom_std=float(truth.std()), screened=screened, unscreened=plain,
                               screened_solves=screened[-1][0], screened_best=target, unscreened_solves_to_match=reached,
                               unscreened_best=plain[-1][1],
                               solves_ratio=(reached/screened[-1][0]) if reached else None,
                               solves_ratio_at_least=None if reached else plain[-1][0]/screened[-1][0],
                               surrogate_fom_rmse=float(np.sqrt(np.mean(errors**2))))
        log(f'campaign {name}: screened best {target:.4f} in {screened[-1][0]} solves; unscreened matched at '
            f'{reached} (best {plain[-1][1]:.4f} in {plain[-1][0]})')
    return dict(samples=n, training_solves=n, parameters=surrogate.net.parameter_count(), epochs=epochs,
                train_seconds=train_seconds, final_train_mse=history[-1], heldout_rmse=error,
                refine_epochs=refine_epochs, campaigns=campaigns)


def _main(argv=None):
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser(prog='python -m torchfdtd_amd.surrogate')
    sub = ap.add_subparsers(dest='command', required=True)
    g = sub.add_parser('generate', help='FDTD transmission of seeded designs (batched cohorts)')
    g.add_argument('--start', type=int, default=0)
    g.add_argument('--count', type=int, default=1024)
    g.add_argument('--steps', type=int, default=2000)
    g.add_argument('--out', required=True)
    s = sub.add_parser('study', help='train on generated data, then the screened and unscreened design loops')
    s.add_argument('--data', nargs='+', required=True)
    s.add_argument('--epochs', type=int, default=150)
    s.add_argument('--iterations', type=int, default=6)
    s.add_argument('--baseline-iterations', type=int, default=48)
    s.add_argument('--refine-epochs', type=int, default=4)
    s.add_argument('--steps', type=int, default=2000)
    s.add_argument('--out', required=True)
    args = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    if args.command == 'generate':
        t, reference = generate(args.start, args.count, steps=args.steps, log=log)
        np.savez(args.out, transmission=t.astype(np.float32), reference_flux=reference, start=args.start, steps=args.steps)
        log(f'generate: wrote {args.out}')
        return 0
    parts = sorted((np.load(p) for p in args.data), key=lambda z: int(z['start']))
    data = np.concatenate([z['transmission'] for z in parts])
    result = design_study(data, epochs=args.epochs, iterations=args.iterations, baseline_iterations=args.baseline_iterations,
                          refine_epochs=args.refine_epochs, steps=args.steps, log=log)
    with open(args.out, 'w') as f:
        json.dump(result, f, indent=1, default=float)
    summary = {k: v for k, v in result.items() if k != 'campaigns'}
    summary['campaigns'] = {name: {k: v for k, v in c.items() if k not in ('screened', 'unscreened')}
                            for name, c in result['campaigns'].items()}
    print(json.dumps(summary, default=float))
    return 0


if __name__ == '__main__':
    raise SystemExit(_main())
