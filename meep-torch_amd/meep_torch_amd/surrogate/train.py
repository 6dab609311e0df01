"""Dataset generation and training of the spectral operator on the GPU.

Every matrix product of training runs in the project's BF16 matrix-core
kernel and the update in its Adam kernel (SpectralOperator.forward /
backward / adam_step); the loss gradient itself is a two-line elementwise
expression on the output spectra."""
import math
import time

import torch

from .fno import SpectralOperator
from .problem import sample_designs


def generate_dataset(problem, count, seed=0, batch=64, log=None):
    """Seeded designs -> (rho (count, d, d) projected densities, T (count,
    nfreq) FDTD transmission), both on the device."""
    t0 = time.perf_counter()
    theta = sample_designs(count, seed, problem.d, problem.device)
    with torch.no_grad():
        rho = problem.project(theta)
    T = problem.transmission(rho, batch=batch, log=log)
    torch.cuda.synchronize()
    if log:
        log(f"[surrogate] dataset: {count} FDTD solves in {time.perf_counter() - t0:.1f} s")
    return rho, T


def spectrum_error(pred, T):
    """Relative L2 error per sample."""
    return (pred - T).norm(dim=-1) / T.norm(dim=-1).clamp_min(1e-12)


def mse_step(net, rho, T, lr, wd=0.0, clip=1.0):
    """One Adam step on the mean squared spectrum error; the gradient is
    rescaled to global norm <= clip first."""
    pred = net.forward(rho)
    diff = pred - T
    dout = diff * (2.0 / diff.numel())
    net.backward(dout, param_grads=True, input_grad=False)
    if clip:
        gn = net.grads.norm()
        net.grads.mul_(torch.clamp(clip / (gn + 1e-12), max=1.0))
    net.adam_step(lr, wd=wd)
    return float((diff * diff).mean())


def flip_half(rho, gen):
    """Mirror a random half of the batch across the guide axis (y): the
    transmission problem is symmetric under y -> n-1-y (guides, source and
    monitor are centred), so T is unchanged and the batch gains variety."""
    m = torch.rand(rho.shape[0], generator=gen, device=rho.device) < 0.5
    return torch.where(m.view(-1, 1, 1), rho.flip(2), rho)


def train(net, rho, T, epochs=20, batch=64, lr=1e-3, lr_min=3e-5, val=None, seed=0, log=None, wd=1e-5,
          augment=True):
    """Adam on mean squared spectrum error, cosine-decayed learning rate, L2
    weight decay wd, y-mirror augmentation.  Returns the per-epoch history."""
    N = rho.shape[0]
    g = torch.Generator(device=rho.device).manual_seed(int(seed))
    steps = epochs * (N // batch)
    hist, k = [], 0
    for ep in range(epochs):
        perm = torch.randperm(N, generator=g, device=rho.device)
        tot = 0.0
        for s in range(0, N - batch + 1, batch):
            idx = perm[s:s + batch]
            lr_k = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * k / max(1, steps)))
            xb = flip_half(rho[idx], g) if augment else rho[idx]
            tot += mse_step(net, xb, T[idx], lr_k, wd)
            k += 1
        rec = {"epoch": ep + 1, "train_mse": tot / max(1, N // batch)}
        if val is not None:
            rec.update(evaluate(net, *val))
        hist.append(rec)
        if log:
            log(f"[surrogate] epoch {ep + 1}: " + ", ".join(f"{a}={b:.4g}" for a, b in rec.items() if a != "epoch"))
    return hist


def evaluate(net, rho, T, batch=64):
    errs = []
    with torch.no_grad():
        for s in range(0, rho.shape[0], batch):
            errs.append(spectrum_error(net.forward(rho[s:s + batch]), T[s:s + batch]))
    e = torch.cat(errs)
    return {"val_err_median": float(e.median()), "val_err_p90": float(e.quantile(0.9)),
            "val_frac_below_2pct": float((e < 0.02).float().mean())}


def build_surrogate(problem, n_train=20000, n_val=512, epochs=20, seed=0, log=None, **net_kw):
    """Generate the dataset with the FDTD engine and train a fresh network."""
    rho, T = generate_dataset(problem, n_train + n_val, seed=seed, log=log)
    net = SpectralOperator(n=problem.d, n_out=len(problem.freqs), seed=seed, device=problem.device, **net_kw)
    t0 = time.perf_counter()
    hist = train(net, rho[:n_train], T[:n_train], epochs=epochs, val=(rho[n_train:], T[n_train:]), seed=seed,
                 log=log)
    torch.cuda.synchronize()
    if log:
        log(f"[surrogate] trained {net.n_params} params, {epochs} epochs in {time.perf_counter() - t0:.1f} s")
    return net, (rho, T), hist
