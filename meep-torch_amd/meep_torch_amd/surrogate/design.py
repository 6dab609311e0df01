"""Inverse-design loops: plain FDTD (forward + adjoint every iteration) and
surrogate-assisted (network forward + input gradient, FDTD verification
every `verify_every` iterations, fallback to FDTD gradients while the
verified spectrum error exceeds `tol`).

Both maximise FOM = mean_f T(f) over theta in [0, 1] with Adam (projected
back into the box after each step) from the same start, and report the
FOM of the final design measured by FDTD."""
import time

import torch

from .fno import predict
from .train import mse_step, spectrum_error


class _Adam:
    """Adam ascent on a single tensor, clamped to [0, 1]."""

    def __init__(self, x, lr, b1=0.9, b2=0.999, eps=1e-8):
        self.x, self.lr, self.b1, self.b2, self.eps = x, lr, b1, b2, eps
        self.m = torch.zeros_like(x)
        self.v = torch.zeros_like(x)
        self.t = 0

    def ascend(self, g):
        self.t += 1
        self.m.mul_(self.b1).add_(g, alpha=1 - self.b1)
        self.v.mul_(self.b2).addcmul_(g, g, value=1 - self.b2)
        mh = self.m / (1 - self.b1 ** self.t)
        vh = self.v / (1 - self.b2 ** self.t)
        self.x.add_(self.lr * mh / (vh.sqrt() + self.eps)).clamp_(0.0, 1.0)


def start_design(problem, seed=0):
    g = torch.Generator(device=problem.device).manual_seed(int(seed))
    return (0.5 + 0.05 * torch.randn(problem.d, problem.d, generator=g, device=problem.device)).clamp(0, 1)


def fdtd_loop(problem, iters=500, lr=0.02, theta0=None, log=None):
    theta = (theta0 if theta0 is not None else start_design(problem)).clone()
    opt = _Adam(theta, lr)
    hist = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for it in range(iters):
        T, fom, g = problem.solve_with_gradient(theta)
        hist.append(fom)
        opt.ascend(g)
        if log and (it + 1) % 50 == 0:
            log(f"[design/fdtd] iter {it + 1}: FOM {fom:.4f}")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    T, fom, _ = problem.solve_with_gradient(theta)
    return {"theta": theta, "final_fom": fom, "history": hist, "wall_s": wall, "iters_per_s": iters / wall}


def surrogate_loop(problem, net, iters=500, lr=0.02, theta0=None, verify_every=10, tol=0.02, data=None,
                   finetune_steps=8, finetune_batch=32, finetune_lr=3e-4, log=None):
    """Surrogate-assisted design.  Every `verify_every` iterations the FDTD
    forward + adjoint of the current design gives its true spectrum and
    gradient; that gradient drives the step, and the network's spectrum
    error there decides the mode of the following iterations: 'surrogate'
    (network forward + input gradient) if the error is <= tol, else 'fdtd'
    (an FDTD forward + adjoint every iteration) until a later verification
    passes.  Every FDTD solve of the loop is also a labelled sample: the
    network is fine-tuned on the samples gathered along the trajectory
    (mixed with replay from its training set), so it learns the region the
    design is moving through.  A verification that follows surrogate steps
    and finds the FDTD figure of merit below the last verified one rolls the
    design back to that point (and restarts the optimizer moments):
    surrogate steps never lose verified progress.  Segments driven by FDTD
    gradients are never rolled back, so the loop is at worst the plain FDTD
    loop plus the fine-tuning cost."""
    theta = (theta0 if theta0 is not None else start_design(problem)).clone()
    opt = _Adam(theta, lr)
    mode = "fdtd"
    new_rho, new_T = [], []
    gen = torch.Generator(device=problem.device).manual_seed(1234)
    stats = {"verifications": 0, "failed": 0, "fdtd_iters": 0, "surrogate_iters": 0, "errors": [],
             "rollbacks": 0}
    best_theta, best_fom = theta.clone(), None
    since = 0   # surrogate steps since the last verification
    hist = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for it in range(iters):
        verify = it % verify_every == 0
        if verify or mode == "fdtd":
            T, fom, g = problem.solve_with_gradient(theta)
            with torch.no_grad():
                rho = problem.project(theta)
            rolled = False
            if verify:
                with torch.no_grad():
                    err = float(spectrum_error(net.forward(rho[None]), T[None])[0])
                stats["verifications"] += 1
                stats["errors"].append(err)
                mode = "surrogate" if err <= tol else "fdtd"
                stats["failed"] += int(err > tol)
                if since and best_fom is not None and fom < best_fom:
                    theta.copy_(best_theta)
                    opt.m.zero_()
                    opt.v.zero_()
                    opt.t = 0
                    mode = "fdtd"
                    rolled = True
                    stats["rollbacks"] += 1
                else:
                    best_theta.copy_(theta)
                    best_fom = fom
                since = 0
            new_rho.append(rho)
            new_T.append(T)
            _finetune(net, new_rho, new_T, data, finetune_steps, finetune_batch, finetune_lr, gen)
            stats["fdtd_iters"] += 1
            hist.append(fom)
            if rolled:
                continue
        else:
            th = theta.detach().clone().requires_grad_(True)
            Ts = predict(net, problem.project(th)[None])[0]
            fom_s = Ts.mean()
            fom_s.backward()
            g = th.grad
            stats["surrogate_iters"] += 1
            since += 1
            hist.append(float(fom_s.detach()))
        opt.ascend(g)
        if log and (it + 1) % 50 == 0:
            log(f"[design/surrogate] iter {it + 1}: FOM {hist[-1]:.4f} mode {mode} "
                f"last verified err {stats['errors'][-1]:.4f}")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    T, fom, _ = problem.solve_with_gradient(theta)
    return {"theta": theta, "final_fom": fom, "history": hist, "wall_s": wall, "iters_per_s": iters / wall,
            **stats}


def _finetune(net, new_rho, new_T, data, steps, batch, lr, gen):
    """A few Adam steps on the trajectory samples mixed with replayed
    training samples (so the network keeps the training distribution)."""
    R = torch.stack(new_rho)
    T = torch.stack(new_T)
    for _ in range(steps):
        k = min(len(new_rho), batch // 2)
        pick = torch.randint(0, R.shape[0], (k,), generator=gen, device=R.device)
        rb, tb = [R[pick]], [T[pick]]
        if data is not None:
            rp = torch.randint(0, data[0].shape[0], (batch - k,), generator=gen, device=R.device)
            rb.append(data[0][rp])
            tb.append(data[1][rp])
        mse_step(net, torch.cat(rb), torch.cat(tb), lr)
