"""Learned column correction for the coarse (T31 L8) dry primitive-equation model.

The coarse model drifts from the truncated fine solution because it misses what the
unresolved scales do to the resolved ones (and damps its smallest resolved scales harder).
Training pairs come from fine (T127 L8) runs of the same model.  At a fine state X(t), the
target is the difference of the two runs' tendencies over a window W (one hour):

    [trunc(X_fine(t + W)) - trunc(X_fine(t))] / W  -  [X_coarse(t + W) - trunc(X(t))] / W

where X_coarse starts from trunc(X(t)), i.e. the spectrally truncated fine-run tendency
minus the coarse tendency, on the coarse grid per column as tendencies of (u, v, T) on the
8 levels and of ln ps.  The input is the coarse column at t (u, v, T on the 8 levels, ln ps
repeated on each level).  A 32 -> 128 -> 128 -> 32 GELU MLP on the engine's kernels learns
it; the plain spectral tendency stays as the fallback and the yardstick: scores compare the
coarse model with and without the correction against the truncated fine run, on runs not
used for training.

Initial states: zonal jets at +-30 deg with a 35 m/s peak and N(0, 1 K) temperature noise
at the coarse resolution (degrees up to the coarse truncation, 1 K root mean square), one
seed per run; the finer scales then come from the fine model's own dynamics.
"""
from __future__ import annotations

import math
import time

import numpy as np
import torch

from .dycore import DryPrimitiveModel, truncate
from .mlp import ColumnMLP

NL = 8  # levels


def _grid_tendencies(model: DryPrimitiveModel, tend: dict) -> torch.Tensor:
    """Grid tendencies (E, 32, npoints): u, v, T per level, then ln ps repeated per level."""
    E, L = model.E, model.L
    U, V = model.uv(tend["vor"], tend["div"])
    g = model.synth(torch.cat([U, V]), unscale=True)
    T = model.synth(torch.cat([tend["temp"], tend["lnps"]]))
    ut, vt = g[:E * L].view(E, L, -1), g[E * L:].view(E, L, -1)
    Tt, qt = T[:E * L].view(E, L, -1), T[E * L:].view(E, 1, -1).expand(E, L, -1)
    return torch.cat([ut, vt, Tt, qt], dim=1)


def _grid_state(grid: dict) -> torch.Tensor:
    u, v, T, q = grid["u"], grid["v"], grid["T"], grid["lnps"]
    E, L = u.shape[0], u.shape[1]
    return torch.cat([u, v, T, q.expand(E, L, -1)], dim=1)


def _columns(x: torch.Tensor) -> torch.Tensor:
    """(E, 32, npoints) -> (E * npoints, 32) rows."""
    return x.permute(0, 2, 1).reshape(-1, x.shape[1]).contiguous()


def _truncated(fine: DryPrimitiveModel, coarse: DryPrimitiveModel) -> dict:
    return {k: truncate(v, fine.spectrum, coarse.spectrum) for k, v in fine.state.items()}


def generate(seeds, fine_trunc=127, coarse_trunc=31, spinup_hours=48.0, sample_every_hours=6.0,
             nsamples=2, window_hours=1.0, members=8, device="cuda:0", log=None):
    """Runs the fine model from each seed and returns (inputs, targets) as float32 column
    tensors (rows = coarse grid columns of every sample)."""
    fine = DryPrimitiveModel(fine_trunc, NL, nmembers=members, device=device)
    coarse = DryPrimitiveModel(coarse_trunc, NL, nmembers=members, device=device)
    xs, ys = [], []
    W = window_hours * 3600.0
    spin = int(round(spinup_hours * 3600 / fine.dt))
    wf, wc = int(round(W / fine.dt)), int(round(W / coarse.dt))
    gap = max(int(round(sample_every_hours * 3600 / fine.dt)) - wf, 0)
    for b in range(0, len(seeds), members):
        chunk = list(seeds[b:b + members])
        n_real = len(chunk)
        while len(chunk) < members:  # pad the last group with repeats (dropped below)
            chunk.append(chunk[-1])
        fine.set_state(fine.jet_state(chunk, noise_trunc=coarse_trunc))
        fine.run(spin)
        for s in range(nsamples):
            if s:
                fine.run(gap)
            st0 = _truncated(fine, coarse)
            _, cgrid = coarse.tendencies(st0, with_grid=True)
            fine.run(wf)
            st1 = _truncated(fine, coarse)
            coarse.set_state(st0)
            coarse.run(wc)
            diff = {k: (st1[k] - coarse.state[k]) / W for k in st0}
            xs.append(_columns(_grid_state(cgrid)[:n_real]))
            ys.append(_columns(_grid_tendencies(coarse, diff)[:n_real]))
        if log:
            log(f"generated seeds {b}..{b + n_real - 1}")
    fine.close()
    coarse.close()
    return torch.cat(xs), torch.cat(ys)


def forecast_score(net, nx, ny, seeds, spinup_hours=48.0, window_hours=1.0, members=8, device="cuda:0",
                   variants=None):
    """Error of the coarse model's forecast over the window from trunc(X(t)) against
    trunc(X_fine(t + W)), with the correction relative to without it (mean squared
    spectral coefficient error of vor, div, temp and ln ps, each normalised by the plain
    run's error).  variants: {name: corrector keyword arguments} scored from the same
    states (returns {name: ratios}); default one corrector with its default arguments."""
    fine = DryPrimitiveModel(127, NL, nmembers=members, device=device)
    coarse = DryPrimitiveModel(31, NL, nmembers=members, device=device)
    W = window_hours * 3600.0
    seeds = list(seeds)[:members]
    while len(seeds) < members:
        seeds.append(seeds[-1])
    fine.set_state(fine.jet_state(seeds, noise_trunc=31))
    fine.run(int(round(spinup_hours * 3600 / fine.dt)))
    st0 = _truncated(fine, coarse)
    fine.run(int(round(W / fine.dt)))
    ref = _truncated(fine, coarse)
    err = {}
    runs = {"plain": None}
    for name, kw in (variants or {"corrected": {}}).items():
        runs[name] = corrector(net, nx, ny, **kw)
    for tag, corr in runs.items():
        coarse.set_state(st0)
        coarse.run(int(round(W / coarse.dt)), correction=corr)
        err[tag] = {k: float((coarse.state[k] - ref[k]).abs().pow(2).mean()) for k in ref}
    fine.close()
    coarse.close()
    ratios = {t: {k: err[t][k] / err["plain"][k] for k in err["plain"]} for t in runs if t != "plain"}
    return ratios if variants else ratios["corrected"]


# How the correction couples into the model (corrector keyword arguments): all of its
# divergence, none of it, or only the divergence's departure from the column mean (the column
# mean drives the surface-pressure tendency through continuity), with or without its own
# ln ps tendency, with the temperature correction's column mean or without it, or without
# the temperature correction.
CORRECTOR_VARIANTS = {
    "uvT": {"fields": ("vor", "div", "temp")},
    "vorT": {"fields": ("vor", "temp")},
    "uvT_baroclinic_div": {"fields": ("vor", "div", "temp"), "baroclinic_div": True},
    "uvTq_baroclinic_div": {"fields": ("vor", "div", "temp", "lnps"), "baroclinic_div": True},
    "uvT_baroclinic": {"fields": ("vor", "div", "temp"), "baroclinic_div": True, "baroclinic_temp": True},
    "uv_baroclinic_div": {"fields": ("vor", "div"), "baroclinic_div": True},
}


def choose_corrector(net, nx, ny, seeds, **kw):
    """Scores every CORRECTOR_VARIANTS entry on the forecast from `seeds` (validation runs,
    neither training nor held-out) and returns (name, kwargs, scores): the lowest mean
    ratio among the variants whose ln ps ratio is at most 1, else the lowest ln ps ratio."""
    scores = forecast_score(net, nx, ny, seeds, variants=CORRECTOR_VARIANTS, **kw)
    ok = [n for n, r in scores.items() if r["lnps"] <= 1.0]
    if ok:
        name = min(ok, key=lambda n: float(np.mean(list(scores[n].values()))))
    else:
        name = min(scores, key=lambda n: scores[n]["lnps"])
    return name, CORRECTOR_VARIANTS[name], scores


class Normaliser:
    def __init__(self, x: torch.Tensor):
        self.mean = x.mean(0)
        self.std = x.std(0).clamp_min(1e-30)

    def __call__(self, x):
        return ((x - self.mean) / self.std).contiguous()

    def inverse(self, y):
        return y * self.std + self.mean


def init_params(seed: int) -> np.ndarray:
    """He-normal weights, zero biases, in the MLP's parameter layout."""
    rng = np.random.default_rng(seed)
    parts = []
    for shape in ((128, 32), (128,), (128, 128), (128,), (32, 128), (32,)):
        if len(shape) == 2:
            parts.append(rng.standard_normal(shape).ravel() * math.sqrt(2.0 / shape[1]))
        else:
            parts.append(np.zeros(shape))
    return np.concatenate(parts).astype(np.float32)


def fit(x: torch.Tensor, y: torch.Tensor, steps=2000, batch=4096, lr=1e-3, seed=0, log=None):
    """Adam on the engine's MLP kernels; returns (net, input normaliser, target normaliser,
    losses)."""
    nx, ny = Normaliser(x), Normaliser(y)
    xn, yn = nx(x), ny(y)
    net = ColumnMLP(init_params(seed), max_batch=batch, device=x.device)
    g = torch.Generator(device="cpu").manual_seed(seed)
    losses = []
    n = xn.shape[0]
    for it in range(steps):
        idx = torch.randint(0, n, (batch,), generator=g).to(x.device)
        l = net.train_step(xn[idx].contiguous(), yn[idx].contiguous(), lr * (0.5 * (1 + math.cos(math.pi * it / steps))))
        if it % 100 == 0 or it == steps - 1:
            losses.append(float(l.item()))
            if log:
                log(f"step {it} loss {losses[-1]:.4f}")
    return net, nx, ny, losses


def predict(net: ColumnMLP, nx: Normaliser, ny: Normaliser, x: torch.Tensor, chunk=65536) -> torch.Tensor:
    out = []
    for i in range(0, x.shape[0], chunk):
        xi = nx(x[i:i + chunk])
        out.append(ny.inverse(net(xi)))
    return torch.cat(out)


def ab_score(net, nx, ny, x, y):
    """Per channel group (u, v, T, ln ps): MSE of (target - prediction) over MSE of the
    target, i.e. the corrected tendency's error relative to the plain tendency's."""
    p = predict(net, nx, ny, x)
    r = y - p
    out = {}
    for name, sl in (("u", slice(0, 8)), ("v", slice(8, 16)), ("T", slice(16, 24)), ("lnps", slice(24, 32))):
        out[name] = float((r[:, sl] ** 2).mean() / (y[:, sl] ** 2).mean())
    out["all_normalised"] = float(((ny(y) - ny(p)) ** 2).mean() / (ny(y) ** 2).mean())
    return out


def corrector(net, nx, ny, fields=("vor", "div", "temp"), baroclinic_div=False, baroclinic_temp=False):
    """A `correction` callback for DryPrimitiveModel.step: adds the MLP's column tendencies
    (u, v -> vor, div via curl / divergence of (u, v) / cos(lat); T; ln ps) to the
    spectral tendencies of `fields`.  ln ps is left to the plain tendency by default.
    baroclinic_div: the divergence correction minus its mean over the (equally spaced)
    sigma levels, so that it does not change the surface-pressure tendency.
    baroclinic_temp: the same for the temperature correction, so that it does not change
    the column's mean temperature (whose geopotential drives the column-mean divergence)."""
    def apply(model, tend, grid):
        E, L, npts = model.E, model.L, model.grid.npoints
        x = _columns(_grid_state(grid))
        y = predict(net, nx, ny, x).view(E, npts, 4 * L).permute(0, 2, 1)
        ut, vt, Tt = y[:, :L], y[:, L:2 * L], y[:, 2 * L:3 * L]
        qt = y[:, 3 * L:].mean(dim=1, keepdim=True)
        ci = model.coslat_inv
        s = model.analyse(torch.cat([(ut * ci).reshape(E * L, -1), (vt * ci).reshape(E * L, -1),
                                     Tt.reshape(E * L, -1), qt.reshape(E, -1)]))
        EL = E * L
        us, vs = s[:EL], s[EL:2 * EL]
        keep = model.ops.keep
        if "vor" in fields:
            tend["vor"] += model.ops.curl(us, vs) / model.a
        if "div" in fields:
            dv = model.ops.divergence(us, vs) / model.a
            if baroclinic_div:
                dv = dv.view(E, L, -1)
                dv = (dv - dv.mean(dim=1, keepdim=True)).view(EL, -1)
            tend["div"] += dv
        if "temp" in fields:
            tt = s[2 * EL:3 * EL] * keep
            if baroclinic_temp:
                tt = tt.view(E, L, -1)
                tt = (tt - tt.mean(dim=1, keepdim=True)).view(EL, -1)
            tend["temp"] += tt
        if "lnps" in fields:
            tq = s[3 * EL:] * keep
            tq[:, 0] = 0
            tend["lnps"] += tq
    return apply
