"""Learned pivot proposal for FW-BW (no counterpart in the reference).

KaSpan seeds forward-backward search with the vertex of maximum
in-degree x out-degree (scc/trim_1_local.hpp:32-45, scc/pivot.hpp:53-74).
Here a small MLP (6 -> 32 -> 32 -> 1) scores the top degree-product
candidates and the best score seeds FW-BW; the plain rule's pivot is always
among the candidates and remains the fallback. Pivot choice changes time,
never labels.

In the multi-pivot rounds the candidates of a partition are its vertices
in the plain rule's top half-octave of in*out, and the score picks among
them. The scorer is trained on pivot outcomes the rounds themselves log
(opts.pivot_log): by default log1p of the SCC each pivot decided, with
pivots drawn at random among the same candidates; optionally the SCC a
FW-BW from every vertex decides (one solve gives all) or the partition's
shrinkage. Forward and gradient (both BF16 MFMA with FP32 accumulation)
and Adam run as libkspn kernels on the GPU; inputs and target are
standardised during training and the scaling is folded back into W1, b1, w3
and b3.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import torch

from . import BidiGraph, Context, _lib, _stream_ptr, scc, scc_sizes

PARAMS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pivot_mlp_params.npy")
N_PARAMS = _lib.MLP_PARAMS


def init_params(seed=0) -> torch.Tensor:
    """He-initialised parameter vector (W1, b1, W2, b2, w3, b3)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS, np.float32)
    p[0:192] = rng.normal(0, np.sqrt(2 / 6), 192)
    p[224:1248] = rng.normal(0, np.sqrt(2 / 32), 1024)
    p[1280:1312] = rng.normal(0, np.sqrt(1 / 32), 32)
    return torch.from_numpy(p).cuda()


def load_params(path=PARAMS_PATH) -> torch.Tensor | None:
    if not os.path.exists(path):
        return None
    return torch.from_numpy(np.load(path, allow_pickle=False).astype(np.float32)).cuda()


def features(g: BidiGraph, vertices: torch.Tensor | None = None, ctx: Context | None = None) -> torch.Tensor:
    ctx = ctx or Context.default()
    count = g.n if vertices is None else vertices.numel()
    out = torch.empty((count, 6), dtype=torch.float32, device="cuda")
    gv = g.view()
    _lib.check(_lib.lib().kspn_pivot_features(ctx.ptr, C.byref(gv), None if vertices is None else C.c_void_p(vertices.data_ptr()),
                                              count, C.c_void_p(out.data_ptr()), _stream_ptr()), "kspn_pivot_features")
    return out


def forward(params: torch.Tensor, x: torch.Tensor, ctx: Context | None = None) -> torch.Tensor:
    ctx = ctx or Context.default()
    x = x.contiguous()
    out = torch.empty(x.shape[0], dtype=torch.float32, device="cuda")
    _lib.check(_lib.lib().kspn_mlp_forward(ctx.ptr, C.c_void_p(params.data_ptr()), C.c_void_p(x.data_ptr()), x.shape[0],
                                           C.c_void_p(out.data_ptr()), _stream_ptr()), "kspn_mlp_forward")
    return out


def grad(params: torch.Tensor, x: torch.Tensor, target: torch.Tensor, ctx: Context | None = None):
    """(loss, gradient) of 0.5*mean((MLP(x) - target)^2): matrix products on MFMA with BF16 operands and
    FP32 accumulation (the sum of squared errors and the gradient come back in FP32)."""
    ctx = ctx or Context.default()
    g = torch.empty(N_PARAMS, dtype=torch.float32, device="cuda")
    loss = torch.empty(1, dtype=torch.float32, device="cuda")
    x = x.contiguous()
    target = target.contiguous()
    _lib.check(_lib.lib().kspn_mlp_grad(ctx.ptr, C.c_void_p(params.data_ptr()), C.c_void_p(x.data_ptr()),
                                        C.c_void_p(target.data_ptr()), x.shape[0], C.c_void_p(g.data_ptr()),
                                        C.c_void_p(loss.data_ptr()), _stream_ptr()), "kspn_mlp_grad")
    return loss, g


def candidates(g: BidiGraph, k_top=4096, k_random=4096, seed=0) -> torch.Tensor:
    """top-k_top vertices by in*out degree plus k_random uniform vertices with both degrees > 0"""
    outd = (g.fw_off[1:] - g.fw_off[:-1])
    ind = (g.bw_off[1:] - g.bw_off[:-1])
    prod = outd * ind
    alive = torch.nonzero(prod > 0).flatten()
    if alive.numel() == 0:
        return alive.to(torch.int32)
    top = alive[torch.topk(prod[alive], min(k_top, alive.numel())).indices]
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    rnd = alive[torch.randint(0, alive.numel(), (min(k_random, alive.numel()),), device="cuda", generator=gen)]
    return torch.unique(torch.cat([top, rnd])).to(torch.int32)


def pivot_log(g: BidiGraph, rounds=24, pivot=_lib.PIVOT_RANDOM, seed=0, cap=1 << 20, **opts) -> torch.Tensor:
    """Solve g with multi-pivot FW-BW and return the logged pivot outcomes as an int64 [records, 8] tensor:
    pivot, round, partition size, SCC size, forward-only, backward-only, unreached, 0."""
    log = torch.zeros(cap * _lib.LOG_WORDS, dtype=torch.int32, device="cuda")
    _, st = scc(g, fwbw_rounds=rounds, pivot=pivot, seed=seed, pivot_log=log, **opts)
    rec = log[: st.pivot_log_count * _lib.LOG_WORDS].view(-1, _lib.LOG_WORDS)
    return rec.to(torch.int64) & 0xFFFFFFFF


def outcome_samples(g: BidiGraph, rounds=24, seed=0, min_size=32, target="log", sample=1 << 19, gamma=0.0):
    """(features, target) pairs from pivot outcomes on g.

    target="log" (default): the multi-pivot rounds' own outcome log, with
    pivots drawn at random among each partition's candidates (the plain
    rule's top half-octave of in*out, the set the learned rule picks from);
    the target is log1p of the SCC each logged pivot decided, less
    ``gamma`` * log1p of its forward set (the SCC and what it reaches in the
    partition), so gamma > 0 prefers pivots whose search decides much of
    what it traverses.
    target="size": every live vertex is a candidate pivot whose outcome is the
    SCC a FW-BW from it decides; one solve gives all of them (kspn_scc_sizes)
    and the target is log1p(SCC size), for a seeded sample of ``sample`` live
    vertices. target="shrinkage": the log as for "log"; per partition of at
    least ``min_size`` vertices the target is 1 - (largest part) / partition size.
    """
    if target == "log":
        rec = pivot_log(g, rounds=rounds, pivot=_lib.PIVOT_RANDOM, seed=seed)
        x = features(g, rec[:, 0].to(torch.int32).contiguous())
        t = torch.log1p(rec[:, 3].float())
        if gamma:
            t = t - gamma * torch.log1p((rec[:, 3] + rec[:, 4]).float())
        return x, t
    if target == "size":
        labels, _ = scc(g)
        sizes = scc_sizes(labels)
        ind = (g.bw_off[1:] - g.bw_off[:-1])
        outd = (g.fw_off[1:] - g.fw_off[:-1])
        live = torch.nonzero((ind > 0) & (outd > 0)).flatten()
        gen = torch.Generator(device="cpu").manual_seed(seed)
        if live.numel() > sample:
            live = live[torch.randperm(live.numel(), generator=gen)[:sample].to(live.device)]
        v = live.to(torch.int32).contiguous()
        return features(g, v), torch.log1p(sizes[v.long()].float())
    rec = pivot_log(g, rounds=rounds, pivot=_lib.PIVOT_RANDOM, seed=seed)
    rec = rec[rec[:, 2] >= min_size]
    size = rec[:, 2].float()
    largest = torch.maximum(torch.maximum(rec[:, 4], rec[:, 5]), rec[:, 6]).float()
    t = 1.0 - largest / size
    x = features(g, rec[:, 0].to(torch.int32).contiguous())
    return x, t


def fold_scaling(params: torch.Tensor, mu: torch.Tensor, sd: torch.Tensor, tm: float, ts: float) -> torch.Tensor:
    """Parameters of a model trained on (x - mu) / sd -> (t - tm) / ts, as a model of x -> t."""
    p = params.clone()
    w1 = p[0:192].view(32, 6)
    b1 = p[192:224]
    b1 -= (w1 * (mu / sd)).sum(dim=1)
    w1 /= sd
    p[1280:1312] *= ts
    p[1312] = p[1312] * ts + tm
    return p


def train(x: torch.Tensor, t: torch.Tensor, params: torch.Tensor | None = None, epochs=400, lr=3e-3, seed=0,
          ctx: Context | None = None, lr_final=None):
    """Full-batch Adam on the GPU kernels with standardised inputs and target
    and a cosine learning-rate decay to ``lr_final`` (lr / 20); returns the
    folded parameters and the losses on the standardised problem."""
    ctx = ctx or Context.default()
    mu = x.mean(dim=0)
    sd = x.std(dim=0).clamp_min(1e-3)
    tm = float(t.mean())
    ts = max(float(t.std()), 1e-3)
    xn = ((x - mu) / sd).contiguous()
    tn = ((t - tm) / ts).contiguous()
    params = init_params(seed) if params is None else params.clone()
    state = torch.zeros(2 * N_PARAMS, dtype=torch.float32, device="cuda")
    lr_final = lr / 20 if lr_final is None else lr_final
    losses = []
    for step in range(1, epochs + 1):
        loss, gr = grad(params, xn, tn, ctx)
        lr_t = lr_final + 0.5 * (lr - lr_final) * (1 + np.cos(np.pi * (step - 1) / max(epochs - 1, 1)))
        _lib.check(_lib.lib().kspn_mlp_adam(ctx.ptr, C.c_void_p(params.data_ptr()), C.c_void_p(gr.data_ptr()),
                                            C.c_void_p(state.data_ptr()), step, float(lr_t), _stream_ptr()),
                   "kspn_mlp_adam")
        if step == 1 or step % 50 == 0 or step == epochs:
            losses.append(float(loss.item()) / max(x.shape[0], 1))
    return fold_scaling(params, mu, sd, tm, ts), losses
