"""Training loop on the device, mirroring ``gpjax.fit`` (``gpjax/fit.py:44-180``).

``fit(model=..., objective=..., train_data=..., num_iters=..., learning_rate=...)``
maximises ``objective`` (e.g. ``objectives.conjugate_mll``) by Adam on the
unconstrained parameters: positive / non-negative parameters go through
softplus, real ones through the identity (``DEFAULT_BIJECTION``,
``gpjax/parameters.py:140-146``; ``fit.py:136-143``).  The loss is
``-objective`` as in the reference's examples.  Adam follows optax's ``adam``
with its defaults (b1 = 0.9, b2 = 0.999, eps = 1e-8, eps_root = 0; Kingma & Ba
2015): m <- b1 m + (1-b1) g;  v <- b2 v + (1-b2) g^2;
u <- u - lr * (m / (1-b1^t)) / (sqrt(v / (1-b2^t)) + eps).

All state (unconstrained values, moments, step count, loss history) lives on
the GPU.  With ``capture_graph=True`` one complete step -- bijection, forward,
the hand-written backward, Adam -- is captured once in a HIP graph (through
``torch.cuda.CUDAGraph``) after ``warmup`` eager steps and replayed for the
remaining iterations, so an optimisation step costs one graph launch instead of
one host dispatch per kernel.  Returns ``(model, history)``.
"""

from __future__ import annotations

import math

import torch

from .parameters import DEFAULT_BIJECTION, Parameter


def _collect_parameters(model) -> list[Parameter]:
    seen, out = set(), []

    def walk(obj, depth=0):
        if depth > 6 or obj is None:
            return
        if isinstance(obj, Parameter):
            if id(obj) not in seen:
                seen.add(id(obj))
                out.append(obj)
            return
        if isinstance(obj, (int, float, str, bool, torch.Tensor)):
            return
        if isinstance(obj, (list, tuple)):
            for v in obj:
                walk(v, depth + 1)
            return
        for v in getattr(obj, "__dict__", {}).values():
            walk(v, depth + 1)

    walk(model)
    return out


class _AdamState:
    def __init__(self, params: list[torch.Tensor], lr: float, b1: float, b2: float, eps: float):
        self.lr, self.b1, self.b2, self.eps = lr, b1, b2, eps
        dev = params[0].device
        self.m = [torch.zeros_like(p) for p in params]
        self.v = [torch.zeros_like(p) for p in params]
        self.t = torch.zeros((), dtype=torch.float64, device=dev)

    def update(self, params: list[torch.Tensor], grads: list[torch.Tensor]) -> None:
        self.t.add_(1.0)
        c1 = 1.0 - torch.pow(torch.full_like(self.t, self.b1), self.t)
        c2 = 1.0 - torch.pow(torch.full_like(self.t, self.b2), self.t)
        for p, g, m, v in zip(params, grads, self.m, self.v):
            m.mul_(self.b1).add_(g, alpha=1.0 - self.b1)
            v.mul_(self.b2).addcmul_(g, g, value=1.0 - self.b2)
            p.sub_(self.lr * (m / c1) / (torch.sqrt(v / c2) + self.eps))


def fit(*, model, objective, train_data, num_iters: int = 100, learning_rate: float = 0.01,
        b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8, trainable=None,
        capture_graph: bool = False, warmup: int = 3):
    """Adam on -objective(model, train_data); returns (model, history [num_iters])."""
    if num_iters < 1:
        raise ValueError("num_iters must be >= 1")
    params = trainable if trainable is not None else _collect_parameters(model)
    if not params:
        raise ValueError("model has no trainable Parameter")
    fwd = [DEFAULT_BIJECTION.get(p.tag, DEFAULT_BIJECTION["real"])[0] for p in params]
    inv = [DEFAULT_BIJECTION.get(p.tag, DEFAULT_BIJECTION["real"])[1] for p in params]
    with torch.no_grad():
        raw = [i(p.value.detach()).clone() for p, i in zip(params, inv)]
    dev = raw[0].device
    opt = _AdamState(raw, learning_rate, b1, b2, eps)
    history = torch.zeros(num_iters, dtype=torch.float64, device=dev)
    counter = torch.zeros((), dtype=torch.int64, device=dev)
    leaves = [torch.empty_like(r).requires_grad_(True) for r in raw]
    for p, leaf in zip(params, leaves):
        p.value = leaf

    def step():
        with torch.no_grad():
            for leaf, r, f in zip(leaves, raw, fwd):
                leaf.copy_(f(r))
                leaf.grad = None
        loss = -objective(model, train_data)
        loss.backward()
        with torch.no_grad():
            # chain rule through the bijection: d loss / d raw = g * f'(raw)
            grads = []
            for leaf, r, f in zip(leaves, raw, fwd):
                if f is DEFAULT_BIJECTION["real"][0]:
                    grads.append(leaf.grad)
                else:
                    grads.append(leaf.grad * torch.sigmoid(r))
            opt.update(raw, grads)
            history.index_copy_(0, counter.reshape(1), loss.detach().reshape(1))
            counter.add_(1)

    if not capture_graph or num_iters <= warmup:
        for _ in range(num_iters):
            step()
    else:
        side = torch.cuda.Stream(device=dev)
        side.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(side):
            for _ in range(warmup):
                step()
        torch.cuda.current_stream(dev).wait_stream(side)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            step()
        # the capture recorded one step without executing it
        for _ in range(num_iters - warmup):
            graph.replay()
    with torch.no_grad():
        for p, r, f in zip(params, raw, fwd):
            p.value = f(r).detach().clone().requires_grad_(True)
    return model, history
