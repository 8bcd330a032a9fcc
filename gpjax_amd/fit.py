"""Training loop on the device, mirroring ``gpjax.fit`` (``gpjax/fit.py:44-180``).

``fit(model=..., objective=..., train_data=..., num_iters=..., learning_rate=...)``
minimises ``objective(model, train_data)`` as given -- like the reference, which
minimises the objective it is handed (``fit.py:148-152,160``); its examples pass
``lambda p, d: -conjugate_mll(p, d)`` -- by Adam on the unconstrained
parameters: positive / non-negative parameters go through softplus, real ones
through the identity (``DEFAULT_BIJECTION``, ``gpjax/parameters.py:140-146``;
``fit.py:136-143``).  Adam follows optax's ``adam`` with its defaults (b1 = 0.9,
b2 = 0.999, eps = 1e-8, eps_root = 0; Kingma & Ba 2015).

All state (unconstrained values, moments, step count, loss history) lives on
the GPU in flat float64 buffers, and the trainable parameters' tensors are
views of one of them.  A step is: the objective's forward and its hand-written
backward, one concatenation of the gradients, and one ``gpk_adam_step`` launch
that applies the softplus chain rule, the Adam update and the bijection and
records the loss.  With ``capture_graph=True`` one complete step is captured
once in a HIP graph (through ``torch.cuda.CUDAGraph``) after ``warmup`` eager
steps and replayed for the remaining iterations, so a step costs one graph
launch instead of one host dispatch per kernel.  Returns ``(model, history)``.
"""

from __future__ import annotations

import torch

from . import _lib
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


def _uses_softplus(p: Parameter) -> bool:
    return p.tag in ("positive", "non_negative")


class DeviceAdam:
    """Flat device state of the optimiser: unconstrained values ``raw``, moments
    ``m`` and ``v``, the step count, the softplus mask and the constrained values
    ``value`` that the parameters' tensors view."""

    def __init__(self, params: list[Parameter], lr: float, b1: float, b2: float, eps: float):
        self.params, self.lr, self.b1, self.b2, self.eps = params, lr, b1, b2, eps
        dev = params[0].value.device
        if dev.type != "cuda":
            raise ValueError("fit runs on the GPU (the Adam step is a libgpk kernel): "
                             f"parameters live on {dev}")
        self.device = dev
        self.shapes = [p.value.shape for p in params]
        self.sizes = [p.value.numel() for p in params]
        self.count = sum(self.sizes)
        with torch.no_grad():
            raws = []
            for p in params:
                inv = DEFAULT_BIJECTION.get(p.tag, DEFAULT_BIJECTION["real"])[1]
                raws.append(inv(p.value.detach().to(dev, torch.float64)).reshape(-1))
            self.raw = torch.cat(raws).contiguous()
            self.mask = torch.cat([torch.full((n,), 1 if _uses_softplus(p) else 0,
                                              dtype=torch.uint8, device=dev)
                                   for p, n in zip(params, self.sizes)])
            fwd = [DEFAULT_BIJECTION.get(p.tag, DEFAULT_BIJECTION["real"])[0] for p in params]
            self.value = torch.cat([f(r) for f, r in zip(fwd, self.raw.split(self.sizes))])
        self.m = torch.zeros_like(self.raw)
        self.v = torch.zeros_like(self.raw)
        self.t = torch.zeros(1, dtype=torch.float64, device=dev)
        # the parameters' tensors: views of `value`, leaves of every step's graph
        self.leaves = [seg.view(shape).requires_grad_(True)
                       for seg, shape in zip(self.value.split(self.sizes), self.shapes)]
        for p, leaf in zip(params, self.leaves):
            p.value = leaf

    def step(self, loss: torch.Tensor, history: torch.Tensor | None) -> None:
        """Gradients of ``loss`` must be on the leaves already."""
        g = torch.cat([(leaf.grad if leaf.grad is not None else torch.zeros_like(leaf))
                       .reshape(-1) for leaf in self.leaves])
        lv = loss.detach().reshape(1)
        if lv.dtype != torch.float64:
            lv = lv.to(torch.float64)
        hist_ptr, hist_len = (0, 0) if history is None else (_lib.ptr(history), history.numel())
        _lib.call("gpk_adam_step", self.count, _lib.ptr(g), _lib.ptr(self.mask),
                  _lib.ptr(self.raw), _lib.ptr(self.m), _lib.ptr(self.v), _lib.ptr(self.t),
                  float(self.lr), float(self.b1), float(self.b2), float(self.eps),
                  _lib.ptr(self.value), _lib.ptr(lv.contiguous()), hist_ptr, hist_len,
                  _lib.stream_ptr(self.device))


def fit(*, model, objective, train_data, num_iters: int = 100, learning_rate: float = 0.01,
        b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8, trainable=None,
        capture_graph: bool = False, warmup: int = 3):
    """Adam on objective(model, train_data); returns (model, history [num_iters])."""
    if num_iters < 1:
        raise ValueError("num_iters must be >= 1")
    if capture_graph and warmup < 1:
        # The first evaluation on a thread creates libgpk's own streams and
        # events; that must not happen inside the capture.
        raise ValueError("capture_graph needs warmup >= 1 eager steps before the capture")
    params = trainable if trainable is not None else _collect_parameters(model)
    if not params:
        raise ValueError("model has no trainable Parameter")
    opt = DeviceAdam(params, learning_rate, b1, b2, eps)
    dev = opt.device
    history = torch.zeros(num_iters, dtype=torch.float64, device=dev)

    def step():
        for leaf in opt.leaves:
            leaf.grad = None
        loss = objective(model, train_data)
        loss.backward()
        with torch.no_grad():
            opt.step(loss, history)

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
        for p, leaf in zip(params, opt.leaves):
            p.value = leaf.detach().clone().requires_grad_(True)
    return model, history
