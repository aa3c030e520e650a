"""Adam on the library's HIP kernels, for the GCN training loop.

The reference trains with torch.optim.Adam over two parameter groups
(gcn_custom.py:116-119: weight decay 5e-4 on the first layer, none on the
second).  ``Adam`` here has the same constructor, update and state
(``step``, ``exp_avg``, ``exp_avg_sq`` per parameter, the step a float32
device tensor as in a capturable torch optimizer) and runs one
``armsp_adam_f32`` call — two kernel launches — per 16 parameters that share
lr, betas and eps.  The update is torch.optim.Adam's with amsgrad and maximize
off; the counters stay on the device, so a step can be captured in a HIP graph.
"""
from __future__ import annotations

import torch

from . import _lib
from .kernels import _stream


class Adam(torch.optim.Optimizer):
    """torch.optim.Adam(params, lr, betas, eps, weight_decay) for float32 CUDA
    parameters with dense gradients.  Parameters without a gradient are skipped
    (their step counters do not advance), as torch does."""

    def __init__(self, params, lr: float = 1e-3, betas=(0.9, 0.999), eps: float = 1e-8, weight_decay: float = 0.0):
        if not 0.0 <= lr:
            raise ValueError(f"Invalid learning rate: {lr}")
        if not 0.0 <= eps:
            raise ValueError(f"Invalid epsilon value: {eps}")
        if not 0.0 <= betas[0] < 1.0 or not 0.0 <= betas[1] < 1.0:
            raise ValueError(f"Invalid beta parameters: {betas}")
        if not 0.0 <= weight_decay:
            raise ValueError(f"Invalid weight_decay value: {weight_decay}")
        super().__init__(params, dict(lr=lr, betas=tuple(betas), eps=eps, weight_decay=weight_decay,
                                      capturable=True))
        # (device, hyper-parameters, offset) -> (tensor pointers and sizes, ctypes argument array)
        self._args: dict = {}

    def _state(self, p: torch.Tensor) -> dict:
        st = self.state[p]
        if not st:
            st["step"] = torch.zeros((), dtype=torch.float32, device=p.device)
            st["exp_avg"] = torch.zeros_like(p, memory_format=torch.contiguous_format)
            st["exp_avg_sq"] = torch.zeros_like(p, memory_format=torch.contiguous_format)
        return st

    @torch.no_grad()
    def step(self, closure=None):
        loss = None
        if closure is not None:
            with torch.enable_grad():
                loss = closure()
        batches: dict = {}
        for group in self.param_groups:
            key = (float(group["lr"]), float(group["betas"][0]), float(group["betas"][1]), float(group["eps"]))
            wd = float(group["weight_decay"])
            for p in group["params"]:
                if p.grad is None:
                    continue
                g = p.grad
                if not (p.is_cuda and g.is_cuda):
                    raise RuntimeError("armsparse_amd.optim.Adam: parameters and gradients must be on the GPU "
                                       "(no CPU fallback)")
                if p.dtype != torch.float32 or g.dtype != torch.float32 or g.is_sparse:
                    raise RuntimeError("armsparse_amd.optim.Adam: float32 parameters with dense gradients only")
                if not (p.is_contiguous() and g.is_contiguous()):
                    raise RuntimeError("armsparse_amd.optim.Adam: parameters and gradients must be contiguous")
                st = self._state(p)
                batches.setdefault((p.device, key), []).append(
                    (p, g, st["exp_avg"], st["exp_avg_sq"], st["step"], wd))
        for (dev, hyper), items in batches.items():
            if dev.index != torch.cuda.current_device():
                with torch.cuda.device(dev):
                    self._launch(dev, hyper, items)
            else:
                self._launch(dev, hyper, items)
        return loss

    def _launch(self, dev, hyper, items):
        stream = _stream(dev)
        for lo in range(0, len(items), _lib.ADAM_MAX_TENSORS):
            part = items[lo:lo + _lib.ADAM_MAX_TENSORS]
            key = tuple((p.data_ptr(), g.data_ptr(), m.data_ptr(), v.data_ptr(), s.data_ptr(), p.numel(), wd)
                        for p, g, m, v, s, wd in part)
            slot = (dev, hyper, lo)
            cached = self._args.get(slot)
            if cached is None or cached[0] != key:  # the argument array only changes with the tensors
                arr = (_lib.AdamTensor * len(part))(*(_lib.AdamTensor(*k) for k in key))
                cached = self._args[slot] = (key, arr)
            _lib.call("armsp_adam_f32", cached[1], len(part), *hyper, stream)
