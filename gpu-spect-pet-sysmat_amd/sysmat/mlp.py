"""Scatter-surrogate MLP (12 -> 256 x4 -> 1, SiLU) on the project's own gfx950 kernels.

Training and inference run in csrc/mlp.hip (BF16 MFMA GEMMs with fused SiLU
epilogues, Adam, register-resident inference); this module only holds device
tensors (PyTorch-ROCm as the allocator and stream provider) and sequences the
C ABI calls.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np

from ._lib import check, load, ptr

N_IN = 12


def _stream_handle(stream=None):
    import torch
    s = torch.cuda.current_stream() if stream is None else stream
    h = int(s.cuda_stream)
    return ctypes.c_void_p(h) if h else None


def param_count() -> int:
    return int(load().sm_mlp_param_count())


def packed_bytes() -> int:
    return int(load().sm_mlp_packed_bytes())


class Mlp:
    """One network: FP32 parameters and Adam moments in device memory."""

    def __init__(self, seed: int = 0, device: str = "cuda", params=None):
        import torch
        self.L = load()
        self.device = torch.device(device)
        n = param_count()
        if params is None:
            self.params = torch.empty(n, dtype=torch.float32, device=self.device)
            check(self.L.sm_mlp_init(ptr(self.params), int(seed), _stream_handle()))
        else:
            p = torch.as_tensor(np.asarray(params, np.float32) if not torch.is_tensor(params) else params)
            self.params = p.to(self.device, torch.float32).contiguous().clone()
            assert self.params.numel() == n
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.step = 0
        self._ws = None
        self._ws_batch = 0
        self._loss = torch.zeros(1, dtype=torch.float32, device=self.device)

    def _workspace(self, batch: int):
        import torch
        if self._ws is None or self._ws_batch != batch:
            nbytes = int(self.L.sm_mlp_workspace_bytes(batch))
            check(nbytes)
            self._ws = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
            self._ws_batch = batch
        return self._ws

    @staticmethod
    def _batch_for(n: int) -> int:
        return max(64, (n + 63) // 64 * 64)

    def forward(self, x):
        """Training-path forward of x [n, 12] (device FP32) -> y [n]."""
        import torch
        x = x.contiguous()
        n = x.shape[0]
        B = self._batch_for(n)
        ws = self._workspace(B)
        y = torch.empty(n, dtype=torch.float32, device=self.device)
        check(self.L.sm_mlp_forward(ptr(self.params), ptr(x), n, B, ptr(ws), ws.numel(), ptr(y), _stream_handle()))
        return y

    def backward(self, dy):
        """Gradient of sum(dy * y) for the rows of the last forward()."""
        import torch
        n = dy.shape[0]
        B = self._batch_for(n)
        ws = self._workspace(B)
        g = torch.empty_like(self.params)
        check(self.L.sm_mlp_backward(ptr(self.params), ptr(dy.contiguous()), n, B, ptr(ws), ws.numel(), ptr(g),
                                     _stream_handle()))
        return g

    def train_step(self, x, t, w=None, lr: float = 1e-3, batch: Optional[int] = None, want_loss: bool = False):
        n = x.shape[0]
        B = batch or self._batch_for(n)
        ws = self._workspace(B)
        self.step += 1
        check(self.L.sm_mlp_train_step(ptr(self.params), ptr(self.m), ptr(self.v), ptr(x), ptr(t),
                                       None if w is None else ptr(w), n, B, float(lr), self.step, ptr(ws), ws.numel(),
                                       ptr(self._loss) if want_loss else None, _stream_handle()))
        return self._loss if want_loss else None

    def train_run(self, x, t, w=None, steps: int = 1000, lr: float = 1e-3, lr_floor: float = 0.0, seed: int = 0,
                  batch: int = 4096, want_loss: bool = False):
        """``steps`` Adam steps on batches of ``batch`` rows drawn on the device from the n
        training pairs (x [n, 12], t [n], optional weights w [n]) with a cosine learning-rate
        schedule from ``lr`` held at ``lr_floor``; returns the last batch's loss (device
        scalar) when ``want_loss``."""
        losses = train_members([self], x, t, w, steps=steps, lr=lr, lr_floor=lr_floor, seeds=[seed], batch=batch,
                               want_loss=want_loss)
        return losses[0] if want_loss else None

    def pack(self, out=None):
        import torch
        if out is None:
            out = torch.empty(packed_bytes(), dtype=torch.uint8, device=self.device)
        check(self.L.sm_mlp_pack(ptr(self.params), ptr(out), _stream_handle()))
        return out


def train_members(nets: Sequence[Mlp], x, t, w=None, steps: int = 1000, lr: float = 1e-3, lr_floor: float = 0.0,
                  seeds: Sequence[int] = (0,), batch: int = 4096, want_loss: bool = False):
    """Train the networks of an ensemble together (each on its own stream, its own batch
    seed, all at the same step count): the per-step kernels are short and latency-bound,
    so the members' chains overlap on the GPU.  Every member's result is the one it would
    reach alone.  Returns the members' last-batch losses (device scalars) when ``want_loss``."""
    L = load()
    nets = list(nets)
    k = len(nets)
    if len(seeds) != k:
        raise ValueError("one seed per member")
    step0 = nets[0].step
    if any(nt.step != step0 for nt in nets):
        raise ValueError("members must be at the same step count")
    x, t = x.contiguous(), t.contiguous()
    w = None if w is None else w.contiguous()
    wss = [nt._workspace(batch) for nt in nets]
    arr = ctypes.c_void_p * k
    params = arr(*[nt.params.data_ptr() for nt in nets])
    ms = arr(*[nt.m.data_ptr() for nt in nets])
    vs = arr(*[nt.v.data_ptr() for nt in nets])
    wsp = arr(*[ws.data_ptr() for ws in wss])
    losses = arr(*[nt._loss.data_ptr() for nt in nets]) if want_loss else None
    sd = (ctypes.c_ulonglong * k)(*[int(s) for s in seeds])
    check(L.sm_mlp_train_run(k, params, ms, vs, ptr(x), ptr(t), None if w is None else ptr(w), int(x.shape[0]),
                             int(batch), int(steps), float(lr), float(lr_floor), sd, int(step0), wsp,
                             min(ws.numel() for ws in wss), losses, _stream_handle()))
    for nt in nets:
        nt.step += steps
    return [nt._loss for nt in nets] if want_loss else None


def pack_members(members: Sequence[Mlp]):
    """Consecutive packed inference images of the members (one device buffer)."""
    import torch
    pb = packed_bytes()
    out = torch.empty(pb * len(members), dtype=torch.uint8, device=members[0].device)
    for i, m in enumerate(members):
        m.pack(out[i * pb:(i + 1) * pb])
    return out


def infer(packed, n_members: int, x):
    """Fused register-resident inference: y [n_members, n]."""
    import torch
    x = x.contiguous()
    n = x.shape[0]
    y = torch.empty((n_members, n), dtype=torch.float32, device=x.device)
    check(load().sm_mlp_infer(ptr(packed), int(n_members), ptr(x), n, ptr(y), _stream_handle()))
    return y
