"""Gradient-tuned per-block predictor coefficients (the differentiation form).

The reference picks each block's 1D4R coefficient t by scoring five discrete
candidates with its H1-PSNR proxy (compress.tpp:90-144).  Within one round the
coarse neighbours are fixed, so every target's pre-quantization residual is
linear in the block's parameters (r = A t + B for 1D4R; r = a (S5 - S1) +
b (S3 - S1) + S1/2 - x0 for 1D6R with free weights), and the residual energy
E_b = sum r^2 is a quadratic that the GPU evaluates (asz_round_energy) and
differentiates with a hand-written adjoint kernel (asz_round_energy_adjoint).

* RoundEnergy: a torch.autograd.Function over those two entry points.
* smooth_h1_bits: the smooth part of the H1 proxy, n_b/2 log2(E_b/n_b) bits
  per block (the entropy of a Gaussian residual of that energy, up to a
  constant), differentiable through E.
* newton_t: projected Newton steps per block on [t_min, t_max] from autograd
  gradients (the secant Hessian is exact for a quadratic), snapped to the
  1/t_precision grid.
* compare_tuned_vs_plain: compression ratio, max error and t-selection time of
  the in-kernel gradient form (ASZ_TSEL_GRAD: Newton on the sampled energy,
  grid snap, exact-proxy judgement) against the discrete scan.

Everything runs on the HIP library; torch supplies device memory and autograd.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np
import torch

from . import _lib, codec


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


class RoundState:
    """The field as round `round` of a plan reads it: pivots and earlier
    rounds reconstructed (with the given t), targets still original."""

    def __init__(self, ctx: codec.Context, field, plan, params, round_, t_int, quintic=False):
        self.ctx, self.plan, self.params, self.round, self.quintic = ctx, plan, params, int(round_), bool(quintic)
        f = torch.as_tensor(np.ascontiguousarray(field, np.float32)).cuda()
        self.work = f.clone()
        self.dims = np.asarray(f.shape, np.uint64)
        mn, mx = ctx.range(self.work.data_ptr(), self.work.numel())
        ti = np.ascontiguousarray(t_int, np.int64)
        L = _lib.lib()
        ctx.after_torch()
        ctx._check(L.asz_prepare_round(ctx._c, _p(self.work), self.dims.ctypes.data_as(ctypes.c_void_p),
                                       ctypes.byref(plan), ctypes.byref(params), mn, mx, self.round,
                                       ti.ctypes.data_as(ctypes.c_void_p)))
        _, _, bn = ctx.geometry(tuple(int(x) for x in self.dims), plan, params.block_pivot_ratio)
        self.nblocks = int(np.prod(bn))

    @property
    def nparam(self):
        return 2 if self.quintic else 1

    def energy(self, theta: torch.Tensor) -> torch.Tensor:
        th = theta.detach().to(device="cuda", dtype=torch.float32).contiguous()
        E = torch.empty(self.nblocks, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        L = _lib.lib()
        self.ctx.after_torch()
        self.ctx._check(L.asz_round_energy(self.ctx._c, _p(self.work), self.dims.ctypes.data_as(ctypes.c_void_p),
                                           ctypes.byref(self.plan), ctypes.byref(self.params), self.round,
                                           int(self.quintic), _p(th), _p(E)))
        return E

    def adjoint(self, theta: torch.Tensor, gE: torch.Tensor) -> torch.Tensor:
        th = theta.detach().to(device="cuda", dtype=torch.float32).contiguous()
        ge = gE.detach().to(device="cuda", dtype=torch.float32).contiguous()
        g = torch.empty_like(th)
        torch.cuda.synchronize()
        L = _lib.lib()
        self.ctx.after_torch()
        self.ctx._check(L.asz_round_energy_adjoint(
            self.ctx._c, _p(self.work), self.dims.ctypes.data_as(ctypes.c_void_p), ctypes.byref(self.plan),
            ctypes.byref(self.params), self.round, int(self.quintic), _p(th), _p(ge), _p(g)))
        return g


class RoundEnergy(torch.autograd.Function):
    """E = RoundEnergy.apply(theta, state): per-block residual energy (float64)
    of state's round; theta is (nblocks,) t values or (nblocks, 2) 1D6R
    weights.  The backward pass is the GPU adjoint kernel."""

    @staticmethod
    def forward(ctx, theta, state: RoundState):
        ctx.state = state
        ctx.save_for_backward(theta)
        return state.energy(theta)

    @staticmethod
    def backward(ctx, gE):
        (theta,) = ctx.saved_tensors
        g = ctx.state.adjoint(theta, gE).reshape(theta.shape).to(theta.dtype)
        return g, None


def smooth_h1_bits(E: torch.Tensor, n: torch.Tensor, eb: float) -> torch.Tensor:
    """Smooth part of the H1 proxy: n/2 log2(E / (n (2 eb)^2)) bits per block."""
    return 0.5 * n * torch.log2(E / (n * (2.0 * eb) ** 2))


def newton_t(state: RoundState, t0=0.25, t_min=1 / 16, t_max=6 / 16, t_precision=32, steps=3, delta=1 / 64):
    """Projected Newton on each block's energy from autograd gradients.
    Returns (t snapped to the grid, t before snapping)."""
    t = torch.full((state.nblocks,), float(t0), dtype=torch.float32, device="cuda")
    for _ in range(steps):
        t1 = t.detach().clone().requires_grad_(True)
        RoundEnergy.apply(t1, state).sum().backward()
        g1 = t1.grad.detach()
        t2 = (t.detach() + delta).requires_grad_(True)
        RoundEnergy.apply(t2, state).sum().backward()
        hess = (t2.grad.detach() - g1) / delta  # exact for a quadratic energy
        step = torch.where(hess > 0, g1 / hess, torch.zeros_like(g1))
        t = torch.clamp(t.detach() - step, t_min, t_max)
    snapped = torch.clamp(torch.round(t * t_precision) / t_precision, t_min, t_max)
    return snapped, t


def compare_tuned_vs_plain(field: torch.Tensor, rel: float, plan, block_pivot_ratio: int, family="cubic",
                           device=0):
    """Compress `field` (a CUDA tensor) with the discrete t scan and with the
    gradient-tuned form; returns CR, max |error| / eb and the t-selection
    kernel time of each, the tuned/plain ratios, and the agreement of t.
    Block rounds are off here, so that every round's t selection is a launch
    of its own and its kernel time is measured."""
    ctx = codec.Context(device)
    ctx.set_coarse_rounds(0)
    try:
        shape = tuple(field.shape)
        n = field.numel()
        params = codec.make_params(rel_eb=float(np.float32(rel)), block_pivot_ratio=block_pivot_ratio, family=family)
        mn, mx = ctx.range(field.data_ptr(), n)
        slope, *_ = ctx.default_slope(field.data_ptr(), shape, params, mn, mx)
        params.default_slope = slope
        pl = codec.make_plan(plan)
        abs_eb = float(np.float32(rel) * (np.float32(mx) - np.float32(mn)))
        out = {}
        ctx.set_kernel_events(True)
        for mode in ("gpu", "grad"):
            w = field.clone()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            res = ctx.compress_device(w.data_ptr(), shape, pl, params, mn, mx, tsel=mode, copy_stream=True,
                                      want_t=True)
            wall = time.perf_counter() - t0
            st = res.stats
            frame = codec.zstd_compress(res.stream)
            err = float((w.double() - field.double()).abs().max())
            out[mode] = {"cr": n * 4 / len(frame), "max_err_over_eb": err / abs_eb,
                         "tsel_ms": float(sum(st.tsel_round_ms[r] for r in range(pl.layers))),
                         "compress_wall_ms": wall * 1e3, "t": res.t}
        ctx.set_kernel_events(False)
        agree = float(np.mean(out["gpu"]["t"] == out["grad"]["t"]))
        for m in out.values():
            del m["t"]
        return {"plain": out["gpu"], "tuned": out["grad"],
                "cr_tuned_over_plain": out["grad"]["cr"] / out["gpu"]["cr"],
                "tsel_time_tuned_over_plain": out["grad"]["tsel_ms"] / max(out["gpu"]["tsel_ms"], 1e-12),
                "t_agreement": agree}
    finally:
        ctx.close()


# Plans the reference s1_compress chose (auto plan generation, cubic family,
# launcher block pivot ratio) for the 512^3 half-smooth / half-rough field of
# fields.config5_field at REL 1e-3 and 1e-4.
C5_PLANS = {
    1e-3: (3, ["cubic2_012", "cubic1_012", "cubic0_012", "cubic2_012", "cubic2_012", "cubic1_012", "cubic0_012",
               "cubic1_012", "cubic1_012", "cubic2_012", "cubic0_021", "cubic0_120", "cubic1_021", "cubic2_012"]),
    1e-4: (96, ["cubic0_012", "cubic1_012", "cubic0_012", "cubic0_012", "cubic2_012", "cubic2_012", "cubic1_012",
                "cubic0_012", "cubic2_012", "cubic0_012", "cubic1_120", "cubic0_120", "cubic2_012", "cubic1_021"]),
}


def config5_comparison(rels=(1e-3, 1e-4), shape=(512, 512, 512), seed=5):
    """Tuned vs plain on the half-smooth / half-rough 512^3 field."""
    from . import fields
    field = fields.config5_field(shape, seed=seed, device="cuda").contiguous()
    out = {}
    for rel in rels:
        bpr, plan = C5_PLANS[rel]
        out[f"{rel:g}"] = compare_tuned_vs_plain(field, rel, plan, bpr)
    return out


if __name__ == "__main__":
    import json
    print(json.dumps(config5_comparison(), indent=1))
