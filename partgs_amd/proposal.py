"""Learned block proposal: a point-set MLP that predicts superquadric block
parameters from a scene's surface points, trained and run on the MI355X
matrix cores (bf16 operands, FP32 accumulation) through include/set_proposal.h.

Network: per point 3 -> 64 -> 128 -> 256 with ReLU, max-pool over the
points, then 256 -> 128 (ReLU) -> 16 x 12 (sq_s(3), sq_r(4), sq_t(3),
sq_eps(2) per block, the layout of partgs_amd.blocks.pack_params).

Forward: one fused kernel for the shared MLP and the pool (it never writes the
per-point activations), then two GEMMs for the head.  Backward: the pool
routes each channel's gradient to its argmax point only, so the shared MLP is
re-run and differentiated on the <= 256 argmax points of each scene; every
product is a bf16 MFMA GEMM.  Parameters are FP32 masters with bf16 copies,
updated by the Adam kernel.

There is no counterpart in the reference (which initialises blocks by
k-means, partgs_amd.init); the plain initialisation stays the fallback.
"""
from __future__ import annotations

import ctypes as C
import math

import torch

from . import _lib

H1, H2, H3, H4 = 64, 128, 256, 128
N_BLOCKS, N_PARAMS = 16, 12
OUT = N_BLOCKS * N_PARAMS

_bound = False


def _lib_sp():
    global _bound
    lib = _lib.load()
    if not _bound:
        vp, i, f, l_ = C.c_void_p, C.c_int, C.c_float, C.c_long
        lib.sp_gemm_bf16.argtypes = [i, i, i, vp, l_, l_, vp, l_, l_, vp, f, i, i, vp, l_, vp, vp]
        lib.sp_gemm_bf16.restype = i
        lib.sp_point_pool.argtypes = [i, i, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
        lib.sp_point_pool.restype = i
        lib.sp_to_bf16.argtypes = [vp, vp, l_, vp]
        lib.sp_to_bf16.restype = i
        lib.sp_adam.argtypes = [vp, vp, vp, vp, vp, l_, f, f, f, f, i, vp]
        lib.sp_adam.restype = i
        _bound = True
    return lib


def _chk(rc, what):
    if rc != 0:
        raise RuntimeError(f"set proposal: {what} failed ({rc})")


def _stream(t):
    return _lib.stream_handle(t.device)


def to_bf16(x: torch.Tensor) -> torch.Tensor:
    """bf16 bit patterns (int16 tensor) of an FP32 tensor, round to nearest even."""
    x = x.contiguous().float()
    y = torch.empty(x.shape, dtype=torch.int16, device=x.device)
    _chk(_lib_sp().sp_to_bf16(x.data_ptr(), y.data_ptr(), x.numel(), _stream(x)), "to_bf16")
    return y


def gemm(A, B, *, M, N, K, a_strides, b_strides, bias=None, relu=False, alpha=1.0, out=None, accumulate=False,
         want_bf16=False):
    """C (M x N, FP32) = act(alpha * A(m,k) B(n,k) + bias); A, B int16 bf16 tensors."""
    dev = A.device
    C_ = out if out is not None else torch.empty((M, N), dtype=torch.float32, device=dev)
    Cb = torch.empty((M, N), dtype=torch.int16, device=dev) if want_bf16 else None
    rc = _lib_sp().sp_gemm_bf16(M, N, K, A.data_ptr(), a_strides[0], a_strides[1], B.data_ptr(), b_strides[0],
                                b_strides[1], bias.data_ptr() if bias is not None else None, float(alpha),
                                int(relu), int(accumulate), C_.data_ptr(), N,
                                Cb.data_ptr() if Cb is not None else None, _stream(A))
    _chk(rc, "gemm")
    return (C_, Cb) if want_bf16 else C_


def linear(X_bf, W_bf, bias=None, relu=False, want_bf16=False):
    """X (R x K) . W^T (W is N x K), both bf16, contiguous."""
    R, K = X_bf.shape
    N = W_bf.shape[0]
    return gemm(X_bf, W_bf, M=R, N=N, K=K, a_strides=(K, 1), b_strides=(K, 1), bias=bias, relu=relu,
                want_bf16=want_bf16)


def grad_weight(dY_bf, X_bf):
    """dW (N x K) = dY^T X for dY (R x N), X (R x K), both bf16."""
    R, N = dY_bf.shape
    K = X_bf.shape[1]
    return gemm(dY_bf, X_bf, M=N, N=K, K=R, a_strides=(1, N), b_strides=(1, K))


def grad_input(dY_bf, W_bf):
    """dX (R x K) = dY W for dY (R x N), W (N x K), both bf16."""
    R, N = dY_bf.shape
    K = W_bf.shape[1]
    return gemm(dY_bf, W_bf, M=R, N=K, K=N, a_strides=(N, 1), b_strides=(1, K))


PARAM_SHAPES = {"W1": (H1, 3), "b1": (H1,), "W2": (H2, H1), "b2": (H2,), "W3": (H3, H2), "b3": (H3,),
                "W4": (H4, H3), "b4": (H4,), "W5": (OUT, H4), "b5": (OUT,)}
BF16_PARAMS = ("W2", "W3", "W4", "W5")


class ProposalNet:
    """FP32 master parameters with bf16 copies and Adam state, on one device."""

    def __init__(self, seed: int = 0, device="cuda"):
        self.device = torch.device(device)
        g = torch.Generator(device="cpu").manual_seed(seed)
        self.p = {}
        for name, shape in PARAM_SHAPES.items():
            if name.startswith("W"):
                fan_in = shape[1]
                bound = math.sqrt(6.0 / fan_in)  # He-uniform for ReLU layers
                if name == "W5":
                    bound = math.sqrt(1.0 / fan_in)
                t = (torch.rand(shape, generator=g) * 2 - 1) * bound
            else:
                t = torch.zeros(shape)
            self.p[name] = t.to(self.device).contiguous()
        self.m = {k: torch.zeros_like(v) for k, v in self.p.items()}
        self.v = {k: torch.zeros_like(v) for k, v in self.p.items()}
        self.t = 0
        self.refresh_bf16()

    def refresh_bf16(self):
        self.pb = {k: to_bf16(self.p[k]) for k in BF16_PARAMS}

    def state_dict(self):
        return {k: v.detach().cpu().clone() for k, v in self.p.items()}

    def load_state_dict(self, sd):
        for k in PARAM_SHAPES:
            self.p[k] = sd[k].to(self.device, torch.float32).contiguous()
        self.refresh_bf16()

    # -- forward ----------------------------------------------------------------------
    def pool(self, pts: torch.Tensor):
        """(values B x 256, argmax B x 256) of the shared MLP max-pooled over each scene."""
        pts = pts.contiguous().float()
        B, N, _ = pts.shape
        val = torch.empty((B, H3), dtype=torch.float32, device=pts.device)
        idx = torch.empty((B, H3), dtype=torch.int32, device=pts.device)
        scratch = torch.empty((B, H3), dtype=torch.int64, device=pts.device)
        p, pb = self.p, self.pb
        rc = _lib_sp().sp_point_pool(B, N, pts.data_ptr(), p["W1"].data_ptr(), p["b1"].data_ptr(),
                                     pb["W2"].data_ptr(), p["b2"].data_ptr(), pb["W3"].data_ptr(),
                                     p["b3"].data_ptr(), val.data_ptr(), idx.data_ptr(), scratch.data_ptr(),
                                     _stream(pts))
        _chk(rc, "point_pool")
        return val, idx

    def forward(self, pts: torch.Tensor):
        """Block parameters B x 16 x 12 and the cache the backward needs."""
        g, idx = self.pool(pts)
        g_bf = to_bf16(g)
        z1, z1_bf = linear(g_bf, self.pb["W4"], self.p["b4"], relu=True, want_bf16=True)
        out = linear(z1_bf, self.pb["W5"], self.p["b5"])
        return out.view(-1, N_BLOCKS, N_PARAMS), dict(pts=pts, g=g, g_bf=g_bf, idx=idx, z1=z1, z1_bf=z1_bf)

    # -- backward ---------------------------------------------------------------------
    def backward(self, dout: torch.Tensor, cache: dict) -> dict:
        """Gradients of every parameter for dL/d(out) (B x 16 x 12)."""
        p, pb = self.p, self.pb
        B = dout.shape[0]
        dout = dout.reshape(B, OUT).contiguous().float()
        grads = {}
        d_bf = to_bf16(dout)
        grads["W5"] = grad_weight(d_bf, cache["z1_bf"])
        grads["b5"] = dout.sum(0)
        dz1 = grad_input(d_bf, pb["W5"]) * (cache["z1"] > 0)
        dz1_bf = to_bf16(dz1)
        grads["W4"] = grad_weight(dz1_bf, cache["g_bf"])
        grads["b4"] = dz1.sum(0)
        dg = grad_input(dz1_bf, pb["W4"]) * (cache["g"] > 0)  # B x 256
        # the pool sends channel c's gradient to its argmax point: one row per (scene, channel)
        pts = cache["pts"]
        N = pts.shape[1]
        rows = (torch.arange(B, device=pts.device)[:, None] * N + cache["idx"].long()).reshape(-1)
        X = pts.reshape(-1, 3)[rows].contiguous()  # R x 3
        R = X.shape[0]
        # layer 1 in FP32 with the fused kernel's operation order
        h1 = ((X[:, 0:1] * p["W1"][:, 0] + X[:, 1:2] * p["W1"][:, 1]) + X[:, 2:3] * p["W1"][:, 2]) + p["b1"]
        h1 = torch.relu(h1)
        h1_bf = to_bf16(h1)
        h2, h2_bf = linear(h1_bf, pb["W2"], p["b2"], relu=True, want_bf16=True)
        d3 = torch.zeros((R, H3), dtype=torch.float32, device=pts.device)
        d3[torch.arange(R, device=pts.device), torch.arange(H3, device=pts.device).repeat(B)] = dg.reshape(-1)
        d3_bf = to_bf16(d3)
        grads["W3"] = grad_weight(d3_bf, h2_bf)
        grads["b3"] = dg.sum(0)
        dh2 = grad_input(d3_bf, pb["W3"]) * (h2 > 0)
        dh2_bf = to_bf16(dh2)
        grads["W2"] = grad_weight(dh2_bf, h1_bf)
        grads["b2"] = dh2.sum(0)
        dh1 = grad_input(dh2_bf, pb["W2"]) * (h1 > 0)
        grads["W1"] = dh1.t() @ X  # 64 x 3: K = 3 is below one MFMA step; FP32 on the VALU
        grads["b1"] = dh1.sum(0)
        return grads

    def adam_step(self, grads: dict, lr: float = 1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.t += 1
        lib = _lib_sp()
        for k, prm in self.p.items():
            g = grads[k].contiguous().float()
            pbk = self.pb.get(k)
            rc = lib.sp_adam(prm.data_ptr(), pbk.data_ptr() if pbk is not None else None, g.data_ptr(),
                             self.m[k].data_ptr(), self.v[k].data_ptr(), prm.numel(), lr, beta1, beta2, eps,
                             self.t, _stream(prm))
            _chk(rc, "adam")
