"""Warm-start surrogate: a 3-D U-Net that predicts steady (v, rho - 1) from the
geometry, on the hand-written BF16 MFMA convolutions of conv3d.hip.

No reference counterpart: the reference always starts at rest (init,
SP:160-170).  The network's prediction initialises f at equilibrium through
init_from_macros (a warm start); a cold start stays the fallback.

Network: 4 levels of 16-32-64-128 channels, two 3x3x3 convolutions with ReLU
per level, 2x2x2 max pooling down, nearest-neighbour upsampling and channel
concatenation up, a 1x1x1 head.  Activations and weights are BF16 in the
kernels, accumulation FP32; master weights and Adam state are FP32.  Tensors
are channels-last (N, X, Y, Z, C) throughout.

Inputs (2 channels): solid mask, Euclidean distance to the nearest solid cell
/ DIST_SCALE.  Outputs (4 channels): (vx, vy, vz) / (fx / nu_eff) and
(rho - 1) / (fx * RHO_SCALE): fields in units of the body force, so one
network serves any small force.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib

CIN_PAD = 16        # input channels are zero-padded to the kernels' multiple of 16
HEAD_PAD = 16       # the 4 output channels are computed as 16 (12 zero rows)
DIST_SCALE = 8.0
RHO_SCALE = 100.0
CHANNELS = (16, 32, 64, 128)
BORDER = 16         # cells next to a volume face that the loss ignores and inference pads around


def _torch():
    import torch
    return torch


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream(dev):
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


# ---- the convolution as an autograd function ----------------------------------------------

CONV_RELU, CONV_PERIODIC = 1, 2  # lbm3d_conv3d_fwd flags


def conv_fwd(x, w_bf16, bias, cout, ks, relu, periodic=False):
    """x (N,X,Y,Z,Cin) bf16 contiguous, w_bf16 (cout, kpad) bf16 -> (N,X,Y,Z,cout) bf16.
    periodic: each volume is one period of a periodic field (taps wrap instead of
    reading zero padding)."""
    torch = _torch()
    n, d, h, wd, cin = x.shape
    y = torch.empty((n, d, h, wd, cout), dtype=torch.bfloat16, device=x.device)
    flags = (CONV_RELU if relu else 0) | (CONV_PERIODIC if periodic else 0)
    _lib.lib().call("lbm3d_conv3d_fwd", _stream(x.device), _vp(x), _vp(w_bf16),
                    _vp(bias) if bias is not None else None, _vp(y), n, d, h, wd, cin, cout, ks, flags)
    return y


def conv_dgrad(dy, w_bf16, cin, ks):
    torch = _torch()
    n, d, h, wd, cout = dy.shape
    lib = _lib.lib()
    kpad_t = lib.cdll.lbm3d_conv_kpad(cout, ks)
    wt = torch.empty((cin, kpad_t), dtype=torch.bfloat16, device=dy.device)
    s = _stream(dy.device)
    lib.call("lbm3d_conv3d_wflip", s, _vp(w_bf16), _vp(wt), cin, cout, ks)
    dx = torch.empty((n, d, h, wd, cin), dtype=torch.bfloat16, device=dy.device)
    lib.call("lbm3d_conv3d_dgrad", s, _vp(dy), _vp(wt), _vp(dx), n, d, h, wd, cin, cout, ks)
    return dx


def conv_wgrad(x, dy, ks):
    torch = _torch()
    n, d, h, wd, cin = x.shape
    cout = dy.shape[-1]
    lib = _lib.lib()
    kpad = lib.cdll.lbm3d_conv_kpad(cin, ks)
    ws_bytes = lib.size("lbm3d_conv3d_wgrad_workspace_bytes", n, d, h, wd, cin, cout, ks)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=x.device)
    dw = torch.empty((cout, kpad), dtype=torch.float32, device=x.device)
    lib.call("lbm3d_conv3d_wgrad", _stream(x.device), _vp(x), _vp(dy), _vp(dw), n, d, h, wd, cin, cout, ks,
             _vp(ws), ws_bytes)
    return dw


def relu_bias_grad(g, y=None):
    """(dz, db) for y = relu(z) with z = conv + b: dz = g * (y > 0) (or g
    without y) in BF16 and db = sum of dz over the voxels in FP32, in one pass
    (lbm3d_relu_bias_grad).  g, y: (..., C) BF16 channels-last."""
    torch = _torch()
    g = g.contiguous()
    c = g.shape[-1]
    vox = g.numel() // c
    lib = _lib.lib()
    ws_bytes = lib.size("lbm3d_relu_bias_grad_workspace_bytes", vox, c)
    ws = torch.empty(max(ws_bytes, 4), dtype=torch.uint8, device=g.device)
    dz = torch.empty_like(g)
    db = torch.empty(c, dtype=torch.float32, device=g.device)
    lib.call("lbm3d_relu_bias_grad", _stream(g.device), _vp(g), _vp(y.contiguous()) if y is not None else None,
             _vp(dz), _vp(db), vox, c, _vp(ws), ws_bytes)
    return dz, db


def _make_conv_function():
    torch = _torch()

    class Conv3dMFMA(torch.autograd.Function):
        """y = act(conv3d(x, w) + b) with the MFMA kernels in both directions."""

        @staticmethod
        def forward(ctx, x, w, b, ks, relu, need_dx, periodic):
            x = x.contiguous()
            wb = w.to(torch.bfloat16).contiguous()
            y = conv_fwd(x, wb, b, w.shape[0], ks, relu, periodic)
            ctx.save_for_backward(x, wb, y if relu else None)
            ctx.ks, ctx.relu, ctx.need_dx, ctx.periodic = ks, relu, need_dx, periodic
            return y

        @staticmethod
        def backward(ctx, gy):
            if ctx.periodic:
                raise NotImplementedError("periodic convolutions are forward-only (inference)")
            x, wb, y = ctx.saved_tensors
            g, db = relu_bias_grad(gy.contiguous().to(torch.bfloat16), y if ctx.relu else None)
            dw = conv_wgrad(x, g, ctx.ks)
            dx = conv_dgrad(g, wb, x.shape[-1], ctx.ks) if ctx.need_dx else None
            return dx, dw, db, None, None, None, None

    return Conv3dMFMA


_CONV = None


def conv3d(x, w, b, ks: int = 3, relu: bool = True, need_dx: bool = True, periodic: bool = False):
    global _CONV
    if _CONV is None:
        _CONV = _make_conv_function()
    return _CONV.apply(x, w, b, ks, relu, need_dx, periodic)


# ---- the network ------------------------------------------------------------------------------

def _pool(x):
    n, d, h, w, c = x.shape
    return x.view(n, d // 2, 2, h // 2, 2, w // 2, 2, c).amax(dim=(2, 4, 6))


def _up(x):
    n, d, h, w, c = x.shape
    return x[:, :, None, :, None, :, None, :].expand(n, d, 2, h, 2, w, 2, c).reshape(n, 2 * d, 2 * h, 2 * w, c)


class UNet3D:
    """Parameters as FP32 master tensors; forward through the MFMA convolutions."""

    def __init__(self, device, seed: int = 0, channels=CHANNELS):
        torch = _torch()
        self.device = device
        self.channels = tuple(channels)
        c1, c2, c3, c4 = self.channels
        specs = [("e1a", CIN_PAD, c1, 3), ("e1b", c1, c1, 3),
                 ("e2a", c1, c2, 3), ("e2b", c2, c2, 3),
                 ("e3a", c2, c3, 3), ("e3b", c3, c3, 3),
                 ("e4a", c3, c4, 3), ("e4b", c4, c4, 3),
                 ("d3a", c4 + c3, c3, 3), ("d3b", c3, c3, 3),
                 ("d2a", c3 + c2, c2, 3), ("d2b", c2, c2, 3),
                 ("d1a", c2 + c1, c1, 3), ("d1b", c1, c1, 3),
                 ("head", c1, HEAD_PAD, 1)]
        self.specs = {name: (cin, cout, ks) for name, cin, cout, ks in specs}
        gen = torch.Generator().manual_seed(seed)
        self.params = {}
        lib = _lib.lib()
        for name, cin, cout, ks in specs:
            kpad = lib.cdll.lbm3d_conv_kpad(cin, ks)
            k = ks ** 3 * cin
            fan_in = k if name != "e1a" else ks ** 3 * 2
            std = math.sqrt(2.0 / fan_in) if name != "head" else math.sqrt(1.0 / fan_in)
            core = torch.randn(cout, ks ** 3, cin, generator=gen) * std
            if name == "e1a":  # only the 2 real input channels carry weights
                core[:, :, 2:] = 0
            w = torch.zeros(cout, kpad)
            w[:, :k] = core.reshape(cout, k)
            if name == "head":
                w[4:] = 0
            self.params[name + ".w"] = w.to(device).requires_grad_(True)
            self.params[name + ".b"] = torch.zeros(cout, device=device).requires_grad_(True)
        # weights that must stay zero: input channels 2..15 of the first layer
        cin, cout, ks = self.specs["e1a"]
        kpad = lib.cdll.lbm3d_conv_kpad(cin, ks)
        keep = torch.zeros(cout, kpad, device=device)
        keep[:, : ks ** 3 * cin] = (torch.arange(ks ** 3 * cin, device=device) % cin < 2).float()
        self._e1a_keep = keep

    def parameters(self):
        return list(self.params.values())

    def _c(self, name, x, relu=True, need_dx=True, periodic=False):
        cin, cout, ks = self.specs[name]
        return conv3d(x, self.params[name + ".w"], self.params[name + ".b"], ks, relu, need_dx, periodic)

    def forward(self, x, periodic: bool = False):
        """x (N,X,Y,Z,16) bf16 with channels 0-1 the inputs -> (N,X,Y,Z,4) float32.
        X, Y, Z must be multiples of 8.  periodic: every volume is one period of a
        periodic geometry, and the convolutions wrap around its faces instead of
        zero padding (forward only)."""
        torch = _torch()
        p = periodic
        e1 = self._c("e1b", self._c("e1a", x, need_dx=False, periodic=p), periodic=p)
        e2 = self._c("e2b", self._c("e2a", _pool(e1), periodic=p), periodic=p)
        e3 = self._c("e3b", self._c("e3a", _pool(e2), periodic=p), periodic=p)
        e4 = self._c("e4b", self._c("e4a", _pool(e3), periodic=p), periodic=p)
        d3 = self._c("d3b", self._c("d3a", torch.cat([_up(e4), e3], dim=-1), periodic=p), periodic=p)
        d2 = self._c("d2b", self._c("d2a", torch.cat([_up(d3), e2], dim=-1), periodic=p), periodic=p)
        d1 = self._c("d1b", self._c("d1a", torch.cat([_up(d2), e1], dim=-1), periodic=p), periodic=p)
        out = self._c("head", d1, relu=False, periodic=p)
        return out[..., :4].float()

    __call__ = forward

    def zero_head_padding(self):
        """Keep the unused head rows and input-channel weights exactly zero."""
        torch = _torch()
        with torch.no_grad():
            self.params["head.w"][4:] = 0
            self.params["head.b"][4:] = 0
            self.params["e1a.w"].mul_(self._e1a_keep)

    def state_dict(self):
        return {k: v.detach().cpu() for k, v in self.params.items()}

    def load_state_dict(self, sd):
        torch = _torch()
        with torch.no_grad():
            for k, v in sd.items():
                self.params[k].copy_(v.to(self.device))


# ---- inputs and targets -----------------------------------------------------------------------------

def periodic_distance(solid, margin: int = 24, device=None):
    """Euclidean distance of every cell to the nearest solid cell under periodic
    wrap, exact up to `margin` cells and capped there (lbm3d_distance_periodic:
    the separable squared-distance transform, one min-plus pass per axis).
    `solid` is a (X, Y, Z) or (N, X, Y, Z) array or device tensor (nonzero =
    solid); returns a float32 tensor of that shape on the GPU."""
    torch = _torch()
    if torch.is_tensor(solid):
        s = (solid != 0).to(torch.int8)
        if device is not None:
            s = s.to(device)
    else:
        s = torch.as_tensor(np.asarray(solid) != 0, device=device).to(torch.int8)
    if s.device.type != "cuda":
        raise RuntimeError("periodic_distance runs on the GPU (lbm3d_distance_periodic)")
    s = s.contiguous()
    shape = s.shape
    vol = shape if s.dim() == 4 else (1,) + tuple(shape)
    dist = torch.empty(shape, dtype=torch.float32, device=s.device)
    work = torch.empty_like(dist)
    _lib.lib().call("lbm3d_distance_periodic", _stream(s.device), _vp(s), _vp(dist), _vp(work), *vol, margin)
    return dist


def make_inputs(solid, device=None):
    """(X, Y, Z, 2) float32 tensor on the device: solid mask and distance / DIST_SCALE.
    `solid` is a numpy array or a device tensor (nonzero = solid)."""
    torch = _torch()
    if torch.is_tensor(solid):
        s = solid.to(device) if device is not None else solid
        s = s != 0
    else:
        s = torch.as_tensor(np.asarray(solid) != 0, device=device)
    return torch.stack([s.float(), periodic_distance(s) / DIST_SCALE], dim=-1)


def make_targets(rho, v, solid, fx, nu_eff):
    """(X, Y, Z, 4) float32 normalised targets; zero on solid cells."""
    fluid = (np.asarray(solid) == 0)[..., None]
    vt = np.asarray(v, np.float32) * (nu_eff / fx)
    rt = (np.asarray(rho, np.float32) - 1.0)[..., None] / (fx * RHO_SCALE)
    return np.where(fluid, np.concatenate([vt, rt], axis=-1), 0.0).astype(np.float32)


def fields_from_prediction(pred, fluid, fx, nu_eff, density=False):
    """Inverse of make_targets on the device: (rho, v, predicted rho) from a
    (X,Y,Z,4) prediction, zero velocity off the fluid.

    By default the warm start takes the predicted velocity with rho = 1, the
    cold start's density.  Every isolated pore keeps its initial mass for good,
    and the steady velocity of a pore scales as 1/rho, so a raw predicted
    density would pin the mean density -- and the permeability, to about the
    prediction's density error -- away from the cold start's.  With
    density=True the returned rho is the predicted density shifted to mean 1
    over the fluid cells: the sample's total mass is the cold start's, and only
    the split of that mass between disconnected pores can differ."""
    torch = _torch()
    p = pred.float()
    fl = fluid.to(p.dtype)
    v = p[..., :3] * (fx / nu_eff) * fl[..., None]
    rho_pred = 1.0 + p[..., 3] * (fx * RHO_SCALE) * fl
    if not density:
        return torch.ones_like(rho_pred), v, rho_pred
    n = fl.sum().clamp_min(1.0)
    dev = (rho_pred - 1.0) * fl
    rho = 1.0 + (dev - dev.sum() / n) * fl
    return rho, v, rho_pred


def to_network_input(inp):
    """(N,X,Y,Z,2) float32 tensor -> (N,X,Y,Z,16) bf16 (channels 2..15 zero)."""
    torch = _torch()
    t = torch.zeros(inp.shape[:4] + (CIN_PAD,), dtype=torch.bfloat16, device=inp.device)
    t[..., :2] = inp.to(torch.bfloat16)
    return t


def periodic_pad(t, pad):
    """(X,Y,Z,...) tensor -> (X+lo+hi, ...) per axis, wrapped periodically; `pad`
    is one width for both sides of every axis or a (lo, hi) pair per axis."""
    torch = _torch()
    for axis in range(3):
        lo, hi = pad[axis] if isinstance(pad, (tuple, list)) else (pad, pad)
        n = t.shape[axis]
        idx = (torch.arange(-lo, n + hi, device=t.device) % n)
        t = t.index_select(axis, idx)
    return t


def inference_pads(shape, wrap: bool):
    """(lo, hi) periodic padding per axis for inference: none when the network
    wraps, else BORDER cells before the volume and at least BORDER after it,
    the padded extent a multiple of 8 (the network pools three times)."""
    if wrap:
        return [(0, 0)] * 3
    return [(BORDER, BORDER + (-(n + 2 * BORDER)) % 8) for n in shape]


def _device_solid(solid, device):
    """(X, Y, Z) bool tensor on the device (True = solid) from an array or tensor."""
    torch = _torch()
    if torch.is_tensor(solid):
        return solid.to(device) != 0
    return torch.from_numpy(np.ascontiguousarray(np.asarray(solid) != 0)).to(device)


def predict(net, solids, fx, nu_eff, batch: int = 2, wrap=None, density=False):
    """Warm-start (rho, v) device tensors predicted for periodic geometries of
    one shape (numpy arrays or device tensors), `batch` volumes per forward pass.
    rho is 1, or with density=True the predicted density at mean 1 over the
    fluid (fields_from_prediction).

    wrap=True runs the network's convolutions periodically on the volume itself
    (every extent a multiple of 8): every cell sees its true periodic
    surroundings.  wrap=False pads each volume periodically by at least BORDER
    cells (to a multiple of 8), runs the zero-padded network and keeps the
    centre.  The default wraps when the extents allow it."""
    torch = _torch()
    out = []
    if not len(solids):
        return out
    shape = tuple(solids[0].shape)
    if wrap is None:
        wrap = all(n % 8 == 0 for n in shape)
    if wrap and any(n % 8 for n in shape):
        raise ValueError(f"periodic inference needs extents that are multiples of 8, got {shape}")
    pads = inference_pads(shape, wrap)
    with torch.no_grad():
        for b0 in range(0, len(solids), batch):
            sol = torch.stack([_device_solid(s, net.device) for s in solids[b0:b0 + batch]])
            dist = periodic_distance(sol)
            inp = torch.stack([sol.float(), dist / DIST_SCALE], dim=-1)
            if not wrap:
                inp = torch.stack([periodic_pad(t, pads) for t in inp])
            pred = net(to_network_input(inp), periodic=wrap)
            (X, Y, Z), (px, py, pz) = shape, (p[0] for p in pads)
            for i in range(sol.shape[0]):
                core = pred[i, px:px + X, py:py + Y, pz:pz + Z]
                out.append(fields_from_prediction(core, ~sol[i], fx, nu_eff, density)[:2])
    return out
