"""Short training-step or inference loop of the warm-start U-Net for kernel profiles.

    python taichi_lbm3d_amd/profile_unet.py [--size 128] [--batch 2] [--iters 20]
    python taichi_lbm3d_amd/profile_unet.py --infer [--size 192] [--batch 2] [--iters 20]
    python taichi_lbm3d_amd/profile_unet.py --layers [--size 192] [--batch 2] [--iters 10]
    python taichi_lbm3d_amd/profile_unet.py --wgrad-layers [--size 128] [--batch 2] [--iters 10]

Random inputs and targets of the training shape (batch x size^3, 2 input and
4 output channels); each iteration is one forward, backward and Adam step,
as in warmstart.train.  With --infer each iteration is one forward pass with
periodic convolutions and no autograd, as in surrogate.predict.  With
--layers every convolution of the network is timed alone at its inference
shape (one JSON line per layer: time, TFLOP/s, activation GB/s); with
--wgrad-layers every weight gradient at its training shape.  Prints the
mean iteration time and the convolution FLOP rate it implies (GEMM FLOPs of
every 3x3x3 and 1x1x1 convolution: forward, plus data and weight gradients
when training).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

if __package__ in (None, ""):  # run as a file
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def conv_flops(net, size, batch, train=True):
    """FLOPs of one training step's convolutions (forward, dgrad of all but the
    first, wgrad), or of one forward pass with train=False."""
    total = 0
    for name, (cin, cout, ks) in net.specs.items():
        vox = batch * (size >> LEVEL[name[:2]]) ** 3
        real_cin = 2 if name == "e1a" else cin
        real_cout = 4 if name == "head" else cout
        f = 2 * vox * ks ** 3 * real_cin * real_cout
        total += f * ((2 if name == "e1a" else 3) if train else 1)
    return total


LEVEL = {"e1": 0, "e2": 1, "e3": 2, "e4": 3, "d3": 2, "d2": 1, "d1": 0, "he": 0}


def time_layers(args):
    """Each convolution of the network alone, forward, periodic, at its
    inference shape (batch x (size >> level)^3)."""
    import torch
    from taichi_lbm3d_amd import surrogate as sg
    dev = torch.device("cuda", 0)
    net = sg.UNet3D(dev, seed=0)
    total = 0.0
    for name, (cin, cout, ks) in net.specs.items():
        n = args.size >> LEVEL[name[:2]]
        x = (torch.randn((args.batch, n, n, n, cin), device=dev) * 0.5).to(torch.bfloat16)
        wb = net.params[name + ".w"].to(torch.bfloat16).contiguous()
        b = net.params[name + ".b"]
        for _ in range(3):
            sg.conv_fwd(x, wb, b, cout, ks, True, periodic=True)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.iters):
            sg.conv_fwd(x, wb, b, cout, ks, True, periodic=True)
        e1.record()
        torch.cuda.synchronize()
        us = e0.elapsed_time(e1) / args.iters * 1e3
        vox = args.batch * n ** 3
        total += us
        print(json.dumps({"layer": name, "cin": cin, "cout": cout, "edge": n, "us": round(us, 1),
                          "tflops": round(2.0 * vox * ks ** 3 * cin * cout / us / 1e6, 1),
                          "gbs": round(vox * (cin + cout) * 2 / us / 1e3, 1)}), flush=True)
    print(json.dumps({"layers_total_us": round(total, 1)}))


def time_wgrad_layers(args):
    """Each convolution's weight gradient alone at its training shape
    (batch x (size >> level)^3), random x and output gradient."""
    import torch
    from taichi_lbm3d_amd import surrogate as sg
    dev = torch.device("cuda", 0)
    net = sg.UNet3D(dev, seed=0)
    total = 0.0
    for name, (cin, cout, ks) in net.specs.items():
        n = args.size >> LEVEL[name[:2]]
        x = (torch.randn((args.batch, n, n, n, cin), device=dev) * 0.5).to(torch.bfloat16)
        gy = (torch.randn((args.batch, n, n, n, cout), device=dev) * 0.5).to(torch.bfloat16)
        for _ in range(3):
            sg.conv_wgrad(x, gy, ks)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.iters):
            sg.conv_wgrad(x, gy, ks)
        e1.record()
        torch.cuda.synchronize()
        us = e0.elapsed_time(e1) / args.iters * 1e3
        total += us
        print(json.dumps({"wgrad_layer": name, "cin": cin, "cout": cout, "edge": n, "us": round(us, 1),
                          "tflops": round(2.0 * args.batch * n ** 3 * ks ** 3 * cin * cout / us / 1e6, 1)}),
              flush=True)
    print(json.dumps({"wgrad_layers_total_us": round(total, 1)}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--size", type=int, default=None, help="volume edge (default 128, or 192 with --infer)")
    ap.add_argument("--batch", type=int, default=2)
    ap.add_argument("--iters", type=int, default=20)
    ap.add_argument("--infer", action="store_true", help="time periodic forward passes instead of training steps")
    ap.add_argument("--layers", action="store_true", help="time each convolution alone at its inference shape")
    ap.add_argument("--wgrad-layers", action="store_true",
                    help="time each weight gradient alone at its training shape")
    args = ap.parse_args()
    args.size = args.size or (192 if (args.infer or args.layers) else 128)
    if args.layers:
        return time_layers(args)
    if args.wgrad_layers:
        return time_wgrad_layers(args)
    import torch
    from taichi_lbm3d_amd import surrogate as sg
    dev = torch.device("cuda", 0)
    net = sg.UNet3D(dev, seed=0)
    opt = torch.optim.Adam(net.parameters(), lr=1e-3)
    g = torch.Generator(device=dev).manual_seed(0)
    shape = (args.batch,) + (args.size,) * 3
    x = torch.zeros(shape + (sg.CIN_PAD,), dtype=torch.bfloat16, device=dev)
    x[..., :2] = torch.rand(shape + (2,), generator=g, device=dev).to(torch.bfloat16)
    y = torch.randn(shape + (4,), generator=g, device=dev)

    def it():
        if args.infer:
            with torch.no_grad():
                net(x, periodic=True)
            return
        out = net(x)
        loss = ((out - y) ** 2).mean()
        opt.zero_grad(set_to_none=True)
        loss.backward()
        opt.step()

    for _ in range(3):
        it()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.iters):
        it()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.iters
    fl = conv_flops(net, args.size, args.batch, train=not args.infer)
    print(json.dumps({"mode": "infer" if args.infer else "train", "size": args.size, "batch": args.batch,
                      "ms_per_iter": round(dt * 1e3, 2),
                      "conv_tflop_per_iter": round(fl / 1e12, 3), "conv_tflops": round(fl / dt / 1e12, 1)}))


if __name__ == "__main__":
    main()
