"""Short training-step or inference loop of the warm-start U-Net for kernel profiles.

    python taichi_lbm3d_amd/profile_unet.py [--size 128] [--batch 2] [--iters 20]
    python taichi_lbm3d_amd/profile_unet.py --infer [--size 192] [--batch 2] [--iters 20]

Random inputs and targets of the training shape (batch x size^3, 2 input and
4 output channels); each iteration is one forward, backward and Adam step,
as in warmstart.train.  With --infer each iteration is one forward pass with
periodic convolutions and no autograd, as in surrogate.predict.  Prints the
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
    level = {"e1": 0, "e2": 1, "e3": 2, "e4": 3, "d3": 2, "d2": 1, "d1": 0, "he": 0}
    for name, (cin, cout, ks) in net.specs.items():
        vox = batch * (size >> level[name[:2]]) ** 3
        real_cin = 2 if name == "e1a" else cin
        real_cout = 4 if name == "head" else cout
        f = 2 * vox * ks ** 3 * real_cin * real_cout
        total += f * ((2 if name == "e1a" else 3) if train else 1)
    return total


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--size", type=int, default=None, help="volume edge (default 128, or 192 with --infer)")
    ap.add_argument("--batch", type=int, default=2)
    ap.add_argument("--iters", type=int, default=20)
    ap.add_argument("--infer", action="store_true", help="time periodic forward passes instead of training steps")
    args = ap.parse_args()
    args.size = args.size or (192 if args.infer else 128)
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
