"""GPU tests of the BF16 MFMA convolutions against a plain PyTorch FP32 CPU
convolution of the same bf16-rounded operands (forward, data gradient, weight
gradient), and of the U-Net built on them."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def rand_bf16(shape, seed, scale=1.0):
    torch = _torch()
    g = torch.Generator().manual_seed(seed)
    return (torch.randn(shape, generator=g) * scale).to(torch.bfloat16)


def torch_weight(w, cin, cout, ks):
    """[cout][kpad] (k = tap*cin + ci, tap = kd*ks*ks + kh*ks + kw) -> torch [cout, cin, ks, ks, ks]."""
    taps = ks ** 3
    return w[:, : taps * cin].float().view(cout, taps, cin).permute(0, 2, 1).reshape(cout, cin, ks, ks, ks)


def cpu_conv(x, w, b, cin, cout, ks, relu):
    torch = _torch()
    import torch.nn.functional as F
    xt = x.float().permute(0, 4, 1, 2, 3)
    y = F.conv3d(xt, torch_weight(w, cin, cout, ks), b, padding=ks // 2)
    if relu:
        y = torch.relu(y)
    return y.permute(0, 2, 3, 4, 1)


SHAPES = [
    # n, d, h, w, cin, cout, ks
    (2, 8, 6, 10, 16, 32, 3),
    (1, 5, 7, 9, 48, 16, 3),
    (1, 4, 4, 6, 192, 64, 3),
    (2, 6, 5, 4, 32, 128, 3),
    (1, 8, 8, 8, 16, 16, 1),
    (1, 3, 4, 5, 96, 48, 3),
]


def make_case(shape, seed):
    torch = _torch()
    from taichi_lbm3d_amd import _lib
    n, d, h, w, cin, cout, ks = shape
    kpad = _lib.lib().cdll.lbm3d_conv_kpad(cin, ks)
    x = rand_bf16((n, d, h, w, cin), seed)
    wt = torch.zeros(cout, kpad, dtype=torch.bfloat16)
    wt[:, : ks ** 3 * cin] = rand_bf16((cout, ks ** 3 * cin), seed + 1, scale=1.0 / np.sqrt(ks ** 3 * cin))
    b = (torch.randn(cout, generator=torch.Generator().manual_seed(seed + 2)) * 0.1).float()
    return x, wt, b


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("relu", [False, True])
def test_conv_forward_matches_cpu_fp32(gpu, shape, relu):
    from taichi_lbm3d_amd.surrogate import conv_fwd
    n, d, h, w, cin, cout, ks = shape
    x, wt, b = make_case(shape, 10)
    y = conv_fwd(x.to(gpu), wt.to(gpu), b.to(gpu), cout, ks, relu).float().cpu()
    ref = cpu_conv(x, wt, b, cin, cout, ks, relu)
    err = (y - ref).abs().max().item()
    # FP32 accumulation on both sides; the GPU rounds its output to bf16 (8 bits)
    assert err <= 2 ** -8 * ref.abs().max().item() + 1e-6, err
    rel = ((y - ref).norm() / ref.norm()).item()
    assert rel < 4e-3, rel


@pytest.mark.parametrize("shape", SHAPES + [(1, 1, 2, 3, 16, 16, 3), (1, 2, 9, 17, 16, 32, 3)])
def test_conv_forward_periodic_matches_circular_cpu_fp32(gpu, shape):
    """Periodic forward convolution against a CPU convolution of the circularly
    padded input (one-voxel volumes wrap onto themselves)."""
    import torch.nn.functional as F
    from taichi_lbm3d_amd.surrogate import conv_fwd
    n, d, h, w, cin, cout, ks = shape
    x, wt, b = make_case(shape, 15)
    y = conv_fwd(x.to(gpu), wt.to(gpu), b.to(gpu), cout, ks, True, periodic=True).float().cpu()
    xt = x.float().permute(0, 4, 1, 2, 3)
    p = ks // 2
    if p:
        # circular padding by one voxel, axis by axis (valid for extents of 1 too)
        for dim in (2, 3, 4):
            m = xt.shape[dim]
            idx = _torch().tensor([m - 1] + list(range(m)) + [0])
            xt = xt.index_select(dim, idx)
    ref = F.conv3d(xt, torch_weight(wt, cin, cout, ks), b, padding=0).relu().permute(0, 2, 3, 4, 1)
    err = (y - ref).abs().max().item()
    assert err <= 2 ** -8 * ref.abs().max().item() + 1e-6, err


@pytest.mark.parametrize("shape", SHAPES)
def test_conv_data_gradient_matches_cpu_autograd(gpu, shape):
    torch = _torch()
    from taichi_lbm3d_amd.surrogate import conv_dgrad
    n, d, h, w, cin, cout, ks = shape
    x, wt, b = make_case(shape, 20)
    gy = rand_bf16((n, d, h, w, cout), 23)
    xr = x.float().requires_grad_(True)
    y = cpu_conv(xr, wt, b, cin, cout, ks, False)
    y.backward(gy.float())
    ref = xr.grad
    dx = conv_dgrad(gy.to(gpu), wt.to(gpu), cin, ks).float().cpu()
    err = (dx - ref).abs().max().item()
    assert err <= 2 ** -8 * ref.abs().max().item() + 1e-6, err


# the last two shapes leave hundreds of partial sets to fold (and give each 3x3x3
# workgroup two bricks, the last one a single brick)
@pytest.mark.parametrize("shape", SHAPES + [(1, 42, 72, 100, 16, 32, 3), (2, 20, 24, 40, 32, 16, 1)])
def test_conv_weight_gradient_matches_cpu_autograd(gpu, shape):
    torch = _torch()
    from taichi_lbm3d_amd.surrogate import conv_wgrad
    n, d, h, w, cin, cout, ks = shape
    x, wt, b = make_case(shape, 30)
    gy = rand_bf16((n, d, h, w, cout), 33)
    wr = torch_weight(wt, cin, cout, ks).clone().requires_grad_(True)
    import torch.nn.functional as F
    y = F.conv3d(x.float().permute(0, 4, 1, 2, 3), wr, b, padding=ks // 2)
    y.backward(gy.float().permute(0, 4, 1, 2, 3))
    ref = wr.grad.permute(0, 2, 3, 4, 1).reshape(cout, ks ** 3 * cin)
    dw = conv_wgrad(x.to(gpu), gy.to(gpu), ks).cpu()
    assert dw.shape[1] == wt.shape[1]
    assert torch.all(dw[:, ks ** 3 * cin:] == 0)
    err = (dw[:, : ks ** 3 * cin] - ref).abs().max().item()
    assert err <= 1e-4 * ref.abs().max().item(), err
    # deterministic: the partial sums fold in a fixed order
    dw2 = conv_wgrad(x.to(gpu), gy.to(gpu), ks).cpu()
    assert torch.equal(dw, dw2)


def test_unet_gradients_match_cpu_fp32(gpu):
    """The whole network: loss and parameter gradients of the MFMA path against
    a CPU FP32 restatement of the same network with the same weights."""
    torch = _torch()
    import torch.nn.functional as F
    from taichi_lbm3d_amd.surrogate import UNet3D, CIN_PAD
    net = UNet3D(gpu, seed=3)
    g = torch.Generator().manual_seed(5)
    x = torch.zeros(1, 16, 16, 16, CIN_PAD)
    x[..., :2] = torch.rand(1, 16, 16, 16, 2, generator=g)
    x = x.to(torch.bfloat16)
    target = torch.randn(1, 16, 16, 16, 4, generator=g)
    out = net(x.to(gpu))
    loss = ((out - target.to(gpu)) ** 2).mean()
    loss.backward()

    cpu = {k: v.detach().cpu().to(torch.bfloat16).float().requires_grad_(True) for k, v in net.params.items()}

    def c(name, t, relu=True):
        cin, cout, ks = net.specs[name]
        y = F.conv3d(t.permute(0, 4, 1, 2, 3), torch_weight(cpu[name + ".w"], cin, cout, ks), cpu[name + ".b"],
                     padding=ks // 2).permute(0, 2, 3, 4, 1)
        return torch.relu(y) if relu else y

    def pool(t):
        n, d, h, w, ch = t.shape
        return t.view(n, d // 2, 2, h // 2, 2, w // 2, 2, ch).amax(dim=(2, 4, 6))

    def up(t):
        return t.repeat_interleave(2, 1).repeat_interleave(2, 2).repeat_interleave(2, 3)

    xf = x.float()
    e1 = c("e1b", c("e1a", xf))
    e2 = c("e2b", c("e2a", pool(e1)))
    e3 = c("e3b", c("e3a", pool(e2)))
    e4 = c("e4b", c("e4a", pool(e3)))
    d3 = c("d3b", c("d3a", torch.cat([up(e4), e3], -1)))
    d2 = c("d2b", c("d2a", torch.cat([up(d3), e2], -1)))
    d1 = c("d1b", c("d1a", torch.cat([up(d2), e1], -1)))
    ref = c("head", d1, relu=False)[..., :4]
    lref = ((ref - target) ** 2).mean()
    lref.backward()
    assert abs(loss.item() - lref.item()) < 0.03 * lref.item()
    assert ((out.detach().cpu() - ref).norm() / ref.norm()).item() < 0.03
    for k in net.params:
        ga = net.params[k].grad.detach().cpu().flatten()
        gb = cpu[k].grad.flatten()
        if gb.norm() == 0:
            continue
        cos = float(ga @ gb / (ga.norm() * gb.norm()))
        assert cos > 0.99, (k, cos)
        assert abs(ga.norm() / gb.norm() - 1) < 0.05, (k, float(ga.norm() / gb.norm()))


def test_unet_periodic_inference_is_translation_equivariant(gpu):
    """With periodic convolutions, rolling a periodic input by a multiple of 8
    voxels rolls the prediction by the same amount, bit for bit."""
    torch = _torch()
    from taichi_lbm3d_amd import surrogate as sg
    net = sg.UNet3D(gpu, seed=3)
    g = torch.Generator().manual_seed(4)
    x = torch.zeros((1, 32, 24, 16, sg.CIN_PAD), dtype=torch.bfloat16)
    x[..., :2] = torch.rand((1, 32, 24, 16, 2), generator=g).to(torch.bfloat16)
    x = x.to(gpu)
    with torch.no_grad():
        y = net(x, periodic=True)
        yr = net(torch.roll(x, shifts=(8, 16, 8), dims=(1, 2, 3)).contiguous(), periodic=True)
    assert torch.equal(torch.roll(y, shifts=(8, 16, 8), dims=(1, 2, 3)), yr)
    # and zero padding breaks that symmetry at the faces
    with torch.no_grad():
        y0 = net(x)
    assert not torch.equal(y0, y)


@pytest.mark.parametrize("shape,por,margin", [((24, 24, 24), 0.3, 24), ((40, 32, 48), 0.6, 24),
                                              ((48, 48, 48), 0.9, 24), ((10, 12, 9), 0.5, 24),
                                              ((20, 16, 24), 0.7, 5)])
def test_distance_kernel_matches_oracle(gpu, shape, por, margin):
    """lbm3d_distance_periodic equals the scipy-EDT oracle exactly, batched and
    single, on cubic and ragged periodic packs (one smaller than the margin)."""
    from oracle.geo_ref import periodic_distance_edt
    from taichi_lbm3d_amd.geometry import sphere_pack
    from taichi_lbm3d_amd.surrogate import periodic_distance
    torch = _torch()
    packs = [sphere_pack(shape, porosity=por, rmin=2, rmax=5, seed=11 + i, periodic=(True,) * 3) for i in range(2)]
    batch = periodic_distance(torch.as_tensor(np.stack(packs)).to(gpu), margin=margin).cpu().numpy()
    for i, s in enumerate(packs):
        want = periodic_distance_edt(s, margin)
        np.testing.assert_array_equal(periodic_distance(s, margin=margin, device=gpu).cpu().numpy(), want)
        np.testing.assert_array_equal(batch[i], want)
    empty = np.zeros(shape, np.int8)
    np.testing.assert_array_equal(periodic_distance(empty, margin=margin, device=gpu).cpu().numpy(),
                                  np.full(shape, float(margin), np.float32))


@pytest.mark.parametrize("vox,c,relu", [(1, 16, True), (7, 32, True), (1000, 64, False), (2 * 33 ** 3, 16, True),
                                        (5 * 17 ** 3, 128, True), (300_000, 32, True)])
def test_relu_bias_grad_matches_torch(gpu, vox, c, relu):
    """Fused ReLU backward and bias gradient: dz equals PyTorch's g * (y > 0)
    bit for bit (signed zeros included; NaN where masking a non-finite g gives
    NaN), db
    equals the FP32 sum over voxels to rounding, and the result repeats
    exactly (fixed summation order)."""
    torch = _torch()
    from taichi_lbm3d_amd.surrogate import relu_bias_grad
    gen = torch.Generator().manual_seed(vox + c)
    g = torch.randn((vox, c), generator=gen).to(torch.bfloat16)
    y = torch.randn((vox, c), generator=gen).to(torch.bfloat16)
    if vox >= 1000:  # special values at fixed places
        g[0, 0], g[1, 1], g[2, 2], g[3, 3] = float("nan"), float("inf"), -float("inf"), -0.0
        y[0, 0], y[1, 1], y[2, 2], y[4, 4], y[5, 5] = -1.0, 0.0, 1.0, float("nan"), -0.0
    want = (g * (y > 0)) if relu else g
    dz, db = relu_bias_grad(g.to(gpu), y.to(gpu) if relu else None)
    dz, db = dz.cpu(), db.cpu()
    # NaN where PyTorch has NaN (its payload is platform-defined), every other
    # element bit for bit, signed zeros included
    nan = torch.isnan(want)
    assert torch.equal(torch.isnan(dz), nan)
    assert torch.equal(dz.view(torch.int16)[~nan], want.view(torch.int16)[~nan])
    ref = want.double().sum(0)
    scale = want.double().abs().sum(0)
    fin = torch.isfinite(ref)
    assert torch.all((db.double()[fin] - ref[fin]).abs() <= 1e-5 * scale[fin] + 1e-30)
    assert torch.equal(torch.isnan(db), torch.isnan(ref.float()))
    dz2, db2 = relu_bias_grad(g.to(gpu), y.to(gpu) if relu else None)
    assert torch.equal(db2.cpu().view(torch.int32), db.view(torch.int32))
