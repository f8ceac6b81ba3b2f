"""GPU: the adjoint backward behind torch.autograd (the cone-flat geometry of the
256x256x64, 720-view benchmark workload).

* dot-product test on 32 seeded pairs: <A x, y> = <x, A^T y> (FP64 sums of FP32
  results), relative error <= 1e-5;
* autograd gradient of 0.5 ||A x - y||^2 equals A^T (A x - y) from the kernels directly;
* central finite differences along 64 random unit directions of
  0.5 ||A x - y||^2 + 1e-3 TV(x), losses summed in FP64; tolerance 1e-3 relative.  Half
  the directions are white noise plus a constant offset (equal energy, h = 1e-3), half
  are sums of eight random low-frequency cosine modes with integer frequencies, so
  exactly zero mean (h = 1e-2): their derivative owes nothing to the constant component,
  and its error is taken relative to the rms of those derivatives.  (For a zero-mean
  white direction the change of the loss over +-h is of the order of the FP32 rounding
  noise of the 23.6M projected samples, so the quotient would measure that noise.)
* double backward (Hessian-vector product) equals A^T A v;
* per-iteration checkpointing of an unrolled reconstruction gives the same gradients as
  storing every activation.
"""
import numpy as np
import pytest
import torch

import xrecon_amd
from xrecon_amd.recon import UnrolledGradientDescent, data_loss, tv

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cfg(cuda):
    proj = xrecon_amd.Projector3d((256, 256, 64), (512, 64), 1.0, 1.0, 500.0, 1000.0, scan_type="flat")
    angles = torch.arange(720, dtype=torch.float32) * np.float32(2 * np.pi / 720)
    return proj, angles


def test_dot_product_32_pairs(cfg, cuda):
    proj, angles = cfg
    g = torch.Generator(device=cuda).manual_seed(1234)
    worst = 0.0
    for k in range(32):
        x = torch.rand((1, 1, 64, 256, 256), device=cuda, generator=g)
        y = torch.rand((1, 1, 720, 64, 512), device=cuda, generator=g)
        lhs = (proj.projection(x, angles).double() * y.double()).sum().item()
        rhs = (x.double() * proj.projection_t(y, angles).double()).sum().item()
        worst = max(worst, abs(lhs - rhs) / abs(rhs))
    print(f"[gpu] dot-product worst relative error over 32 pairs: {worst:.2e}", flush=True)
    assert worst <= 1e-5


def test_autograd_gradient_is_adjoint(cfg, cuda):
    proj, angles = cfg
    g = torch.Generator(device=cuda).manual_seed(7)
    x = torch.rand((1, 1, 64, 256, 256), device=cuda, generator=g).requires_grad_(True)
    y = torch.rand((1, 1, 720, 64, 512), device=cuda, generator=g)
    loss = data_loss(proj, x, y, angles)
    (grad,) = torch.autograd.grad(loss, x)
    with torch.no_grad():
        want = proj.projection_t(proj.projection(x, angles) - y, angles)
    assert torch.equal(grad, want)


def test_central_finite_differences_64_directions(cfg, cuda):
    proj, angles = cfg
    g = torch.Generator(device=cuda).manual_seed(99)
    x = torch.rand((1, 1, 64, 256, 256), device=cuda, generator=g).requires_grad_(True)
    y = proj.projection(0.5 * torch.rand((1, 1, 64, 256, 256), device=cuda, generator=g), angles).detach()

    def f(z):
        r = (proj.projection(z, angles) - y).double()
        return 0.5 * (r * r).sum() + 1e-3 * tv(z.double())

    (grad,) = torch.autograd.grad(f(x), x)
    white, smooth = [], []
    zz, yy, xx = torch.meshgrid(*(torch.arange(n, device=cuda, dtype=torch.float32) / n for n in (64, 256, 256)),
                                indexing="ij")
    with torch.no_grad():
        for k in range(64):
            if k % 2 == 0:
                d, h = torch.randn((1, 1, 64, 256, 256), device=cuda, generator=g) + 1.0, 1e-3
            else:
                d, h = torch.zeros((64, 256, 256), device=cuda), 1e-2
                for m in range(8):
                    kx, ky, kz = (int(v) for v in torch.randint(-3, 4, (3,), generator=torch.Generator().manual_seed(k * 8 + m)))
                    if kx == ky == kz == 0:
                        kx = 1
                    d += torch.cos(2 * np.pi * (kx * xx + ky * yy + kz * zz) + 2 * np.pi * (k + 0.37 * m))
                d = d.view(1, 1, 64, 256, 256)
                assert abs(d.double().mean().item()) < 1e-5
            d = d / d.norm()
            fd = (f(x + h * d) - f(x - h * d)).item() / (2 * h)
            an = (grad.double() * d.double()).sum().item()
            (white if k % 2 == 0 else smooth).append((fd, an))
    # white + constant: relative to each derivative; smooth zero-mean: relative to their rms
    # (a random low-frequency direction can be nearly orthogonal to the gradient, leaving a
    # derivative at the level of the loss's own FP32 rounding noise)
    e_white = [abs(fd - an) / abs(an) for fd, an in white]
    rms = float(np.sqrt(np.mean([an * an for _, an in smooth])))
    e_smooth = [abs(fd - an) / rms for fd, an in smooth]
    print(f"[gpu] central FD error: white+constant median {np.median(e_white):.2e} worst {max(e_white):.2e}; "
          f"smooth zero-mean (vs rms {rms:.3e}) median {np.median(e_smooth):.2e} worst {max(e_smooth):.2e}", flush=True)
    assert max(e_white) <= 1e-3 and max(e_smooth) <= 1e-3


def test_double_backward_hessian_vector(cuda):
    proj = xrecon_amd.Projector3d((32, 32, 8), (48, 8), 1.0, 1.0, 500.0, 1000.0)
    angles = torch.arange(36, dtype=torch.float32) * np.float32(2 * np.pi / 36)
    g = torch.Generator(device=cuda).manual_seed(3)
    x = torch.rand((1, 1, 8, 32, 32), device=cuda, generator=g).requires_grad_(True)
    y = torch.rand((1, 1, 36, 8, 48), device=cuda, generator=g)
    v = torch.rand((1, 1, 8, 32, 32), device=cuda, generator=g)
    (gx,) = torch.autograd.grad(data_loss(proj, x, y, angles), x, create_graph=True)
    (hv,) = torch.autograd.grad((gx * v).sum(), x)
    with torch.no_grad():
        want = proj.projection_t(proj.projection(v, angles), angles)
    assert torch.allclose(hv, want, rtol=1e-5, atol=1e-6 * want.abs().max().item())


def test_checkpointed_unrolled_gradients_match(cuda):
    proj = xrecon_amd.Projector3d((32, 32, 8), (48, 8), 1.0, 1.0, 500.0, 1000.0)
    angles = torch.arange(36, dtype=torch.float32) * np.float32(2 * np.pi / 36)
    g = torch.Generator(device=cuda).manual_seed(4)
    xt = torch.rand((2, 1, 8, 32, 32), device=cuda, generator=g)
    y = proj.projection(xt, angles).detach()
    grads = []
    for ckpt in (True, False):
        net = UnrolledGradientDescent(proj, angles, iterations=4, step=1e-4, use_checkpoint=ckpt).to(cuda)
        out = net(y)
        loss = ((out - xt) ** 2).sum()
        loss.backward()
        grads.append(net.steps.grad.clone())
    assert torch.equal(grads[0], grads[1])
