"""
The autograd adjoint: gradcheck against finite differences (FP64, a 33x33
vertex pair), backward bit-identical to the reference's
regrid_from_weights(*transpose_weights(w)) order (the oracle over the swapped
triples), FP32, and the least-squares inversion driven by it.
"""

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle
from regridding_amd import _util, workloads
from regridding_amd._conservative import build_2d, build_2d_batched
from regridding_amd.inversion import least_squares, lipschitz
from regridding_amd.torch_ops import regrid_apply, regrid_apply_views

pytestmark = pytest.mark.gpu


def _pair(n_in=33, n_out=33):
    gi = workloads.warped_grid(n_in, 0.3, 0.05, 0.05)
    go = workloads.warped_grid(n_out, 0.1, 0.0, 0.0, -0.9, 0.9)
    ci, co, *_ = _util.normalize_coordinates(gi, go, perturb=True)
    return ci, co


def test_gradcheck_33x33_fp64():
    ci, co = _pair()
    w = build_2d(ci, co)
    x = torch.randn(2, w.n_in, dtype=torch.float64, device="cuda", requires_grad=True)
    assert torch.autograd.gradcheck(lambda v: regrid_apply(w, v), (x,), eps=1e-6, atol=1e-9)
    y = torch.randn(2, w.n_out, dtype=torch.float64, device="cuda", requires_grad=True)
    assert torch.autograd.gradcheck(lambda v: regrid_apply(w, v, transpose=True), (y,),
                                    eps=1e-6, atol=1e-9)


def test_backward_is_transpose_weights_apply():
    ci, co = _pair(41, 29)
    w = build_2d(ci, co)
    ii, io, ww = w.export()
    rng = np.random.default_rng(3)
    x = torch.tensor(rng.normal(size=(3, w.n_in)), device="cuda", requires_grad=True)
    gy = rng.normal(size=(3, w.n_out))
    y = regrid_apply(w, x)
    assert np.array_equal(y.detach().cpu().numpy(), oracle.apply(ii, io, ww, x.detach().cpu().numpy(), w.n_out))
    y.backward(torch.tensor(gy, device="cuda"))
    # transpose_weights swaps the index arrays; applying them accumulates
    # every input cell in ascending output order
    assert np.array_equal(x.grad.cpu().numpy(), oracle.apply(io, ii, ww, gy, w.n_in))


def _views(n_views=3, n_in=33):
    gi = workloads.warped_grid(n_in, 0.3, 0.05, 0.05)
    pairs = []
    for k in range(n_views):
        go = workloads.warped_grid(21 + 6 * k, 0.1 * k, 0.02, 0.0, -0.9, 0.9)
        ci, co, *_ = _util.normalize_coordinates(gi, go, perturb=True)
        pairs.append((ci, co))
    return build_2d_batched(pairs)


def test_views_forward_and_backward_match_single_views():
    ws = _views()
    rng = np.random.default_rng(11)
    x = torch.tensor(rng.normal(size=ws[0].n_in), device="cuda", requires_grad=True)
    y = regrid_apply_views(ws, x)
    ref = torch.cat([regrid_apply(w, x.detach()) for w in ws])
    assert torch.equal(y.detach(), ref)  # the same gather per element, bit for bit
    g = torch.tensor(rng.normal(size=y.numel()), device="cuda")
    y.backward(g)
    parts = torch.split(g, [w.n_out for w in ws])
    adj = torch.stack([regrid_apply(w, p, transpose=True) for w, p in zip(ws, parts)])
    # each element's adjoint is bit-identical; the sum over elements is torch's
    assert torch.equal(x.grad, adj.sum(0))


def test_views_gradcheck_fp64():
    ws = _views(2, 17)
    x = torch.randn(ws[0].n_in, dtype=torch.float64, device="cuda", requires_grad=True)
    assert torch.autograd.gradcheck(lambda v: regrid_apply_views(ws, v), (x,), eps=1e-6,
                                    atol=1e-9)


def test_views_fp32_and_errors():
    ws = _views()
    x = torch.randn(ws[0].n_in, dtype=torch.float64, device="cuda")
    out = regrid_apply_views(ws, x.float())
    ref = regrid_apply_views(ws, x).float()
    assert torch.max(torch.abs(out - ref)) <= 1e-5 * torch.max(torch.abs(ref))
    with pytest.raises(ValueError):
        regrid_apply_views(ws, x[:-1])
    with pytest.raises(ValueError):
        regrid_apply_views(ws, x.reshape(1, -1))


def test_adjoint_identity_and_fp32():
    ci, co = _pair(50, 37)
    w = build_2d(ci, co)
    x = torch.randn(4, w.n_in, dtype=torch.float64, device="cuda")
    y = torch.randn(4, w.n_out, dtype=torch.float64, device="cuda")
    lhs = (regrid_apply(w, x) * y).sum()
    rhs = (x * regrid_apply(w, y, transpose=True)).sum()
    assert abs(float(lhs - rhs)) <= 1e-12 * float(lhs.abs() + rhs.abs())
    x32 = x.float().requires_grad_(True)
    out = regrid_apply(w, x32)
    ref = regrid_apply(w, x).float()
    assert torch.max(torch.abs(out - ref)) <= 1e-5 * torch.max(torch.abs(ref))
    out.sum().backward()
    g64 = regrid_apply(w, torch.ones(4, w.n_out, dtype=torch.float64, device="cuda"), transpose=True)
    assert torch.max(torch.abs(x32.grad.double() - g64)) <= 1e-5 * torch.max(torch.abs(g64))


def test_inversion_reduces_misfit_and_error():
    grid, image, views, rng = workloads.inversion_problem(n_truth=65, n_obs=49, n_views=4)
    pairs = []
    for v in views:
        ci, co, *_ = _util.normalize_coordinates(grid, v, perturb=True)
        pairs.append((ci, co))
    hs = build_2d_batched(pairs)
    truth = torch.tensor(image.ravel(), device="cuda")
    obs = []
    for w in hs:
        lam = regrid_apply(w, truth).cpu().numpy()
        obs.append(torch.tensor(rng.poisson(np.maximum(lam, 0)).astype(np.float64), device="cuda"))
    views_t = list(zip(hs, obs))
    L = lipschitz(views_t, truth.numel(), device="cuda")
    assert L > 0
    x0 = torch.full_like(truth, float(np.mean(image)))
    x, losses = least_squares(views_t, x0, iters=60)
    assert losses[-1] < 0.05 * losses[0]
    # the one-launch loss of all views equals the sum of the per-view losses
    from regridding_amd.inversion import _fused, _loss

    fused = _fused(views_t)
    assert fused is not None
    a, b = float(_loss(views_t, x, fused)), float(_loss(views_t, x, None))
    assert abs(a - b) <= 1e-12 * abs(b)
    err0 = float(torch.linalg.norm(x0 - truth))
    err = float(torch.linalg.norm(x - truth))
    assert err < 0.5 * err0


def test_free_waits_for_applies_on_another_stream():
    """Weights applied on a side stream and freed at once: the free returns
    only after the side stream's applies, whose results are intact (the CSR
    memory goes back to the pool of the stream it was built on)."""
    ci, co = _pair(513, 385)
    h = build_2d(ci, co)
    x = torch.tensor(np.random.default_rng(8).poisson(30, (16, h.n_in)).astype(np.float64),
                     device="cuda")
    expect = regrid_apply(h, x)
    torch.cuda.synchronize()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        ys = [regrid_apply(h, x) for _ in range(40)]
    h.free()  # no synchronisation in between
    assert side.query(), "the free returned while applies on the side stream were queued"
    for y in ys:
        assert torch.equal(y, expect)
