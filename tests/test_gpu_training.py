"""Loss seeding and the fused Adam step on gfx950 against torch's own
l1_loss / mse_loss and torch.optim.Adam (fp32 references of the same ops)."""
import pytest
import torch

from gaussian_splatting_3d_amd import training

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


@pytest.mark.parametrize("kind", ["l1", "l2"])
@pytest.mark.parametrize("n", [1, 1000, 1297 * 840 * 3])
def test_photometric_loss_matches_torch(kind, n):
    g = torch.Generator().manual_seed(n)
    out = torch.rand(n, generator=g).to(DEV).requires_grad_(True)
    gt = torch.rand(n, generator=g).to(DEV)
    gt[: n // 10] = out.detach()[: n // 10]  # exact zeros of the difference
    fn = torch.nn.functional.l1_loss if kind == "l1" else torch.nn.functional.mse_loss
    ref = fn(out, gt)
    ref.backward()
    loss, grad = training.photometric_loss(out.detach(), gt, kind)
    assert abs(float(loss) - float(ref)) <= 1e-6 * max(abs(float(ref)), 1e-30)
    assert float((grad - out.grad).abs().max()) <= 1e-6 * float(out.grad.abs().max())


def _groups(seed):
    g = torch.Generator().manual_seed(seed)
    shapes = [(5000, 3), (5000, 4), (5000, 3), (5000, 3, 16), (5000,)]
    return [torch.randn(s, generator=g).to(DEV) for s in shapes], [1.6e-4, 1e-3, 5e-3, 2.5e-3, 5e-2]


def _close(a, b, tol=2e-6):
    return float((a - b).abs().max()) <= tol * max(float(b.abs().max()), 1e-30)


def test_fused_adam_matches_torch_adam_over_steps():
    ps, lrs = _groups(1)
    ref = [p.clone().requires_grad_(True) for p in ps]
    opt = torch.optim.Adam([{"params": [p], "lr": lr} for p, lr in zip(ref, lrs)], betas=(0.9, 0.99), foreach=False)
    mine = [p.clone() for p in ps]
    fa = training.FusedAdam(list(zip(mine, lrs)))
    gen = torch.Generator().manual_seed(7)
    for _ in range(4):
        grads = [torch.randn(p.shape, generator=gen).to(DEV) * 0.1 for p in ps]
        for p, gr in zip(ref, grads):
            p.grad = gr.clone()
        opt.step()
        fa.step(grads)
        for a, b in zip(mine, ref):
            assert _close(a, b.detach())


@pytest.mark.parametrize("fresh", [False, True])
def test_fused_adam_ragged_and_unaligned_groups(fresh):
    """Groups whose length is not a multiple of four (the last elements take
    the scalar path) and a group starting 4 bytes into its storage (no
    16-byte vectors at all) match torch's Adam like the aligned ones."""
    g = torch.Generator().manual_seed(5)
    base = torch.randn(10_003, generator=g).to(DEV)
    ps = [torch.randn(4_999, 3, generator=g).to(DEV), base[1:], torch.randn(7, generator=g).to(DEV)]
    lrs = [1e-3, 5e-3, 2e-2]
    ref = [p.clone().requires_grad_(True) for p in ps]
    mine = [p if k == 1 else p.clone() for k, p in enumerate(ps)]  # the offset view is updated in place
    assert mine[1].data_ptr() % 16 == 4
    fa = training.FusedAdam(list(zip(mine, lrs)), fresh=fresh)
    opt = torch.optim.Adam([{"params": [p], "lr": lr} for p, lr in zip(ref, lrs)], betas=(0.9, 0.99), foreach=False)
    gen = torch.Generator().manual_seed(13)
    for _ in range(3):
        grads = [torch.randn(p.shape, generator=gen).to(DEV) * 0.1 for p in ps]
        if fresh:
            opt = torch.optim.Adam([{"params": [p], "lr": lr} for p, lr in zip(ref, lrs)], betas=(0.9, 0.99),
                                   foreach=False)
        for p, gr in zip(ref, grads):
            p.grad = gr.clone()
        opt.step()
        fa.step(grads)
        for a_, b in zip(mine, ref):
            assert _close(a_, b.detach())
    assert float(base[0]) == float(torch.randn(10_003, generator=torch.Generator().manual_seed(5))[0])  # untouched


def test_fused_adam_fresh_matches_rebuilt_optimizer():
    """The reference builds a new Adam every iteration (main_sh.py:238)."""
    ps, lrs = _groups(2)
    ref = [p.clone().requires_grad_(True) for p in ps]
    mine = [p.clone() for p in ps]
    fa = training.FusedAdam(list(zip(mine, lrs)), fresh=True)
    gen = torch.Generator().manual_seed(9)
    for _ in range(3):
        grads = [torch.randn(p.shape, generator=gen).to(DEV) for p in ps]
        opt = torch.optim.Adam([{"params": [p], "lr": lr} for p, lr in zip(ref, lrs)], betas=(0.9, 0.99),
                               foreach=False)
        for p, gr in zip(ref, grads):
            p.grad = gr.clone()
        opt.step()
        fa.step(grads)
        for a, b in zip(mine, ref):
            assert _close(a, b.detach())
