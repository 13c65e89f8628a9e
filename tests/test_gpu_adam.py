"""GPU parity of the flat-buffer Adam kernel against torch.optim.Adam (the
optimiser of exp_runner.py:97) over several steps with a changing learning
rate; |err| <= 1e-6 + 1e-5 |ref| (same fp32 recurrence, different op order)."""
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 7, 4096 + 3, 1_000_003])
def test_adam_matches_torch(n, dev):
    from supernormal_amd._lib import call, ptr, stream_of
    g = torch.Generator(device="cpu").manual_seed(n)
    p0 = torch.randn(n, generator=g)
    ref = p0.clone().to(dev).requires_grad_(True)
    opt = torch.optim.Adam([ref], lr=1e-2)
    p = p0.clone().to(dev)
    m = torch.zeros_like(p)
    v = torch.zeros_like(p)
    for step in range(1, 6):
        grad = torch.randn(n, generator=g).to(dev) * (0.1 if step % 2 else 3.0)
        lr = 1e-2 * (0.5 + 0.1 * step)
        for gp in opt.param_groups:
            gp["lr"] = lr
        ref.grad = grad.clone()
        opt.step()
        call("sn_adam_step", ptr(p), ptr(grad), ptr(m), ptr(v), n, lr, None, 0.9, 0.999, 1e-8,
             step, None, stream_of(dev))
        torch.testing.assert_close(p, ref.detach(), rtol=1e-5, atol=1e-6)
    st = opt.state[ref]
    torch.testing.assert_close(m, st["exp_avg"], rtol=1e-5, atol=1e-7)
    torch.testing.assert_close(v, st["exp_avg_sq"], rtol=1e-5, atol=1e-9)


def test_adam_device_lr_and_step(dev):
    """lr and the step count read from device memory (graph replay) give the
    host-argument result."""
    from supernormal_amd._lib import call, ptr, stream_of
    n = 1000
    p = torch.ones(n, device=dev)
    q = torch.ones(n, device=dev)
    m1, v1, m2, v2 = (torch.zeros(n, device=dev) for _ in range(4))
    lr = torch.tensor([0.03], device=dev)
    step = torch.zeros(1, dtype=torch.int32, device=dev)
    for t in range(1, 4):
        grad = torch.full((n,), 0.5 * t, device=dev)
        step += 1
        call("sn_adam_step", ptr(p), ptr(grad), ptr(m1), ptr(v1), n, 0.0, ptr(lr), 0.9, 0.999,
             1e-8, 0, ptr(step), stream_of(dev))
        call("sn_adam_step", ptr(q), ptr(grad), ptr(m2), ptr(v2), n, 0.03, None, 0.9, 0.999,
             1e-8, t, None, stream_of(dev))
        torch.testing.assert_close(p, q, rtol=2e-7, atol=0)
