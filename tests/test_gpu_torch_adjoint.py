"""The differentiable CG (torch.autograd.Function over the C ABI): the adjoint
gradient against central finite differences (FP64) and against the FP64 CPU
reference gradient built from the oracle (forward CG, adjoint CG, closed-form
edge gradient dL/dtheta_e = -(lam_i - lam_j)(x_i - x_j))."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup(g, B, dtype, seed=0):
    import torch
    from cumat_amd.problems import grid_laplacian
    from cumat_amd.torch_cg import EdgeLaplacian
    rng = np.random.default_rng(seed)
    P = grid_laplacian((g, g), lambda s: np.ones(s))
    lap = EdgeLaplacian(P, "cuda")
    theta = np.exp(0.5 * rng.standard_normal((B, P["nedges"])))  # theta = exp(N(0, 0.5^2))
    b = rng.uniform(-1, 1, (B, P["n"]))
    xs = rng.standard_normal((B, P["n"]))
    T = lambda a: torch.tensor(a, dtype=dtype, device="cuda")  # noqa: E731
    return P, lap, T(theta), T(b), T(xs), theta, b, xs


def _ref_grad(orc, P, theta, b, xs):
    """FP64 CPU reference: forward CG, adjoint CG, closed-form edge gradient."""
    from cumat_amd.problems import grid_laplacian
    grads, lams, xsol = [], [], []
    for k in range(theta.shape[0]):
        it = iter([theta[k]])
        off = {"o": 0}

        def sampler(shape, th=theta[k]):
            m = int(np.prod(shape))
            out = th[off["o"]:off["o"] + m].reshape(shape)
            off["o"] += m
            return out
        Pk = grid_laplacian(tuple(int(np.sqrt(P["n"])) for _ in range(2)), sampler)
        x, _, _ = orc.pcg(Pk["JA"], Pk["IA"], Pk["vals"], b[k], P["n"], tol=1e-14, max_iter=20000)
        lam, _, _ = orc.pcg(Pk["JA"], Pk["IA"], Pk["vals"], x - xs[k], P["n"], tol=1e-14, max_iter=20000)
        grads.append(orc.edge_grad(P["ei"], P["ej"], lam, x))
        lams.append(lam)
        xsol.append(x)
        del it
    return np.stack(grads), np.stack(lams), np.stack(xsol)


def test_adjoint_fp64_vs_fd_and_reference(gpu, orc):
    import torch
    from cumat_amd.torch_cg import cg_solve
    P, lap, theta, b, xs, th_np, b_np, xs_np = _setup(24, 3, torch.float64)
    theta.requires_grad_(True)
    b.requires_grad_(True)
    x = cg_solve(theta, b, lap, tol=1e-13, max_iter=20000)
    loss = 0.5 * ((x - xs) ** 2).sum()
    loss.backward()
    g_theta = theta.grad.detach().cpu().numpy()
    g_b = b.grad.detach().cpu().numpy()
    ref, lam_ref, x_ref = _ref_grad(orc, P, th_np, b_np, xs_np)
    assert np.linalg.norm(x.detach().cpu().numpy() - x_ref) / np.linalg.norm(x_ref) < 1e-10
    assert np.linalg.norm(g_theta - ref) / np.linalg.norm(ref) < 1e-8
    assert np.linalg.norm(g_b - lam_ref) / np.linalg.norm(lam_ref) < 1e-8
    # central differences on 32 random directions, FP64
    rng = np.random.default_rng(99)

    def L(th):
        with torch.no_grad():
            xx = cg_solve(th, b.detach(), lap, tol=1e-14, max_iter=20000)
            return float(0.5 * ((xx - xs) ** 2).sum())
    worst = 0.0
    for _ in range(32):
        d = torch.tensor(rng.standard_normal(th_np.shape), dtype=torch.float64, device="cuda")
        h = 1e-4
        fd = (L(theta.detach() + h * d) - L(theta.detach() - h * d)) / (2 * h)
        an = float((theta.grad * d).sum())
        worst = max(worst, abs(fd - an) / max(abs(an), 1e-12))
    assert worst < 1e-6, worst


def test_adjoint_fp32_against_fp64_reference(gpu, orc):
    """FP32 forward/adjoint (tol 1e-7): gradient within 2e-3 of the FP64
    reference (FP32 CG cannot resolve the iterate below ~1e-6 relative)."""
    import torch
    from cumat_amd.torch_cg import cg_solve
    P, lap, theta, b, xs, th_np, b_np, xs_np = _setup(32, 4, torch.float32, seed=3)
    theta.requires_grad_(True)
    x = cg_solve(theta, b, lap, tol=1e-7, max_iter=20000)
    loss = 0.5 * ((x - xs) ** 2).sum()
    loss.backward()
    ref, _, _ = _ref_grad(orc, P, th_np.astype(np.float32).astype(np.float64), b_np.astype(np.float32).astype(np.float64),
                          xs_np.astype(np.float32).astype(np.float64))
    g = theta.grad.detach().cpu().numpy().astype(np.float64)
    assert np.linalg.norm(g - ref) / np.linalg.norm(ref) < 2e-3
