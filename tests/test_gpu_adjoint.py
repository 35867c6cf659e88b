"""Differentiable trajectory (N = 256 FENE/WCA/RPY chain, block-Lanczos
noise recorded once and held fixed, 200 EM steps, checkpoints every 20):
the hand-written adjoint gradient of the final Rg^2 with respect to
(LJ eps, LJ sigma, FENE k) against central finite differences (h = 1e-5) of
(a) the GPU forward and (b) the FP64 CPU oracle forward with the same
replayed displacements."""
import numpy as np
import pytest

from helpers import fene_wca_chain, gpu_system, rel

pytestmark = pytest.mark.gpu

NB, NSTEPS, S, CKPT, H = 256, 200, 4, 20, 1e-5
THETA0 = np.array([1.0, 2.0, 7.5])


@pytest.fixture(scope="module")
def c5():
    m = fene_wca_chain(NB, 256)
    g = gpu_system(m, mob_fctr="LANC", nts_mobsam=S, lanc_mxitr=50, lanc_tol=1e-6, lanc_reorth=True)
    g.set_rng(2025)
    disp = g.run_record(NSTEPS)
    return m, g, disp


def _gpu_loss(m, g, disp, theta):
    g.set_coordinates(m["crd"])
    g.reset_run()
    g.set_theta(*theta)
    g.run_replay(disp)
    return g.rgsq()


def _oracle_loss(oracle, m, disp, theta):
    eps, sig, k = theta
    so = oracle.System(NB, m["bonds"], [2], [[k, 3.0, 0.0]], vdw_styles=[1],
                       vdw_params=[[eps, sig, 2.0 ** (1.0 / 6.0) * sig]], rcutoff=2.0 ** (1.0 / 6.0) * sig)
    out = oracle.bd_run_replay(so, m["crd"], disp, lhdia=True, nts_mobsam=S)
    return oracle.rgsq(out)


def test_adjoint_matches_gpu_finite_differences(c5):
    m, g, disp = c5
    assert np.abs(disp).max() > 0
    g.set_coordinates(m["crd"])
    g.reset_run()
    g.set_theta(*THETA0)
    loss, grad = g.adjoint_rgsq(disp, ckpt=CKPT)
    assert abs(loss - _gpu_loss(m, g, disp, THETA0)) < 1e-12 * loss
    fd = np.zeros(3)
    for a in range(3):
        tp, tm = THETA0.copy(), THETA0.copy()
        tp[a] += H
        tm[a] -= H
        fd[a] = (_gpu_loss(m, g, disp, tp) - _gpu_loss(m, g, disp, tm)) / (2 * H)
    print("adjoint", grad, "fd", fd, flush=True)
    assert rel(grad, fd) < 1e-5


def test_adjoint_matches_oracle_fp64_gradient(oracle, c5):
    m, g, disp = c5
    g.set_coordinates(m["crd"])
    g.reset_run()
    g.set_theta(*THETA0)
    loss, grad = g.adjoint_rgsq(disp, ckpt=CKPT)
    assert abs(loss - _oracle_loss(oracle, m, disp, THETA0)) < 1e-9 * loss
    fd = np.zeros(3)
    for a in range(3):
        tp, tm = THETA0.copy(), THETA0.copy()
        tp[a] += H
        tm[a] -= H
        fd[a] = (_oracle_loss(oracle, m, disp, tp) - _oracle_loss(oracle, m, disp, tm)) / (2 * H)
    print("adjoint", grad, "oracle fd", fd, flush=True)
    assert rel(grad, fd) < 1e-5


def test_autograd_function(c5):
    import torch
    from brownpak_amd.autograd import rgsq_trajectory
    m, g, disp = c5
    theta = torch.tensor(THETA0, dtype=torch.float64, requires_grad=True)
    loss = rgsq_trajectory(theta, g, m["crd"], disp, ckpt=CKPT)
    loss.backward()
    g.set_coordinates(m["crd"])
    g.reset_run()
    g.set_theta(*THETA0)
    _, grad = g.adjoint_rgsq(disp, ckpt=CKPT)
    assert np.allclose(theta.grad.numpy(), grad, rtol=1e-12, atol=0)


# ---------------------------------------------------------------- FP32
# The configuration as specified: FP32 mobility products (drift D.F and the
# adjoint's D.mu), FP64 Lanczos tridiagonal, FP64 coordinates and adjoint
# accumulation.  Bounds: the FP32 products carry ~1e-7 relative error, which
# the 200-step trajectory carries into the loss (< 1e-5 relative against the
# oracle's FP64 replay) and the gradient (< 1e-4 against the oracle's FP64
# central differences).  Central differences of the FP32 forward itself
# need a step large enough that FP32 rounding of the loss (~1e-7) does not
# dominate: h = 1e-3 gives about 1e-4 relative from rounding and 1e-6 from
# truncation; bound 1e-3.
H32 = 1e-3


@pytest.fixture(scope="module")
def c5_fp32():
    m = fene_wca_chain(NB, 256)
    g = gpu_system(m, mob_fctr="LANC", nts_mobsam=S, lanc_mxitr=50, lanc_tol=1e-6, lanc_reorth=True,
                   precision="fp32")
    g.set_rng(2025)
    disp = g.run_record(NSTEPS)
    return m, g, disp


def test_adjoint_fp32_matches_oracle_fp64_gradient(oracle, c5_fp32):
    m, g, disp = c5_fp32
    g.set_coordinates(m["crd"])
    g.reset_run()
    g.set_theta(*THETA0)
    loss, grad = g.adjoint_rgsq(disp, ckpt=CKPT)
    lref = _oracle_loss(oracle, m, disp, THETA0)
    fd = np.zeros(3)
    for a in range(3):
        tp, tm = THETA0.copy(), THETA0.copy()
        tp[a] += H
        tm[a] -= H
        fd[a] = (_oracle_loss(oracle, m, disp, tp) - _oracle_loss(oracle, m, disp, tm)) / (2 * H)
    print(f"fp32 loss rel {abs(loss - lref) / lref:.2e}; adjoint {grad} oracle fd {fd} rel {rel(grad, fd):.2e}",
          flush=True)
    assert abs(loss - lref) < 1e-5 * lref
    assert rel(grad, fd) < 1e-4


def test_adjoint_fp32_matches_fp32_finite_differences(c5_fp32):
    m, g, disp = c5_fp32
    g.set_coordinates(m["crd"])
    g.reset_run()
    g.set_theta(*THETA0)
    loss, grad = g.adjoint_rgsq(disp, ckpt=CKPT)
    fd = np.zeros(3)
    for a in range(3):
        tp, tm = THETA0.copy(), THETA0.copy()
        tp[a] += H32
        tm[a] -= H32
        fd[a] = (_gpu_loss(m, g, disp, tp) - _gpu_loss(m, g, disp, tm)) / (2 * H32)
    print(f"fp32 adjoint {grad} fp32 fd {fd} rel {rel(grad, fd):.2e}", flush=True)
    assert rel(grad, fd) < 1e-3
