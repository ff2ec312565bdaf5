"""Hand-written adjoint of the LBM step (zf_adjoint_*) against central finite
differences, through the C ABI.

Objective: time-averaged |wall shear stress| over a wall patch of a tube
with a pulsatile velocity inlet and a pressure outlet.  Variables: inlet
amplitude and mean velocity, uniform body force, initial populations and a
per-node force field.  FP64 storage: gradients agree with central
differences (h = 1e-4 relative for parameters) to 1e-6 relative.  FP32
storage: the FP32 adjoint agrees with the FP64 finite-difference gradient
to 1e-4 relative.  The forward objective agrees with the oracle's LBM step
and WSS operator to 1e-12.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from zebraflow_amd import geometry
from tests.test_gpu_wss import node_normals
from zebraflow_amd.engine import (B_PRESSURE_OUTLET, B_VELOCITY_INLET, STORE_F32DEV, STORE_F64,
                                  Engine, lattice_config)

pytestmark = pytest.mark.gpu

U0, AMP, PERIOD, G = 0.02, 0.3, 40.0, (2e-6, 1e-6, 0.0)
N, T0, T1, MU = 50, 20, 50, 0.1


def case(store, checkpoint=8, field=False):
    b = geometry.tube((24, 14, 14), axis=0, r=5.0, inlet=B_VELOCITY_INLET, outlet=B_PRESSURE_OUTLET)
    cfg = lattice_config(*b.shape, 0.8, storage=store, inlet_u0=U0, inlet_amp=AMP,
                         inlet_period=PERIOD, inlet_dir=(1, 0, 0), body_force=G)
    eng = Engine(cfg)
    eng.upload_lattice(b)
    wx, wt, wn = geometry.wall_mesh(b)
    nrm = node_normals(wx, wt, wn)
    patch = (wx[:, 0] > 8) & (wx[:, 0] < 16)
    rng = np.random.default_rng(11)
    domid, dl = po.dom_setup(b)
    n = len(dl)
    f0 = po.feq(1 + 0.001 * rng.standard_normal(n), 0.005 * rng.standard_normal((3, n)))
    force = 1e-5 * rng.standard_normal((3, n)) if field else None
    if field:
        eng.upload_force(force)
    eng.adjoint_setup(wx[patch], nrm[patch], MU, T0, T1, N, checkpoint)
    return b, eng, f0, force, (wx[patch], nrm[patch])


def objective(eng, f0, u0=U0, amp=AMP, g=G, force=None):
    eng.upload_populations(f0)
    if force is not None:
        eng.upload_force(force)
    eng.set_forcing(u0, amp, g)
    return eng.adjoint_forward(0, N)


def fd(fun, x, h):
    return (fun(x + h) - fun(x - h)) / (2 * h)


def test_forward_objective_matches_oracle():
    print("case adjoint forward objective", flush=True)
    b, eng, f0, _, (px, pn) = case(STORE_F64)
    J = objective(eng, f0)
    o = po.LbmOracle(b, dict(wall_mode=1, visc_scale=0.3, force_dt=1.0, inlet_u0=U0, inlet_amp=AMP,
                             inlet_period=PERIOD, inlet_dir=(1, 0, 0), outlet_rho=1.0,
                             body_force=G))
    o.set_populations(f0)
    p = po.ibm_params(b.shape)
    acc = 0.0
    for k in range(1, N + 1):
        o.step(k)
        if T0 <= k <= T1:
            w, _, _ = po.wss(p, b, o.domid, (o.D, o.U, o.V, o.W), px, pn, MU)
            acc += np.linalg.norm(w, axis=1).sum()
    ref = acc / ((T1 - T0 + 1) * len(px))
    assert ref > 0
    np.testing.assert_allclose(J, ref, rtol=1e-12)
    eng.close()


def test_parameter_gradients_match_finite_differences_fp64():
    print("case adjoint parameters fp64", flush=True)
    _, eng, f0, _, _ = case(STORE_F64)
    objective(eng, f0)
    grad = eng.adjoint_backward()
    fds = [fd(lambda a: objective(eng, f0, amp=a), AMP, 1e-4 * AMP),
           fd(lambda u: objective(eng, f0, u0=u), U0, 1e-4 * U0)]
    # body force: J responds weakly (dJ/dg_y, dJ/dg_z ~ 1e-4 dJ/dg_x), so
    # the step is 1 % of |g| to keep the difference above FP64 noise in J
    h = 1e-2 * abs(G[0])
    for a in range(3):
        fds.append(fd(lambda x: objective(eng, f0, g=tuple(x if i == a else G[i] for i in range(3))),
                      G[a], h))
    fds = np.array(fds)
    print("adjoint", grad, "fd", fds, flush=True)
    np.testing.assert_allclose(grad, fds, rtol=1e-6, atol=1e-9 * np.abs(fds).max())
    eng.close()


def test_initial_state_gradient_and_checkpoint_independence():
    print("case adjoint initial state, checkpoints 8 and 50", flush=True)
    _, eng, f0, _, _ = case(STORE_F64, checkpoint=8)
    objective(eng, f0)
    grad8 = eng.adjoint_backward()
    lam8, _ = eng.adjoint_download()
    rng = np.random.default_rng(5)
    for trial in range(2):
        d = rng.standard_normal(f0.shape)
        d -= d.mean(0)  # mass-neutral perturbation of every node
        dd = fd(lambda t: objective(eng, f0 + t * d), 0.0, 1e-6)
        np.testing.assert_allclose((lam8 * d).sum(), dd, rtol=1e-6)
    _, eng2, _, _, _ = case(STORE_F64, checkpoint=50)
    objective(eng2, f0)
    grad50 = eng2.adjoint_backward()
    lam50, _ = eng2.adjoint_download()
    np.testing.assert_array_equal(grad8, grad50)
    np.testing.assert_array_equal(lam8, lam50)
    eng.close()
    eng2.close()


def test_force_field_gradient():
    print("case adjoint force field", flush=True)
    _, eng, f0, force, _ = case(STORE_F64, field=True)
    objective(eng, f0, force=force)
    eng.adjoint_backward()
    _, gF = eng.adjoint_download()
    rng = np.random.default_rng(8)
    d = rng.standard_normal(force.shape)
    dd = fd(lambda t: objective(eng, f0, force=force + t * d), 0.0, 1e-7)
    np.testing.assert_allclose((gF * d).sum(), dd, rtol=1e-6)
    eng.close()


def test_fp32_adjoint_matches_fp64_gradient():
    print("case adjoint fp32 vs fp64", flush=True)
    _, e64, f0, _, _ = case(STORE_F64)
    objective(e64, f0)
    g64 = e64.adjoint_backward()
    _, e32, _, _, _ = case(STORE_F32DEV)
    J32 = objective(e32, f0)
    g32 = e32.adjoint_backward()
    J64 = objective(e64, f0)
    np.testing.assert_allclose(J32, J64, rtol=1e-5)
    np.testing.assert_allclose(g32, g64, rtol=1e-4, atol=1e-4 * np.abs(g64).max())
    e64.close()
    e32.close()


def test_torch_autograd_function():
    """The adjoint as a torch.autograd.Function of (initial populations,
    inlet amplitude), checked against the direct adjoint calls."""
    torch = pytest.importorskip("torch")
    print("case adjoint autograd", flush=True)
    _, eng, f0, _, _ = case(STORE_F64)

    class WssObjective(torch.autograd.Function):
        @staticmethod
        def forward(ctx, f, amp):
            return torch.tensor(objective(eng, f.detach().cpu().numpy(), amp=float(amp)),
                                dtype=torch.float64)

        @staticmethod
        def backward(ctx, gout):
            g = eng.adjoint_backward()
            lam, _ = eng.adjoint_download()
            return (gout * torch.from_numpy(lam), gout * torch.tensor(g[0], dtype=torch.float64))

    f = torch.tensor(f0, dtype=torch.float64, requires_grad=True)
    amp = torch.tensor(AMP, dtype=torch.float64, requires_grad=True)
    J = WssObjective.apply(f, amp)
    (2.0 * J).backward()
    objective(eng, f0)
    g = eng.adjoint_backward()
    lam, _ = eng.adjoint_download()
    np.testing.assert_allclose(amp.grad.item(), 2.0 * g[0], rtol=1e-14)
    np.testing.assert_allclose(f.grad.numpy(), 2.0 * lam, rtol=1e-14, atol=0)
    eng.close()
