"""Adjoint through the RBC iterations (zf_adjoint_* with cells) against
central finite differences, through the C ABI.

A tube with a pulsatile velocity inlet and two pre-stretched RBCs (162-node
meshes, lattice-unit moduli), membrane/IBM every second iteration.
Objective: time-averaged |WSS| over a wall patch.  Variables: inlet
amplitude, shear modulus es and bending modulus kb (the lattice-unit
set-up rebuilt with the perturbed modulus), initial cell positions.

The forward pass in adjoint mode is zf_step's iteration (checked against
zf_step itself).  FP64 storage (FP64 throughout, IBM included): gradients
against central differences, 1e-4 relative.  FP32 storage against the
FP64 gradients.

The forward map has jump discontinuities, as the reference's has: the
interpolation and spreading stencils cut at |d| = 2 DX where the cosine
weight is not zero, and bending switches off above |cos dihedral| =
cos_cut.  A difference whose two runs fall on opposite sides of a jump
measures the jump, not the derivative.  The case therefore sets cos_cut
above every dihedral these cells reach (0.9999; the largest is 0.99965),
and the steps are small enough that no stencil cut is crossed: the inlet
amplitude at h = 1e-4 relative (the cells move by ~1e-5 lu), the moduli at
h = 1e-6 relative, the positions at 1e-8 lu.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from zebraflow_amd import geometry
from tests.test_gpu_wss import node_normals
from zebraflow_amd.engine import (B_PRESSURE_OUTLET, B_VELOCITY_INLET, STORE_F32DEV, STORE_F64,
                                  Engine, lattice_config, lattice_ibm_config,
                                  lattice_membrane_config)

pytestmark = pytest.mark.gpu

U0, AMP, PERIOD = 0.02, 0.3, 40.0
ES, KB = 5e-3, 2e-3
N, T0, T1, MU = 40, 10, 40, 0.1


class Case:
    def __init__(self, store=STORE_F64, es=ES, kb=KB, checkpoint=8, interval=2):
        b = geometry.tube((32, 20, 20), axis=0, r=8.0, inlet=B_VELOCITY_INLET,
                          outlet=B_PRESSURE_OUTLET)
        cfg = lattice_config(*b.shape, 0.8, storage=store, inlet_u0=U0, inlet_amp=AMP,
                             inlet_period=PERIOD, inlet_dir=(1, 0, 0))
        eng = Engine(cfg)
        eng.upload_lattice(b)
        xyz, tris = geometry.rbc_mesh(radius=3.0, level=2)
        cells = geometry.place_cells(b, xyz, 3.0, 1.0, 2, 3)
        assert len(cells) == 2
        # stretched about their centroids: membrane forces act from the start
        c = cells.mean(1, keepdims=True)
        self.x0 = c + (cells - c) * np.array([1.06, 0.96, 1.0])
        mc = lattice_membrane_config(xyz, tris, 2, es=es, kb=kb)
        mc.cos_cut = 0.9999
        eng.membrane_setup(mc, tris, xyz)
        eng.upload_cells(self.x0)
        eng.ibm_setup(lattice_ibm_config(rbc_interval=interval))
        wx, wt, wn = geometry.wall_mesh(b)
        nrm = node_normals(wx, wt, wn)
        patch = (wx[:, 0] > 8) & (wx[:, 0] < 24)
        eng.adjoint_setup(wx[patch], nrm[patch], MU, T0, T1, N, checkpoint, cells=True)
        rng = np.random.default_rng(11)
        _, dl = po.dom_setup(b)
        n = len(dl)
        self.f0 = po.feq(1 + 0.001 * rng.standard_normal(n), 0.005 * rng.standard_normal((3, n)))
        self.zero = np.zeros((3, n))
        self.b, self.eng, self.probes = b, eng, (wx[patch], nrm[patch])

    def reset(self, amp=AMP, x0=None):
        e = self.eng
        e.upload_populations(self.f0)
        e.upload_force(self.zero)
        e.upload_cells(self.x0 if x0 is None else x0)
        e.set_forcing(U0, amp)

    def J(self, amp=AMP, x0=None):
        self.reset(amp, x0)
        return self.eng.adjoint_forward(0, N)

    def grads(self):
        g = self.eng.adjoint_backward()
        dp, gx = self.eng.adjoint_cell_gradients()
        return g, dp, gx

    def close(self):
        self.eng.close()


def fd(fun, x, h):
    return (fun(x + h) - fun(x - h)) / (2 * h)


def test_recorded_forward_is_zf_step():
    print("case cells adjoint forward vs zf_step", flush=True)
    cs = Case()
    J = cs.J()
    xa, _, _ = cs.eng.cells()
    cs.reset()
    px, pn = cs.probes
    acc = 0.0
    for k in range(1, N + 1):
        cs.eng.full_step(k)
        if T0 <= k <= T1:
            w, _, _ = cs.eng.wss(px, pn, MU)
            acc += np.linalg.norm(w, axis=1).sum()
    xb, _, _ = cs.eng.cells()
    np.testing.assert_array_equal(xa, xb)
    np.testing.assert_allclose(J, acc / ((T1 - T0 + 1) * len(px)), rtol=1e-12)
    assert np.abs(xa - cs.x0).max() > 1e-3  # the cells moved
    cs.close()


def test_parameter_gradients_fp64():
    print("case cells adjoint parameters fp64", flush=True)
    cs = Case()
    J0 = cs.J()
    g, dp, _ = cs.grads()
    fd_amp = fd(lambda a: cs.J(amp=a), AMP, 1e-4 * AMP)
    m = cs.eng.membrane()
    assert (m["guard_fail"] == 0).all()

    def J_mod(es, kb):
        c = Case(es=es, kb=kb)
        v = c.J()
        c.close()
        return v

    fds = []
    for h in (1e-4, 1e-6):
        fds.append(((J_mod(ES * (1 + h), KB) - J_mod(ES * (1 - h), KB)) / (2 * h),
                    (J_mod(ES, KB * (1 + h)) - J_mod(ES, KB * (1 - h))) / (2 * h)))
    print("J", J0, "amp", g[0], fd_amp, "ln es, ln kb", dp, "fd h=1e-4", fds[0], "h=1e-6", fds[1],
          flush=True)
    np.testing.assert_allclose(g[0], fd_amp, rtol=1e-4)
    np.testing.assert_allclose(dp, fds[1], rtol=1e-4)
    cs.close()


def test_position_gradient_fp64():
    print("case cells adjoint initial positions fp64", flush=True)
    cs = Case()
    cs.J()
    _, _, gx = cs.grads()
    rng = np.random.default_rng(2)
    for trial in range(2):
        d = rng.standard_normal(cs.x0.shape)
        dd = fd(lambda t: cs.J(x0=cs.x0 + t * d), 0.0, 1e-8)
        print("x0 direction", (gx * d).sum(), dd, flush=True)
        np.testing.assert_allclose((gx * d).sum(), dd, rtol=1e-4)
    cs.close()


def test_checkpoint_interval_does_not_change_the_gradient():
    print("case cells adjoint checkpoints 1 and 40", flush=True)
    out = []
    for K in (1, 40):
        cs = Case(checkpoint=K)
        cs.J()
        g, dp, gx = cs.grads()
        lam, gF = cs.eng.adjoint_download()
        out.append((g, dp, gx, lam, gF))
        cs.close()
    for a, b in zip(*out):
        np.testing.assert_array_equal(a, b)


def test_fp32_matches_fp64_gradient():
    print("case cells adjoint fp32 vs fp64", flush=True)
    res = {}
    for store in (STORE_F64, STORE_F32DEV):
        cs = Case(store=store)
        J = cs.J()
        g, dp, gx = cs.grads()
        res[store] = (J, g, dp, gx)
        cs.close()
    (J64, g64, dp64, gx64), (J32, g32, dp32, gx32) = res[STORE_F64], res[STORE_F32DEV]
    print("fp64", J64, g64[0], dp64, "fp32", J32, g32[0], dp32, "x0 rel",
          np.abs(gx32 - gx64).max() / np.abs(gx64).max(), flush=True)
    np.testing.assert_allclose(J32, J64, rtol=1e-5)
    np.testing.assert_allclose(g32[0], g64[0], rtol=1e-3)
    np.testing.assert_allclose(dp32, dp64, rtol=1e-3, atol=1e-3 * np.abs(dp64).max())
    np.testing.assert_allclose(gx32, gx64, rtol=1e-3, atol=1e-3 * np.abs(gx64).max())


def test_torch_autograd_function():
    """The adjoint as a torch.autograd.Function of (inlet amplitude, initial
    positions, es, kb), against the direct calls."""
    torch = pytest.importorskip("torch")
    print("case cells adjoint autograd", flush=True)
    cs = Case()

    class CellWss(torch.autograd.Function):
        @staticmethod
        def forward(ctx, amp, x0, es, kb):
            assert float(es) == ES and float(kb) == KB  # the case is built for these moduli
            return torch.tensor(cs.J(amp=float(amp), x0=x0.detach().numpy()), dtype=torch.float64)

        @staticmethod
        def backward(ctx, gout):
            g, dp, gx = cs.grads()
            t = lambda v: torch.tensor(v, dtype=torch.float64)
            return gout * t(g[0]), gout * torch.from_numpy(gx), gout * t(dp[0] / ES), gout * t(dp[1] / KB)

    amp = torch.tensor(AMP, dtype=torch.float64, requires_grad=True)
    x0 = torch.tensor(cs.x0, dtype=torch.float64, requires_grad=True)
    es = torch.tensor(ES, dtype=torch.float64, requires_grad=True)
    kb = torch.tensor(KB, dtype=torch.float64, requires_grad=True)
    (3.0 * CellWss.apply(amp, x0, es, kb)).backward()
    cs.J()
    g, dp, gx = cs.grads()
    np.testing.assert_allclose(amp.grad.item(), 3.0 * g[0], rtol=1e-14)
    np.testing.assert_allclose(es.grad.item(), 3.0 * dp[0] / ES, rtol=1e-14)
    np.testing.assert_allclose(kb.grad.item(), 3.0 * dp[1] / KB, rtol=1e-14)
    np.testing.assert_allclose(x0.grad.numpy(), 3.0 * gx, rtol=1e-14)
    cs.close()
