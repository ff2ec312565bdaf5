"""Oracle LBM step (a4-a6) pinned by what the mathematics fixes.

The reference holds no known-answer test for its time step, so the oracle's
restatement of ZebraflowWT.c:6101-6145/:6292-6433/:6601-6626 is pinned here by
closed forms and invariants chosen so that a dropped term, a wrong sign or a
wrong index fails one of them:
  - rest and uniform-flow equilibria are fixed points (feq, pull index);
  - total momentum grows by exactly N*DT*F per step (Guo source and shift);
  - mass is conserved with bounce-back walls (BB index/direction);
  - a shear wave decays at exp(-nu k^2 t), nu = (tau-1/2)/3 (collision);
  - plane Poiseuille flow reaches g H^2/(8 nu) (walls + forcing, 2nd order).
"""
import numpy as np
import pytest

from oracle import pyoracle as po


def periodic_box(shape, b=1):
    return np.full(shape, b, np.int32)


def lattice_params(tau, **kw):
    p = dict(visc_scale=tau - 0.5, force_dt=1.0, wall_mode=1, outlet_rho=1.0,
             inlet_period=1.0, t_scale=5e-7, pressure_scale=10000.0 / 3)
    p.update(kw)
    return p


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_rest_and_uniform_flow_are_fixed_points(dtype):
    b = periodic_box((6, 5, 7))
    for u0 in [(0.0, 0.0, 0.0), (0.03, -0.01, 0.02)]:
        o = po.LbmOracle(b, lattice_params(0.8, periodic=(1, 1, 1)), dtype=dtype)
        n = o.n
        f0 = po.feq(np.full(n, 1.02), np.tile(np.array(u0)[:, None], (1, n)))
        o.set_populations(f0)
        for it in range(1, 6):
            o.step(it)
        tol = 1e-14 if dtype == np.float64 else 2e-7
        np.testing.assert_allclose(o.populations, f0, rtol=0, atol=tol)


def test_momentum_grows_by_dt_force():
    b = periodic_box((8, 6, 5))
    g = np.array([2e-5, -1e-5, 3e-6])
    for dt in (1.0, 0.5):
        o = po.LbmOracle(b, lattice_params(0.9, periodic=(1, 1, 1), body_force=g, force_dt=dt))
        rng = np.random.default_rng(3)
        u = 0.01 * rng.standard_normal((3, o.n))
        o.set_populations(po.feq(1 + 0.01 * rng.standard_normal(o.n), u))
        mom0 = po.E[:, :3].T.astype(float) @ o.populations.sum(1)
        mass0 = o.populations.sum()
        steps = 7
        for it in range(1, steps + 1):
            o.step(it)
        mom = po.E[:, :3].T.astype(float) @ o.populations.sum(1)
        np.testing.assert_allclose(mom - mom0, steps * o.n * dt * g, rtol=1e-9, atol=1e-13)
        assert abs(o.populations.sum() - mass0) < 1e-11


def test_mass_conserved_with_bounce_back():
    from zebraflow_amd.geometry import tube
    b = tube((6, 18, 18), axis=0, r=6.5, periodic=True)
    o = po.LbmOracle(b, lattice_params(0.8, periodic=(1, 0, 0), body_force=(1e-4, 0, 0)))
    interior = (o.bnode() > 0) & (o.bnode() < 10)
    m0 = o.populations[:, interior].sum()
    for it in range(1, 41):
        o.step(it)
    assert abs(o.populations[:, interior].sum() - m0) < 1e-10 * m0
    # a planted mistake in the oracle's BB direction would break this
    assert np.all(np.isfinite(o.populations))


def test_shear_wave_decay_rate():
    L = 32
    b = periodic_box((4, L, 3))
    tau = 0.8
    nu = (tau - 0.5) / 3
    o = po.LbmOracle(b, lattice_params(tau, periodic=(1, 1, 1)))
    j = o.domlist[:, 1] - 1
    k = 2 * np.pi / L
    u0 = 1e-3
    u = np.zeros((3, o.n))
    u[0] = u0 * np.sin(k * j)
    o.set_populations(po.feq(np.ones(o.n), u))
    steps = 200
    for it in range(1, steps + 1):
        o.step(it)
    ux = (po.E[:, 0] @ o.populations) / o.populations.sum(0)
    amp = 2 * np.mean(ux * np.sin(k * j))
    # the first step relaxes the non-equilibrium part; compare decay rates
    expect = u0 * np.exp(-nu * k * k * steps)
    assert abs(amp / expect - 1) < 0.01, (amp, expect)


def test_plane_poiseuille_profile():
    H = 16                                   # fluid layers between the walls
    b = np.full((3, H + 2, 3), 1, np.int32)
    b[:, 0, :] = 0
    b[:, -1, :] = 0
    tau, g = 0.8, 1e-6
    nu = (tau - 0.5) / 3
    o = po.LbmOracle(b, lattice_params(tau, periodic=(1, 0, 1), body_force=(g, 0, 0)))
    for it in range(1, 6001):
        o.step(it)
    fluid = o.bnode() == 1
    y = (o.domlist[fluid, 1] - 1) - 0.5      # halfway wall at y = 0
    # velocity with the half-force shift
    rho = o.populations[:, fluid].sum(0)
    ux = ((po.E[:, 0] @ o.populations[:, fluid]) + 0.5 * g) / rho
    exact = g / (2 * nu) * y * (H - y)
    np.testing.assert_allclose(ux, exact, rtol=2e-3, atol=1e-3 * exact.max())


@pytest.mark.parametrize("code", range(1001, 1010))
def test_cardiac_pressure_waveforms_are_continuous_and_periodic(code):
    """ZebraflowWT.c:6318-6421: each waveform is a rise (sin^2 to 23 % of the
    0.36 s cycle, or sin to 30 %) joined to a cosine fall; the pieces meet at
    the peak and at the cycle end, so the pressure is continuous, periodic
    and spans [120 + base, 120 + base + amp] mmHg-scaled units.  A wrong rise
    fraction or fall phase breaks one of the joins."""
    T = 0.36
    t = np.linspace(0.0, 3 * T, 200001)
    p = np.array([po.lib().zfo_bc_pressure(code, x) for x in t])
    dp = np.abs(np.diff(p))
    assert dp.max() < 1e-2 * (p.max() - p.min())  # no jumps at 1.6e-6 s sampling
    np.testing.assert_allclose(
        [po.lib().zfo_bc_pressure(code, x + T) for x in t[:2000:97]],
        p[:2000:97], rtol=0, atol=1e-9)
    lo, hi = p.min(), p.max()
    assert lo >= 120.0 and hi > lo
    # the minimum is reached at the cycle boundary and the peak once per cycle
    peaks = np.nonzero((p[1:-1] > p[:-2]) & (p[1:-1] >= p[2:]))[0]
    assert len(peaks) == 3
