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
    # one cycle later, to the float resolution of the time (about 3e-8 s here)
    np.testing.assert_allclose(
        [po.lib().zfo_bc_pressure(code, x + T) for x in t[:2000:97]],
        p[:2000:97], rtol=0, atol=2e-3)
    lo, hi = p.min(), p.max()
    assert lo >= 120.0 and hi > lo
    # the minimum is reached at the cycle boundary and the peak once per cycle
    # (float times and pressures make the crest a plateau of equal values:
    # count the crests as runs of samples within 1e-3 of the peak value)
    top = np.nonzero(p >= hi - 1e-3 * (hi - lo))[0]
    crests = 1 + int((np.diff(t[top]) > T / 4).sum())
    assert crests == 3


def test_cardiac_pressure_waveforms_match_the_reference_bit_for_bit():
    """The oracle's waveforms against the reference's own branch chain
    (ZebraflowWT.c:6318-6421, compiled from the reference file and evaluated
    with its float variables by oracle/_ref/ref_waveform; vectors in
    tests/golden/waveforms_ref.npz): every code at 462 iterations over more
    than a cycle, its branch joins and wraps, and late float times.  The
    pressure is bit-identical, and so is D_bc = pressure/LBM_PRESSURE_SCALE + 1
    rounded to float as the reference stores it."""
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "waveforms_ref.npz"))
    assert set(np.unique(g["code"]).tolist()) == set(range(1001, 1010))
    t_scale, pscale = 0.5 * 1e-6, (1.0 / 3.0) * 10.0 * 1000.0  # T_SCALE, LBM_PRESSURE_SCALE
    got = np.array([po.lib().zfo_bc_pressure(int(c), float(np.float32(int(it) * t_scale)))
                    for it, c in zip(g["iter"], g["code"])], np.float32)
    np.testing.assert_array_equal(got, g["pressure"])
    d_bc = (got.astype(np.float64) / pscale + 1.0).astype(np.float32)
    np.testing.assert_array_equal(d_bc, g["d_bc"])
    # the samples reach both branches of every shape and the peak values
    for c in range(1001, 1010):
        p = g["pressure"][g["code"] == c]
        base, amp = {1001: (54, 82), 1002: (48, 68), 1003: (11, 23), 1004: (18, 27),
                     1005: (23, 27), 1006: (29, 66), 1007: (28.5, 60), 1008: (29, 66),
                     1009: (28.5, 60)}[c]
        assert p.min() >= 120 + base - 1e-3 and abs(p.max() - (120 + base + amp)) < 0.05


# ---- the main lattice's boundary pass (ZebraflowWT.c:6292-6433) -------------

def _tier_neighbour(b, i, j, k):
    """BCNLIST of one BC node by the reference's tier rule, from the B field
    alone (LBM_BC_SETUP, ZebraflowWT.c:1506-1589): the first of c = 1..18
    whose node is interior (0 < B < 1000), else the first with B > 1000, else
    the first wall (B == 0); None without any."""
    im, jm, km = b.shape
    for test in (lambda v: 0 < v < 1000, lambda v: v > 1000, lambda v: v == 0):
        for e in po.E[:18]:
            ii, jj, kk = i + e[0], j + e[1], k + e[2]
            if 1 <= ii <= im and 1 <= jj <= jm and 1 <= kk <= km and test(b[ii - 1, jj - 1, kk - 1]):
                return ii, jj, kk
    return None


def _smooth_fields(shape, seed):
    """A density and a velocity field over the box, varying at every node."""
    ii, jj, kk = np.meshgrid(*[np.arange(1, n + 1) for n in shape], indexing="ij")
    ph = np.random.default_rng(seed).uniform(0, 2 * np.pi, 8)
    rho = 1 + 0.01 * np.sin(0.9 * ii + ph[0]) + 0.02 * np.cos(0.6 * jj + ph[1]) \
        + 0.015 * np.sin(0.8 * kk + ph[2])
    u = np.stack([0.01 * np.sin(0.5 * jj + ph[3]) + 0.004 * np.cos(0.7 * kk),
                  0.008 * np.cos(0.4 * ii + ph[4]) * np.sin(0.3 * kk + ph[5]),
                  0.006 * np.sin(0.35 * ii + 0.45 * jj + ph[6])])
    return rho, u


def _brute_force_bc(b, params, it, expect):
    """One oracle step from f = feq(rho(x), u(x)) at every listed node; every
    BC node with a neighbour must then hold feq(D_bc, U_bc) as `expect`
    returns it from (code, D_m, U_m) of its tier neighbour m, where D_m, U_m
    are recomputed here: for an interior m, the moments of its pulled
    populations sum_c feq_c(rho, u)(x_m - e_c) with the force shift DT F/2
    (ZebraflowWT.c:6113-6135); for a BC or wall m, the moments of its
    populations before the step (the previous macro pass, :6606-6614)."""
    o = po.LbmOracle(b, params)
    rho, u = _smooth_fields(b.shape, 5)
    d = o.domlist - 1
    o.set_populations(po.feq(rho[d[:, 0], d[:, 1], d[:, 2]], u[:, d[:, 0], d[:, 1], d[:, 2]]))
    o.step(it)
    # pulled moments over the box (sources of interior nodes are all listed)
    feq_box = po.feq(rho.ravel(), u.reshape(3, -1)).reshape((19,) + b.shape)
    D = np.zeros(b.shape)
    J = np.zeros((3,) + b.shape)
    for c in range(19):
        src = np.roll(feq_box[c], shift=tuple(po.E[c]), axis=(0, 1, 2))  # f_c(x - e_c)
        D += src
        for a in range(3):
            J[a] += po.E[c][a] * src
    g = np.asarray(params.get("body_force", (0, 0, 0)), float)
    U = (J + 0.5 * params["force_dt"] * g[:, None, None, None]) / D
    checked = {}
    for i, j, k in zip(*np.nonzero((b == 0) | (b >= 1000))):
        i, j, k = i + 1, j + 1, k + 1
        code = int(b[i - 1, j - 1, k - 1])
        m = _tier_neighbour(b, i, j, k)
        if m is None:
            continue
        mi = (m[0] - 1, m[1] - 1, m[2] - 1)
        if 0 < b[mi] < 10:
            dm, um = D[mi], U[(slice(None),) + mi]
        else:
            dm, um = rho[mi], u[(slice(None),) + mi]
        dbc, ubc = expect(code, dm, um)
        q = o.domid[i - 1, j - 1, k - 1] - 1
        want = po.feq(np.array([dbc]), np.asarray(ubc, float)[:, None])[:, 0]
        np.testing.assert_allclose(o.populations[:, q], want, rtol=0, atol=2e-15,
                                   err_msg=f"node {(i, j, k)} code {code}")
        checked[code] = checked.get(code, 0) + 1
    return checked


def test_boundary_pass_against_brute_force():
    """Walls (0 and 4998) take feq(D_m, 0), the velocity inlet 2001
    feq(D_m, u_in(iter) dir) and the pressure outlet 3001 feq(rho_out, U_m),
    with m each node's tier neighbour and D_m, U_m of this step (interior m,
    force shift included) or of the previous macro pass (BC m).  A density
    taken from the BC node instead of m, an inlet phase from another
    iteration, a velocity without the force shift or from the BC node, or
    another neighbour tier each fail it."""
    from zebraflow_amd import geometry
    from zebraflow_amd.engine import B_PRESSURE_OUTLET, B_VELOCITY_INLET
    b = geometry.tube((14, 12, 13), axis=0, r=4.6, inlet=B_VELOCITY_INLET,
                      outlet=B_PRESSURE_OUTLET)
    b[(b == 0) & (np.arange(13)[None, None, :] < 5)] = 4998
    u0, amp, period, it = 0.02, 0.5, 7.0, 5
    params = lattice_params(0.8, wall_mode=0, inlet_u0=u0, inlet_amp=amp, inlet_period=period,
                            inlet_dir=(0.8, 0.6, 0.0), outlet_rho=1.003,
                            body_force=(2e-4, -1e-4, 5e-5), force_dt=0.5)
    uin = u0 * (1 + amp * np.sin(2 * np.pi * it / period))

    def expect(code, dm, um):
        if code in (0, 4998):
            return dm, np.zeros(3)
        if code == B_VELOCITY_INLET:
            return dm, uin * np.array([0.8, 0.6, 0.0])
        assert code == B_PRESSURE_OUTLET
        return 1.003, um

    checked = _brute_force_bc(b, params, it, expect)
    assert set(checked) == {0, 4998, B_VELOCITY_INLET, B_PRESSURE_OUTLET}
    assert min(checked.values()) >= 20
    # some BC nodes' tier neighbour is itself a BC or wall node (the second
    # and third tiers are exercised, not only interior neighbours)
    second = sum(1 for i, j, k in zip(*np.nonzero((b == 0) | (b >= 1000)))
                 if (m := _tier_neighbour(b, i + 1, j + 1, k + 1)) is not None
                 and not 0 < b[m[0] - 1, m[1] - 1, m[2] - 1] < 1000)
    assert second > 0


@pytest.mark.parametrize("it", [0, 1999 * 62, 165600, 309601, 1999 * 200, 1234567])
def test_pressure_codes_boundary_pass_against_brute_force(it):
    """Codes 1001-1009 take feq(pressure(iter)/LBM_PRESSURE_SCALE + 1, U_m)
    (ZebraflowWT.c:6318-6427), the pressure taken from the reference's own
    waveform vectors (tests/golden/waveforms_ref.npz, written by the
    reference's branch chain) at this iteration, not from the oracle."""
    import os
    from zebraflow_amd import geometry
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "waveforms_ref.npz"))
    sel = g["iter"] == it
    assert sel.sum() == 9
    pref = dict(zip(g["code"][sel].tolist(), g["pressure"][sel].astype(np.float64).tolist()))
    b = geometry.tube((12, 14, 14), axis=0, r=5.2, inlet=1001, outlet=1001)
    # every code on the inlet and outlet planes, in patches
    plane = np.nonzero(b[0] == 1001)
    for n, (j, k) in enumerate(zip(*plane)):
        b[0, j, k] = 1001 + n % 9
        b[-1, j, k] = 1001 + (n + 4) % 9
    params = lattice_params(0.9, wall_mode=0, body_force=(1e-4, 0, 0), force_dt=0.5)
    pscale = params["pressure_scale"]

    def expect(code, dm, um):
        if code == 0:
            return dm, np.zeros(3)
        return pref[code] / pscale + 1.0, um

    checked = _brute_force_bc(b, params, it, expect)
    assert set(range(1001, 1010)) <= set(checked)
