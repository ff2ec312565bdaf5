"""GPU time step vs the FP64 oracle on the same seeded inputs (through the
C ABI).  Tolerances, per storage form, on populations f and on velocity:

  FP64           |f - f_oracle| <= 1e-13 (same arithmetic, different order)
  FP32 deviation |f - f_oracle| <= 2e-7, rel. L2 velocity error <= 1e-5
  FP16 deviation rel. L2 velocity error <= 1e-2 (10-bit mantissa storage)
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from zebraflow_amd import geometry
from zebraflow_amd.engine import (B_PRESSURE_OUTLET, B_VELOCITY_INLET, STORE_F16DEV, STORE_F32DEV,
                                  STORE_F64, Engine, lattice_config, reference_config)

pytestmark = pytest.mark.gpu

STORES = {"f64": STORE_F64, "f32dev": STORE_F32DEV, "f16dev": STORE_F16DEV}


def oracle_params(cfg):
    return dict(wall_mode=cfg.wall_mode, visc_scale=cfg.visc_scale, force_dt=cfg.force_dt,
                body_force=list(cfg.body_force), inlet_u0=cfg.inlet_u0, inlet_amp=cfg.inlet_amp,
                inlet_period=cfg.inlet_period, inlet_dir=list(cfg.inlet_dir),
                outlet_rho=cfg.outlet_rho, t_scale=cfg.t_scale, pressure_scale=cfg.pressure_scale,
                periodic=tuple(cfg.periodic))


def case(name, store):
    if name == "tube_io":
        b = geometry.tube((24, 20, 20), axis=0, r=7.0, inlet=B_VELOCITY_INLET,
                          outlet=B_PRESSURE_OUTLET)
        cfg = lattice_config(*b.shape, 0.8, storage=store, inlet_u0=0.02, inlet_amp=0.5,
                             inlet_period=20.0, inlet_dir=(1, 0, 0))
        return b, cfg, 30, None
    if name == "ref_eqwall":
        b = geometry.tube((20, 18, 18), axis=0, r=6.0, inlet=1001, outlet=1003)
        cfg = reference_config(*b.shape, storage=store)
        return b, cfg, 20, 1e-7
    if name == "periodic_force":
        b = geometry.tube((16, 16, 24), axis=2, r=6.5, periodic=True)
        cfg = lattice_config(*b.shape, 0.8, storage=store, periodic=(0, 0, 1),
                             body_force=(0, 0, 1e-4))
        return b, cfg, 30, None
    if name == "pair":
        b = geometry.vessel_pair((96, 48, 40), 5.0, 2.5, 4.0, 3, 3)
        cfg = lattice_config(*b.shape, 0.8, storage=store, inlet_u0=0.03, inlet_amp=0.5,
                             inlet_period=4000.0, inlet_dir=(1, 0, 0))
        return b, cfg, 25, 2e-6
    raise KeyError(name)


def run_pair(name, store):
    b, cfg, steps, force_amp = case(name, store)
    print(f"case {name} store {store} box {b.shape} steps {steps}", flush=True)
    ora = po.LbmOracle(b, oracle_params(cfg))
    rng = np.random.default_rng(20261203)
    eng = Engine(cfg)
    eng.upload_lattice(b)
    assert eng.dommax == ora.n
    # seeded non-equilibrium start so every population matters
    rho = 1 + 0.002 * rng.standard_normal(ora.n)
    u = 0.01 * rng.standard_normal((3, ora.n))
    f0 = po.feq(rho, u) * (1 + 0.01 * rng.standard_normal((19, ora.n)))
    ora.set_populations(f0)
    eng.upload_populations(f0)
    force = None
    if force_amp:
        force = force_amp * rng.standard_normal((3, ora.n))
        ora.force[...] = force
        eng.upload_force(force)
    for it in range(1, steps + 1):
        rbc = 1 if (force_amp and it == steps // 2) else 0
        ora.step(it, rbc_step=rbc)
        eng.step(it, rbc_step=rbc)
    fg = eng.populations()
    mg = eng.macro()
    if force_amp:
        np.testing.assert_array_equal(eng.force(), 0.0)
    eng.close()
    return ora, fg, mg


def vel_rel_l2(ora, mg):
    uo = np.stack([ora.U, ora.V, ora.W])
    ug = mg[1:4]
    return np.linalg.norm(ug - uo) / np.linalg.norm(uo)


@pytest.mark.parametrize("name", ["tube_io", "ref_eqwall", "periodic_force", "pair"])
def test_f64_matches_oracle(name):
    ora, fg, mg = run_pair(name, STORE_F64)
    np.testing.assert_allclose(fg, ora.populations, rtol=0, atol=1e-13)
    np.testing.assert_allclose(mg[0], ora.D, rtol=0, atol=1e-13)
    np.testing.assert_allclose(mg[1], ora.U, rtol=0, atol=1e-13)


@pytest.mark.parametrize("name", ["tube_io", "ref_eqwall", "periodic_force", "pair"])
def test_f32dev_matches_oracle(name):
    ora, fg, mg = run_pair(name, STORE_F32DEV)
    np.testing.assert_allclose(fg, ora.populations, rtol=0, atol=2e-7)
    assert vel_rel_l2(ora, mg) <= 1e-5


@pytest.mark.parametrize("name", ["tube_io", "periodic_force"])
def test_f16dev_matches_oracle(name):
    ora, fg, mg = run_pair(name, STORE_F16DEV)
    assert vel_rel_l2(ora, mg) <= 1e-2
    np.testing.assert_allclose(mg[0], ora.D, rtol=0, atol=2e-3)


@pytest.mark.parametrize("store", ["f64", "f32dev"])
def test_mass_and_momentum_sums(store):
    """Device FP64 conservation sums: mass is invariant with bounce-back in a
    periodic tube and momentum grows by N*F per step along the tube."""
    b = geometry.tube((16, 16, 20), axis=2, r=6.5, periodic=True)
    g = 2e-5
    cfg = lattice_config(*b.shape, 0.8, storage=STORES[store], periodic=(0, 0, 1),
                         body_force=(0, 0, g))
    eng = Engine(cfg)
    eng.upload_lattice(b)
    s0 = eng.conservation_sums()
    n = eng.n_interior
    assert abs(s0[0] - n) < 1e-9 * n
    eng.run(0, 10)
    s1 = eng.conservation_sums()
    tol = 1e-12 if store == "f64" else 1e-6
    assert abs(s1[0] - s0[0]) <= tol * n
    # z-momentum: +N g per step minus what the walls absorb (bounded, same sign)
    assert 0 < s1[3] <= 10 * n * g * (1 + 1e-6)
    eng.close()


@pytest.mark.parametrize("store", ["f64", "f32dev", "f16dev"])
def test_cache_policies_bit_identical(store):
    """The stream kernel's cache policy (which directions are loaded and
    stored non-temporally) and its block order change only how memory is
    accessed: populations after 12 steps with a force field are bit-identical
    across every policy.  The 128x64x64 vessel (about 268k fluid nodes, 1048
    blocks of 256) holds a whole round of 8 x 128 blocks of the chunked
    order, the order used with partial caching."""
    from zebraflow_amd.engine import (FLAG_LINEAR_BLOCKS, FLAG_NT_POPULATIONS,
                                      flag_cached_dirs)
    b = geometry.tube((128, 64, 64), axis=0, r=26.0, inlet=B_VELOCITY_INLET,
                      outlet=B_PRESSURE_OUTLET)
    rng = np.random.default_rng(20261204)
    variants = [0, flag_cached_dirs(19), flag_cached_dirs(18), flag_cached_dirs(8),
                flag_cached_dirs(7), flag_cached_dirs(0), FLAG_NT_POPULATIONS,
                FLAG_LINEAR_BLOCKS | flag_cached_dirs(8), FLAG_LINEAR_BLOCKS]
    ref = None
    f0 = force = None
    for flags in variants:
        cfg = lattice_config(*b.shape, 0.8, storage=STORES[store], inlet_u0=0.03, inlet_amp=0.5,
                             inlet_period=40.0, inlet_dir=(1, 0, 0), flags=flags)
        eng = Engine(cfg)
        eng.upload_lattice(b)
        if f0 is None:
            n = eng.dommax
            assert eng.n_interior > 8 * 128 * 256
            rho = 1 + 0.002 * rng.standard_normal(n)
            u = 0.01 * rng.standard_normal((3, n))
            f0 = po.feq(rho, u) * (1 + 0.01 * rng.standard_normal((19, n)))
            force = 1e-5 * rng.standard_normal((3, n))
        eng.upload_populations(f0)
        eng.upload_force(force)
        for it in range(1, 13):
            eng.step(it, rbc_step=0)
        f = eng.populations()
        eng.close()
        print(f"store {store} flags {flags:#x}: max |f| {np.abs(f).max():.6f}", flush=True)
        if ref is None:
            ref = f
        else:
            np.testing.assert_array_equal(f, ref, err_msg=f"flags {flags:#x}")


@pytest.mark.parametrize("store", [STORE_F64, STORE_F32DEV])
@pytest.mark.parametrize("iter0", [0, 150000, 165590, 300000, 359990, 700000, 1234560])
def test_pressure_codes_over_the_cycle(store, iter0):
    """All nine cardiac pressure codes (ZebraflowWT.c:6318-6427) on the inlet
    and outlet planes of a tube with equilibrium walls, 12 steps from
    iterations spread over the 0.36 s cycle (720,000 steps of 5e-7 s): rise
    and fall of every shape, the joins at 23 % and 30 %, a cycle wrap, and a
    late float time; the device against the oracle, whose waveforms are the
    reference's to the bit."""
    print(f"case pressure codes store {store} iter0 {iter0}", flush=True)
    b = geometry.tube((18, 16, 16), axis=0, r=6.0, inlet=1001, outlet=1001)
    plane = np.nonzero(b[0] == 1001)
    for n, (j, k) in enumerate(zip(*plane)):
        b[0, j, k] = 1001 + n % 9
        b[-1, j, k] = 1001 + (n + 4) % 9
    cfg = reference_config(*b.shape, storage=store)
    ora = po.LbmOracle(b, oracle_params(cfg))
    eng = Engine(cfg)
    eng.upload_lattice(b)
    rng = np.random.default_rng(iter0 + 11)
    f0 = po.feq(1.04 + 0.002 * rng.standard_normal(ora.n), 1e-3 * rng.standard_normal((3, ora.n)))
    ora.set_populations(f0)
    eng.upload_populations(f0)
    for it in range(iter0 + 1, iter0 + 13):
        ora.step(it)
        eng.step(it)
    fg = eng.populations()
    eng.close()
    tol = 1e-13 if store == STORE_F64 else 2e-7
    np.testing.assert_allclose(fg, ora.populations, rtol=0, atol=tol)
