"""Diagnostics of the resident state (gt_hip_state_diagnostics) against the
reference's compute_diagnostics (src/integrator.cpp:18-33) restated here in
FP64 numpy on the downloaded state, against closed forms for a circular
binary, and per system for a batch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def reference_diagnostics(pos, vel, mass, pot):
    # sums in the reference's order of terms per particle
    k = np.sum(0.5 * mass * np.einsum("ij,ij->i", vel, vel))
    w = np.sum(0.5 * mass * pot)
    p = np.sum(mass[:, None] * vel, axis=0)
    l = np.sum(mass[:, None] * np.cross(pos, vel), axis=0)
    m = np.sum(mass)
    c = np.sum(mass[:, None] * pos, axis=0) / m if m > 0 else np.zeros(3)
    return k, w, p, l, c


def close(a, b, rtol=1e-12, atol=1e-14):
    return np.allclose(np.asarray(a, float), np.asarray(b, float), rtol=rtol, atol=atol)


def test_diagnostics_of_a_galaxy(hip):
    import gt_ics
    m = gt_ics.model(disk_n=20000, bulge_n=5000, halo_n=25000, seed=1)
    pos, vel, mass, _ = gt_ics.galaxy(m)
    cfg = dict(theta=0.75, eps=0.025)
    hip.state_set(pos, vel, mass)
    with pytest.raises(ValueError):  # potentials not computed yet
        hip.state_diagnostics()
    hip.prime(**cfg)
    for step in range(3):
        d = hip.state_diagnostics(time=0.01 * step)[0]
        p, v, a, phi = hip.state_get()
        k, w, mom, l, c = reference_diagnostics(p, v, mass, phi)
        assert d["time"] == 0.01 * step
        assert close(d["kinetic"], k) and close(d["potential"], w)
        scale = np.sum(mass) * np.max(np.abs(v))
        assert np.allclose(d["momentum"], mom, rtol=1e-12, atol=1e-12 * scale)
        assert close(d["angular_momentum"], l, atol=1e-12 * scale * np.max(np.abs(p)))
        assert close(d["centre_of_mass"], c, atol=1e-12 * np.max(np.abs(p)))
        hip.steps(5, 0.01, **cfg)
    assert d["potential"] < 0 < d["kinetic"]


def test_diagnostics_closed_form_binary(hip):
    # two unit masses at +-1 on x moving at +-v on y: K = v^2, L_z = 2 v,
    # W = 0.5 (m phi_1 + m phi_2) = -G m m / r_soft, P = 0, centre at 0
    v = 0.3
    pos = np.array([[1.0, 0, 0], [-1.0, 0, 0]])
    vel = np.array([[0, v, 0], [0, -v, 0]])
    mass = np.ones(2)
    eps = 0.01
    hip.state_set(pos, vel, mass)
    hip.prime(theta=0.5, eps=eps, exact=True)
    d = hip.state_diagnostics()[0]
    assert close(d["kinetic"], v * v)
    assert close(d["angular_momentum"], [0, 0, 2 * v])
    assert close(d["momentum"], [0, 0, 0])
    assert close(d["centre_of_mass"], [0, 0, 0])
    # the evaluation is FP32: the potential to single precision
    assert abs(d["potential"] + 1.0 / np.sqrt(4 + eps * eps)) < 1e-6


def test_diagnostics_per_system(hip, oracle):
    a, ma = oracle.test_system(0, 3000, 51)
    b, mb = oracle.test_system(1, 5000, 52)
    pos = np.concatenate([a, b + 100.0])
    mass = np.concatenate([ma, mb])
    rng = np.random.default_rng(5)
    vel = 0.2 * rng.standard_normal(pos.shape)
    hip.state_set(pos, vel, mass)
    hip.state_set_systems([0, 3000, 8000])
    hip.prime(theta=0.7, eps=0.01)
    with pytest.raises(ValueError):
        hip.state_diagnostics(systems=1)
    ds = hip.state_diagnostics(systems=2)
    p, v, _, phi = hip.state_get()
    for d, sl in zip(ds, (slice(0, 3000), slice(3000, 8000))):
        k, w, mom, l, c = reference_diagnostics(p[sl], v[sl], mass[sl], phi[sl])
        assert close(d["kinetic"], k) and close(d["potential"], w)
        assert np.allclose(d["momentum"], mom, rtol=1e-12, atol=1e-13)
        assert np.allclose(d["angular_momentum"], l, rtol=1e-12, atol=1e-10)
        assert close(d["centre_of_mass"], c, atol=1e-12)
