import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libifgf_amd.so")


def sphere(n, seed):
    g = np.random.default_rng(seed).standard_normal((n, 3))
    return g / np.linalg.norm(g, axis=1)[:, None]


def spheroid(n, seed, c=0.5):
    """Points on the spheroid (1,1,c), area-uniform by rejection."""
    rng = np.random.default_rng(seed)
    out = []
    while sum(len(o) for o in out) < n:
        g = rng.standard_normal((2 * n, 3))
        p = g / np.linalg.norm(g, axis=1)[:, None]
        q = p * np.array([1.0, 1.0, c])
        # surface element of the map from the unit sphere
        dens = np.sqrt((c * p[:, 0]) ** 2 + (c * p[:, 1]) ** 2 + p[:, 2] ** 2)
        keep = rng.random(2 * n) < dens / dens.max()
        out.append(q[keep])
    return np.concatenate(out)[:n]


def perturbed_sphere(n, seed, lmax=8, amp=0.1):
    """r = 1 + amp sum_{l<=lmax, |m|<=l} a_lm Y_lm(theta, phi) with real spherical harmonics
    and a_lm ~ N(0,1)/(l+1), at n random directions (the perturbed-sphere geometry)."""
    from scipy.special import sph_harm_y
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((n, 3))
    p = g / np.linalg.norm(g, axis=1)[:, None]
    theta = np.arccos(np.clip(p[:, 2], -1.0, 1.0))
    phi = np.arctan2(p[:, 1], p[:, 0])
    r = np.ones(n)
    for l in range(lmax + 1):
        for m in range(-l, l + 1):
            a = rng.standard_normal() / (l + 1)
            y = sph_harm_y(l, abs(m), theta, phi)
            if m > 0:
                yr = np.sqrt(2.0) * (-1) ** m * y.real
            elif m < 0:
                yr = np.sqrt(2.0) * (-1) ** m * y.imag
            else:
                yr = y.real
            r += amp * a * yr
    return p * r[:, None]


def torus(n, seed, R=1.0, r=0.4):
    """Points area-uniform on the torus (R, r) (rejection on the tube angle) and the analytic
    unit normals."""
    rng = np.random.default_rng(seed)
    us, vs = [], []
    while sum(len(u) for u in us) < n:
        u = rng.uniform(0, 2 * np.pi, 2 * n)
        v = rng.uniform(0, 2 * np.pi, 2 * n)
        keep = rng.random(2 * n) < (R + r * np.cos(v)) / (R + r)
        us.append(u[keep])
        vs.append(v[keep])
    u = np.concatenate(us)[:n]
    v = np.concatenate(vs)[:n]
    x = np.stack([(R + r * np.cos(v)) * np.cos(u), (R + r * np.cos(v)) * np.sin(u),
                  r * np.sin(v)], axis=1)
    nrm = np.stack([np.cos(v) * np.cos(u), np.cos(v) * np.sin(u), np.sin(v)], axis=1)
    return x, nrm


def cdens(n, seed):
    r = np.random.default_rng(seed)
    return r.standard_normal(n) + 1j * r.standard_normal(n)


@pytest.fixture(scope="session")
def gpu():
    """Fails (does not skip) when the library cannot see a HIP device."""
    import ifgf_amd._lib as L
    n = L.lib().ifgf_device_count()
    assert n > 0, "no HIP device visible to libifgf_amd.so"
    return n
