"""Pins the CPU oracle (oracle/kpt_oracle.cpp) with known answers, closed forms and
invariants, chosen so that a dropped term, a wrong sign, index or draw role fails
one of them. No GPU is needed."""
import numpy as np
import pytest

import kpt
import kpt_oracle as O
from helpers import SQRT3, const_field, copy_parts, random_particles, rel_err, uniform_edges

# Random123 known-answer vectors for philox4x32_10 (Salmon et al. 2011).
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,want", PHILOX_KAT)
def test_philox_known_answers(oracle_mod, kpt_mod, ctr, key, want):
    assert tuple(O.philox4x32_10(ctr, key)) == want
    assert tuple(kpt.philox4x32_10(ctr, key)) == want  # host copy in libkpt


def test_uniform_sqrt3_range_and_moments(oracle_mod):
    assert O.uniform_sqrt3(0) > -SQRT3
    assert O.uniform_sqrt3(0xFFFFFFFF) < SQRT3
    # midpoint of the 32-bit lattice maps to ~0
    assert abs(O.uniform_sqrt3(0x7FFFFFFF) + O.uniform_sqrt3(0x80000000)) < 1e-15


# ---------------------------------------------------------------------------
# coefficient packing
# ---------------------------------------------------------------------------
def _trig_field(nx, ny, x0, W, y0, H):
    xe = uniform_edges(nx, x0, x0 + W)
    ye = uniform_edges(ny, y0, y0 + H)
    i = np.arange(nx)
    j = np.arange(ny)
    xn = x0 + i * W / nx
    yn = y0 + j * H / ny
    X, Y = np.meshgrid(xn, yn)
    kx, ky = 2 * np.pi / W, 4 * np.pi / H
    planes = [
        np.sin(kx * (X - x0)) + 2.0,            # Bx
        np.cos(ky * (Y - y0)),                  # By
        0.5 * np.sin(kx * (X - x0)) * np.cos(ky * (Y - y0)),  # Bz
        np.cos(kx * (X - x0)),                  # Vx
        np.sin(ky * (Y - y0)),                  # Vy
        np.sin(kx * (X - x0) + ky * (Y - y0)),  # Vz
    ]
    return xe, ye, planes, (kx, ky)


@pytest.mark.parametrize("model", [16, 21])
def test_packing_trig_closed_form(oracle_mod, model):
    """Central differences of sin/cos on the periodic grid equal f'(x) sin(kh)/(kh),
    at every stored node including the periodic ghosts -1 and n."""
    nx, ny, x0, W, y0, H = 32, 24, -1.0, 2.0, -0.5, 1.0
    xe, ye, planes, (kx, ky) = _trig_field(nx, ny, x0, W, y0, H)
    f = O.Field(xe, ye, planes, model)
    c = f.coefficients()
    hx, hy = W / nx, H / ny
    ii = np.arange(-1, nx + 1)
    jj = np.arange(-1, ny + 1)
    X, Y = np.meshgrid(x0 + ii * hx, y0 + jj * hy)
    sx, sy = np.sin(kx * hx) / (kx * hx), np.sin(ky * hy) / (ky * hy)
    ax, ay = kx * (X - x0), ky * (Y - y0)
    exact = {
        "Vx": np.cos(ax), "Vy": np.sin(ay), "Vz": np.sin(ax + ay),
        "Bx": np.sin(ax) + 2.0, "By": np.cos(ay), "Bz": 0.5 * np.sin(ax) * np.cos(ay),
        "dVxdx": -kx * np.sin(ax) * sx, "dVxdy": 0 * X,
        "dVydx": 0 * X, "dVydy": ky * np.cos(ay) * sy,
        "dBxdx": kx * np.cos(ax) * sx, "dBxdy": 0 * X,
        "dBydx": 0 * X, "dBydy": -ky * np.sin(ay) * sy,
        "dBzdx": 0.5 * kx * np.cos(ax) * sx * np.cos(ay),
        "dBzdy": -0.5 * ky * np.sin(ax) * np.sin(ay) * sy,
    }
    # d/dx of sin(ax+ay) by central difference in x only: cos(ax+ay) * sin(kx hx)/hx
    exact["dVzdx"] = kx * np.cos(ax + ay) * sx
    exact["dVzdy"] = ky * np.cos(ax + ay) * sy
    bm = lambda X_, Y_: np.sqrt((np.sin(kx * (X_ - x0)) + 2) ** 2 + np.cos(ky * (Y_ - y0)) ** 2
                                + (0.5 * np.sin(kx * (X_ - x0)) * np.cos(ky * (Y_ - y0))) ** 2)
    exact["Bmag"] = bm(X, Y)
    exact["dBmagdx"] = (bm(X + hx, Y) - bm(X - hx, Y)) / (2 * hx)
    exact["dBmagdy"] = (bm(X, Y + hy) - bm(X, Y - hy)) / (2 * hy)
    if model == 16:
        names = ["Vx", "Vy", "Bx", "By", "Bz", "Bmag", "dVxdx", "dVydy", "dBxdx", "dBxdy",
                 "dBydx", "dBydy", "dBzdx", "dBzdy", "dBmagdx", "dBmagdy"]
    else:
        names = ["Vx", "Vy", "Vz", "Bx", "By", "Bz", "Bmag", "dVxdx", "dVxdy", "dVydx", "dVydy",
                 "dVzdx", "dVzdy", "dBxdx", "dBxdy", "dBydx", "dBydy", "dBzdx", "dBzdy",
                 "dBmagdx", "dBmagdy"]
    for k, name in enumerate(names):
        np.testing.assert_allclose(c[:, :, k], exact[name], rtol=0, atol=1e-12, err_msg=name)


def test_packing_periodic_ghost_identity(oracle_mod):
    rng = np.random.default_rng(3)
    nx, ny = 17, 11
    xe, ye = uniform_edges(nx, 0.0, 1.7), uniform_edges(ny, -0.3, 0.8)
    planes = [rng.standard_normal((ny, nx)) for _ in range(6)]
    c = O.Field(xe, ye, planes, 21).coefficients()
    np.testing.assert_array_equal(c[:, 0, :7], c[:, nx, :7])       # node -1 == node n-1
    np.testing.assert_array_equal(c[:, nx + 1, :7], c[:, 1, :7])   # node n == node 0
    np.testing.assert_array_equal(c[0, :, :7], c[ny, :, :7])
    np.testing.assert_allclose(c[:, 0, 7:], c[:, nx, 7:], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(c[0, :, 7:], c[ny, :, 7:], rtol=1e-12, atol=1e-12)


# ---------------------------------------------------------------------------
# interpolation
# ---------------------------------------------------------------------------
def test_interpolation_bilinear_exact_and_periodic(oracle_mod):
    nx, ny, x0, W, y0, H = 20, 16, -1.0, 2.0, -0.5, 1.0
    xe, ye = uniform_edges(nx, x0, x0 + W), uniform_edges(ny, y0, y0 + H)
    hx, hy = W / nx, H / ny
    I, J = np.meshgrid(np.arange(nx), np.arange(ny))
    g = 0.3 + 0.7 * I - 0.2 * J + 0.05 * I * J  # bilinear in the node indices
    planes = [g, 2 * g, 1.0 + 0 * g, -g, 0 * g, 0 * g]
    f = O.Field(xe, ye, planes, 21)
    rng = np.random.default_rng(1)
    for _ in range(200):
        # sampling coordinate q = x + dx/2 with x inside the interior cells
        fi = rng.uniform(1.0, nx - 2.0)
        fj = rng.uniform(1.0, ny - 2.0)
        q = (x0 + fi * hx + 0.5 * hx, y0 + fj * hy + 0.5 * hy)
        out = f.sample(*q)
        want = 0.3 + 0.7 * fi - 0.2 * fj + 0.05 * fi * fj
        assert abs(out[0] + want) < 1e-12 * max(1, abs(want))  # Vx = -g
        assert abs(out[3] - want) < 1e-12 * max(1, abs(want))  # Bx = g
    # node values are returned at node positions
    for (i, j) in [(0, 0), (nx - 1, ny - 1), (5, 7)]:
        out = f.sample(x0 + i * hx + 0.5 * hx, y0 + j * hy + 0.5 * hy)
        assert abs(out[3] - g[j, i]) < 1e-12
    # periodicity
    for _ in range(50):
        qx, qy = rng.uniform(x0, x0 + W), rng.uniform(y0, y0 + H)
        a = f.sample(qx, qy)
        b = f.sample(qx + W, qy - H)
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12)


# ---------------------------------------------------------------------------
# the push: closed forms with injected noise
# ---------------------------------------------------------------------------
def run(lo, hi, s, parts, noise):
    """Push with an injected noise table padded by two zero rows (a last sub-step
    truncated to the interval end may need one more row than dt_out / dt)."""
    pad = np.zeros((noise.shape[0], noise.shape[1] + 2, 4))
    pad[:, : noise.shape[1]] = noise
    s2 = s.copy()
    s2.max_particle_steps_per_interval = min(s2.max_particle_steps_per_interval, pad.shape[1])
    return O.push_interval(lo, hi, s2, 0.0, 0, 1, parts, noise=pad)


def _box(nx=32, ny=32):
    return uniform_edges(nx, -1.0, 1.0), uniform_edges(ny, -1.0, 1.0)


def _host(**kw):
    h = kpt.default_host_settings()
    h.magnetic_dependency = 0
    h.momentum_dependency = 0
    h.p0 = 1.0
    h.kpara0 = 0.01
    h.kperp_over_kpara = 0.1
    h.mhd_dt = 0.01
    h.dt_min_rel = 0.1  # 10 sub-steps of 1e-3 per interval
    h.dt_max_rel = 0.1
    for k, v in kw.items():
        setattr(h, k, v)
    return h


def test_push_pure_diffusion_closed_form(oracle_mod):
    """Uniform B along (cos a, sin a), V = 0: every step is dt_min and
    dx = sqrt(dt) (r1 sqrt(2 kperp) + r3 sqrt(2 (kpar - kperp)) bx/b),
    dy = sqrt(dt) (r2 sqrt(2 kperp) + r3 sqrt(2 (kpar - kperp)) by/b)."""
    xe, ye = _box()
    a = 0.3
    planes = const_field(32, 32, bx=2 * np.cos(a), by=2 * np.sin(a), bz=0.0)
    lo = O.Field(xe, ye, planes, 16)
    h = _host()
    s = kpt.make_device_settings(h, xe, ye)
    n, steps = 64, 10
    rng = np.random.default_rng(5)
    parts = random_particles(n, s, rng)
    parts["x"][:] = rng.uniform(-0.2, 0.2, n)
    parts["y"][:] = rng.uniform(-0.2, 0.2, n)
    noise = rng.uniform(-SQRT3, SQRT3, size=(n, steps, 4))
    x0, y0, p0 = parts["x"].copy(), parts["y"].copy(), parts["p"].copy()
    taken, fails = run(lo, lo, s, parts, noise)
    assert fails == 0 and taken == n * steps
    dt = s.dt_min
    kpar, kperp = h.kpara0, h.kpara0 * h.kperp_over_kpara
    r = noise
    dx = np.sqrt(dt) * (r[:, :, 0] * np.sqrt(2 * kperp) + r[:, :, 2] * np.sqrt(2 * (kpar - kperp)) * np.cos(a))
    dy = np.sqrt(dt) * (r[:, :, 1] * np.sqrt(2 * kperp) + r[:, :, 2] * np.sqrt(2 * (kpar - kperp)) * np.sin(a))
    np.testing.assert_allclose(parts["x"], x0 + dx.sum(1), rtol=0, atol=1e-13)
    np.testing.assert_allclose(parts["y"], y0 + dy.sum(1), rtol=0, atol=1e-13)
    np.testing.assert_array_equal(parts["p"], p0)
    np.testing.assert_allclose(parts["t"], s.mhd_dt, rtol=1e-13)
    assert np.all(parts["step_dt"] == dt)


def test_push_diffusion_variance_philox(oracle_mod):
    """Philox noise: Var(dx) = 2 kpar t along B, 2 kperp t across (B along x)."""
    xe, ye = _box()
    lo = O.Field(xe, ye, const_field(32, 32, bx=1.0), 16)
    h = _host(mhd_dt=0.05, dt_min_rel=1e-2, dt_max_rel=1e-2)
    s = kpt.make_device_settings(h, xe, ye)
    n = 40000
    parts = random_particles(n, s, np.random.default_rng(2))
    parts["x"][:] = 0.0
    parts["y"][:] = 0.0
    O.push_interval(lo, lo, s, 0.0, 3, 99, parts)
    t = s.mhd_dt
    vx, vy = parts["x"].var(), parts["y"].var()
    kpar, kperp = h.kpara0, h.kpara0 * h.kperp_over_kpara
    assert abs(vx / (2 * kpar * t) - 1) < 0.03
    assert abs(vy / (2 * kperp * t) - 1) < 0.03
    assert abs(parts["x"].mean()) < 4 * np.sqrt(2 * kpar * t / n)


def test_push_adiabatic_closed_form(oracle_mod):
    """Vx = c (x - xc) in the interior: dp/dt = -p c/3, so p_n = p_0 (1 - c dt/3)^n
    and x_n - xc = (x_0 - xc)(1 + c dt)^n (noise zero), until the 1/4 p0 floor."""
    nx = ny = 64
    xe, ye = _box(nx, ny)
    hx = 2.0 / nx
    xn = -1.0 + np.arange(nx) * hx
    c = 3.0
    Vx = np.tile(c * xn, (ny, 1))
    planes = const_field(nx, ny, bx=1.0)
    planes[3] = Vx
    lo = O.Field(xe, ye, planes, 16)
    h = _host(kpara0=1e-12, kperp_over_kpara=0.0, mhd_dt=0.02, dt_min_rel=0.05, dt_max_rel=0.05)
    s = kpt.make_device_settings(h, xe, ye)
    parts = random_particles(4, s, np.random.default_rng(0))
    parts["x"][:] = [-0.2, -0.05, 0.1, 0.3]
    parts["y"][:] = 0.0
    parts["p"][:] = 1.0
    noise = np.zeros((4, 20, 4))
    x0 = parts["x"].copy()
    run(lo, lo, s, parts, noise)
    dt = s.dt_min
    nsteps = round(s.mhd_dt / dt)
    # the sampled x is wrapped x + dx/2 with value i at x0 + i*dx: field is c*x exactly
    np.testing.assert_allclose(parts["p"], (1 - c * dt / 3) ** nsteps, rtol=1e-12)
    np.testing.assert_allclose(parts["x"], x0 * (1 + c * dt) ** nsteps, rtol=1e-11, atol=1e-14)


def test_push_momentum_floor(oracle_mod):
    nx = ny = 32
    xe, ye = _box(nx, ny)
    xn = -1.0 + np.arange(nx) * (2.0 / nx)
    planes = const_field(nx, ny, bx=1.0)
    planes[3] = np.tile(30.0 * xn, (ny, 1))  # strong expansion: p drops fast
    lo = O.Field(xe, ye, planes, 16)
    h = _host(kpara0=1e-12, kperp_over_kpara=0.0, mhd_dt=0.5, dt_min_rel=0.02, dt_max_rel=0.02)
    s = kpt.make_device_settings(h, xe, ye)
    parts = random_particles(2, s, np.random.default_rng(0))
    parts["x"][:] = 0.0
    parts["y"][:] = 0.0
    parts["p"][:] = 1.0
    run(lo, lo, s, parts, np.zeros((2, 50, 4)))
    np.testing.assert_array_equal(parts["p"], 0.25 * h.p0)


def test_push_dt_limiter_momentum_candidate(oracle_mod):
    """All three rates nonzero: dt = min(diffusion, advection, 0.1 p/|dp/dt|),
    clamped to [dt_min, dt_max]. At a node where V is small but div V = c + 0.5
    is large, the momentum candidate 0.1 * 3 / |div V| wins and is the stored step_dt."""
    nx = ny = 64
    xe, ye = _box(nx, ny)
    xn = -1.0 + np.arange(nx) * (2.0 / nx)
    yn = -1.0 + np.arange(ny) * (2.0 / ny)
    planes = const_field(nx, ny, bx=1.0, by=0.5)
    c = 20.0
    planes[3] = np.tile(c * xn, (ny, 1)) + 0.01
    planes[4] = np.tile((0.5 * yn)[:, None], (1, nx)) + 0.01
    lo = O.Field(xe, ye, planes, 16)
    h = _host(kpara0=1e-3, kperp_over_kpara=0.5, mhd_dt=1.0, dt_min_rel=1e-6, dt_max_rel=1.0)
    s = kpt.make_device_settings(h, xe, ye)
    s.max_particle_steps_per_interval = 1
    parts = random_particles(1, s, np.random.default_rng(0))
    parts["x"][:] = 0.0
    parts["y"][:] = 0.0
    run(lo, lo, s, parts, np.zeros((1, 1, 4)))
    want = 0.1 * 3.0 / (c + 0.5)
    diff = (0.5 * (2.0 / nx)) ** 2 / (2 * h.kpara0)
    adv = 2 * h.kpara0 * h.kperp_over_kpara / 0.01 ** 2
    assert want < diff and want < adv
    assert abs(parts["step_dt"][0] / want - 1) < 1e-12
    # with a tiny dt_max the clamp wins
    s.dt_max = 1e-3
    parts["x"][:] = 0.0
    parts["y"][:] = 0.0
    parts["t"][:] = 0.0
    parts["status"][:] = 0  # the one-step guard marked it
    run(lo, lo, s, parts, np.zeros((1, 1, 4)))
    assert parts["step_dt"][0] == 1e-3


def test_push_drift_closed_form(oracle_mod):
    """B = (0, 0, Bz(y)) with Bz linear: the curvature/gradient drift is
    vdx = -(1/(3q)) / D(p) * Bz'/Bz^2, vdy = 0, with D(p) = sqrt((d1 p0/p)^2 + (d2 p0^2/p^2)^2)."""
    nx = ny = 64
    xe, ye = _box(nx, ny)
    yn = -1.0 + np.arange(ny) * (2.0 / ny)
    g = 0.4
    planes = const_field(nx, ny, bx=0.0, bz=1.0)
    planes[2] = np.tile((1.0 + g * yn)[:, None], (1, nx))
    lo = O.Field(xe, ye, planes, 16)
    h = _host(kpara0=1e-6, kperp_over_kpara=0.0, drift_param1=2.0, drift_param2=0.5, charge=-1.0,
              mhd_dt=0.01, dt_min_rel=0.1, dt_max_rel=0.1)
    s = kpt.make_device_settings(h, xe, ye)
    parts = random_particles(3, s, np.random.default_rng(0))
    parts["x"][:] = 0.0
    parts["y"][:] = [-0.3, 0.0, 0.25]
    parts["p"][:] = [0.5, 1.0, 2.0]
    y0 = parts["y"].copy()
    run(lo, lo, s, parts, np.zeros((3, 10, 4)))
    p = parts["p"]
    D = np.sqrt((h.drift_param1 * h.p0 / p) ** 2 + (h.drift_param2 * h.p0 ** 2 / p ** 2) ** 2)
    bz = 1.0 + g * y0
    vdx = (1.0 / (3 * h.charge)) / D * (g / bz ** 2 - 2 * bz * g / bz ** 3)
    np.testing.assert_allclose(parts["x"], vdx * s.mhd_dt, rtol=1e-10)
    np.testing.assert_array_equal(parts["y"], y0)


@pytest.mark.parametrize("gk", [2.0, 5.0 / 3.0])
def test_push_focused_streaming_and_scattering(oracle_mod, gk):
    """Reduced focused branch, uniform B along x, V = 0: dx/dt = v0 (p/p0) mu, and
    one zero-noise step moves mu by dt * dD/dmu of D(mu) = duu0 (1 - mu^2)(|mu|^(g-1) + h0).
    At g = 2 that is the finite-difference derivative of D. For other g the
    reference's derivative (src/main.cpp:1557-1571) lacks the (g - 1) factor of
    the |mu|^(g-2) term, so it exceeds dD/dmu by (2 - g) duu0 (1 - mu^2) |mu|^(g-2) sgn(mu).
    The pitch noise enters as r4 * sqrt(2 D dt); the third draw is ignored."""
    xe, ye = _box()
    lo = O.Field(xe, ye, const_field(32, 32, bx=1.5), 21)
    h = _host(focused_transport=1, particle_v0=2.0, duu0=0.3, gamma_turb=gk,
              mhd_dt=1e-3, dt_min_rel=1.0, dt_max_rel=1.0)
    s = kpt.make_device_settings(h, xe, ye)
    s.max_particle_steps_per_interval = 1
    mu0 = np.array([0.5, -0.3, 0.8, 0.05])
    n = len(mu0)
    base = random_particles(n, s, np.random.default_rng(0))
    base["x"][:] = 0.0
    base["y"][:] = 0.0
    base["p"][:] = 1.0
    base["mu"][:] = mu0
    a = copy_parts(base)
    run(lo, lo, s, a, np.zeros((n, 1, 4)))
    dt = s.dt_min
    np.testing.assert_allclose(a["x"], h.particle_v0 * mu0 * dt, rtol=1e-12)
    h0 = s.pitch_angle_scattering_h0
    D = lambda m: h.duu0 * (1 - m * m) * (np.abs(m) ** (gk - 1) + h0)
    eps = 1e-7
    dD = (D(mu0 + eps) - D(mu0 - eps)) / (2 * eps)
    deficit = (2 - gk) * h.duu0 * (1 - mu0 ** 2) * np.abs(mu0) ** (gk - 2) * np.sign(mu0)
    np.testing.assert_allclose((a["mu"] - mu0) / dt, dD + deficit, rtol=1e-6)
    b = copy_parts(base)
    nz = np.zeros((n, 1, 4))
    nz[:, 0, 3] = 1.0
    run(lo, lo, s, b, nz)
    np.testing.assert_allclose(b["mu"] - a["mu"], np.sqrt(2 * D(mu0) * dt), rtol=1e-10)
    c = copy_parts(base)
    nz = np.zeros((n, 1, 4))
    nz[:, 0, 2] = 1.0
    run(lo, lo, s, c, nz)
    np.testing.assert_array_equal(c["mu"], a["mu"])
    np.testing.assert_array_equal(c["x"], a["x"])


def test_push_failure_zero_field_and_guard(oracle_mod):
    xe, ye = _box()
    h = _host()
    s = kpt.make_device_settings(h, xe, ye)
    zero = O.Field(xe, ye, const_field(32, 32, bx=0.0), 16)
    parts = random_particles(8, s, np.random.default_rng(0))
    steps, fails = O.push_interval(zero, zero, s, 0.0, 0, 1, parts)
    assert steps == 0 and fails == 8 and np.all(parts["status"] == 2)
    ok = O.Field(xe, ye, const_field(32, 32, bx=1.0), 16)
    parts = random_particles(8, s, np.random.default_rng(0))
    s.max_particle_steps_per_interval = 3  # interval needs 100 steps
    steps, fails = O.push_interval(ok, ok, s, 0.0, 0, 1, parts)
    assert steps == 24 and fails == 8 and np.all(parts["status"] == 2)


def test_push_time_blend_is_linear(oracle_mod):
    """Advection by Vx only, lower frame Vx = a, upper Vx = b: with negligible
    kappa and zero noise, x(t_end) - x(t0) = dt * sum_k (a + (b - a) rt_k)."""
    xe, ye = _box()
    lo = O.Field(xe, ye, const_field(32, 32, bx=1.0, vx=0.2), 16)
    hi = O.Field(xe, ye, const_field(32, 32, bx=1.0, vx=0.6), 16)
    h = _host(kpara0=1e-14, kperp_over_kpara=0.0, mhd_dt=0.1, dt_min_rel=0.1, dt_max_rel=0.1)
    s = kpt.make_device_settings(h, xe, ye)
    parts = random_particles(2, s, np.random.default_rng(0))
    parts["x"][:] = 0.0
    parts["y"][:] = 0.0
    run(lo, hi, s, parts, np.zeros((2, 10, 4)))
    dt = s.dt_min
    rts = np.arange(10) * dt / s.mhd_dt
    np.testing.assert_allclose(parts["x"], np.sum(dt * (0.2 + 0.4 * rts)), rtol=1e-12)


# ---------------------------------------------------------------------------
# split, histogram, summary, inject
# ---------------------------------------------------------------------------
def _split_settings():
    h = kpt.default_host_settings()
    h.p0 = 1.0
    h.split_ratio = 1.2
    h.pmin_split_over_p0 = 2.0
    h.pmax = 10.0
    xe, ye = _box()
    return kpt.make_device_settings(h, xe, ye)


def test_split_conserves_weight_and_orders_children(oracle_mod):
    s = _split_settings()
    rng = np.random.default_rng(4)
    n, cap = 500, 2000
    parts = kpt.empty_particles(cap)
    parts["status"][:] = 2
    sub = random_particles(n, s, rng, p_range=(0.5, 20.0))
    for k in sub:
        parts[k][:n] = sub[k]
    parts["split_level"][:n] = rng.integers(0, 4, n)
    parts["weight"][:n] = 0.5 ** parts["split_level"][:n]
    parts["status"][:n][rng.uniform(size=n) < 0.1] = 2
    before = copy_parts(parts)
    active = before["status"][:n] == 0
    w_before = before["weight"][:n][active].sum()
    thr = s.pmin_split * s.split_ratio ** before["split_level"][:n]
    cand = active & (before["p"][:n] > thr) & (before["p"][:n] <= s.pmax)
    new_n, split, skipped = O.split(s, parts, n, cap)
    assert split == cand.sum() and skipped == 0 and new_n == n + split
    act_after = parts["status"][:new_n] == 0
    assert parts["weight"][:new_n][act_after].sum() == pytest.approx(w_before, rel=0, abs=1e-12)
    kids = np.arange(n, new_n)
    parents = np.flatnonzero(cand)
    np.testing.assert_array_equal(parts["x"][kids], before["x"][parents])
    np.testing.assert_array_equal(parts["split_level"][kids], before["split_level"][parents] + 1)
    np.testing.assert_array_equal(parts["weight"][kids], 0.5 ** (before["split_level"][parents] + 1))
    # child ids: parent path with bit `old level` set
    pid = before["id"][parents]
    want = (pid & np.uint64(0xFFFFFFFF)) | (((pid >> np.uint64(32)) | (np.uint64(1) << before["split_level"][parents].astype(np.uint64))) << np.uint64(32))
    np.testing.assert_array_equal(parts["id"][kids], want)
    assert len(np.unique(parts["id"][:new_n])) == new_n


# SplitMix64 (Steele, Lea and Flood 2014) from state 0: its first outputs are the
# finalizer of 1 and 2 golden gammas.
SPLITMIX64_SEED0 = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
GOLDEN_GAMMA = 0x9E3779B97F4A7C15


def test_split_child_id_deep_levels(oracle_mod):
    """Below 32 levels a child's id sets the parent's path bit; from level 32 on it
    is the splitmix64 finalizer of pid + (level + 1) * golden gamma, pinned by
    the generator's published outputs from state 0."""
    m64 = (1 << 64) - 1
    for k, want in enumerate(SPLITMIX64_SEED0):
        level = 32 + k
        # pid + (level + 1) g = (k + 1) g  (mod 2^64)
        pid = (-(level - k) * GOLDEN_GAMMA) & m64
        assert O.split_child_id(pid, level) == want
    root, path = 123456789, 0b1011
    pid = root | (path << 32)
    for level in (0, 5, 31):
        assert O.split_child_id(pid, level) == root | ((path | (1 << level)) << 32)
    # a tree split 40 levels deep: every particle keeps its own id
    ids = [root]
    for level in range(40):
        ids += [O.split_child_id(i, level) for i in ids[-(2 ** min(level, 10)):]]
    assert len(set(ids)) == len(ids)


def test_split_capacity_clamp(oracle_mod):
    s = _split_settings()
    n, cap = 10, 13
    parts = kpt.empty_particles(cap)
    parts["status"][:] = 2
    sub = random_particles(n, s, np.random.default_rng(0))
    for k in sub:
        parts[k][:n] = sub[k]
    parts["p"][:n] = 5.0  # all split candidates
    new_n, split, skipped = O.split(s, parts, n, cap)
    assert (new_n, split, skipped) == (13, 3, 7)


def test_histogram_totals_and_edges(oracle_mod):
    s = _split_settings()
    rng = np.random.default_rng(9)
    n = 5000
    parts = random_particles(n, s, rng, p_range=(0.005, 20.0))
    parts["weight"][:] = 0.5 ** rng.integers(0, 5, n)
    parts["status"][rng.uniform(size=n) < 0.05] = 2
    h = O.histogram_logp(parts, n, 128, 0.01, 10.0)
    act = parts["status"] == 0
    inr = act & (parts["p"] >= 0.01) & (parts["p"] <= 10.0)
    assert h.sum() == pytest.approx(parts["weight"][inr].sum(), abs=1e-12)
    # numpy binning of the same rule
    idx = np.floor((np.log(parts["p"][inr]) - np.log(0.01)) * (128 / (np.log(10.0) - np.log(0.01))))
    idx = np.clip(idx.astype(int), 0, 127)
    np.testing.assert_allclose(h, np.bincount(idx, parts["weight"][inr], minlength=128), atol=1e-12)


def test_summary_matches_numpy(oracle_mod):
    s = _split_settings()
    rng = np.random.default_rng(1)
    n = 3000
    parts = random_particles(n, s, rng)
    parts["step_dt"][:] = rng.uniform(1e-6, 1e-4, n)
    parts["split_level"][:] = rng.integers(0, 7, n)
    parts["status"][rng.uniform(size=n) < 0.2] = 2
    sm = O.summary(parts, n)
    a = parts["status"] == 0
    assert sm["stored"] == n and sm["active"] == a.sum() and sm["inactive"] == n - a.sum()
    assert sm["p_min"] == parts["p"][a].min() and sm["p_max"] == parts["p"][a].max()
    assert sm["p_avg"] == pytest.approx(parts["p"][a].mean(), rel=1e-13)
    assert sm["dt_avg"] == pytest.approx(parts["step_dt"][a].mean(), rel=1e-13)
    assert sm["max_split_level"] == parts["split_level"][a].max()


def test_inject_region_and_determinism(oracle_mod):
    s = _split_settings()
    o = kpt.InjectOpts()
    o.use_region, o.fixed_time = 1, 0
    o.region_x_lo, o.region_x_hi, o.region_y_lo, o.region_y_hi = -0.5, 0.25, 0.1, 0.2
    cap = 20000
    a = kpt.empty_particles(cap)
    b = kpt.empty_particles(cap)
    n1 = O.inject(s, o, 15000, 1.0, 7, 42, 0, a, 0, cap)
    n2 = O.inject(s, o, 15000, 1.0, 7, 42, 0, b, 0, cap)
    assert n1 == n2 == 15000
    for k in a:
        np.testing.assert_array_equal(a[k], b[k])
    x, y, t, mu = a["x"][:n1], a["y"][:n1], a["t"][:n1], a["mu"][:n1]
    assert x.min() >= -0.5 and x.max() < 0.25 and y.min() >= 0.1 and y.max() < 0.2
    assert t.min() >= 1.0 and t.max() < 1.0 + s.mhd_dt
    assert abs(mu.mean()) < 0.02 and np.all(np.abs(mu) < s.mu_max)
    assert np.all(a["p"][:n1] == s.p0) and np.all(a["weight"][:n1] == 1.0)
    # capacity clamp
    assert O.inject(s, o, 10000, 1.0, 8, 42, n1, a, n1, cap) == cap - n1
