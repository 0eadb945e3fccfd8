"""Pins the oracle's force, integrator, break and skin functions with closed
forms and invariants that do not retype the oracle's arithmetic:

* contact: Hertz law E*scale*delta^1.5 along the normal, linear attraction
  zone, zero beyond R+a, action/reaction, pure damping term;
* beam: the Euler-Bernoulli element stiffnesses (EA/L with A=W^2, 12EI/L^3,
  6EI/L^2, 4EI/L, 2EI/L with I=W^4/12, GJ/L with J=W^4/6), rigid translation
  and rigid rotation produce no internal-force increment, orthonormal frame,
  equal and opposite end forces, equivariance under rotations about z;
* integrator: exact constant-acceleration recurrence, terminal drift F*dt/c,
  buoyancy switch at the waterline, clamp;
* break: threshold crossing at load == maxLoad, broken links stay broken;
* skin: heuristic ignores negative Diff, exact test = max displacement.
"""
import ctypes as C
import math

import numpy as np
import pytest

from hidem_amd._abi import Links, Particles

E, W = 1e9, 0.5
FP = dict(invDt=1e4, attractDistance=0.01, attractScaleForce=1e-5, damping=1e4, dampingRotation=1e4,
          scaleE=1e9, E=E, W2=W * W, W4=W ** 4, G=E / 2)


def fparams(orc, **kw):
    d = dict(FP)
    d.update(kw)
    return orc.ForceParams(*[d[f] for f, _ in orc.ForceParams._fields_])


def contact(orc, dist_vec, self7, other7, fp):
    s = (C.c_double * 7)(*self7)
    o = (C.c_double * 7)(*other7)
    f = (C.c_double * 6)()
    orc.lib().hdmo_contact_pair(1.0, *dist_vec, s, o, C.byref(fp), f)
    return np.array(list(f))


def beam(orc, dist_vec, self7, other7, fp, ct=None, mask=1.0):
    s = (C.c_double * 7)(*self7)
    o = (C.c_double * 7)(*other7)
    c = (C.c_double * 6)(*(ct if ct is not None else [0] * 6))
    f = (C.c_double * 6)()
    orc.lib().hdmo_beam_side(mask, *dist_vec, s, o, C.byref(fp), c, f)
    return np.array(list(c)), np.array(list(f))


R0 = [0, 0, 0, 0, 0, 0, 0.5]


@pytest.mark.parametrize("delta", [1e-3, 0.02, 0.1])
def test_contact_hertz_overlap(orc, delta):
    fp = fparams(orc, damping=0.0, dampingRotation=0.0)
    d = 1.0 - delta
    f = contact(orc, (d, 0, 0), R0, R0, fp)
    expect = FP["scaleE"] * delta ** 1.5
    assert f[0] == pytest.approx(expect, rel=1e-13)
    assert np.all(f[1:] == 0)
    # arbitrary direction: force along the unit normal
    n = np.array([0.3, -0.5, 0.8])
    n /= np.linalg.norm(n)
    f = contact(orc, tuple(d * n), R0, R0, fp)
    np.testing.assert_allclose(f[:3], expect * n, rtol=1e-12, atol=0)


def test_contact_attraction_zone_and_cutoff(orc):
    fp = fparams(orc, damping=0.0, dampingRotation=0.0)
    a = FP["attractDistance"]
    gap = 0.4 * a
    f = contact(orc, (1.0 + gap, 0, 0), R0, R0, fp)
    assert f[0] == pytest.approx(FP["scaleE"] * FP["attractScaleForce"] * (-gap), rel=1e-13)
    assert f[0] < 0  # pulls towards the partner
    f = contact(orc, (1.0 + a, 0, 0), R0, R0, fp)
    assert np.all(f == 0)
    f = contact(orc, (1.0 + 3 * a, 0, 0), R0, R0, fp)
    assert np.all(f == 0)


def test_contact_action_reaction_and_damping(orc):
    fp = fparams(orc)
    rng = np.random.default_rng(3)
    for _ in range(20):
        dvec = rng.normal(size=3)
        dvec *= rng.uniform(0.9, 1.005) / np.linalg.norm(dvec)
        s = list(rng.normal(scale=1e-5, size=6)) + [0.5]
        o = list(rng.normal(scale=1e-5, size=6)) + [0.5]
        fij = contact(orc, tuple(dvec), s, o, fp)
        fji = contact(orc, tuple(-dvec), o, s, fp)
        np.testing.assert_allclose(fij, -fji, rtol=1e-12, atol=1e-9)
    # pure damping inside the active zone: -c (v_i - v_j)
    s = [1e-4, 0, 0, 2e-4, 0, 0, 0.5]
    f = contact(orc, (1.005, 0, 0), s, R0, fp)
    hertz_attr = FP["scaleE"] * FP["attractScaleForce"] * (-0.005)
    assert f[0] == pytest.approx(-FP["damping"] * 1e-4 * FP["invDt"] + hertz_attr, rel=1e-12)
    assert f[3] == pytest.approx(-FP["dampingRotation"] * 2e-4 * FP["invDt"], rel=1e-12)


# ---- beam --------------------------------------------------------------------
L = 1.04
I = W ** 4 / 12.0
J = W ** 4 / 6.0  # the reference's torsion constant (K2 = G W^4 / (6 L))


def zero_damping(orc):
    return fparams(orc, damping=0.0, dampingRotation=0.0)


def sd(dx=(0, 0, 0), dth=(0, 0, 0)):
    return list(dx) + list(dth) + [0.5]


def test_beam_axial(orc):
    eps = 1e-5
    ct, f = beam(orc, (L, 0, 0), sd(dx=(eps, 0, 0)), sd(), zero_damping(orc))
    assert ct[0] == pytest.approx(E * W * W / L * eps, rel=1e-13)   # EA/L
    assert f[0] == pytest.approx(-E * W * W / L * eps, rel=1e-13)   # restoring
    assert np.all(ct[1:] == 0) and np.all(f[1:] == 0)


def test_beam_shear_and_bending_stiffness(orc):
    eps = 1e-5
    ct, f = beam(orc, (L, 0, 0), sd(dx=(0, eps, 0)), sd(), zero_damping(orc))
    assert ct[1] == pytest.approx(12 * E * I / L ** 3 * eps, rel=1e-13)
    assert abs(ct[5]) == pytest.approx(6 * E * I / L ** 2 * eps, rel=1e-13)
    phi = 1e-5
    ct, f = beam(orc, (L, 0, 0), sd(dth=(0, 0, phi)), sd(), zero_damping(orc))
    assert abs(ct[1]) == pytest.approx(6 * E * I / L ** 2 * phi, rel=1e-13)
    assert ct[5] == pytest.approx(4 * E * I / L * phi, rel=1e-13)
    ct, f = beam(orc, (L, 0, 0), sd(), sd(dth=(0, 0, phi)), zero_damping(orc))
    assert ct[5] == pytest.approx(2 * E * I / L * phi, rel=1e-13)
    # the y-bending plane has the same magnitudes
    ct, f = beam(orc, (L, 0, 0), sd(dth=(0, phi, 0)), sd(), zero_damping(orc))
    assert ct[4] == pytest.approx(4 * E * I / L * phi, rel=1e-13)
    assert abs(ct[2]) == pytest.approx(6 * E * I / L ** 2 * phi, rel=1e-13)


def test_beam_torsion(orc):
    phi = 1e-5
    ct, f = beam(orc, (L, 0, 0), sd(dth=(phi, 0, 0)), sd(), zero_damping(orc))
    assert ct[3] == pytest.approx(FP["G"] * J / L * phi, rel=1e-13)
    assert f[3] == pytest.approx(-FP["G"] * J / L * phi, rel=1e-13)


def random_dir(rng):
    v = rng.normal(size=3)
    return v / np.linalg.norm(v)


def test_beam_rigid_translation_no_strain(orc):
    rng = np.random.default_rng(5)
    for _ in range(10):
        d = tuple(L * random_dir(rng))
        u = list(rng.normal(scale=1e-5, size=3))
        ct, f = beam(orc, d, sd(dx=u), sd(dx=u), zero_damping(orc))
        assert np.max(np.abs(ct)) <= 1e-13 * E * W * W * 1e-5


def test_beam_rigid_rotation_no_strain(orc):
    # A small rigid rotation phi about axis k moves p by phi k x d and turns
    # both ends by phi; the reference stores rotations with the opposite sign
    # (theta = -phi) -- in that convention the element sees no deformation.
    rng = np.random.default_rng(6)
    scale = E * W ** 4 / L ** 2 * 1e-6
    for _ in range(10):
        d = L * random_dir(rng)
        k = random_dir(rng)
        phi = 1e-6
        dxp = phi * np.cross(k, d)
        th = -phi * k
        ct, f = beam(orc, tuple(d), sd(dx=dxp, dth=th), sd(dth=th), zero_damping(orc))
        assert np.max(np.abs(ct)) <= 1e-9 * scale


def frame(d):
    x, y, z = d
    li = math.sqrt(x * x + y * y + z * z)
    lp = math.sqrt(x * x + y * y)
    return np.array([[x / li, y / li, z / li], [-y / lp, x / lp, 0], [-z * x / (li * lp), -z * y / (li * lp), lp / li]])


def test_beam_frame_orthonormal_and_end_balance(orc):
    rng = np.random.default_rng(7)
    for _ in range(20):
        d = L * random_dir(rng)
        T = frame(d)
        np.testing.assert_allclose(T @ T.T, np.eye(3), atol=1e-14)
        up = list(rng.normal(scale=1e-5, size=6)) + [0.5]
        uq = list(rng.normal(scale=1e-5, size=6)) + [0.5]
        ctp, fp_ = beam(orc, tuple(d), up, uq, zero_damping(orc))
        ctq, fq = beam(orc, tuple(-d), uq, up, zero_damping(orc))
        # equal and opposite end forces; the element carries no net moment.
        # Stored moments follow the reversed rotation convention (see the rigid
        # rotation test), so the physical balance M_p + M_q + d x F_p = 0 reads
        # -(M_p + M_q) + d x F_p = 0 in stored quantities.
        np.testing.assert_allclose(fp_[:3], -fq[:3], rtol=1e-10, atol=1e-9)
        moment = fp_[3:] + fq[3:] - np.cross(d, fp_[:3])
        assert np.max(np.abs(moment)) <= 1e-9 * max(1.0, np.max(np.abs(fp_)))


def test_beam_equivariant_about_z(orc):
    rng = np.random.default_rng(8)
    d = L * random_dir(rng)
    up = rng.normal(scale=1e-5, size=6)
    uq = rng.normal(scale=1e-5, size=6)
    ang = 0.7
    Rz = np.array([[math.cos(ang), -math.sin(ang), 0], [math.sin(ang), math.cos(ang), 0], [0, 0, 1]])
    ct1, f1 = beam(orc, tuple(d), list(up) + [0.5], list(uq) + [0.5], zero_damping(orc))
    rot = lambda u: list(Rz @ u[:3]) + list(Rz @ u[3:])  # noqa: E731
    ct2, f2 = beam(orc, tuple(Rz @ d), rot(up) + [0.5], rot(uq) + [0.5], zero_damping(orc))
    np.testing.assert_allclose(ct2, ct1, rtol=1e-9, atol=1e-6)
    np.testing.assert_allclose(f2[:3], Rz @ f1[:3], rtol=1e-9, atol=1e-6)
    np.testing.assert_allclose(f2[3:], Rz @ f1[3:], rtol=1e-9, atol=1e-6)


def test_beam_mask_only_gates_force(orc):
    # broken links keep integrating CT (HiDEMDefaultFunctions.h:181-202), force is masked
    ct, f = beam(orc, (L, 0, 0), sd(dx=(1e-5, 0, 0)), sd(), zero_damping(orc), mask=0.0)
    assert ct[0] != 0 and np.all(f == 0)


# ---- integrator ------------------------------------------------------------------
def one_particle(z=10.0, radius=0.5):
    p = Particles.zeros(1)
    p.arrays["xCord"][0] = 5.0
    p.arrays["zCord"][0] = z
    p.arrays["radius"][0] = radius
    return p


def iparams(orc, **kw):
    dt = kw.pop("dt", 1e-4)
    d = dict(variant=0, rhoIce=910.0, g=9.81, lift=(1030 - 910) / 910 * 9.81, waterLevel=0.0, airDrag=0.0,
             waterDrag=0.0, invDt=1 / dt, invDt2=1 / (2 * dt), dtSqr=dt * dt, invDtSqr=1 / (dt * dt), dt=dt,
             bedDrag=0, bedInteract=0, bedStiffness=0, frictionCap=0, bedZ=-1e300, bedXEnd=1e300, deepZ=-1e300,
             frictionXEnd=-1e300, maxVel=1e300, maxRot=1e300, velocityLimit=0)
    d.update(kw)
    ip = orc.IntegParams()
    for k, v in d.items():
        setattr(ip, k, v)
    return ip


def integrate(orc, p, ip, steps):
    ps = orc.particles_struct(p)
    for _ in range(steps):
        orc.lib().hdmo_integrate(C.byref(ps), C.byref(ip))


@pytest.mark.parametrize("variant", [0, 1])
def test_free_fall_exact(orc, variant):
    p = one_particle()
    dt, N = 1e-4, 200
    integrate(orc, p, iparams(orc, variant=variant, dt=dt), N)
    expect = 10.0 - 9.81 * dt * dt * N * (N + 1) / 2
    # Diff is re-derived as x_new - x_prev (MainIceBlock.cpp:394), so each step
    # carries ~ulp(10) of rounding into the velocity; N^2 ulp bounds the drift.
    ulp = np.spacing(10.0)
    assert abs(p["zCord"][0] - expect) <= 2 * N * N * ulp
    assert p["zDiff"][0] == pytest.approx(-9.81 * dt * dt * N, abs=4 * N * ulp)


def test_terminal_drift(orc):
    p = one_particle()
    c, dt = 5e6, 1e-4
    integrate(orc, p, iparams(orc, airDrag=c, waterDrag=c, dt=dt), 2000)
    m = 910.0
    assert p["zDiff"][0] == pytest.approx(-9.81 * m * dt / c, rel=1e-9)


def test_buoyancy_switch(orc):
    p = one_particle(z=-1.0)
    integrate(orc, p, iparams(orc, waterLevel=0.0), 1)
    assert p["zDiff"][0] > 0  # lift below the waterline
    p = one_particle(z=1.0)
    integrate(orc, p, iparams(orc, waterLevel=0.0), 1)
    assert p["zDiff"][0] < 0


def test_clamp(orc):
    p = one_particle()
    p.type[0] = 1
    p.arrays["zDiff"][0] = 0.3
    integrate(orc, p, iparams(orc), 5)
    assert p["zCord"][0] == 10.0 and p["zDiff"][0] == 0.0


def test_glacier_velocity_limit_and_bed(orc):
    p = one_particle(z=10.0)
    p.arrays["zDiff"][0] = -0.01  # 100 m/s downward
    ip = iparams(orc, variant=1, maxVel=20.0, maxRot=5.0, velocityLimit=1)
    integrate(orc, p, ip, 1)
    assert p["zDiff"][0] == pytest.approx(-20.0 * 1e-4, rel=1e-12)
    # bed penalty pushes a particle whose top is below the bed upward
    p = one_particle(z=-1.0)
    ip = iparams(orc, variant=1, bedZ=0.0, bedXEnd=1e300, bedStiffness=1e9, waterLevel=-100.0)
    integrate(orc, p, ip, 1)
    assert p["zDiff"][0] > 0


# ---- break ---------------------------------------------------------------------
def test_break_threshold(orc):
    p = Particles.zeros(2)
    p.arrays["xCord"][:] = [0.0, 1.0]
    p.arrays["radius"][:] = 0.5
    l = Links.zeros(1)
    l.a[0], l.b[0] = 0, 1
    l.initialLen[0] = 1.0
    thr = 3e-4
    l.maxLoad[0] = thr
    bp = orc.BreakParams(0.01, 0.0)
    ps, ls = orc.particles_struct(p), orc.links_struct(l)
    p.arrays["xCord"][1] = 1.0 + 0.9 * thr
    assert orc.lib().hdmo_break(C.byref(ps), C.byref(ls), C.byref(bp)) == 0 and l.state[0] == 1
    # bend term adds bendBreak * max |dtheta|
    p.arrays["yTheta"][1] = 0.2 * thr / 0.01
    assert orc.lib().hdmo_break(C.byref(ps), C.byref(ls), C.byref(bp)) == 1 and l.state[0] == 0
    p.arrays["xCord"][1] = 1.0
    p.arrays["yTheta"][1] = 0.0
    assert orc.lib().hdmo_break(C.byref(ps), C.byref(ls), C.byref(bp)) == 0 and l.state[0] == 0


def test_break_at_equality(orc):
    p = Particles.zeros(2)
    p.arrays["xCord"][:] = [0.0, 1.5]
    l = Links.zeros(1)
    l.a[0], l.b[0] = 0, 1
    l.initialLen[0] = 1.0
    l.maxLoad[0] = 0.5  # load = (1.5 - 1)/1 = 0.5 exactly
    bp = orc.BreakParams(0.0, 0.0)
    ps, ls = orc.particles_struct(p), orc.links_struct(l)
    assert orc.lib().hdmo_break(C.byref(ps), C.byref(ls), C.byref(bp)) == 1


# ---- skin ----------------------------------------------------------------------
def test_skin_heuristic_and_exact(orc):
    p = Particles.zeros(3)
    p.arrays["xDiff"][:] = [-5.0, 0.003, 0.001]
    p.arrays["yDiff"][:] = [-1.0, -2.0, -3.0]
    p.arrays["zDiff"][:] = [0.004, 0.0, 0.002]
    ps = orc.particles_struct(p)
    h = orc.lib().hdmo_skin_heuristic(C.byref(ps), 10)
    assert h == pytest.approx(10 * math.hypot(0.003, 0.004), rel=1e-15)
    x0 = np.zeros(3)
    p.arrays["xCord"][:] = [0.0, 0.3, 0.0]
    p.arrays["zCord"][:] = [0.0, 0.4, 0.1]
    dp = C.POINTER(C.c_double)
    e = orc.lib().hdmo_skin_exact(C.byref(ps), x0.ctypes.data_as(dp), x0.ctypes.data_as(dp), x0.ctypes.data_as(dp))
    assert e == pytest.approx(0.5, rel=1e-15)
