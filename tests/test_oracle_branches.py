"""Closed-form pins of the oracle's branches that no recorded reference trace covers:
bond laws 1 and 2 (2D_Simulation.wgsl:375-432), planet and mouse gravity (:445-453),
external forces (:440-441), the circular wall (2D_LOM.wgsl:130-145), the collision
interval counter (2D_LOM.wgsl:155-163) and the Coulomb cap of the contact law
(2D_Simulation.wgsl:361-368).  Each expectation is computed here from the mathematics
of the law (in f64, or f32 where the law is a single rounding), not by re-running the
oracle's own expressions.  CPU only."""
import math

import numpy as np
import pytest

import orc
from psim import scenes
from psim.compute import assemble_bonds

F = np.float32


def _scene(n, gravity=False):
    """A 2 x 2 granular box reused as a container for n <= 4 hand-placed discs: no
    gravity, no friction, material kn = ks = 2e5, density 1, dt 1e-3."""
    s, st = scenes.granular_box(n_side=2, seed=0, spacing=20.0, restitution=0.0)
    s.physics.gravity = gravity
    s.physics.planet_mode = False
    s.physics.friction_coefficient = 0.0
    s.simulation.timestep = F(1e-3)
    s.simulation.hor_bound = s.simulation.vert_bound = 100.0
    st.p_count = n
    st.radii[:] = 0.0
    st.pos[:] = 0.0
    st.vel[:] = 0.0
    return s, st


def _mass(r, rho=1.0):
    return F(F(F(F(rho) * F(np.pi)) * F(r)) * F(r))


def _pulled_pair(mode, kn_b, sn, cd=1.0):
    """Discs 0 and 1 (r = 0.5) touching at x = -0.5 and 0.5, bonded (bond law `mode`),
    moved apart at 0.5 units/s each by prescribed velocities (fixity 1)."""
    s, st = _scene(2)
    s.physics.bonds = mode
    s.physics.bond_tearing = True
    s.physics.bond_normal_stiffness = F(kn_b)
    s.physics.bond_shear_stiffness = F(kn_b)
    s.physics.bond_normal_strength = F(sn)
    s.physics.bond_shear_strength = F(1e9)
    s.physics.contact_damping = F(cd)
    st.radii[:2] = 0.5
    st.pos[:4] = [-0.5, 0.0, 0.5, 0.0]
    st.vel[:4] = [-0.5, 0.0, 0.5, 0.0]
    fx = st.fixity.reshape(-1, 6)
    fx[:2, 0:3] = 1
    st.bonds, st.contacts = assemble_bonds(st, np.array([[0, 1]], dtype=np.int32), st.max_contacts)
    return s, st


def _gap(st):
    return float(np.float64(st.pos[2]) - np.float64(st.pos[0]))


def test_normal_bonds_closed_form():
    """Bond law 1 (normal_bonds): the bonded pair pulls together with
    bond_normal_stiffness * (|d| - (r_a + r_b)) along the line of centres, and tears
    (code -10 + sign = -9) at the first step where that force reaches the normal
    strength; from then on no force acts (the discs are apart)."""
    kn_b, sn = 1.0e3, 40.0
    s, st = _pulled_pair(1, kn_b, sn)
    e = orc.OracleEngine(st, s, mode=orc.MODE_CANONICAL)
    m = float(_mass(0.5))
    broke_at = None
    for k in range(1, 81):
        e.step(1)
        disp = _gap(e.st) - 1.0
        idx = int(e.st.bonds[0])
        if broke_at is None and kn_b * disp >= sn * (1 - 1e-6):
            broke_at = k
        if broke_at is None:
            assert idx == 1
            assert abs(e.st.acc[0] - kn_b * disp / m) <= 1e-5 * kn_b * disp / m + 1e-6, k
            assert abs(e.st.acc[2] + kn_b * disp / m) <= 1e-5 * kn_b * disp / m + 1e-6, k
        else:
            assert idx == -9, k
            assert e.st.acc[0] == 0.0 and e.st.acc[2] == 0.0, k
    assert broke_at is not None and 20 < broke_at < 80


def test_linear_contact_bonds_closed_form():
    """Bond law 2 (linear_contact_bonds): normal force -(|d| - (r_a + r_b)) kn/2 (the two
    bond stiffnesses in series), scaled by contact_damping; data[0..2] are assigned
    (-nf, ft, moment), not added; tears when the tension passes the normal strength."""
    kn_b, sn, cd = 2.0e3, 30.0, 0.5
    s, st = _pulled_pair(2, kn_b, sn, cd)
    e = orc.OracleEngine(st, s, mode=orc.MODE_CANONICAL)
    m = float(_mass(0.5))
    kn = kn_b / 2.0
    broke = False
    for k in range(1, 81):
        e.step(1)
        disp = _gap(e.st) - 1.0
        tension = disp * kn
        if not broke and int(e.st.bonds[0]) < 0:
            broke = True
            assert int(e.st.bonds[0]) == -9 and tension > sn * (1 - 1e-6)
        if not broke:
            want = cd * tension / m
            assert abs(e.st.acc[0] - want) <= 1e-5 * want + 1e-6, k
            assert abs(e.st.data[0] - tension) <= 1e-5 * tension + 1e-7, k    # -nf, assigned
            assert e.st.data[1] == 0.0 and e.st.data[2] == 0.0
        else:
            assert e.st.acc[0] == 0.0
    assert broke


@pytest.mark.parametrize("mouse", [False, True])
def test_planet_gravity_closed_form(mouse):
    """Planet mode: a disc at rest is pulled toward the centre (the origin, or the mouse
    point with mouse gravity) at 9.81 * gravity_acc, whatever its mass; a disc at the
    centre feels nothing."""
    s, st = _scene(3, gravity=True)
    s.physics.planet_mode = True
    s.physics.mouse_gravity = mouse
    s.physics.gravity_acceleration = F(2.0)
    c = (1.0, 1.0) if mouse else (0.0, 0.0)
    s.world_pos = c
    st.radii[:3] = [0.5, 1.0, 0.7]
    st.pos[:6] = [3.0, 4.0, -6.0, 2.0, c[0], c[1]]
    e = orc.OracleEngine(st, s, mode=orc.MODE_CANONICAL)
    e.step(1)
    g = 9.81 * 2.0
    for i in range(2):
        dx, dy = c[0] - float(st.pos[2 * i]), c[1] - float(st.pos[2 * i + 1])
        l = math.hypot(dx, dy)
        assert abs(e.st.acc[2 * i] - g * dx / l) <= 2e-6 * g
        assert abs(e.st.acc[2 * i + 1] - g * dy / l) <= 2e-6 * g
    assert e.st.acc[4] == 0.0 and e.st.acc[5] == 0.0


def test_external_forces_closed_form():
    """External force and moment (forces[id].{x, y, rot}): a = F / m and alpha = M / I
    with m = rho pi r^2 and I = m r^2 / 2, each one f32 rounding of the law."""
    s, st = _scene(1)
    st.radii[0] = 0.8
    f = st.forces.reshape(-1, 6)
    f[0, :3] = [3.0, -7.0, 0.25]
    e = orc.OracleEngine(st, s, mode=orc.MODE_CANONICAL)
    e.step(1)
    m = _mass(0.8)
    inertia = F(F(F(0.5) * m) * F(0.8)) * F(0.8)
    assert e.st.acc[0] == F(F(3.0) / m) and e.st.acc[1] == F(F(-7.0) / m)
    assert e.st.rot_acc[0] == F(F(0.25) / inertia)


def test_circular_wall_closed_form():
    """Round walls: a disc driven through the circle of radius R is put back on it (its
    centre at R - r from the middle) and leaves with minus half its normal velocity;
    the tangential part is dropped (the LOM sets the velocity to -0.5 (v.n) n)."""
    for ang in (math.pi / 2, math.pi / 4):
        s, st = _scene(1)
        s.simulation.round_walls = True
        s.simulation.wall_radius = F(10.0)
        st.radii[0] = 0.5
        nx, ny = math.cos(ang), math.sin(ang)
        st.pos[:2] = [F(9.0 * nx), F(9.0 * ny)]
        v = np.array([30.0 * nx + 5.0 * ny, 30.0 * ny - 5.0 * nx])   # outward plus a tangential part
        st.vel[:2] = v.astype(np.float32)
        e = orc.OracleEngine(st, s, mode=orc.MODE_CANONICAL)
        checked = False
        for k in range(60):
            prev = e.st.pos[:2].astype(np.float64)
            e.step(1)
            p = e.st.pos[:2].astype(np.float64)
            if e.st.vel[0] * nx + e.st.vel[1] * ny < 0:      # the step of the bounce
                # the wall normal where the drift would have taken the disc
                q = prev + v * 1e-3
                n = q / np.hypot(*q)
                got = e.st.vel[:2].astype(np.float64)
                assert np.allclose(got, -0.5 * float(np.dot(v, n)) * n, rtol=1e-4, atol=1e-4)
                assert abs(float(np.hypot(*p)) - 9.5) < 2e-5
                checked = True
                break
            assert np.hypot(*p) <= 9.5 + 1e-5, k      # never past the wall
        assert checked


@pytest.mark.parametrize("interval", [1, 3])
def test_collision_interval_closed_form(interval):
    """The narrow phase runs when invocation 0's counter reaches collision_interval
    (2D_LOM.wgsl:155-163): from a zero counter that is step interval, 2 interval, ...
    Two discs closing head-on get their contact slot at the first such step on which
    they overlap, and no force before it."""
    s, st = _scene(2)
    s.physics.collision_interval = interval
    st.radii[:2] = 0.5
    st.pos[:4] = [-0.52, 0.0, 0.52, 0.0]
    st.vel[:4] = [1.0, 0.0, -1.0, 0.0]
    e = orc.OracleEngine(st, s, mode=orc.MODE_CANONICAL)
    first_overlap = None
    for k in range(1, 40):
        e.step(1)
        overlap = _gap(e.st) < 1.0
        if overlap and first_overlap is None:
            first_overlap = k
        slot = e.st.contacts.view(np.int32).reshape(-1, 14, 6)[0, :, 1]
        has = bool((slot == 1).any())
        expect_from = None if first_overlap is None else \
            next(q for q in range(first_overlap, first_overlap + interval) if q % interval == 0)
        if expect_from is None or k < expect_from:
            assert not has and e.st.acc[0] == 0.0, k
        else:
            assert has, k
            break
    assert first_overlap is not None


def test_coulomb_cap_two_discs():
    """Incremental tangential spring with the Coulomb cap (linear_model): disc 1 held at a
    fixed overlap with disc 0 slides past it at constant speed; each step adds ks_eff times
    the relative tangential displacement to ft until |ft| reaches mu |fn| = mu kn_eff d,
    where it stays; the moment on disc 0 is -r_a ft.  Expectations from the positions in
    f64 (fn = kn_eff (r_a + r_b - |d|), t = perpendicular of the unit normal)."""
    s, st = _scene(2)
    s.physics.friction_coefficient = F(0.3)
    st.radii[:2] = 0.5
    st.pos[:4] = [0.0, 0.0, 0.99, 0.0]
    st.vel[:4] = [0.0, 0.0, 0.0, 0.2]
    fx = st.fixity.reshape(-1, 6)
    fx[:2, 0:3] = 1
    e = orc.OracleEngine(st, s, mode=orc.MODE_CANONICAL)
    kn = ks = 1e5
    ft = 0.0
    capped = False
    for k in range(1, 60):
        e.step(1)
        p = e.st.pos[:4].astype(np.float64)
        d = p[0:2] - p[2:4]
        l = float(np.hypot(*d))
        n = d / l
        t = np.array([-n[1], n[0]])
        fn = kn * (1.0 - l)
        ut = float(np.dot(np.array([0.0, 0.2 * 1e-3]), t))     # disc 1 moves, disc 0 stays
        ft = max(min(ft + ut * ks, 0.3 * fn), -0.3 * fn)
        capped = capped or abs(ft) >= 0.3 * fn * (1 - 1e-9)
        assert abs(e.st.data[0] - fn) <= 1e-3 * fn, k
        assert abs(e.st.data[1] - ft) <= 1e-3 * abs(ft) + 1e-6, k
        assert abs(e.st.data[2] + 0.5 * e.st.data[1]) <= 1e-6 * abs(e.st.data[1]) + 1e-9, k
    assert capped
