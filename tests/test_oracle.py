"""Pins the CPU oracle: against the reference's own recorded experiment traces
(data/*.csv exported by the reference while running its saved scenes), against
closed forms, and against brute force.  CPU only."""
import ctypes as C
import math
import os

import numpy as np
import pytest

import orc
from psim import abi, scenes
from psim.scripts import ScriptReplay
from psim.settings import Settings
from psim.state import State

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _load(scene):
    s = Settings()
    st = State.load_file(os.path.join(GOLDEN, "saved_states", scene + ".bin"), s)
    return s, st


def _replay(scene, mode):
    s, st = _load(scene)
    eng = orc.OracleEngine(st, s, mode=mode)
    rep = ScriptReplay(eng, st.scripts)
    return {p.replace("\\", "/").split("/")[-1][:-4]: rows for p, rows in rep.run()}


@pytest.fixture(scope="module")
def traces():
    return np.load(os.path.join(GOLDEN, "traces.npz"))


def test_texp1_bit_exact(traces):
    """TEXP_1 (two discs, parallel bond pulled apart at 1e-3 u/s, f64 distances):
    every value of every row equals the reference's export."""
    got = _replay("TEXPS", orc.MODE_REFERENCE)["TEXP_1"]
    ref = traces["TEXP_1"]
    assert got.shape[0] == ref.shape[0] == 101
    assert np.array_equal(got[:, :11], ref[:, :11])


@pytest.mark.parametrize("mode", [orc.MODE_REFERENCE, orc.MODE_CANONICAL])
@pytest.mark.parametrize("scene,cols,tol", [
    ("TEXPS", range(1, 11), 1e-3),        # all columns incl. force/moment sums
    ("PBEXPS", range(1, 7), 1e-4),        # kinematics (this export predates the bond data columns)
    ("Chain_EXP", range(1, 7), 1e-4),
    ("Chain_TEXP", range(1, 11), 1e-3),
    ("CHTEXPDET", range(1, 11), 1e-3),
])
def test_traces_follow_reference(traces, scene, cols, tol, mode):
    """All 13 experiments of each scene: the oracle's trace follows the
    reference's within `tol` of each column's range (the residue is the
    reference GPU's own float evaluation order)."""
    got = _replay(scene, mode)
    assert len(got) >= 13
    for stem, rows in got.items():
        ref = traces[stem]
        n = min(len(ref), len(rows))
        assert n >= 100
        for c in cols:
            scale = max(np.abs(ref[:n, c]).max(), 1e-30)
            err = np.abs(ref[:n, c] - rows[:n, c]).max() / scale
            assert err < tol, (stem, c, err)


def test_chtexpdet_reproduces_determinism_runs(traces):
    """CHTEXPDET.bin (the chain tension experiments, f64 distances): its first
    experiment's position and velocity columns equal the reference's first recorded
    run (data/Chain_DET1.csv) bit for bit, and lie as close to the second recorded run
    (Chain_DET2.csv) as that run lies to the first."""
    got = _replay("CHTEXPDET", orc.MODE_CANONICAL)["Chain_T_1"]
    d1, d2 = traces["Chain_DET1"], traces["Chain_DET2"]
    n = min(len(got), len(d1), len(d2))
    assert n == 501
    assert np.array_equal(got[:n, 1:7].astype(np.float32), d1[:n, 1:7].astype(np.float32))
    for c in range(1, 7):
        spread = np.abs(d2[:n, c] - d1[:n, c]).max()
        assert np.abs(d2[:n, c] - got[:n, c]).max() <= spread


def test_texp_closed_form(traces):
    """Parallel-bond tension: Δfn = kn A v dt per step, break at the closed-form step
    floor(σ A / (kn A v dt)) + 1 (2D_Simulation.wgsl:296, :322)."""
    rows = _replay("TEXPS", orc.MODE_CANONICAL)["TEXP_1"]
    kn, A, v, dt = 1e11, 2 * 0.5 * 0.05, 1e-3, 4.999000339012127e-06
    dfn = kn * A * v * dt
    k_break = math.floor(1e4 * A / dfn) + 1
    assert k_break == 21
    fn = rows[1:, 7]
    np.testing.assert_allclose(fn[: k_break - 1], dfn * np.arange(1, k_break), rtol=1e-5)
    assert np.all(fn[k_break - 1:] == 0.0)


def _granular(n_side=16, spacing=1.8, seed=4):
    """A box under the reference's own contact law (no dashpot)."""
    return scenes.granular_box(n_side=n_side, seed=seed, spacing=spacing, restitution=0.0)


def test_canonical_equals_reference_pairs_and_forces():
    """The canonical slot order (what the HIP path implements) and the WGSL's own
    slot rules give the same contact set and per-particle forces within 1e-5 of
    the force scale after one step."""
    s, st = _granular()
    a = orc.OracleEngine(st, s, mode=orc.MODE_REFERENCE, grid=s.grid_info())
    b = orc.OracleEngine(st, s, mode=orc.MODE_CANONICAL)
    a.step(1)
    b.step(1)
    n = st.p_count
    ca = a.canonical_contacts().view(np.int32).reshape(-1, 14, 6)[:n, :, 1]
    cb = b.canonical_contacts().view(np.int32).reshape(-1, 14, 6)[:n, :, 1]
    assert np.array_equal(ca, cb)
    assert (ca >= 0).sum() > n // 4
    fa = a.st.acc[: 2 * n].astype(np.float64)
    fb = b.st.acc[: 2 * n].astype(np.float64)
    assert np.max(np.abs(fa - fb)) <= 1e-5 * np.max(np.abs(fa))
    np.testing.assert_allclose(b.st.pos[: 2 * n], a.st.pos[: 2 * n], rtol=1e-5, atol=1e-6)


def test_multistep_energy_matches_reference_mode():
    """Chaotic multi-step run: kinetic energy of canonical vs reference slot order
    within 1% over 200 steps."""
    s, st = _granular(spacing=1.7)
    a = orc.OracleEngine(st, s, mode=orc.MODE_REFERENCE, grid=s.grid_info())
    b = orc.OracleEngine(st, s, mode=orc.MODE_CANONICAL)
    mats = abi.materials_array(s.materials)
    n = st.p_count
    for _ in range(4):
        a.step(50)
        b.step(50)
        ke = [orc.lib().orc_kinetic_energy(n, e.st.vel.ctypes.data_as(orc.FP), e.st.rot_vel.ctypes.data_as(orc.FP),
                                           e.st.radii.ctypes.data_as(orc.FP),
                                           e.st.material_pointers.ctypes.data_as(orc.IP), mats, len(mats))
              for e in (a, b)]
        assert abs(ke[0] - ke[1]) <= 0.01 * abs(ke[0]) + 1e-12


def test_free_flight_and_wall_bounce():
    """One disc, no contacts: v_n = v0 - g n dt exactly as the two half-kicks sum,
    x advances by v dt; at the floor the velocity flips with factor -0.5
    (2D_LOM.wgsl:108-112)."""
    s, st = scenes.granular_box(n_side=1, seed=0, spacing=4.0, rmin=0.5, rmax=0.5)
    st.pos[:2] = [0.0, 0.0]
    st.vel[:2] = [0.0, -3.0]
    s.simulation.timestep = 1e-3
    e = orc.OracleEngine(st, s, mode=orc.MODE_CANONICAL)
    g = np.float32(np.float32(9.81) * np.float32(1.0))
    vy = np.float32(-3.0)
    ay = np.float32(0.0)
    dt = np.float32(1e-3)
    y = np.float32(0.0)
    bounced = False
    for k in range(400):
        e.step(1)
        # restatement of the two half-kicks and the drift in f32
        iv = np.float32(vy + np.float32(np.float32(ay * dt) * np.float32(0.5)))
        dy = np.float32(iv * dt)
        ny = np.float32(y + dy)
        if ny - np.float32(0.5) < np.float32(-2.0):
            iv = np.float32(-iv * np.float32(0.5))
            y = np.float32(y + np.float32(np.float32(-2.0) - np.float32(ny - np.float32(0.5))))
            bounced = True
        y = np.float32(y + dy)
        mass = np.float32(np.float32(np.float32(1.0) * np.float32(np.pi)) * np.float32(0.5)) * np.float32(0.5)
        fy = np.float32(np.float32(0.0) + np.float32(-g * mass))
        ay = np.float32(fy / mass)
        vy = np.float32(iv + np.float32(np.float32(np.float32(0.5) * ay) * dt))
        assert e.st.pos[1] == y and e.st.vel[1] == vy, k
    assert bounced


def test_energy_conserved_frictionless():
    """μ = 0, no gravity, no walls reached: kinetic + spring energy conserved to O(dt)."""
    s, st = _granular(n_side=8, spacing=1.75, seed=7)
    s.physics.gravity = False
    s.physics.friction_coefficient = 0.0
    s.simulation.hor_bound = s.simulation.vert_bound = 200.0
    st.vel[: 2 * st.p_count] = np.random.default_rng(1).normal(0, 2.0, 2 * st.p_count).astype(np.float32)
    e = orc.OracleEngine(st, s, mode=orc.MODE_CANONICAL)
    mats = abi.materials_array(s.materials)
    n = st.p_count

    def energy():
        ke = orc.lib().orc_kinetic_energy(n, e.st.vel.ctypes.data_as(orc.FP), e.st.rot_vel.ctypes.data_as(orc.FP),
                                          e.st.radii.ctypes.data_as(orc.FP),
                                          e.st.material_pointers.ctypes.data_as(orc.IP), mats, len(mats))
        p = e.st.pos[: 2 * n].reshape(-1, 2).astype(np.float64)
        r = e.st.radii[:n].astype(np.float64)
        d = np.sqrt(((p[:, None, :] - p[None, :, :]) ** 2).sum(-1))
        ov = np.maximum(r[:, None] + r[None, :] - d, 0.0)
        np.fill_diagonal(ov, 0.0)
        kn = 1.0 / (1.0 / 2e5 + 1.0 / 2e5)
        return ke + 0.5 * kn * (ov ** 2).sum() / 2
    e0 = energy()
    e.step(2000)
    e1 = energy()
    assert abs(e1 - e0) < 0.02 * e0


@pytest.mark.parametrize("factor,gap", [(1.05, 0.0), (1.0, 0.08)])
def test_regen_bonds_bruteforce(factor, gap):
    """State::regen_bonds restated (oracle) equals an independent numpy pair scan, with
    the reference's rule (factor 1.05) and with a gap rule (every pair closer than
    0.08, the bonded-assembly rule); for the gap rule the pairs also agree with the
    f64 geometry away from the threshold's f32 rounding."""
    rng = np.random.default_rng(3)
    n = 60
    pos = rng.uniform(-6, 6, 2 * n).astype(np.float32)
    r = rng.uniform(0.8, 1.2, n).astype(np.float32)
    bonds = np.zeros(3 * 400, dtype=np.int32)
    cont = np.zeros(14 * 6 * n, dtype=np.float32)
    nb = orc.lib().orc_regen_bonds(n, pos.ctypes.data_as(orc.FP), r.ctypes.data_as(orc.FP), 6, 14,
                                   bonds.ctypes.data_as(orc.IP), 400, cont.ctypes.data_as(orc.FP), factor, gap)
    want = []
    for i in range(n):
        c = 0
        for j in range(i + 1, n):
            dx = np.float32(pos[2 * j] - pos[2 * i])
            dy = np.float32(pos[2 * j + 1] - pos[2 * i + 1])
            d = np.sqrt(np.float32(dx * dx + dy * dy))
            if d < np.float32(np.float32(r[i] + r[j]) * np.float32(factor)) + np.float32(gap):
                if c < 6:
                    want.append((i, j, d))
                    c += 1
            if gap > 0 and c < 6:
                g64 = np.hypot(float(pos[2 * j]) - float(pos[2 * i]), float(pos[2 * j + 1]) - float(pos[2 * i + 1])) \
                    - float(r[i]) - float(r[j])
                if abs(g64 - gap) > 1e-5:
                    assert ((i, j, d) in want) == (g64 < gap)
    assert nb == len(want) > 20
    tab = cont.view(np.int32).reshape(n, 14, 6)
    for k, (i, j, d) in enumerate(want):
        assert bonds[3 * k] == 1 and bonds[3 * k + 2] == np.float32(d).view(np.int32)
        assert k in tab[i, :, 5] and k in tab[j, :, 5]


def test_stage_closed_forms_bruteforce():
    """Stage 1-3 closed forms against a direct Python evaluation on a tiny grid."""
    rng = np.random.default_rng(9)
    n = 200
    pos = rng.uniform(-9, 9, 2 * n).astype(np.float32)
    r = rng.uniform(0.2, 0.5, n).astype(np.float32)
    r[5] = 0.0
    g = abi.DemGrid(np.float32(1.25), 0, 16, 14)
    keys = np.zeros(n, dtype=np.int32)
    orc.lib().orc_cell_keys(n, n, pos.ctypes.data_as(orc.FP), r.ctypes.data_as(orc.FP), C.byref(g),
                            keys.ctypes.data_as(orc.IP))
    cs = np.float32(1.25)
    inv = np.float32(np.float32(1.0) / cs)
    bx = np.float32(np.float32(-cs * np.float32(16)) * np.float32(0.5))
    by = np.float32(np.float32(cs * np.float32(14)) * np.float32(0.5))
    for i in range(n):
        if r[i] == 0:
            assert keys[i] == 16 * 14
            continue
        cx = min(max(int(np.float32(np.float32(pos[2 * i] - bx) * inv)), 0), 15)
        cy = min(max(int(np.float32(np.float32(by - pos[2 * i + 1]) * inv)), 0), 13)
        assert keys[i] == cy * 16 + cx
    order = np.zeros(n, dtype=np.int32)
    start = np.zeros(16 * 14 + 1, dtype=np.int32)
    orc.lib().orc_sort_cells(n, keys.ctypes.data_as(orc.IP), 16 * 14, order.ctypes.data_as(orc.IP),
                             start.ctypes.data_as(orc.IP))
    assert np.array_equal(order, np.lexsort((np.arange(n), keys)))
    assert np.array_equal(start, np.searchsorted(np.sort(keys), np.arange(16 * 14 + 1)))


def test_cell_list_pairs_equal_bruteforce():
    """The O(N) cell-list pair search (used on millions of discs) against the O(N^2)
    oracle scan and an independent numpy evaluation of the f32 predicate
    sqrt(dx*dx + dy*dy) < r_b + r_a; clustered, polydisperse discs, one inactive."""
    rng = np.random.default_rng(21)
    n = 1500
    c = rng.uniform(-20, 20, (30, 2))
    p = (c[rng.integers(0, 30, n)] + rng.normal(0, 2.5, (n, 2))).astype(np.float32)
    pos = p.reshape(-1).copy()
    r = rng.uniform(0.1, 0.9, n).astype(np.float32)
    r[7] = 0.0
    got = []
    for fn in (orc.lib().orc_contact_pairs, orc.lib().orc_contact_pairs_grid):
        cnt = np.zeros(n, dtype=np.int32)
        part = np.zeros(n * 64, dtype=np.int32)
        tot = fn(n, pos.ctypes.data_as(orc.FP), r.ctypes.data_as(orc.FP), 64, cnt.ctypes.data_as(orc.IP),
                 part.ctypes.data_as(orc.IP))
        assert tot == cnt.sum() > 0
        got.append((cnt, part.reshape(n, 64)))
    assert np.array_equal(got[0][0], got[1][0]) and np.array_equal(got[0][1], got[1][1])
    dx = (p[None, :, 0] - p[:, None, 0]).astype(np.float32)
    dy = (p[None, :, 1] - p[:, None, 1]).astype(np.float32)
    d = np.sqrt((dx * dx + dy * dy).astype(np.float32))
    hit = (d < (r[None, :] + r[:, None]).astype(np.float32)) & (r[None, :] > 0) & (r[:, None] > 0)
    np.fill_diagonal(hit, False)
    for a in range(n):
        want = np.nonzero(hit[a])[0]
        assert got[1][0][a] == len(want) and np.array_equal(got[1][1][a, :len(want)], want)


@pytest.mark.parametrize("e", [0.2, 0.5, 0.9])
def test_dashpot_restitution_closed_form(e):
    """The contact dashpot (an extension: the configurations' spring-dashpot law) against
    its closed form: two equal discs meeting head on rebound with relative speed e times
    the approach speed, zeta = -ln e / sqrt(pi^2 + ln^2 e) being the damping ratio whose
    damped half-period gives exp(-pi zeta / sqrt(1 - zeta^2)) = e.  The discrete
    integrator is within 1 % at 600 steps per contact; without a dashpot the collision
    is elastic (the reference's law)."""
    def collide(rest):
        s, st = scenes.granular_box(n_side=2, seed=0)
        s.physics.gravity = False
        s.physics.friction_coefficient = 0.0
        s.physics.restitution = rest
        s.simulation.timestep = 1e-5
        s.simulation.hor_bound = s.simulation.vert_bound = 100.0
        st.p_count = 2
        st.radii[:] = 0.0
        st.radii[:2] = 0.5
        st.pos[:4] = [-0.5005, 0.0, 0.5005, 0.0]
        st.vel[:4] = [1.0, 0.0, -1.0, 0.0]
        eng = orc.OracleEngine(st, s, mode=orc.MODE_CANONICAL)
        eng.step(2000)
        v = eng.st.vel[:4].astype(np.float64)
        assert abs(eng.st.pos[0]) > 0.5005 and not (eng.st.contacts.view(np.int32)[1] >= 0)   # separated
        return (v[2] - v[0]) / 2.0
    assert abs(collide(0.0) - 1.0) < 1e-3
    assert abs(collide(e) - e) < 0.01 * e


def _crowded(big_r=2.0, n_small=20, slots_preload=()):
    """A disc of radius big_r ringed by n_small discs of radius 0.3, every one
    overlapping it; no gravity, no friction, at rest.  Disc 0 is the big one; its table
    is preloaded with plain contacts to the small discs in `slots_preload`."""
    s, st = scenes.granular_box(n_side=5, seed=0, restitution=0.0)
    s.physics.gravity = False
    s.physics.friction_coefficient = 0.0
    s.simulation.hor_bound = s.simulation.vert_bound = 64.0
    n = 1 + n_small
    st.p_count = n
    st.radii[:] = 0.0
    st.radii[0] = big_r
    st.radii[1:n] = 0.3
    ang = 2 * np.pi * np.arange(n_small) / n_small
    st.pos[:2 * n] = 0.0
    st.pos[2:2 * n:2] = ((big_r + 0.25) * np.cos(ang)).astype(np.float32)
    st.pos[3:2 * n:2] = ((big_r + 0.25) * np.sin(ang)).astype(np.float32)
    c = st.contacts.view(np.int32).reshape(-1, st.max_contacts, 6)
    for k, b in enumerate(slots_preload):
        c[0, k] = [0, b, 0, -1, -1, -1]
        c[b, 0] = [b, 0, 0, -1, -1, -1]
    return s, st


def test_contact_slots_capped_at_14_existing_kept():
    """The 14-slot table of 2D_Simulation.wgsl:125-227: a disc touching 20 others holds
    14 of them.  Slots already occupied are never evicted (a contact is deleted only when
    its partner is no longer found, :185-204) and new partners only take free slots
    (:206-227), here in ascending id: preloaded with the 14 largest ids, the table keeps
    them; from empty, it takes the 14 smallest.  Every small disc holds the big one."""
    for preload, want in ((list(range(7, 21)), list(range(7, 21))), ((), list(range(1, 15)))):
        s, st = _crowded(slots_preload=preload)
        eng = orc.OracleEngine(st, s, mode=orc.MODE_CANONICAL)
        eng.step(1)
        c = eng.canonical_contacts().view(np.int32).reshape(-1, 14, 6)
        assert c[0, :, 1].tolist() == want
        for b in range(1, 21):
            assert c[b, 0, 1] == 0 and c[b, 1, 1] == -1


def test_parallel_oracle_equals_serial():
    """The canonical mode runs LOM, the pair search and the Sim pass over OpenMP threads,
    with bond breaks queued and written in particle-id order: struck chains with an
    r = 8 striker (a large body in the pair search) through breakage, identical bit for
    bit on 1 and 6 threads."""
    s, st, strengths = scenes.bonded_chains(chains=8, per_chain=64, seed=5, striker_r=8.0, striker_v=40.0,
                                            striker_density=10.0, strength_scale=30.0)
    scenes.chain_bonds(st, 8, 64, strengths)
    runs = []
    prev = orc.lib().orc_set_threads(1)
    try:
        for threads in (1, 6):
            orc.lib().orc_set_threads(threads)
            e = orc.OracleEngine(st.copy(), s, mode=orc.MODE_CANONICAL)
            e.step(600)
            runs.append((e.st.copy(), e.canonical_contacts().copy()))
    finally:
        orc.lib().orc_set_threads(prev)
    (a, ca), (b, cb) = runs
    for name in ("pos", "vel", "acc", "rot", "rot_vel", "rot_acc", "data", "bonds"):
        assert np.array_equal(getattr(a, name).view(np.int32), getattr(b, name).view(np.int32)), name
    assert np.array_equal(ca.view(np.int32), cb.view(np.int32))
    assert (a.bonds[0:3 * a.n_bonds:3] < 0).sum() > 5
