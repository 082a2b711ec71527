"""Parity of the HIP path (through the C ABI) with the CPU oracle and the
reference's recorded traces.  Needs an MI355X."""
import os

import numpy as np
import pytest

import orc
from psim import scenes
from psim.compute import ComputeProg
from psim.scripts import ScriptReplay
from psim.settings import Settings
from psim.state import State

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
STATE_FIELDS = ("pos", "vel", "acc", "rot", "rot_vel", "rot_acc")


def _n(name, n):
    return 2 * n if name in ("pos", "vel", "acc") else n


def assert_bits_equal(got: State, want: State, n, fields=STATE_FIELDS):
    for name in fields:
        a = getattr(got, name)[: _n(name, n)]
        b = getattr(want, name)[: _n(name, n)]
        if not np.array_equal(a.view(np.int32), b.view(np.int32)):
            bad = np.nonzero(a.view(np.int32) != b.view(np.int32))[0]
            raise AssertionError("%s differs at %d entries, first %d: %r vs %r (max abs diff %g)" % (
                name, len(bad), bad[0], a[bad[0]], b[bad[0]], float(np.max(np.abs(a - b)))))


def test_device_visible():
    from psim import abi
    assert abi.load_library().dem_device_count() >= 1


@pytest.mark.parametrize("steps,restitution", [(1, 0.5), (7, 0.5), (50, 0.5), (50, 0.0)])
def test_granular_box_bit_exact(steps, restitution):
    """4,096-disc box settle: positions, velocities, accelerations identical to the
    oracle (canonical contact order) bit for bit after `steps` steps, with the contact
    dashpot of restitution 0.5 and with the reference's undamped law."""
    s, st = scenes.granular_box(seed=0, restitution=restitution)
    prog = ComputeProg(s, st.copy(), reorder_interval=4)
    eng = orc.OracleEngine(st, s, mode=orc.MODE_CANONICAL)
    prog.step(steps)
    eng.step(steps)
    got = prog.update_state()
    assert_bits_equal(got, eng.st, st.p_count)
    assert np.array_equal(got.data[: 7 * st.p_count].view(np.int32), eng.st.data[: 7 * st.p_count].view(np.int32))
    c = prog.counters()
    assert c["contact_overflow"] == 0 and c["scan_timeouts"] == 0
    prog.close()


@pytest.mark.parametrize("zero", [(17, 300), (0, 17, 300)])
def test_zero_radius_discs_and_wall_bounces_bit_exact(zero):
    """Discs of radius 0 are inactive (2D_LOM.wgsl:83, 2D_Simulation.wgsl:118) and keep
    their state; with disc 0 among them the reference's invocation 0 never updates
    coll_cont, so no narrow phase runs at all.  Fast random velocities drive the outer
    discs into the walls: a bounce halves and reverses the velocity after the
    displacement was formed, so those drifts keep their displacement as an explicit
    record.  Three dem_step calls: state identical to the oracle."""
    s, st = scenes.granular_box(n_side=24, seed=11, spacing=1.9)
    rng = np.random.default_rng(11)
    n = st.p_count
    st.vel[: 2 * n] = rng.uniform(-200.0, 200.0, 2 * n).astype(np.float32)
    st.rot_vel[:n] = rng.uniform(-5.0, 5.0, n).astype(np.float32)
    for k in zero:
        st.radii[k] = 0.0
    prog = ComputeProg(s, st.copy(), reorder_interval=4)
    eng = orc.OracleEngine(st, s, mode=orc.MODE_CANONICAL)
    for k in (3, 10, 25):
        prog.step(k)
        eng.step(k)
    got = prog.update_state()
    assert_bits_equal(got, eng.st, n)
    for k in zero:   # inactive discs never move
        assert got.pos[2 * k] == st.pos[2 * k] and got.pos[2 * k + 1] == st.pos[2 * k + 1]
    prog.close()


@pytest.mark.parametrize("block", [64, 128])
def test_smaller_blocks_bit_exact(block):
    """The autotuner may pick 64- or 128-slot blocks: scene padding, the step kernel's
    block argument, re-sort units of 64 blocks and the rebuild kernels all follow the
    block size.  A compressed box with re-sorts every 3 steps, identical to the oracle
    (state and contact table) over two calls."""
    s, st = scenes.granular_box(n_side=40, seed=7, spacing=1.75)
    n = st.p_count
    prog = ComputeProg(s, st.copy(), reorder_interval=3, block_size=block)
    eng = orc.OracleEngine(st, s, mode=orc.MODE_CANONICAL)
    for k in (9, 21):
        prog.step(k)
        eng.step(k)
    got = prog.update_state()
    assert_bits_equal(got, eng.st, n)
    want = eng.canonical_contacts().view(np.int32).reshape(-1, 14, 6)[:n]
    assert np.array_equal(got.contacts.view(np.int32).reshape(-1, 14, 6)[:n], want)
    assert prog.counters()["reorders"] >= 5
    prog.close()


def test_granular_contacts_dense_bit_exact():
    """A compressed box (spacing 1.8 < 2 r_max) so that most discs touch several
    neighbours; 30 steps bit-exact including the contact table."""
    s, st = scenes.granular_box(n_side=32, seed=5, spacing=1.7)
    prog = ComputeProg(s, st.copy(), reorder_interval=3)
    eng = orc.OracleEngine(st, s, mode=orc.MODE_CANONICAL, slots=16)
    prog.step(30)
    eng.step(30)
    got = prog.update_state()
    n = st.p_count
    assert_bits_equal(got, eng.st, n)
    # contact table: canonical layout (occupied slots ascending by partner, -1 words after)
    want = eng.canonical_contacts().view(np.int32).reshape(-1, 16, 6)[:n, :14]
    have = got.contacts.view(np.int32).reshape(-1, 14, 6)[:n]
    assert np.array_equal(have, want)
    assert prog.counters()["contacts"] > n // 2  # really in contact
    prog.close()


def test_reads_between_calls_keep_contact_history():
    """Contacts of the list path are kept beside the neighbour list on the device (a mask
    and one tangential force per candidate); every read from the host writes them back
    as partner-id lists, which the next step merges into the list form again.  Ten calls
    of five steps with the state and the contact table read after each equal one call of
    fifty steps and the oracle bit for bit, contact table included (the compressed box of
    the test above, a list rebuild every few steps and a re-sort every 7)."""
    s, st = scenes.granular_box(n_side=32, seed=5, spacing=1.7)
    n = st.p_count
    a = ComputeProg(s, st.copy(), reorder_interval=7)
    b = ComputeProg(s, st.copy(), reorder_interval=7)
    eng = orc.OracleEngine(st, s, mode=orc.MODE_CANONICAL, slots=16)
    for call in range(10):
        a.step(5)
        eng.step(5)
        got = a.update_state()
        assert_bits_equal(got, eng.st, n)
        want = eng.canonical_contacts().view(np.int32).reshape(-1, 16, 6)[:n, :14]
        assert np.array_equal(got.contacts.view(np.int32).reshape(-1, 14, 6)[:n], want), call
        c, p = a.debug_contacts(16)
        assert int(c[:n].sum()) == int((want[:, :, 1] >= 0).sum()), call
    b.step(50)
    one = b.update_state()
    assert_bits_equal(one, got, n)
    assert np.array_equal(one.contacts.view(np.int32), got.contacts.view(np.int32))
    assert a.counters()["contacts"] > n // 2
    a.close()
    b.close()


@pytest.mark.parametrize("reorder", [11, -1])
def test_lists_around_the_slot_limit_keep_contact_history(reorder):
    """Cells 1.9 times the default: about two-thirds of the neighbour lists fit the 14
    slots (the list path, contacts kept beside the list), the rest take the general path
    (15-16 candidates) or the stencil fallback (more); with discs thrown at ~150 units/s
    the lists are rebuilt every few dozen steps, by re-sorts (every 11 steps) or by the
    displacement check alone (no re-sorts), and discs cross between the paths at the
    rebuilds (contacts moved beside the new list, or turned back into partner-id lists
    and merged again).  Bit-exact with the oracle after each of three 50-step calls,
    contact table included."""
    s, st = scenes.granular_box(n_side=32, seed=5, spacing=1.7)
    n = st.p_count
    st.vel[: 2 * n] = np.random.default_rng(5).normal(0.0, 150.0, 2 * n).astype(np.float32)
    g = ComputeProg.default_grid(s, st)
    grid = ComputeProg.default_grid(s, st, cell_size=1.9 * g.cell_size)
    prog = ComputeProg(s, st.copy(), grid=grid, reorder_interval=reorder)
    eng = orc.OracleEngine(st, s, mode=orc.MODE_CANONICAL, slots=16)
    for call in range(3):
        prog.step(50)
        eng.step(50)
        got = prog.update_state()
        assert_bits_equal(got, eng.st, n)
        want = eng.canonical_contacts().view(np.int32).reshape(-1, 16, 6)[:n, :14]
        assert np.array_equal(got.contacts.view(np.int32).reshape(-1, 14, 6)[:n], want), call
    c = prog.counters()
    assert c["list_overflow"] > 0 and c["broad_phases"] >= 3 and c["contacts"] > 0
    assert c["contact_overflow"] == 0
    if reorder < 0:
        assert c["reorders"] == 0
    prog.close()


def test_stage_views_match_closed_forms():
    """Stages 1-4: cell keys, (cell, id) order, cell starts and contact pairs equal
    the oracle's closed forms / brute force."""
    import ctypes as C
    s, st = scenes.granular_box(n_side=32, seed=3, spacing=1.8)
    prog = ComputeProg(s, st.copy())
    prog.step(3)
    got = prog.update_state()
    n = st.p_count
    keys, order, start = prog.debug_broad_phase()
    g = prog.grid
    L = orc.lib()
    k_want = np.zeros(n, dtype=np.int32)
    L.orc_cell_keys(n, n, got.pos.ctypes.data_as(orc.FP), got.radii.ctypes.data_as(orc.FP), C.byref(g),
                    k_want.ctypes.data_as(orc.IP))
    assert np.array_equal(keys, k_want)
    o_want = np.zeros(n, dtype=np.int32)
    s_want = np.zeros(g.w * g.h + 1, dtype=np.int32)
    assert L.orc_sort_cells(n, k_want.ctypes.data_as(orc.IP), g.w * g.h, o_want.ctypes.data_as(orc.IP),
                            s_want.ctypes.data_as(orc.IP)) == 0
    assert np.array_equal(order, o_want)
    assert np.array_equal(start, s_want)
    # contact pairs of the last narrow phase (positions of the last LOM = current)
    counts, partners = prog.debug_contacts(16)
    c_want = np.zeros(n, dtype=np.int32)
    p_want = np.zeros(n * 16, dtype=np.int32)
    tot = L.orc_contact_pairs(n, got.pos.ctypes.data_as(orc.FP), got.radii.ctypes.data_as(orc.FP), 16,
                              c_want.ctypes.data_as(orc.IP), p_want.ctypes.data_as(orc.IP))
    assert tot > 0
    assert np.array_equal(counts, c_want)
    assert np.array_equal(partners, p_want.reshape(n, 16))
    prog.close()


def _replay(engine_factory, scene):
    s = Settings()
    st = State.load_file(os.path.join(GOLDEN, "saved_states", scene + ".bin"), s)
    eng = engine_factory(st, s)
    rep = ScriptReplay(eng, st.scripts)
    return {p.replace("\\", "/").split("/")[-1][:-4]: rows for p, rows in rep.run()}


@pytest.mark.parametrize("scene", ["TEXPS", "PBEXPS", "Chain_TEXP", "Chain_EXP", "CHTEXPDET"])
def test_experiment_replay_matches_oracle_and_traces(scene):
    """The embedded experiment scripts replayed on the GPU (f64 distances, as the
    fixtures were recorded): every exported row equals the oracle's bit for bit, and
    the position/velocity columns follow the reference's CSV."""
    gpu = _replay(lambda st, s: ComputeProg(s, st, record_data=True), scene)
    cpu = _replay(lambda st, s: orc.OracleEngine(st, s, mode=orc.MODE_CANONICAL), scene)
    traces = np.load(os.path.join(GOLDEN, "traces.npz"))
    assert gpu.keys() == cpu.keys()
    for stem in gpu:
        assert np.array_equal(gpu[stem], cpu[stem]), stem
        if stem in traces.files:
            ref = traces[stem]
            n = min(len(ref), len(gpu[stem]))
            for c in range(1, 7):
                scale = max(np.abs(ref[:n, c]).max(), 1e-30)
                assert np.abs(ref[:n, c] - gpu[stem][:n, c]).max() / scale < 1e-3, (stem, c)


def _run_both(s, st, calls, **kw):
    prog = ComputeProg(s, st.copy(), **kw)
    eng = orc.OracleEngine(st, s, mode=orc.MODE_CANONICAL)
    for k in calls:
        prog.step(k)
        eng.step(k)
    got = prog.update_state()
    c = prog.counters()
    prog.close()
    return got, eng, c


def test_neighbour_list_overflow_takes_stencil_fallback():
    """One r = 4 disc in a dense bed of r ~ U[0.3, 0.45] discs: the big disc has more
    neighbour-list candidates than a list holds and is searched over the cell list of
    the last rebuild instead; the run stays bit-exact with the oracle."""
    s, st = scenes.granular_box(n_side=24, seed=11, spacing=1.0, rmin=0.3, rmax=0.45)
    n = st.p_count
    p = st.pos[: 2 * n].reshape(-1, 2)
    d = np.hypot(p[:, 0], p[:, 1])
    big = int(np.argmin(d))
    keep = (d > 4.0 + 0.5) | (np.arange(n) == big)
    st2 = scenes._new_state(int(keep.sum()))
    idx = np.nonzero(keep)[0]
    st2.pos[: 2 * len(idx)] = p[idx].ravel()
    st2.radii[: len(idx)] = st.radii[idx]
    j = int(np.nonzero(idx == big)[0][0])
    st2.pos[2 * j: 2 * j + 2] = 0.0
    st2.radii[j] = 4.0
    s.setup.particles = st2.p_count
    got, eng, c = _run_both(s, st2, (60, 120, 120), reorder_interval=16)
    assert_bits_equal(got, eng.st, st2.p_count)
    assert c["list_overflow"] > 0 and c["broad_phases"] >= 1 and c["contact_overflow"] == 0


@pytest.mark.parametrize("reorder", [8, -1])
def test_fast_motion_rebuilds_lists_often(reorder):
    """Discs at ~30 units/s (dt 1e-4): the lists are rebuilt every few dozen steps
    as particles cross 0.45 skin; every step stays bit-exact with the oracle.  With
    re-sorts every 8 steps the rebuilds alternate between those that follow a re-sort
    (the lists built over the moved storage) and those raised by the displacement
    check alone; without re-sorts every rebuild is of the second kind."""
    s, st = scenes.granular_box(n_side=24, seed=12, spacing=1.9)
    n = st.p_count
    st.vel[: 2 * n] = np.random.default_rng(12).normal(0.0, 30.0, 2 * n).astype(np.float32)
    got, eng, c = _run_both(s, st, (100, 100, 100), reorder_interval=reorder)
    assert_bits_equal(got, eng.st, n)
    assert np.array_equal(got.data[: 7 * n].view(np.int32), eng.st.data[: 7 * n].view(np.int32))
    assert c["broad_phases"] >= 5
    if reorder < 0:
        assert c["reorders"] == 0


@pytest.mark.parametrize("scene", ["box", "chains"])
def test_graph_replays_equal_direct_launches(scene):
    """dem_step through HIP graphs (graph = 1) against direct launches, both after every
    one of ten dem_step(8) calls: bit-identical states.  Re-sorts every 16 steps put
    different launch patterns into different calls, so graphs are captured and
    replayed; the host reads the state back between replays."""
    if scene == "box":
        s, st = scenes.granular_box(n_side=32, seed=21, spacing=1.9)
    else:
        s, st, strengths = scenes.bonded_chains(chains=8, per_chain=32, seed=5, striker_r=1.0, striker_v=40.0,
                                                striker_density=50.0, strength_scale=50.0)
        scenes.chain_bonds(st, 8, 32, strengths)
    n = st.p_count
    a = ComputeProg(s, st.copy(), reorder_interval=16, graph=True)
    b = ComputeProg(s, st.copy(), reorder_interval=16, graph=False)
    for call in range(10):
        a.step(8)
        b.step(8)
        ga, gb = a.update_state(), b.update_state()
        assert_bits_equal(ga, gb, n)
        assert np.array_equal(ga.contacts.view(np.int32), gb.contacts.view(np.int32)), call
    ca = a.counters()
    # one graph per distinct launch pattern (buffer parities, re-sort position), replayed
    # when a pattern recurs
    assert ca["graph_captures"] >= 2 and ca["graph_replays"] >= 1
    assert ca["graph_captures"] + ca["graph_replays"] == 10
    assert b.counters()["graph_captures"] == 0
    a.close()
    b.close()


@pytest.mark.parametrize("preload", [False, True])
def test_contact_slot_cap_matches_oracle(preload):
    """A disc of radius 2 (a large body among r = 0.3 discs) touching 20 discs holds 14
    contacts, as the reference's table does (2D_Simulation.wgsl:125): slots already held
    are kept (preloaded with the 14 largest ids) and new partners take the free ones in
    ascending id.  State and contact table equal the oracle after 1 and 25 more steps;
    the dropped pairs are counted."""
    from test_oracle import _crowded
    s, st = _crowded(slots_preload=list(range(7, 21)) if preload else ())
    n = st.p_count
    prog = ComputeProg(s, st.copy(), reorder_interval=4)
    eng = orc.OracleEngine(st, s, mode=orc.MODE_CANONICAL)
    for k in (1, 25):
        prog.step(k)
        eng.step(k)
        got = prog.update_state()
        assert_bits_equal(got, eng.st, n)
        want = eng.canonical_contacts().view(np.int32).reshape(-1, 14, 6)[:n]
        assert np.array_equal(got.contacts.view(np.int32).reshape(-1, 14, 6)[:n], want)
    assert prog.counters()["contact_overflow"] > 0
    prog.close()


def test_wide_cells_general_path_bit_exact():
    """Cells three times wider than needed: each neighbour list holds more candidates than
    the slots allow (or overflows into the stencil fallback), so every disc takes the
    general path.  A compressed box, identical to the oracle after 20 steps."""
    s, st = scenes.granular_box(n_side=24, seed=13, spacing=1.75)
    n = st.p_count
    g = ComputeProg.default_grid(s, st)
    grid = ComputeProg.default_grid(s, st, cell_size=3.0 * g.cell_size)
    prog = ComputeProg(s, st.copy(), grid=grid, reorder_interval=5)
    eng = orc.OracleEngine(st, s, mode=orc.MODE_CANONICAL)
    prog.step(20)
    eng.step(20)
    got = prog.update_state()
    assert_bits_equal(got, eng.st, n)
    want = eng.canonical_contacts().view(np.int32).reshape(-1, 14, 6)[:n]
    assert np.array_equal(got.contacts.view(np.int32).reshape(-1, 14, 6)[:n], want)
    assert prog.counters()["list_overflow"] > 0
    prog.close()


@pytest.mark.parametrize("mouse", [False, True])
def test_planet_gravity_round_walls_bit_exact(mouse):
    """Planet-mode gravity (the reference's default, settings.rs:339) toward the origin
    or the mouse point, inside a circular wall (round_bounds): 900 discs fall together
    and collide; state and contact table identical to the oracle after 60 and 300 steps."""
    s, st = scenes.granular_box(n_side=30, seed=17, spacing=2.0)
    s.physics.planet_mode = True
    s.physics.mouse_gravity = mouse
    s.physics.gravity_acceleration = 20.0
    s.world_pos = (7.0, -5.0) if mouse else (0.0, 0.0)
    s.simulation.round_walls = True
    s.simulation.wall_radius = 43.0
    s.simulation.timestep = 5e-4
    n = st.p_count
    prog = ComputeProg(s, st.copy(), reorder_interval=16)
    eng = orc.OracleEngine(st, s, mode=orc.MODE_CANONICAL)
    for k in (60, 240):
        prog.step(k)
        eng.step(k)
        got = prog.update_state()
        assert_bits_equal(got, eng.st, n)
        want = eng.canonical_contacts().view(np.int32).reshape(-1, 14, 6)[:n]
        assert np.array_equal(got.contacts.view(np.int32).reshape(-1, 14, 6)[:n], want)
    r = np.hypot(got.pos[0:2 * n:2], got.pos[1:2 * n:2]) + got.radii[:n]
    assert r.max() <= 43.0 + 1e-3        # held by the circular wall
    assert prog.counters()["contacts"] > n // 8   # discs in contact
    prog.close()


def test_external_forces_and_collision_interval_bit_exact():
    """External forces and moments on a third of the discs (forces[id].{x, y, rot}) and a
    narrow phase every third step (collision_interval 3: contacts carry over in between);
    state, data words and contact table identical to the oracle over three calls."""
    s, st = scenes.granular_box(n_side=24, seed=19, spacing=1.8)
    n = st.p_count
    s.physics.collision_interval = 3
    rng = np.random.default_rng(19)
    f = st.forces.reshape(-1, 6)
    sel = rng.random(n) < 1.0 / 3.0
    f[:n][sel, 0] = rng.normal(0, 50.0, sel.sum())
    f[:n][sel, 1] = rng.normal(0, 50.0, sel.sum())
    f[:n][sel, 2] = rng.normal(0, 5.0, sel.sum())
    prog = ComputeProg(s, st.copy(), reorder_interval=8)
    eng = orc.OracleEngine(st, s, mode=orc.MODE_CANONICAL)
    for k in (7, 11, 40):
        prog.step(k)
        eng.step(k)
        got = prog.update_state()
        assert_bits_equal(got, eng.st, n)
        assert np.array_equal(got.data[: 7 * n].view(np.int32), eng.st.data[: 7 * n].view(np.int32))
        want = eng.canonical_contacts().view(np.int32).reshape(-1, 14, 6)[:n]
        assert np.array_equal(got.contacts.view(np.int32).reshape(-1, 14, 6)[:n], want)
    prog.close()
