"""Serving loop (continuous batching) on the GPU: every scene served in a mix of
other scenes - admitted while others are mid-run, its steps split over ticks, its
neighbours in the launches changing - reproduces its solo run and the oracle's run
(the reference algorithm on the CPU) bit for bit.  Needs an MI355X."""
import os

import numpy as np
import pytest

import orc
from psim import abi, scenes
from psim.compute import ComputeProg
from psim.serve import Server, footprint
from psim.settings import Settings
from psim.state import State

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
FIELDS = ("pos", "vel", "acc", "rot", "rot_vel", "rot_acc", "data", "contacts", "bonds")


def _len(st, name):
    n = st.p_count
    return {"pos": 2 * n, "vel": 2 * n, "acc": 2 * n, "data": 7 * n, "contacts": st.max_contacts * 6 * n,
            "bonds": 3 * st.n_bonds}.get(name, n)


def assert_same(got: State, want: State, tag):
    for name in FIELDS:
        k = _len(want, name)
        a = np.asarray(getattr(got, name))[:k].view(np.int32)
        b = np.asarray(getattr(want, name))[:k].view(np.int32)
        if not np.array_equal(a, b):
            bad = np.nonzero(a != b)[0]
            raise AssertionError("%s: %s differs at %d of %d words (first %d)" % (tag, name, len(bad), k, bad[0]))


def mixed_scenes():
    """Granular boxes of several sizes, frictions and packings (f32), the reference's
    parallel-bond experiment state (f64 distances, bond mode 3) and bonded chains struck
    by an r = 8 disc (a large body), through the impact."""
    out = []
    for k, (n_side, mu, spacing) in enumerate([(16, 0.5, 1.8), (24, 0.2, 1.9), (8, 0.8, 1.7), (32, 0.5, 2.0),
                                               (12, 0.3, 1.75)]):
        s, st = scenes.granular_box(n_side=n_side, seed=20 + k, spacing=spacing, mu=mu)
        out.append(("box%d" % k, s, st, 30 + 11 * k))
    s = Settings()
    st = State.load_file(os.path.join(GOLDEN, "saved_states", "PBEXPS.bin"), s)
    out.append(("PBEXPS", s, st, 45))
    # bonded chains struck by config 1's r = 8 disc (a large body on its scene's list)
    s, st, strengths = scenes.bonded_chains(chains=4, per_chain=24, seed=31, striker_r=8.0, striker_v=40.0,
                                            striker_density=10.0, strength_scale=40.0)
    scenes.chain_bonds(st, 4, 24, strengths)
    out.append(("striker", s, st, 300))
    return out


def oracle(s, st, n):
    """The scene advanced n steps by the oracle (canonical slot order), as a State with
    the oracle's contact table."""
    eng = orc.OracleEngine(st.copy(), s, mode=orc.MODE_CANONICAL)
    eng.step(n)
    out = eng.st.copy()
    out.contacts = eng.canonical_contacts().copy()
    return out


def solo(s, st, n):
    prog = ComputeProg(s, st.copy(), record_data=True)
    prog.step(n)
    got = prog.update_state().copy()
    prog.close()
    return got


@pytest.mark.parametrize("graph", [False, True])
def test_batched_equals_solo_with_continuous_arrivals(graph):
    """Scenes arriving while others run reproduce their solo runs and the oracle's runs;
    with `graph` the ticks' launches run as HIP graphs (captured once per launch
    pattern, replayed after)."""
    sc = mixed_scenes()
    want = [solo(s, st, n) for _, s, st, n in sc]
    ref = [oracle(s, st, n) for _, s, st, n in sc]
    srv = Server(slot_capacity=8192, cell_capacity=1 << 16, bond_capacity=256, scene_capacity=16, substeps=7,
                 graph=graph)
    tickets, outs = [], []
    # two scenes start; the others arrive one per tick while those run
    for i in range(len(sc)):
        name, s, st, n = sc[i]
        outs.append(st.copy())
        tickets.append(srv.submit(s, outs[i], n))
        if i >= 1:
            srv.tick()
    for _ in range(1000):
        if all(srv.status(t) == abi.DEM_REQ_DONE for t in tickets):
            break
        srv.tick()
    st = srv.stats()
    assert st["completed"] == len(sc) and st["queued"] == 0 and st["resident"] == 0
    assert st["p99_s"] >= st["p50_s"] > 0.0
    for i, t in enumerate(tickets):
        got, lat = srv.take(t, outs[i])
        assert lat > 0.0
        assert_same(got, want[i], sc[i][0])
        assert_same(got, ref[i], sc[i][0] + " vs oracle")
    srv.close()


def test_stream_scenes_match_oracle():
    """Eight scenes of the serving configuration's distribution (scenes.serving_stream:
    granular boxes and r = 8-struck bonded chains with their own friction, restitution
    and bond strengths), at reduced sizes (n log-uniform in [1k, 4k], 150-400 steps)
    so that the oracle finishes in seconds, all submitted at once to pools that hold
    about half of them, served with HIP graphs: each result equals the oracle's."""
    stream = scenes.serving_stream(n_scenes=8, seed=11, n_max=4096, steps_min=150, steps_max=400)
    slots = sum((st.p_count + 255) // 256 * 256 for _, _, st, _ in stream)
    cells = sum(ComputeProg.default_grid(s, st).w * ComputeProg.default_grid(s, st).h + 1 for _, s, st, _ in stream)
    bonds = sum(st.n_bonds for _, _, st, _ in stream)
    srv = Server(slot_capacity=slots // 2 + 4096, cell_capacity=cells, bond_capacity=max(bonds, 1),
                 scene_capacity=8, substeps=25, graph=True)
    outs = [st.copy() for _, _, st, _ in stream]
    tickets = [srv.submit(s, outs[i], n) for i, (_, s, st, n) in enumerate(stream)]
    for _ in range(2000):
        if all(srv.status(t) == abi.DEM_REQ_DONE for t in tickets):
            break
        srv.tick()
    kinds = []
    for i, t in enumerate(tickets):
        _, s, st, n = stream[i]
        got, _ = srv.take(t, outs[i])
        assert_same(got, oracle(s, st, n), "stream scene %d (%d discs, %d steps)" % (i, st.p_count, n))
        kinds.append((st.p_count, st.n_bonds, float(s.physics.restitution)))
    srv.close()
    print("served and matched:", kinds, flush=True)
    assert sum(1 for k in kinds if k[1] > 0) == 4


def test_admission_by_footprint():
    """Pools that hold two of the scenes at a time: the others wait in the queue,
    are admitted as regions free up, and still reproduce their solo runs."""
    sc = []
    for k in range(5):
        s, st = scenes.granular_box(n_side=32, seed=40 + k, spacing=1.8)
        sc.append((s, st, 25 + 5 * k))
    g = ComputeProg.default_grid(sc[0][0], sc[0][1])
    slots, cells, _ = footprint(sc[0][1], g)
    srv = Server(slot_capacity=2 * slots, cell_capacity=2 * cells + 8, scene_capacity=8, substeps=10)
    outs = [st.copy() for _, st, _ in sc]
    tickets = [srv.submit(s, outs[i], n) for i, (s, st, n) in enumerate(sc)]
    max_res = 0
    for _ in range(1000):
        if all(srv.status(t) == abi.DEM_REQ_DONE for t in tickets):
            break
        srv.tick()
        max_res = max(max_res, sum(srv.status(t) == abi.DEM_REQ_RESIDENT for t in tickets))
        max_res = max(max_res, int(srv.stats()["resident"]))
    assert max_res == 2
    for i, t in enumerate(tickets):
        got, _ = srv.take(t, outs[i])
        assert_same(got, solo(*sc[i]), "scene %d" % i)
    srv.close()


def test_oversized_scene_refused():
    s, st = scenes.granular_box(n_side=20, seed=1)     # 400 discs: two 256-slot blocks
    srv = Server(slot_capacity=256, cell_capacity=1 << 12, scene_capacity=2, substeps=4)
    with pytest.raises(abi.DemError) as e:
        srv.submit(s, st, 10)
    assert e.value.code == abi.DEM_ECAPACITY
    srv.close()


def test_failed_admission_does_not_block_the_queue():
    """A request the engine refuses at admission (particle 0 lists more contacts than a
    particle can hold) sits between valid ones: it becomes DEM_REQ_FAILED, its take
    returns the error, and the scenes around it are served and equal their solo runs."""
    good = [scenes.granular_box(n_side=16, seed=50 + k, spacing=1.8) + (20 + 5 * k,) for k in range(2)]
    s_bad, st0 = scenes.granular_box(n_side=16, seed=52, spacing=1.8)
    st_bad = State(st0.capacity, 20)
    for name in ("pos", "vel", "acc", "rot", "rot_vel", "rot_acc", "forces", "radii", "fixity", "material_pointers"):
        setattr(st_bad, name, getattr(st0, name).copy())
    st_bad.p_count = st0.p_count
    c = st_bad.contacts.view(np.int32).reshape(-1, 20, 6)
    c[0, :17, 0] = 0
    c[0, :17, 1] = np.arange(1, 18)       # 17 plain contacts on particle 0
    c[0, :17, 2] = 0
    srv = Server(slot_capacity=4096, cell_capacity=1 << 14, scene_capacity=4, substeps=10, bypass=0)
    outs = [good[0][1].copy(), st_bad.copy(), good[1][1].copy()]
    t0 = srv.submit(good[0][0], outs[0], good[0][2])
    tb = srv.submit(s_bad, outs[1], 10)
    t1 = srv.submit(good[1][0], outs[2], good[1][2])
    for _ in range(100):
        if srv.status(t0) == abi.DEM_REQ_DONE and srv.status(t1) == abi.DEM_REQ_DONE:
            break
        srv.tick()
    assert srv.status(tb) == abi.DEM_REQ_FAILED
    assert srv.stats()["failed"] == 1
    with pytest.raises(abi.DemError) as e:
        srv.take(tb, outs[1])
    assert e.value.code == abi.DEM_ECAPACITY and "admission" in str(e.value)
    for t, out, (s, st, n) in ((t0, outs[0], good[0]), (t1, outs[2], good[1])):
        got, _ = srv.take(t, out)
        assert_same(got, solo(s, st, n), "scene around the failed one")
    srv.close()


def test_submit_refuses_what_admission_never_could():
    """A grid finer than 2 r_max and more materials than the pool holds are refused at
    submit (nothing queued), instead of blocking the queue."""
    s, st = scenes.granular_box(n_side=8, seed=3)
    srv = Server(slot_capacity=1024, cell_capacity=1 << 14, scene_capacity=1, substeps=4)
    g = ComputeProg.default_grid(s, st)
    g.cell_size = 1.0           # radii up to 1.0: cells must be wider than 2.0
    with pytest.raises(abi.DemError) as e:
        srv.submit(s, st, 5, grid=g)
    assert e.value.code == abi.DEM_EINVAL
    s.materials = s.materials * 9   # 9 materials; the pool holds 8 per scene
    with pytest.raises(abi.DemError) as e:
        srv.submit(s, st, 5)
    assert e.value.code == abi.DEM_ECAPACITY
    assert srv.stats()["queued"] == 0
    srv.close()
