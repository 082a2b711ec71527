"""Edge cases of the step on the GPU: an empty scene, scenes whose sizes are not a
multiple of the block (the pads are inactive slots) from one disc up, and inputs the
library refuses with a status code instead of running.  Needs an MI355X."""
import numpy as np
import pytest

import orc
from psim import abi, scenes
from psim.compute import ComputeProg

pytestmark = pytest.mark.gpu

FIELDS = ("pos", "vel", "acc", "rot", "rot_vel", "rot_acc", "data")


def _len(name, n):
    return {"pos": 2 * n, "vel": 2 * n, "acc": 2 * n, "data": 7 * n}.get(name, n)


def _scene(n, seed):
    """n discs (r ~ U[0.5, 1.0]) on sites 1.4 apart in a 16 x 16 walled box, most of them
    overlapping at the start, so that contacts, wall bounces and gravity all act within
    200 steps."""
    s, st = scenes.granular_box(n_side=4, seed=seed, spacing=2.0)
    s.simulation.hor_bound = np.float32(16.0)
    s.simulation.vert_bound = np.float32(16.0)
    s.setup.particles = n
    st2 = scenes._new_state(n)
    rng = np.random.default_rng(seed)
    k = 0
    for y in np.arange(-6.6, 6.5, 1.4):          # 10 x 10 sites
        for x in np.arange(-6.6, 6.5, 1.4):
            if k == n:
                break
            st2.pos[2 * k] = np.float32(x + rng.uniform(-0.1, 0.1))
            st2.pos[2 * k + 1] = np.float32(y + rng.uniform(-0.1, 0.1))
            st2.radii[k] = np.float32(rng.uniform(0.5, 1.0))
            k += 1
    assert k == n
    return s, st2


def test_empty_scene_steps_and_reads_back():
    """No particles: uploads, steps, downloads, and reports zero energy and counters."""
    s, st = _scene(1, 1)
    st0 = scenes._new_state(0)
    s.setup.particles = 0
    prog = ComputeProg(s, st0, capacity=16)
    try:
        prog.step(10)
        got = prog.update_state()
        assert got.p_count == 0
        assert prog.kinetic_energy() == 0.0
        c = prog.counters()
        assert c["contacts"] == 0 and c["steps"] == 10
    finally:
        prog.close()


@pytest.mark.parametrize("n", [1, 2, 63, 65, 100])
def test_ragged_sizes_bit_exact(n):
    """Scenes of 1, 2, 63, 65 and 100 discs (the pad slots up to the block are inactive):
    200 steps of falling, bouncing and colliding equal the oracle bit for bit, the contact
    table included."""
    s, st = _scene(n, 10 + n)
    prog = ComputeProg(s, st.copy(), record_data=True, reorder_interval=16)
    eng = orc.OracleEngine(st, s, mode=orc.MODE_CANONICAL)
    for k in (1, 99, 100):
        prog.step(k)
        eng.step(k)
    got = prog.update_state()
    prog.close()
    for name in FIELDS:
        a = getattr(got, name)[: _len(name, n)].view(np.int32)
        b = getattr(eng.st, name)[: _len(name, n)].view(np.int32)
        assert np.array_equal(a, b), (n, name)
    want = eng.canonical_contacts().view(np.int32).reshape(-1, 14, 6)[:n]
    assert np.array_equal(got.contacts.view(np.int32).reshape(-1, 14, 6)[:n], want)


def test_refused_inputs():
    """A cell side below 2 r_max, more particles than the capacity (rounded up to the
    block size), a negative step count: each returns its status code (EINVAL,
    ECAPACITY, EINVAL) with a message, and the context stays usable."""
    s, st = _scene(65, 3)
    g = ComputeProg.default_grid(s, st)
    small = abi.DemGrid(np.float32(1.0), 0, g.w, g.h)   # discs reach r = 1: 2 r_max > 1
    with pytest.raises(abi.DemError) as e:
        ComputeProg(s, st.copy(), grid=small)
    assert e.value.code == abi.DEM_EINVAL and str(e.value)
    prog = ComputeProg(s, st.copy(), capacity=65)   # slots: the capacity rounded up to the block, 256
    try:
        big = scenes._new_state(257)
        rng = np.random.default_rng(4)
        big.pos[: 2 * 257] = rng.uniform(-7.0, 7.0, 2 * 257).astype(np.float32)
        big.radii[:257] = 0.5
        with pytest.raises(abi.DemError) as e:
            prog.restore(big)
        assert e.value.code == abi.DEM_ECAPACITY
        with pytest.raises(abi.DemError) as e:
            prog.step(-1)
        assert e.value.code == abi.DEM_EINVAL
        prog.restore(st)     # the context still works
        prog.step(5)
        assert prog.update_state().p_count == 65
    finally:
        prog.close()
