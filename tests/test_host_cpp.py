"""The C++ host mirror (physics-sim_amd/host: State, Settings, ComputeProg over the C
ABI) against the Python mirror: the same saved-state bytes, the same uniform words,
grids and bond slots; on a GPU the same simulated bytes after the same dem_step calls."""
import glob
import os
import sys

import numpy as np
import pytest

from psim import abi, scenes
from psim.compute import ComputeProg, assemble_bonds
from psim.settings import Settings
from psim.state import State

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "physics-sim_amd", "host"))
GOLDEN = os.path.join(ROOT, "tests", "golden", "saved_states")
FIXTURES = sorted(glob.glob(os.path.join(GOLDEN, "*.bin")))

ph = pytest.importorskip("_psim_host", reason="C++ host mirror not built (run __graft_entry__.build())")


def _bytes(path):
    return open(path, "rb").read()


@pytest.mark.parametrize("path", FIXTURES)
def test_codec_roundtrip_bytes(path):
    b = _bytes(path)
    assert ph.codec_roundtrip(b) == b


@pytest.mark.parametrize("path", FIXTURES)
def test_state_load_save_equals_python(path):
    b = _bytes(path)
    s = Settings()
    st = State.load(b, s)
    assert ph.state_roundtrip(b) == st.save(s)


@pytest.mark.parametrize("path", FIXTURES)
def test_uniform_and_grids_equal_python(path):
    b = _bytes(path)
    s = Settings()
    st = State.load(b, s)
    assert ph.collision_settings(b) == bytes(s.collision_settings())
    (g_ref, g_def) = ph.grids(b)
    g = s.grid_info()
    d = ComputeProg.default_grid(s, st)
    assert g_ref == (g.cell_size, g.cell_cap, g.w, g.h)
    assert g_def == (d.cell_size, d.cell_cap, d.w, d.h)


def test_assemble_bonds_equals_python():
    s, st = scenes.granular_box(n_side=24, seed=5, spacing=2.0, rmin=0.8, rmax=1.2)
    import orc
    n = st.p_count
    bonds_o = np.zeros(3 * 6 * n, dtype=np.int32)
    cont_o = np.zeros(14 * 6 * n, dtype=np.float32)
    nb = orc.lib().orc_regen_bonds(n, st.pos.ctypes.data_as(orc.FP), st.radii.ctypes.data_as(orc.FP), 6, 14,
                                   bonds_o.ctypes.data_as(orc.IP), 6 * n, cont_o.ctypes.data_as(orc.FP), 1.05, 0.0)
    c = cont_o.view(np.int32).reshape(n, 14, 6)
    pairs = sorted((c[i, k, 5], i, c[i, k, 1]) for i in range(n) for k in range(14) if c[i, k, 0] == i and c[i, k, 1] > i)
    pairs = np.array([(i, j) for _, i, j in pairs], dtype=np.int32)
    assert len(pairs) == nb > n
    bonds_py, cont_py = assemble_bonds(st, pairs, 14)
    b_cpp, c_cpp = ph.assemble_bonds(st.save(s), pairs.tobytes(), 14)
    assert b_cpp == bonds_py.tobytes()
    assert c_cpp == cont_py.tobytes()


def test_compute_prog_fails_loudly_without_device():
    if abi.load_library().dem_device_count() > 0:
        pytest.skip("a device is visible")
    with pytest.raises(ph.DemError, match="device"):
        ph.run(_bytes(os.path.join(GOLDEN, "PBEXPS.bin")), [1], False)


def _python_run(b, calls, regen):
    s = Settings()
    st = State.load(b, s)
    prog = ComputeProg(s, st, bond_capacity=max(st.p_count * s.setup.max_bonds, 1) if regen else None)
    if regen:
        prog.regen_bonds(s)
    for k in calls:
        prog.step(k)
    got = prog.update_state()
    out = (got.save(s), prog.kinetic_energy(), prog.broken_bond_count())
    prog.close()
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("scene,calls,regen", [("TEXPS", [40, 60], False), ("Chain_TEXP", [25, 25, 50], False),
                                               ("box", [30, 30], False), ("box", [20, 40], True)])
def test_cpp_client_equals_python_client(scene, calls, regen):
    """A C++ client (State::load, ComputeProg::new/step/update_state, State::save) and
    the Python one write the same bytes after the same dem_step calls."""
    if scene == "box":
        s, st = scenes.granular_box(n_side=20, seed=7, spacing=1.9, rmin=0.8, rmax=1.0)
        b = st.save(s)
    else:
        b = _bytes(os.path.join(GOLDEN, scene + ".bin"))
    want = _python_run(b, calls, regen)
    got = ph.run(b, calls, regen)
    assert got[0] == want[0]
    assert got[1] == want[1] and got[2] == want[2]


def test_grids_set_large_bodies_aside_like_python():
    """A striker of r = 8 among r = 0.5 chain discs: both clients size the cells for the
    chain discs (the striker pairs through the large-body list)."""
    s, st, strengths = scenes.bonded_chains(chains=4, per_chain=16, seed=2, striker_r=8.0)
    b = st.save(s)
    st2 = State.load(b, Settings())
    (g_ref, g_def) = ph.grids(b)
    d = ComputeProg.default_grid(s, st2)
    assert g_def == (d.cell_size, d.cell_cap, d.w, d.h)
    assert d.cell_size < 1.5          # 2 x 0.5 + margin + skin, not 2 x 8
