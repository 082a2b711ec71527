"""Host-side checks that need no GPU: the C ABI library exports, the saved-state
codec, the settings packing and the scene generators."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

from psim import abi, flatbuf, scenes
from psim.compute import assemble_bonds, min_cell_size
from psim.settings import Settings
from psim.state import State

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def header_functions():
    txt = open(os.path.join(ROOT, "include", "dem.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(dem_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_header_symbol():
    lib = abi.load_library()
    names = header_functions()
    assert len(names) >= 18
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(abi.EXPORTS)


def test_library_fails_loudly_without_device():
    lib = abi.load_library()
    if lib.dem_device_count() > 0:
        pytest.skip("a device is visible")
    s, st = scenes.granular_box(n_side=4)
    from psim.compute import ComputeProg
    with pytest.raises(abi.DemError):
        ComputeProg(s, st)
    # the raw entry point reports DEM_EDEVICE with a message, no abort
    p = abi.DemParams()
    p.u = s.collision_settings()
    mats = abi.materials_array(s.materials)
    p.materials = C.cast(mats, C.POINTER(abi.DemMaterial))
    p.n_materials = 1
    p.grid = s.box_grid()
    ctx = C.c_void_p()
    assert lib.dem_create(C.byref(p), 16, 0, C.byref(ctx)) == abi.DEM_EDEVICE
    assert b"device" in lib.dem_last_error()


def test_abi_struct_sizes():
    assert C.sizeof(abi.DemUniform) == 116
    assert C.sizeof(abi.DemMaterial) == 24
    assert C.sizeof(abi.DemGrid) == 16


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "saved_states", "*.bin"))))
def test_flatbuffer_roundtrip_bytes(path):
    """load -> save reproduces the reference writer's bytes exactly."""
    b = open(path, "rb").read()
    st = flatbuf.read_state(b)
    assert flatbuf.write_state(st.arrays, st.settings_raw, st.scripts) == b


def test_flatbuffer_reference_states_if_present():
    """Every saved state shipped with the reference (when the checkout is mounted)."""
    files = sorted(glob.glob("/root/reference/saved_states/*.bin"))
    if not files:
        pytest.skip("reference checkout not mounted")
    for f in files:
        b = open(f, "rb").read()
        st = flatbuf.read_state(b)
        assert flatbuf.write_state(st.arrays, st.settings_raw, st.scripts) == b, f


def test_state_load_save_roundtrip():
    s = Settings()
    path = os.path.join(GOLDEN, "saved_states", "PBEXPS.bin")
    b = open(path, "rb").read()
    st = State.load(b, s)
    assert st.p_count == 2 and s.simulation.use_f64 and s.physics.bonds == 3
    assert st.save(s) == b


def test_collision_settings_word_order():
    s = Settings()
    s.simulation.hor_bound, s.simulation.vert_bound = 7.0, 5.0
    s.simulation.timestep = 0.25
    s.physics.friction_coefficient = 0.125
    s.setup.particles = 42
    u = s.collision_settings()
    words = np.frombuffer(bytes(u), dtype=np.int32)
    f = words.view(np.float32)
    assert len(words) == 29
    assert f[1] == 7.0 and f[2] == 5.0 and f[10] == 0.125 and f[19] == 0.25 and words[28] == 42
    assert words[6] == 1 and words[7] == 1 and words[9] == 1   # defaults: gravity, planet mode, collisions


def test_grid_info_matches_reference_formula():
    import orc
    s = Settings()
    s.simulation.hor_bound, s.simulation.vert_bound = 3.2679124, 2.0
    s.setup.max_radius, s.setup.min_radius, s.setup.variable_rad = 0.05, 0.02, True
    s.setup.particles = 100
    g = s.grid_info()
    want = abi.DemGrid()
    orc.lib().orc_grid_capacity(C.c_float(3.2679124), C.c_float(2.0), C.c_float(0.05), C.c_float(0.02), 1, 100,
                                C.byref(want))
    assert (g.cell_size, g.cell_cap, g.w, g.h) == (want.cell_size, want.cell_cap, want.w, want.h)


def test_min_cell_size_covers_rounding():
    cs = min_cell_size(0.525, 4096.0)
    assert cs > 1.05 and cs < 1.06


def test_granular_box_no_overlap_and_inside():
    s, st = scenes.granular_box()
    n = st.p_count
    p = st.pos[: 2 * n].reshape(-1, 2)
    r = st.radii[:n]
    assert n == 4096 and r.min() >= 0.5 and r.max() <= 1.0
    assert np.all(np.abs(p) + r[:, None] <= 64.0)
    # lattice cells of side 2 keep every disc inside its own cell
    cell = np.floor((p + 64.0) / 2.0)
    assert len({tuple(c) for c in cell}) == n


def test_column_collapse_small():
    """The deposited pile: inside the box, radii in range, no pair closer than the
    designed overlap (5e-4 of the radius sum, recomputed in f64 from the f32
    positions, up to their rounding), three to five contacts per disc by the brute-force
    pair search, and most discs above the floor resting on two discs below them."""
    import orc
    s, st = scenes.column_collapse(n=4096, per_row=64, width=256.0, height=128.0)
    n = st.p_count
    r = st.radii[:n]
    assert r.min() >= 0.175 and r.max() <= 0.525
    p = st.pos[: 2 * n].reshape(-1, 2)
    assert np.all(p[:, 1] - r >= -64.0) and np.all(np.abs(p[:, 0]) + r <= 128.0)
    cnt = np.zeros(n, dtype=np.int32)
    part = np.zeros(n * 16, dtype=np.int32)
    tot = orc.lib().orc_contact_pairs(n, st.pos.ctypes.data_as(orc.FP), st.radii.ctypes.data_as(orc.FP), 16,
                                      cnt.ctypes.data_as(orc.IP), part.ctypes.data_as(orc.IP))
    assert 3.0 < tot / n < 5.0
    part = part.reshape(n, 16)
    pd = p.astype(np.float64)
    worst = 0.0
    on_floor = 0
    supports = [0, 0, 0]   # discs above the floor touching 0, 1, >= 2 discs below them
    for i in range(n):
        js = part[i, :cnt[i]]
        d = np.hypot(pd[js, 0] - pd[i, 0], pd[js, 1] - pd[i, 1])
        worst = max(worst, float(np.max(1.0 - d / (r[js].astype(np.float64) + r[i]), initial=0.0)))
        if pd[i, 1] - r[i] < -64.0 + 1e-3:
            on_floor += 1
            continue
        below = int(np.sum(pd[js, 1] < pd[i, 1]))
        supports[min(below, 2)] += 1
    assert 0 < on_floor < n // 10
    # drop-and-roll leaves most discs in a valley; a few balance on one disc
    assert supports[2] >= 0.8 * (n - on_floor) and supports[0] <= 0.01 * n
    # f32 rounding of coordinates below 128 moves a distance by at most ~2.2e-5, i.e.
    # 6.2e-5 of the smallest radius sum (0.35)
    assert worst < 5e-4 + 7e-5


def test_assemble_bonds_slots_sorted_by_partner():
    st = State(4)
    st.p_count = 4
    st.pos[:8] = [0, 0, 1, 0, 2, 0, 1, 1]
    pairs = np.array([[0, 1], [1, 2], [1, 3]], dtype=np.int32)
    bonds, contacts = assemble_bonds(st, pairs, 14)
    assert bonds.reshape(-1, 3)[:, 0].tolist() == [1, 1, 1]
    assert bonds.reshape(-1, 3)[:, 2].view(np.float32).tolist() == [1.0, 1.0, 1.0]
    t = contacts.view(np.int32).reshape(4, 14, 6)
    assert t[1, :3, 1].tolist() == [0, 2, 3] and t[1, :3, 5].tolist() == [0, 1, 2]
    assert t[0, 0, :2].tolist() == [0, 1] and t[3, 0, 1] == 1 and t[3, 1, 1] == -1


def test_large_body_split():
    """The large-body rule (radii in decreasing order, the first k <= 64 with
    r_{k-1} > 2 r_k sets the top k apart): brute force over small radius sets."""
    from psim.compute import small_radius

    def brute(r):
        r = sorted([x for x in r if x > 0], reverse=True)
        for k in range(1, min(len(r), 65)):
            if r[k - 1] > 2 * r[k]:
                return r[k]
        return r[0] if r else 0.0

    rng = np.random.default_rng(7)
    cases = [[8.0] + [0.5] * 100, [1.0] + [0.5] * 10,            # a striker; exactly 2x is no gap
             [8.0, 6.0] + [0.5] * 50, [3.0] * 70 + [1.0] * 10,    # two large bodies; 70 > 64 large
             [0.175, 0.35, 0.525] * 20, [0.0, 0.0, 4.0, 1.0],     # no gap; inactive discs
             []]
    cases += [list(rng.choice([0.5, 1.0, 2.5, 7.0], size=int(rng.integers(1, 90)))) for _ in range(200)]
    for c in cases:
        r = np.asarray(c, dtype=np.float32)
        assert small_radius(r) == np.float32(brute([float(x) for x in r])), c[:8]
