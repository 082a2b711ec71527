"""The bonded-assembly tensile test (scenes.bonded_assembly) on the GPU: the relaxed
packing's bond set equals the oracle's State::regen_bonds with the gap rule, the pull
equals the oracle bit for bit while bonds tear, and the device observables - the grip
group sums (dem_group_sums) and the kinetic energy (dem_kinetic_energy) - equal the
same sums evaluated on the oracle's state.  Needs an MI355X."""
import numpy as np
import pytest

import orc
from psim import scenes
from psim.compute import ComputeProg

pytestmark = pytest.mark.gpu

FIELDS = ("pos", "vel", "acc", "rot", "rot_vel", "rot_acc")
_SCENE = {}


def _n(name, n):
    return 2 * n if name in ("pos", "vel", "acc") else n


def assert_bits(a, b, what):
    a, b = np.asarray(a).view(np.int32), np.asarray(b).view(np.int32)
    if not np.array_equal(a, b):
        bad = np.nonzero(a != b)[0]
        raise AssertionError("%s differs at %d of %d words (first %d)" % (what, len(bad), len(a), bad[0]))


def _assembly():
    """48 x 48 discs (2,304) of the full-size assembly's construction."""
    if "s" not in _SCENE:
        _SCENE["s"] = scenes.bonded_assembly(rows=48, per_row=48, seed=3)
    s, st, bottom, top, info = _SCENE["s"]
    return s, st.copy(), bottom, top, info


def _group_expect(st, s, ids):
    """The group sums evaluated on a host state: f32 mass as the step forms it, f64
    products and sums."""
    mats = np.asarray(s.materials, dtype=np.float32).reshape(-1, 6)
    dens = mats[st.material_pointers[ids], 3]
    r = st.radii[ids]
    m = (dens * np.float32(np.pi) * r * r).astype(np.float64)
    acc = st.acc.reshape(-1, 2)[ids].astype(np.float64)
    pos = st.pos.reshape(-1, 2)[ids].astype(np.float64)
    vel = st.vel.reshape(-1, 2)[ids].astype(np.float64)
    d0 = np.abs(st.data.reshape(-1, 7)[ids, 0].astype(np.float64))
    return np.array([(m * acc[:, 0]).sum(), (m * acc[:, 1]).sum(), pos[:, 0].sum(), pos[:, 1].sum(),
                     vel[:, 0].sum(), vel[:, 1].sum(), d0.sum(), len(ids)])


def _dev_group(prog, ids):
    g = prog.group_sums(ids)
    return np.array([g["force"][0], g["force"][1], g["pos"][0], g["pos"][1], g["vel"][0], g["vel"][1],
                     g["normal_force"], g["count"]])


def test_assembly_packing_and_bonds_match_oracle():
    """The relaxed packing is dense (area fraction above 0.83, every disc within 0.08 of
    four or more others on average) and its bonds - built on the device with the gap
    rule |d| < r_a + r_b + 0.08 - are the oracle's State::regen_bonds with that rule:
    the same bonds in the same order, the same rest lengths and endpoint slots."""
    s, st, bottom, top, info = _assembly()
    n = st.p_count
    nb = st.n_bonds
    bonds_o = np.zeros(3 * 6 * n, dtype=np.int32)
    cont_o = np.zeros(14 * 6 * n, dtype=np.float32)
    nb_o = orc.lib().orc_regen_bonds(n, st.pos.ctypes.data_as(orc.FP), st.radii.ctypes.data_as(orc.FP), 6, 14,
                                     bonds_o.ctypes.data_as(orc.IP), 6 * n, cont_o.ctypes.data_as(orc.FP),
                                     1.0, np.float32(0.1 * 0.8))
    assert nb == nb_o
    b = st.bonds[: 3 * nb].reshape(-1, 3)
    w = bonds_o[: 3 * nb].reshape(-1, 3)
    assert np.array_equal(b[:, 0], w[:, 0]) and np.array_equal(b[:, 2], w[:, 2])
    assert np.abs(b[:, 1].astype(np.int64) - w[:, 1].astype(np.int64)).max() <= 1
    have = st.contacts.view(np.int32).reshape(-1, 14, 6)[:n]
    assert np.array_equal(have, cont_o.view(np.int32).reshape(-1, 14, 6)[:n])
    r = st.radii[:n].astype(np.float64)
    phi = np.pi * (r ** 2).sum() / (info["width"] * (info["L0"] + 2 * r.mean()))
    z = 2.0 * nb / n
    print("assembly 48 x 48: %d bonds, %.2f per disc, area fraction %.3f" % (nb, z, phi), flush=True)
    assert phi > 0.83 and z > 4.0


def test_tensile_pull_matches_oracle_and_observables():
    """The pull over three 800-step calls: state, break codes and contact table equal
    the oracle; bonds tear; the grip group sums and the kinetic energy from the device
    equal the sums over the oracle's state (f64 sums of equal terms: 1e-12 relative),
    and the grip stress is tensile."""
    s, st, bottom, top, info = _assembly()
    n, nb = st.p_count, st.n_bonds
    prog = ComputeProg(s, st.copy(), bond_capacity=nb, record_data=True, reorder_interval=8)
    eng = orc.OracleEngine(st.copy(), s, mode=orc.MODE_CANONICAL)
    stress = []
    for k in (800, 800, 800):
        prog.step(k)
        eng.step(k)
        o = eng.st
        for ids in (bottom, top):
            np.testing.assert_allclose(_dev_group(prog, ids), _group_expect(o, s, ids), rtol=1e-12, atol=1e-9)
        stress.append(scenes.grip_stress(prog, bottom, top, info["width"]))
        ke = prog.kinetic_energy()
        ke_o = orc.lib().orc_kinetic_energy(n, o.vel.ctypes.data_as(orc.FP), o.rot_vel.ctypes.data_as(orc.FP),
                                            o.radii.ctypes.data_as(orc.FP), o.material_pointers.ctypes.data_as(orc.IP),
                                            eng.mats, len(eng.mats))
        assert abs(ke - ke_o) <= 1e-12 * ke_o
    got = prog.update_state()
    for name in FIELDS:
        assert_bits(getattr(got, name)[: _n(name, n)], getattr(eng.st, name)[: _n(name, n)], name)
    assert_bits(got.data[: 7 * n], eng.st.data[: 7 * n], "data")
    assert_bits(got.bonds[: 3 * nb], eng.st.bonds[: 3 * nb], "bond words (break codes)")
    want = eng.canonical_contacts().view(np.int32).reshape(-1, 14, 6)[:n]
    assert_bits(got.contacts.view(np.int32).reshape(-1, 14, 6)[:n], want, "contact table")
    broken = prog.broken_bond_count()
    rebuilds = prog.counters()["broad_phases"]
    strain = scenes.grip_strain(prog, bottom, top, info["L0"])
    sel = np.zeros(n, dtype=np.int32)
    sel[top] = 1
    dev = prog.datum(sel)
    prog.close()
    got.selections[:] = 0
    got.selections[top] = 1
    np.testing.assert_allclose(dev, got.get_datum(), rtol=1e-12, atol=1e-12)
    print("assembly pull 2,400 steps: strain %.3g, grip stress %s, %d of %d bonds broken, %d list rebuilds"
          % (strain, ", ".join("%.4g" % x for x in stress), broken, nb, rebuilds), flush=True)
    assert 0 < broken < nb
    assert rebuilds <= 2400 // 8 + 20    # the re-sorts (every 8 steps) and the drift
    assert stress[0] > 0.0
    assert abs(strain - 2400 * 1e-6) < 2e-4


def test_group_sums_edge_cases():
    """No ids: zeros and a count of 0; an id outside [0, N) is refused with DEM_EINVAL;
    a repeated id counts twice."""
    from psim import abi
    s, st, bottom, top, info = _assembly()
    prog = ComputeProg(s, st, bond_capacity=st.n_bonds, record_data=False)
    try:
        prog.step(3)
        z = prog.group_sums(np.zeros(0, dtype=np.int32))
        assert z["count"] == 0 and z["force"] == (0.0, 0.0) and z["pos"] == (0.0, 0.0)
        with pytest.raises(abi.DemError) as ei:
            prog.group_sums([0, st.p_count])
        assert ei.value.code == abi.DEM_EINVAL
        one = prog.group_sums([5])
        two = prog.group_sums([5, 5])
        assert two["count"] == 2 and two["pos"][1] == 2.0 * one["pos"][1] and two["normal_force"] == 0.0
    finally:
        prog.close()


def test_pull_rebuilds_follow_drift_not_breaks():
    """16,384 discs pulled for 3,000 steps with the storage re-sort off: bonds break and
    their discs separate every few steps, and each such record's death puts the partner
    into its owner's neighbour list, so the lists are rebuilt only when a disc has moved
    0.45 skin (the grips move 1.1e-4 per step: about one rebuild per 1,000 steps), not
    after every death."""
    s, st, bottom, top, info = scenes.bonded_assembly(rows=128, per_row=128, seed=5)
    prog = ComputeProg(s, st, bond_capacity=st.n_bonds, record_data=False, reorder_interval=-1)
    try:
        prog.step(3000)
        c = prog.counters()
        broken = prog.broken_bond_count()
    finally:
        prog.close()
    print("16k pull, 3,000 steps: %d bonds broken, %d list rebuilds" % (broken, c["broad_phases"]), flush=True)
    assert broken > 100
    assert c["broad_phases"] < 20
