"""Bonded-particle path on the GPU: struck chains with per-bond Weibull strengths in
bond modes 1-3 (intact laws, tearing, broken-bond codes, contacts that take over
from broken bonds) and the bond builder on the cell list, against the oracle (the
gripped tensile assembly is in test_gpu_assembly).  Needs an MI355X."""
import numpy as np
import pytest

import orc
from psim import scenes
from psim.compute import ComputeProg

pytestmark = pytest.mark.gpu

FIELDS = ("pos", "vel", "acc", "rot", "rot_vel", "rot_acc")


def _n(name, n):
    return 2 * n if name in ("pos", "vel", "acc") else n


def assert_bits(a, b, what):
    a, b = np.asarray(a).view(np.int32), np.asarray(b).view(np.int32)
    if not np.array_equal(a, b):
        bad = np.nonzero(a != b)[0]
        raise AssertionError("%s differs at %d of %d words (first %d)" % (what, len(bad), len(a), bad[0]))


def chain_scene(mode, chains=8, per_chain=32, scale=50.0, seed=3):
    """Chains of touching r = 0.5 discs bonded to their neighbours, ends pinned,
    struck from below by a dense r = 1 disc at 40 units/s (dt 5e-5)."""
    s, st, strengths = scenes.bonded_chains(chains=chains, per_chain=per_chain, seed=seed, striker_r=1.0,
                                            striker_v=40.0, striker_density=50.0, strength_scale=scale)
    s.physics.bonds = mode
    scenes.chain_bonds(st, chains, per_chain, strengths)
    return s, st


@pytest.mark.parametrize("mode", [3, 2, 1])
def test_struck_chains_match_oracle(mode):
    s, st = chain_scene(mode)
    n, nb = st.p_count, st.n_bonds
    prog = ComputeProg(s, st.copy(), record_data=True, reorder_interval=8)
    eng = orc.OracleEngine(st, s, mode=orc.MODE_CANONICAL)
    for k in (120, 240, 240):       # several calls: the state crosses call boundaries
        prog.step(k)
        eng.step(k)
    got = prog.update_state()
    for name in FIELDS:
        assert_bits(getattr(got, name)[: _n(name, n)], getattr(eng.st, name)[: _n(name, n)], name)
    assert_bits(got.data[: 7 * n], eng.st.data[: 7 * n], "data")
    assert_bits(got.bonds[: 3 * nb], eng.st.bonds[: 3 * nb], "bond words (break codes)")
    want = eng.canonical_contacts().view(np.int32).reshape(-1, 14, 6)[:n]
    assert_bits(got.contacts.view(np.int32).reshape(-1, 14, 6)[:n], want, "contact table")
    broken = int((eng.st.bonds[0:3 * nb:3] < 0).sum())
    assert prog.broken_bond_count() == broken
    if mode == 3:
        assert broken > 0      # the strike tears parallel bonds
    prog.close()


@pytest.mark.parametrize("mode,block", [(3, 256), (1, 256), (3, 64)])
def test_struck_chains_without_data_match_oracle(mode, block):
    """Without data recording (the bond pass keeps no data log per record), at 256- and
    64-slot blocks: state, break codes and contact table equal the oracle, with the
    striker's neighbours in contact, so contacts merge with bond records."""
    s, st = chain_scene(mode)
    n, nb = st.p_count, st.n_bonds
    prog = ComputeProg(s, st.copy(), record_data=False, reorder_interval=8, block_size=block)
    eng = orc.OracleEngine(st, s, mode=orc.MODE_CANONICAL)
    in_contact = 0
    for k in (120, 240, 240):
        prog.step(k)
        eng.step(k)
        in_contact = max(in_contact, prog.counters()["contacts"])
    got = prog.update_state()
    for name in FIELDS:
        assert_bits(getattr(got, name)[: _n(name, n)], getattr(eng.st, name)[: _n(name, n)], name)
    assert_bits(got.bonds[: 3 * nb], eng.st.bonds[: 3 * nb], "bond words (break codes)")
    want = eng.canonical_contacts().view(np.int32).reshape(-1, 14, 6)[:n]
    assert_bits(got.contacts.view(np.int32).reshape(-1, 14, 6)[:n], want, "contact table")
    assert in_contact > 0          # contacts merged with bond records
    prog.close()


def _assert_same(got, eng, n, nb):
    for name in FIELDS:
        assert_bits(getattr(got, name)[: _n(name, n)], getattr(eng.st, name)[: _n(name, n)], name)
    assert_bits(got.data[: 7 * n], eng.st.data[: 7 * n], "data")
    assert_bits(got.bonds[: 3 * nb], eng.st.bonds[: 3 * nb], "bond words (break codes)")
    want = eng.canonical_contacts().view(np.int32).reshape(-1, 14, 6)[:n]
    assert_bits(got.contacts.view(np.int32).reshape(-1, 14, 6)[:n], want, "contact table")


def test_uploaded_broken_bonds_match_oracle():
    """Bonds uploaded broken while both endpoints' contact slots still hold them (a
    state saved between the two endpoints' passes, or edited): both endpoints take the
    bond's word at the first step, a bond law's code (-10) and any other negative word
    (-3) alike.  Then a torn state downloaded mid-run and restored continues equal to
    the oracle started from the same state."""
    s, st = chain_scene(3)
    n, nb = st.p_count, st.n_bonds
    cut = np.arange(5, nb, 17)
    st.bonds[3 * cut] = np.where(np.arange(len(cut)) % 2 == 0, -10, -3).astype(np.int32)
    prog = ComputeProg(s, st.copy(), record_data=True, reorder_interval=8)
    eng = orc.OracleEngine(st.copy(), s, mode=orc.MODE_CANONICAL)
    prog.step(120)
    eng.step(120)
    mid = prog.update_state()
    _assert_same(mid, eng, n, nb)
    assert prog.broken_bond_count() >= len(cut)
    prog.restore(mid.copy())
    eng = orc.OracleEngine(mid.copy(), s, mode=orc.MODE_CANONICAL)
    for k in (1, 239):
        prog.step(k)
        eng.step(k)
    _assert_same(prog.update_state(), eng, n, nb)
    prog.close()


def test_bond_builder_matches_oracle():
    """dem_build_bond_pairs (pair search on the cell list) + host assembly equal the
    oracle's State::regen_bonds: the same bonds in the same order, the same rest
    lengths, the same endpoint slots; the bond angle within one ulp (numpy's and
    libm's atan2)."""
    s, st = scenes.granular_box(n_side=48, seed=5, spacing=2.0, rmin=0.8, rmax=1.2)
    n = st.p_count
    max_bonds, slots = 6, 14
    s.setup.max_bonds, s.setup.max_contacts = max_bonds, slots
    prog = ComputeProg(s, st.copy(), bond_capacity=max_bonds * n)
    got = prog.regen_bonds(s)
    nb = got.n_bonds
    bonds_o = np.zeros(3 * max_bonds * n, dtype=np.int32)
    cont_o = np.zeros(slots * 6 * n, dtype=np.float32)
    nb_o = orc.lib().orc_regen_bonds(n, st.pos.ctypes.data_as(orc.FP), st.radii.ctypes.data_as(orc.FP), max_bonds,
                                     slots, bonds_o.ctypes.data_as(orc.IP), max_bonds * n,
                                     cont_o.ctypes.data_as(orc.FP), 1.05, 0.0)
    assert nb == nb_o > n          # a dense packing: more bonds than discs
    b = got.bonds[: 3 * nb].reshape(-1, 3)
    w = bonds_o[: 3 * nb].reshape(-1, 3)
    assert np.array_equal(b[:, 0], w[:, 0])
    assert np.array_equal(b[:, 2], w[:, 2])                       # rest length bits
    ulp = np.abs(b[:, 1].astype(np.int64) - w[:, 1].astype(np.int64))
    assert ulp.max() <= 1                                         # angle bits
    have = got.contacts.view(np.int32).reshape(-1, slots, 6)[:n]
    assert np.array_equal(have, cont_o.view(np.int32).reshape(-1, slots, 6)[:n])
    # after the re-upload the device holds the bonds: one step later every bond is intact
    prog.step(1)
    after = prog.update_state()
    assert prog.broken_bond_count() == 0 and np.all(after.bonds[0:3 * nb:3] == 1)
    prog.close()


def test_struck_chains_r8_striker_match_oracle():
    """Config 1's striker radius: an r = 8 dense disc among r = 0.5 chain discs (grid cell
    2 r_max + skin, so the striker's neighbour list overflows and it searches the
    stencil of the cell list every step).  Through the impact the state, break codes and
    contact table equal the oracle, with no contact dropped on either side."""
    s, st, strengths = scenes.bonded_chains(chains=8, per_chain=48, seed=9, striker_r=8.0, striker_v=40.0,
                                            striker_density=10.0, strength_scale=50.0)
    scenes.chain_bonds(st, 8, 48, strengths)
    n, nb = st.p_count, st.n_bonds
    prog = ComputeProg(s, st.copy(), record_data=True, reorder_interval=32)
    eng = orc.OracleEngine(st, s, mode=orc.MODE_CANONICAL)
    for k in (200, 300, 300):
        prog.step(k)
        eng.step(k)
    got = prog.update_state()
    for name in FIELDS:
        assert_bits(getattr(got, name)[: _n(name, n)], getattr(eng.st, name)[: _n(name, n)], name)
    assert_bits(got.bonds[: 3 * nb], eng.st.bonds[: 3 * nb], "bond words (break codes)")
    want = eng.canonical_contacts().view(np.int32).reshape(-1, 14, 6)[:n]
    assert_bits(got.contacts.view(np.int32).reshape(-1, 14, 6)[:n], want, "contact table")
    c = prog.counters()
    assert c["contact_overflow"] == 0 and c["list_overflow"] > 0
    striker = got.contacts.view(np.int32).reshape(-1, 14, 6)[n - 1]
    assert (striker[:, 1] >= 0).sum() > 0          # the striker reached the chains
    prog.close()


def test_config1_full_size_matches_oracle():
    """The struck-chain configuration at full size: 64 chains of 256 r = 0.5 discs
    (16,320 parallel bonds, Weibull(6, 50) strengths, ends pinned) struck by an r = 8
    disc of density 10 at 40 units/s, dt 5e-5, 20,000 steps.  The state, break codes,
    contact table (the striker's 14-slot cap included) and kinetic energy equal the
    oracle's at steps 5,000 and 20,000."""
    s, st, strengths = scenes.bonded_chains()
    scenes.chain_bonds(st, 64, 256, strengths)
    n, nb = st.p_count, st.n_bonds
    assert (n, nb) == (16385, 16320)
    prog = ComputeProg(s, st.copy(), record_data=True, reorder_interval=64)
    eng = orc.OracleEngine(st, s, mode=orc.MODE_CANONICAL)
    done = 0
    for k in (5000, 15000):
        prog.step(k)
        eng.step(k)
        done += k
        got = prog.update_state()
        for name in FIELDS:
            assert_bits(getattr(got, name)[: _n(name, n)], getattr(eng.st, name)[: _n(name, n)],
                        "%s at step %d" % (name, done))
        assert_bits(got.data[: 7 * n], eng.st.data[: 7 * n], "data at step %d" % done)
        assert_bits(got.bonds[: 3 * nb], eng.st.bonds[: 3 * nb], "bond words at step %d" % done)
        want = eng.canonical_contacts().view(np.int32).reshape(-1, 14, 6)[:n]
        have = got.contacts.view(np.int32).reshape(-1, 14, 6)[:n]
        assert_bits(have, want, "contact table at step %d" % done)
        ke = prog.kinetic_energy()
        o = eng.st
        ke_o = orc.lib().orc_kinetic_energy(n, o.vel.ctypes.data_as(orc.FP), o.rot_vel.ctypes.data_as(orc.FP),
                                            o.radii.ctypes.data_as(orc.FP), o.material_pointers.ctypes.data_as(orc.IP),
                                            eng.mats, len(eng.mats))
        assert abs(ke - ke_o) <= 1e-12 * ke_o
        striker = int((have[n - 1, :, 1] >= 0).sum())
        print("config 1 at step %d: %d of %d bonds broken, striker holds %d slots, kinetic energy %.6g"
              % (done, int((got.bonds[0:3 * nb:3] < 0).sum()), nb, striker, ke), flush=True)
    c = prog.counters()
    prog.close()
    assert int((got.bonds[0:3 * nb:3] < 0).sum()) > 0
    print("counters:", c, flush=True)
