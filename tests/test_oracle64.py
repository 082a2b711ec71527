"""Pins of the f64 oracle ("ref64", oracle/dem_oracle64.cpp): it agrees with the f32
oracle to f32 rounding on a dense box, its dashpot rebounds at the closed-form
restitution, and its bond law 1 tears at the closed-form step.  CPU only."""
import math

import numpy as np

import orc
from psim import scenes

from test_oracle_branches import _gap, _mass, _pulled_pair, _scene


def test_ref64_agrees_with_f32_oracle():
    """A deposited pile of 1,024 discs in contact (spring-dashpot contacts): 100 steps of
    the f32 oracle and of ref64 from the same state - positions within 1e-3 (the f32
    run's rounding grows with the chaos of the packing), kinetic energy within 1e-3."""
    s, st = scenes.column_collapse(n=1024, per_row=32, width=128.0, height=64.0)
    n = st.p_count
    e = orc.OracleEngine(st, s, mode=orc.MODE_CANONICAL)
    e.step(100)
    r64 = orc.run64(st, s, 100)
    assert np.abs(e.st.pos[: 2 * n].astype(np.float64) - r64["pos"]).max() < 1e-3
    ke = orc.lib().orc_kinetic_energy(n, e.st.vel.ctypes.data_as(orc.FP), e.st.rot_vel.ctypes.data_as(orc.FP),
                                      e.st.radii.ctypes.data_as(orc.FP), e.st.material_pointers.ctypes.data_as(orc.IP),
                                      e.mats, len(e.mats))
    assert abs(ke - r64["ke"][-1]) < 1e-3 * r64["ke"][-1]
    assert np.count_nonzero(r64["data"].reshape(-1, 7)[:, 0]) > n // 2


def test_ref64_restitution_closed_form():
    """Two equal discs meeting head on rebound at the restitution of the dashpot."""
    for e_ in (0.3, 0.7):
        s, st = _scene(2)
        s.physics.restitution = e_
        s.simulation.timestep = 1e-5
        st.radii[:2] = 0.5
        st.pos[:4] = [-0.5005, 0.0, 0.5005, 0.0]
        st.vel[:4] = [1.0, 0.0, -1.0, 0.0]
        v = orc.run64(st, s, 2000)["vel"]
        assert abs((v[2] - v[0]) / 2.0 - e_) < 0.01 * e_


def test_ref64_normal_bond_tears_at_closed_form_step():
    """Bond law 1 pulled apart at 1 unit/s: intact while kn (|d| - 1) < strength, torn
    (-9) from the first step at which it reaches it."""
    kn_b, sn = 1.0e3, 40.0
    s, st = _pulled_pair(1, kn_b, sn)
    k_star = None
    for k in range(30, 60):
        r = orc.run64(st, s, k)
        disp = float(r["pos"][2] - r["pos"][0]) - 1.0
        if kn_b * disp >= sn:
            k_star = k
            assert r["bonds"][0] == -9
            before = orc.run64(st, s, k - 1)
            assert before["bonds"][0] == 1
            break
    assert k_star is not None
