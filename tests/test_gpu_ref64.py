"""The f32 HIP path against the f64 oracle ("ref64", oracle/dem_oracle64.cpp): one-step
positions, forces and contact sums, kinetic-energy traces of 1,000-step runs within 1 %,
and the bond-breakage count of struck chains equal.  The errors found are printed.
(The f32 path equals the f32 oracle bit for bit in the other GPU tests; these bound its
distance from exact arithmetic.)  Needs an MI355X."""
import numpy as np
import pytest

import orc
from psim import scenes
from psim.compute import ComputeProg

pytestmark = pytest.mark.gpu


def _pile():
    """4,096 discs of the headline column's radius law, deposited in contact (about four
    contacts per disc), configuration 0's contact law (kn_eff = 1e5, mu = 0.5, e = 0.5)."""
    return scenes.column_collapse(n=4096, per_row=64, width=256.0, height=128.0)


@pytest.mark.parametrize("scene,pre", [("pile", 20), ("pile", 500), ("config0", 12000)])
def test_one_step_against_ref64(scene, pre):
    """One step on the GPU and in f64 from the same state (the pile after 20 and 500
    steps; configuration 0 after 12,000, when its lattice has landed).  A one-step
    position differs by one f32 rounding of the drift.  A contact force in f32 is
    resolved no better than kn_eff times the spacing of f32 positions (ulp(|x|), ~7.6e-6
    at |x| ~ 100, against overlaps of ~1e-4 to 1e-2): each particle's force error is
    bounded by 10 kn_eff ulp(|x|) per contact (measured: p99 ~0.5-0.8, max ~1-6)."""
    s, st = _pile() if scene == "pile" else scenes.granular_box(seed=0)
    n = st.p_count
    prog = ComputeProg(s, st.copy(), reorder_interval=8)
    prog.step(pre)
    s0 = prog.update_state().copy()
    prog.step(1)
    got = prog.update_state()
    prog.close()
    r64 = orc.run64(s0, s, 1)
    pos0 = s0.pos[: 2 * n].reshape(-1, 2)
    ulp = np.spacing(np.abs(pos0).max(axis=1).astype(np.float32)).astype(np.float64)
    dp = np.abs(got.pos[: 2 * n].astype(np.float64) - r64["pos"]).reshape(-1, 2).max(axis=1)
    mass = np.pi * s0.radii[:n].astype(np.float64) ** 2          # density 1
    df = np.hypot(*(got.acc[: 2 * n].astype(np.float64) - r64["acc"]).reshape(-1, 2).T) * mass
    fmax = (np.hypot(*r64["acc"].reshape(-1, 2).T) * mass).max()
    ncon = (s0.contacts.view(np.int32).reshape(-1, 14, 6)[:n, :, 1] >= 0).sum(axis=1)
    ratio = df / (1e5 * ulp * np.maximum(ncon, 1))
    print("f32 GPU vs ref64, %s after %d steps (%.2f contacts per disc): one-step position within %.2f ulp; "
          "force error %.3g of the largest force, %.3g (max) / %.3g (p99) kn_eff ulp per contact"
          % (scene, pre, ncon.mean(), (dp / ulp).max(), df.max() / fmax, ratio.max(), np.percentile(ratio, 99)),
          flush=True)
    assert ncon.sum() > 0
    assert (dp / ulp).max() <= 1.0
    assert ratio.max() < 10.0


@pytest.mark.parametrize("scene", ["config0", "pile"])
def test_kinetic_energy_trace_against_ref64(scene):
    """1,000 steps: configuration 0 as specified (the 4,096-disc box from its lattice) and
    the pile in contact; the device's kinetic energy every 50 steps against the f64 run's
    (the largest relative difference is printed)."""
    s, st = scenes.granular_box(seed=0) if scene == "config0" else _pile()
    prog = ComputeProg(s, st.copy(), reorder_interval=64, record_data=False)
    ke, nc = [], []
    for _ in range(20):
        prog.step(50)
        ke.append(prog.kinetic_energy())
        nc.append(prog.counters()["contacts"] / st.p_count)
    got = prog.update_state()
    prog.close()
    pos = got.pos[: 2 * st.p_count].astype(np.float64)
    r64 = orc.run64(st, s, 1000)
    k64 = r64["ke"][49::50]
    rel = np.abs(np.array(ke) - k64) / k64
    print("f32 GPU vs ref64, %s 1000 steps (%.2f contacts per disc on average): kinetic energy within %.3g, "
          "positions within %.3g" % (scene, float(np.mean(nc)), rel.max(), np.abs(pos - r64["pos"]).max()),
          flush=True)
    assert rel.max() < 1e-2


def test_breakage_count_against_ref64():
    """Bonded chains (parallel bonds, Weibull strengths) struck by an r = 8 disc, 800 steps:
    the number of broken bonds equals the f64 run's."""
    s, st, strengths = scenes.bonded_chains(chains=8, per_chain=48, seed=5, striker_r=8.0, striker_v=40.0,
                                            striker_density=10.0, strength_scale=30.0)
    scenes.chain_bonds(st, 8, 48, strengths)
    prog = ComputeProg(s, st.copy(), record_data=False, reorder_interval=32)
    prog.step(800)
    broken = prog.broken_bond_count()
    prog.close()
    r64 = orc.run64(st, s, 800)
    b64 = int((r64["bonds"][: st.n_bonds] < 0).sum())
    print("f32 GPU vs ref64: %d vs %d broken bonds of %d" % (broken, b64, st.n_bonds), flush=True)
    assert b64 > 0 and abs(broken - b64) <= max(1, b64 // 100)
