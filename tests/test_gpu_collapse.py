"""The polydisperse column collapse (the headline workload, 4,194,304 discs at full
size) on the GPU against the oracle: a reduced column bit for bit through a contact-rich
window with neighbour-list rebuilds and storage re-sorts, and the full-size column's
contact pairs against the oracle's cell-list pair search.  Needs an MI355X."""
import numpy as np
import pytest

import orc
from psim import scenes
from psim.autotune import cell_size_for
from psim.compute import ComputeProg

from test_gpu_parity import assert_bits_equal

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("restitution", [0.5, 0.0])
def test_column_collapse_matches_oracle(restitution):
    """65,536 discs (256 mean diameters wide, same radius law, deposited in contact) in a
    512 x 256 box on the tuned grid (0.92 x the default cell: a skin of ~0.02 x 2 r_max),
    re-sorted every 16 steps: state, data words and contact table identical to the
    oracle (canonical contact order) after each of three dem_step calls, the window
    holding > 1 contact per disc, rebuilds and re-sorts.  With the contact dashpot of the
    configurations (restitution 0.5) and with the reference's undamped law."""
    n = 65536
    s, st = scenes.column_collapse(n=n, per_row=256, width=512.0, height=256.0, restitution=restitution)
    grid = ComputeProg.default_grid(s, st, cell_size=cell_size_for(s, st, 0.92))
    prog = ComputeProg(s, st.copy(), grid=grid, reorder_interval=16)
    eng = orc.OracleEngine(st, s, mode=orc.MODE_CANONICAL)
    c0 = prog.counters()
    for k in (40, 60, 100):
        prog.step(k)
        eng.step(k)
        got = prog.update_state()
        print("collapse e=%.1f: %d steps, contacts/disc %.3f" % (restitution, k, prog.counters()["contacts"] / n),
              flush=True)
        assert_bits_equal(got, eng.st, n)
        assert np.array_equal(got.data[: 7 * n].view(np.int32), eng.st.data[: 7 * n].view(np.int32))
        want = eng.canonical_contacts().view(np.int32).reshape(-1, 14, 6)[:n]
        assert np.array_equal(got.contacts.view(np.int32).reshape(-1, 14, 6)[:n], want)
        assert prog.counters()["contacts"] > n      # more than one contact per disc
    c = prog.counters()
    assert c["broad_phases"] - c0["broad_phases"] >= 3 and c["reorders"] - c0["reorders"] >= 3
    assert c["contact_overflow"] == 0 and c["scatter_errors"] == 0
    prog.close()


def test_column_collapse_full_size_contact_pairs():
    """4,194,304 discs after 100 steps of the collapse (about 3 contacts per disc): the
    contact lists of the last narrow phase equal the oracle's O(N) cell-list pair search
    over the downloaded positions, partner for partner; no overflow, nothing NaN."""
    s, st = scenes.column_collapse()
    n = st.p_count
    grid = ComputeProg.default_grid(s, st, cell_size=cell_size_for(s, st, 0.92))
    prog = ComputeProg(s, st, grid=grid, record_data=False, reorder_interval=256)
    prog.step(100)
    got = prog.update_state()
    counts, partners = prog.debug_contacts(16)
    c = prog.counters()
    prog.close()
    assert np.all(np.isfinite(got.pos[: 2 * n])) and np.all(np.isfinite(got.vel[: 2 * n]))
    assert c["contact_overflow"] == 0 and c["list_overflow"] == 0 and c["scatter_errors"] == 0
    w_cnt = np.zeros(n, dtype=np.int32)
    w_part = np.zeros(n * 16, dtype=np.int32)
    tot = orc.lib().orc_contact_pairs_grid(n, got.pos.ctypes.data_as(orc.FP), got.radii.ctypes.data_as(orc.FP), 16,
                                           w_cnt.ctypes.data_as(orc.IP), w_part.ctypes.data_as(orc.IP))
    print("full-size collapse: %d pairs, %.3f contacts/disc" % (tot, tot / n), flush=True)
    assert tot > 2 * n
    assert np.array_equal(counts, w_cnt)
    assert np.array_equal(partners, w_part.reshape(n, 16))
