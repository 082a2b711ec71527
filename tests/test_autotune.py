"""The autotuner's choice rule, on timings supplied in place of device runs (CPU)."""
import pytest

import psim.autotune as autotune


@pytest.fixture
def fake_timings(monkeypatch):
    def use(table):
        monkeypatch.setattr(autotune, "time_config", lambda s, st, b, f, steps, warmup: table[(b, f)])
        monkeypatch.setattr(autotune, "cell_size_for", lambda s, st, f: f)
        blocks = sorted({k[0] for k in table})
        factors = sorted({k[1] for k in table})
        return autotune.tune(None, None, blocks, factors)
    return use


def test_fastest_block_at_a_thin_skin(fake_timings):
    # the 4M collapse: the thin skin wins by far; blocks within 2 % of each other
    r = fake_timings({(64, 0.92): 0.2122, (64, 1.0): 0.2281, (128, 0.92): 0.2135, (128, 1.0): 0.2299,
                      (256, 0.92): 0.2161, (256, 1.0): 0.2322})
    assert (r["block"], r["cell_factor"]) == (64, 0.92)
    assert r["ms_per_step"] == 0.2122


def test_default_cell_preferred_within_the_tie(fake_timings):
    # a thin skin only 1 % faster: the default cell, at its fastest block
    r = fake_timings({(128, 0.92): 1.00, (128, 1.0): 1.012, (256, 0.92): 1.005, (256, 1.0): 1.01})
    assert (r["block"], r["cell_factor"]) == (256, 1.0)


def test_wider_cell_when_clearly_faster(fake_timings):
    r = fake_timings({(256, 1.0): 1.05, (256, 1.25): 1.0})
    assert (r["block"], r["cell_factor"]) == (256, 1.25)
