"""Block-size and cell-size autotuner for the DEM step.

The two launch parameters that change the step's speed without changing its results
are the threads per block of the particle kernels and the grid's cell side (any cell
of at least 2 r_max plus the rounding margin finds the same contacts; what a cell
leaves above that is the neighbour-list skin: bigger cells mean fewer cells, longer
lists and rarer rebuilds).  `tune` times a short burst of
a scene for every pair on the device (HIP events around every launch) and returns
the fastest; `Tuner` caches the choice per scene class (kind and size bucket), so a
serving loop tunes once per class, not per scene.
"""
from __future__ import annotations

import math
from typing import Dict, Iterable, Optional, Tuple

from .compute import ComputeProg, default_cell_size, small_radius
from .settings import Settings
from .state import State

BLOCKS = (64, 128, 256)
# relative to the default cell (skin 0.1 x 2 r_max): 0.92 leaves a skin of about
# 0.02 x 2 r_max (more rebuilds, fewer list candidates), 1.25 a wider one
CELL_FACTORS = (0.92, 1.0, 1.25)


def cell_size_for(settings: Settings, state: State, factor: float) -> float:
    """`factor` times the default cell side (2 r_max + margin + neighbour-list skin)."""
    n = state.p_count
    rmax = float(max(state.radii[:n].max(), 1e-6)) if n else float(settings.setup.max_radius)
    rs = max(small_radius(state.radii[:n]), 1e-6) if n else rmax
    ext = max(settings.simulation.hor_bound, settings.simulation.vert_bound) + 8.0 * rmax
    return default_cell_size(rs, ext) * factor


def time_config(settings: Settings, state: State, block: int, factor: float, steps: int = 60, warmup: int = 5) -> float:
    """Mean milliseconds per step of `steps` steps (after `warmup`) at this block
    size and cell factor, from HIP events on the context's stream."""
    grid = ComputeProg.default_grid(settings, state, cell_size=cell_size_for(settings, state, factor))
    prog = ComputeProg(settings, state.copy(), grid=grid, record_data=False, block_size=block)
    try:
        prog.step(warmup)
        prog.sync()
        prog.timing(True)
        prog.step(steps)
        prog.sync()
        t = prog.timing(False)
    finally:
        prog.close()
    return t["total_ms"] / steps


# configurations this close to the fastest count as ties (the timing's run-to-run spread)
TIE = 0.02


def tune(settings: Settings, state: State, blocks: Iterable[int] = BLOCKS,
         factors: Iterable[float] = CELL_FACTORS, steps: int = 60, warmup: int = 5) -> Dict:
    """The cell factor: the default cell (factor 1) when a configuration with it is
    within TIE of the fastest (in a batch one scene's thin skin rebuilds every scene's
    lists), else the fastest configuration's; then the fastest block at that cell (the
    block changes only the speed, never the results)."""
    table = {}
    for b in blocks:
        for f in factors:
            table[(b, f)] = time_config(settings, state, b, f, steps, warmup)
    best = min(table.values())
    ties = [kv for kv in table.items() if kv[1] <= best * (1.0 + TIE)]
    f = 1.0 if any(k[1] == 1.0 for k, _ in ties) else min(ties, key=lambda kv: kv[1])[0][1]
    (b, f), ms = min(((k, v) for k, v in table.items() if k[1] == f), key=lambda kv: kv[1])
    return {"block": b, "cell_factor": f, "cell_size": cell_size_for(settings, state, f), "ms_per_step": ms,
            "table": {"%d/%.2f" % k: v for k, v in table.items()}}


class Tuner:
    """Per-class cache of tuned cell factors (and the block size of the first class,
    since one server runs one block size).  A class is (kind, log2 of the particle
    count)."""

    def __init__(self, blocks: Iterable[int] = BLOCKS, factors: Iterable[float] = CELL_FACTORS,
                 steps: int = 60, warmup: int = 5):
        self.blocks, self.factors = tuple(blocks), tuple(factors)
        self.steps, self.warmup = steps, warmup
        self.cache: Dict[Tuple[str, int], Dict] = {}

    @staticmethod
    def key(kind: str, state: State) -> Tuple[str, int]:
        return kind, int(math.log2(max(state.p_count, 1)))

    def choose(self, kind: str, settings: Settings, state: State, block: Optional[int] = None) -> Dict:
        k = self.key(kind, state)
        if k not in self.cache:
            blocks = (block,) if block else self.blocks
            self.cache[k] = tune(settings, state, blocks, self.factors, self.steps, self.warmup)
        return self.cache[k]

    def grid_for(self, kind: str, settings: Settings, state: State, block: Optional[int] = None):
        c = self.choose(kind, settings, state, block)
        return ComputeProg.default_grid(settings, state,
                                        cell_size=cell_size_for(settings, state, c["cell_factor"]))
