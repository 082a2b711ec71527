"""`ComputeProg`: the host object that stands in for the reference's
`WGPUComputeProg` (src/wgpu_prog.rs:677-696) on top of the C ABI (include/dem.h).

Reference method          -> here
  WGPUComputeProg::new      ComputeProg(settings, state)        dem_create + dem_upload_state
  compute (gen_per_frame)   compute(settings) / step(n)         dem_step
  update_state              update_state() -> State             dem_download_state
  restore                   restore(state)                      dem_upload_state
  collision_settings update set_params(settings)                dem_set_params
  State::get_datum          datum(selections)                   dem_get_datum
  State::regen_bonds        regen_bonds(settings)               dem_build_bond_pairs on the device

There is no CPU path: every call goes to libpsim.so, and constructing a
ComputeProg without a HIP device raises.
"""
from __future__ import annotations

import ctypes as C
import math
import sys
from typing import Optional

import numpy as np

from . import abi
from .scripts import apply_set_properties
from .settings import Settings, f32
from .state import State


class ComputeProg:
    def __init__(self, settings: Settings, state: State, capacity: Optional[int] = None,
                 bond_capacity: Optional[int] = None, grid: Optional[abi.DemGrid] = None,
                 use_f64: Optional[bool] = None, record_data: bool = True, reorder_interval: int = 64,
                 block_size: int = 256, graph: bool = False):
        self.lib = abi.load_library()
        if self.lib.dem_device_count() <= 0:
            raise abi.DemError(abi.DEM_EDEVICE, "no HIP device visible")
        self.settings = settings
        self.use_f64 = settings.simulation.use_f64 if use_f64 is None else bool(use_f64)
        self.record_data = record_data
        self.reorder_interval = reorder_interval
        self.block_size = block_size
        self.graph = bool(graph)
        self.grid = grid if grid is not None else self.default_grid(settings, state)
        self.capacity = int(capacity if capacity is not None else max(state.p_count, 1))
        nb = state.n_bonds
        self.bond_capacity = int(bond_capacity if bond_capacity is not None else max(nb, 1))
        self._mats = None
        params = self._params(settings)
        ctx = C.c_void_p()
        abi.check(self.lib.dem_create(C.byref(params), self.capacity, self.bond_capacity, C.byref(ctx)))
        self.ctx = ctx
        self.state = state
        self.restore(state)

    # -- configuration -----------------------------------------------------
    @staticmethod
    def default_grid(settings: Settings, state: State, cell_size: Optional[float] = None) -> abi.DemGrid:
        """Cells over the walled box plus one cell of margin, of side
        default_cell_size() (2 r_max, the rounding margin and a neighbour-list skin of
        SKIN_FRACTION * 2 r_max) or `cell_size` (at least min_cell_size()).  Every such
        grid gives the same contacts; the cell side only trades cells for candidates
        and rebuild frequency."""
        n = state.p_count
        rmax = float(np.max(state.radii[:n])) if n else float(settings.setup.max_radius)
        rmax = max(rmax, 1e-6)
        rs = max(small_radius(state.radii[:n]), 1e-6) if n else rmax   # large bodies aside
        ext = max(settings.simulation.hor_bound, settings.simulation.vert_bound) + 8.0 * rmax
        if cell_size is not None:
            cs = max(min_cell_size(rs, ext), float(cell_size))
        else:
            cs = default_cell_size(rs, ext)
        return settings.box_grid(cell_size=cs)

    def _params(self, settings: Settings) -> abi.DemParams:
        p = abi.DemParams()
        p.u = settings.collision_settings()
        self._mats = abi.materials_array(settings.materials)
        p.materials = C.cast(self._mats, C.POINTER(abi.DemMaterial))
        p.n_materials = len(self._mats)
        p.grid = self.grid
        p.use_f64 = int(self.use_f64)
        p.record_data = int(self.record_data)
        p.reorder_interval = int(self.reorder_interval)
        p.block_size = int(self.block_size)
        p.graph = int(self.graph)
        p.restitution = float(settings.physics.restitution)
        return p

    def set_params(self, settings: Settings = None):
        if settings is not None:
            self.settings = settings
        p = self._params(self.settings)
        abi.check(self.lib.dem_set_params(self.ctx, C.byref(p)))

    # -- the step ------------------------------------------------------------
    def compute(self, settings: Settings = None):
        """WGPUComputeProg::compute: gen_per_frame steps, asynchronous."""
        s = settings or self.settings
        self.step(s.simulation.gen_per_frame)

    def step(self, n: int = 1):
        abi.check(self.lib.dem_step(self.ctx, int(n)))

    def sync(self):
        abi.check(self.lib.dem_sync(self.ctx))

    # -- state transfer ------------------------------------------------------
    def restore(self, state: Optional[State] = None):
        """WGPUComputeProg::restore: uploads the host state (copied)."""
        if state is not None:
            self.state = state
        st = self.state
        v = st.view(st.p_count, with_data=True)
        abi.check(self.lib.dem_upload_state(self.ctx, C.byref(v)))

    def update_state(self) -> State:
        """WGPUComputeProg::update_state: downloads into self.state (synchronous)."""
        st = self.state
        v = st.view(st.p_count, with_data=self.record_data)
        abi.check(self.lib.dem_download_state(self.ctx, C.byref(v)))
        return st

    # -- observables -------------------------------------------------------------
    def kinetic_energy(self) -> float:
        out = C.c_double()
        abi.check(self.lib.dem_kinetic_energy(self.ctx, C.byref(out)))
        return out.value

    def broken_bond_count(self) -> int:
        out = C.c_int64()
        abi.check(self.lib.dem_broken_bond_count(self.ctx, C.byref(out)))
        return out.value

    def counters(self) -> dict:
        out = (C.c_int64 * 12)()
        abi.check(self.lib.dem_counters(self.ctx, out, 12))
        keys = ["steps", "broad_phases", "contact_overflow", "bond_overflow", "contacts", "bond_endpoints",
                "reorders", "scan_timeouts", "scatter_errors", "list_overflow", "graph_captures", "graph_replays"]
        return {k: int(out[i]) for i, k in enumerate(keys)}

    def datum(self, selections) -> Optional[np.ndarray]:
        if selections is None:
            return None
        n = self.state.p_count
        sel = np.zeros(n, dtype=np.int32)
        k = min(len(selections), n)
        sel[:k] = selections[:k]
        if not sel.any():
            return None
        out = (C.c_double * 10)()
        abi.check(self.lib.dem_get_datum(self.ctx, abi.iptr(sel), out))
        return np.array(out[:], dtype=np.float64)

    def timing(self, enable: Optional[bool] = None):
        ms = (C.c_double * 6)()
        cnt = (C.c_int64 * 6)()
        abi.check(self.lib.dem_last_step_timing(self.ctx, -1 if enable is None else int(enable), ms, cnt))
        return {"total_ms": ms[0], "step_ms": ms[1], "scan_ms": ms[2], "scatter_ms": ms[3], "reorder_ms": ms[4],
                "bonds_ms": ms[5], "launches": cnt[0], "n_step": cnt[1], "n_scan": cnt[2], "n_scatter": cnt[3],
                "n_reorder": cnt[4], "n_bonds": cnt[5]}

    # -- parity views ------------------------------------------------------------
    def debug_broad_phase(self):
        n = self.state.p_count
        g = self.grid
        keys = np.zeros(n, dtype=np.int32)
        order = np.zeros(n, dtype=np.int32)
        start = np.zeros(g.w * g.h + 1, dtype=np.int32)
        abi.check(self.lib.dem_debug_broad_phase(self.ctx, abi.iptr(keys), abi.iptr(order), abi.iptr(start)))
        return keys, order, start

    def debug_contacts(self, max_per: int = 16):
        n = self.state.p_count
        counts = np.zeros(n, dtype=np.int32)
        partners = np.zeros(n * max_per, dtype=np.int32)
        abi.check(self.lib.dem_debug_contacts(self.ctx, max_per, abi.iptr(counts), abi.iptr(partners)))
        return counts, partners.reshape(n, max_per)

    def bond_pairs(self, max_bonds: int = 6, factor: float = 1.05, cap: Optional[int] = None,
                   gap: float = 0.0) -> np.ndarray:
        n = self.state.p_count
        cap = int(cap if cap is not None else max(n * max_bonds, 1))
        pairs = np.zeros(2 * cap, dtype=np.int32)
        nout = C.c_int32()
        abi.check(self.lib.dem_build_bond_pairs(self.ctx, int(max_bonds), C.c_float(factor), C.c_float(gap),
                                                abi.iptr(pairs), cap, C.byref(nout)))
        return pairs[: 2 * nout.value].reshape(-1, 2)

    def regen_bonds(self, settings: Optional[Settings] = None, factor: float = 1.05,
                    strengths: Optional[np.ndarray] = None, gap: float = 0.0) -> State:
        """State::regen_bonds (state.rs:291-367) with the pair search on the device:
        bond (i, j), j > i, when |pj - pi| < (ri + rj) * factor + gap (the reference:
        factor 1.05, gap 0), at most max_bonds per lower index in increasing j;
        bond = {1, atan2(dx, dy) bits, |d| bits}; each endpoint gets
        {a, b, 0, 0, 0, bond index} in its first free slot.  The bonds and slots are
        assembled on the host and the scene is re-uploaded."""
        s = settings or self.settings
        st = self.update_state()
        pairs = self.bond_pairs(s.setup.max_bonds, factor, gap=gap)
        st.bonds, st.contacts = assemble_bonds(st, pairs, s.setup.max_contacts)
        st.bond_strengths = strengths
        nb = st.n_bonds
        if nb > self.bond_capacity:
            raise abi.DemError(abi.DEM_ECAPACITY, "ComputeProg was created with bond_capacity %d < %d bonds"
                               % (self.bond_capacity, nb))
        self.restore(st)
        return st

    def group_sums(self, ids) -> dict:
        """Sums over the particles `ids` computed on the device (dem_group_sums): the
        net force on the group (Σ m a, m the step's f32 mass), Σ position, Σ velocity,
        Σ |normal-force datum| and the count.  The grip force and displacement of a
        tensile test; the reference reads the state back for the same sums
        (State::get_datum, state.rs:708-737)."""
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        out = (C.c_double * 8)()
        abi.check(self.lib.dem_group_sums(self.ctx, abi.iptr(ids), len(ids), out))
        return {"force": (out[0], out[1]), "pos": (out[2], out[3]), "vel": (out[4], out[5]),
                "normal_force": out[6], "count": int(out[7])}

    # -- experiment-script engine interface (psim.scripts.ScriptReplay) ----------
    @property
    def dt(self):
        return np.float32(self.settings.simulation.timestep)

    def snapshot(self) -> State:
        return self.update_state().copy()

    def restore_snapshot(self, snap: State):
        st = snap.copy()
        st.data[:] = 0.0
        self.restore(st)

    def set_properties(self, words, selections):
        st = self.update_state()
        moved = apply_set_properties(st, words, selections)
        if moved:
            raise NotImplementedError("position sets (del_pos override) are not exposed through the C ABI")
        self.restore(st)

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.dem_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        # No HIP calls while the interpreter shuts down: the runtime (and a profiler
        # hooked into it) may already be torn down, and process exit frees the device
        # memory anyway.  Callers close explicitly.
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass


# Neighbour-list skin of the default grid, as a fraction of 2 r_max: the lists are
# rebuilt when a particle has moved 0.45 skin since the last build.
SKIN_FRACTION = 0.1


LB_MAX = 64


def small_radius(radii) -> float:
    """Largest radius once the large bodies are set aside: with the radii in decreasing
    order r_0 >= r_1 >= ..., the first k <= 64 with r_{k-1} > 2 r_k makes the top k
    discs large bodies (they pair through a list, not through the cells) and r_k the
    answer; without such a gap it is r_0.  The library applies the same rule."""
    r = np.asarray(radii, dtype=np.float32)
    r = r[r > 0]
    if len(r) == 0:
        return 0.0
    k = min(len(r), LB_MAX + 1)
    top = -np.sort(-np.partition(r, len(r) - k)[len(r) - k:])
    for i in range(1, len(top)):
        if top[i - 1] > np.float32(2.0) * top[i]:
            return float(top[i])
    return float(top[0])


def default_cell_size(rmax: float, extent: float) -> float:
    return float(np.float32(min_cell_size(rmax, extent) + SKIN_FRACTION * 2.0 * rmax))


def min_cell_size(rmax: float, extent: float) -> float:
    """Smallest cell side for which the 3x3 stencil finds every overlapping pair:
    2 r_max plus room for the rounding of (x - base) / cell_size, whose absolute
    error is a few ulp of the grid extent."""
    ulp = np.spacing(np.float32(extent))
    return float(np.float32(2.0 * rmax * (1.0 + 1e-4) + 8.0 * float(ulp)))


def assemble_bonds(st: State, pairs: np.ndarray, slots: int):
    """Bond triples and endpoint slots for the (i, j) pairs, in regen_bonds order."""
    pos = st.pos.reshape(-1, 2)
    cap = st.capacity
    contacts = np.full(cap * slots * 6, -1, dtype=np.int32)
    if len(pairs) == 0:
        return np.full(1, -1, dtype=np.int32), contacts.view(np.float32)
    i = pairs[:, 0].astype(np.int64)
    j = pairs[:, 1].astype(np.int64)
    dx = (pos[j, 0] - pos[i, 0]).astype(np.float32)
    dy = (pos[j, 1] - pos[i, 1]).astype(np.float32)
    mag = np.sqrt((dx * dx + dy * dy).astype(np.float32)).astype(np.float32)
    # atan2 of the f32 unit vector (state.rs:306-310), evaluated in f64 and rounded once:
    # numpy's f32 arctan2 is off by up to 2 ulp
    nx = (dx / mag).astype(np.float32)
    ny = (dy / mag).astype(np.float32)
    ang = np.arctan2(nx.astype(np.float64), ny.astype(np.float64)).astype(np.float32)
    nb = len(pairs)
    bonds = np.empty(3 * nb, dtype=np.int32)
    bonds[0::3] = 1
    bonds[1::3] = ang.view(np.int32)
    bonds[2::3] = mag.view(np.int32)
    # endpoint slots: each particle's bonds fill its slots in creation order, which is
    # ascending partner id (partners below it were bonded first, as their own pairs)
    owner = np.concatenate([i, j])
    partner = np.concatenate([j, i])
    bid = np.concatenate([np.arange(nb), np.arange(nb)])
    order = np.lexsort((partner, owner))
    owner, partner, bid = owner[order], partner[order], bid[order]
    first = np.searchsorted(owner, owner, side="left")
    rank = np.arange(len(owner)) - first
    keep = rank < slots
    c = contacts.reshape(cap, slots, 6)
    o, r = owner[keep], rank[keep]
    c[o, r, 0] = o
    c[o, r, 1] = partner[keep]
    c[o, r, 2] = 0
    c[o, r, 3] = 0
    c[o, r, 4] = 0
    c[o, r, 5] = bid[keep]
    return bonds, contacts.view(np.float32)
