"""Client of the serving loop (dem_server_* in include/dem.h): continuous batching of
independent scenes on one GPU.

The reference runs one scene per window; serving many scenes has no counterpart
there.  A `Server` owns one set of device pools.  `submit` queues a scene (settings +
State + step count, copied at once); every `tick` admits queued scenes whose
footprint fits, advances all resident scenes by up to `substeps` steps in shared
launches and retires the finished ones; `take` returns a finished scene's State.
A scene's result equals its solo run through `ComputeProg` bit for bit.
"""
from __future__ import annotations

import ctypes as C
import sys
import time
from typing import Optional

import numpy as np

from . import abi
from .compute import ComputeProg
from .settings import Settings
from .state import State


def scene_params(settings: Settings, state: State, grid: Optional[abi.DemGrid] = None,
                 use_f64: Optional[bool] = None, record_data: bool = True, reorder_interval: int = 64,
                 block_size: int = 256):
    """dem_params for one scene (as ComputeProg builds them); returns (params, keep-alive)."""
    p = abi.DemParams()
    p.u = settings.collision_settings()
    mats = abi.materials_array(settings.materials)
    p.materials = C.cast(mats, C.POINTER(abi.DemMaterial))
    p.n_materials = len(mats)
    p.grid = grid if grid is not None else ComputeProg.default_grid(settings, state)
    p.use_f64 = int(settings.simulation.use_f64 if use_f64 is None else use_f64)
    p.record_data = int(record_data)
    p.reorder_interval = int(reorder_interval)
    p.block_size = int(block_size)
    p.restitution = float(settings.physics.restitution)
    return p, mats


def footprint(state: State, grid: abi.DemGrid, block_size: int = 256):
    """Pool units a scene takes: (slots, cells, bonds)."""
    n = max(state.p_count, 1)
    return ((n + block_size - 1) // block_size * block_size, grid.w * grid.h + 1, state.n_bonds)


class Server:
    def __init__(self, slot_capacity: int, cell_capacity: int, bond_capacity: int = 0, scene_capacity: int = 64,
                 substeps: int = 50, max_resident: int = 0, bypass: int = 8, record_data: bool = True,
                 block_size: int = 256, contact_slots: int = 14, record_capacity: int = 0, graph: bool = False):
        self.lib = abi.load_library()
        if self.lib.dem_device_count() <= 0:
            raise abi.DemError(abi.DEM_EDEVICE, "no HIP device visible")
        sp = abi.DemServerParams()
        sp.slot_capacity, sp.cell_capacity, sp.bond_capacity = int(slot_capacity), int(cell_capacity), int(bond_capacity)
        sp.record_capacity = int(record_capacity)
        sp.scene_capacity, sp.material_capacity = int(scene_capacity), 0
        sp.block_size, sp.record_data, sp.substeps = int(block_size), int(record_data), int(substeps)
        sp.max_resident, sp.bypass, sp.contact_slots = int(max_resident), int(bypass), int(contact_slots)
        sp.graph = int(graph)
        self.params = sp
        self.block_size = int(block_size)
        self.record_data = bool(record_data)
        h = C.c_void_p()
        abi.check(self.lib.dem_server_create(C.byref(sp), C.byref(h)))
        self.h = h
        self._shape = {}
        self.failures = {}

    def submit(self, settings: Settings, state: State, nsteps: int, grid: Optional[abi.DemGrid] = None,
               use_f64: Optional[bool] = None, reorder_interval: int = 64) -> int:
        p, _keep = scene_params(settings, state, grid, use_f64, self.record_data, reorder_interval, self.block_size)
        v = state.view(state.p_count, with_data=self.record_data)
        t = C.c_int64()
        abi.check(self.lib.dem_server_submit(self.h, C.byref(p), C.byref(v), int(nsteps), C.byref(t)))
        self._shape[t.value] = (state.p_count, state.n_bonds, state.max_contacts)
        return t.value

    def tick(self) -> int:
        done = C.c_int32()
        abi.check(self.lib.dem_server_tick(self.h, C.byref(done)))
        return done.value

    def status(self, ticket: int) -> int:
        s = C.c_int32()
        abi.check(self.lib.dem_server_status(self.h, int(ticket), C.byref(s)))
        return s.value

    def take(self, ticket: int, state: State):
        """Copies the finished scene into `state` (same particle count); returns
        (state, latency in seconds from submit to completion)."""
        v = state.view(state.p_count, with_data=self.record_data)
        lat = C.c_double()
        abi.check(self.lib.dem_server_take(self.h, int(ticket), C.byref(v), C.byref(lat)))
        self._shape.pop(ticket, None)
        return state, lat.value

    def stats(self) -> dict:
        out = (C.c_double * 10)()
        abi.check(self.lib.dem_server_stats(self.h, out, 10))
        keys = ("completed", "resident", "queued", "p50_s", "p99_s", "scenes_per_s", "particle_steps", "ticks",
                "step_launches", "failed")
        return {k: out[i] for i, k in enumerate(keys)}

    def run_stream(self, requests, poll: float = 0.0):
        """Replays a request stream in real time: `requests` is a list of
        (arrival offset s, settings, state, nsteps[, grid]); each is submitted once its
        arrival time has passed, ticks run back to back (new arrivals join at the next
        tick), finished scenes are taken into their State.  Returns (tickets,
        latencies, wall s)."""
        order = sorted(range(len(requests)), key=lambda i: requests[i][0])
        tickets = [None] * len(requests)
        lat = [None] * len(requests)
        pending = {}
        t0 = time.perf_counter()
        k = 0
        while k < len(order) or pending:
            now = time.perf_counter() - t0
            while k < len(order) and requests[order[k]][0] <= now:
                i = order[k]
                s, st, n = requests[i][1:4]
                grid = requests[i][4] if len(requests[i]) > 4 else None
                tickets[i] = self.submit(s, st, n, grid=grid)
                pending[tickets[i]] = i
                k += 1
            if not pending:
                time.sleep(max(0.0, min(requests[order[k]][0] - now, 0.01)) if poll == 0.0 else poll)
                continue
            self.tick()
            for t in list(pending):
                st_t = self.status(t)
                if st_t == abi.DEM_REQ_DONE:
                    i = pending.pop(t)
                    _, lat[i] = self.take(t, requests[i][2])
                elif st_t == abi.DEM_REQ_FAILED:
                    i = pending.pop(t)
                    try:
                        self.take(t, requests[i][2])
                    except abi.DemError as e:
                        self.failures[t] = e
        return tickets, lat, time.perf_counter() - t0

    def close(self):
        if getattr(self, "h", None):
            self.lib.dem_server_destroy(self.h)
            self.h = None

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
