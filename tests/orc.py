"""ctypes access to the CPU oracle (oracle/liborc.so) for the tests.

The oracle is a checker: tests compare the HIP path against it, never the other
way round, and nothing in physics-sim_amd/ imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from psim import abi
from psim.scripts import apply_set_properties
from psim.state import State

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORC_PATH = os.path.join(ROOT, "oracle", "liborc.so")

MODE_REFERENCE = 0
MODE_CANONICAL = 1

FP = abi.FP
IP = abi.IP


class OrcState(C.Structure):
    _fields_ = [("n", C.c_int32), ("pos", FP), ("vel", FP), ("acc", FP), ("rot", FP), ("rot_vel", FP),
                ("rot_acc", FP), ("forces", FP), ("radii", FP), ("fixity", IP), ("material_pointers", IP),
                ("bonds", IP), ("n_bonds", C.c_int32), ("bond_strengths", FP), ("contacts", FP),
                ("slots", C.c_int32), ("data", FP), ("del_pos", FP), ("del_rot", FP), ("coll_cont", IP),
                ("grid", IP), ("grid_len", C.c_int64)]


_orc = None


def lib():
    global _orc
    if _orc is None:
        if not os.path.exists(ORC_PATH):
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        L = C.CDLL(ORC_PATH)
        L.orc_run.argtypes = [C.POINTER(abi.DemUniform), C.POINTER(abi.DemMaterial), C.c_int32,
                              C.POINTER(abi.DemGrid), C.POINTER(OrcState), C.c_int32, C.c_int32, C.c_int32,
                              C.c_float]
        L.orc_canonicalize.argtypes = [FP, C.c_int32, C.c_int32]
        L.orc_grid_capacity.argtypes = [C.c_float, C.c_float, C.c_float, C.c_float, C.c_int32, C.c_int32,
                                        C.POINTER(abi.DemGrid)]
        L.orc_cell_keys.argtypes = [C.c_int32, C.c_int32, FP, FP, C.POINTER(abi.DemGrid), IP]
        L.orc_sort_cells.argtypes = [C.c_int32, IP, C.c_int32, IP, IP]
        L.orc_contact_pairs.argtypes = [C.c_int32, FP, FP, C.c_int32, IP, IP]
        L.orc_contact_pairs.restype = C.c_int64
        L.orc_contact_pairs_grid.argtypes = [C.c_int32, FP, FP, C.c_int32, IP, IP]
        L.orc_contact_pairs_grid.restype = C.c_int64
        L.orc_regen_bonds.argtypes = [C.c_int32, FP, FP, C.c_int32, C.c_int32, IP, C.c_int64, FP, C.c_float,
                                      C.c_float]
        L.orc_regen_bonds.restype = C.c_int64
        L.orc_set_threads.argtypes = [C.c_int32]
        L.orc_kinetic_energy.argtypes = [C.c_int32, FP, FP, FP, IP, C.POINTER(abi.DemMaterial), C.c_int32]
        L.orc_kinetic_energy.restype = C.c_double
        L.orc_time_reference.argtypes = [C.POINTER(abi.DemUniform), C.POINTER(abi.DemMaterial), C.c_int32,
                                         C.POINTER(abi.DemGrid), C.POINTER(OrcState), C.c_int32, C.c_int32,
                                         C.c_float]
        L.orc_time_reference.restype = C.c_double
        D = C.POINTER(C.c_double)
        L.orc64_run.argtypes = [C.POINTER(abi.DemUniform), C.POINTER(abi.DemMaterial), C.c_int32, C.c_int32,
                                FP, FP, FP, FP, FP, FP, FP, FP, IP, IP, IP, C.c_int32, FP, FP, C.c_int32,
                                C.c_int32, C.c_float, D, D, D, D, D, D, D, IP, D]
        _orc = L
    return _orc


def run64(state: State, settings, nsteps: int) -> dict:
    """The f64 oracle ("ref64", oracle/dem_oracle64.cpp): `nsteps` steps from the f32
    `state` in double throughout; returns the final state arrays (f64), the bond index
    words and the kinetic energy after every step."""
    n = state.p_count
    u = settings.collision_settings()
    u.particles = n
    mats = abi.materials_array(settings.materials)
    out = {k: np.zeros(m * n, dtype=np.float64) for k, m in
           (("pos", 2), ("vel", 2), ("acc", 2), ("rot", 1), ("rot_vel", 1), ("rot_acc", 1), ("data", 7))}
    out["bonds"] = np.zeros(max(state.n_bonds, 1), dtype=np.int32)
    out["ke"] = np.zeros(max(nsteps, 1), dtype=np.float64)
    D = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    s = state
    rc = lib().orc64_run(C.byref(u), mats, len(mats), n, abi.fptr(s.pos), abi.fptr(s.vel), abi.fptr(s.acc),
                         abi.fptr(s.rot), abi.fptr(s.rot_vel), abi.fptr(s.rot_acc), abi.fptr(s.forces),
                         abi.fptr(s.radii), abi.iptr(s.fixity), abi.iptr(s.material_pointers), abi.iptr(s.bonds),
                         s.n_bonds, abi.fptr(s.bond_strengths) if s.bond_strengths is not None else None,
                         abi.fptr(s.contacts), s.max_contacts, int(nsteps),
                         C.c_float(settings.physics.restitution), D(out["pos"]), D(out["vel"]), D(out["acc"]),
                         D(out["rot"]), D(out["rot_vel"]), D(out["rot_acc"]), D(out["data"]),
                         out["bonds"].ctypes.data_as(IP), D(out["ke"]))
    assert rc == 0, rc
    return out


class OracleEngine:
    """Runs the oracle on a host State, keeping the device-only buffers (del_pos,
    del_rot, data, coll_cont, grid) between calls the way the reference's device
    buffers persist between compute() calls."""

    def __init__(self, state: State, settings, mode=MODE_CANONICAL, use_f64=None, slots=None, grid=None):
        self.settings = settings
        self.mode = mode
        self.use_f64 = settings.simulation.use_f64 if use_f64 is None else use_f64
        self.slots = slots or state.max_contacts
        self.st = None
        self.grid = grid if grid is not None else settings.grid_info()
        self.mats = abi.materials_array(settings.materials)
        self.coll_cont = np.zeros(4, dtype=np.int32)
        self.restore(state)

    @property
    def dt(self):
        return np.float32(self.settings.simulation.timestep)

    def restore(self, state: State):
        """WGPUComputeProg::restore (wgpu_prog.rs:1235-1262): vel is uploaded into
        del_pos and rot_vel into del_rot; data and grid are cleared; coll_cont keeps
        its words."""
        st = state.copy()
        cap = st.capacity
        if self.slots != st.max_contacts:
            c = st.contacts.view(np.int32).reshape(cap, st.max_contacts, 6)
            nc = np.full((cap, self.slots, 6), -1, dtype=np.int32)
            k = min(self.slots, st.max_contacts)
            nc[:, :k] = c[:, :k]
            st.contacts = nc.reshape(-1).view(np.float32).copy()
            st.max_contacts = self.slots
        st.data = np.zeros(7 * cap, dtype=np.float32)
        self.del_pos = st.vel.copy()
        self.del_rot = st.rot_vel.copy()
        self.st = st
        g = self.grid
        self.grid_buf = np.zeros(max(1, g.w * g.h * g.cell_cap), dtype=np.int32) if self.mode == MODE_REFERENCE \
            else np.zeros(1, dtype=np.int32)

    def snapshot(self) -> State:
        return self.st.copy()

    restore_snapshot = restore

    def _view(self):
        s = self.st
        v = OrcState()
        v.n = s.capacity
        v.pos, v.vel, v.acc = abi.fptr(s.pos), abi.fptr(s.vel), abi.fptr(s.acc)
        v.rot, v.rot_vel, v.rot_acc = abi.fptr(s.rot), abi.fptr(s.rot_vel), abi.fptr(s.rot_acc)
        v.forces, v.radii = abi.fptr(s.forces), abi.fptr(s.radii)
        v.fixity, v.material_pointers = abi.iptr(s.fixity), abi.iptr(s.material_pointers)
        v.bonds = abi.iptr(s.bonds)
        v.n_bonds = s.n_bonds
        v.bond_strengths = abi.fptr(s.bond_strengths) if s.bond_strengths is not None else None
        v.contacts = abi.fptr(s.contacts)
        v.slots = self.slots
        v.data = abi.fptr(s.data)
        v.del_pos, v.del_rot = abi.fptr(self.del_pos), abi.fptr(self.del_rot)
        v.coll_cont = abi.iptr(self.coll_cont)
        v.grid = abi.iptr(self.grid_buf)
        v.grid_len = len(self.grid_buf)
        return v

    def step(self, n=1):
        u = self.settings.collision_settings()
        u.particles = self.st.p_count
        v = self._view()
        rc = lib().orc_run(C.byref(u), self.mats, len(self.mats), C.byref(self.grid), C.byref(v), self.mode,
                           int(bool(self.use_f64)), int(n), C.c_float(self.settings.physics.restitution))
        assert rc == 0, rc

    def set_properties(self, words, selections):
        moved = apply_set_properties(self.st, words, selections, self.del_pos, self.del_rot)
        if moved:
            self.coll_cont[3] = 1

    def datum(self, selections):
        if selections is None:
            return None
        s = self.st
        sel = np.zeros(s.capacity, dtype=np.int32)
        k = min(len(selections), s.capacity)
        sel[:k] = selections[:k]
        s.selections = sel
        return s.get_datum()

    def canonical_contacts(self):
        c = self.st.contacts.copy()
        lib().orc_canonicalize(abi.fptr(c), self.st.capacity, self.slots)
        return c
