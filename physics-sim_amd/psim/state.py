"""Host mirror of the reference's `State` (src/state.rs:69-93).

Arrays are kept in the reference layouts and at capacity length (PREALLOC times
the particle count for states built here, the stored length for loaded ones);
`p_count` is the active count.  `load`/`save` go through the flatbuffer codec
(state.rs:369-493, 583-706); `get_datum` is state.rs:708-737.
"""
from __future__ import annotations

import copy
from typing import Optional

import numpy as np

from . import abi, flatbuf
from .settings import Settings

PREALLOC = 2          # state.rs:95
DATA_SIZE = 7         # state.rs:67


def _f(n):
    return np.zeros(n, dtype=np.float32)


def _i(n, v=0):
    return np.full(n, v, dtype=np.int32)


class State:
    def __init__(self, capacity: int, max_contacts: int = 14):
        c = int(capacity)
        self.p_count = 0
        self.max_contacts = max_contacts
        self.pos = _f(2 * c)
        self.vel = _f(2 * c)
        self.acc = _f(2 * c)
        self.rot = _f(c)
        self.rot_vel = _f(c)
        self.rot_acc = _f(c)
        self.forces = _f(6 * c)
        self.radii = _f(c)
        self.fixity = _i(6 * c)
        self.bonds = _i(1, -1)
        self.bond_strengths: Optional[np.ndarray] = None
        self.material_pointers = _i(c)
        self.selections = _i(c)
        self.groups = _i(c)
        self.contacts = _i(6 * max_contacts * c, -1).view(np.float32)
        self.data = _f(DATA_SIZE * c)
        self.scripts = ""
        self.settings_raw: Optional[bytes] = None

    @property
    def capacity(self) -> int:
        return len(self.radii)

    @property
    def n_bonds(self) -> int:
        if len(self.bonds) < 3:
            return 0
        return len(self.bonds) // 3

    def copy(self) -> "State":
        return copy.deepcopy(self)

    # ---- flatbuffer load/save -------------------------------------------------
    @classmethod
    def load(cls, buf: bytes, settings: Settings, init: bool = True) -> "State":
        """State::load (state.rs:583-706): arrays at their stored length, p_count from
        the stored settings; min/max radius recomputed over the whole radii vector."""
        saved = flatbuf.read_state(buf)
        if init:
            settings.from_saved(saved)
        n = saved.settings()["setup"]["particles"]
        a = saved.arrays
        cap = len(a["radii"])
        st = cls.__new__(cls)
        st.max_contacts = settings.setup.max_contacts
        st.p_count = n
        st.pos = a["pos"].astype(np.float32)
        st.vel = a["vel"].astype(np.float32)
        st.acc = a["acc"].astype(np.float32)
        st.rot = a["rot"].astype(np.float32)
        st.rot_vel = a["rot_vel"].astype(np.float32)
        st.rot_acc = a["rot_acc"].astype(np.float32)
        st.forces = a["forces"].astype(np.float32)
        st.radii = a["radii"].astype(np.float32)
        st.fixity = a["fixity"].astype(np.int32)
        st.bonds = a["bonds"].astype(np.int32)
        st.bond_strengths = None
        st.contacts = a["contacts"].astype(np.float32)
        st.material_pointers = a["material_pointers"].astype(np.int32)
        st.groups = a["groups"].astype(np.int32) if "groups" in a else _i(cap)
        st.selections = _i(max(cap, n))
        st.data = _f(DATA_SIZE * max(cap, n))
        st.scripts = saved.scripts or ""
        st.settings_raw = saved.settings_raw
        r = st.radii
        settings.setup.min_radius = float(r.min()) if len(r) else 0.0
        settings.setup.max_radius = float(r.max()) if len(r) else 0.0
        settings.setup.variable_rad = settings.setup.min_radius != settings.setup.max_radius
        return st

    @classmethod
    def load_file(cls, path, settings: Settings, init: bool = True) -> "State":
        with open(path, "rb") as f:
            return cls.load(f.read(), settings, init)

    def save(self, settings: Settings, scripts: Optional[str] = None) -> bytes:
        """State::save (state.rs:369-493)."""
        raw = settings.to_settings_struct(base=self.settings_raw)
        arrays = dict(pos=self.pos, vel=self.vel, acc=self.acc, rot=self.rot, rot_vel=self.rot_vel,
                      rot_acc=self.rot_acc, forces=self.forces, radii=self.radii, fixity=self.fixity,
                      bonds=self.bonds, contacts=self.contacts, material_pointers=self.material_pointers,
                      materials=np.asarray(settings.materials, dtype=np.float32), groups=self.groups)
        return flatbuf.write_state(arrays, raw, self.scripts if scripts is None else scripts)

    # ---- observables ----------------------------------------------------------
    def get_datum(self) -> Optional[np.ndarray]:
        """State::get_datum (state.rs:708-737): means over selected particles, except
        index 6 (Σ|data[0]|) which is a sum."""
        sel = np.nonzero(self.selections[: self.capacity] != 0)[0]
        if len(sel) == 0:
            return None
        d = self.data.reshape(-1, DATA_SIZE).astype(np.float64)
        p = self.pos.reshape(-1, 2).astype(np.float64)
        sums = np.zeros(10)
        for i in sel:
            sums += [p[i, 0], p[i, 1], d[i, 4], d[i, 5], float(self.rot[i]), d[i, 6], abs(d[i, 0]),
                     d[i, 1], d[i, 2], d[i, 3]]
        out = sums / len(sel)
        out[6] = sums[6]
        return out

    # ---- C ABI view -----------------------------------------------------------
    def view(self, n: Optional[int] = None, with_data: bool = True) -> abi.DemStateView:
        n = self.p_count if n is None else n
        for name in ("pos", "vel", "acc", "rot", "rot_vel", "rot_acc", "forces", "radii", "data", "contacts"):
            arr = getattr(self, name)
            if arr.dtype != np.float32 or not arr.flags.c_contiguous:
                setattr(self, name, np.ascontiguousarray(arr, dtype=np.float32))
        for name in ("fixity", "material_pointers", "bonds"):
            arr = getattr(self, name)
            if arr.dtype != np.int32 or not arr.flags.c_contiguous:
                setattr(self, name, np.ascontiguousarray(arr, dtype=np.int32))
        v = abi.DemStateView()
        v.n = n
        v.pos, v.vel, v.acc = abi.fptr(self.pos), abi.fptr(self.vel), abi.fptr(self.acc)
        v.rot, v.rot_vel, v.rot_acc = abi.fptr(self.rot), abi.fptr(self.rot_vel), abi.fptr(self.rot_acc)
        v.forces, v.radii = abi.fptr(self.forces), abi.fptr(self.radii)
        v.fixity, v.material_pointers = abi.iptr(self.fixity), abi.iptr(self.material_pointers)
        v.bonds = abi.iptr(self.bonds)
        v.n_bonds = self.n_bonds
        if self.bond_strengths is not None:
            self.bond_strengths = np.ascontiguousarray(self.bond_strengths, dtype=np.float32)
            v.bond_strengths = abi.fptr(self.bond_strengths)
        v.contacts = abi.fptr(self.contacts)
        v.contact_slots = self.max_contacts
        v.data = abi.fptr(self.data) if with_data else None
        return v
