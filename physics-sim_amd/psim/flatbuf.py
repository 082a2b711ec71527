"""Saved-state codec for the reference's flatbuffer `State` table.

Schema: new_schema.fbs:1-105 (root table `State`, 16 fields, inline 212-byte
`Settings` struct) and the older schema.fbs (15 fields, 204-byte `Settings`, no
`scripts`, no `use_f64`, no `bond_tearing`).  Field slots and struct offsets are the
ones of the generated code (new_schema_generated.rs:2500-2515 and the struct
accessors at :15, :1003, :1383, :1795, :2367).

`write_state` lays bytes out the way the Rust writer does (State::save,
state.rs:369-493): vectors created in the order pos, vel, acc, rot, rot_vel,
rot_acc, forces, radii, fixity, bonds, contacts, material_pointers, materials,
groups, then the scripts string; the table adds `settings` first and the other
fields in reverse schema order (new_schema_generated.rs:2522-2542); the buffer is
finished without a file identifier.  The builder below follows the published
FlatBuffers back-to-front building algorithm (alignment padding before each
object, vtable written after the table and de-duplicated against earlier ones).
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field
from typing import Dict, Optional

import numpy as np

# Table fields in slot order (voffset = 4 + 2*slot).
FIELDS = ["pos", "vel", "acc", "rot", "rot_vel", "rot_acc", "forces", "radii", "fixity",
          "bonds", "contacts", "material_pointers", "materials", "groups", "scripts", "settings"]
INT_FIELDS = {"fixity", "bonds", "material_pointers", "groups"}

SETTINGS_SIZE_NEW = 212
SETTINGS_SIZE_OLD = 204

# (name, offset, format) of the nested structs, new schema (new_schema.fbs:1-84).
VIEW_FIELDS = [
    ("circular_particles", 0, "?"), ("render_rot", 1, "?"), ("render_bonds", 2, "?"),
    ("render_outline", 3, "?"), ("render_bp_grid", 4, "?"), ("color_code_rot", 5, "?"),
    ("use_particle_color_outline", 6, "?"), ("outline_color_r", 8, "f"), ("outline_color_g", 12, "f"),
    ("outline_color_b", 16, "f"), ("background_color_r", 20, "f"), ("background_color_g", 24, "f"),
    ("background_color_b", 28, "f"), ("color_source", 32, "i"), ("dim_slow_particles", 36, "?"),
    ("max_brightness_vel", 40, "f"), ("crt_res", 44, "i"), ("grain", 48, "?"),
    ("grain_strength", 52, "f"), ("grain_size", 56, "i"), ("sobel", 60, "?"), ("colored_sobel", 61, "?"),
    ("invert", 62, "?"), ("chrom_ab", 63, "?"), ("abb_strength", 64, "f"),
    ("bond_highlight_strength", 68, "f"), ("render_unbonded_contacts", 72, "?"), ("lighting", 73, "?"),
    ("show_hit_tex", 74, "?"),
]
SETUP_FIELDS = [
    ("particles", 0, "i"), ("max_radius", 4, "f"), ("variable_rad", 8, "?"), ("holeyness", 12, "f"),
    ("max_h_velocity", 16, "f"), ("min_h_velocity", 20, "f"), ("max_v_velocity", 24, "f"),
    ("min_v_velocity", 28, "f"), ("grid_width", 32, "f"), ("hex_grid", 36, "?"),
]
SIM_FIELDS = [
    ("timestep", 0, "f"), ("gen_per_frame", 4, "i"), ("auto_width", 8, "?"), ("width", 12, "f"),
    ("height", 16, "f"), ("maintain_ar", 20, "?"), ("round_walls", 21, "?"), ("wall_radius", 24, "f"),
    ("d3", 28, "?"), ("x_timesteps", 32, "i"), ("use_f64", 36, "?"),
]
PHYS_FIELDS = [
    ("gravity", 0, "?"), ("g_force", 4, "f"), ("planet_mode", 8, "?"), ("mouse_gravity", 9, "?"),
    ("collisions", 10, "?"), ("collision_interval", 12, "i"), ("friction_coef", 16, "f"), ("bond", 20, "i"),
    ("bond_tearing", 24, "?"), ("bond_normal_stiffness", 28, "f"), ("bond_shear_stiffness", 32, "f"),
    ("bond_normal_strength", 36, "f"), ("bond_shear_strength", 40, "f"),
    ("moment_contribution_factor", 44, "f"), ("local_damping", 48, "?"), ("local_damping_alpha", 52, "f"),
]
# (struct, offset inside Settings, size)
SETTINGS_PARTS_NEW = [("view", 0, 76, VIEW_FIELDS), ("setup", 76, 40, SETUP_FIELDS),
                      ("sim", 116, 40, SIM_FIELDS), ("physics", 156, 56, PHYS_FIELDS)]


class FlatbufferError(ValueError):
    pass


@dataclass
class SavedState:
    """Decoded `State` table.  Vectors are numpy arrays at their stored length
    (capacity, PREALLOC * N in files written by the app)."""
    arrays: Dict[str, np.ndarray] = field(default_factory=dict)
    scripts: Optional[str] = None
    settings_raw: bytes = b""
    schema: str = "new"

    def settings(self) -> Dict[str, Dict[str, object]]:
        return decode_settings(self.settings_raw)


def decode_settings(raw: bytes) -> Dict[str, Dict[str, object]]:
    if len(raw) != SETTINGS_SIZE_NEW:
        raise FlatbufferError("settings struct is %d bytes, expected %d" % (len(raw), SETTINGS_SIZE_NEW))
    out: Dict[str, Dict[str, object]] = {}
    for name, off, size, fields in SETTINGS_PARTS_NEW:
        part = raw[off:off + size]
        d = {}
        for fname, foff, fmt in fields:
            if fmt == "?":
                d[fname] = part[foff] != 0
            else:
                d[fname] = struct.unpack_from("<" + fmt, part, foff)[0]
        out[name] = d
    return out


def encode_settings(values: Dict[str, Dict[str, object]], base: Optional[bytes] = None) -> bytes:
    raw = bytearray(base if base is not None else bytes(SETTINGS_SIZE_NEW))
    for name, off, size, fields in SETTINGS_PARTS_NEW:
        part = values.get(name, {})
        for fname, foff, fmt in fields:
            if fname not in part:
                continue
            v = part[fname]
            if fmt == "?":
                raw[off + foff] = 1 if v else 0
            else:
                struct.pack_into("<" + fmt, raw, off + foff, v)
    return bytes(raw)


def read_state(buf: bytes) -> SavedState:
    """Parses a `State` buffer (new schema; old-schema buffers are recognised by
    their 15-slot vtable and 204-byte settings struct)."""
    if len(buf) < 8:
        raise FlatbufferError("buffer too short")
    root = struct.unpack_from("<I", buf, 0)[0]
    if root + 4 > len(buf):
        raise FlatbufferError("root offset out of range")
    vt = root - struct.unpack_from("<i", buf, root)[0]
    if vt < 0 or vt + 4 > len(buf):
        raise FlatbufferError("vtable out of range")
    vsz, tsz = struct.unpack_from("<HH", buf, vt)
    if vsz < 4 or vsz % 2 or vt + vsz > len(buf):
        raise FlatbufferError("bad vtable size")
    nslots = (vsz - 4) // 2
    offs = struct.unpack_from("<%dH" % nslots, buf, vt + 4)
    old = nslots == 15 and tsz != 0 and offs[14] != 0 and _settings_len_guess(buf, root, offs) == SETTINGS_SIZE_OLD
    names = FIELDS if not old else FIELDS[:14] + ["settings"]
    st = SavedState(schema="old" if old else "new")
    for slot, o in enumerate(offs):
        if o == 0 or slot >= len(names):
            continue
        name = names[slot]
        p = root + o
        if name == "settings":
            size = SETTINGS_SIZE_OLD if old else SETTINGS_SIZE_NEW
            if p + size > len(buf):
                raise FlatbufferError("settings struct out of range")
            st.settings_raw = bytes(buf[p:p + size])
            continue
        vp = p + struct.unpack_from("<I", buf, p)[0]
        if vp + 4 > len(buf):
            raise FlatbufferError("vector %s out of range" % name)
        n = struct.unpack_from("<I", buf, vp)[0]
        if name == "scripts":
            if vp + 4 + n > len(buf):
                raise FlatbufferError("string out of range")
            st.scripts = bytes(buf[vp + 4:vp + 4 + n]).decode("utf-8")
            continue
        if vp + 4 + 4 * n > len(buf):
            raise FlatbufferError("vector %s out of range" % name)
        dt = "<i4" if name in INT_FIELDS else "<f4"
        st.arrays[name] = np.frombuffer(buf, dtype=dt, count=n, offset=vp + 4).copy()
    return st


def _settings_len_guess(buf, root, offs):
    # Old schema: Settings is the last slot and the table ends 204 bytes after it.
    vt = root - struct.unpack_from("<i", buf, root)[0]
    tsz = struct.unpack_from("<H", buf, vt + 2)[0]
    return tsz - offs[14]


def read_state_file(path) -> SavedState:
    with open(path, "rb") as f:
        return read_state(f.read())


class _Builder:
    """Back-to-front FlatBuffers builder (little endian), enough for one table of
    vectors, a string and an inline struct."""

    def __init__(self):
        self.buf = bytearray()   # written bytes, stored reversed-in-position: buf[-used:]
        self.min_align = 1
        self.vtables = []        # revpos of written vtables
        self.fields = []         # (voffset, revpos of value)
        self.table_start = 0

    # `used` is the distance from the end of the finished buffer.
    @property
    def used(self):
        return len(self.buf)

    def _pad(self, n):
        self.buf[0:0] = bytes(n)

    def align(self, size, alignment):
        self.min_align = max(self.min_align, alignment)
        pad = (-(self.used + size)) % alignment
        self._pad(pad)

    def _prepend(self, data: bytes):
        self.buf[0:0] = data

    def push_scalar(self, fmt, v):
        sz = struct.calcsize("<" + fmt)
        self.align(sz, sz)
        self._prepend(struct.pack("<" + fmt, v))
        return self.used

    def create_vector(self, arr: np.ndarray, elem_fmt: str):
        data = np.ascontiguousarray(arr).astype("<" + ("i4" if elem_fmt == "i" else "f4")).tobytes()
        self.align(len(data), 4)
        self._prepend(data)
        return self.push_scalar("I", len(arr))

    def create_string(self, s: str):
        b = s.encode("utf-8")
        self.align(len(b) + 1, 4)
        self._prepend(b + b"\x00")
        return self.push_scalar("I", len(b))

    def start_table(self):
        self.fields = []
        self.table_start = self.used

    def add_offset(self, voff, target_revpos):
        self.align(4, 4)
        rel = self.used + 4 - target_revpos
        self._prepend(struct.pack("<I", rel))
        self.fields.append((voff, self.used))

    def add_struct(self, voff, data: bytes, alignment=4):
        self.align(len(data), alignment)
        self._prepend(data)
        self.fields.append((voff, self.used))

    def end_table(self):
        obj = self.push_scalar("i", 0)  # placeholder soffset to the vtable
        max_voff = max((v for v, _ in self.fields), default=2)
        vt_len = max_voff + 2
        table_size = obj - self.table_start
        vt = bytearray(vt_len)
        struct.pack_into("<HH", vt, 0, vt_len, table_size)
        for voff, rev in self.fields:
            struct.pack_into("<H", vt, voff, obj - rev)
        vt = bytes(vt)
        found = None
        for rev in self.vtables:
            start = len(self.buf) - rev
            if bytes(self.buf[start:start + vt_len]) == vt:
                found = rev
                break
        if found is None:
            self._prepend(vt)
            found = self.used
            self.vtables.append(found)
        obj_pos = len(self.buf) - obj
        struct.pack_into("<i", self.buf, obj_pos, found - obj)
        return obj

    def finish(self, root_revpos):
        self.align(4, self.min_align)
        self.align(4, 4)
        rel = self.used + 4 - root_revpos
        self._prepend(struct.pack("<I", rel))
        return bytes(self.buf)


def write_state(arrays: Dict[str, np.ndarray], settings_raw: bytes, scripts: Optional[str] = "") -> bytes:
    """Serialises a `State` table with the reference writer's field order."""
    if len(settings_raw) != SETTINGS_SIZE_NEW:
        raise FlatbufferError("settings struct must be %d bytes" % SETTINGS_SIZE_NEW)
    b = _Builder()
    order = ["pos", "vel", "acc", "rot", "rot_vel", "rot_acc", "forces", "radii", "fixity",
             "bonds", "contacts", "material_pointers", "materials", "groups"]
    off = {}
    for name in order:
        if name not in arrays or arrays[name] is None:
            continue
        off[name] = b.create_vector(np.asarray(arrays[name]), "i" if name in INT_FIELDS else "f")
    if scripts is not None:
        off["scripts"] = b.create_string(scripts)
    b.start_table()
    b.add_struct(4 + 2 * FIELDS.index("settings"), settings_raw)
    for name in reversed(FIELDS[:-1]):
        if name in off:
            b.add_offset(4 + 2 * FIELDS.index(name), off[name])
    root = b.end_table()
    return b.finish(root)
