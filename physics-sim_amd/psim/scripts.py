"""Replay of the reference's embedded experiment scripts.

The saved states used for validation carry a JSON list of scripts
(new_schema.fbs:101; ScriptManager::from_json, scripts.rs:123).  The actions that
drive the physics are replayed here with the reference's semantics:

  Backup          update_state + State::save of the current state (scripts.rs:314-317)
  Restore         State::load + restore of that snapshot; data and the record are
                  cleared (scripts.rs:319-323)
  Select          replaces the selection buffer with the given list (scripts.rs:310-312)
  Set_Properties  Set_Properties.wgsl:68-90 on the selected particles; the 28
                  parameters are the shader's Input struct in order, booleans as 0/1
  Call_Script     runs another script to completion, then continues (scripts.rs:325-328)
  Record          starts/stops recording (scripts.rs:335-342)
  Advance n       Client::advance (client.rs:931-960): while recording, n times
                  {step, generation+1, record a row}, then one more step that is
                  neither counted nor recorded; otherwise n steps
  Export          hands the recorded rows to a callback (Settings::save_data columns)
  Wait            ends the current frame

Frames: the reference runs script actions inside its render loop
(client.rs:1106-1421): an Advance is performed at the start of the next frame, and
each frame ends by recording one row while recording is on (client.rs:1420-1421).
Here every Wait ends exactly one frame.  Real waits can span several frames; the
rows such extra frames add are the only part of an exported trace this model does
not reproduce, and they never fall between two per-step rows.

Row columns (Settings::save_data, settings.rs:2014-2052): Timestamp (= f32 dt times
generation), mean x, mean y, mean v*x, mean v*y, mean rot, mean ω*, Σ|Σfn|, mean Σft,
mean ΣM, mean data3.
"""
from __future__ import annotations

import json
from typing import Callable, List, Optional

import numpy as np


def _param_value(p):
    kind = next(iter(p))
    name, value = p[kind]
    return kind, name, value


def set_properties_words(params) -> List[float]:
    """The 28-word Input uniform of Set_Properties.wgsl:1-30 from script parameters
    (Parameter::as_f32, scripts.rs:828-836: booleans and integers are bit-cast)."""
    out = []
    for p in params:
        kind, _, v = _param_value(p)
        if kind == "Float":
            out.append(("f", float(v)))
        elif kind == "Integer":
            out.append(("i", int(v)))
        elif kind == "Boolean":
            out.append(("i", 1 if v else 0))
        else:
            out.append(("f", -1.0))
    return out


def apply_set_properties(state, words, selections, del_pos=None, del_rot=None):
    """Set_Properties.wgsl:68-90 on host arrays (state in reference layout).  A
    position or rotation set also writes the displacement into del_pos / del_rot
    (when given) and returns True: the caller then raises the del_pos override
    flag (coll_cont[3] = 1) for the next LOM pass."""
    def w(k):
        return words[k][1]
    flags = [int(w(k)) if words[k][0] == "i" else 0 for k in range(14)]
    (set_x_pos, set_y_pos, set_rot, set_x_vel, set_y_vel, set_rot_vel, set_x_force, set_y_force,
     set_rot_force, set_radius, set_x_fix, set_y_fix, set_rot_fix, set_material) = flags
    x_pos, y_pos, rot, x_vel, y_vel, rot_vel, x_force, y_force, rot_force, radius = \
        [np.float32(w(k)) for k in range(14, 24)]
    x_fix, y_fix, rot_fix, material = [int(w(k)) for k in range(24, 28)]
    moved = False
    n = state.p_count
    for i in range(min(len(selections), n)):
        if selections[i] == 0:
            continue
        if set_x_pos == 1:
            if del_pos is not None:
                del_pos[2 * i] = np.float32(x_pos - state.pos[2 * i])
            state.pos[2 * i] = x_pos
            moved = True
        if set_y_pos == 1:
            if del_pos is not None:
                del_pos[2 * i + 1] = np.float32(y_pos - state.pos[2 * i + 1])
            state.pos[2 * i + 1] = y_pos
            moved = True
        if set_rot == 1:
            if del_rot is not None:
                del_rot[i] = np.float32(rot - state.rot[i])
            state.rot[i] = rot
        if set_x_vel == 1:
            state.vel[2 * i] = x_vel
        if set_y_vel == 1:
            state.vel[2 * i + 1] = y_vel
        if set_rot_vel == 1:
            state.rot_vel[i] = rot_vel
        if set_radius == 1:
            state.radii[i] = radius
        if set_x_force == 1:
            state.forces[6 * i] = x_force
        if set_y_force == 1:
            state.forces[6 * i + 1] = y_force
        if set_rot_force == 1:
            state.forces[6 * i + 2] = rot_force
        if set_x_fix == 1:
            state.fixity[6 * i] = x_fix
            state.fixity[6 * i + 3] = x_fix
        if set_y_fix == 1:
            state.fixity[6 * i + 1] = y_fix
            state.fixity[6 * i + 4] = y_fix
        if set_rot_fix == 1:
            state.fixity[6 * i + 2] = rot_fix
            state.fixity[6 * i + 5] = rot_fix
        if set_material == 1:
            state.material_pointers[i] = material
        if set_x_fix == 1 or set_y_fix == 1 or set_rot_fix == 1:
            selections[i] = 2
    return moved


class ScriptReplay:
    """Runs the first (auto-run) script of a saved state against an engine.

    The engine provides: snapshot() -> opaque, restore_snapshot(snap),
    set_properties(words, selections) -> None, step(n), datum(selections) -> array
    of 10 or None, and the attribute `dt` (f32 time step)."""

    def __init__(self, engine, scripts_json: str):
        self.engine = engine
        self.scripts = json.loads(scripts_json)
        self.generation = 0
        self.recording = False
        self.rows: List[List[float]] = []
        self.selections = None
        self.backup = None
        self.pending_advance = 0
        self.exports = []

    def _collect(self):
        d = self.engine.datum(self.selections)
        if d is None:
            self.rows = []
            return
        t = float(np.float32(np.float32(self.engine.dt) * np.float32(self.generation)))
        self.rows.append([t] + [float(x) for x in d])

    def _advance(self, ticks):
        if self.recording:
            for _ in range(ticks):
                self.engine.step(1)
                self.generation += 1
                self._collect()
            self.engine.step(1)
        else:
            self.engine.step(ticks)
            self.generation += ticks

    def run(self, on_export: Optional[Callable] = None, script_index: int = 0):
        # Actions are flattened into a stream; a Wait ends the frame.
        stack = [(script_index, 0)]
        while stack:
            si, ai = stack.pop()
            actions = self.scripts[si]["actions"]
            while ai < len(actions):
                act = actions[ai]
                ai += 1
                name, params = act["name"], act["parameters"]
                if name == "Backup":
                    self.backup = self.engine.snapshot()
                elif name == "Restore":
                    if self.backup is not None:
                        self.engine.restore_snapshot(self.backup)
                    self.rows = []
                elif name == "Select":
                    _, _, lst = _param_value(params[0])
                    self.selections = np.array(lst if isinstance(lst, list) else [lst], dtype=np.int32)
                elif name == "Set_Properties":
                    words = set_properties_words(params)
                    self.engine.set_properties(words, self.selections)
                elif name == "Call_Script":
                    _, _, k = _param_value(params[0])
                    stack.append((si, ai))
                    stack.append((int(k), 0))
                    break
                elif name == "Record":
                    _, _, v = _param_value(params[0])
                    self.recording = bool(v)
                elif name == "Advance":
                    _, _, n = _param_value(params[0])
                    self.pending_advance = int(n)
                elif name == "Wait":
                    # end of frame: record, then the next frame starts with the advance
                    if self.recording:
                        self._collect()
                    if self.pending_advance:
                        n = self.pending_advance
                        self.pending_advance = 0
                        self._advance(n)
                elif name == "Export":
                    _, _, path = _param_value(params[0])
                    self.exports.append((path, np.array(self.rows, dtype=np.float64)))
                    if on_export:
                        on_export(path, np.array(self.rows, dtype=np.float64))
                elif name == "Simulate" or name == "Select_All" or name == "Goto" or name == "None":
                    pass
        return self.exports
