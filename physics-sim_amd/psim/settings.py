"""Host mirror of the reference's `Settings` (src/settings.rs:102-254).

Only the parts the time step reads are kept: setup, simulation and physics
settings plus the material table.  `collision_settings()` packs the 29-word
physics uniform exactly as Settings::collision_settings does
(settings.rs:2086-2119) and `grid_info()` is the reference grid sizing
(settings.rs:1955-1974 = grid_capacity, wgpu_prog.rs:698-717).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List

import numpy as np

from . import abi

BOND_TYPES = {0: "Unbonded", 1: "Normal_Bonds", 2: "Linear_Contact_Bond", 3: "Parallel_Linear_Contact_Bond"}


def f32(x) -> float:
    return float(np.float32(x))


@dataclass
class SetupSettings:  # settings.rs:139-156, defaults :300-317
    particles: int = 256
    workgroup_size: int = 256
    max_radius: float = 0.025
    min_radius: float = f32(0.025 / 1.7)
    variable_rad: bool = False
    holeyness: float = 1.7
    max_bonds: int = 6
    max_contacts: int = 14
    max_h_velocity: float = 0.0
    min_h_velocity: float = 0.0
    max_v_velocity: float = 0.0
    min_v_velocity: float = 0.0
    grid_width: float = 32.0
    hex_grid: bool = False

    @property
    def workgroups(self) -> int:  # set_particles, settings.rs:445-448
        return int(math.ceil(self.particles / self.workgroup_size))


@dataclass
class SimulationSettings:  # settings.rs:158-174, defaults :319-335
    timestep: float = f32(1.0 / 12600.0)
    gen_per_frame: int = 105
    walls: bool = True
    hor_bound: float = f32(2.0 * 1.333)
    vert_bound: float = 2.0
    round_walls: bool = False
    wall_radius: float = 1.0
    use_f64: bool = False
    x_timesteps: int = 1


@dataclass
class PhysicsSettings:  # settings.rs:176-197, defaults :336-357
    gravity: bool = True
    gravity_acceleration: float = 1.0
    planet_mode: bool = True
    mouse_gravity: bool = False
    collisions: bool = True
    collision_interval: int = 1
    friction_coefficient: float = 0.5
    bonds: int = 0
    bond_tearing: bool = False
    bond_normal_stiffness: float = 10.0
    bond_shear_stiffness: float = 10.0
    bond_normal_strength: float = 0.5
    bond_shear_strength: float = 0.5
    contact_damping: float = 0.2
    bond_damping: float = 0.2
    drag: float = 1.0
    moment_contribution_factor: float = 1.0
    local_damping: bool = False
    local_damping_alpha: float = 0.1
    # Extension, not a reference setting (not in the uniform or the saved state): the
    # restitution of the contact dashpot the benchmark configurations call for
    # (dem_params.restitution); 0 = no dashpot, the reference's contact law.
    restitution: float = 0.0


DEFAULT_MATERIALS = [1.0, 1.0, 1.0, 0.01, 100.0, 50.0, 1.0, 0.0, 0.0, 0.01, 100.0, 50.0]  # settings.rs:266


@dataclass
class Settings:
    setup: SetupSettings = field(default_factory=SetupSettings)
    simulation: SimulationSettings = field(default_factory=SimulationSettings)
    physics: PhysicsSettings = field(default_factory=PhysicsSettings)
    materials: List[float] = field(default_factory=lambda: list(DEFAULT_MATERIALS))
    wall_friction: float = 0.0
    world_pos: tuple = (0.0, 0.0)

    def set_particles(self, n: int):
        self.setup.particles = int(n)

    def collision_settings(self) -> abi.DemUniform:
        """The 29-word physics uniform, settings.rs:2086-2119."""
        s, p = self.simulation, self.physics
        return abi.uniform_from_dict(dict(
            walls=int(s.walls), hor_bound=s.hor_bound, vert_bound=s.vert_bound, round_bounds=int(s.round_walls),
            bound_radius=s.wall_radius, wall_friction=self.wall_friction, gravity=int(p.gravity),
            planet_mode=int(p.planet_mode), bonds=int(p.bonds), collisions=int(p.collisions),
            friction_coefficient=p.friction_coefficient, gravity_acc=p.gravity_acceleration,
            bond_normal_stiffness=p.bond_normal_stiffness, bonds_tear=int(p.bond_tearing),
            bond_normal_strength=p.bond_normal_strength, contact_damping=p.contact_damping,
            bond_damping=p.bond_damping, drag=p.drag, bond_shear_strength=p.bond_shear_strength,
            dT=s.timestep, bond_shear_stiffness=p.bond_shear_stiffness, gravity_x=self.world_pos[0],
            gravity_y=self.world_pos[1], mouse_gravity=int(p.mouse_gravity),
            moment_contribution_factor=p.moment_contribution_factor, collision_interval=p.collision_interval,
            local_damping=int(p.local_damping), local_damping_alpha=p.local_damping_alpha,
            particles=self.setup.particles))

    def grid_info(self) -> abi.DemGrid:
        """Reference grid: cells of 2*max_radius over twice the box (settings.rs:1955-1974)."""
        width = f32(self.simulation.hor_bound * np.float32(2.0))
        height = f32(self.simulation.vert_bound * np.float32(2.0))
        max_rad = f32(np.float32(self.setup.max_radius) * np.float32(2.0))
        min_rad = self.setup.min_radius if self.setup.variable_rad else self.setup.max_radius
        w = int(math.ceil(np.float32(width) / np.float32(max_rad)))
        h = int(math.ceil(np.float32(height) / np.float32(max_rad)))
        cap = int(math.ceil(np.float32(np.float32(max_rad) / np.float32(min_rad) + np.float32(1.0)) ** 2))
        cap = min(cap, self.setup.particles) + 2
        return abi.DemGrid(max_rad, cap, w, h)

    def box_grid(self, cell_size: float = None, margin_cells: int = 1) -> abi.DemGrid:
        """A grid that covers the walled box (plus `margin_cells` on every side)
        instead of twice the box; cell_size defaults to 2*max_radius."""
        cs = f32(2.0 * self.setup.max_radius) if cell_size is None else f32(cell_size)
        w = int(math.ceil(self.simulation.hor_bound / cs)) + 2 * margin_cells
        h = int(math.ceil(self.simulation.vert_bound / cs)) + 2 * margin_cells
        w += w % 2
        h += h % 2
        return abi.DemGrid(cs, 0, w, h)

    # --- flatbuffer Settings struct (new_schema.fbs:1-84) ---
    def from_saved(self, saved) -> "Settings":
        """Applies the settings stored in a saved state (State::load, state.rs:583-706)."""
        d = saved.settings()
        st, sm, ph = d["setup"], d["sim"], d["physics"]
        self.setup.particles = st["particles"]
        self.setup.max_radius = st["max_radius"]
        self.setup.variable_rad = st["variable_rad"]
        self.setup.holeyness = st["holeyness"]
        self.setup.max_h_velocity = st["max_h_velocity"]
        self.setup.min_h_velocity = st["min_h_velocity"]
        self.setup.max_v_velocity = st["max_v_velocity"]
        self.setup.min_v_velocity = st["min_v_velocity"]
        self.setup.grid_width = st["grid_width"]
        self.setup.hex_grid = st["hex_grid"]
        self.simulation.timestep = sm["timestep"]
        self.simulation.gen_per_frame = sm["gen_per_frame"]
        self.simulation.hor_bound = sm["width"]
        self.simulation.vert_bound = sm["height"]
        self.simulation.round_walls = sm["round_walls"]
        self.simulation.wall_radius = sm["wall_radius"]
        self.simulation.x_timesteps = sm["x_timesteps"]
        self.simulation.use_f64 = sm["use_f64"]
        self.physics.gravity = ph["gravity"]
        self.physics.gravity_acceleration = ph["g_force"]
        self.physics.planet_mode = ph["planet_mode"]
        self.physics.mouse_gravity = ph["mouse_gravity"]
        self.physics.collisions = ph["collisions"]
        self.physics.collision_interval = ph["collision_interval"]
        self.physics.friction_coefficient = ph["friction_coef"]
        b = ph["bond"]
        self.physics.bonds = 0 if b <= 0 else (b if b <= 3 else 3)  # BondType::from_i32, settings.rs:2214
        self.physics.bond_tearing = ph["bond_tearing"]
        self.physics.bond_normal_stiffness = ph["bond_normal_stiffness"]
        self.physics.bond_shear_stiffness = ph["bond_shear_stiffness"]
        self.physics.bond_normal_strength = ph["bond_normal_strength"]
        self.physics.bond_shear_strength = ph["bond_shear_strength"]
        self.physics.moment_contribution_factor = ph["moment_contribution_factor"]
        self.physics.local_damping = ph["local_damping"]
        self.physics.local_damping_alpha = ph["local_damping_alpha"]
        if "materials" in saved.arrays:
            self.materials = [float(x) for x in saved.arrays["materials"]]
        return self

    def to_settings_struct(self, base: bytes = None) -> bytes:
        from . import flatbuf
        s, p, st = self.simulation, self.physics, self.setup
        return flatbuf.encode_settings({
            "setup": dict(particles=st.particles, max_radius=st.max_radius, variable_rad=st.variable_rad,
                          holeyness=st.holeyness, max_h_velocity=st.max_h_velocity,
                          min_h_velocity=st.min_h_velocity, max_v_velocity=st.max_v_velocity,
                          min_v_velocity=st.min_v_velocity, grid_width=st.grid_width, hex_grid=st.hex_grid),
            "sim": dict(timestep=s.timestep, gen_per_frame=s.gen_per_frame, width=s.hor_bound,
                        height=s.vert_bound, round_walls=s.round_walls, wall_radius=s.wall_radius,
                        x_timesteps=s.x_timesteps, use_f64=s.use_f64),
            "physics": dict(gravity=p.gravity, g_force=p.gravity_acceleration, planet_mode=p.planet_mode,
                            mouse_gravity=p.mouse_gravity, collisions=p.collisions,
                            collision_interval=p.collision_interval, friction_coef=p.friction_coefficient,
                            bond=p.bonds, bond_tearing=p.bond_tearing,
                            bond_normal_stiffness=p.bond_normal_stiffness,
                            bond_shear_stiffness=p.bond_shear_stiffness,
                            bond_normal_strength=p.bond_normal_strength,
                            bond_shear_strength=p.bond_shear_strength,
                            moment_contribution_factor=p.moment_contribution_factor,
                            local_damping=p.local_damping, local_damping_alpha=p.local_damping_alpha),
        }, base=base)
