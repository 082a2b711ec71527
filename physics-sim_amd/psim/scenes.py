"""Seeded scene generators for the benchmark configurations.

The reference builds scenes with `setup::grid` (src/setup.rs:24-50: a square or hex
lattice, radii and velocities from the unseeded `rand::thread_rng`).  These
generators build the same kind of lattice scenes from a seeded PCG64 stream
(numpy.random.default_rng(seed)), so every run of a configuration starts from the
same state.

  granular_box     4,096 discs, r ~ U[0.5, 1.0], jittered 64 x 64 lattice of
                   spacing 2.0 filling a 128 x 128 walled box, gravity -9.81,
                   effective kn = 1e5 (material kn = ks = 2e5: kn_eff is the series
                   combination 1/(1/kn_a + 1/kn_b), 2D_Simulation.wgsl:345-346),
                   mu = 0.5, density 1, dt = 1e-4, contact restitution 0.5 (the
                   dashpot extension; 0 gives the reference's undamped law).
  column_collapse  polydisperse column collapsing into a wide box; the full size is
                   4,194,304 discs in a 4096 x 2048 box.  Radii are log-normal
                   (median 0.35, sigma 0.25, clipped to [0.175, 0.525]): the stated
                   radii (median 1.0, clip [0.5, 1.5]) scaled by 0.35, because at
                   full size 4.19M discs of median radius 1 cover 1.75 times the
                   box.  A packed pile 2048 mean diameters wide, deposited disc by
                   disc (each resting on two discs below it), centred on the floor.
  bonded_chains    horizontal chains of r = 0.5 discs at spacing 1.0, neighbours
                   bonded (parallel bonds), ends pinned, hit by a dense r = 8 disc.
  bonded_assembly  dense packing of r ~ U[0.8, 1.2] discs bonded where the gap is
                   below 0.08 (0.1 r_min), top and bottom rows pulled apart.
"""
from __future__ import annotations

import math

import numpy as np

from .settings import Settings, f32
from .state import State


def _base_settings(n, width, height, dt, rmax, rmin, kn=2e5, ks=2e5, density=1.0, mu=0.5) -> Settings:
    s = Settings()
    s.setup.particles = n
    s.setup.max_radius = f32(rmax)
    s.setup.min_radius = f32(rmin)
    s.setup.variable_rad = rmin != rmax
    s.simulation.hor_bound = f32(width)
    s.simulation.vert_bound = f32(height)
    s.simulation.timestep = f32(dt)
    s.simulation.gen_per_frame = 1
    s.physics.gravity = True
    s.physics.planet_mode = False
    s.physics.gravity_acceleration = 1.0
    s.physics.collisions = True
    s.physics.collision_interval = 1
    s.physics.friction_coefficient = f32(mu)
    s.physics.bonds = 0
    s.materials = [1.0, 1.0, 1.0, float(density), float(kn), float(ks)]
    return s


def _new_state(n, max_contacts=14) -> State:
    st = State(max(n, 1), max_contacts)
    st.p_count = n
    return st


def granular_box(n_side: int = 64, seed: int = 0, spacing: float = 2.0, rmin: float = 0.5, rmax: float = 1.0,
                 dt: float = 1e-4, mu: float = 0.5, restitution: float = 0.5):
    """Plain granular box settle (4,096 discs at n_side = 64); contacts with the
    spring-dashpot law at `restitution` (0: the reference's undamped law)."""
    rng = np.random.default_rng(seed)
    n = n_side * n_side
    box = n_side * spacing
    s = _base_settings(n, box, box, dt, rmax, rmin, mu=mu)
    s.physics.restitution = float(restitution)
    st = _new_state(n)
    r = rng.uniform(rmin, rmax, n).astype(np.float32)
    ii, jj = np.meshgrid(np.arange(n_side), np.arange(n_side), indexing="xy")
    cx = (-box / 2 + spacing / 2 + ii.ravel() * spacing).astype(np.float64)
    cy = (-box / 2 + spacing / 2 + jj.ravel() * spacing).astype(np.float64)
    room = (spacing / 2 - r).astype(np.float64)
    jx = rng.uniform(-1.0, 1.0, n) * room
    jy = rng.uniform(-1.0, 1.0, n) * room
    st.pos[0::2][:n] = (cx + jx).astype(np.float32)
    st.pos[1::2][:n] = (cy + jy).astype(np.float32)
    st.radii[:n] = r
    return s, st


_PILE = None


def _pile_lib():
    global _PILE
    if _PILE is None:
        import ctypes as C
        import os
        path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "host", "libpsim_pile.so")
        if not os.path.exists(path):
            raise RuntimeError("libpsim_pile.so not built (%s); run __graft_entry__.build()" % path)
        L = C.CDLL(path)
        D = C.POINTER(C.c_double)
        L.psim_deposit_pile.argtypes = [C.c_int32, C.POINTER(C.c_float), C.c_double, C.c_double, C.c_double,
                                        C.c_double, D, C.c_uint64, D, D]
        _PILE = L
    return _PILE


def deposit_pile(radii: np.ndarray, left: float, width: float, floor: float, overlap: float = 5e-4,
                 seed: int = 0, load: float = 0.0):
    """A packed pile (drop-and-roll deposition, host/pile.cpp): disc after disc, each
    dropped on a line that sweeps the column [left, left + width] in layers and placed
    at the lowest spot near that line where it rests on two discs (or a disc and the
    floor), overlapping them by max(overlap (r_i + r_j), d_i) and nothing more.
    With load = 0, d_i = 0.  Otherwise the pile is deposited twice: the first pass gives
    each disc's depth z below the pile's surface, and in the second a disc is pressed
    into its two supports by the force that depth puts on them, d_i = load * z * r_i
    (load = 1.155 phi rho g / kn for a bulk fraction phi under gravity g; 1.155 =
    1 / cos 30 deg for two supports), so that the pile starts near its static
    compression.  Returns f64 (x, y)."""
    import ctypes as C
    r = np.ascontiguousarray(radii, dtype=np.float32)
    n = len(r)
    D = C.POINTER(C.c_double)

    def run(delta):
        x = np.empty(n, dtype=np.float64)
        y = np.empty(n, dtype=np.float64)
        rc = _pile_lib().psim_deposit_pile(n, r.ctypes.data_as(C.POINTER(C.c_float)), left, width, floor, overlap,
                                           None if delta is None else delta.ctypes.data_as(D), seed,
                                           x.ctypes.data_as(D), y.ctypes.data_as(D))
        if rc != 0:
            raise ValueError("deposit_pile: bad arguments")
        return x, y
    x, y = run(None)
    if load > 0.0 and n:
        # surface height over bins of four diameters, then each disc's depth below it
        bw = 8.0 * float(r.max())
        b = np.minimum(((x - left) / bw).astype(np.int64), int(width / bw))
        top = np.full(int(width / bw) + 1, -np.inf)
        np.maximum.at(top, b, y + r)
        depth = np.maximum(top[b] - y, 0.0)
        x, y = run(np.ascontiguousarray(load * depth * r, dtype=np.float64))
    return x, y


def column_collapse(n: int = 4194304, per_row: int = 2048, seed: int = 2, width: float = 4096.0,
                    height: float = 2048.0, dt: float = 1e-4, mu: float = 0.5, overlap: float = 5e-4,
                    restitution: float = 0.5):
    """Large polydisperse granular collapse (4,194,304 discs at the default size): a
    packed column as wide as `per_row` mean diameters, deposited on the floor of the
    box and centred in it (deposit_pile: every disc resting on two others, pressed into
    them by the weight above it, so that the pile starts in contact near its static
    compression); its flanks are vertical, so it collapses sideways under gravity.  Radii log-normal,
    median 0.35, sigma 0.25, clipped to [0.175, 0.525]: the stated radii scaled by 0.35,
    since 4.19M discs of median radius 1 cover 1.75 times the 4096 x 2048 box.  The
    contact law of the granular box: spring-dashpot at `restitution`."""
    rng = np.random.default_rng(seed)
    rmed, sigma, lo, hi = 0.35, 0.25, 0.175, 0.525
    r = np.clip(rmed * np.exp(sigma * rng.standard_normal(n)), lo, hi).astype(np.float32)
    w = 2.0 * float(np.mean(r)) * per_row
    if w + 2 * hi > width:
        raise ValueError("column wider than the box: %.1f" % w)
    x, y = deposit_pile(r, -w / 2.0, w, -height / 2.0, overlap, seed)
    if y.max() + hi > height / 2:
        raise ValueError("column taller than the box: top %.1f" % (y.max() + hi))
    s = _base_settings(n, width, height, dt, hi, lo, mu=mu)
    s.physics.restitution = float(restitution)
    st = _new_state(n)
    st.pos[0::2][:n] = x.astype(np.float32)
    st.pos[1::2][:n] = y.astype(np.float32)
    st.radii[:n] = r
    return s, st


def weibull(rng, k, lam, n):
    return (lam * rng.weibull(k, n)).astype(np.float32)


def bonded_chains(chains: int = 64, per_chain: int = 256, seed: int = 1, striker_r: float = 8.0,
                  striker_v: float = 40.0, striker_density: float = 10.0, dt: float = 5e-5,
                  kn: float = 1e5, ks: float = 1e5, strength_k: float = 6.0, strength_scale: float = 50.0):
    """Bonded chains under impact.  Returns (settings, state, strengths); bonds are
    built on the device with ComputeProg.regen_bonds (factor 1.05 of r_a + r_b)."""
    rng = np.random.default_rng(seed)
    r = 0.5
    gap = 2.0
    n = chains * per_chain + 1
    width = per_chain * 1.0 + 40.0
    height = chains * gap + 4 * striker_r + 40.0
    s = _base_settings(n, width, height, dt, striker_r, r, kn=kn, ks=ks)
    s.physics.bonds = 3
    s.physics.bond_tearing = True
    s.physics.bond_normal_stiffness = f32(kn)
    s.physics.bond_shear_stiffness = f32(ks)
    s.physics.gravity = False
    s.materials = [1.0, 1.0, 1.0, 1.0, float(2 * kn), float(2 * ks),
                   1.0, 0.0, 0.0, float(striker_density), float(2 * kn), float(2 * ks)]
    st = _new_state(n)
    k = np.arange(chains * per_chain)
    c, i = k // per_chain, k % per_chain
    x = -per_chain / 2 + 0.5 + i * 1.0
    y = -chains * gap / 2 + c * gap
    st.pos[0:2 * (n - 1):2] = x.astype(np.float32)
    st.pos[1:2 * (n - 1):2] = y.astype(np.float32)
    st.radii[: n - 1] = r
    ends = (i == 0) | (i == per_chain - 1)
    fx = st.fixity.reshape(-1, 6)
    fx[: n - 1][ends, 0:3] = 1
    fx[: n - 1][ends, 3:6] = 1
    # striker: below the chains' middle, moving up
    st.pos[2 * (n - 1)] = 0.0
    st.pos[2 * (n - 1) + 1] = np.float32(-chains * gap / 2 - striker_r - 1.0)
    st.radii[n - 1] = striker_r
    st.vel[2 * (n - 1) + 1] = striker_v
    st.material_pointers[n - 1] = 1
    nb_est = chains * (per_chain - 1)
    strengths = np.empty(2 * nb_est, dtype=np.float32)
    strengths[0::2] = weibull(rng, strength_k, strength_scale, nb_est)
    strengths[1::2] = weibull(rng, strength_k, strength_scale, nb_est)
    return s, st, strengths


def chain_bonds(st: State, chains: int, per_chain: int, strengths: np.ndarray = None):
    """Bonds between consecutive discs of every chain (the pairs State::regen_bonds
    finds for chains of touching discs: |d| = 1.0 < 1.05 (r_a + r_b)); fills
    st.bonds, st.contacts and st.bond_strengths."""
    from .compute import assemble_bonds
    k = np.arange(chains * per_chain).reshape(chains, per_chain)
    pairs = np.stack([k[:, :-1].ravel(), k[:, 1:].ravel()], axis=1).astype(np.int32)
    st.bonds, st.contacts = assemble_bonds(st, pairs, st.max_contacts)
    if strengths is not None:
        st.bond_strengths = np.ascontiguousarray(strengths[: 2 * len(pairs)], dtype=np.float32)
    return st


def serving_stream(n_scenes: int = 96, rate: float = 50.0, seed: int = 4, n_min: int = 1024, n_max: int = 16384,
                   steps_min: int = 200, steps_max: int = 1000):
    """A stream of independent scenes in the style of the serving configuration:
    Poisson arrivals at `rate` scenes/s; half granular boxes (plain contacts, f32),
    half bonded chains struck by a small dense disc (parallel bonds, Weibull(k = 6)
    strengths); particle counts log-uniform in [n_min, n_max]; per-scene seeded
    friction mu ~ U[0.2, 0.8] and bond strength scale ~ U[20, 80]; step counts
    uniform in [steps_min, steps_max].  Scene i uses seed `seed * 100003 + i`.
    Returns a list of (arrival offset s, settings, state, steps)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n_scenes))
    t -= t[0]
    out = []
    for i in range(n_scenes):
        r = np.random.default_rng(seed * 100003 + i)
        n = int(round(math.exp(r.uniform(math.log(n_min), math.log(n_max)))))
        mu = float(r.uniform(0.2, 0.8))
        steps = int(r.integers(steps_min, steps_max + 1))
        if i % 2 == 0:
            side = max(4, int(round(math.sqrt(n))))
            s, st = granular_box(n_side=side, seed=int(r.integers(1 << 30)), spacing=2.0, mu=mu)
        else:
            per_chain = 64
            chains = max(1, n // per_chain)
            scale = float(r.uniform(20.0, 80.0))
            s, st, strengths = bonded_chains(chains=chains, per_chain=per_chain, seed=int(r.integers(1 << 30)),
                                             striker_r=1.0, striker_v=40.0, striker_density=50.0,
                                             strength_scale=scale)
            s.physics.friction_coefficient = f32(mu)
            chain_bonds(st, chains, per_chain, strengths)
        out.append((float(t[i]), s, st, steps))
    return out


def bonded_assembly(n_side: int = 32, seed: int = 3, spacing: float = 2.0, rmin: float = 0.95, rmax: float = 1.05,
                    pull_v: float = 0.5, dt: float = 1e-4, kn: float = 1e5, ks: float = 1e5, mu: float = 0.5):
    """Bonded-particle assembly for a tensile test: an n_side x n_side lattice of discs
    whose neighbours nearly touch (r ~ U[rmin, rmax] at spacing 2.0, so neighbour gaps
    lie within +-0.1), parallel bonds (mode 3, tearing on), no gravity; the bottom row
    and the top row are grips: their y velocity is fixed (fixity y = 1) at -pull_v and
    +pull_v.  Bonds come from the bond builder (State::regen_bonds rule, factor 1.05)
    with per-bond strengths drawn afterwards (Weibull, see `assembly_strengths`).
    Returns (settings, state, bottom-grip ids, top-grip ids)."""
    rng = np.random.default_rng(seed)
    n = n_side * n_side
    box = n_side * spacing * 4.0
    s = _base_settings(n, box, box, dt, rmax, rmin, kn=2 * kn, ks=2 * ks, mu=mu)
    s.physics.gravity = False
    s.physics.bonds = 3
    s.physics.bond_tearing = True
    s.physics.bond_normal_stiffness = f32(kn)
    s.physics.bond_shear_stiffness = f32(ks)
    st = _new_state(n)
    ii, jj = np.meshgrid(np.arange(n_side), np.arange(n_side), indexing="xy")
    st.pos[0:2 * n:2] = (-n_side * spacing / 2 + spacing / 2 + ii.ravel() * spacing).astype(np.float32)
    st.pos[1:2 * n:2] = (-n_side * spacing / 2 + spacing / 2 + jj.ravel() * spacing).astype(np.float32)
    st.radii[:n] = rng.uniform(rmin, rmax, n).astype(np.float32)
    bottom = np.nonzero(jj.ravel() == 0)[0]
    top = np.nonzero(jj.ravel() == n_side - 1)[0]
    fx = st.fixity.reshape(-1, 6)
    for rows, v in ((bottom, -pull_v), (top, pull_v)):
        fx[rows, 1] = 1
        fx[rows, 4] = 1
        st.vel[2 * rows + 1] = np.float32(v)
    return s, st, bottom, top


def assembly_strengths(n_bonds: int, seed: int = 3, k: float = 4.0, scale: float = 30.0):
    """Per-bond {normal, shear} strengths, Weibull(k, scale) each."""
    rng = np.random.default_rng(seed + 7919)
    out = np.empty(2 * n_bonds, dtype=np.float32)
    out[0::2] = weibull(rng, k, scale, n_bonds)
    out[1::2] = weibull(rng, k, scale, n_bonds)
    return out
