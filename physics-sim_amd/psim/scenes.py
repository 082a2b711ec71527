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
  bonded_assembly  1,048,576 discs, r ~ U[0.8, 1.2]: a seeded triangular lattice at
                   area fraction 0.86 relaxed on the device into a dense disordered
                   packing, every pair closer than 0.08 (0.1 r_min) bonded (parallel
                   bonds, Weibull(4, 30) strengths), bottom and top rows gripped and
                   pulled apart at 1e-3 strain per 1,000 steps from the affine
                   velocity field of that rate.  Needs a GPU.
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


def serving_stream(n_scenes: int = 64, rate: float = 50.0, seed: int = 4, n_min: int = 1024, n_max: int = 262144,
                   steps_min: int = 2000, steps_max: int = 20000, striker_r: float = 8.0):
    """A stream of independent scenes drawn from the serving configuration's
    distribution: Poisson arrivals at `rate` scenes/s; half granular boxes in the style
    of the granular-box configuration (plain contacts, f32), half bonded chains in the
    style of the struck-chain configuration (chains of 256 r = 0.5 discs, parallel
    bonds, Weibull(k = 6) strengths, ends pinned, struck by an r = `striker_r` disc of
    density 10 at 40 units/s); particle counts log-uniform in [n_min, n_max]; per-scene
    seeded friction mu ~ U[0.2, 0.8], contact restitution ~ U[0.2, 0.9] and bond
    strength scale ~ U[20, 80]; step counts uniform in [steps_min, steps_max].  Scene i
    uses seed `seed * 100003 + i`.  Returns a list of (arrival offset s, settings,
    state, steps)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n_scenes))
    t -= t[0]
    out = []
    for i in range(n_scenes):
        r = np.random.default_rng(seed * 100003 + i)
        n = int(round(math.exp(r.uniform(math.log(n_min), math.log(n_max)))))
        mu = float(r.uniform(0.2, 0.8))
        steps = int(r.integers(steps_min, steps_max + 1))
        e = float(r.uniform(0.2, 0.9))
        if i % 2 == 0:
            side = max(4, int(round(math.sqrt(n))))
            s, st = granular_box(n_side=side, seed=int(r.integers(1 << 30)), spacing=2.0, mu=mu, restitution=e)
        else:
            per_chain = 256
            chains = max(1, int(round(n / per_chain)))
            scale = float(r.uniform(20.0, 80.0))
            s, st, strengths = bonded_chains(chains=chains, per_chain=per_chain, seed=int(r.integers(1 << 30)),
                                             striker_r=striker_r, striker_v=40.0, striker_density=10.0,
                                             strength_scale=scale)
            s.physics.friction_coefficient = f32(mu)
            s.physics.restitution = e
            chain_bonds(st, chains, per_chain, strengths)
        out.append((float(t[i]), s, st, steps))
    return out


def assembly_packing(rows: int = 1024, per_row: int = 1024, seed: int = 3, rmin: float = 0.8, rmax: float = 1.2,
                     phi: float = 0.86, dt: float = 1e-4):
    """The unrelaxed packing of the bonded assembly: `rows` x `per_row` discs,
    r ~ U[rmin, rmax] (seeded), on a triangular lattice whose spacing a gives the
    area fraction `phi` (a^2 = 2 pi <r^2> / (sqrt(3) phi)); at phi = 0.86, above the
    jamming fraction of this radius law (about 0.84), neighbours overlap by up to
    2 r_max - a.  The settings are those of the relaxation: walls at the lattice's
    extent, no gravity, frictionless contacts with restitution 0.05 (kn_eff 1e5).
    Row 0 is the bottom row, row rows-1 the top; particle id = row * per_row + column.
    Returns (settings, state, spacing)."""
    rng = np.random.default_rng(seed)
    n = rows * per_row
    r = rng.uniform(rmin, rmax, n).astype(np.float32)
    a = math.sqrt(2.0 * math.pi * float(np.mean(r.astype(np.float64) ** 2)) / (math.sqrt(3.0) * phi))
    width = (per_row + 0.5) * a
    height = rows * a * math.sqrt(3.0) / 2.0
    s = _base_settings(n, width, height, dt, rmax, rmin, mu=0.0)
    s.physics.gravity = False
    s.physics.restitution = 0.05
    st = _new_state(n)
    j, i = np.divmod(np.arange(n), per_row)
    st.pos[0:2 * n:2] = (-width / 2 + a * (i + 0.5 + 0.5 * (j % 2))).astype(np.float32)
    st.pos[1:2 * n:2] = (-height / 2 + a * math.sqrt(3.0) / 2.0 * (j + 0.5)).astype(np.float32)
    st.radii[:n] = r
    return s, st, a


def bonded_assembly(rows: int = 1024, per_row: int = 1024, seed: int = 3, rmin: float = 0.8, rmax: float = 1.2,
                    phi: float = 0.89, relax_steps: int = 3000, quench_every: int = 500, gap_frac: float = 0.1,
                    kn: float = 1500.0, ks: float = 1500.0, mu: float = 0.5, dt: float = 1e-4,
                    strain_per_1000: float = 1e-3, strength_k: float = 4.0, strength_scale: float = 30.0,
                    margin: float = 0.25, progress=None):
    """Bonded-particle assembly for a tensile test (1,048,576 discs at the default size).
    `progress(done, total)`, if given, is called after each quench of the relaxation.

    1. assembly_packing: a seeded triangular lattice of r ~ U[0.8, 1.2] discs at area
       fraction 0.86, relaxed on the device (ComputeProg: frictionless, no gravity,
       walls at the lattice's extent) for `relax_steps` steps, velocities zeroed every
       `quench_every` steps: a dense disordered packing.
    2. Every pair closer than gap_frac * rmin (0.08) is bonded (the device bond
       builder, rule |d| < r_a + r_b + 0.08, at most 6 bonds per lower id as in
       State::regen_bonds); a bond's rest length is the distance at bonding, so the
       assembly starts free of force.  Parallel bonds (mode 3, tearing on),
       strengths {normal, shear} ~ Weibull(strength_k, strength_scale) per bond.
       Bond stiffness kn = ks = 1500: a bond tears in tension at kn * delta = strength,
       so the median bond fails near 1 % strain of its length (30 / (1500 * 2)), the
       failure strain of the reference's own tensile fixture
       (saved_states/PFC_assembly_test.bin: strength 1e7 over stiffness 8e11 on bonds
       1.24e-3 long).  Contacts: the same linear law with material kn = ks = 2 kn
       (series stiffness kn), mu = 0.5.
    3. Grips: the bottom and top lattice rows, their y velocity prescribed (fixity y).
       They move apart at the strain rate `strain_per_1000` per 1,000 steps; every
       disc starts with the affine velocity field of that rate, v_y = rate (y - y_mid),
       so the pull starts without a shock at the grips (a grip started alone from rest
       sends a strain of v / c through the bonds: at this size v_grip is about 9 and
       the bonds' wave speed about 40, far above their failure strain).  The grip
       rows move at exactly -+v_grip.
    The test box is larger than the specimen by `margin` of its size on each side.
    Needs a GPU (steps 1-2 run through ComputeProg).  Returns (settings, state,
    bottom ids, top ids, info) with info = {bonds, spacing, L0 (grip separation),
    width, strain_rate (per unit time), v_grip}."""
    from .compute import ComputeProg
    s0, st, a = assembly_packing(rows, per_row, seed, rmin, rmax, phi, dt)
    n = st.p_count
    s0.setup.max_bonds = 6
    prog = ComputeProg(s0, st, bond_capacity=6 * n, record_data=False, reorder_interval=64)
    try:
        done = 0
        while done < relax_steps:
            k = min(quench_every, relax_steps - done)
            prog.step(k)
            done += k
            cur = prog.update_state()
            cur.vel[:] = 0.0
            cur.rot_vel[:] = 0.0
            prog.restore(cur)
            if progress is not None:
                progress(done, relax_steps)
        built = prog.regen_bonds(s0, factor=1.0, gap=gap_frac * rmin)
    finally:
        prog.close()
    st = built.copy()
    nb = st.n_bonds
    width0, height0 = float(s0.simulation.hor_bound), float(s0.simulation.vert_bound)
    s = _base_settings(n, width0 * (1 + 2 * margin), height0 * (1 + 2 * margin), dt, rmax, rmin,
                       kn=kn, ks=ks, mu=mu)
    s.physics.gravity = False
    s.physics.bonds = 3
    s.physics.bond_tearing = True
    s.physics.bond_normal_stiffness = f32(kn)
    s.physics.bond_shear_stiffness = f32(ks)
    s.setup.max_bonds = 6
    st.vel[:] = 0.0
    st.rot_vel[:] = 0.0
    st.acc[:] = 0.0
    st.rot_acc[:] = 0.0
    st.bond_strengths = assembly_strengths(nb, seed, strength_k, strength_scale)
    bottom = np.arange(per_row)
    top = np.arange((rows - 1) * per_row, rows * per_row)
    y = st.pos[1:2 * n:2].astype(np.float64)
    yb, yt = float(y[bottom].mean()), float(y[top].mean())
    rate = strain_per_1000 / (1000.0 * dt)              # strain per unit time
    ym = 0.5 * (yb + yt)
    st.vel[1:2 * n:2] = (rate * (y - ym)).astype(np.float32)
    v_grip = 0.5 * rate * (yt - yb)
    st.vel[2 * bottom + 1] = np.float32(-v_grip)
    st.vel[2 * top + 1] = np.float32(v_grip)
    fx = st.fixity.reshape(-1, 6)
    for rows_ in (bottom, top):
        fx[rows_, 1] = 1
        fx[rows_, 4] = 1
    info = {"bonds": nb, "spacing": a, "L0": yt - yb, "width": width0, "strain_rate": rate,
            "v_grip": v_grip, "relax_steps": relax_steps}
    return s, st, bottom, top, info


def grip_stress(prog, bottom, top, width: float) -> float:
    """Tensile stress on the grips from the device group sums: the grips' net force
    (Σ m a_y over each grip row, the pull of the specimen on it) averaged over the two
    grips, per unit width (discs are unit-thickness slabs)."""
    fb = prog.group_sums(bottom)["force"][1]
    ft = prog.group_sums(top)["force"][1]
    return 0.5 * (fb - ft) / width


def grip_strain(prog, bottom, top, L0: float) -> float:
    """Engineering strain between the grips: mean y of the top row minus that of the
    bottom row, over its value at the start."""
    gb, gt = prog.group_sums(bottom), prog.group_sums(top)
    return (gt["pos"][1] / gt["count"] - gb["pos"][1] / gb["count"]) / L0 - 1.0


def assembly_strengths(n_bonds: int, seed: int = 3, k: float = 4.0, scale: float = 30.0):
    """Per-bond {normal, shear} strengths, Weibull(k, scale) each."""
    rng = np.random.default_rng(seed + 7919)
    out = np.empty(2 * n_bonds, dtype=np.float32)
    out[0::2] = weibull(rng, k, scale, n_bonds)
    out[1::2] = weibull(rng, k, scale, n_bonds)
    return out
