"""dem_kinetic_energy_noalloc (the kinetic-energy observable without allocation) against
dem_kinetic_energy, through the C ABI: two calls in a row, the second on a smaller
scene, and the rejection of a scene beyond the static scratch."""
import ctypes as C

import numpy as np
import pytest

from psim import abi, scenes
from psim.compute import ComputeProg

pytestmark = pytest.mark.gpu


def _ke_pair(prog):
    lib = prog.lib
    a, b = C.c_double(), C.c_double()
    abi.check(lib.dem_kinetic_energy(prog.ctx, C.byref(a)))
    abi.check(lib.dem_kinetic_energy_noalloc(prog.ctx, C.byref(b)))
    return a.value, b.value


def test_noalloc_kinetic_energy_matches_allocating_entry_point():
    big = ComputeProg(*scenes.granular_box(n_side=64, seed=11), record_data=False)
    small = ComputeProg(*scenes.granular_box(n_side=20, seed=12), record_data=False)
    try:
        big.step(40)
        small.step(25)
        e_big, n_big = _ke_pair(big)          # first call: 4,096 discs
        e_small, n_small = _ke_pair(small)    # second call in a row: 400 discs, same scratch
        print("noalloc KE: big %.17g vs %.17g, small %.17g vs %.17g" % (e_big, n_big, e_small, n_small), flush=True)
        assert e_big > 0 and e_small > 0
        # f64 sums of the same per-id energies in two fixed orders (host loop vs device tree)
        assert abs(n_big - e_big) <= 1e-12 * e_big
        assert abs(n_small - e_small) <= 1e-12 * e_small
        # the smaller scene's sum does not see the larger scene's entries left in the scratch
        assert abs(n_small - e_big) > 1e-6 * e_big
        # and the larger scene again after the smaller one
        e2, n2 = _ke_pair(big)
        assert e2 == e_big and abs(n2 - e2) <= 1e-12 * e2
    finally:
        big.close()
        small.close()


def test_noalloc_kinetic_energy_rejects_scene_beyond_scratch():
    lib = abi.load_library()
    n = abi.DEM_KE_NOALLOC_MAX + 1
    s, st = scenes.granular_box(n_side=8, seed=1)
    p = abi.DemParams()
    p.u = s.collision_settings()
    mats = abi.materials_array(s.materials)
    p.materials = C.cast(mats, C.POINTER(abi.DemMaterial))
    p.n_materials = len(mats)
    p.grid = ComputeProg.default_grid(s, st)
    p.record_data = 0
    p.reorder_interval = -1
    p.block_size = 256
    ctx = C.c_void_p()
    abi.check(lib.dem_create(C.byref(p), n, 1, C.byref(ctx)))
    try:
        f = lambda k: np.zeros(k, dtype=np.float32)
        i = lambda k: np.zeros(k, dtype=np.int32)
        arrs = dict(pos=f(2 * n), vel=f(2 * n), acc=f(2 * n), rot=f(n), rot_vel=f(n), rot_acc=f(n), forces=f(6 * n),
                    radii=f(n), fixity=i(6 * n), material_pointers=i(n))
        v = abi.DemStateView()
        v.n = n
        for k, a in arrs.items():
            setattr(v, k, abi.iptr(a) if a.dtype == np.int32 else abi.fptr(a))
        abi.check(lib.dem_upload_state(ctx, C.byref(v)))
        out = C.c_double(-1.0)
        rc = lib.dem_kinetic_energy_noalloc(ctx, C.byref(out))
        assert rc == abi.DEM_ECAPACITY
        assert b"DEM_KE_NOALLOC_MAX" in lib.dem_last_error()
        assert out.value == -1.0
        # the allocating entry point still serves the scene (all discs inactive: 0)
        e = C.c_double(-1.0)
        abi.check(lib.dem_kinetic_energy(ctx, C.byref(e)))
        assert e.value == 0.0
    finally:
        lib.dem_destroy(ctx)
