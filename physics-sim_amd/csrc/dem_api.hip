// dem_api.hip - the C ABI of include/dem.h over psim::Engine.
#include <exception>
#include <new>
#include <string>

#include "../../include/dem.h"
#include "dem_engine.h"

namespace psim {
// A fixed buffer, so that setting an error allocates nothing (the no-allocation entry
// points report their errors through it too).
static thread_local char g_last_error[1024];
void set_error(const char* msg) {
    size_t k = 0;
    for (; msg && msg[k] && k + 1 < sizeof(g_last_error); k++) g_last_error[k] = msg[k];
    g_last_error[k] = 0;
}
void set_error(const std::string& msg) { set_error(msg.c_str()); }
int init_static_scratch(int device);
}  // namespace psim

struct dem_ctx {
    psim::Engine eng;
    int32_t sid = -1;
};

using psim::set_error;

#define DEM_GUARD(body)                                             \
    try {                                                           \
        body                                                        \
    } catch (const std::bad_alloc&) {                               \
        set_error("host allocation failed");                        \
        return DEM_ENOMEM;                                          \
    } catch (const std::exception& ex) {                            \
        set_error(std::string("internal error: ") + ex.what());     \
        return DEM_EDEVICE;                                         \
    } catch (...) {                                                 \
        set_error("internal error");                                \
        return DEM_EDEVICE;                                         \
    }

extern "C" {

const char* dem_last_error(void) { return psim::g_last_error; }

const char* dem_version(void) { return "psim-dem 0.1 (gfx950)"; }

int dem_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int dem_create(const dem_params* params, int32_t capacity, int32_t bond_capacity, dem_ctx** out) {
    DEM_GUARD({
        if (!params || !out) { set_error("dem_create: NULL argument"); return DEM_EINVAL; }
        *out = nullptr;
        if (capacity <= 0) { set_error("dem_create: capacity must be positive"); return DEM_EINVAL; }
        if (bond_capacity < 0) { set_error("dem_create: negative bond capacity"); return DEM_EINVAL; }
        if (dem_device_count() <= 0) { set_error("dem_create: no HIP device"); return DEM_EDEVICE; }
        dem_ctx* c = new dem_ctx();
        const int32_t block = params->block_size > 0 ? params->block_size : 256;
        const int64_t slots = ((int64_t)capacity + block - 1) / block * block;
        const int64_t cells = (int64_t)params->grid.w * params->grid.h + 1;
        // bond endpoint records: two per bond plus room for broken-bond / odd slots
        const int64_t recs = 2 * (int64_t)bond_capacity + capacity;
        int rc = c->eng.init(slots, cells, bond_capacity, recs, params->n_materials > 8 ? params->n_materials : 8, 1,
                             block, true);
        if (rc == DEM_OK) rc = c->eng.add_scene(*params, capacity, bond_capacity, (int32_t)recs, &c->sid);
        if (rc != DEM_OK) {
            delete c;
            return rc;
        }
        // the no-allocation observables' device scratch: its address on this device
        rc = psim::init_static_scratch(c->eng.device());
        if (rc != DEM_OK) {
            delete c;
            return rc;
        }
        c->eng.enable_graphs(params->graph != 0);
        *out = c;
        return DEM_OK;
    })
}

int dem_destroy(dem_ctx* ctx) {
    DEM_GUARD({
        if (!ctx) return DEM_OK;
        delete ctx;
        return DEM_OK;
    })
}

int dem_set_params(dem_ctx* ctx, const dem_params* params) {
    DEM_GUARD({
        if (!ctx || !params) { set_error("dem_set_params: NULL argument"); return DEM_EINVAL; }
        const int rc = ctx->eng.set_params(ctx->sid, *params);
        if (rc == DEM_OK) ctx->eng.enable_graphs(params->graph != 0);
        return rc;
    })
}

int dem_upload_state(dem_ctx* ctx, const dem_state_view* state) {
    DEM_GUARD({
        if (!ctx || !state) { set_error("dem_upload_state: NULL argument"); return DEM_EINVAL; }
        return ctx->eng.upload(ctx->sid, *state);
    })
}

int dem_step(dem_ctx* ctx, int32_t n) {
    DEM_GUARD({
        if (!ctx) { set_error("dem_step: NULL ctx"); return DEM_EINVAL; }
        if (n < 0) { set_error("dem_step: negative step count"); return DEM_EINVAL; }
        return ctx->eng.step({ctx->sid}, {n});
    })
}

int dem_download_state(dem_ctx* ctx, dem_state_view* state) {
    DEM_GUARD({
        if (!ctx || !state) { set_error("dem_download_state: NULL argument"); return DEM_EINVAL; }
        return ctx->eng.download(ctx->sid, *state);
    })
}

int dem_sync(dem_ctx* ctx) {
    DEM_GUARD({
        if (!ctx) { set_error("dem_sync: NULL ctx"); return DEM_EINVAL; }
        return ctx->eng.sync();
    })
}

int dem_kinetic_energy(dem_ctx* ctx, double* out) {
    DEM_GUARD({
        if (!ctx || !out) { set_error("dem_kinetic_energy: NULL argument"); return DEM_EINVAL; }
        return ctx->eng.kinetic_energy(ctx->sid, out);
    })
}

int dem_kinetic_energy_noalloc(dem_ctx* ctx, double* out) {
    // no DEM_GUARD: nothing here allocates or throws
    if (!ctx || !out) {
        set_error("dem_kinetic_energy_noalloc: NULL argument");
        return DEM_EINVAL;
    }
    return ctx->eng.kinetic_energy_noalloc(ctx->sid, out);
}

int dem_broken_bond_count(dem_ctx* ctx, int64_t* out) {
    DEM_GUARD({
        if (!ctx || !out) { set_error("dem_broken_bond_count: NULL argument"); return DEM_EINVAL; }
        return ctx->eng.broken_bonds(ctx->sid, out);
    })
}

int dem_counters(dem_ctx* ctx, int64_t* out, int32_t n_out) {
    DEM_GUARD({
        if (!ctx || !out || n_out < 0) { set_error("dem_counters: bad argument"); return DEM_EINVAL; }
        return ctx->eng.counters(ctx->sid, out, n_out);
    })
}

int dem_get_datum(dem_ctx* ctx, const int32_t* select, double* out) {
    DEM_GUARD({
        if (!ctx || !select || !out) { set_error("dem_get_datum: NULL argument"); return DEM_EINVAL; }
        return ctx->eng.datum(ctx->sid, select, out);
    })
}

int dem_debug_broad_phase(dem_ctx* ctx, int32_t* cell_keys, int32_t* sorted_ids, int32_t* cell_start) {
    DEM_GUARD({
        if (!ctx) { set_error("dem_debug_broad_phase: NULL ctx"); return DEM_EINVAL; }
        return ctx->eng.debug_broad_phase(ctx->sid, cell_keys, sorted_ids, cell_start);
    })
}

int dem_debug_contacts(dem_ctx* ctx, int32_t max_per, int32_t* counts, int32_t* partners) {
    DEM_GUARD({
        if (!ctx || !counts || !partners || max_per <= 0) { set_error("dem_debug_contacts: bad argument"); return DEM_EINVAL; }
        return ctx->eng.debug_contacts(ctx->sid, max_per, counts, partners);
    })
}

int dem_build_bond_pairs(dem_ctx* ctx, int32_t max_bonds, float factor, float gap, int32_t* pairs,
                         int32_t bond_cap, int32_t* n_out) {
    DEM_GUARD({
        if (!ctx || !pairs || !n_out || bond_cap < 0) { set_error("dem_build_bond_pairs: bad argument"); return DEM_EINVAL; }
        return ctx->eng.bond_pairs(ctx->sid, max_bonds, factor, gap, pairs, bond_cap, n_out);
    })
}

int dem_group_sums(dem_ctx* ctx, const int32_t* ids, int32_t n_ids, double* out) {
    DEM_GUARD({
        if (!ctx || !out || (n_ids > 0 && !ids)) { set_error("dem_group_sums: NULL argument"); return DEM_EINVAL; }
        return ctx->eng.group_sums(ctx->sid, ids, n_ids, out);
    })
}

int dem_last_step_timing(dem_ctx* ctx, int32_t enable, double* ms_out, int64_t* launches_out) {
    DEM_GUARD({
        if (!ctx) { set_error("dem_last_step_timing: NULL ctx"); return DEM_EINVAL; }
        psim::Engine& e = ctx->eng;
        if (ms_out) {
            ms_out[0] = e.t_total; ms_out[1] = e.t_step; ms_out[2] = e.t_scan; ms_out[3] = e.t_scatter;
            ms_out[4] = e.t_reorder; ms_out[5] = e.t_bonds;
        }
        if (launches_out) {
            launches_out[0] = e.n_step + e.n_scan + e.n_scatter + e.n_reorder + e.n_bonds;
            launches_out[1] = e.n_step; launches_out[2] = e.n_scan; launches_out[3] = e.n_scatter;
            launches_out[4] = e.n_reorder; launches_out[5] = e.n_bonds;
        }
        if (enable >= 0) e.enable_timing(enable != 0);
        return DEM_OK;
    })
}

}  // extern "C"
