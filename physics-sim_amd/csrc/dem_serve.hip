// dem_serve.hip - serving loop: continuous batching of independent scenes on one GPU.
//
// Requests (a scene's parameters, its initial state and a step count) queue on the
// host.  Each tick admits queued scenes into one Engine's pools while their footprint
// (slots, cells, bonds, bond records, materials) fits, advances every resident scene
// by up to `substeps` steps in the same kernel launches (per-scene offsets, grids and
// flags from the device scene table), and retires the scenes that finished: their
// state is downloaded into host staging and their pool regions are freed for the next
// admissions.  Scenes admitted while others are mid-run share the next launches.
//
// A scene's result does not depend on its neighbours in a launch, on where its region
// sits in the pools, or on how its steps are split over ticks: every particle sums its
// partners in ascending id and the storage order never enters the arithmetic.  So a
// scene served in a mix reproduces its solo run bit for bit.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <exception>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/dem.h"
#include "dem_engine.h"

using psim::set_error;

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <typename T>
void take_copy(std::vector<T>& dst, const T* src, size_t n) {
    if (src) dst.assign(src, src + n);
    else dst.clear();
}

template <typename T>
T* ptr_or_null(std::vector<T>& v) {
    return v.empty() ? nullptr : v.data();
}

struct Request {
    int64_t ticket = 0;
    dem_params params{};
    std::vector<dem_material> mats;
    int32_t n = 0, n_bonds = 0, contact_slots = 0;
    int32_t nsteps = 0, done = 0;
    std::vector<float> pos, vel, acc, rot, rot_vel, rot_acc, forces, radii, bond_strengths, contacts, data;
    std::vector<int32_t> fixity, matptr, bonds;
    int32_t sid = -1;
    int status = DEM_REQ_QUEUED;
    int error = DEM_OK;           // DEM_REQ_FAILED: the admission's error code and message
    std::string error_text;
    double t_submit = 0, t_admit = 0, t_done = 0;

    dem_state_view view() {
        dem_state_view v{};
        v.n = n;
        v.pos = ptr_or_null(pos);
        v.vel = ptr_or_null(vel);
        v.acc = ptr_or_null(acc);
        v.rot = ptr_or_null(rot);
        v.rot_vel = ptr_or_null(rot_vel);
        v.rot_acc = ptr_or_null(rot_acc);
        v.forces = ptr_or_null(forces);
        v.radii = ptr_or_null(radii);
        v.fixity = ptr_or_null(fixity);
        v.material_pointers = ptr_or_null(matptr);
        v.bonds = ptr_or_null(bonds);
        v.n_bonds = n_bonds;
        v.bond_strengths = ptr_or_null(bond_strengths);
        v.contacts = ptr_or_null(contacts);
        v.contact_slots = contact_slots;
        v.data = ptr_or_null(data);
        return v;
    }
    int64_t particle_steps() const { return (int64_t)n * nsteps; }
};

double quantile(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    // nearest rank
    size_t k = (size_t)std::ceil(q * (double)v.size());
    if (k < 1) k = 1;
    if (k > v.size()) k = v.size();
    return v[k - 1];
}

}  // namespace

struct dem_server {
    psim::Engine eng;
    dem_server_params prm{};
    int64_t next_ticket = 1;
    std::deque<int64_t> queue;
    std::vector<int64_t> resident;
    std::unordered_map<int64_t, Request> reqs;
    std::vector<double> latencies;   // submit -> done, seconds, of every finished scene
    int64_t completed = 0, failed = 0, ticks = 0, launches_steps = 0;
    int64_t particle_steps_done = 0;
    double t_first_submit = -1.0, t_last_done = 0.0;
};

// variadic: a body may hold top-level commas (declarations, initialiser lists)
#define SRV_GUARD(...)                                              \
    try {                                                           \
        __VA_ARGS__                                                 \
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

namespace {

// A request the engine refused for a reason other than room (its parameters or its
// state): it leaves the queue as DEM_REQ_FAILED with the error, and serving goes on.
void fail(dem_server* s, Request& r, int rc) {
    r.status = DEM_REQ_FAILED;
    r.error = rc;
    r.error_text = dem_last_error();
    r.t_done = now_s();
    s->failed++;
}

// Admits queued requests in arrival order.  When the head does not fit, up to
// `bypass` later requests may be tried (smaller scenes fill the gaps); the head keeps
// its place and is tried first again on the next tick.  A request that fails for any
// other reason is dropped from the queue as failed.
int admit(dem_server* s) {
    const int32_t max_res = s->prm.max_resident > 0 ? s->prm.max_resident : INT32_MAX;
    size_t qi = 0;
    int32_t skipped = 0;
    while (qi < s->queue.size() && (int32_t)s->resident.size() < max_res) {
        Request& r = s->reqs.at(s->queue[qi]);
        const int32_t recs = 2 * r.n_bonds + r.n;
        int32_t sid = -1;
        int rc = s->eng.add_scene(r.params, std::max(r.n, 1), r.n_bonds, recs, &sid);
        if (rc == DEM_EBUSY) {
            if (++skipped > s->prm.bypass) break;
            qi++;
            continue;
        }
        if (rc == DEM_OK) {
            dem_state_view v = r.view();
            rc = s->eng.upload(sid, v);
            if (rc != DEM_OK) {
                const std::string why = dem_last_error();
                s->eng.remove_scene(sid);
                set_error(why);
            }
        }
        if (rc == DEM_EDEVICE || rc == DEM_ENOMEM) return rc;   // the device, not the request
        if (rc != DEM_OK) {
            fail(s, r, rc);
            s->queue.erase(s->queue.begin() + (std::ptrdiff_t)qi);
            continue;
        }
        r.sid = sid;
        r.status = DEM_REQ_RESIDENT;
        r.t_admit = now_s();
        s->resident.push_back(r.ticket);
        s->queue.erase(s->queue.begin() + (std::ptrdiff_t)qi);
    }
    return DEM_OK;
}

// Downloads and frees every finished scene.  The resident list is rebuilt as it goes:
// after a failure the scenes not yet retired, the failed one among them, stay resident.
int retire(dem_server* s) {
    std::vector<int64_t> keep;
    int rc = DEM_OK;
    for (int64_t t : s->resident) {
        Request& r = s->reqs.at(t);
        if (rc != DEM_OK || r.done < r.nsteps) {
            keep.push_back(t);
            continue;
        }
        dem_state_view v = r.view();
        rc = s->eng.download(r.sid, v);
        if (rc == DEM_OK) rc = s->eng.remove_scene(r.sid);
        if (rc != DEM_OK) {
            keep.push_back(t);
            continue;
        }
        r.sid = -1;
        r.status = DEM_REQ_DONE;
        r.t_done = now_s();
        s->latencies.push_back(r.t_done - r.t_submit);
        s->completed++;
        s->particle_steps_done += r.particle_steps();
        s->t_last_done = r.t_done;
    }
    s->resident.swap(keep);
    return rc;
}

}  // namespace

extern "C" {

int dem_server_create(const dem_server_params* p, dem_server** out) {
    SRV_GUARD({
        if (!p || !out) { set_error("dem_server_create: NULL argument"); return DEM_EINVAL; }
        *out = nullptr;
        if (p->slot_capacity <= 0 || p->cell_capacity <= 0 || p->scene_capacity <= 0 || p->substeps <= 0) {
            set_error("dem_server_create: capacities and substeps must be positive");
            return DEM_EINVAL;
        }
        if (dem_device_count() <= 0) { set_error("dem_server_create: no HIP device"); return DEM_EDEVICE; }
        dem_server* s = new dem_server();
        s->prm = *p;
        if (s->prm.bypass < 0) s->prm.bypass = 0;
        const int64_t recs = p->record_capacity > 0 ? p->record_capacity : 2 * p->bond_capacity + p->slot_capacity;
        const int32_t mats = p->material_capacity > 0 ? p->material_capacity : 8 * p->scene_capacity;
        int rc = s->eng.init(p->slot_capacity, p->cell_capacity, std::max<int64_t>(p->bond_capacity, 1), recs, mats,
                             p->scene_capacity, p->block_size > 0 ? p->block_size : 256, p->record_data != 0);
        if (rc != DEM_OK) {
            delete s;
            return rc;
        }
        s->eng.enable_graphs(p->graph != 0);
        *out = s;
        return DEM_OK;
    })
}

int dem_server_destroy(dem_server* s) {
    SRV_GUARD({
        delete s;
        return DEM_OK;
    })
}

int dem_server_submit(dem_server* s, const dem_params* params, const dem_state_view* st, int32_t nsteps,
                      int64_t* ticket) {
    SRV_GUARD({
        if (!s || !params || !st || !ticket) { set_error("dem_server_submit: NULL argument"); return DEM_EINVAL; }
        if (nsteps < 0) { set_error("dem_server_submit: negative step count"); return DEM_EINVAL; }
        if (st->n <= 0 || !st->pos || !st->vel || !st->acc || !st->rot || !st->rot_vel || !st->rot_acc ||
            !st->forces || !st->radii || !st->fixity || !st->material_pointers) {
            set_error("dem_server_submit: state view needs n > 0 and every per-particle array");
            return DEM_EINVAL;
        }
        // what admission would refuse for good is refused here, not queued
        if (int rc = psim::validate_params(*params)) return rc;
        if (std::max(params->n_materials, 8) > s->eng.material_capacity()) {
            set_error("dem_server_submit: more materials than the server's material pool holds");
            return DEM_ECAPACITY;
        }
        {
            float rsmax = 0.0f;   // the largest radius, large bodies aside
            psim::split_large(st->radii, st->n, &rsmax);
            if (int rc = psim::check_cell_size(params->grid, rsmax)) return rc;
        }
        {
            // a scene that can never fit the pools is refused now rather than queued forever
            const int64_t blk = s->prm.block_size > 0 ? s->prm.block_size : 256;
            const int64_t slots = ((int64_t)st->n + blk - 1) / blk * blk;
            const int64_t cells = (int64_t)params->grid.w * params->grid.h + 1;
            const int64_t recs = s->prm.record_capacity > 0 ? s->prm.record_capacity
                                                            : 2 * s->prm.bond_capacity + s->prm.slot_capacity;
            if (slots > s->prm.slot_capacity || cells > s->prm.cell_capacity ||
                (int64_t)std::max(st->n_bonds, 0) > s->prm.bond_capacity ||
                2 * (int64_t)std::max(st->n_bonds, 0) + st->n > recs) {
                set_error("dem_server_submit: the scene is larger than the server's pools");
                return DEM_ECAPACITY;
            }
        }
        Request r;
        r.ticket = s->next_ticket++;
        r.params = *params;
        r.mats.assign(params->materials, params->materials + params->n_materials);
        r.params.materials = r.mats.data();
        r.n = st->n;
        r.n_bonds = st->n_bonds;
        r.contact_slots = st->contacts ? st->contact_slots : 0;
        r.nsteps = nsteps;
        const size_t n = (size_t)st->n;
        take_copy(r.pos, st->pos, 2 * n);
        take_copy(r.vel, st->vel, 2 * n);
        take_copy(r.acc, st->acc, 2 * n);
        take_copy(r.rot, st->rot, n);
        take_copy(r.rot_vel, st->rot_vel, n);
        take_copy(r.rot_acc, st->rot_acc, n);
        take_copy(r.forces, st->forces, 6 * n);
        take_copy(r.radii, st->radii, n);
        take_copy(r.fixity, st->fixity, 6 * n);
        take_copy(r.matptr, st->material_pointers, n);
        take_copy(r.bonds, st->bonds, 3 * (size_t)std::max(st->n_bonds, 0));
        take_copy(r.bond_strengths, st->bond_strengths, 2 * (size_t)std::max(st->n_bonds, 0));
        take_copy(r.contacts, st->contacts, (size_t)r.contact_slots * 6 * n);
        take_copy(r.data, st->data, 7 * n);
        r.t_submit = now_s();
        if (s->t_first_submit < 0) s->t_first_submit = r.t_submit;
        const int64_t t = r.ticket;
        // the result must fit where it will be written: keep room for a full table
        if (r.contact_slots == 0 && s->prm.contact_slots > 0) {
            r.contact_slots = s->prm.contact_slots;
            r.contacts.assign((size_t)r.contact_slots * 6 * n, 0.0f);
            int32_t* w = reinterpret_cast<int32_t*>(r.contacts.data());
            std::fill(w, w + r.contacts.size(), -1);
        }
        s->reqs.emplace(t, std::move(r));
        s->reqs.at(t).params.materials = s->reqs.at(t).mats.data();
        s->queue.push_back(t);
        *ticket = t;
        return DEM_OK;
    })
}

int dem_server_tick(dem_server* s, int32_t* completed) {
    SRV_GUARD({
        if (!s) { set_error("dem_server_tick: NULL server"); return DEM_EINVAL; }
        const int64_t before = s->completed;
        int rc = admit(s);
        if (rc != DEM_OK) return rc;
        if (!s->resident.empty()) {
            std::vector<int32_t> sids, ns;
            for (int64_t t : s->resident) {
                Request& r = s->reqs.at(t);
                const int32_t k = std::min(s->prm.substeps, r.nsteps - r.done);
                sids.push_back(r.sid);
                ns.push_back(k);
            }
            rc = s->eng.step(sids, ns);
            if (rc != DEM_OK) return rc;
            for (size_t i = 0; i < s->resident.size(); i++) s->reqs.at(s->resident[i]).done += ns[i];
            s->launches_steps += *std::max_element(ns.begin(), ns.end());
            rc = s->eng.sync();
            if (rc != DEM_OK) return rc;
            rc = retire(s);
            if (rc != DEM_OK) return rc;
        }
        s->ticks++;
        if (completed) *completed = (int32_t)(s->completed - before);
        return DEM_OK;
    })
}

int dem_server_status(dem_server* s, int64_t ticket, int32_t* status) {
    SRV_GUARD({
        if (!s || !status) { set_error("dem_server_status: NULL argument"); return DEM_EINVAL; }
        auto it = s->reqs.find(ticket);
        if (it == s->reqs.end()) { set_error("dem_server_status: unknown ticket"); return DEM_EINVAL; }
        *status = it->second.status;
        return DEM_OK;
    })
}

int dem_server_take(dem_server* s, int64_t ticket, dem_state_view* out, double* latency_s) {
    SRV_GUARD({
        if (!s || !out) { set_error("dem_server_take: NULL argument"); return DEM_EINVAL; }
        auto it = s->reqs.find(ticket);
        if (it == s->reqs.end()) { set_error("dem_server_take: unknown ticket"); return DEM_EINVAL; }
        Request& r = it->second;
        if (r.status == DEM_REQ_FAILED) {
            // the admission's error, once: the ticket is forgotten
            const int rc = r.error;
            set_error("dem_server_take: the scene failed admission: " + r.error_text);
            if (latency_s) *latency_s = r.t_done - r.t_submit;
            s->reqs.erase(it);
            return rc;
        }
        if (r.status != DEM_REQ_DONE) { set_error("dem_server_take: scene not finished"); return DEM_ESTATE; }
        if (out->n != r.n) { set_error("dem_server_take: view size differs from the scene"); return DEM_EINVAL; }
        const size_t n = (size_t)r.n;
        auto put = [](void* dst, const void* src, size_t bytes) {
            if (dst && src && bytes) std::memcpy(dst, src, bytes);
        };
        put(out->pos, r.pos.data(), 8 * n);
        put(out->vel, r.vel.data(), 8 * n);
        put(out->acc, r.acc.data(), 8 * n);
        put(out->rot, r.rot.data(), 4 * n);
        put(out->rot_vel, r.rot_vel.data(), 4 * n);
        put(out->rot_acc, r.rot_acc.data(), 4 * n);
        put(out->forces, r.forces.data(), 24 * n);
        put(out->radii, r.radii.data(), 4 * n);
        put(out->fixity, r.fixity.data(), 24 * n);
        put(out->material_pointers, r.matptr.data(), 4 * n);
        if (out->bonds && !r.bonds.empty()) {
            if (out->n_bonds < r.n_bonds) { set_error("dem_server_take: bond array too small"); return DEM_ECAPACITY; }
            put(out->bonds, r.bonds.data(), 12 * (size_t)r.n_bonds);
        }
        if (out->contacts && !r.contacts.empty()) {
            if (out->contact_slots != r.contact_slots) {
                set_error("dem_server_take: contact_slots differs from the scene's table");
                return DEM_EINVAL;
            }
            put(out->contacts, r.contacts.data(), r.contacts.size() * 4);
        }
        if (out->data && !r.data.empty()) put(out->data, r.data.data(), 28 * n);
        if (latency_s) *latency_s = r.t_done - r.t_submit;
        s->reqs.erase(it);
        return DEM_OK;
    })
}

int dem_server_stats(dem_server* s, double* out, int32_t n_out) {
    SRV_GUARD({
        if (!s || !out) { set_error("dem_server_stats: NULL argument"); return DEM_EINVAL; }
        const double span = s->completed && s->t_first_submit >= 0 ? s->t_last_done - s->t_first_submit : 0.0;
        const double v[] = {
            (double)s->completed,
            (double)s->resident.size(),
            (double)s->queue.size(),
            quantile(s->latencies, 0.50),
            quantile(s->latencies, 0.99),
            span > 0 ? (double)s->completed / span : 0.0,
            (double)s->particle_steps_done,
            (double)s->ticks,
            (double)s->launches_steps,
            (double)s->failed,
        };
        const int32_t m = std::min<int32_t>(n_out, (int32_t)(sizeof(v) / sizeof(v[0])));
        for (int32_t i = 0; i < m; i++) out[i] = v[i];
        return DEM_OK;
    })
}

}  // extern "C"
