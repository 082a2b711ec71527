// compute_prog.cpp - ComputeProg of the C++ host mirror: WGPUComputeProg's compute half
// (wgpu_prog.rs:677-696) driven through the C ABI of include/dem.h.  There is no CPU
// path: without libpsim.so's device the constructor throws.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <string>

#include "psim_host.h"

namespace psim {
namespace host {

namespace {

void check(int rc) {
    if (rc != DEM_OK) throw DemError(rc, std::string("libpsim: ") + dem_last_error());
}

}  // namespace

constexpr double kSkinFraction = 0.1;
constexpr int kLargeMax = 64;

float small_radius(const std::vector<float>& radii, int32_t n) {
    // the library's large-body rule: radii in decreasing order, the first k <= 64 with
    // r_{k-1} > 2 r_k sets the top k apart and r_k is the answer (else r_0)
    std::vector<float> r;
    for (int32_t i = 0; i < n; i++)
        if (radii[i] > 0.0f) r.push_back(radii[i]);
    if (r.empty()) return 0.0f;
    const size_t k = std::min<size_t>(r.size(), kLargeMax + 1);
    std::partial_sort(r.begin(), r.begin() + (long)k, r.end(), std::greater<float>());
    for (size_t i = 1; i < k; i++)
        if (r[i - 1] > 2.0f * r[i]) return r[i];
    return r[0];
}

float min_cell_size(float rmax, float extent) {
    // 2 r_max plus room for the rounding of (x - base) / cell_size, a few ulp of the extent
    const float e = extent;
    const double ulp = (double)std::nextafter(e, INFINITY) - (double)e;
    return (float)(2.0 * (double)rmax * (1.0 + 1e-4) + 8.0 * ulp);
}

dem_grid ComputeProg::default_grid(const Settings& settings, const State& state, float cell_size) {
    const int32_t n = state.p_count;
    double rmax = n ? (double)*std::max_element(state.radii.begin(), state.radii.begin() + n)
                    : (double)settings.setup.max_radius;
    rmax = std::max(rmax, 1e-6);
    // large bodies pair through a list: the cells follow the other discs
    const double rs = n ? std::max((double)small_radius(state.radii, n), 1e-6) : rmax;
    const double ext = std::max((double)settings.simulation.hor_bound, (double)settings.simulation.vert_bound) + 8.0 * rmax;
    const float lo = min_cell_size((float)rs, (float)ext);
    // default: room for a neighbour-list skin of kSkinFraction * 2 r_small
    const float cs = cell_size > 0.0f ? std::max(lo, cell_size) : (float)((double)lo + kSkinFraction * 2.0 * rs);
    return settings.box_grid(cs);
}

dem_params ComputeProg::params(const Settings& s) {
    dem_params p{};
    p.u = s.collision_settings();
    mats_.clear();
    for (size_t k = 0; k + 6 <= s.materials.size(); k += 6) {
        const float* m = s.materials.data() + k;
        mats_.push_back(dem_material{m[0], m[1], m[2], m[3], m[4], m[5]});
    }
    p.materials = mats_.data();
    p.n_materials = (int32_t)mats_.size();
    p.grid = grid_;
    p.use_f64 = use_f64_ ? 1 : 0;
    p.record_data = opt_.record_data ? 1 : 0;
    p.reorder_interval = opt_.reorder_interval;
    p.block_size = opt_.block_size;
    p.graph = opt_.graph ? 1 : 0;
    p.restitution = s.physics.restitution;
    return p;
}

ComputeProg::ComputeProg(const Settings& settings, State& state, const Options& opt) : opt_(opt) {
    if (dem_device_count() <= 0) throw DemError(DEM_EDEVICE, "no HIP device visible");
    use_f64_ = opt.use_f64 < 0 ? settings.simulation.use_f64 : opt.use_f64 != 0;
    grid_ = default_grid(settings, state, opt.cell_size);
    const int32_t capacity = opt.capacity > 0 ? opt.capacity : std::max(state.p_count, 1);
    bond_capacity_ = opt.bond_capacity > 0 ? opt.bond_capacity : std::max(state.n_bonds(), 1);
    const dem_params p = params(settings);
    check(dem_create(&p, capacity, bond_capacity_, &ctx_));
    restore(state);
}

ComputeProg::~ComputeProg() {
    if (ctx_) dem_destroy(ctx_);
}

void ComputeProg::compute(const Settings& settings) { step(settings.simulation.gen_per_frame); }

void ComputeProg::step(int32_t n) { check(dem_step(ctx_, n)); }

void ComputeProg::sync() { check(dem_sync(ctx_)); }

State& ComputeProg::update_state() {
    dem_state_view v = state_->view(opt_.record_data);
    check(dem_download_state(ctx_, &v));
    return *state_;
}

void ComputeProg::restore(State& state) {
    state_ = &state;
    dem_state_view v = state.view(true);
    check(dem_upload_state(ctx_, &v));
}

void ComputeProg::set_params(const Settings& settings) {
    const dem_params p = params(settings);
    check(dem_set_params(ctx_, &p));
}

double ComputeProg::kinetic_energy() {
    double e = 0.0;
    check(dem_kinetic_energy(ctx_, &e));
    return e;
}

int64_t ComputeProg::broken_bond_count() {
    int64_t k = 0;
    check(dem_broken_bond_count(ctx_, &k));
    return k;
}

State& ComputeProg::regen_bonds(const Settings& settings, float factor, float gap) {
    State& st = update_state();
    const int32_t n = st.p_count;
    const int32_t cap = std::max(n * settings.setup.max_bonds, 1);
    std::vector<int32_t> pairs(2 * (size_t)cap);
    int32_t nout = 0;
    check(dem_build_bond_pairs(ctx_, settings.setup.max_bonds, factor, gap, pairs.data(), cap, &nout));
    pairs.resize(2 * (size_t)nout);
    assemble_bonds(st, pairs, settings.setup.max_contacts, st.bonds, st.contacts);
    st.bond_strengths.clear();
    if (st.n_bonds() > bond_capacity_)
        throw DemError(DEM_ECAPACITY, "ComputeProg was created with bond_capacity " + std::to_string(bond_capacity_) +
                                          " < " + std::to_string(st.n_bonds()) + " bonds");
    restore(st);
    return st;
}

void assemble_bonds(const State& st, const std::vector<int32_t>& pairs, int32_t slots, std::vector<int32_t>& bonds,
                    std::vector<float>& contacts) {
    const int32_t cap = st.capacity();
    std::vector<int32_t> words((size_t)cap * slots * 6, -1);
    const size_t nb = pairs.size() / 2;
    auto words_to_floats = [&]() {
        contacts.resize(words.size());
        std::memcpy(contacts.data(), words.data(), 4 * words.size());
    };
    if (nb == 0) {
        bonds.assign(1, -1);
        words_to_floats();
        return;
    }
    bonds.resize(3 * nb);
    for (size_t k = 0; k < nb; k++) {
        const int32_t i = pairs[2 * k], j = pairs[2 * k + 1];
        const float dx = st.pos[2 * (size_t)j] - st.pos[2 * (size_t)i];
        const float dy = st.pos[2 * (size_t)j + 1] - st.pos[2 * (size_t)i + 1];
        const float xx = dx * dx;
        const float yy = dy * dy;
        const float mag = std::sqrt(xx + yy);
        const float nx = dx / mag, ny = dy / mag;
        // atan2 of the f32 unit vector (state.rs:306-310), evaluated in f64, rounded once
        const float ang = (float)std::atan2((double)nx, (double)ny);
        bonds[3 * k] = 1;
        std::memcpy(&bonds[3 * k + 1], &ang, 4);
        std::memcpy(&bonds[3 * k + 2], &mag, 4);
    }
    // each particle's bonds fill its slots in creation order = ascending partner id
    struct End { int32_t owner, partner, bid; };
    std::vector<End> ends;
    ends.reserve(2 * nb);
    for (size_t k = 0; k < nb; k++) {
        ends.push_back({pairs[2 * k], pairs[2 * k + 1], (int32_t)k});
        ends.push_back({pairs[2 * k + 1], pairs[2 * k], (int32_t)k});
    }
    std::stable_sort(ends.begin(), ends.end(), [](const End& a, const End& b) {
        return a.owner != b.owner ? a.owner < b.owner : a.partner < b.partner;
    });
    int32_t rank = 0;
    for (size_t q = 0; q < ends.size(); q++) {
        rank = (q > 0 && ends[q].owner == ends[q - 1].owner) ? rank + 1 : 0;
        if (rank >= slots) continue;
        int32_t* w = words.data() + ((size_t)ends[q].owner * slots + rank) * 6;
        w[0] = ends[q].owner;
        w[1] = ends[q].partner;
        w[2] = 0;
        w[3] = 0;
        w[4] = 0;
        w[5] = ends[q].bid;
    }
    words_to_floats();
}

}  // namespace host
}  // namespace psim
