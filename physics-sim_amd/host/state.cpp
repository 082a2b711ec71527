// state.cpp - Settings and State of the C++ host mirror (settings.rs, state.rs).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iterator>

#include "psim_host.h"

namespace psim {
namespace host {

namespace {

// Offsets inside the 212-byte Settings struct (new_schema.fbs:1-84): view at 0,
// setup at 76, simulation at 116, physics at 156.
constexpr size_t SETUP = 76, SIM = 116, PHYS = 156;

template <typename T>
T get(const std::vector<uint8_t>& raw, size_t at) {
    T v;
    std::memcpy(&v, raw.data() + at, sizeof(T));
    return v;
}
bool getb(const std::vector<uint8_t>& raw, size_t at) { return raw[at] != 0; }
template <typename T>
void put(std::vector<uint8_t>& raw, size_t at, T v) {
    std::memcpy(raw.data() + at, &v, sizeof(T));
}
void putb(std::vector<uint8_t>& raw, size_t at, bool v) { raw[at] = v ? 1 : 0; }

}  // namespace

dem_uniform Settings::collision_settings() const {
    const SimulationSettings& s = simulation;
    const PhysicsSettings& p = physics;
    dem_uniform u{};
    u.walls = s.walls ? 1 : 0;
    u.hor_bound = s.hor_bound;
    u.vert_bound = s.vert_bound;
    u.round_bounds = s.round_walls ? 1 : 0;
    u.bound_radius = s.wall_radius;
    u.wall_friction = wall_friction;
    u.gravity = p.gravity ? 1 : 0;
    u.planet_mode = p.planet_mode ? 1 : 0;
    u.bonds = p.bonds;
    u.collisions = p.collisions ? 1 : 0;
    u.friction_coefficient = p.friction_coefficient;
    u.gravity_acc = p.gravity_acceleration;
    u.bond_normal_stiffness = p.bond_normal_stiffness;
    u.bonds_tear = p.bond_tearing ? 1 : 0;
    u.bond_normal_strength = p.bond_normal_strength;
    u.contact_damping = p.contact_damping;
    u.bond_damping = p.bond_damping;
    u.drag = p.drag;
    u.bond_shear_strength = p.bond_shear_strength;
    u.dT = s.timestep;
    u.bond_shear_stiffness = p.bond_shear_stiffness;
    u.gravity_x = world_pos[0];
    u.gravity_y = world_pos[1];
    u.mouse_gravity = p.mouse_gravity ? 1 : 0;
    u.moment_contribution_factor = p.moment_contribution_factor;
    u.collision_interval = p.collision_interval;
    u.local_damping = p.local_damping ? 1 : 0;
    u.local_damping_alpha = p.local_damping_alpha;
    u.particles = setup.particles;
    return u;
}

dem_grid Settings::grid_info() const {
    // cells of 2 * max_radius over twice the box; per-cell capacity from the radius ratio
    const float width = simulation.hor_bound * 2.0f;
    const float height = simulation.vert_bound * 2.0f;
    const float max_rad = setup.max_radius * 2.0f;
    const float min_rad = setup.variable_rad ? setup.min_radius : setup.max_radius;
    dem_grid g{};
    g.cell_size = max_rad;
    g.w = (int32_t)std::ceil(width / max_rad);
    g.h = (int32_t)std::ceil(height / max_rad);
    const float t = max_rad / min_rad + 1.0f;
    const float t2 = t * t;
    g.cell_cap = std::min((int32_t)std::ceil(t2), setup.particles) + 2;
    return g;
}

dem_grid Settings::box_grid(float cell_size, int margin_cells) const {
    dem_grid g{};
    g.cell_size = cell_size;
    g.cell_cap = 0;
    g.w = (int32_t)std::ceil((double)simulation.hor_bound / (double)cell_size) + 2 * margin_cells;
    g.h = (int32_t)std::ceil((double)simulation.vert_bound / (double)cell_size) + 2 * margin_cells;
    g.w += g.w % 2;
    g.h += g.h % 2;
    return g;
}

void Settings::from_saved(const SavedState& s) {
    const std::vector<uint8_t>& r = s.settings_raw;
    if (r.size() != 212) throw DemError(DEM_EINVAL, "saved state: settings struct is not the 212-byte kind");
    setup.particles = get<int32_t>(r, SETUP + 0);
    setup.max_radius = get<float>(r, SETUP + 4);
    setup.variable_rad = getb(r, SETUP + 8);
    setup.holeyness = get<float>(r, SETUP + 12);
    setup.max_h_velocity = get<float>(r, SETUP + 16);
    setup.min_h_velocity = get<float>(r, SETUP + 20);
    setup.max_v_velocity = get<float>(r, SETUP + 24);
    setup.min_v_velocity = get<float>(r, SETUP + 28);
    setup.grid_width = get<float>(r, SETUP + 32);
    setup.hex_grid = getb(r, SETUP + 36);
    simulation.timestep = get<float>(r, SIM + 0);
    simulation.gen_per_frame = get<int32_t>(r, SIM + 4);
    simulation.hor_bound = get<float>(r, SIM + 12);
    simulation.vert_bound = get<float>(r, SIM + 16);
    simulation.round_walls = getb(r, SIM + 21);
    simulation.wall_radius = get<float>(r, SIM + 24);
    simulation.x_timesteps = get<int32_t>(r, SIM + 32);
    simulation.use_f64 = getb(r, SIM + 36);
    physics.gravity = getb(r, PHYS + 0);
    physics.gravity_acceleration = get<float>(r, PHYS + 4);
    physics.planet_mode = getb(r, PHYS + 8);
    physics.mouse_gravity = getb(r, PHYS + 9);
    physics.collisions = getb(r, PHYS + 10);
    physics.collision_interval = get<int32_t>(r, PHYS + 12);
    physics.friction_coefficient = get<float>(r, PHYS + 16);
    const int32_t b = get<int32_t>(r, PHYS + 20);
    physics.bonds = b <= 0 ? 0 : std::min(b, 3);   // BondType::from_i32 (settings.rs:2214)
    physics.bond_tearing = getb(r, PHYS + 24);
    physics.bond_normal_stiffness = get<float>(r, PHYS + 28);
    physics.bond_shear_stiffness = get<float>(r, PHYS + 32);
    physics.bond_normal_strength = get<float>(r, PHYS + 36);
    physics.bond_shear_strength = get<float>(r, PHYS + 40);
    physics.moment_contribution_factor = get<float>(r, PHYS + 44);
    physics.local_damping = getb(r, PHYS + 48);
    physics.local_damping_alpha = get<float>(r, PHYS + 52);
    if (const std::vector<float>* m = s.f("materials")) materials = *m;
}

std::vector<uint8_t> Settings::to_settings_struct(const std::vector<uint8_t>& base) const {
    std::vector<uint8_t> r = base.size() == 212 ? base : std::vector<uint8_t>(212, 0);
    put<int32_t>(r, SETUP + 0, setup.particles);
    put<float>(r, SETUP + 4, setup.max_radius);
    putb(r, SETUP + 8, setup.variable_rad);
    put<float>(r, SETUP + 12, setup.holeyness);
    put<float>(r, SETUP + 16, setup.max_h_velocity);
    put<float>(r, SETUP + 20, setup.min_h_velocity);
    put<float>(r, SETUP + 24, setup.max_v_velocity);
    put<float>(r, SETUP + 28, setup.min_v_velocity);
    put<float>(r, SETUP + 32, setup.grid_width);
    putb(r, SETUP + 36, setup.hex_grid);
    put<float>(r, SIM + 0, simulation.timestep);
    put<int32_t>(r, SIM + 4, simulation.gen_per_frame);
    put<float>(r, SIM + 12, simulation.hor_bound);
    put<float>(r, SIM + 16, simulation.vert_bound);
    putb(r, SIM + 21, simulation.round_walls);
    put<float>(r, SIM + 24, simulation.wall_radius);
    put<int32_t>(r, SIM + 32, simulation.x_timesteps);
    putb(r, SIM + 36, simulation.use_f64);
    putb(r, PHYS + 0, physics.gravity);
    put<float>(r, PHYS + 4, physics.gravity_acceleration);
    putb(r, PHYS + 8, physics.planet_mode);
    putb(r, PHYS + 9, physics.mouse_gravity);
    putb(r, PHYS + 10, physics.collisions);
    put<int32_t>(r, PHYS + 12, physics.collision_interval);
    put<float>(r, PHYS + 16, physics.friction_coefficient);
    put<int32_t>(r, PHYS + 20, physics.bonds);
    putb(r, PHYS + 24, physics.bond_tearing);
    put<float>(r, PHYS + 28, physics.bond_normal_stiffness);
    put<float>(r, PHYS + 32, physics.bond_shear_stiffness);
    put<float>(r, PHYS + 36, physics.bond_normal_strength);
    put<float>(r, PHYS + 40, physics.bond_shear_strength);
    put<float>(r, PHYS + 44, physics.moment_contribution_factor);
    putb(r, PHYS + 48, physics.local_damping);
    put<float>(r, PHYS + 52, physics.local_damping_alpha);
    return r;
}

State State::load(const uint8_t* buf, size_t len, Settings& settings) {
    const SavedState saved = read_state(buf, len);
    settings.from_saved(saved);
    const int32_t n = get<int32_t>(saved.settings_raw, SETUP + 0);
    auto fv = [&](const char* name) {
        const std::vector<float>* v = saved.f(name);
        if (!v) throw DemError(DEM_EINVAL, std::string("saved state: no ") + name + " vector");
        return *v;
    };
    auto iv = [&](const char* name) {
        const std::vector<int32_t>* v = saved.i(name);
        if (!v) throw DemError(DEM_EINVAL, std::string("saved state: no ") + name + " vector");
        return *v;
    };
    State st;
    st.max_contacts = settings.setup.max_contacts;
    st.p_count = n;
    st.pos = fv("pos");
    st.vel = fv("vel");
    st.acc = fv("acc");
    st.rot = fv("rot");
    st.rot_vel = fv("rot_vel");
    st.rot_acc = fv("rot_acc");
    st.forces = fv("forces");
    st.radii = fv("radii");
    st.fixity = iv("fixity");
    st.bonds = iv("bonds");
    st.contacts = fv("contacts");
    st.material_pointers = iv("material_pointers");
    const int32_t cap = (int32_t)st.radii.size();
    if (const std::vector<int32_t>* g = saved.i("groups")) st.groups = *g;
    else st.groups.assign(cap, 0);
    st.selections.assign(std::max(cap, n), 0);
    st.data.assign(7 * (size_t)std::max(cap, n), 0.0f);
    st.scripts = saved.scripts;
    st.settings_raw = saved.settings_raw;
    if (!st.radii.empty()) {
        settings.setup.min_radius = *std::min_element(st.radii.begin(), st.radii.end());
        settings.setup.max_radius = *std::max_element(st.radii.begin(), st.radii.end());
    } else {
        settings.setup.min_radius = settings.setup.max_radius = 0.0f;
    }
    settings.setup.variable_rad = settings.setup.min_radius != settings.setup.max_radius;
    return st;
}

State State::load_file(const std::string& path, Settings& settings) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw DemError(DEM_EINVAL, "cannot open " + path);
    const std::vector<uint8_t> buf((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    return load(buf.data(), buf.size(), settings);
}

std::vector<uint8_t> State::save(const Settings& settings) const {
    SavedState s;
    s.settings_raw = settings.to_settings_struct(settings_raw);
    s.fvec = {{"pos", pos},         {"vel", vel},         {"acc", acc},         {"rot", rot},
              {"rot_vel", rot_vel}, {"rot_acc", rot_acc}, {"forces", forces},   {"radii", radii},
              {"contacts", contacts}, {"materials", settings.materials}};
    s.ivec = {{"fixity", fixity}, {"bonds", bonds}, {"material_pointers", material_pointers}, {"groups", groups}};
    s.has_scripts = true;
    s.scripts = scripts;
    return write_state(s);
}

bool State::get_datum(double out[10]) const {
    double sums[10] = {0};
    int64_t count = 0;
    const int32_t cap = capacity();
    for (int32_t i = 0; i < cap && i < (int32_t)selections.size(); i++) {
        if (selections[i] == 0) continue;
        const float* d = data.data() + 7 * (size_t)i;
        const double v[10] = {pos[2 * i], pos[2 * i + 1], d[4], d[5], rot[i], d[6], std::fabs((double)d[0]),
                              d[1], d[2], d[3]};
        for (int k = 0; k < 10; k++) sums[k] += v[k];
        count++;
    }
    if (count == 0) return false;
    for (int k = 0; k < 10; k++) out[k] = sums[k] / (double)count;
    out[6] = sums[6];
    return true;
}

dem_state_view State::view(bool with_data) {
    dem_state_view v{};
    v.n = p_count;
    v.pos = pos.data();
    v.vel = vel.data();
    v.acc = acc.data();
    v.rot = rot.data();
    v.rot_vel = rot_vel.data();
    v.rot_acc = rot_acc.data();
    v.forces = forces.data();
    v.radii = radii.data();
    v.fixity = fixity.data();
    v.material_pointers = material_pointers.data();
    v.bonds = bonds.data();
    v.n_bonds = n_bonds();
    v.bond_strengths = bond_strengths.empty() ? nullptr : bond_strengths.data();
    v.contacts = contacts.data();
    v.contact_slots = max_contacts;
    v.data = with_data ? data.data() : nullptr;
    return v;
}

}  // namespace host
}  // namespace psim
