// psim_host.h - C++ host mirror of the reference's State / Settings / compute program
// (src/state.rs, src/settings.rs, src/wgpu_prog.rs) over the C ABI of include/dem.h.
//
// The reference's host code is Rust; this mirror keeps its names and argument
// meaning so a C++ caller drives the HIP time step the way the reference's client
// drives WGPUComputeProg:
//
//   Settings                 settings.rs:102-254 (the parts the time step reads)
//     collision_settings()   settings.rs:2086-2119 (29-word physics uniform)
//     grid_info()            settings.rs:1955-1974 (reference grid sizing)
//   State                    state.rs:69-93, load state.rs:583-706, save state.rs:369-493,
//                            get_datum state.rs:708-737
//   ComputeProg              WGPUComputeProg (wgpu_prog.rs:677-696): new, compute,
//                            update_state, restore, regen_bonds (state.rs:291-367)
//
// Errors: the reference panics on a wgpu failure; here every failing call throws
// DemError carrying the C ABI's status code and dem_last_error() text.
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/dem.h"

namespace psim {
namespace host {

struct DemError : std::runtime_error {
    int code;
    DemError(int c, const std::string& msg) : std::runtime_error(msg), code(c) {}
};

// ---------------------------------------------------------------- saved states
// The flatbuffer `State` table (new_schema.fbs:1-105; the older schema.fbs, without
// scripts and with a 204-byte Settings struct, is read too).
struct SavedState {
    // float vectors: pos vel acc rot rot_vel rot_acc forces radii contacts materials
    // int vectors:   fixity bonds material_pointers groups
    std::vector<std::pair<std::string, std::vector<float>>> fvec;
    std::vector<std::pair<std::string, std::vector<int32_t>>> ivec;
    bool has_scripts = false;
    std::string scripts;
    std::vector<uint8_t> settings_raw;   // 212 bytes (new schema) or 204 (old)
    bool old_schema = false;

    const std::vector<float>* f(const std::string& name) const;
    const std::vector<int32_t>* i(const std::string& name) const;
};

SavedState read_state(const uint8_t* buf, size_t len);
// Byte layout of the reference writer (state.rs:369-493): vectors in creation order,
// the table's fields added settings first, then in reverse schema order.
std::vector<uint8_t> write_state(const SavedState& s);

// ---------------------------------------------------------------- settings
struct SetupSettings {         // settings.rs:139-156, defaults :300-317
    int32_t particles = 256;
    float max_radius = 0.025f;
    float min_radius = 0.025f / 1.7f;
    bool variable_rad = false;
    float holeyness = 1.7f;
    int32_t max_bonds = 6;
    int32_t max_contacts = 14;
    float max_h_velocity = 0, min_h_velocity = 0, max_v_velocity = 0, min_v_velocity = 0;
    float grid_width = 32.0f;
    bool hex_grid = false;
};

struct SimulationSettings {    // settings.rs:158-174, defaults :319-335
    float timestep = 1.0f / 12600.0f;
    int32_t gen_per_frame = 105;
    bool walls = true;
    float hor_bound = 2.0f * 1.333f;
    float vert_bound = 2.0f;
    bool round_walls = false;
    float wall_radius = 1.0f;
    bool use_f64 = false;
    int32_t x_timesteps = 1;
};

struct PhysicsSettings {       // settings.rs:176-197, defaults :336-357
    bool gravity = true;
    float gravity_acceleration = 1.0f;
    bool planet_mode = true;
    bool mouse_gravity = false;
    bool collisions = true;
    int32_t collision_interval = 1;
    float friction_coefficient = 0.5f;
    int32_t bonds = 0;
    bool bond_tearing = false;
    float bond_normal_stiffness = 10.0f, bond_shear_stiffness = 10.0f;
    float bond_normal_strength = 0.5f, bond_shear_strength = 0.5f;
    float contact_damping = 0.2f, bond_damping = 0.2f, drag = 1.0f;
    float moment_contribution_factor = 1.0f;
    bool local_damping = false;
    float local_damping_alpha = 0.1f;
    // extension, not saved: restitution of the contact dashpot (dem_params.restitution);
    // 0 = none, the reference's contact law
    float restitution = 0.0f;
};

struct Settings {
    SetupSettings setup;
    SimulationSettings simulation;
    PhysicsSettings physics;
    std::vector<float> materials{1.0f, 1.0f, 1.0f, 0.01f, 100.0f, 50.0f,
                                 1.0f, 0.0f, 0.0f, 0.01f, 100.0f, 50.0f};   // settings.rs:266
    float wall_friction = 0.0f;
    float world_pos[2] = {0.0f, 0.0f};

    dem_uniform collision_settings() const;
    dem_grid grid_info() const;
    // A grid over the walled box plus `margin_cells` on each side (even w and h).
    dem_grid box_grid(float cell_size, int margin_cells = 1) const;
    void from_saved(const SavedState& s);
    std::vector<uint8_t> to_settings_struct(const std::vector<uint8_t>& base) const;
};

// ---------------------------------------------------------------- state
struct State {
    int32_t p_count = 0;
    int32_t max_contacts = 14;
    std::vector<float> pos, vel, acc, rot, rot_vel, rot_acc, forces, radii, contacts, data;
    std::vector<int32_t> fixity, bonds, material_pointers, groups, selections;
    std::vector<float> bond_strengths;   // optional {normal, shear} per bond (extension)
    std::string scripts;
    std::vector<uint8_t> settings_raw;

    int32_t capacity() const { return (int32_t)radii.size(); }
    int32_t n_bonds() const { return bonds.size() < 3 ? 0 : (int32_t)(bonds.size() / 3); }

    // State::load: arrays at their stored length, p_count from the stored settings,
    // min/max radius recomputed; applies the stored settings to `settings`.
    static State load(const uint8_t* buf, size_t len, Settings& settings);
    static State load_file(const std::string& path, Settings& settings);
    std::vector<uint8_t> save(const Settings& settings) const;
    // State::get_datum over `selections`: false when nothing is selected.
    bool get_datum(double out[10]) const;
    dem_state_view view(bool with_data = true);
};

// ---------------------------------------------------------------- compute
float min_cell_size(float rmax, float extent);

class ComputeProg {
public:
    struct Options {
        int32_t capacity = 0;        // 0: the state's particle count
        int32_t bond_capacity = 0;   // 0: the state's bond count
        bool record_data = true;
        int32_t reorder_interval = 64;
        int32_t block_size = 256;
        float cell_size = 0.0f;      // 0: just above 2 r_max (default_grid)
        int use_f64 = -1;            // -1: settings.simulation.use_f64
        bool graph = false;          // replay repeated dem_step launch sequences as HIP graphs
    };

    // WGPUComputeProg::new (buffers) + restore
    ComputeProg(const Settings& settings, State& state, const Options& opt);
    ComputeProg(const Settings& settings, State& state) : ComputeProg(settings, state, Options()) {}
    ~ComputeProg();
    ComputeProg(const ComputeProg&) = delete;
    ComputeProg& operator=(const ComputeProg&) = delete;

    static dem_grid default_grid(const Settings& settings, const State& state, float cell_size = 0.0f);

    void compute(const Settings& settings);          // gen_per_frame steps
    void step(int32_t n);
    void sync();
    State& update_state();                            // download into the bound State
    void restore(State& state);                       // upload (binds `state`)
    void set_params(const Settings& settings);
    double kinetic_energy();
    int64_t broken_bond_count();
    // State::regen_bonds with the pair search on the device (dem_build_bond_pairs):
    // bonded iff |d| < (ri + rj) * factor + gap (the reference: factor 1.05, gap 0)
    State& regen_bonds(const Settings& settings, float factor = 1.05f, float gap = 0.0f);

    const dem_grid& grid() const { return grid_; }

private:
    dem_params params(const Settings& s);
    dem_ctx* ctx_ = nullptr;
    State* state_ = nullptr;
    dem_grid grid_{};
    Options opt_;
    std::vector<dem_material> mats_;
    int32_t bond_capacity_ = 0;
    bool use_f64_ = false;
};

// Bond triples and endpoint slots for (i, j) pairs in regen_bonds order
// (state.rs:306-330): {1, atan2(nx, ny) bits, |d| bits}; each endpoint's first free
// slot gets {a, b, 0, 0, 0, bond index}.
void assemble_bonds(const State& st, const std::vector<int32_t>& pairs, int32_t slots, std::vector<int32_t>& bonds,
                    std::vector<float>& contacts);

}  // namespace host
}  // namespace psim
