// dem_engine.h - host side of the DEM library: device pools, scenes and the step driver.
//
// An Engine owns one set of device pools on one HIP stream and advances any number
// of independent scenes in the same launches.  The C ABI's dem_ctx is an Engine
// holding a single scene; the serving loop uses one Engine for many scenes.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "dem_device.h"

namespace psim {

void set_error(const std::string& msg);
void set_error(const char* msg);

// Large bodies of a radius set (rule in include/dem.h, dem_grid): their count, and the
// largest radius of the other discs in *rsmax.
int32_t split_large(const float* radii, int32_t n, float* rsmax);
// Parameter checks shared by scene creation, dem_set_params and the serving queue.
int validate_params(const dem_params& p);
// Whether the grid's cells are wide enough for the 3x3 stencil: 2 rsmax + the rounding
// margin of the cell index (4 ulp of the grid extent).
int check_cell_size(const dem_grid& g, float rsmax);
// 2 zeta of the contact dashpot for restitution e (dem_params.restitution); 0: none.
float dashpot_factor(float e);

struct DevSet {
    float4* PV[2] = {nullptr, nullptr};   // by parity: per slot {position, velocity record}, 32 B
    float4* P[2] = {nullptr, nullptr};    // PV[k]: positions (stride 32 B), by position parity (cur_)
    float4* V[2] = {nullptr, nullptr};    // PV[k] + 16 B: velocity records, by list parity (lcur_ == cur_)
    float4* DX[2] = {nullptr, nullptr};
    float4* A = nullptr;
    float* TH = nullptr;     // rotation (read and written by the particle's own thread only)
    int32_t* M = nullptr;    // material | flags | contact count
    int32_t* ID = nullptr;   // particle id within the scene
    float4* EXT = nullptr;
    float4* DATA0 = nullptr;
    float4* DATA1 = nullptr;
    int32_t* NL = nullptr;   // neighbour lists [NLMAX][cap], row 0 packs the count
    float2* XR = nullptr;    // positions at the list build
    int2* BRS = nullptr;     // bond record range of each slot
};

struct ContactLists {
    int32_t* CP = nullptr;
    float* CF = nullptr;
};

struct SceneHost {
    dem_params params{};
    std::vector<dem_material> mats;
    bool live = false;
    // pool regions
    int64_t p0 = 0;
    int32_t np = 0;          // slots in use (multiple of the block size)
    int32_t slot_cap = 0;    // slots reserved
    int32_t n = 0;           // particles
    int64_t c0 = 0;
    int32_t cell_cap = 0;    // cells reserved (w*h + 1 must fit)
    int64_t b0 = 0;
    int32_t nb = 0, bond_cap = 0;
    int64_t r0 = 0;
    int32_t nrec = 0, rec_cap = 0;
    int32_t m0 = 0, nm = 0, mat_cap = 0;
    // host copies of arrays the step never changes
    std::vector<int32_t> fixity;
    std::vector<float> forces;
    std::vector<int32_t> matptr;
    std::vector<int32_t> bond_angle, bond_len;
    bool has_ext = false;
    bool has_strength = false;
    // coll_cont words (2D_LOM.wgsl:154-163, 2D_Simulation.wgsl:258-261)
    int32_t coll0 = 0, coll1 = 0;
    bool del_override = false;
    bool r0_zero = false;    // particle 0 has radius 0: invocation 0 never updates coll_cont
    float rmax = 0.0f;       // largest radius (set at upload; radii never change in a run)
    float rsmax = 0.0f;      // largest radius below the large bodies (rmax without them)
    std::vector<int32_t> large;   // particle ids of the large bodies
    int64_t step = 0;        // Sim halves done since creation (bond-break epochs)
    // counters
    int64_t steps = 0, broad_phases = 0, reorders = 0;
    int32_t since_reorder = 1 << 30;
};

// One planned launch of a step sequence.  Operations are built zero-filled (padding
// included) so that two plans compare bytewise: a repeated plan replays its graph.
enum OpKind { OP_STEP, OP_RB_STORE, OP_RB_BIN, OP_SCAN, OP_RB_SCATTER, OP_BUILD_NL, OP_PARTITION, OP_MOVE, OP_BONDS,
              OP_RSLOTS, OP_LARGE };
static const char* const kOpNames[] = {"step kernel",      "rebuild: flag",  "rebuild: bin",
                                       "rebuild: scan",    "rebuild: scatter", "rebuild: lists",
                                       "re-sort: partition", "re-sort: move", "bond pass",
                                       "re-sort: record slots", "large-body search"};
struct LaunchOp {
    int32_t kind, timed;
    int32_t grid, block, nlb, ncells, value, f64, bonds;
    int32_t* rb;
    const int32_t* gate;
    StepArgs sa;
    ReorderArgs ra;
    DevArrays da;
};

struct GraphEntry {
    std::vector<LaunchOp> ops;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t used = 0;
};

class Engine {
public:
    Engine() = default;
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    // Pools: particle slots, cells, bonds, bond records, materials, scenes.
    int init(int64_t slot_cap, int64_t cell_cap, int64_t bond_cap, int64_t rec_cap, int32_t mat_cap,
             int32_t scene_cap, int32_t block, bool with_data);

    // Scene lifetime (regions carved from the pools, first fit).
    int add_scene(const dem_params& p, int32_t capacity, int32_t bond_cap, int32_t rec_cap, int32_t* sid_out);
    int remove_scene(int32_t sid);
    int set_params(int32_t sid, const dem_params& p);
    int upload(int32_t sid, const dem_state_view& v);
    int download(int32_t sid, dem_state_view& v);

    // Advances the scenes in `sids` by `nsteps[i]` steps each in shared launches;
    // scenes that finish early sit out the remaining launches.
    int step(const std::vector<int32_t>& sids, const std::vector<int32_t>& nsteps);

    int kinetic_energy(int32_t sid, double* out);
    // the same sum without allocating: partial results in the library's static device
    // scratch (at most DEM_KE_NOALLOC_MAX slots), reduced on the device in a fixed order
    int kinetic_energy_noalloc(int32_t sid, double* out);
    int broken_bonds(int32_t sid, int64_t* out);
    int counters(int32_t sid, int64_t* out, int32_t n_out);
    int datum(int32_t sid, const int32_t* select, double* out);
    int debug_broad_phase(int32_t sid, int32_t* keys, int32_t* sorted_ids, int32_t* cell_start);
    int debug_contacts(int32_t sid, int32_t max_per, int32_t* counts, int32_t* partners);
    int bond_pairs(int32_t sid, int32_t max_bonds, float factor, float gap, int32_t* pairs, int32_t cap,
                   int32_t* n_out);
    int group_sums(int32_t sid, const int32_t* ids, int32_t n_ids, double* out);
    int sync();

    // timing of the last step() call
    void enable_timing(bool on) { timing_ = on; }
    // capture step sequences in HIP graphs and replay them when a sequence repeats
    void enable_graphs(bool on) {
        if (!on) clear_graphs();
        use_graph_ = on;
    }
    int64_t graph_captures() const { return graph_captures_; }
    int64_t graph_replays() const { return graph_replays_; }
    double t_total = 0, t_step = 0, t_scan = 0, t_scatter = 0, t_reorder = 0, t_bonds = 0;
    int64_t n_step = 0, n_scan = 0, n_scatter = 0, n_reorder = 0, n_bonds = 0;

    hipStream_t stream() const { return stream_; }
    int device() const { return device_; }
    int block() const { return block_; }
    const SceneHost& scene(int32_t sid) const { return scenes_[sid]; }
    int32_t scene_count() const { return (int32_t)scenes_.size(); }
    int64_t slot_cap() const { return slot_cap_; }
    int32_t material_capacity() const { return mat_cap_; }
    int64_t free_slots() const;

private:
    int check(hipError_t e, const char* what);
    DevArrays arrays(int set) const;
    StepArgs make_args(const int2* lf) const;
    int write_scene_table();
    void plan_reorder(const int2* lf_dev, std::vector<LaunchOp>& ops);
    static LaunchOp& new_op(std::vector<LaunchOp>& ops, int kind, int timed);
    hipError_t run_op(const LaunchOp& op, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
    int run_ops(const std::vector<LaunchOp>& ops);
    void clear_graphs();
    int rebuild_chain(const int2* lf_dev, bool resort);
    float skin_of(const SceneHost& sh) const;
    int launch_grid_blocks() const;   // blocks covering the whole slot pool
    int active_blocks() const;        // blocks covering the slots below the high-water mark
    void update_high_water();
    int alloc_region(std::vector<std::pair<int64_t, int64_t>>& used, int64_t cap, int64_t need, int64_t align,
                     int64_t* out);
    SceneDev scene_dev(const SceneHost& sh, int32_t sid) const;

    int device_ = 0;
    hipStream_t stream_ = nullptr;
    int block_ = 256;
    int cus_ = 256;   // compute units of the device
    // pinned host staging for uploads and downloads (grown on demand)
    char* stage_ = nullptr;
    size_t stage_cap_ = 0;
    char* staging(size_t bytes);
    bool with_data_ = true;
    int64_t slot_cap_ = 0, cell_cap_ = 0, bond_cap_ = 0, rec_cap_ = 0;
    // ends of the highest live slot and cell regions: launches and the cell scan stop there
    int64_t hw_slots_ = 0, hw_cells_ = 0;
    int32_t mat_cap_ = 0, scene_cap_ = 0;
    int set_ = 0;   // which DevSet is current
    int cur_ = 0;   // which position buffer is current (flips with lcur_ after every launch)
    DevSet sets_[2];
    ContactLists lists_[2];
    int lcur_ = 0;  // which contact-list (and velocity) buffer is current
    int2* BOFF_ = nullptr;      // per slot (scene p0 + particle id): bond record range [x, y)
    int4* BR_ = nullptr;
    float4* BF_ = nullptr;
    int2* BS_ = nullptr;
    float4* BLS_ = nullptr;
    int2* BRB_ = nullptr;
    dem_material* MAT_ = nullptr;
    int32_t *KEY_ = nullptr, *RANK_ = nullptr, *CNT_ = nullptr, *START_ = nullptr, *SIDX_ = nullptr, *INV_ = nullptr;
    int32_t* RB_ = nullptr;
    float* FT_ = nullptr;     // [KMAX][cap] tangential forces of list candidates (PF_ALIGNED slots)
    int32_t* LB_ = nullptr;   // large-body ids, LB_MAX per scene
    int32_t* LBH_ = nullptr;  // small discs overlapping each large body (LBH_MAX per large body)
    int32_t* LBHC_ = nullptr; // their counts
    int32_t* RO_ = nullptr;   // bond pass: record owners, their slots and results
    int32_t *RS_ = nullptr, *RP_ = nullptr;
    float4 *RF_ = nullptr, *RD_ = nullptr;
    int64_t hw_recs_ = 0;     // end of the highest live record region
    int32_t *ORDER_ = nullptr, *NEWSLOT_ = nullptr;   // re-sort permutation
    bool rb_force_ = true;        // rebuild the neighbour lists after the next drift
    unsigned long long* STAT_ = nullptr;
    int32_t* TSUM_ = nullptr;
    int32_t ntiles_ = 0;
    SceneDev* d_scenes_ = nullptr;
    int32_t* d_blk_scene_ = nullptr;
    int2* d_lf_ = nullptr;        // launch flags ring
    int64_t lf_cap_ = 0;          // entries (launches * scene_cap)
    double* d_ke_ = nullptr;      // per-id kinetic energy scratch
    // group sums: 64 x 7 block partials + 7 totals (GROUP_SCRATCH_DOUBLES), then the ids
    static constexpr int GROUP_SCRATCH_DOUBLES = 64 * 7 + 8;
    void* d_group_ = nullptr;
    int64_t group_bytes_ = 0;
    unsigned long long* d_cnt_ = nullptr;
    std::vector<SceneHost> scenes_;
    std::vector<std::pair<int64_t, int64_t>> used_slots_, used_cells_, used_bonds_, used_recs_, used_mats_;
    bool timing_ = false;
    // graphs of step sequences (at most kGraphCache, least recently used dropped)
    static constexpr size_t kGraphMaxOps = 4096;
    static constexpr size_t kGraphCache = 32;
    bool use_graph_ = false;
    std::vector<GraphEntry> graphs_;
    uint64_t graph_clock_ = 0;
    int64_t graph_captures_ = 0, graph_replays_ = 0;
    std::vector<hipEvent_t> ev_pool_;
    bool table_dirty_ = true;
};

}  // namespace psim
