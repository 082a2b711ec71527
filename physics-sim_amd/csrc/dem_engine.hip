// dem_engine.hip - pools, scene regions, upload/download and the step driver.
#include "dem_engine.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <numeric>

namespace psim {

hipError_t launch_step(const StepArgs& args, int nblocks, int block, bool f64, bool bonds, bool large,
                       hipStream_t st, hipEvent_t ev0, hipEvent_t ev1);
hipError_t launch_bin(const StepArgs& args, int nblocks, int block, hipStream_t st);
int scan_tiles(int n);
hipError_t launch_scan(const DevArrays& a, int ncells, const int32_t* gate, hipStream_t st,
                       hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
hipError_t launch_rb_bin(const StepArgs& args, int nlb, int grid, int block, hipStream_t st,
                         hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
hipError_t launch_rb_scatter(const StepArgs& args, int nlb, int grid, int block, hipStream_t st,
                             hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
hipError_t launch_build_nl(const StepArgs& args, int nlb, int grid, int block, hipStream_t st,
                           hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
hipError_t launch_rb_store(int32_t* rb, int v, hipStream_t st,
                           hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);

hipError_t launch_scatter(const StepArgs& args, int nblocks, int block, hipStream_t st);
hipError_t launch_cell_sort(const StepArgs& args, int cell_lo, int cell_hi, int p0, hipStream_t st);
hipError_t launch_reorder_move(const ReorderArgs& r, int nblocks, int block, hipStream_t st,
                               hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
hipError_t launch_partition(const ReorderArgs& r, int nblocks, int block, hipStream_t st,
                            hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
hipError_t launch_bond_forces(const StepArgs& args, int nrec, int lbs, bool f64, hipStream_t st,
                              hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
hipError_t launch_record_slots(const StepArgs& args, int nrec, int lbs, hipStream_t st,
                               hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
hipError_t launch_kinetic(const DevArrays& a, int cur, int m0, int nm, int p0, int np, double* out, hipStream_t st);
hipError_t launch_count_contacts(const int32_t* m, int p0, int np, unsigned long long* out, hipStream_t st);
hipError_t launch_contacts_by_id(const StepArgs& args, int p0, int np, hipStream_t st);
hipError_t launch_bond_pairs(const StepArgs& args, int nblocks, int block, int reach, int maxb, float factor,
                             float gap, int32_t* out_cnt, int32_t* out_j, hipStream_t st);
hipError_t launch_large_search(const StepArgs& args, int nblocks, hipStream_t st, hipEvent_t ev0 = nullptr,
                               hipEvent_t ev1 = nullptr);
hipError_t launch_group_sums(const DevArrays& a, int cur, int m0, int nm, int p0, const int32_t* ids, int n_ids,
                             int with_data, double* partials, double* out, hipStream_t st);
hipError_t launch_ke_reduce(double* buf, int n, double* host_out, hipStream_t st);
double* static_scratch(int device);

namespace {
inline int32_t fbits(float f) { int32_t i; std::memcpy(&i, &f, 4); return i; }
inline float bitsf(int32_t i) { float f; std::memcpy(&f, &i, 4); return f; }
// Host view of one half of interleaved {position, velocity} records (staging buffers
// laid out as on the device, so that each copy is one contiguous transfer).
struct PVHalf {
    float4* b;
    float4& operator[](size_t i) const { return b[2 * i]; }
};
inline int64_t round_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

}  // namespace

// Large bodies of a radius set: with the radii in decreasing order r_0 >= r_1 >= ...,
// the first k <= LB_MAX with r_{k-1} > 2 r_k makes the top k discs large bodies and
// r_k the largest small radius; without such a gap there are none (k = 0, rsmax = r_0).
int32_t split_large(const float* radii, int32_t n, float* rsmax) {
    std::vector<float> top;
    top.reserve(LB_MAX + 2);
    for (int32_t i = 0; i < n; i++) {
        if (!(radii[i] > 0.0f)) continue;
        top.push_back(radii[i]);
        std::push_heap(top.begin(), top.end(), std::greater<float>());   // min-heap of the largest
        if ((int32_t)top.size() > LB_MAX + 1) {
            std::pop_heap(top.begin(), top.end(), std::greater<float>());
            top.pop_back();
        }
    }
    std::sort(top.begin(), top.end(), std::greater<float>());
    *rsmax = top.empty() ? 0.0f : top[0];
    for (size_t k = 1; k < top.size(); k++)
        if (top[k - 1] > 2.0f * top[k]) {
            *rsmax = top[k];
            return (int32_t)k;
        }
    return 0;
}

namespace {

// Carves typed, 256-byte aligned arrays out of one buffer; with a null base it only
// adds up the size.
struct Carve {
    char* base = nullptr;
    size_t off = 0;
    template <class T>
    T* take(size_t count) {
        off = (off + 255) & ~(size_t)255;
        T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
        off += count * sizeof(T);
        return p;
    }
};

template <typename T>
hipError_t dalloc(T** p, int64_t n) {
    *p = nullptr;
    if (n <= 0) n = 1;
    return hipMalloc((void**)p, sizeof(T) * (size_t)n);
}

struct Slot {  // 2D_Simulation.wgsl:19-26
    int32_t a, b;
    float ft, fbt, tb;
    int32_t bonded;
};
}  // namespace

Engine::~Engine() {
    if (stream_) hipStreamSynchronize(stream_);
    clear_graphs();   // before the buffers their nodes refer to
    for (auto& s : sets_) {
        hipFree(s.PV[0]); hipFree(s.PV[1]); hipFree(s.DX[0]); hipFree(s.DX[1]);
        hipFree(s.A); hipFree(s.TH); hipFree(s.M); hipFree(s.ID); hipFree(s.EXT); hipFree(s.DATA0); hipFree(s.DATA1);
    }
    for (auto& s : sets_) { hipFree(s.NL); hipFree(s.XR); hipFree(s.BRS); }
    for (auto& l : lists_) { hipFree(l.CP); hipFree(l.CF); }
    hipFree(FT_);
    hipFree(BOFF_); hipFree(BR_); hipFree(BF_); hipFree(BS_); hipFree(BLS_); hipFree(BRB_); hipFree(MAT_);
    hipFree(KEY_); hipFree(RANK_); hipFree(CNT_); hipFree(START_); hipFree(SIDX_); hipFree(INV_);
    hipFree(RB_); hipFree(ORDER_); hipFree(NEWSLOT_); hipFree(RO_); hipFree(RF_); hipFree(RD_);
    hipFree(RS_); hipFree(RP_); hipFree(LB_); hipFree(LBH_); hipFree(LBHC_);
    hipFree(STAT_); hipFree(TSUM_); hipFree(d_scenes_); hipFree(d_blk_scene_); hipFree(d_lf_);
    hipFree(d_ke_); hipFree(d_cnt_);
    if (d_group_) hipFree(d_group_);
    if (stage_) hipHostFree(stage_);
    for (auto e : ev_pool_) hipEventDestroy(e);
    if (stream_) hipStreamDestroy(stream_);
}

int Engine::check(hipError_t e, const char* what) {
    if (e == hipSuccess) return DEM_OK;
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? DEM_ENOMEM : DEM_EDEVICE;
}

int Engine::init(int64_t slot_cap, int64_t cell_cap, int64_t bond_cap, int64_t rec_cap, int32_t mat_cap,
                 int32_t scene_cap, int32_t block, bool with_data) {
    if (block <= 0) block = 256;
    if (block % 64 || block > 256) {
        set_error("block_size must be a multiple of 64 in [64, 256]");
        return DEM_EINVAL;
    }
    block_ = block;
    with_data_ = with_data;
    slot_cap_ = round_up(std::max<int64_t>(slot_cap, block), block);
    cell_cap_ = std::max<int64_t>(cell_cap, 2);
    bond_cap_ = std::max<int64_t>(bond_cap, 1);
    rec_cap_ = std::max<int64_t>(rec_cap, 1);
    mat_cap_ = std::max(mat_cap, 1);
    scene_cap_ = std::max(scene_cap, 1);
    // kernels address every pool with 32-bit byte offsets: each must stay below 4 GiB
    // (the contact lists, KMAX words per slot, bound the particle count at 2^26)
    const int64_t lim = (int64_t)1 << 32;
    if (slot_cap_ * KMAX * 4 >= lim || slot_cap_ * NLMAX * 4 >= lim || slot_cap_ * 16 >= lim || cell_cap_ * 4 + 8 >= lim ||
        rec_cap_ * 16 >= lim || bond_cap_ * 8 >= lim) {
        set_error("pool too large for 32-bit offsets (at most 2^26 - 1 particle slots per context)");
        return DEM_EINVAL;
    }
    int rc;
    if ((rc = check(hipGetDevice(&device_), "hipGetDevice"))) return rc;
    if ((rc = check(hipDeviceGetAttribute(&cus_, hipDeviceAttributeMultiprocessorCount, device_), "CU count")))
        return rc;
    if ((rc = check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate"))) return rc;
    const int64_t S = slot_cap_;
    for (auto& s : sets_) {
        hipError_t e = hipSuccess;
        e = e ? e : dalloc(&s.PV[0], 2 * S); e = e ? e : dalloc(&s.PV[1], 2 * S);
        for (int k = 0; k < 2; k++) {
            s.P[k] = s.PV[k];
            s.V[k] = s.PV[k] ? s.PV[k] + 1 : nullptr;
        }
        e = e ? e : dalloc(&s.DX[0], S); e = e ? e : dalloc(&s.DX[1], S);
        e = e ? e : dalloc(&s.A, S); e = e ? e : dalloc(&s.TH, S); e = e ? e : dalloc(&s.M, S); e = e ? e : dalloc(&s.ID, S);
        e = e ? e : dalloc(&s.EXT, S);
        e = e ? e : dalloc(&s.NL, S * NLMAX); e = e ? e : dalloc(&s.XR, S); e = e ? e : dalloc(&s.BRS, S);
        if (with_data) { e = e ? e : dalloc(&s.DATA0, S); e = e ? e : dalloc(&s.DATA1, S); }
        if ((rc = check(e, "hipMalloc particle pools"))) return rc;
    }
    for (auto& l : lists_) {
        hipError_t e = hipSuccess;
        e = e ? e : dalloc(&l.CP, S * KMAX); e = e ? e : dalloc(&l.CF, S * KMAX);
        if ((rc = check(e, "hipMalloc contact lists"))) return rc;
    }
    if ((rc = check(dalloc(&FT_, S * KMAX), "hipMalloc contact lists"))) return rc;
    hipError_t e = hipSuccess;
    e = e ? e : dalloc(&BOFF_, S); e = e ? e : dalloc(&BR_, rec_cap_); e = e ? e : dalloc(&BF_, rec_cap_);
    e = e ? e : dalloc(&RO_, rec_cap_); e = e ? e : dalloc(&RF_, rec_cap_);
    e = e ? e : dalloc(&RS_, rec_cap_); e = e ? e : dalloc(&RP_, rec_cap_);
    e = e ? e : dalloc(&RD_, with_data ? 2 * rec_cap_ : 2);
    e = e ? e : dalloc(&BS_, bond_cap_); e = e ? e : dalloc(&BLS_, bond_cap_); e = e ? e : dalloc(&BRB_, rec_cap_);
    e = e ? e : dalloc(&MAT_, mat_cap_);
    e = e ? e : dalloc(&KEY_, S); e = e ? e : dalloc(&RANK_, S); e = e ? e : dalloc(&CNT_, cell_cap_ + 1);
    e = e ? e : dalloc(&START_, cell_cap_ + 2); e = e ? e : dalloc(&SIDX_, S); e = e ? e : dalloc(&INV_, S);
    e = e ? e : dalloc(&ORDER_, S); e = e ? e : dalloc(&NEWSLOT_, S);
    e = e ? e : dalloc(&RB_, 2);
    e = e ? e : dalloc(&STAT_, 8);
    ntiles_ = scan_tiles((int)cell_cap_);
    e = e ? e : dalloc(&TSUM_, ntiles_ + 1);
    e = e ? e : dalloc(&d_scenes_, scene_cap_);
    e = e ? e : dalloc(&LB_, (int64_t)scene_cap_ * LB_MAX);
    e = e ? e : dalloc(&LBH_, (int64_t)scene_cap_ * LB_MAX * LBH_MAX);
    e = e ? e : dalloc(&LBHC_, (int64_t)scene_cap_ * LB_MAX);
    const int64_t nblk = launch_grid_blocks();
    e = e ? e : dalloc(&d_blk_scene_, nblk);
    lf_cap_ = (int64_t)scene_cap_ * 4096;
    e = e ? e : dalloc(&d_lf_, lf_cap_);
    e = e ? e : dalloc(&d_ke_, S);
    e = e ? e : dalloc(&d_cnt_, 4);
    if ((rc = check(e, "hipMalloc pools"))) return rc;
    e = hipMemsetAsync(CNT_, 0, sizeof(int32_t) * (cell_cap_ + 1), stream_);
    e = e ? e : hipMemsetAsync(START_, 0, sizeof(int32_t) * (cell_cap_ + 2), stream_);
    e = e ? e : hipMemsetAsync(STAT_, 0, sizeof(unsigned long long) * 8, stream_);
    // the binning accumulates tile totals; the slot lists start as valid slot indices
    // (every index read from them stays inside the pools even before the first build)
    e = e ? e : hipMemsetAsync(TSUM_, 0, sizeof(int32_t) * (ntiles_ + 1), stream_);
    e = e ? e : hipMemsetAsync(SIDX_, 0, sizeof(int32_t) * S, stream_);
    e = e ? e : hipMemsetAsync(ORDER_, 0, sizeof(int32_t) * S, stream_);
    e = e ? e : hipMemsetAsync(NEWSLOT_, 0, sizeof(int32_t) * S, stream_);
    e = e ? e : hipMemsetAsync(RB_, 0, 2 * sizeof(int32_t), stream_);
    e = e ? e : hipMemsetAsync(FT_, 0, sizeof(float) * S * KMAX, stream_);
    e = e ? e : hipMemsetAsync(RO_, 0xff, sizeof(int32_t) * rec_cap_, stream_);   // no records
    e = e ? e : hipMemsetAsync(RS_, 0xff, sizeof(int32_t) * rec_cap_, stream_);
    e = e ? e : hipMemsetAsync(RP_, 0, sizeof(int32_t) * rec_cap_, stream_);
    for (auto& st : sets_) e = e ? e : hipMemsetAsync(st.BRS, 0, sizeof(int2) * S, stream_);
    for (auto& st : sets_) e = e ? e : hipMemsetAsync(st.NL, 0, sizeof(int32_t) * S, stream_);   // empty lists
    e = e ? e : hipMemsetAsync(BOFF_, 0, sizeof(int2) * S, stream_);
    for (auto& s : sets_) e = e ? e : hipMemsetAsync(s.M, 0xff, sizeof(int32_t) * S, stream_);
    for (auto& s : sets_) e = e ? e : hipMemsetAsync(s.ID, 0xff, sizeof(int32_t) * S, stream_);
    for (auto& s : sets_)
        for (int b = 0; b < 2; b++) {
            e = e ? e : hipMemsetAsync(s.PV[b], 0, 2 * sizeof(float4) * S, stream_);
            e = e ? e : hipMemsetAsync(s.DX[b], 0, sizeof(float4) * S, stream_);
        }
    std::vector<int32_t> blk(nblk, -1);
    e = e ? e : hipMemcpyAsync(d_blk_scene_, blk.data(), sizeof(int32_t) * nblk, hipMemcpyHostToDevice, stream_);
    e = e ? e : hipStreamSynchronize(stream_);
    return check(e, "pool init");
}

char* Engine::staging(size_t bytes) {
    if (bytes > stage_cap_) {
        // every earlier copy from or into the buffer has completed (each call syncs)
        if (hipStreamSynchronize(stream_) != hipSuccess) return nullptr;
        if (stage_) hipHostFree(stage_);
        stage_ = nullptr;
        stage_cap_ = 0;
        const size_t cap = std::max(bytes, (size_t)(1 << 20));
        if (hipHostMalloc((void**)&stage_, cap, hipHostMallocDefault) != hipSuccess) {
            stage_ = nullptr;
            return nullptr;
        }
        stage_cap_ = cap;
    }
    return stage_;
}

int Engine::launch_grid_blocks() const {
    return (int)((slot_cap_ + block_ - 1) / block_);
}

int Engine::active_blocks() const {
    return (int)((hw_slots_ + block_ - 1) / block_);
}

void Engine::update_high_water() {
    hw_slots_ = 0;
    hw_cells_ = 0;
    hw_recs_ = 0;
    for (const SceneHost& sh : scenes_) {
        if (!sh.live) continue;
        hw_slots_ = std::max<int64_t>(hw_slots_, sh.p0 + sh.slot_cap);
        hw_cells_ = std::max<int64_t>(hw_cells_, sh.c0 + sh.cell_cap);
        if (sh.rec_cap) hw_recs_ = std::max<int64_t>(hw_recs_, sh.r0 + sh.rec_cap);
    }
}

int64_t Engine::free_slots() const {
    int64_t used = 0;
    for (auto& r : used_slots_) used += r.second;
    return slot_cap_ - used;
}

int Engine::alloc_region(std::vector<std::pair<int64_t, int64_t>>& used, int64_t cap, int64_t need, int64_t align,
                         int64_t* out) {
    std::sort(used.begin(), used.end());
    int64_t at = 0;
    for (auto& r : used) {
        if (round_up(at, align) + need <= r.first) break;
        at = std::max(at, r.first + r.second);
    }
    at = round_up(at, align);
    if (at + need > cap) return DEM_EBUSY;
    used.push_back({at, need});
    *out = at;
    return DEM_OK;
}

static void release_region(std::vector<std::pair<int64_t, int64_t>>& used, int64_t at) {
    for (size_t i = 0; i < used.size(); i++)
        if (used[i].first == at) { used.erase(used.begin() + i); return; }
}

int validate_params(const dem_params& p) {
    if (!p.materials || p.n_materials <= 0) { set_error("params: no materials"); return DEM_EINVAL; }
    if (!(p.grid.cell_size > 0.0f) || p.grid.w <= 0 || p.grid.h <= 0) {
        set_error("params: grid needs cell_size > 0, w > 0, h > 0");
        return DEM_EINVAL;
    }
    if ((int64_t)p.grid.w * p.grid.h > (int64_t)INT32_MAX / 4) { set_error("params: grid too large"); return DEM_EINVAL; }
    if (p.u.particles < 0) { set_error("params: negative particle count"); return DEM_EINVAL; }
    return DEM_OK;
}

int check_cell_size(const dem_grid& g, float rsmax) {
    const float cs = g.cell_size;
    const float ext = cs * (float)std::max(g.w, g.h);
    const float need = 2.0f * rsmax + 4.0f * (std::nextafter(ext, INFINITY) - ext);
    if (cs < need) {
        set_error("grid cell_size " + std::to_string(cs) + " is below 2*max radius (large bodies aside) + rounding "
                  "margin " + std::to_string(need));
        return DEM_EINVAL;
    }
    return DEM_OK;
}

int Engine::add_scene(const dem_params& p, int32_t capacity, int32_t bond_cap, int32_t rec_cap, int32_t* sid_out) {
    int rc = validate_params(p);
    if (rc) return rc;
    if (capacity <= 0) { set_error("capacity must be positive"); return DEM_EINVAL; }
    SceneHost sh;
    sh.params = p;
    sh.mats.assign(p.materials, p.materials + p.n_materials);
    sh.params.materials = nullptr;
    sh.slot_cap = (int32_t)round_up(capacity, block_);
    sh.cell_cap = p.grid.w * p.grid.h + 1;
    sh.bond_cap = std::max(bond_cap, 0);
    sh.rec_cap = std::max(rec_cap, 0);
    sh.mat_cap = std::max(p.n_materials, 8);
    int64_t at;
    if (alloc_region(used_slots_, slot_cap_, sh.slot_cap, block_, &at)) { set_error("no room for the particles"); return DEM_EBUSY; }
    sh.p0 = at;
    if (alloc_region(used_cells_, cell_cap_, sh.cell_cap, 1, &at)) {
        release_region(used_slots_, sh.p0);
        set_error("no room for the grid cells");
        return DEM_EBUSY;
    }
    sh.c0 = at;
    if (sh.bond_cap && alloc_region(used_bonds_, bond_cap_, sh.bond_cap, 1, &at)) {
        release_region(used_slots_, sh.p0); release_region(used_cells_, sh.c0);
        set_error("no room for the bonds");
        return DEM_EBUSY;
    }
    sh.b0 = sh.bond_cap ? at : 0;
    if (sh.rec_cap && alloc_region(used_recs_, rec_cap_, sh.rec_cap, 1, &at)) {
        release_region(used_slots_, sh.p0); release_region(used_cells_, sh.c0);
        if (sh.bond_cap) release_region(used_bonds_, sh.b0);
        set_error("no room for the bond records");
        return DEM_EBUSY;
    }
    sh.r0 = sh.rec_cap ? at : 0;
    if (alloc_region(used_mats_, mat_cap_, sh.mat_cap, 1, &at)) {
        release_region(used_slots_, sh.p0); release_region(used_cells_, sh.c0);
        if (sh.bond_cap) release_region(used_bonds_, sh.b0);
        if (sh.rec_cap) release_region(used_recs_, sh.r0);
        set_error("no room for the materials");
        return DEM_EBUSY;
    }
    sh.m0 = (int32_t)at;
    sh.nm = p.n_materials;
    sh.live = true;
    // reuse a dead scene id or append
    int32_t sid = -1;
    for (size_t i = 0; i < scenes_.size(); i++)
        if (!scenes_[i].live) { sid = (int32_t)i; break; }
    if (sid < 0) {
        if ((int32_t)scenes_.size() >= scene_cap_) {
            release_region(used_slots_, sh.p0); release_region(used_cells_, sh.c0);
            if (sh.bond_cap) release_region(used_bonds_, sh.b0);
            if (sh.rec_cap) release_region(used_recs_, sh.r0);
            release_region(used_mats_, sh.m0);
            set_error("scene table full");
            return DEM_EBUSY;
        }
        sid = (int32_t)scenes_.size();
        scenes_.push_back(sh);
    } else {
        scenes_[sid] = sh;
    }
    update_high_water();
    hipError_t e = hipMemcpyAsync(MAT_ + sh.m0, sh.mats.data(), sizeof(dem_material) * sh.nm, hipMemcpyHostToDevice,
                                  stream_);
    if ((rc = check(e, "upload materials"))) return rc;
    table_dirty_ = true;
    *sid_out = sid;
    return DEM_OK;
}

int Engine::remove_scene(int32_t sid) {
    if (sid < 0 || sid >= (int32_t)scenes_.size() || !scenes_[sid].live) { set_error("no such scene"); return DEM_EINVAL; }
    SceneHost& sh = scenes_[sid];
    release_region(used_slots_, sh.p0);
    release_region(used_cells_, sh.c0);
    if (sh.bond_cap) release_region(used_bonds_, sh.b0);
    if (sh.rec_cap) {
        release_region(used_recs_, sh.r0);
        // the bond pass skips the region until a scene uploads records into it
        hipError_t e = hipMemsetAsync(RO_ + sh.r0, 0xff, sizeof(int32_t) * sh.rec_cap, stream_);
        if (e == hipSuccess) e = hipMemsetAsync(RS_ + sh.r0, 0xff, sizeof(int32_t) * sh.rec_cap, stream_);
        if (int rc = check(e, "remove scene")) return rc;
    }
    release_region(used_mats_, sh.m0);
    sh = SceneHost();
    update_high_water();
    table_dirty_ = true;
    return DEM_OK;
}

int Engine::set_params(int32_t sid, const dem_params& p) {
    if (sid < 0 || sid >= (int32_t)scenes_.size() || !scenes_[sid].live) { set_error("no such scene"); return DEM_EINVAL; }
    int rc = validate_params(p);
    if (rc) return rc;
    SceneHost& sh = scenes_[sid];
    if (p.grid.w * p.grid.h + 1 > sh.cell_cap) { set_error("grid larger than the scene's cell reservation"); return DEM_ECAPACITY; }
    if (p.n_materials > sh.mat_cap) { set_error("too many materials for the scene's reservation"); return DEM_ECAPACITY; }
    if (sh.n > 0 && (rc = check_cell_size(p.grid, sh.rsmax))) return rc;
    rb_force_ = true;   // the grid (and with it the skin) may have changed
    const int32_t n_keep = sh.params.u.particles;
    sh.params = p;
    sh.params.materials = nullptr;
    sh.params.u.particles = n_keep;
    sh.mats.assign(p.materials, p.materials + p.n_materials);
    sh.nm = p.n_materials;
    hipError_t e = hipMemcpyAsync(MAT_ + sh.m0, sh.mats.data(), sizeof(dem_material) * sh.nm, hipMemcpyHostToDevice,
                                  stream_);
    if ((rc = check(e, "upload materials"))) return rc;
    table_dirty_ = true;
    return DEM_OK;
}

// Neighbour-list skin of a scene: what the cell side leaves above 2 r_small (the
// largest radius but the large bodies') plus the
// rounding margin of the cell index (4 ulp of the grid extent).  Zero skin is correct
// (the lists are rebuilt after every drift), only slower.
float Engine::skin_of(const SceneHost& sh) const {
    const float cs = sh.params.grid.cell_size;
    const float ext = cs * (float)std::max(sh.params.grid.w, sh.params.grid.h);
    const float need = 2.0f * sh.rsmax + 4.0f * (std::nextafter(ext, INFINITY) - ext);
    return std::max(cs - need, 0.0f);
}

// 2 zeta of a linear dashpot whose binary collisions rebound with restitution e:
// zeta = -ln e / sqrt(pi^2 + ln^2 e), evaluated in f64 and rounded once; 0 outside (0, 1).
float dashpot_factor(float e) {
    if (!(e > 0.0f && e < 1.0f)) return 0.0f;
    const double le = std::log((double)e);
    return (float)(2.0 * (-le / std::sqrt(M_PI * M_PI + le * le)));
}

SceneDev Engine::scene_dev(const SceneHost& sh, int32_t sid) const {
    SceneDev d{};
    d.u = sh.params.u;
    d.u.particles = sh.n;
    d.p0 = (int32_t)sh.p0;
    d.np = sh.np;
    d.c0 = (int32_t)sh.c0;
    d.w = sh.params.grid.w;
    d.h = sh.params.grid.h;
    d.cs = sh.params.grid.cell_size;
    d.base_x = -d.cs * (float)d.w * 0.5f;
    d.base_y = d.cs * (float)d.h * 0.5f;
    d.inv_cs = 1.0f / d.cs;
    d.half_w = d.u.hor_bound / 2.0f;
    d.half_h = d.u.vert_bound / 2.0f;
    d.b0 = (int32_t)sh.b0;
    d.nb = sh.nb;
    d.m0 = sh.m0;
    d.nm = sh.nm;
    d.use_f64 = sh.params.use_f64;
    d.record = (sh.params.record_data && with_data_) ? 1 : 0;
    d.has_strength = sh.has_strength ? 1 : 0;
    d.n_real = sh.n;
    d.skin = skin_of(sh);
    d.trig2 = (0.45f * d.skin) * (0.45f * d.skin);
    d.pscale = d.skin > 0.0f ? (float)VW_PATH_TRIG / (0.45f * d.skin) : 1.0e30f;
    d.rsmax = sh.rsmax;
    d.lb0 = sid * LB_MAX;
    d.nlb = (int32_t)sh.large.size();
    d.cn2 = dashpot_factor(sh.params.restitution);
    if (!sh.mats.empty()) {
        d.rho1 = sh.mats[0].density;
        d.kn1 = sh.mats[0].normal_stiffness;
        d.ks1 = sh.mats[0].shear_stiffness;
        // the series stiffness of two material-0 discs, the f32 expression of
        // 2D_Simulation.wgsl:345-346 evaluated once (every contact of a one-material scene
        // forms the same value)
        const float kn = d.kn1, ks = d.ks1;
        d.kne1 = 1.0f / (1.0f / kn + 1.0f / kn);
        d.kse1 = 1.0f / (1.0f / ks + 1.0f / ks);
    }
    return d;
}

int Engine::write_scene_table() {
    if (!table_dirty_) return DEM_OK;
    std::vector<SceneDev> tab(scene_cap_);
    std::vector<int32_t> blk(launch_grid_blocks(), -1);
    for (size_t i = 0; i < scenes_.size(); i++) {
        const SceneHost& sh = scenes_[i];
        if (!sh.live) continue;
        tab[i] = scene_dev(sh, (int32_t)i);
        for (int64_t b = sh.p0 / block_; b < (sh.p0 + sh.np) / block_; b++) blk[b] = (int32_t)i;
    }
    hipError_t e = hipMemcpyAsync(d_scenes_, tab.data(), sizeof(SceneDev) * scene_cap_, hipMemcpyHostToDevice, stream_);
    e = e ? e : hipMemcpyAsync(d_blk_scene_, blk.data(), sizeof(int32_t) * blk.size(), hipMemcpyHostToDevice, stream_);
    e = e ? e : hipStreamSynchronize(stream_);
    int rc = check(e, "scene table");
    if (!rc) table_dirty_ = false;
    return rc;
}

DevArrays Engine::arrays(int set) const {
    const DevSet& s = sets_[set];
    DevArrays a{};
    a.P[0] = s.P[0]; a.P[1] = s.P[1];
    a.V = s.V[lcur_]; a.Vn = s.V[lcur_ ^ 1]; a.DX = s.DX[lcur_]; a.DXn = s.DX[lcur_ ^ 1];
    a.A = s.A; a.TH = s.TH; a.M = s.M; a.ID = s.ID; a.EXT = s.EXT; a.DATA0 = s.DATA0; a.DATA1 = s.DATA1;
    const ContactLists& lc = lists_[lcur_];
    const ContactLists& ln = lists_[lcur_ ^ 1];
    a.CP = lc.CP; a.CF = lc.CF;
    a.CPn = ln.CP; a.CFn = ln.CF;
    a.FT = FT_;
    a.cap = slot_cap_;
    a.BOFF = BOFF_; a.BR = BR_; a.BF = BF_; a.BS = BS_; a.BLS = BLS_; a.BRB = BRB_; a.MAT = MAT_;
    a.KEY = KEY_; a.RANK = RANK_; a.CNT = CNT_; a.START = START_; a.SIDX = SIDX_; a.INV = INV_;
    a.NL = s.NL; a.XR = s.XR; a.RB = RB_;
    a.RO = RO_; a.RF = RF_; a.RD = RD_; a.RS = RS_; a.RP = RP_; a.BRS = s.BRS; a.LB = LB_;
    a.LBH = LBH_; a.LBHC = LBHC_;
    a.STAT = STAT_; a.TSUM = TSUM_; a.ntiles = ntiles_;
    return a;
}

StepArgs Engine::make_args(const int2* lf) const {
    StepArgs s{};
    s.scenes = d_scenes_;
    s.blk_scene = d_blk_scene_;
    s.lf = lf;
    s.a = arrays(set_);
    s.cur = cur_;
    s.one_scene = -1;
    int live = 0;
    for (size_t i = 0; i < scenes_.size(); i++)
        if (scenes_[i].live) {
            live++;
            s.one_scene = (int)i;
        }
    if (live != 1) s.one_scene = -1;
    return s;
}

// ---------------------------------------------------------------------------
// Upload: reference layouts (state.rs:69-93) -> device layout, storage order = id order.
int Engine::upload(int32_t sid, const dem_state_view& v) {
    if (sid < 0 || sid >= (int32_t)scenes_.size() || !scenes_[sid].live) { set_error("no such scene"); return DEM_EINVAL; }
    // a slot's position and velocity records of one parity share a record: the two
    // parities flip together
    if (cur_ != lcur_) { set_error("position and list parities differ"); return DEM_ESTATE; }
    SceneHost& sh = scenes_[sid];
    const int32_t n = v.n;
    if (n < 0 || n > sh.slot_cap) { set_error("particle count exceeds the scene capacity"); return DEM_ECAPACITY; }
    if (n > 0 && (!v.pos || !v.vel || !v.acc || !v.rot || !v.rot_vel || !v.rot_acc || !v.forces || !v.radii ||
                  !v.fixity || !v.material_pointers)) {
        set_error("state view is missing a particle array");
        return DEM_EINVAL;
    }
    if (v.n_bonds > 0 && !v.bonds) { set_error("n_bonds > 0 but bonds is NULL"); return DEM_EINVAL; }
    if (v.n_bonds > sh.bond_cap) { set_error("more bonds than the scene's bond capacity"); return DEM_ECAPACITY; }
    if (v.contacts && v.contact_slots <= 0) { set_error("contacts given with contact_slots <= 0"); return DEM_EINVAL; }
    const int32_t np = (int32_t)round_up(std::max(n, 1), block_);
    sh.n = n;
    sh.np = np;
    sh.params.u.particles = n;
    const int32_t nmat = sh.nm;
    // the device layout is built in pinned staging memory and copied from there
    Carve cv;
    auto carve = [&](Carve& c) {
        c.take<float4>(2 * (size_t)np); c.take<float4>(np); c.take<float4>(np); c.take<int32_t>(np);
        c.take<int32_t>(np); c.take<float4>(np); c.take<float4>(np); c.take<float4>(np); c.take<int32_t>(np);
        c.take<int32_t>((size_t)np * KMAX); c.take<float>((size_t)np * KMAX); c.take<int32_t>(np);
        c.take<int32_t>(np); c.take<int2>(np); c.take<float>(np);
    };
    carve(cv);
    cv.base = staging(cv.off);
    if (!cv.base) { set_error("upload: pinned staging allocation failed"); return DEM_ENOMEM; }
    cv.off = 0;
    float4* PV = cv.take<float4>(2 * (size_t)np);   // {position, velocity record} per slot
    const PVHalf P{PV}, V{PV + 1};
    float4* DX = cv.take<float4>(np);
    float4* A = cv.take<float4>(np);
    int32_t* M = cv.take<int32_t>(np);    // flag words
    int32_t* ID = cv.take<int32_t>(np);
    float4* EXT = cv.take<float4>(np);
    float4* D0 = cv.take<float4>(np);
    float4* D1 = cv.take<float4>(np);
    int32_t* CC = cv.take<int32_t>(np);
    int32_t* CP = cv.take<int32_t>((size_t)np * KMAX);   // rows beyond a slot's count are never read
    float* CF = cv.take<float>((size_t)np * KMAX);
    int32_t* INV = cv.take<int32_t>(np);
    int32_t* SIDX = cv.take<int32_t>(np);
    int2* BRANGE = cv.take<int2>(np);
    float* TH = cv.take<float>(np);
    std::fill(CC, CC + np, 0);
    int32_t rows = 0;   // contact-list rows in use
    sh.fixity.assign(v.fixity, v.fixity + 6 * (size_t)n);
    sh.forces.assign(v.forces, v.forces + 6 * (size_t)n);
    sh.matptr.assign(v.material_pointers, v.material_pointers + n);
    sh.has_ext = false;
    // large bodies (a striker among small discs) pair through the scene's list
    float rsmax = 0.0f;
    const int32_t nlarge = n > 0 ? split_large(v.radii, n, &rsmax) : 0;
    sh.large.clear();
    for (int32_t i = 0; i < n && nlarge > 0; i++)
        if (v.radii[i] > rsmax) sh.large.push_back(i);
    for (int32_t i = 0; i < np; i++) {
        INV[i] = (int32_t)sh.p0 + i;
        SIDX[i] = (int32_t)sh.p0 + i;
        if (i >= n) {
            P[i] = DX[i] = V[i] = A[i] = EXT[i] = D0[i] = D1[i] = make_float4(0, 0, 0, 0);
            P[i].w = bitsf(i);
            TH[i] = 0.0f;
            M[i] = (int32_t)PF_INACTIVE;
            ID[i] = i;
            continue;
        }
        const float r = v.radii[i];
        // WGPUComputeProg::restore uploads vel into del_pos and rot_vel into del_rot
        // (wgpu_prog.rs:1247-1248): an explicit displacement record
        // the position record carries the particle id (its neighbours read it with the
        // position); the rotation has its own array
        P[i] = make_float4(v.pos[2 * i], v.pos[2 * i + 1], r, bitsf(i));
        TH[i] = v.rot[i];
        DX[i] = make_float4(v.vel[2 * i], v.vel[2 * i + 1], v.rot_vel[i], 0.0f);
        float vw;
        std::memcpy(&vw, &VW_EXPLICIT, sizeof(vw));
        V[i] = make_float4(v.vel[2 * i], v.vel[2 * i + 1], v.rot_vel[i], vw);
        A[i] = make_float4(v.acc[2 * i], v.acc[2 * i + 1], v.rot_acc[i], 0.0f);
        const float* f = v.forces + 6 * (size_t)i;
        EXT[i] = make_float4(f[0], f[1], f[2], 0.0f);
        const int32_t* fx = v.fixity + 6 * (size_t)i;
        uint32_t flags = 0;
        int32_t mat = v.material_pointers[i];
        if (mat < 0 || mat >= nmat) mat = 0;
        flags |= (uint32_t)mat & PF_MAT_MASK;
        if (fx[0] == 0) flags |= PF_KICK_X;
        if (fx[1] == 0) flags |= PF_KICK_Y;
        if (fx[2] == 0) flags |= PF_KICK_R;
        if (fx[0] != 1) flags |= PF_WALL_X;
        if (fx[1] != 1) flags |= PF_WALL_Y;
        if (fx[2] != 1) flags |= PF_WALL_R;
        if (fbits(f[0]) || fbits(f[1]) || fbits(f[2])) { flags |= PF_EXT; sh.has_ext = true; }
        if (r == 0.0f) flags |= PF_INACTIVE;
        if (nlarge > 0 && r > rsmax) flags |= PF_LARGE;
        M[i] = (int32_t)flags;
        ID[i] = i;
        if (v.data) {
            const float* d = v.data + 7 * (size_t)i;
            D0[i] = make_float4(d[0], d[1], d[2], d[3]);
            D1[i] = make_float4(d[4], d[5], d[6], 0.0f);
        } else {
            D0[i] = D1[i] = make_float4(0, 0, 0, 0);
        }
    }
    sh.r0_zero = n > 0 && v.radii[0] == 0.0f;
    {
        // the 3x3 stencil needs cells wider than 2 r_max plus the rounding of
        // (x - base) / cell_size, a few ulp of the grid extent
        float rmax = 0.0f;
        for (int32_t i = 0; i < n; i++) rmax = std::max(rmax, v.radii[i]);
        sh.rmax = rmax;
        sh.rsmax = nlarge > 0 ? rsmax : rmax;
        if (int rc = check_cell_size(sh.params.grid, sh.rsmax)) return rc;
    }
    // bonds: index word, rest length (angle kept on the host)
    sh.nb = v.n_bonds;
    sh.bond_angle.assign(sh.nb, 0);
    sh.bond_len.assign(sh.nb, 0);
    std::vector<int2> BS(std::max(sh.nb, 1));
    std::vector<float4> BLS(std::max(sh.nb, 1));
    for (int32_t k = 0; k < sh.nb; k++) {
        const int32_t idx = v.bonds[3 * k];
        sh.bond_angle[k] = v.bonds[3 * k + 1];
        sh.bond_len[k] = v.bonds[3 * k + 2];
        BS[k] = make_int2(idx, INT32_MIN);
        BLS[k] = v.bond_strengths ? make_float4(bitsf(v.bonds[3 * k + 2]), v.bond_strengths[2 * k],
                                                v.bond_strengths[2 * k + 1], 0.0f)
                                  : make_float4(bitsf(v.bonds[3 * k + 2]), sh.params.u.bond_normal_strength,
                                                sh.params.u.bond_shear_strength, 0.0f);
        // a bond uploaded broken reads as broken before the next step
        if (idx < 0) BLS[k].w = bitsf((int32_t)brk_word(idx, (uint32_t)sh.step));
    }
    sh.has_strength = v.bond_strengths != nullptr;
    // contact table -> contact lists (plain contacts) and bond records (everything else)
    std::vector<int32_t> BOFF(np + 1, 0);
    std::vector<int32_t> RO;   // owner (global particle id) of each record
    std::vector<int4> BR;
    std::vector<float4> BF;
    std::vector<int2> BRB;
    if (v.contacts) {
        const Slot* tab = reinterpret_cast<const Slot*>(v.contacts);
        const int K = v.contact_slots;
        std::vector<std::pair<int32_t, int32_t>> rec, con;  // (b, slot index)
        for (int32_t i = 0; i < n; i++) {
            rec.clear();
            con.clear();
            for (int k = 0; k < K; k++) {
                const Slot& s = tab[(size_t)i * K + k];
                if (s.b == -1) continue;
                if (s.b < 0 || s.b >= n) continue;  // stale word pointing nowhere
                const bool plain = s.bonded == -1 && fbits(s.fbt) == -1 && fbits(s.tb) == -1;
                (plain ? con : rec).push_back({s.b, k});
            }
            std::sort(con.begin(), con.end());
            std::sort(rec.begin(), rec.end());
            con.erase(std::unique(con.begin(), con.end(),
                                  [](auto& x, auto& y) { return x.first == y.first; }), con.end());
            if ((int)(con.size() + rec.size()) > KMAX) {
                // the reference's table holds 14 slots per particle (2D_Simulation.wgsl:125)
                set_error("particle " + std::to_string(i) + " occupies more than 14 contact slots");
                return DEM_ECAPACITY;
            }
            CC[i] = (int32_t)con.size();
            rows = std::max(rows, CC[i]);
            M[i] = (int32_t)(((uint32_t)M[i] & ~PF_CC_MASK) | ((uint32_t)CC[i] << PF_CC_SHIFT));
            for (size_t q = 0; q < con.size(); q++) {
                const Slot& s = tab[(size_t)i * K + con[q].second];
                CP[(size_t)q * np + i] = s.b;
                CF[(size_t)q * np + i] = s.ft;
            }
            BOFF[i] = (int32_t)BR.size();
            for (size_t k = 0; k < rec.size(); k++) RO.push_back((int32_t)sh.p0 + i);
            for (auto& rr : rec) {
                const Slot& s = tab[(size_t)i * K + rr.second];
                int32_t bid = s.bonded >= 0 ? s.bonded : -1;
                BR.push_back(make_int4(s.b, bid, s.bonded, 1));
                BF.push_back(make_float4(s.ft, s.fbt, s.tb, 0.0f));
            }
        }
        for (int32_t i = n; i <= np; i++) BOFF[i] = (int32_t)BR.size();
    }
    if ((int64_t)BR.size() > sh.rec_cap) { set_error("more bond records than the scene's record capacity"); return DEM_ECAPACITY; }
    sh.nrec = (int32_t)BR.size();
    // per particle id: its record range [begin, end) in the pool (one entry per slot, so
    // no scene's ranges touch a neighbour region)
    for (int32_t i = 0; i < np; i++)
        BRANGE[i] = make_int2(BOFF[i] + (int32_t)sh.r0, BOFF[i + 1] + (int32_t)sh.r0);
    // device copies: the current set, the current position and list buffers
    hipError_t e = hipSuccess;
    const DevSet& S = sets_[set_];
    const int64_t p0 = sh.p0;
    auto h2d = [&](void* dst, const void* src, size_t bytes) {
        if (e == hipSuccess && bytes) e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream_);
    };
    // positions and velocity records of one parity share their records (cur_ == lcur_)
    h2d(S.PV[cur_] + 2 * p0, PV, 2 * sizeof(float4) * np);
    h2d(S.DX[lcur_] + p0, DX, sizeof(float4) * np);
    h2d(S.A + p0, A, sizeof(float4) * np);
    h2d(S.TH + p0, TH, sizeof(float) * np);
    h2d(S.M + p0, M, sizeof(int32_t) * np);
    h2d(S.ID + p0, ID, sizeof(int32_t) * np);
    h2d(S.EXT + p0, EXT, sizeof(float4) * np);
    if (with_data_) {
        h2d(S.DATA0 + p0, D0, sizeof(float4) * np);
        h2d(S.DATA1 + p0, D1, sizeof(float4) * np);
    }
    const ContactLists& L = lists_[lcur_];
    if (rows > 0 && e == hipSuccess) {
        // rows [0, rows) of the slot-major lists in one strided copy each
        e = hipMemcpy2DAsync(L.CP + p0, sizeof(int32_t) * slot_cap_, CP, sizeof(int32_t) * np, sizeof(int32_t) * np,
                             rows, hipMemcpyHostToDevice, stream_);
        if (e == hipSuccess)
            e = hipMemcpy2DAsync(L.CF + p0, sizeof(float) * slot_cap_, CF, sizeof(float) * np, sizeof(float) * np,
                                 rows, hipMemcpyHostToDevice, stream_);
    }
    h2d(INV_ + p0, INV, sizeof(int32_t) * np);
    h2d(SIDX_ + p0, SIDX, sizeof(int32_t) * np);
    h2d(BOFF_ + p0, BRANGE, sizeof(int2) * np);
    h2d(S.BRS + p0, BRANGE, sizeof(int2) * np);   // storage order is id order after an upload
    if (!sh.large.empty()) h2d(LB_ + (int64_t)sid * LB_MAX, sh.large.data(), sizeof(int32_t) * sh.large.size());
    if (sh.nrec) {
        h2d(BR_ + sh.r0, BR.data(), sizeof(int4) * sh.nrec);
        BRB.resize(sh.nrec);
        for (int32_t q = 0; q < sh.nrec; q++)   // the bond pass's copy of the record words
            BRB[q] = make_int2(BR[q].y, BR[q].w != 0 ? BR[q].z : REC_DEAD);
        h2d(BRB_ + sh.r0, BRB.data(), sizeof(int2) * sh.nrec);
        h2d(BF_ + sh.r0, BF.data(), sizeof(float4) * sh.nrec);
        h2d(RO_ + sh.r0, RO.data(), sizeof(int32_t) * sh.nrec);
    }
    if (sh.rec_cap > sh.nrec && e == hipSuccess)   // the rest of the region holds no records
        e = hipMemsetAsync(RO_ + sh.r0 + sh.nrec, 0xff, sizeof(int32_t) * (sh.rec_cap - sh.nrec), stream_);
    if (sh.nb) {
        h2d(BS_ + sh.b0, BS.data(), sizeof(int2) * sh.nb);
        h2d(BLS_ + sh.b0, BLS.data(), sizeof(float4) * sh.nb);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    int rc = check(e, "upload");
    if (rc) return rc;
    sh.since_reorder = 1 << 30;  // re-sort storage at the first rebuild
    rb_force_ = true;            // neighbour lists for the new state
    table_dirty_ = true;
    if (sh.nrec) {
        // owner and partner slots of the records (the scene table maps slots to scenes)
        if ((rc = write_scene_table())) return rc;
        if ((rc = check(launch_record_slots(make_args(d_lf_), (int)hw_recs_, block_, stream_), "record slots")))
            return rc;
    }
    return DEM_OK;
}

// ---------------------------------------------------------------------------
void Engine::plan_reorder(const int2* lf_dev, std::vector<LaunchOp>& ops) {
    // Every live scene, binned by the rebuild this re-sort belongs to and with its
    // neighbour lists just built, moves into cell order, each unit of 64 blocks of slots
    // then partitioned stably into particles with list candidates first (the order
    // inside a cell is the binning's; results depend on neither).
    ReorderArgs r{};
    r.scenes = d_scenes_;
    r.blk_scene = d_blk_scene_;
    r.lf = lf_dev;
    r.src = arrays(set_);
    r.dst = arrays(set_ ^ 1);
    // the contact lists and velocity records land in the other list buffer (lcur_
    // flips with set_ below)
    r.dst.CP = r.src.CPn;
    r.dst.CF = r.src.CFn;
    r.dst.V = r.dst.Vn;
    r.dst.DX = r.dst.DXn;
    // and the positions in the other position buffer (cur_ flips too)
    std::swap(r.dst.P[0], r.dst.P[1]);
    r.order = ORDER_;
    r.newslot = NEWSLOT_;
    r.cur = cur_;
    r.with_ext = 0;
    for (const SceneHost& sh : scenes_)
        if (sh.live && sh.has_ext) r.with_ext = 1;  // EXT is read only where PF_EXT is set
    r.with_data = with_data_ ? 1 : 0;
    LaunchOp& a = new_op(ops, OP_PARTITION, 3);
    a.ra = r;
    a.nlb = active_blocks();
    a.block = block_;
    LaunchOp& b = new_op(ops, OP_MOVE, 3);
    b.ra = r;
    b.nlb = active_blocks();
    b.block = block_;
    bool records = false;
    for (const SceneHost& sh : scenes_)
        if (sh.live && sh.nrec > 0) records = true;
    if (records) {
        // the records' owner and partner slots after the move
        LaunchOp& c = new_op(ops, OP_RSLOTS, 3);
        c.sa = make_args(lf_dev);
        c.sa.a = r.dst;
        c.ncells = (int)hw_recs_;
        c.block = block_;
    }
    set_ ^= 1;
    lcur_ ^= 1;
    cur_ ^= 1;
    for (SceneHost& sh : scenes_) {
        if (!sh.live) continue;
        sh.since_reorder = 0;
        sh.reorders++;
    }
}

LaunchOp& Engine::new_op(std::vector<LaunchOp>& ops, int kind, int timed) {
    ops.emplace_back();
    LaunchOp& op = ops.back();
    std::memset(static_cast<void*>(&op), 0, sizeof(LaunchOp));   // padding too: ops compare bytewise
    op.kind = kind;
    op.timed = timed;
    return op;
}

hipError_t Engine::run_op(const LaunchOp& op, hipEvent_t ev0, hipEvent_t ev1) {
    switch (op.kind) {
        case OP_STEP:
            return launch_step(op.sa, op.nlb, op.block, op.f64 != 0, op.bonds != 0, op.value != 0, stream_, ev0, ev1);
        case OP_RB_STORE: return launch_rb_store(op.rb, op.value, stream_, ev0, ev1);
        case OP_RB_BIN: return launch_rb_bin(op.sa, op.nlb, op.grid, op.block, stream_, ev0, ev1);
        case OP_SCAN: return launch_scan(op.da, op.ncells, op.gate, stream_, ev0, ev1);
        case OP_RB_SCATTER: return launch_rb_scatter(op.sa, op.nlb, op.grid, op.block, stream_, ev0, ev1);
        case OP_BUILD_NL: return launch_build_nl(op.sa, op.nlb, op.grid, op.block, stream_, ev0, ev1);
        case OP_PARTITION: return launch_partition(op.ra, op.nlb, op.block, stream_, ev0, ev1);
        case OP_MOVE: return launch_reorder_move(op.ra, op.nlb, op.block, stream_, ev0, ev1);
        case OP_BONDS: return launch_bond_forces(op.sa, op.ncells, op.block, op.f64 != 0, stream_, ev0, ev1);
        case OP_RSLOTS: return launch_record_slots(op.sa, op.ncells, op.block, stream_, ev0, ev1);
        case OP_LARGE: return launch_large_search(op.sa, op.grid, stream_, ev0, ev1);
        default: return hipErrorInvalidValue;
    }
}

static bool same_ops(const std::vector<LaunchOp>& a, const std::vector<LaunchOp>& b) {
    return a.size() == b.size() && std::memcmp(a.data(), b.data(), sizeof(LaunchOp) * a.size()) == 0;
}

void Engine::clear_graphs() {
    for (GraphEntry& g : graphs_) {
        if (g.exec) hipGraphExecDestroy(g.exec);
        if (g.graph) hipGraphDestroy(g.graph);
    }
    graphs_.clear();
}

// Runs the planned launches: eagerly (timing, graphs off, long sequences), or as a HIP
// graph, captured the first time a sequence occurs and replayed when it repeats.
int Engine::run_ops(const std::vector<LaunchOp>& ops) {
    const bool graph = use_graph_ && !timing_ && ops.size() <= kGraphMaxOps;
    if (!graph) {
        struct Ev { int timed, kind; hipEvent_t a, b; };
        std::vector<Ev> evs;
        size_t ev_used = 0;
        auto ev = [&]() -> hipEvent_t {
            if (ev_used >= ev_pool_.size()) {
                hipEvent_t x;
                hipEventCreate(&x);
                ev_pool_.push_back(x);
            }
            return ev_pool_[ev_used++];
        };
        hipEvent_t t0 = nullptr, t1 = nullptr;
        if (timing_) { t0 = ev(); t1 = ev(); hipEventRecord(t0, stream_); }
        for (const LaunchOp& op : ops) {
            hipEvent_t a = nullptr, b = nullptr;
            // every launch's events carry the kernel's own start and end
            if (timing_) { a = ev(); b = ev(); }
            const hipError_t e = run_op(op, a, b);
            if (timing_) evs.push_back({op.timed, op.kind, a, b});
            if (int rc = check(e, kOpNames[op.kind])) return rc;
        }
        if (timing_) {
            hipEventRecord(t1, stream_);
            hipEventSynchronize(t1);
            float ms = 0;
            hipEventElapsedTime(&ms, t0, t1);
            t_total = ms;
            t_step = t_scan = t_scatter = t_reorder = t_bonds = 0;
            n_step = n_scan = n_scatter = n_reorder = n_bonds = 0;
            for (auto& x : evs) {
                float m = 0;
                hipEventElapsedTime(&m, x.a, x.b);
                switch (x.timed) {
                    case 0: t_step += m; n_step += x.kind == OP_STEP ? 1 : 0; break;
                    case 1: t_scan += m; n_scan++; break;
                    case 2: t_scatter += m; n_scatter++; break;
                    case 4: t_bonds += m; n_bonds++; break;
                    default: t_reorder += m; n_reorder++; break;
                }
            }
        }
        return DEM_OK;
    }
    GraphEntry* hit = nullptr;
    for (GraphEntry& g : graphs_)
        if (same_ops(g.ops, ops)) hit = &g;
    if (!hit) {
        if (graphs_.size() >= kGraphCache) {
            // drop the least recently used
            size_t lru = 0;
            for (size_t i = 1; i < graphs_.size(); i++)
                if (graphs_[i].used < graphs_[lru].used) lru = i;
            hipGraphExecDestroy(graphs_[lru].exec);
            hipGraphDestroy(graphs_[lru].graph);
            graphs_.erase(graphs_.begin() + (long)lru);
        }
        GraphEntry g;
        g.ops = ops;
        hipError_t e = hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal);
        if (int rc = check(e, "graph capture")) return rc;
        hipError_t le = hipSuccess;
        for (const LaunchOp& op : ops)
            if (le == hipSuccess) le = run_op(op);
        e = hipStreamEndCapture(stream_, &g.graph);
        if (int rc = check(le, "graph capture: launch")) {
            if (g.graph) hipGraphDestroy(g.graph);
            return rc;
        }
        if (int rc = check(e, "graph capture: end")) return rc;
        e = hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0);
        if (int rc = check(e, "graph instantiate")) {
            hipGraphDestroy(g.graph);
            return rc;
        }
        graphs_.push_back(std::move(g));
        hit = &graphs_.back();
        graph_captures_++;
    } else {
        graph_replays_++;
    }
    hit->used = ++graph_clock_;
    return check(hipGraphLaunch(hit->exec, stream_), "graph launch");
}

int Engine::step(const std::vector<int32_t>& sids, const std::vector<int32_t>& nsteps) {
    if (sids.size() != nsteps.size()) { set_error("step: size mismatch"); return DEM_EINVAL; }
    int rc = write_scene_table();
    if (rc) return rc;
    int32_t maxn = 0;
    for (size_t i = 0; i < sids.size(); i++) {
        const int32_t sid = sids[i];
        if (sid < 0 || sid >= (int32_t)scenes_.size() || !scenes_[sid].live) { set_error("step: no such scene"); return DEM_EINVAL; }
        if (nsteps[i] < 0) { set_error("step: negative step count"); return DEM_EINVAL; }
        maxn = std::max(maxn, nsteps[i]);
    }
    if (maxn == 0) return DEM_OK;
    const int64_t nlaunch = (int64_t)maxn + 1;
    // per-launch, per-scene flags, computed on the host from the coll_cont rules
    std::vector<int2> lf((size_t)nlaunch * scene_cap_, make_int2(0, 0));
    std::vector<std::vector<uint8_t>> bin_after(nlaunch);
    for (size_t i = 0; i < sids.size(); i++) {
        SceneHost& sh = scenes_[sids[i]];
        const int32_t n = nsteps[i];
        if (n == 0) continue;
        const dem_uniform& u = sh.params.u;
        bool narrow_prev = false;
        for (int32_t k = 0; k <= n; k++) {
            uint32_t f = SF_ACTIVE;
            int32_t stepno = (int32_t)(sh.step + k);
            if (k > 0) {
                f |= SF_SIM;
                if (narrow_prev) f |= SF_NARROW;
                // end of Sim_k: invocation 0 resets coll_cont[3] (2D_Simulation.wgsl:258-261)
                if (!sh.r0_zero) sh.del_override = false;
            }
            if (k < n) {
                f |= SF_LOM;
                // LOM_{k+1}, invocation 0: interval counter (2D_LOM.wgsl:155-163)
                if (!sh.r0_zero) {
                    sh.coll1 += 1;
                    if (sh.coll1 >= u.collision_interval) { sh.coll0 = 1; sh.coll1 = 0; }
                    else sh.coll0 = 0;
                }
                if (sh.del_override) f |= SF_DELOVR;
                narrow_prev = u.collisions == 1 && sh.coll0 == 1;
                if (narrow_prev) f |= SF_BIN;
            }
            lf[(size_t)k * scene_cap_ + sids[i]] = make_int2((int)f, stepno);
        }
        sh.step += n;
        sh.steps += n;
    }
    // ring of launch flags on the device
    if (nlaunch * scene_cap_ > lf_cap_) {
        hipFree(d_lf_);
        lf_cap_ = nlaunch * scene_cap_;
        if ((rc = check(dalloc(&d_lf_, lf_cap_), "hipMalloc launch flags"))) return rc;
    }
    hipError_t e = hipMemcpyAsync(d_lf_, lf.data(), sizeof(int2) * lf.size(), hipMemcpyHostToDevice, stream_);
    if ((rc = check(e, "upload launch flags"))) return rc;

    bool any_f64 = false, any_bonds = false, any_large = false;
    int lb_grid = 0;   // one wave per large-body entry of the scenes up to the last with large bodies
    for (int32_t sid : sids) {
        any_f64 |= scenes_[sid].params.use_f64 != 0;
        any_bonds |= scenes_[sid].nrec > 0;
        if (!scenes_[sid].large.empty()) {
            any_large = true;
            lb_grid = std::max(lb_grid, (sid + 1) * LB_MAX);
        }
    }
    const int nblk = active_blocks();
    // the neighbour-list rebuild kernels stride over the slot blocks with one round of
    // workgroups (32 waves per CU), so that a launch whose rebuild flag is clear is cheap
    const int rb_grid = std::max(1, std::min(nblk, cus_ * std::max(1, 2048 / block_)));
    std::vector<LaunchOp> ops;
    ops.reserve((size_t)nlaunch * 6);
    for (int64_t k = 0; k < nlaunch; k++) {
        const int2* lfk = d_lf_ + k * scene_cap_;
        bool lom = false;
        for (size_t i = 0; i < sids.size(); i++)
            if (lf[(size_t)k * scene_cap_ + sids[i]].x & SF_LOM) lom = true;
        bool sim = false;
        for (size_t i = 0; i < sids.size(); i++)
            if (lf[(size_t)k * scene_cap_ + sids[i]].x & SF_SIM) sim = true;
        if (!lom) {
            // No drift, so no rebuild follows this launch, but its position parity flips:
            // the flag that the last rebuild chain consumed (RB[cur]) is cleared here,
            // before any record death of this launch sets both flags.
            LaunchOp& o = new_op(ops, OP_RB_STORE, 2);
            o.rb = RB_ + cur_;
            o.value = 0;
        }
        if (any_bonds && sim) {
            // bond records first: one thread per record, results read by the step kernel
            LaunchOp& bp = new_op(ops, OP_BONDS, 4);
            bp.sa = make_args(lfk);
            bp.ncells = (int)hw_recs_;
            bp.block = block_;
            bp.f64 = any_f64 ? 1 : 0;
        }
        if (any_large && sim) {
            // each large body's overlapping small discs, one wave per large body, for the
            // contact search of its thread in the step kernel (timed with the step kernel)
            LaunchOp& lo = new_op(ops, OP_LARGE, 0);
            lo.sa = make_args(lfk);
            lo.grid = lb_grid;
        }
        LaunchOp& st = new_op(ops, OP_STEP, 0);
        st.sa = make_args(lfk);
        st.sa.flip = 1;
        st.nlb = nblk;
        st.block = block_;
        st.f64 = any_f64 ? 1 : 0;
        st.bonds = any_bonds ? 1 : 0;
        st.value = any_large ? 1 : 0;   // a scene with large bodies
        // The step kernel wrote every live slot's next contact list and velocity record,
        // and its next position (drifted or carried over): both parities flip with every
        // launch, so which position buffer pairs with which velocity buffer never changes
        // (one pairing of the buffers streams measurably faster than the other).
        lcur_ ^= 1;
        cur_ ^= 1;
        if (!lom) continue;
        // Neighbour-list rebuild after every drift.  The kernels run only when the
        // rebuild flag is set: by the step kernel (a particle moved more than 0.45
        // skin since the last build) or here (new state, new parameters, re-sort due).
        // A due re-sort moves every live scene into the order of the fresh cell list.
        // With graphs a due re-sort waits for the first drift of a call, so that calls
        // repeat a few launch patterns (results do not depend on when re-sorts happen).
        const bool may_sort = k == 0 || !(use_graph_ && !timing_);
        bool due = false;
        for (size_t i = 0; i < sids.size(); i++) {
            if (!(lf[(size_t)k * scene_cap_ + sids[i]].x & SF_LOM)) continue;
            SceneHost& sh = scenes_[sids[i]];
            const int32_t R = sh.params.reorder_interval;
            if (may_sort && (R > 0 ? sh.since_reorder >= R : (R == 0 && sh.since_reorder >= (1 << 30)))) due = true;
            sh.since_reorder++;
        }
        const StepArgs ra = make_args(lfk);
        int32_t* rb = RB_ + cur_;   // this drift's flag (the next chain's first kernel clears the other)
        if (rb_force_ || due) {
            LaunchOp& o = new_op(ops, OP_RB_STORE, 2);
            o.rb = rb;
            o.value = 1;
        }
        const int chain[3] = {OP_RB_BIN, OP_SCAN, OP_RB_SCATTER};
        for (int kind : chain) {
            LaunchOp& o = new_op(ops, kind, kind == OP_SCAN ? 1 : 2);
            o.sa = ra;
            o.nlb = nblk;
            o.grid = rb_grid;
            o.block = block_;
            if (kind == OP_SCAN) {
                o.da = arrays(set_);
                o.ncells = (int)hw_cells_;
                o.gate = rb;
            }
        }
        LaunchOp& bl = new_op(ops, OP_BUILD_NL, 2);
        bl.sa = ra;
        bl.nlb = nblk;
        bl.grid = rb_grid;
        bl.block = block_;
        if (due) {
            // the lists move with their particles; the re-sort flips the position parity,
            // so the flag this chain consumed would otherwise be read by the next chain
            plan_reorder(lfk, ops);
            LaunchOp& o = new_op(ops, OP_RB_STORE, 3);
            o.rb = rb;
            o.value = 0;
        }
        rb_force_ = false;
    }
    return run_ops(ops);
}

int Engine::sync() { return check(hipStreamSynchronize(stream_), "hipStreamSynchronize"); }

// ---------------------------------------------------------------------------
int Engine::download(int32_t sid, dem_state_view& v) {
    if (sid < 0 || sid >= (int32_t)scenes_.size() || !scenes_[sid].live) { set_error("no such scene"); return DEM_EINVAL; }
    if (cur_ != lcur_) { set_error("position and list parities differ"); return DEM_ESTATE; }
    const SceneHost& sh = scenes_[sid];
    const int32_t n = sh.n, np = sh.np;
    if (v.n != n) { set_error("view.n does not match the scene's particle count"); return DEM_EINVAL; }
    const DevSet& S = sets_[set_];
    const int64_t p0 = sh.p0;
    // copies land in pinned staging memory
    const size_t nrec = (size_t)std::max(sh.nrec, 1), nbs = (size_t)std::max(sh.nb, 1);
    Carve cv;
    auto carve = [&](Carve& c) {
        c.take<float4>(2 * (size_t)np); c.take<float4>(np); c.take<float4>(np); c.take<float4>(np);
        c.take<int32_t>(np); c.take<int32_t>(np); c.take<int32_t>((size_t)np * KMAX);
        c.take<float>((size_t)np * KMAX); c.take<int2>(np); c.take<int4>(nrec); c.take<float4>(nrec);
        c.take<int2>(nbs); c.take<float>(np);
    };
    carve(cv);
    cv.base = staging(cv.off);
    if (!cv.base) { set_error("download: pinned staging allocation failed"); return DEM_ENOMEM; }
    cv.off = 0;
    float4* PV = cv.take<float4>(2 * (size_t)np);   // {position, velocity record} per slot
    const PVHalf P{PV}, V{PV + 1};
    float4* A = cv.take<float4>(np);
    float4* D0 = cv.take<float4>(np);
    float4* D1 = cv.take<float4>(np);
    int32_t* M = cv.take<int32_t>(np);
    int32_t* INV = cv.take<int32_t>(np);
    int32_t* CP = cv.take<int32_t>((size_t)np * KMAX);
    float* CF = cv.take<float>((size_t)np * KMAX);
    int2* BRANGE = cv.take<int2>(np);
    int4* BR = cv.take<int4>(nrec);
    float4* BF = cv.take<float4>(nrec);
    int2* BS = cv.take<int2>(nbs);
    float* TH = cv.take<float>(np);
    // contact lists kept beside the neighbour lists become partner-id lists first
    hipError_t e = launch_contacts_by_id(make_args(d_lf_), (int)p0, np, stream_);
    auto d2h = [&](void* dst, const void* src, size_t bytes) {
        if (e == hipSuccess && bytes) e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream_);
    };
    d2h(PV, S.PV[cur_] + 2 * p0, 2 * sizeof(float4) * np);   // cur_ == lcur_
    d2h(A, S.A + p0, sizeof(float4) * np);
    d2h(TH, S.TH + p0, sizeof(float) * np);
    d2h(M, S.M + p0, sizeof(int32_t) * np);
    d2h(INV, INV_ + p0, sizeof(int32_t) * np);
    const bool rec = with_data_ && sh.params.record_data && v.data;
    if (rec) {
        d2h(D0, S.DATA0 + p0, sizeof(float4) * np);
        d2h(D1, S.DATA1 + p0, sizeof(float4) * np);
    }
    if (v.contacts) {
        const ContactLists& L = lists_[lcur_];
        // the slot-major lists, KMAX rows in one strided copy each
        if (e == hipSuccess)
            e = hipMemcpy2DAsync(CP, sizeof(int32_t) * np, L.CP + p0, sizeof(int32_t) * slot_cap_,
                                 sizeof(int32_t) * np, KMAX, hipMemcpyDeviceToHost, stream_);
        if (e == hipSuccess)
            e = hipMemcpy2DAsync(CF, sizeof(float) * np, L.CF + p0, sizeof(float) * slot_cap_, sizeof(float) * np,
                                 KMAX, hipMemcpyDeviceToHost, stream_);
        d2h(BRANGE, BOFF_ + p0, sizeof(int2) * np);
        if (sh.nrec) {
            d2h(BR, BR_ + sh.r0, sizeof(int4) * sh.nrec);
            d2h(BF, BF_ + sh.r0, sizeof(float4) * sh.nrec);
        }
    }
    if (sh.nb) d2h(BS, BS_ + sh.b0, sizeof(int2) * sh.nb);
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    int rc = check(e, "download");
    if (rc) return rc;
    for (int32_t i = 0; i < n; i++) {
        const int64_t s = (int64_t)INV[i] - p0;
        if (s < 0 || s >= np) { set_error("corrupt slot map"); return DEM_ESTATE; }
        const float4 p = P[s], vv = V[s], aa = A[s];
        if (v.pos) { v.pos[2 * i] = p.x; v.pos[2 * i + 1] = p.y; }
        if (v.vel) { v.vel[2 * i] = vv.x; v.vel[2 * i + 1] = vv.y; }
        if (v.rot) v.rot[i] = TH[s];
        if (v.rot_vel) v.rot_vel[i] = vv.z;
        if (v.acc) { v.acc[2 * i] = aa.x; v.acc[2 * i + 1] = aa.y; }
        if (v.rot_acc) v.rot_acc[i] = aa.z;
        if (v.radii) v.radii[i] = p.z;
        if (v.forces) std::memcpy(v.forces + 6 * (size_t)i, sh.forces.data() + 6 * (size_t)i, 24);
        if (v.fixity) std::memcpy(v.fixity + 6 * (size_t)i, sh.fixity.data() + 6 * (size_t)i, 24);
        if (v.material_pointers) v.material_pointers[i] = sh.matptr[i];
        if (rec) {
            float* d = v.data + 7 * (size_t)i;
            d[0] = D0[s].x; d[1] = D0[s].y; d[2] = D0[s].z; d[3] = D0[s].w;
            d[4] = D1[s].x; d[5] = D1[s].y; d[6] = D1[s].z;
        }
    }
    if (v.bonds) {
        if (v.n_bonds != sh.nb) { set_error("view.n_bonds does not match the scene"); return DEM_EINVAL; }
        for (int32_t k = 0; k < sh.nb; k++) {
            v.bonds[3 * k] = BS[k].x;
            v.bonds[3 * k + 1] = sh.bond_angle[k];
            v.bonds[3 * k + 2] = sh.bond_len[k];
        }
    }
    bool overflow = false;
    if (v.contacts) {
        const int K = v.contact_slots;
        Slot* tab = reinterpret_cast<Slot*>(v.contacts);
        std::vector<Slot> row;
        for (int32_t i = 0; i < n; i++) {
            row.clear();
            const int64_t s = (int64_t)INV[i] - p0;
            for (int q = 0; q < cc_of((uint32_t)M[s]) && q < KMAX; q++) {
                Slot x;
                x.a = i;
                x.b = CP[(size_t)q * np + s];
                x.ft = CF[(size_t)q * np + s];
                x.fbt = bitsf(-1);
                x.tb = bitsf(-1);
                x.bonded = -1;
                row.push_back(x);
            }
            for (int32_t q = BRANGE[i].x - (int32_t)sh.r0; q < BRANGE[i].y - (int32_t)sh.r0; q++) {
                const int4 br = BR[q];
                if (br.w == 0) continue;
                const float4 bf = BF[q];
                Slot x;
                x.a = i;
                x.b = br.x;
                x.ft = bf.x;
                x.fbt = bf.y;
                x.tb = bf.z;
                x.bonded = br.z;
                row.push_back(x);
            }
            std::stable_sort(row.begin(), row.end(), [](const Slot& x, const Slot& y) { return x.b < y.b; });
            if ((int)row.size() > K) overflow = true;
            Slot* out = tab + (size_t)i * K;
            for (int k = 0; k < K; k++) {
                if (k < (int)row.size()) out[k] = row[k];
                else std::memset(&out[k], 0xff, sizeof(Slot));
            }
        }
    }
    if (overflow) { set_error("a particle has more occupied slots than contact_slots"); return DEM_ECAPACITY; }
    return DEM_OK;
}

// ---------------------------------------------------------------------------
int Engine::kinetic_energy(int32_t sid, double* out) {
    if (sid < 0 || sid >= (int32_t)scenes_.size() || !scenes_[sid].live) { set_error("no such scene"); return DEM_EINVAL; }
    const SceneHost& sh = scenes_[sid];
    std::vector<double> ke(sh.np);
    hipError_t e = launch_kinetic(arrays(set_), cur_, sh.m0, sh.nm, (int)sh.p0, sh.np, d_ke_, stream_);
    if (e == hipSuccess) e = hipMemcpyAsync(ke.data(), d_ke_, sizeof(double) * sh.np, hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    int rc = check(e, "kinetic energy");
    if (rc) return rc;
    double s = 0.0;
    for (int32_t i = 0; i < sh.n; i++) s += ke[i];
    *out = s;
    return DEM_OK;
}

// No allocation on any path: the per-id energies (the kernel of kinetic_energy) and the
// partial sums live in the static scratch, the total is copied into *out.
int Engine::kinetic_energy_noalloc(int32_t sid, double* out) {
    if (sid < 0 || sid >= (int32_t)scenes_.size() || !scenes_[sid].live) {
        set_error("kinetic_energy_noalloc: no such scene");
        return DEM_EINVAL;
    }
    const SceneHost& sh = scenes_[sid];
    if ((int64_t)sh.np > (int64_t)DEM_KE_NOALLOC_MAX) {
        set_error("dem_kinetic_energy_noalloc: the scene has more particle slots than the static scratch holds "
                  "(DEM_KE_NOALLOC_MAX = 4194304)");
        return DEM_ECAPACITY;
    }
    double* buf = static_scratch(device_);
    if (!buf) {
        set_error("dem_kinetic_energy_noalloc: static scratch not resolved on this device");
        return DEM_ESTATE;
    }
    hipError_t e = launch_kinetic(arrays(set_), cur_, sh.m0, sh.nm, (int)sh.p0, sh.np, buf, stream_);
    if (e == hipSuccess) e = launch_ke_reduce(buf, sh.n, out, stream_);
    if (e != hipSuccess) {
        set_error("dem_kinetic_energy_noalloc: HIP error in the kinetic-energy kernels");
        return e == hipErrorOutOfMemory ? DEM_ENOMEM : DEM_EDEVICE;
    }
    return DEM_OK;
}

int Engine::group_sums(int32_t sid, const int32_t* ids, int32_t n_ids, double* out) {
    if (sid < 0 || sid >= (int32_t)scenes_.size() || !scenes_[sid].live) { set_error("no such scene"); return DEM_EINVAL; }
    const SceneHost& sh = scenes_[sid];
    if (n_ids < 0) { set_error("group sums: n_ids < 0"); return DEM_EINVAL; }
    for (int32_t k = 0; k < n_ids; k++)
        if (ids[k] < 0 || ids[k] >= sh.n) {
            set_error("group sums: id " + std::to_string(ids[k]) + " outside [0, " + std::to_string(sh.n) + ")");
            return DEM_EINVAL;
        }
    // ids, then 64 x 7 partials and the 7 totals, in one buffer grown on demand
    const int64_t need = (int64_t)(GROUP_SCRATCH_DOUBLES * sizeof(double)) + (int64_t)n_ids * sizeof(int32_t);
    hipError_t e = hipSuccess;
    if (group_bytes_ < need) {
        if (d_group_) hipFree(d_group_);
        d_group_ = nullptr;
        group_bytes_ = 0;
        e = hipMalloc(&d_group_, (size_t)need);
        if (e != hipSuccess) return check(e, "group sums");
        group_bytes_ = need;
    }
    double* partials = reinterpret_cast<double*>(d_group_);
    double* total = partials + GROUP_SCRATCH_DOUBLES - 8;
    int32_t* d_ids = reinterpret_cast<int32_t*>(partials + GROUP_SCRATCH_DOUBLES);
    double host[8] = {0};
    if (n_ids > 0) {
        e = hipMemcpyAsync(d_ids, ids, sizeof(int32_t) * (size_t)n_ids, hipMemcpyHostToDevice, stream_);
        const int rec = (with_data_ && sh.params.record_data) ? 1 : 0;
        if (e == hipSuccess)
            e = launch_group_sums(arrays(set_), cur_, sh.m0, sh.nm, (int)sh.p0, d_ids, n_ids, rec, partials, total,
                                  stream_);
        if (e == hipSuccess) e = hipMemcpyAsync(host, total, sizeof(double) * 7, hipMemcpyDeviceToHost, stream_);
        if (e == hipSuccess) e = hipStreamSynchronize(stream_);
        int rc = check(e, "group sums");
        if (rc) return rc;
    }
    for (int c = 0; c < 7; c++) out[c] = host[c];
    out[7] = (double)n_ids;
    return DEM_OK;
}

int Engine::broken_bonds(int32_t sid, int64_t* out) {
    if (sid < 0 || sid >= (int32_t)scenes_.size() || !scenes_[sid].live) { set_error("no such scene"); return DEM_EINVAL; }
    const SceneHost& sh = scenes_[sid];
    std::vector<int2> BS(std::max(sh.nb, 1));
    hipError_t e = hipSuccess;
    if (sh.nb) e = hipMemcpyAsync(BS.data(), BS_ + sh.b0, sizeof(int2) * sh.nb, hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    int rc = check(e, "broken bonds");
    if (rc) return rc;
    int64_t c = 0;
    for (int32_t k = 0; k < sh.nb; k++) c += BS[k].x < 0;
    *out = c;
    return DEM_OK;
}

int Engine::counters(int32_t sid, int64_t* out, int32_t n_out) {
    if (sid < 0 || sid >= (int32_t)scenes_.size() || !scenes_[sid].live) { set_error("no such scene"); return DEM_EINVAL; }
    const SceneHost& sh = scenes_[sid];
    unsigned long long st[8] = {0}, nc = 0;
    hipError_t e = hipMemcpyAsync(st, STAT_, sizeof(st), hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess) e = hipMemsetAsync(d_cnt_, 0, sizeof(unsigned long long), stream_);
    if (e == hipSuccess) e = launch_count_contacts(sets_[set_].M, (int)sh.p0, sh.np, d_cnt_, stream_);
    if (e == hipSuccess) e = hipMemcpyAsync(&nc, d_cnt_, sizeof(nc), hipMemcpyDeviceToHost, stream_);
    int64_t nrec_alive = 0;
    std::vector<int4> BR(std::max(sh.nrec, 1));
    if (e == hipSuccess && sh.nrec) e = hipMemcpyAsync(BR.data(), BR_ + sh.r0, sizeof(int4) * sh.nrec, hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    int rc = check(e, "counters");
    if (rc) return rc;
    for (int32_t q = 0; q < sh.nrec; q++) nrec_alive += BR[q].w != 0;
    const int64_t vals[12] = {sh.steps, (int64_t)st[4], (int64_t)st[0], (int64_t)st[1], (int64_t)nc, nrec_alive,
                              sh.reorders, (int64_t)st[2], (int64_t)st[3], (int64_t)st[5], graph_captures_,
                              graph_replays_};
    for (int32_t i = 0; i < n_out && i < 12; i++) out[i] = vals[i];
    return DEM_OK;
}

int Engine::datum(int32_t sid, const int32_t* select, double* out) {
    if (sid < 0 || sid >= (int32_t)scenes_.size() || !scenes_[sid].live) { set_error("no such scene"); return DEM_EINVAL; }
    const SceneHost& sh = scenes_[sid];
    if (!(with_data_ && sh.params.record_data)) { set_error("get_datum needs record_data"); return DEM_ESTATE; }
    const int32_t n = sh.n;
    std::vector<float> pos(2 * (size_t)n), rot(n), data(7 * (size_t)n), vel(2 * (size_t)n), rv(n);
    dem_state_view v{};
    v.n = n;
    v.pos = pos.data();
    v.rot = rot.data();
    v.data = data.data();
    int rc = download(sid, v);
    if (rc) return rc;
    double sums[10] = {0};
    int64_t count = 0;
    for (int32_t i = 0; i < n; i++) {
        if (!select[i]) continue;
        count++;
        const float* d = data.data() + 7 * (size_t)i;
        sums[0] += pos[2 * i]; sums[1] += pos[2 * i + 1];
        sums[2] += d[4]; sums[3] += d[5]; sums[4] += rot[i]; sums[5] += d[6];
        sums[6] += std::fabs((double)d[0]); sums[7] += d[1]; sums[8] += d[2]; sums[9] += d[3];
    }
    if (!count) { set_error("nothing selected"); return DEM_ESTATE; }
    for (int k = 0; k < 10; k++) out[k] = k == 6 ? sums[k] : sums[k] / (double)count;
    return DEM_OK;
}

// Stage views: bin the current positions, scan, scatter and sort each cell by id.
int Engine::debug_broad_phase(int32_t sid, int32_t* keys, int32_t* sorted_ids, int32_t* cell_start) {
    if (sid < 0 || sid >= (int32_t)scenes_.size() || !scenes_[sid].live) { set_error("no such scene"); return DEM_EINVAL; }
    int rc = write_scene_table();
    if (rc) return rc;
    const SceneHost& sh = scenes_[sid];
    // this re-bins one scene over the whole cell scan: the other scenes' cell lists
    // (the neighbour-list fallback reads them) are rebuilt after the next drift
    rb_force_ = true;
    std::vector<int2> lf(scene_cap_, make_int2(0, 0));
    lf[sid] = make_int2((int)(SF_ACTIVE | SF_BIN), 0);
    hipError_t e = hipMemcpyAsync(d_lf_, lf.data(), sizeof(int2) * scene_cap_, hipMemcpyHostToDevice, stream_);
    StepArgs args = make_args(d_lf_);
    const int ncl = sh.params.grid.w * sh.params.grid.h;
    if (e == hipSuccess) e = launch_bin(args, active_blocks(), block_, stream_);
    if (e == hipSuccess) e = launch_scan(arrays(set_), (int)hw_cells_, nullptr, stream_);
    if (e == hipSuccess) e = launch_scatter(args, active_blocks(), block_, stream_);
    if (e == hipSuccess) e = launch_cell_sort(args, (int)sh.c0, (int)(sh.c0 + ncl + 1), (int)sh.p0, stream_);
    std::vector<int32_t> KEY(sh.np), SIDX(sh.np), START(ncl + 2);
    std::vector<int32_t> ID(sh.np);
    if (e == hipSuccess) e = hipMemcpyAsync(KEY.data(), KEY_ + sh.p0, sizeof(int32_t) * sh.np, hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess) e = hipMemcpyAsync(SIDX.data(), SIDX_ + sh.p0, sizeof(int32_t) * sh.np, hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess) e = hipMemcpyAsync(START.data(), START_ + sh.c0, sizeof(int32_t) * (ncl + 1), hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess) e = hipMemcpyAsync(ID.data(), sets_[set_].ID + sh.p0, sizeof(int32_t) * sh.np, hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    if ((rc = check(e, "debug broad phase"))) return rc;
    // the cell lists now describe the current storage; the next step re-bins anyway
    for (int32_t s = 0; s < sh.np; s++) {
        const int32_t id = ID[s];
        if (id < sh.n && keys) keys[id] = KEY[s] - (int32_t)sh.c0;
    }
    if (sorted_ids)
        for (int32_t e2 = 0; e2 < sh.n; e2++) sorted_ids[e2] = ID[SIDX[e2] - sh.p0];
    if (cell_start)
        for (int32_t c = 0; c <= ncl; c++) cell_start[c] = START[c] - START[0];
    return DEM_OK;
}

int Engine::debug_contacts(int32_t sid, int32_t max_per, int32_t* counts, int32_t* partners) {
    if (sid < 0 || sid >= (int32_t)scenes_.size() || !scenes_[sid].live) { set_error("no such scene"); return DEM_EINVAL; }
    const SceneHost& sh = scenes_[sid];
    const int np = sh.np;
    std::vector<int32_t> M(np);
    std::vector<int32_t> CP((size_t)np * KMAX), INV(np);
    hipError_t e = launch_contacts_by_id(make_args(d_lf_), (int)sh.p0, np, stream_);
    if (e == hipSuccess)
        e = hipMemcpyAsync(M.data(), sets_[set_].M + sh.p0, sizeof(int32_t) * np, hipMemcpyDeviceToHost, stream_);
    for (int q = 0; q < KMAX && e == hipSuccess; q++)
        e = hipMemcpyAsync(CP.data() + (size_t)q * np, lists_[lcur_].CP + (int64_t)q * slot_cap_ + sh.p0,
                           sizeof(int32_t) * np, hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess) e = hipMemcpyAsync(INV.data(), INV_ + sh.p0, sizeof(int32_t) * np, hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    int rc = check(e, "debug contacts");
    if (rc) return rc;
    bool over = false;
    for (int32_t i = 0; i < sh.n; i++) {
        const int64_t s = (int64_t)INV[i] - sh.p0;
        const int c = cc_of((uint32_t)M[s]);
        counts[i] = c;
        for (int k = 0; k < max_per; k++)
            partners[(size_t)i * max_per + k] = (k < c && k < KMAX) ? CP[(size_t)k * np + s] : -1;
        if (c > max_per) over = true;
    }
    if (over) { set_error("max_per smaller than a contact count"); return DEM_ECAPACITY; }
    return DEM_OK;
}

int Engine::bond_pairs(int32_t sid, int32_t max_bonds, float factor, float gap, int32_t* pairs, int32_t cap,
                       int32_t* n_out) {
    if (sid < 0 || sid >= (int32_t)scenes_.size() || !scenes_[sid].live) { set_error("no such scene"); return DEM_EINVAL; }
    if (max_bonds <= 0 || max_bonds > 16) { set_error("max_bonds must be in [1, 16]"); return DEM_EINVAL; }
    if (!(factor > 0.0f) || !(gap >= 0.0f) || !std::isfinite(factor) || !std::isfinite(gap)) {
        set_error("bond pairs: factor must be > 0 and gap >= 0, both finite");
        return DEM_EINVAL;
    }
    const SceneHost& sh = scenes_[sid];
    const int n = sh.n;
    // bin the current positions (also refreshes the cell lists)
    int rc = debug_broad_phase(sid, nullptr, nullptr, nullptr);
    if (rc) return rc;
    std::vector<float4> PV(2 * (size_t)sh.np);   // {position, velocity record} per slot
    hipError_t e = hipMemcpyAsync(PV.data(), sets_[set_].PV[cur_] + 2 * sh.p0, 2 * sizeof(float4) * sh.np,
                                  hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    if ((rc = check(e, "bond pairs"))) return rc;
    float rmax = 0.0f;
    for (int i = 0; i < sh.np; i++) rmax = std::max(rmax, PV[2 * (size_t)i].z);
    const int reach = std::max(1, (int)std::ceil(((double)factor * 2.0 * rmax + gap) / sh.params.grid.cell_size));
    int32_t* d_cnt = nullptr;
    int32_t* d_j = nullptr;
    e = dalloc(&d_cnt, std::max(n, 1));
    if (e == hipSuccess) e = dalloc(&d_j, (int64_t)std::max(n, 1) * max_bonds);
    std::vector<int2> lf(scene_cap_, make_int2(0, 0));
    lf[sid] = make_int2((int)SF_ACTIVE, 0);
    if (e == hipSuccess) e = hipMemcpyAsync(d_lf_, lf.data(), sizeof(int2) * scene_cap_, hipMemcpyHostToDevice, stream_);
    if (e == hipSuccess) e = launch_bond_pairs(make_args(d_lf_), active_blocks(), block_, reach, max_bonds, factor, gap, d_cnt, d_j,
                                                  stream_);
    std::vector<int32_t> cnt(std::max(n, 1)), jj((size_t)std::max(n, 1) * max_bonds);
    if (e == hipSuccess) e = hipMemcpyAsync(cnt.data(), d_cnt, sizeof(int32_t) * n, hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess) e = hipMemcpyAsync(jj.data(), d_j, sizeof(int32_t) * (size_t)n * max_bonds, hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    hipFree(d_cnt);
    hipFree(d_j);
    if ((rc = check(e, "bond pairs"))) return rc;
    int64_t k = 0;
    bool over = false;
    for (int i = 0; i < n; i++)
        for (int q = 0; q < cnt[i]; q++) {
            if (k < cap) { pairs[2 * k] = i; pairs[2 * k + 1] = jj[(size_t)i * max_bonds + q]; }
            else over = true;
            k++;
        }
    *n_out = (int32_t)std::min<int64_t>(k, INT32_MAX);
    if (over) { set_error("bond_cap too small"); return DEM_ECAPACITY; }
    return DEM_OK;
}

}  // namespace psim
