// dem_device.h - device-side data layout shared by the kernels and the host context.
//
// One launch can advance several independent scenes ("batch").  Every scene owns a
// contiguous range of particle slots, cells and bonds in the global pools; a block
// never straddles two scenes (scene slot ranges are padded to the block size), so a
// block reads its scene descriptor once.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>
#include "../../include/dem.h"

namespace psim {

// Slots per particle, contacts and live bond records together (the reference's contact
// table, 2D_Simulation.wgsl:125): contact lists hold at most KMAX less the live records.
constexpr int KMAX = 14;
constexpr int NLMAX = 16;        // neighbour-list candidates per particle (more: stencil fallback)

// Row 0 of a neighbour list: the candidate count in the top bits (NL_OVERFLOW: more
// than NLMAX, stencil fallback) and the first candidate's slot in the low 26 bits
// (slots stay below 2^26), so one load gives the count and the first candidate.
constexpr int NL_SHIFT = 26;
constexpr int NL_OVERFLOW = 31;
__host__ __device__ inline int nl_word(int count, int first) {
    return ((count < 0 ? NL_OVERFLOW : count) << NL_SHIFT) | first;
}
__host__ __device__ inline int nl_count(int w) {
    const int c = (int)((unsigned)w >> NL_SHIFT);
    return c == NL_OVERFLOW ? -1 : c;
}
__host__ __device__ inline int nl_first(int w) { return w & ((1 << NL_SHIFT) - 1); }
constexpr int DATA_WORDS = 7;
constexpr int REC_DEAD = INT32_MIN;   // bonded word of a dead bond record in BRB
constexpr int REC_GONE = INT32_MIN + 1;   // dead, and its result already says so (RF.w = REC_RF_DEAD)
// last word of a bond-pass result (RF.w, int bits): the partner id of a live record, or
constexpr int REC_RF_DEAD = -1;       // a dead record (no force)
constexpr int REC_RF_DIED = -2;       // a record that died in this launch (its partner joins the owner's list)

// Break word of a bond: the bits of BLS.w, so that the bond pass reads a bond's break
// state in the same 16-byte load as its rest length and strengths.  0 while intact;
// after a break, bit 31 set, bits 2..30 the step of the break mod 2^29 and bits 0..1
// the code + 12 for the codes -11..-9 of the bond laws (2D_Simulation.wgsl:322-324),
// or 0 for any other code (an uploaded one), which is then read from BS.x.
constexpr uint32_t BRK_SET = 1u << 31;
constexpr uint32_t BRK_STEP_MASK = (1u << 29) - 1u;
__host__ __device__ inline uint32_t brk_word(int code, uint32_t step) {
    const uint32_t c = (code >= -11 && code <= -9) ? (uint32_t)(code + 12) : 0u;
    return BRK_SET | ((step & BRK_STEP_MASK) << 2) | c;
}
// Whether break word w was written before step `step` (a break is seen by the bond's
// other endpoint from the next step on); equal steps mod 2^29 read as the same step.
__host__ __device__ inline bool brk_before(uint32_t w, uint32_t step) {
    return (w & BRK_SET) && ((w >> 2) & BRK_STEP_MASK) != (step & BRK_STEP_MASK);
}

// Word V.w of the velocity record (bits, not a float):
//  VW_EXPLICIT  the displacement of the last drift, {del_x, del_y, del_rot}, is the
//               explicit record DX of the same buffer; otherwise it is the stored
//               velocity times dT (2D_LOM.wgsl:97-99), bit for bit, because no wall
//               changed the velocity after it was formed.  Set after a wall bounce,
//               when the drift kept the loaded displacement (coll_cont[3] != 0), after
//               a force half without a drift (the velocity is kicked again) and at upload.
//  VW_MASK      (bits 16..29) where M has PF_ALIGNED: bit k set when neighbour-list
//               candidate k is a contact, its tangential force in FT row k.
//  VW_PATH      (bits 0..15) distance travelled since the neighbour lists were built, an
//               upper bound in units of 0.45 skin / 2^14 (sum of |dx| + |dy| per drift,
//               rounded up, saturating): the position at the build is read only once it
//               passes 2^14 (a drift adds at most one unit of rounding).
constexpr uint32_t VW_EXPLICIT = 1u << 31;
constexpr int VW_MASK_SHIFT = 16;
constexpr uint32_t VW_MASK = ((1u << KMAX) - 1u) << VW_MASK_SHIFT;
constexpr uint32_t VW_PATH = (1u << 16) - 1u;
constexpr uint32_t VW_PATH_TRIG = 1u << 14;
constexpr uint32_t VW_PATH_SAT = 1u << 15;
static_assert(VW_MASK_SHIFT + KMAX <= 30, "the contact mask sits below VW_EXPLICIT");

// per-particle flag bits (word M bits 16..24; bits 0..15 hold the material id)
enum : uint32_t {
    PF_KICK_X = 1u << 16,   // fixity.x_vel == 0: velocity integrates (2D_LOM.wgsl:88)
    PF_KICK_Y = 1u << 17,
    PF_KICK_R = 1u << 18,
    PF_WALL_X = 1u << 19,   // fixity.x_vel != 1: walls act on x (2D_LOM.wgsl:119)
    PF_WALL_Y = 1u << 20,
    PF_WALL_R = 1u << 21,   // fixity.rot_vel != 1 (only a dead write in the reference)
    PF_EXT = 1u << 22,      // nonzero external force / moment
    PF_INACTIVE = 1u << 23, // r == 0 or beyond the active count
    PF_LARGE = 1u << 24,    // a large body: on the scene's large-body list, not found through cells
    PF_MAT_MASK = 0xffffu,
};
// Bits 25..29 of M: the particle's contact count (entries of its contact list), kept
// beside the flags so that the step kernel reads it with them; rewritten only when it
// changes (neighbours read the word for its material bits, which never change).
constexpr int PF_CC_SHIFT = 25;
constexpr uint32_t PF_CC_MASK = 0x1fu << PF_CC_SHIFT;
__host__ __device__ inline int cc_of(uint32_t mx) { return (int)((mx & PF_CC_MASK) >> PF_CC_SHIFT); }
// Bit 30 of M: the contact list is held beside the neighbour list (the contact mask in
// the velocity word, VW_MASK, and the tangential forces in FT rows of the candidates)
// instead of in CP / CF.  The list path of the step kernel writes this form and the list
// build moves it beside the new list; the re-sort, the other paths, reads from the host
// and a list build that cannot hold it turn it back into the partner-id list
// (contacts_from_aligned).
constexpr uint32_t PF_ALIGNED = 1u << 30;

// per-step flags (one word per scene and step)
enum : uint32_t {
    SF_SIM = 1u,        // run the Sim half (forces, first half-kick)
    SF_LOM = 2u,        // run the LOM half of the next step (second half-kick, drift, walls)
    SF_NARROW = 4u,     // the Sim half runs the narrow phase (collisions && coll_cont[0])
    SF_BIN = 8u,        // the next launch runs a narrow phase (the debug binning bins such scenes)
    SF_DELOVR = 16u,    // the LOM half keeps del_pos/del_rot (coll_cont[3] != 0)
    SF_ACTIVE = 32u,    // scene takes part in this launch
};

struct SceneDev {
    dem_uniform u;
    int32_t p0, np;        // particle slots [p0, p0 + np)
    int32_t c0;            // first cell; cells [c0, c0 + w*h], the last one collects inactive particles
    int32_t w, h;
    float cs, base_x, base_y;
    float inv_cs;          // 1 / cs (f32), the cell index multiplier
    float half_w, half_h;  // hor_bound / 2, vert_bound / 2 (exact halvings)
    int32_t b0, nb;        // bonds [b0, b0 + nb)
    int32_t m0, nm;        // materials [m0, m0 + nm)
    int32_t use_f64;
    int32_t record;        // maintain data[7]
    int32_t has_strength;  // per-bond strengths present
    int32_t n_real;        // particles of the scene; slots [p0+n_real, p0+np) hold inactive padding
    float skin;            // neighbour-list skin: cell side minus (2 r_small + rounding margin)
    float trig2;           // squared displacement since the list build that forces a rebuild
    float pscale;          // 2^14 / (0.45 skin): path units per unit length (1e30 for no skin)
    float rsmax;           // largest radius of the discs that are not large bodies
    int32_t lb0, nlb;      // large bodies: particle ids LB[lb0, lb0 + nlb)
    float rho1, kn1, ks1;  // density and stiffnesses of material 0 (the whole scene when nm == 1)
    float cn2;             // contact dashpot factor 2 zeta of dem_params.restitution (0: none)
    float kne1, kse1;      // 1 / (1 / k + 1 / k) of material 0's normal and shear stiffness (f32)
};

// Large bodies: at most LB_MAX per scene, the discs above the first gap of a factor 2
// in the radii sorted in decreasing order (a striker among small discs).  The grid is
// sized for the others; large bodies pair with everything through the list.
constexpr int LB_MAX = 64;
// Small discs overlapping one large body, found by the large-body search each launch with
// a narrow phase (more: the large body's thread walks the stencil itself).
constexpr int LBH_MAX = 128;

// A device array addressed by 32-bit element index: the byte offset is formed in
// 32 bits, so loads and stores take the SGPR-base + VGPR-offset form and need no
// 64-bit address arithmetic per access.  Every pool is below 4 GiB (the engine
// checks its sizes at allocation).
// STRIDE: the bytes from one element to the next (records interleaved with others).
template <class T, uint32_t STRIDE = (uint32_t)sizeof(T)>
struct Pool {
    T* p = nullptr;
    Pool() = default;
    __host__ __device__ Pool(T* q) : p(q) {}
    template <class I>
    __device__ __forceinline__ T& operator[](I i) const {
        return *reinterpret_cast<T*>(reinterpret_cast<char*>(const_cast<typename std::remove_const<T>::type*>(p)) +
                                     (uint32_t)i * STRIDE);
    }
};

// A slot's position record and velocity record of the same parity sit side by side in
// one 32-byte record {x, y, r, id | vx, vy, w, word}: a neighbour's gather of both (and
// of its id) reads one cache line.
constexpr uint32_t PV_STRIDE = 32;

struct DevArrays {
    // double-buffered neighbour-visible state.  Positions: P[cur] read, P[cur ^ 1]
    // written by a drift.  Velocities and explicit displacements flip after every launch
    // with the contact lists: V / DX read, Vn / DXn written.
    Pool<float4, PV_STRIDE> P[2];   // {x, y, r, id bits}: the id travels with the position its neighbours read
    Pool<float4, PV_STRIDE> V;      // {vx, vy, w, word: VW_EXPLICIT | path}  (after a LOM half: the mid-step velocity)
    Pool<float4, PV_STRIDE> Vn;     // V[k] and P[k] of one parity interleave (PV_STRIDE)
    Pool<float4> DX;        // {del_x, del_y, del_rot, 0}, valid where V.w has VW_EXPLICIT
    Pool<float4> DXn;
    Pool<float4> A;         // {ax, ay, alpha, 0}
    Pool<float> TH;         // rotation, read and written by the particle's own thread only
    Pool<int32_t> M;        // material | flags | contact count (the word the step kernel reads)
    Pool<int32_t> ID;       // particle id within the scene (read for partners only)
    Pool<float4> EXT;       // external {fx, fy, moment, 0}
    Pool<float4> DATA0;     // {Σfn, Σft, ΣM, data3}
    Pool<float4> DATA1;     // {v*x, v*y, w*, 0}
    // contact lists, slot-major [KMAX][cap]: the partner and force columns are
    // double-buffered (every step kernel reads CP/CF and writes CPn/CFn); the count is
    // part of M (cc_of)
    Pool<int32_t> CP;       // partner id (scene-local), ascending
    Pool<float> CF;         // tangential force
    Pool<int32_t> CPn;
    Pool<float> CFn;
    // [KMAX][cap] tangential force of neighbour-list candidate k where M has PF_ALIGNED,
    // read and written in place by the particle's own thread (valid where VW_MASK says)
    Pool<float> FT;
    int64_t cap;       // particle pool size (slot stride of CP / CF)
    // bond endpoint records: ranges indexed by (scene p0 + particle id), each list sorted
    // by partner id; records never move
    Pool<int2> BOFF;        // [cap] record range [x, y) of (scene p0 + particle id)
    Pool<int4> BR;          // {partner id, bond id, bonded, alive}
    Pool<int2> BRB;         // the bond pass's copy: {bond id, bonded, or REC_DEAD once dead, REC_GONE after}
    Pool<float4> BF;        // {ft, fbt, theta_b, 0}
    Pool<int2> BS;          // per bond {index word, step of break} (download and counts)
    Pool<float4> BLS;       // per bond {rest length, normal strength, shear strength, break word bits}
    Pool<const dem_material> MAT;
    // bond pass: owner of each record (global particle id, -1: none) and its results
    Pool<int32_t> LB;       // large-body particle ids, LB_MAX per scene
    Pool<int32_t> LBH;      // [LB pool][LBH_MAX] slots of the small discs overlapping each large body
    Pool<int32_t> LBHC;     // [LB pool] their count, or -1 (more than LBH_MAX)
    Pool<int32_t> RO;
    Pool<int32_t> RS, RP;   // slots of each record's owner (-1: none) and partner
    Pool<int2> BRS;         // [cap] per slot: its record range (moves with the particle)
    Pool<float4> RF;        // {fx, fy, moment, partner id bits (REC_RF_DEAD / REC_RF_DIED: none)}
    Pool<float4> RD;        // [2 per record] data[0..2] updates (with data recording)
    // broad phase
    Pool<int32_t> KEY;      // cell of each slot (global cell index)
    Pool<int32_t> RANK;     // rank inside the cell from the counting atomics
    Pool<int32_t> CNT;      // [ncells_total] counts, zeroed by the scan
    Pool<int32_t> START;    // [ncells_total + 1]
    Pool<int32_t> SIDX;     // slots in cell order
    Pool<int32_t> INV;      // [cap] slot of (scene p0 + particle id)
    // neighbour lists (Verlet lists with a skin), rebuilt on the device when a particle
    // has moved more than 0.45 skin since the last build; part of the particle set
    // (a re-sort moves them, renaming the candidate slots)
    Pool<int32_t> NL;       // [NLMAX][cap] candidate slots within r_a + r_b + skin at the build;
                            // row 0 packs the count with the first slot (nl_word)
    Pool<float2> XR;        // [cap] position at the build
    Pool<int32_t> RB;       // [1] rebuild flag
    // counters
    Pool<unsigned long long> STAT;  // [0] contact overflow, [1] bond overflow, [2] scan timeouts
    // per-tile cell counts of the scan (reduce-then-scan, no inter-block waits)
    Pool<int32_t> TSUM;
    int32_t ntiles;
};

struct StepArgs {
    const SceneDev* scenes;
    const int32_t* blk_scene;  // scene of each block (-1: none)
    const int2* lf;            // per scene {step flags, global step index of the Sim half}
    DevArrays a;
    int cur;                   // which P buffer holds the current positions
    int flip;                  // cur flips after this launch: every slot that does not drift
                               // carries its P over to the other buffer (every launch flips, so the
                               // position and list parities keep one relation)
    int lbs;                   // block size of the step kernel's launch (a kernel argument: the
                               // dispatch packet's copy costs a vector load per wave)
    int one_scene;             // the only live scene, or -1: the step kernel then skips the
                               // block -> scene table load
};

struct ReorderArgs {
    const SceneDev* scenes;
    const int32_t* blk_scene;
    const int2* lf;
    DevArrays src;  // current arrays
    DevArrays dst;  // destination arrays
    int32_t* order;     // [cap] destination slot -> source slot
    int32_t* newslot;   // [cap] source slot -> destination slot
    int cur;
    int with_ext, with_data;
};

}  // namespace psim
