// dem_kernels.hip - gfx950 kernels of the DEM time step.
//
// The reference runs a time step as two WGSL passes, LOM (2D_LOM.wgsl:80-201:
// half-kick, drift, walls, AABB binning) and Sim (2D_Simulation.wgsl:108-478:
// narrow phase, contact/bond laws, first half-kick).  Here one step is:
//
//   k_step_plain, k_step_plain_f64, k_step_bonds (one body, step_slot)
//              fused "Sim half of step k" + "LOM half of step k+1" per particle slot:
//              contact search over the particle's neighbour list (the slots within
//              r_a + r_b + skin at the last rebuild; the 3x3 stencil of that
//              rebuild's cell list when a list overflows), force laws in ascending
//              partner-id order, store_forces, then the next drift, which raises the
//              rebuild flag when the particle has moved more than 0.45 skin since the
//              rebuild (a path bound in the velocity word gates the exact test)
//              (stages 2a, 4, 5).  Per disc it streams P (x, y, r, id), TH (rotation) and V
//              (velocity, contact mask, path word) in and out and reads the flag word and
//              the list word; the contacts' tangential forces sit beside the neighbour
//              list (FT row k for candidate k), read and rewritten in place; the drift's
//              displacement is re-formed as velocity x dT unless an explicit record DX
//              holds it.  The first launch of a dem_step call runs
//              only the LOM half and the last one only the Sim half, so the state
//              between calls is the reference's post-Sim state.
//   neighbour-list rebuild, after every drift, each kernel a no-op unless the flag
//   is set: k_rb_bin (cell key + rank by a per-cell counting atomic, one atomic per
//              run of equal keys in a wave, tile totals likewise), k_scan_tiles
//              (exclusive prefix of the cell counts over tiles of 4096 cells, wave64
//              prefix in registers + LDS across waves; zeroes the counts), k_rb_scatter
//              (slot -> position START[key] + RANK in the cell-ordered slot list) and
//              k_build_nl (the lists, in cell-list order) - stages 1, 2b, 3 of the
//              reference's broad phase.
//   k_partition, k_reorder_move   every few rebuilds: move every per-particle
//              array, the neighbour lists included, into cell-list order, each unit of
//              64 blocks partitioned into particles with list candidates first (k_cell_sort
//              orders each cell by id for the debug view of the broad phase only).
//
// Arithmetic follows the WGSL expression by expression (operand order, no fused
// multiply-add: the library is built with -ffp-contract=off; sqrt and division are
// the correctly rounded sequences), so the HIP path reproduces the CPU restatement
// of the reference bit for bit when both sum contacts in the same order (ascending
// partner id).
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <stdint.h>

#include <algorithm>

#include "dem_device.h"

namespace psim {

constexpr float PI_F = 3.14159265358979323846f;
constexpr float G_F = 9.81f;

__device__ __forceinline__ float lenf(float x, float y) {
    float xx = x * x;
    float yy = y * y;
    return sqrtf(xx + yy);
}
__device__ __forceinline__ double lend(double x, double y) {
    double xx = x * x;
    double yy = y * y;
    return sqrt(xx + yy);
}
__device__ __forceinline__ int sign_i(float x) { return x > 0.0f ? 1 : (x < 0.0f ? -1 : 0); }
__device__ __forceinline__ int f2i(float v) {
    if (!(v == v)) return 0;
    if (v >= 2147483520.0f) return INT32_MAX;
    if (v <= -2147483648.0f) return INT32_MIN;
    return (int)v;
}

// Centre cell of a position in the GridInfo convention (2D_LOM.wgsl:167-178 applied
// to the centre): column from the left edge, row from the top edge, clamped to the
// grid (NaN -> 0).  Multiplies by the f32 reciprocal of the cell side; clamping in
// the float domain before truncation equals truncate-then-clamp.
__device__ __forceinline__ int cell_of(const SceneDev& S, float x, float y, int& cx, int& cy) {
    cx = (int)fminf(fmaxf((x - S.base_x) * S.inv_cs, 0.0f), (float)(S.w - 1));
    cy = (int)fminf(fmaxf((S.base_y - y) * S.inv_cs, 0.0f), (float)(S.h - 1));
    return cy * S.w + cx;
}

struct Self {
    float x, y, r, drot, dx, dy;
    float kn, ks;
    float kne, kse;  // series stiffnesses against material 0 when the scene has one material
    bool one;        //   (the same f32 expression, formed once on the host), else unused
    float rho;      // density (mass of the dashpot)
    float cn2;      // dashpot factor 2 zeta (0: no dashpot, the reference's law)
    float dt;
};

__device__ __forceinline__ uint32_t vword(const float4& v) { return __float_as_uint(v.w); }

// Displacement of the last drift, {del_x, del_y, del_rot, 0}, of the particle in slot j
// whose current velocity record is v: the explicit record where the velocity word says
// so, else the velocity times dT (the drift formed it so and no wall changed the
// velocity afterwards).
__device__ __forceinline__ float4 del_from(const DevArrays& a, int j, const float4& v, float dt) {
    if (vword(v) & VW_EXPLICIT) return a.DX[j];
    return make_float4(v.x * dt, v.y * dt, v.z * dt, 0.0f);
}
__device__ __forceinline__ float4 del_of(const DevArrays& a, int j, float dt) { return del_from(a, j, a.V[j], dt); }

// Normal and shear stiffness and density of the partner in slot j: a one-material scene
// takes them from the scene table (the record of material 0), without a load.
__device__ __forceinline__ float3 partner_k(const DevArrays& a, const SceneDev& S, int j) {
    if (S.nm <= 1) return make_float3(S.kn1, S.ks1, S.rho1);
    const int mb = a.M[j] & PF_MAT_MASK;
    const dem_material MB = a.MAT[S.m0 + (mb < S.nm ? mb : 0)];
    return make_float3(MB.normal_stiffness, MB.shear_stiffness, MB.density);
}

// Carries the velocity record (and an explicit displacement) of slot s over to the
// buffers of the next launch.
__device__ __forceinline__ void carry_velocity(const DevArrays& a, int s) {
    const float4 v = a.V[s];
    a.Vn[s] = v;
    if (vword(v) & VW_EXPLICIT) a.DXn[s] = a.DX[s];
}

struct Data3 {
    float d0, d1, d2;  // data[0..2] of the calling particle
    __device__ void add(float a, float b, float c) {
        d0 += a;
        d1 += b;
        d2 += c;
    }
    __device__ void assign(float a, float b, float c) {
        d0 = a;
        d1 = b;
        d2 = c;
    }
};

// The data[0..2] updates of one bond record, logged in order (at most two) so that the
// owner replays them onto its running sums with the same roundings.
constexpr float DD_ADD = 1.0f, DD_SET = 2.0f;
struct DataLog {
    float4 op[2];   // {a, b, c, DD_ADD | DD_SET}; w = 0: no update
    int n;
    __device__ void add(float a, float b, float c) {
        if (n < 2) op[n] = make_float4(a, b, c, DD_ADD);
        n++;
    }
    __device__ void assign(float a, float b, float c) {
        if (n < 2) op[n] = make_float4(a, b, c, DD_SET);
        n++;
    }
};
__device__ __forceinline__ void replay(Data3& dd, const float4& o) {
    if (o.w == DD_ADD) dd.add(o.x, o.y, o.z);
    else if (o.w == DD_SET) dd.assign(o.x, o.y, o.z);
}

// ---- force laws (2D_Simulation.wgsl:264-432); `a` is the calling particle ----

// distance2 (:273-284): |p_a - p_b| - L.  l = the f32 length lenf(a.x - bx, a.y - by),
// which the f32 variant uses as is.
template <bool F64>
__device__ __forceinline__ float dist2(const Self& a, float bx, float by, float L, float l) {
    if (F64) {
        double dx = (double)a.x - (double)bx, dy = (double)a.y - (double)by;
        return (float)(lend(dx, dy) - (double)L);
    }
    return l - L;
}

// distance (:264-271): |p_a - p_b| - (r_a + r_b); l as for dist2
template <bool F64>
__device__ __forceinline__ float dist1(const Self& a, const float4& pb, float l) {
    if (F64) {
        double dx = (double)a.x - (double)pb.x, dy = (double)a.y - (double)pb.y;
        return (float)(lend(dx, dy) - ((double)a.r + (double)pb.z));
    }
    return l - (a.r + pb.z);
}

// relative tangential displacement, :352-360; db = the partner's {del_x, del_y, del_rot}
__device__ __forceinline__ float rel_tangent(const Self& a, const float4& pb, const float4& db, float tx, float ty) {
    float dra = a.drot * a.r;
    float drb = db.z * pb.z;
    float rtx = db.x - a.dx, rty = db.y - a.dy;
    float rr = drb + dra;
    return (rtx * tx + rty * ty) + rr;
}

// linear_model, :338-373.  L = r_a + r_b (contact) or the bond length.  kb = the
// partner's {kn, ks, density}.  l = |p_a - p_b| in f32, lenf(a.x - pb.x, a.y - pb.y):
// the f32 distance (:273-284) and the normal share it, and a caller that has formed the
// same squares for the overlap test passes its root in (x - y and y - x square to the
// same bits, so the value is identical).  With a.cn2 != 0 the normal force carries a
// dashpot (an extension; the reference's law has none): + 2 zeta sqrt(m_eff kn) times
// the approach speed along the normal, the relative displacement of the last drift over dT.
template <bool F64, class DD>
__device__ __forceinline__ float3 linear_model(float mu, const Self& a, const float4& pb, const float4& db,
                                               const float3& kb, float L, float l, float& ft, DD& dd) {
    float nd = fminf(0.0f, dist2<F64>(a, pb.x, pb.y, L, l));
    if (nd == 0.0f) return make_float3(0.0f, 0.0f, 0.0f);
    // one-material scenes take the series stiffnesses formed once for the scene (a
    // branch uniform over the wave: the divisions are not even issued)
    float kn = a.kne, ks = a.kse;
    if (!a.one) {
        kn = 1.0f / (1.0f / a.kn + 1.0f / kb.x);
        ks = 1.0f / (1.0f / a.ks + 1.0f / kb.y);
    }
    float nf = -nd * kn;
    float dx = a.x - pb.x, dy = a.y - pb.y;
    float nx = dx / l, ny = dy / l;
    if (a.cn2 != 0.0f) {
        float ma = a.rho * PI_F * a.r * a.r;
        float mb = kb.z * PI_F * pb.z * pb.z;
        float me = ma * mb / (ma + mb);
        float cn = a.cn2 * sqrtf(me * kn);
        float ax = db.x - a.dx, ay = db.y - a.dy;
        float va = (ax * nx + ay * ny) / a.dt;
        nf = nf + cn * va;
    }
    float tx = -ny, ty = nx;
    float rt = rel_tangent(a, pb, db, tx, ty);
    float lim = fabsf(nf) * mu;
    // clamp(ft + rt ks, -lim, lim) by comparisons: min/max leave the sign of a zero
    // result open (mu = 0), comparisons keep it, on the CPU and the GPU alike
    const float t = ft + rt * ks;
    ft = t < -lim ? -lim : (t > lim ? lim : t);
    float m = -a.r * ft;
    dd.add(nf, ft, m);
    return make_float3(nx * nf + tx * ft, ny * nf + ty * ft, m);
}

struct BondCtx {
    float L, sn, ss;
};

// Intact-bond laws: linear_parallel_bonds (mode 3, :286-336), linear_contact_bonds
// (mode 2, :375-417), normal_bonds (mode 1, :419-432).  bf = {ft, fbt, theta_b}.
// Sets `code` to the break code when the bond tears.
template <bool F64, class DD>
__device__ __forceinline__ float3 bond_law(const dem_uniform& u, const Self& a, const float4& pb, const float4& db,
                                           const float3& kb, const BondCtx& bc, float l, float4& bf, DD& dd,
                                           int& code) {
    code = 0;
    if (u.bonds == 3) {
        float kn = u.bond_normal_stiffness;
        float ks = u.bond_shear_stiffness;
        float R = 0.5f * fminf(a.r, pb.z);
        float A = 2.0f * R * 1.0f;
        float I = 2.0f / 3.0f * R * R * R;
        float nd = dist2<F64>(a, pb.x, pb.y, bc.L, l);
        float nf = -kn * nd * A;
        float dx = a.x - pb.x, dy = a.y - pb.y;
        float nx = dx / l, ny = dy / l;
        float tx = -ny, ty = nx;
        float rt = rel_tangent(a, pb, db, tx, ty);
        bf.y = bf.y + rt * ks * A;
        float dtb = a.drot - db.z;
        bf.z = bf.z + dtb;
        float m = -kn * I * bf.z - bf.y * a.r;
        float fx = nx * nf + tx * bf.y, fy = ny * nf + ty * bf.y;
        if (u.bonds_tear == 1 &&
            (nf / A - u.moment_contribution_factor * (fabsf(-kn * I * bf.z) * R) / I < -bc.sn ||
             fabsf(bf.y) / A > bc.ss)) {
            code = -10 + sign_i(nd);
            nf = 0.0f;
            bf.y = 0.0f;
            fx = 0.0f;
            fy = 0.0f;
            m = 0.0f;
        }
        dd.add(-nf, bf.y, m);
        float3 lm = linear_model<F64>(u.friction_coefficient, a, pb, db, kb, bc.L, l, bf.x, dd);
        return make_float3(fx + lm.x, fy + lm.y, m + lm.z);
    } else if (u.bonds == 2) {
        float nd = dist1<F64>(a, pb, l);
        float kn = 1.0f / (1.0f / u.bond_normal_stiffness + 1.0f / u.bond_normal_stiffness);
        float ks = 1.0f / (1.0f / u.bond_shear_stiffness + 1.0f / u.bond_shear_stiffness);
        float nf = -nd * kn;
        float dx = a.x - pb.x, dy = a.y - pb.y;
        float nx = dx / l, ny = dy / l;
        float tx = -ny, ty = nx;
        float rt = rel_tangent(a, pb, db, tx, ty);
        bf.x = bf.x + rt * ks;
        float cd = u.contact_damping;
        float fx = cd * (nx * nf + tx * bf.x), fy = cd * (ny * nf + ty * bf.x);
        float m = -a.r * bf.x;
        if (u.bonds_tear == 1 && (nf < -bc.sn || fabsf(bf.x) > bc.ss)) {
            code = -10 + sign_i(nd);
            nf = 0.0f;
            bf.y = 0.0f;
            fx = 0.0f;
            fy = 0.0f;
            m = 0.0f;
        }
        dd.assign(-nf, bf.x, m);
        return make_float3(fx, fy, m);
    } else if (u.bonds == 1) {
        float disp = dist1<F64>(a, pb, l);
        float sf = u.bond_normal_stiffness * disp;
        float dx = pb.x - a.x, dy = pb.y - a.y;
        float fx = 0.0f + sf * (dx / l), fy = 0.0f + sf * (dy / l);
        if (u.bonds_tear == 1 && sf >= bc.sn) {
            code = -10 + sign_i(disp);
            fx = 0.0f;
            fy = 0.0f;
        }
        float3 lm = linear_model<F64>(u.friction_coefficient, a, pb, db, kb, bc.L, l, bf.x, dd);
        return make_float3(lm.x + fx, lm.y + fy, lm.z + 0.0f);
    }
    return make_float3(0.0f, 0.0f, 0.0f);
}

// ---- the fused step kernel ----

// Block -> tile of slots, XCD-aware: the hardware hands consecutive blocks to the 8
// XCDs in turn, each with its own L2; remapping so that XCD k processes one contiguous
// eighth of the slots keeps a block's neighbour rows in the L2 that already holds them.
constexpr int N_XCD = 8;
__device__ __forceinline__ int xcd_block(int bid, int nb) {
    const int q = nb / N_XCD, r = nb % N_XCD;
    const int x = bid % N_XCD, i = bid / N_XCD;
    return x * q + min(x, r) + i;
}


// Cell scan tiles: the exclusive prefix of the cell counts is taken over tiles of
// SCAN_TILE cells, whose totals the binning accumulates in TSUM.
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

__host__ __device__ inline int tile_count(int n) { return (n + SCAN_TILE - 1) / SCAN_TILE; }

// Counting-sort binning: KEY[s] = key and RANK[s] = a unique rank in [0, count)
// among the slots of that cell.  Lanes of the wave that call together and hold
// the same key as their lower neighbour form a run; the run's lowest lane does one
// atomicAdd for the whole run and each lane takes base + its offset in the run.
// In cell-sorted storage neighbours share cells, and same-line atomics from one
// wave are serialised in L2, so this halves the atomics of the binning there.
// Only lanes active at the call take part (the exec mask bounds every run).
__device__ __forceinline__ void bin_slot(const DevArrays& a, int s, int key) {
    const uint64_t live = __ballot(1);
    const int lane = (int)(threadIdx.x & 63);
    const int below = __shfl_up(key, 1, 64);
    const bool head = lane == 0 || !((live >> (lane - 1)) & 1ull) || below != key;
    const uint64_t heads = __ballot(head);
    const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);      // lanes 0..lane
    const int h = 63 - __clzll((long long)(heads & upto));                  // this lane's run head
    const uint64_t above = heads & ~upto;                                   // heads after this lane
    const int end = above ? __ffsll((long long)above) - 1 : 64;             // first lane after the run
    const uint64_t run = (end == 64 ? ~0ull : ((1ull << end) - 1ull)) & ~((1ull << h) - 1ull) & live;
    int base = 0;
    if (head) base = atomicAdd(&a.CNT[key], __popcll(run));
    base = __shfl(base, h, 64);
    a.KEY[s] = key;
    a.RANK[s] = base + __popcll(run & ((1ull << lane) - 1ull));  // run lanes below this one
    // the same aggregation over scan tiles: one atomic per run of lanes in one tile
    const int tile = key / SCAN_TILE;
    const int tbelow = __shfl_up(tile, 1, 64);
    const bool thead = lane == 0 || !((live >> (lane - 1)) & 1ull) || tbelow != tile;
    const uint64_t theads = __ballot(thead);
    const uint64_t tabove = theads & ~upto;
    const int tend = tabove ? __ffsll((long long)tabove) - 1 : 64;
    const int th = 63 - __clzll((long long)(theads & upto));
    const uint64_t trun = (tend == 64 ? ~0ull : ((1ull << tend) - 1ull)) & ~((1ull << th) - 1ull) & live;
    if (thead) atomicAdd(&a.TSUM[tile], __popcll(trun));
}

// Clears the tile totals for the next binning (after the scan has read them).
__device__ __forceinline__ void clear_tile_sums(const DevArrays& a) {
    const int stride = (int)(gridDim.x * blockDim.x);
    for (int i = (int)(blockIdx.x * blockDim.x + threadIdx.x); i < a.ntiles; i += stride) a.TSUM[i] = 0;
}

// Carries the contact list of slot s (c entries) over to the next-list buffers; a list
// held beside the neighbour list (PF_ALIGNED in flag word mw) stays where it is, its
// mask travelling in the velocity word.
__device__ __forceinline__ void copy_list(const DevArrays& a, int s, uint32_t mw) {
    if (mw & PF_ALIGNED) return;
    const int c = cc_of(mw);
    const int cap = (int)a.cap;
    for (int q = 0; q < c && q < KMAX; q++) {
        a.CPn[q * cap + s] = a.CP[q * cap + s];
        a.CFn[q * cap + s] = a.CF[q * cap + s];
    }
}

__device__ __forceinline__ uint32_t contact_mask(const float4& v) {
    return (__float_as_uint(v.w) & VW_MASK) >> VW_MASK_SHIFT;
}

// The contact list of slot s held beside its neighbour list (contact mask m over the
// candidates, tangential forces in FT) written as a partner-id list into CPd / CFd at
// column j: the candidates are sorted by partner id, so the entries come out ascending,
// popc(m) of them (the count in M).
__device__ __forceinline__ void contacts_from_aligned(const DevArrays& a, int s, uint32_t m, Pool<int32_t> CPd,
                                                      Pool<float> CFd, int j) {
    const int cap = (int)a.cap;
    const int w0 = a.NL[s];
    int n = 0;
    for (uint32_t mk = m; mk != 0u; mk &= mk - 1u) {
        const int k = __ffs(mk) - 1;
        const int js = k == 0 ? nl_first(w0) : a.NL[k * cap + s];
        CPd[n * cap + j] = a.ID[js];
        CFd[n * cap + j] = a.FT[k * cap + s];
        n++;
    }
}

__device__ __forceinline__ int wave_min_i(int x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x = min(x, __shfl_xor(x, off, 64));
    return x;
}
__device__ __forceinline__ int wave_max_i(int x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x = max(x, __shfl_xor(x, off, 64));
    return x;
}
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// narrow-phase predicate of 2D_Simulation.wgsl:159, |p_b - p_a| < r_b + r_a, with an
// exact early reject: d^2 >= (r_b + r_a)^2 * 1.0001 implies fl(sqrt(d^2)) > r_b + r_a
__device__ __forceinline__ bool overlaps(const float4& pj, const float4& p) {
    const float dx = pj.x - p.x, dy = pj.y - p.y;
    const float xx = dx * dx;
    const float yy = dy * dy;
    const float d2 = xx + yy;
    const float R = pj.z + p.z;
    if (d2 >= R * R * 1.0001f) return false;
    return sqrtf(d2) < R;
}
// the same predicate, also giving the root it took: lenf(p.x - pj.x, p.y - pj.y)
__device__ __forceinline__ bool overlaps_l(const float4& pj, const float4& p, float& l) {
    const float dx = pj.x - p.x, dy = pj.y - p.y;
    const float xx = dx * dx;
    const float yy = dy * dy;
    const float d2 = xx + yy;
    const float R = pj.z + p.z;
    if (d2 >= R * R * 1.0001f) return false;
    l = sqrtf(d2);
    return l < R;
}

// a live bond record of this launch's bond pass with partner oj among [b_lo, b_hi)
template <bool BONDS>
__device__ __forceinline__ bool bonded_partner(const DevArrays& a, int b_lo, int b_hi, int oj) {
    if (!BONDS) return false;
    bool hit = false;
    for (int q = b_lo; q < b_hi; q++)
        if (__float_as_int(a.RF[q].w) == oj) hit = true;
    return hit;
}

// The loaded words of one bond endpoint record: the record word {bond id, bonded} (the
// bond pass's 8-byte copy; BR keeps the partner id for the step kernel and the lists)
// and, for a live record, its {ft, fbt, theta_b}, the partner's position and velocity
// records and the bond's {rest length, strengths, break word} (bond id clamped into the
// scene's bonds; used only where the id is valid).
struct RecIn {
    int2 rb;
    float4 bf, pb, vb, bl;
};
__device__ __forceinline__ void load_record(const DevArrays& a, const SceneDev& S, int cur, int q, int j, RecIn& in) {
    in.rb = a.BRB[q];
    if (in.rb.y != REC_DEAD && in.rb.y != REC_GONE) {
        in.bf = a.BF[q];
        in.pb = a.P[cur][j];
        in.vb = a.V[j];
        in.bl = a.BLS[S.b0 + min(max(in.rb.x, 0), max(S.nb - 1, 0))];
    }
}

// A dead record costs the bond pass its record word only: the launch after its death
// writes its result as dead once (the death launch left REC_RF_DIED there, which its
// owner has read) and marks it gone; a gone record is skipped.  Returns whether the
// record at q (record word rb) is dead.
__device__ __forceinline__ bool settle_dead_record(const DevArrays& a, int q, int2 rb) {
    if (rb.y == REC_GONE) return true;
    if (rb.y != REC_DEAD) return false;
    a.RF[q] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(REC_RF_DEAD));
    a.BRB[q].y = REC_GONE;
    return true;
}

// One bond endpoint record of the particle `me` (position p, scene S; partner in slot j) in a launch
// whose force half runs: the narrow-phase check of the bonded partner, the break
// state seen from the partner, the bond or contact law; writes the record (and the
// bond's break state) back.  Returns {fx, fy, moment, partner id} (the id's bits in the
// last word: the owner merges its records with its contacts in partner-id order from
// these results alone), {0, 0, 0, -1} when the record is dead, or {0, 0, 0, -2} when it
// dies here (2D_Simulation.wgsl:185-204, 230-262).
template <bool F64, class DD>
__device__ __forceinline__ float4 eval_bond_record(const DevArrays& a, const SceneDev& S, const Self& me,
                                                   const float4& p, bool narrow, int step, int q, int j,
                                                   const RecIn& in, DD& dd) {
    const dem_uniform& u = S.u;
    const int2 rb = in.rb;
    if (rb.y == REC_DEAD || rb.y == REC_GONE) return make_float4(0.0f, 0.0f, 0.0f, __int_as_float(REC_RF_DEAD));
    float4 bf = in.bf;
    const float4 pb = in.pb;
    const int pid = __float_as_int(pb.w);   // the partner's id travels with its position
    const float4 db = del_from(a, j, in.vb, u.dT);
    int bonded = rb.y;
    // |p - pb| in f32: the narrow-phase test and every law below use this one value
    const float l = lenf(pb.x - p.x, pb.y - p.y);
    if (narrow) {
        const bool found = l < (pb.z + p.z);
        if (!found && (bonded < 0 || u.bonds == 0)) {
            a.BR[q].w = 0;
            a.BRB[q].y = REC_DEAD;
            // the partner may come back into contact: the owner's step kernel of this
            // launch puts it into the owner's neighbour list (which leaves out live bond
            // partners), told by the alive word -1
            return make_float4(0.0f, 0.0f, 0.0f, __int_as_float(REC_RF_DIED));
        } else if (!found && bonded > 0 && u.bonds <= 1) {
            bf.x = 0.0f;
        }
    }
    const int bid = rb.x;
    const bool bid_ok = bid >= 0 && bid < S.nb;
    // {rest length, strengths, break word}: one load gives the break state and the law's
    // constants
    float4 bl = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (bonded >= 0 && bid_ok) {
        bl = in.bl;
        // a break recorded by the partner in an earlier step is seen here
        const uint32_t bw = __float_as_uint(bl.w);
        if (brk_before(bw, (uint32_t)step)) bonded = (bw & 3u) ? (int)(bw & 3u) - 12 : a.BS[S.b0 + bid].x;
    }
    const float3 kb = partner_k(a, S, j);
    float3 f;
    if (bonded < 0 || u.bonds == 0) {
        f = linear_model<F64>(u.friction_coefficient, me, pb, db, kb, me.r + pb.z, l, bf.x, dd);
    } else if (u.bonds >= 1 && u.bonds <= 3) {
        BondCtx bc;
        bc.L = bl.x;
        if (S.has_strength && bid_ok) {
            bc.sn = bl.y;
            bc.ss = bl.z;
        } else {
            bc.sn = u.bond_normal_strength;
            bc.ss = u.bond_shear_strength;
        }
        int code = 0;
        f = bond_law<F64>(u, me, pb, db, kb, bc, l, bf, dd, code);
        if (code < 0) {
            bonded = code;
            if (bid_ok) {
                a.BS[S.b0 + bid] = make_int2(code, step);
                a.BLS[S.b0 + bid].w = __uint_as_float(brk_word(code, (uint32_t)step));
            }
        }
    } else {
        f = make_float3(0.0f, 0.0f, 0.0f);
    }
    if (bonded != rb.y) {   // the record word changes only when the bond breaks
        a.BR[q].z = bonded;
        a.BRB[q].y = bonded;
    }
    a.BF[q] = bf;
    return make_float4(f.x, f.y, f.z, __int_as_float(pid));
}

// Candidates of the list path per round of loads (their slots, position and velocity
// records in flight together).
constexpr int STEP_ROUND = 4;
// the rows loaded ahead (1-3 with the own state, 4-7 with the first round) follow rounds of four
static_assert(STEP_ROUND == 4, "rounds of four candidates");

template <bool F64, bool BONDS, bool LARGE>
__device__ __forceinline__ void step_slot(const StepArgs& args) {
    const DevArrays& a = args.a;
    const int blk = xcd_block((int)blockIdx.x, (int)gridDim.x);
    const int s = blk * args.lbs + (int)threadIdx.x;
    // one live scene: its slot range below decides, with no table load on the way.  The
    // scene is the block's (uniform over the wave), and its record is read into registers
    // here, before any store: read through the table later, every field would be loaded
    // again after each store to the lists (the compiler cannot rule out aliasing).
    const int sid = __builtin_amdgcn_readfirstlane(args.one_scene >= 0 ? args.one_scene : args.blk_scene[blk]);
    if (sid < 0) return;
    const SceneDev S = args.scenes[sid];
    if (s < S.p0 || s >= S.p0 + S.np) return;   // whole blocks (np is a multiple of the block)
    const int2 lfw = args.lf[sid];
    const uint32_t fl = (uint32_t)lfw.x;
    const int cur = args.cur, nxt = cur ^ 1;
    if (!(fl & SF_ACTIVE)) {
        copy_list(a, s, (uint32_t)a.M[s]);  // the lists and velocities flip after every launch
        carry_velocity(a, s);
        if (args.flip) a.P[nxt][s] = a.P[cur][s];
        return;
    }
    const dem_uniform& u = S.u;
    const float dt = u.dT;
    const uint32_t pf = (uint32_t)a.M[s];
    const bool act = !(pf & PF_INACTIVE);
    float4 p = a.P[cur][s];
    const float th = a.TH[s];   // rotation, advanced by the drift
    const int nlw = a.NL[s];     // issued with the other own-state loads, used if the narrow phase runs
    const int cap = (int)a.cap;  // KMAX * cap < 2^31 (checked at pool creation)
    // Issued with the own state as well, before the counts that say how many of them are
    // valid are known (rows past a count hold stale words of this slot's column, in
    // bounds): neighbour-list rows 1-3 and the tangential forces of candidates 0-3.
    int nlr[3];
    float ftr[4];
#pragma unroll
    for (int t = 0; t < 3; t++) nlr[t] = a.NL[(t + 1) * cap + s];
#pragma unroll
    for (int t = 0; t < 4; t++) ftr[t] = a.FT[t * cap + s];
    int b_lo = 0, b_hi = 0;
    // bond records: at most four are held in registers (partner ids, alive bits) for
    // the bonded-partner test of the contact search and the merge with the contacts
    // (loaded on first use: list candidates are never live bond partners, since the
    // lists leave those out and take a partner in only when its record dies, so only
    // the stencil fallback and the merge with contacts need the records)
    int rx[4] = {INT32_MAX, INT32_MAX, INT32_MAX, INT32_MAX};
    int ralive = 0;
    bool rloaded = false;
    if (BONDS && act) {
        const int2 br = a.BRS[s];
        b_lo = br.x;
        b_hi = br.y;
    }
    // the partner ids come from this launch's bond-pass results (their last word: the
    // partner id of a live record, negative otherwise); a dead record's -1 sorts first,
    // where it adds nothing
    auto load_records = [&]() {
        if (!BONDS || rloaded) return;
        rloaded = true;
        if (b_hi <= b_lo) return;
        // loads without branches (indices past the range repeat its last record)
        int w[4];
#pragma unroll
        for (int k = 0; k < 4; k++) w[k] = __float_as_int(a.RF[min(b_lo + k, b_hi - 1)].w);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (b_lo + k < b_hi) {
                rx[k] = w[k];
                if (rx[k] >= 0) ralive |= 1 << k;
            }
        }
    };
    const bool rec_regs = !BONDS || b_hi - b_lo <= 4;
    auto is_bonded = [&](int oj) {
        if (!BONDS) return false;
        if (!rec_regs) return bonded_partner<BONDS>(a, b_lo, b_hi, oj);
        load_records();
        bool hit = false;
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (((ralive >> k) & 1) && rx[k] == oj) hit = true;
        return hit;
    };

    // own-state loads issued before the contact search so that their latency
    // overlaps it (the compiler keeps them behind the list stores otherwise)
    float4 v = a.V[s];
    const int cc_old = cc_of(pf);
    const int mat = (int)(pf & PF_MAT_MASK);
    // a scene of one material (the common case) takes the record from the scene table,
    // without waiting for the particle's material word
    float ma_rho = S.rho1, ma_kn = S.kn1, ma_ks = S.ks1;
    if (S.nm > 1) {
        const dem_material MA = a.MAT[S.m0 + (mat < S.nm ? mat : 0)];
        ma_rho = MA.density;
        ma_kn = MA.normal_stiffness;
        ma_ks = MA.shear_stiffness;
    }

    if (!act) {
        // never binned into a real cell, never a partner; state kept in both buffers
        copy_list(a, s, pf);
        if ((fl & SF_SIM) && S.record) {
            const float4 d0 = a.DATA0[s];
            a.DATA0[s] = make_float4(0.0f, 0.0f, 0.0f, d0.w);
        }
        carry_velocity(a, s);
        if (args.flip) a.P[nxt][s] = p;
        return;
    }

    int n = 0;   // entries of the new contact list
    const int oc = min(cc_old, KMAX);
    const bool narrow = (fl & SF_SIM) && (fl & SF_NARROW);
    const int w0 = narrow ? nlw : 0;
    const int nc = nl_count(w0);
    // The list path: every candidate of a list sorted by partner id is tested, forced
    // and appended in that order, so the contacts come out ascending with no sorting;
    // taken while the list cannot overflow the KMAX slots that contacts share with bond
    // records (2D_Simulation.wgsl:125).  Otherwise (stencil fallback, a list that could
    // overflow) the general path collects the partners first and applies the
    // reference's keep order before forcing them.
    const bool stream = narrow && nc > 0 && nc + (b_hi - b_lo) <= KMAX;
    // The list path keeps its contacts beside the neighbour list (PF_ALIGNED): bit k of a
    // mask in the velocity word for candidate k, its tangential force in FT row k, read
    // and rewritten in place, so that no partner ids are stored and no merge by id runs.
    // The other paths take and give partner-id lists (CP -> CPn), converting a list
    // that arrives in the other form first.
    const bool aligned = (pf & PF_ALIGNED) != 0;
    bool out_aligned = aligned;     // the form this launch leaves
    uint32_t vbits = vword(v) & VW_MASK;   // and the mask it leaves in the velocity word

    float ax, ay, al;

    if (fl & SF_SIM) {
        // this particle's displacement in the last drift (before the kick below)
        const float4 dme = del_from(a, s, v, dt);
        Self me;
        me.x = p.x; me.y = p.y; me.r = p.z; me.drot = dme.z; me.dx = dme.x; me.dy = dme.y;
        me.kn = ma_kn; me.ks = ma_ks;
        me.rho = ma_rho; me.cn2 = S.cn2; me.dt = dt;
        me.one = S.nm <= 1; me.kne = S.kne1; me.kse = S.kse1;
        Data3 dd{0.0f, 0.0f, 0.0f};
        float nfx = 0.0f, nfy = 0.0f, nm = 0.0f;

        // bond records: the bond pass of this launch evaluated each one (force, moment,
        // record and break state); the owner adds them in partner order, merged with its
        // contacts
        int bq = b_lo;
        // records that died in this launch's bond pass (bit q - b_lo; `dead_far`: one
        // past the 32nd): their partners join this particle's neighbour list below
        uint32_t dead = 0;
        bool dead_far = false;
        auto died = [&](int q) {
            const int k = q - b_lo;
            if (k < 32) dead |= 1u << k;
            else dead_far = true;
        };
        auto eval_record = [&](int q) {
            const float4 rf = a.RF[q];
            const int rw = __float_as_int(rf.w);
            if (rw < 0) {
                if (rw == REC_RF_DIED) died(q);
                return;
            }
            nfx += rf.x;
            nfy += rf.y;
            nm += rf.z;
            if (S.record) {
                replay(dd, a.RD[2 * q]);
                replay(dd, a.RD[2 * q + 1]);
            }
        };
        // partner id of record q (selects, not an indexed register array, which would
        // live in scratch memory)
        auto rec_id = [&](int q) {
            if (!rec_regs) return __float_as_int(a.RF[q].w);
            load_records();
            const int k = q - b_lo;
            return k == 0 ? rx[0] : (k == 1 ? rx[1] : (k == 2 ? rx[2] : rx[3]));
        };

        if (stream) {
            if (BONDS) load_records();
            // A partner-id list (the first launch after a re-sort, a read from the host,
            // another path, or a list build that could not keep the contacts beside the
            // new list) is read through a window of four entries in registers: ow[0] is
            // the entry under the merge pointer, the window shifts as it advances and
            // loads the entry four ahead.  A list beside the neighbour list needs no search.
            int ow[4], pn = 4;   // pn: old-list index of the next entry to load into the window
            float fw[4];
            const uint32_t omask = aligned ? contact_mask(v) : 0u;
#pragma unroll
            for (int t = 0; t < 4; t++) {
                ow[t] = INT32_MAX;
                fw[t] = 0.0f;
                if (!aligned && t < oc) {
                    ow[t] = a.CP[t * cap + s];
                    fw[t] = a.CF[t * cap + s];
                }
            }
            uint32_t nmask = 0u;
            // List rows 4-7 and their forces, loaded with the first round's records where
            // the list has more than four candidates, so that the second round waits on its
            // records only (in the plain f32 kernel: the bonded, large-body and f64 variants
            // have no registers to spare).
            constexpr bool ROWS_AHEAD = !BONDS && !F64 && !LARGE;
            int nl4[4] = {0, 0, 0, 0};
            float ft4[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            if (ROWS_AHEAD && nc > 4) {
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    nl4[t] = a.NL[(4 + t) * cap + s];
                    ft4[t] = a.FT[min(4 + t, KMAX - 1) * cap + s];
                }
            }
            // four candidates at a time: slots, then their positions, velocity records,
            // ids and tangential forces in one round of loads (before the overlap tests, so
            // that a hit waits on nothing)
            for (int k0 = 0; k0 < nc; k0 += STEP_ROUND) {
                // Loads without branches, so that all of them are in flight together: list
                // rows past the count (rows < NLMAX exist for every slot) and, for a missing
                // candidate, this slot's own records stand in and the test below drops them.
                // Rows 0-3 came with the own state.
                int jj[STEP_ROUND];
                float fq[STEP_ROUND];
#pragma unroll
                for (int t = 0; t < STEP_ROUND; t++) {
                    const int k = k0 + t;
                    int w;
                    if (k == 0) w = nl_first(w0);
                    else if (k == 1) w = nlr[0];
                    else if (k == 2) w = nlr[1];
                    else if (k == 3) w = nlr[2];
                    else if (ROWS_AHEAD && k0 == 4) w = nl4[t];
                    else w = a.NL[min(k, NLMAX - 1) * cap + s];
                    jj[t] = k < nc ? w : -1;
                    fq[t] = k0 == 0 ? ftr[t] : ((ROWS_AHEAD && k0 == 4) ? ft4[t] : a.FT[min(k, KMAX - 1) * cap + s]);
                }
                float4 qq[STEP_ROUND], vq[STEP_ROUND];
                int iq[STEP_ROUND];   // partner ids: the last word of the position record
#pragma unroll
                for (int t = 0; t < STEP_ROUND; t++) {
                    const int jl = jj[t] >= 0 ? jj[t] : s;
                    qq[t] = a.P[cur][jl];
                    vq[t] = a.V[jl];
                    iq[t] = __float_as_int(qq[t].w);
                }
                // keeps the compiler from sinking any of these loads into the branches
                // below, which would wait on each candidate in turn
                asm volatile("" ::: "memory");
                // this round's tangential forces, stored after its bodies: a store issued
                // inside them would hold up every later wait on a load (the vector memory
                // counter counts loads and stores in one queue)
                float hft[STEP_ROUND];
                uint32_t hits = 0u;
#pragma unroll
                for (int t = 0; t < STEP_ROUND; t++) {
                    const int j = jj[t];
                    float lj = 0.0f;
                    hft[t] = 0.0f;
                    if (j < 0 || j == s || !overlaps_l(qq[t], p, lj)) continue;
                    const int oj = iq[t];
                    if (BONDS) {
                        while (bq < b_hi && rec_id(bq) < oj) eval_record(bq++);
                    }
                    float ft;
                    if (aligned) {
                        ft = ((omask >> (k0 + t)) & 1u) ? fq[t] : 0.0f;
                    } else {
                        while (ow[0] < oj) {
                            ow[0] = ow[1]; fw[0] = fw[1];
                            ow[1] = ow[2]; fw[1] = fw[2];
                            ow[2] = ow[3]; fw[2] = fw[3];
                            ow[3] = pn < oc ? a.CP[pn * cap + s] : INT32_MAX;
                            fw[3] = pn < oc ? a.CF[pn * cap + s] : 0.0f;
                            pn++;
                        }
                        ft = ow[0] == oj ? fw[0] : 0.0f;
                    }
                    const float4 db = del_from(a, j, vq[t], dt);
                    const float3 kb = partner_k(a, S, j);
                    const float3 f = linear_model<F64>(u.friction_coefficient, me, qq[t], db, kb, me.r + qq[t].z, lj, ft, dd);
                    nfx += f.x;
                    nfy += f.y;
                    nm += f.z;
                    hft[t] = ft;
                    hits |= 1u << t;
                }
                // candidates below nc <= KMAX: rows of FT and bits of the mask exist for them
#pragma unroll
                for (int t = 0; t < STEP_ROUND; t++) {
                    if (!((hits >> t) & 1u)) continue;
                    a.FT[(k0 + t) * cap + s] = hft[t];
                    nmask |= 1u << (k0 + t);
                    n++;
                }
            }
            out_aligned = true;
            vbits = nmask << VW_MASK_SHIFT;
        } else {
            // the other paths read a partner-id list
            if (aligned) contacts_from_aligned(a, s, contact_mask(v), a.CP, a.CF, s);
            out_aligned = false;
            vbits = 0u;
            if (narrow && nc != 0) {
                // general path: every overlapping partner, from the list or (nc < 0) from the
                // stencil of the last cell list plus the scene's large bodies
                // capacity for contacts: the KMAX slots less the live bond records
                int alive = 0;
                if (BONDS) {
                    if (rec_regs) {
                        load_records();
                        alive = __popc(ralive);
                    } else {
                        for (int q = b_lo; q < b_hi; q++) alive += __float_as_int(a.RF[q].w) >= 0;
                    }
                }
                const int cap_c = max(KMAX - alive, 0);
                // Existing contacts found again are kept (bit per old-list position); new
                // partners go into an ascending list in the next-list buffer, the smallest
                // cap_c of them kept (the reference fills free slots and never evicts).
                uint32_t keep = 0;
                int nnew = 0, found_new = 0;
                auto test = [&](int j, const float4& q, bool listed) {
                    if (j == s || !overlaps(q, p)) return;
                    const int oj = a.ID[j];
                    if (!listed && is_bonded(oj)) return;
                    int at = -1;
                    for (int m = 0; m < oc; m++)
                        if (a.CP[m * cap + s] == oj) at = m;
                    if (at >= 0) {
                        keep |= 1u << at;
                        return;
                    }
                    found_new++;
                    int pos = nnew;
                    while (pos > 0 && a.CPn[(pos - 1) * cap + s] > oj) {
                        if (pos < cap_c) a.CPn[pos * cap + s] = a.CPn[(pos - 1) * cap + s];
                        pos--;
                    }
                    if (pos < cap_c) {
                        a.CPn[pos * cap + s] = oj;
                        if (nnew < cap_c) nnew++;
                    }
                };
                if (nc > 0) {
                    for (int k = 0; k < nc; k++) {
                        const int j = k == 0 ? nl_first(w0) : a.NL[k * cap + s];
                        test(j, a.P[cur][j], true);
                    }
                } else {
                    // a large body: the small discs that overlap it, found by this launch's
                    // large-body search (all 64 lanes of a wave on its stencil)
                    int lbh = -1, lbg = -1;
                    if (LARGE && (pf & PF_LARGE) && S.nlb > 0) {
                        const int me_id = a.ID[s];
                        for (int k = 0; k < S.nlb; k++)
                            if (a.LB[S.lb0 + k] == me_id) lbg = S.lb0 + k;
                        if (lbg >= 0) lbh = a.LBHC[lbg];
                    }
                    if (lbh >= 0) {
                        for (int k = 0; k < lbh; k++) {
                            const int j = a.LBH[lbg * LBH_MAX + k];
                            test(j, a.P[cur][j], false);
                        }
                    } else {
                    // more candidates than a list holds, or a large body whose search
                    // overflowed: the stencil around the particle's cell at the last build,
                    // over the last cell list (cells are wider than 2 r_small + skin, so the
                    // 3x3 stencil holds every small disc that can touch a small disc until
                    // the rebuild; a large body reaches r + r_small + skin)
                    const float2 xr0 = a.XR[s];
                    int cx, cy;
                    cell_of(S, xr0.x, xr0.y, cx, cy);
                    const int pos0 = S.p0 - a.START[S.c0];  // cell-list position -> slot-list index
                    const bool large = LARGE && S.nlb > 0 && p.z > S.rsmax;
                    const int reach = large ? max(1, (int)ceilf((p.z + S.rsmax + S.skin) * S.inv_cs) + 1) : 1;
                    const int xa = max(cx - reach, 0), xb = min(cx + reach, S.w - 1);
                    for (int ry = max(cy - reach, 0); ry <= min(cy + reach, S.h - 1); ry++) {
                        const int e0 = max(a.START[S.c0 + ry * S.w + xa] + pos0, S.p0);
                        const int e1 = min(a.START[S.c0 + ry * S.w + xb + 1] + pos0, S.p0 + S.np);
                        for (int e = e0; e < e1; e += 2) {
                            const int ja = a.SIDX[e];
                            const int jb = e + 1 < e1 ? a.SIDX[e + 1] : -1;
                            const float4 qa = a.P[cur][ja];
                            const float4 qb = jb >= 0 ? a.P[cur][jb] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                            // large bodies come from their list
                            if (!LARGE || S.nlb == 0 || qa.z <= S.rsmax) test(ja, qa, false);
                            if (jb >= 0 && (!LARGE || S.nlb == 0 || qb.z <= S.rsmax)) test(jb, qb, false);
                        }
                    }
                    }
                    // then the scene's large bodies, from their list
                    if (LARGE) {
                        for (int k = 0; k < S.nlb; k++) {
                            const int jl = a.INV[S.p0 + a.LB[S.lb0 + k]];
                            test(jl, a.P[cur][jl], false);   // test skips s itself
                        }
                    }
                }
                // the kept old contacts come first; the new ones fill what is left
                const int nk = __popc(keep);
                const int room = max(cap_c - nk, 0);
                if (nnew > room) nnew = room;
                if (found_new > nnew) atomicAdd(&a.STAT[0], (unsigned long long)(found_new - nnew));
                // merge kept old entries (ascending, from CP) with the new ones (CPn[0, nnew))
                // into CPn[0, nk + nnew), from the back
                int i_new = nnew - 1, i_old = oc - 1;
                for (int w = nk + nnew - 1; w >= 0; w--) {
                    while (i_old >= 0 && !((keep >> i_old) & 1u)) i_old--;
                    const int vo = i_old >= 0 ? a.CP[i_old * cap + s] : INT32_MIN;
                    const int vn = i_new >= 0 ? a.CPn[i_new * cap + s] : INT32_MIN;
                    if (vo > vn) {
                        a.CPn[w * cap + s] = vo;
                        i_old--;
                    } else {
                        a.CPn[w * cap + s] = vn;
                        i_new--;
                    }
                }
                n = nk + nnew;
            } else if (!narrow) {
                // no narrow phase this step: the existing contacts carry over
                n = oc;
                for (int q = 0; q < n; q++) a.CPn[q * cap + s] = a.CP[q * cap + s];
            }
            // forces over the list in CPn, merged with the bond records in partner order;
            // partners are found by id
            int pp = 0;  // read position in the previous list (ascending ids)
            for (int q = 0; q < n; q++) {
                const int ob = a.CPn[q * cap + s];
                if (BONDS) {
                    while (bq < b_hi && rec_id(bq) < ob) eval_record(bq++);
                }
                float ft = 0.0f;
                if (narrow) {
                    while (pp < oc && a.CP[pp * cap + s] < ob) pp++;
                    if (pp < oc && a.CP[pp * cap + s] == ob) ft = a.CF[pp * cap + s];
                } else {
                    ft = a.CF[q * cap + s];
                }
                const int j = a.INV[S.p0 + ob];
                const float4 pb = a.P[cur][j];
                const float4 db = del_of(a, j, dt);
                const float3 kb = partner_k(a, S, j);
                const float f3l = lenf(me.x - pb.x, me.y - pb.y);
                const float3 f = linear_model<F64>(u.friction_coefficient, me, pb, db, kb, me.r + pb.z, f3l, ft, dd);
                nfx += f.x;
                nfy += f.y;
                nm += f.z;
                a.CFn[q * cap + s] = ft;
            }
        }
        if (BONDS) {
            // the records after the last contact, eight results in flight at a time (loads
            // without branches: indices past the range repeat its last record, which the
            // count then drops)
            while (bq < b_hi) {
                constexpr int RR = 8;
                const int m = min(b_hi - bq, RR);
                float4 rf[RR];
#pragma unroll
                for (int k = 0; k < RR; k++) rf[k] = a.RF[bq + min(k, m - 1)];
#pragma unroll
                for (int k = 0; k < RR; k++) {
                    if (k < m && __float_as_int(rf[k].w) >= 0) {
                        nfx += rf[k].x;
                        nfy += rf[k].y;
                        nm += rf[k].z;
                        if (S.record) {
                            replay(dd, a.RD[2 * (bq + k)]);
                            replay(dd, a.RD[2 * (bq + k) + 1]);
                        }
                    } else if (k < m && __float_as_int(rf[k].w) == REC_RF_DIED) {
                        died(bq + k);
                    }
                }
                bq += m;
            }
            // A record that died here (its bond broken and the discs apart) hands the pair
            // to the contact search: the partner goes into this particle's list, in id
            // order, so that the list again holds every disc that can touch it before the
            // next rebuild (adding a candidate keeps a neighbour list valid).  A list that
            // would overflow raises the rebuild flags instead (the rebuild after this
            // launch's drift, or after the next launch's if this one has none); a particle
            // on the stencil fallback (nc < 0) finds the partner in the cells.
            if (dead != 0u || dead_far) {
                if (dead_far || nl_count(nlw) + __popc(dead) > NLMAX) {
                    a.RB[0] = 1;
                    a.RB[1] = 1;
                } else if (nl_count(nlw) >= 0) {
                    // the insertion renumbers the candidates: contacts kept beside the list
                    // leave as a partner-id list
                    if (out_aligned) {
                        contacts_from_aligned(a, s, vbits >> VW_MASK_SHIFT, a.CPn, a.CFn, s);
                        out_aligned = false;
                        vbits = 0u;
                    }
                    int cnt = nl_count(nlw);
                    int first = nl_first(nlw);
                    for (uint32_t mk = dead; mk != 0u; mk &= mk - 1u) {
                        const int oj = a.BR[b_lo + __ffs(mk) - 1].x;
                        const int js = a.INV[S.p0 + oj];
                        int k = cnt - 1;
                        for (; k >= 0; k--) {
                            const int sk = k == 0 ? first : a.NL[k * cap + s];
                            if (a.ID[sk] < oj) break;
                            a.NL[(k + 1) * cap + s] = sk;
                        }
                        if (k < 0) first = js;
                        else a.NL[(k + 1) * cap + s] = js;
                        cnt++;
                    }
                    a.NL[s] = nl_word(cnt, first);
                }
            }
        }
        // single-buffered: written only when it changes
        const uint32_t mnew = (pf & ~(PF_CC_MASK | PF_ALIGNED)) | ((uint32_t)n << PF_CC_SHIFT) |
                              (out_aligned ? PF_ALIGNED : 0u);
        if (mnew != pf) a.M[s] = (int)mnew;

        // 3. store_forces, 2D_Simulation.wgsl:434-478
        const float density = ma_rho;
        const float rr = p.z;
        const float mass = density * PI_F * rr * rr;
        const float inertia = 0.5f * mass * rr * rr;
        float fx, fy, mm;
        if (pf & PF_EXT) {
            const float4 e = a.EXT[s];
            fx = nfx + e.x;
            fy = nfy + e.y;
            mm = nm + e.z;
        } else {
            fx = nfx + 0.0f;
            fy = nfy + 0.0f;
            mm = nm + 0.0f;
        }
        if (u.gravity == 1 && u.planet_mode == 1) {
            float gx = 0.0f, gy = 0.0f;
            if (u.mouse_gravity == 1) {
                gx = u.gravity_x;
                gy = u.gravity_y;
            }
            const float ddx = gx - p.x, ddy = gy - p.y;
            if (ddx != 0.0f || ddy != 0.0f) {
                const float l = lenf(ddx, ddy);
                fx = fx + ddx / l * G_F * u.gravity_acc * mass;
                fy = fy + ddy / l * G_F * u.gravity_acc * mass;
            }
        } else if (u.gravity == 1) {
            const float ga = G_F * u.gravity_acc;
            fx = fx + 0.0f;
            fy = fy + -ga * mass;
        }
        ax = fx / mass;
        ay = fy / mass;
        al = mm / inertia;
        if (pf & PF_KICK_X) v.x = v.x + 0.5f * ax * dt;
        if (pf & PF_KICK_Y) v.y = v.y + 0.5f * ay * dt;
        if (pf & PF_KICK_R) v.z = v.z + 0.5f * al * dt;
        if (S.record) {
            const float4 d0 = a.DATA0[s];
            a.DATA0[s] = make_float4(dd.d0, dd.d1, dd.d2, d0.w);
        }
        if (!(fl & SF_LOM)) {
            // no drift follows: the kicked velocity no longer gives the displacement,
            // which is kept explicitly
            a.Vn[s] = make_float4(v.x, v.y, v.z, __uint_as_float(VW_EXPLICIT | vbits | (vword(v) & VW_PATH)));
            a.DXn[s] = dme;
            a.A[s] = make_float4(ax, ay, al, 0.0f);
            // force pass only, while other scenes drift: positions carry over
            if (args.flip) a.P[nxt][s] = p;
            return;
        }
    } else {
        copy_list(a, s, pf);
        const float4 av = a.A[s];
        ax = av.x;
        ay = av.y;
        al = av.z;
    }

    // ---- LOM half, 2D_LOM.wgsl:85-152 ----
    float ivx = v.x, ivy = v.y, iw = v.z;
    if (pf & PF_KICK_X) ivx += ax * dt * 0.5f;
    if (pf & PF_KICK_Y) ivy += ay * dt * 0.5f;
    if (pf & PF_KICK_R) iw += al * dt * 0.5f;
    if (S.record) a.DATA1[s] = make_float4(ivx, ivy, iw, 0.0f);
    float dpx, dpy, dr;
    // a displacement that the next force half cannot re-form as velocity x dT
    bool expl = (fl & SF_DELOVR) != 0;
    if (expl) {
        const float4 e = a.DX[s];   // explicit: written by the upload or the last force half
        dpx = e.x;
        dpy = e.y;
        dr = e.z;
    } else {
        dpx = ivx * dt;
        dpy = ivy * dt;
        dr = iw * dt;
    }
    float px = p.x, py = p.y;
    const float rr = p.z;
    const float npx = px + dpx, npy = py + dpy;
    if (u.round_bounds == 0) {
        const float yH = S.half_h;
        const float xW = S.half_w;
        if (pf & PF_WALL_Y) {
            if (npy - rr < -yH) {
                ivy = -ivy * 0.5f;
                py += -yH - (npy - rr);
                expl = true;
            } else if (npy + rr > yH) {
                ivy = -ivy * 0.5f;
                py -= (npy + rr) - yH;
                expl = true;
            }
        }
        if (pf & PF_WALL_X) {
            if (npx - rr < -xW) {
                ivx = -ivx * 0.5f;
                px += -xW - (npx - rr);
                expl = true;
            } else if (npx + rr > xW) {
                ivx = -ivx * 0.5f;
                px -= (npx + rr) - xW;
                expl = true;
            }
        }
    } else if (lenf(npx, npy) + rr > u.bound_radius) {
        expl = true;
        const float l = lenf(npx, npy);
        const float nx = npx / l, ny = npy / l;
        const float ddd = dpx * nx + dpy * ny;
        const float dcx = ddd * nx, dcy = ddd * ny;
        const float vv = ivx * nx + ivy * ny;
        const float cvx = vv * nx, cvy = vv * ny;
        if (pf & PF_WALL_X) {
            px = nx * (u.bound_radius - rr) - dcx;
            ivx = -cvx * 0.5f;
        }
        if (pf & PF_WALL_Y) {
            py = ny * (u.bound_radius - rr) - dcy;
            ivy = -cvy * 0.5f;
        }
    }
    px += dpx;
    py += dpy;
    const float rot = th + dr;
    a.TH[s] = rot;
    a.P[nxt][s] = make_float4(px, py, rr, p.w);
    if (expl) a.DXn[s] = make_float4(dpx, dpy, dr, 0.0f);
    // Moved more than 0.45 skin since the neighbour lists were built: rebuild them after
    // this launch (every writer stores 1).  The path word bounds the distance from above
    // (|dx| + |dy| per drift, rounded up), so the build position is read only once the
    // path passes 0.45 skin; the word then restarts from the exact distance.
    uint32_t q = vword(v) & VW_PATH;
    const float inc = fminf((fabsf(px - p.x) + fabsf(py - p.y)) * S.pscale * 1.0001f, 1.0e9f);
    q = min(q + (uint32_t)ceilf(inc), VW_PATH_SAT);
    if (q > VW_PATH_TRIG) {
        const float2 xr = a.XR[s];
        const float mx = px - xr.x, my = py - xr.y;
        const float mxx = mx * mx;
        const float myy = my * my;
        if (mxx + myy > S.trig2) a.RB[nxt] = 1;
        else q = min((uint32_t)ceilf(fminf((fabsf(mx) + fabsf(my)) * S.pscale * 1.0001f, 1.0e9f)), VW_PATH_SAT);
    }
    a.Vn[s] = make_float4(ivx, ivy, iw, __uint_as_float((expl ? VW_EXPLICIT : 0u) | vbits | q));
}

// ---- bond pass: one thread per bond endpoint record ----
// Runs before the step kernel of every launch with a force half.  The owner of record
// q is the particle with global id RO[q] (-1: no record); lbs is the scene table's
// block of slots.  Results: RF[q] = {fx, fy, moment, alive}, RD[2q..2q+1] = the
// record's data[0..2] updates (with data recording).
template <bool F64>
__device__ __forceinline__ void bond_pass_body(const StepArgs& args, const SceneDev& S, int q, int i, int j, int2 lfw,
                                               const float4& p, const float4& vi, uint32_t mi, const RecIn& in) {
    const DevArrays& a = args.a;
    const float4 d = del_from(a, i, vi, S.u.dT);
    const int mat = (int)(mi & PF_MAT_MASK);
    float rho = S.rho1, kn = S.kn1, ks = S.ks1;
    if (S.nm > 1) {
        const dem_material MA = a.MAT[S.m0 + (mat < S.nm ? mat : 0)];
        rho = MA.density;
        kn = MA.normal_stiffness;
        ks = MA.shear_stiffness;
    }
    Self me;
    me.x = p.x; me.y = p.y; me.r = p.z; me.drot = d.z; me.dx = d.x; me.dy = d.y;
    me.kn = kn; me.ks = ks;
    me.rho = rho; me.cn2 = S.cn2; me.dt = S.u.dT;
    me.one = S.nm <= 1; me.kne = S.kne1; me.kse = S.kse1;
    const uint32_t fl = (uint32_t)lfw.x;
    const bool narrow = (fl & SF_SIM) && (fl & SF_NARROW);
    DataLog dl;
    dl.n = 0;
    dl.op[0] = dl.op[1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    a.RF[q] = eval_bond_record<F64>(a, S, me, p, narrow, lfw.y, q, j, in, dl);
    if (S.record) {
        a.RD[2 * q] = dl.op[0];
        a.RD[2 * q + 1] = dl.op[1];
    }
}

// One record q (owner slot i, partner slot j) of scene sid whose record is S.
template <bool F64>
__device__ __forceinline__ void bond_record(const StepArgs& args, const SceneDev& S, int sid, int q, int i, int j) {
    const DevArrays& a = args.a;
    const int2 lfw = args.lf[sid];
    const uint32_t fl = (uint32_t)lfw.x;
    if (!(fl & SF_ACTIVE) || !(fl & SF_SIM)) return;
    const uint32_t mi = (uint32_t)a.M[i];
    if (mi & PF_INACTIVE) return;   // an inactive owner evaluates nothing
    const float4 p = a.P[args.cur][i];
    const float4 vi = a.V[i];
    RecIn in;
    load_record(a, S, args.cur, q, j, in);
    if (F64 && S.use_f64) bond_pass_body<true>(args, S, q, i, j, lfw, p, vi, mi, in);
    else bond_pass_body<false>(args, S, q, i, j, lfw, p, vi, mi, in);
}

// A scene's records are contiguous, so a wave's records nearly always share one scene:
// its record is then read into registers once (a wave that spans scenes reads each
// lane's through the table).  5 waves per SIMD: both variants fit without spills.
template <bool F64>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5, 8))) k_bond_forces(StepArgs args, int nrec, int lbs) {
    const DevArrays& a = args.a;
    const int q = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (q >= nrec) return;
    const int i = a.RS[q];   // owner slot (-1: no record)
    const int2 rb = a.BRB[q];
    if (i < 0) return;
    if (settle_dead_record(a, q, rb)) return;
    const int j = a.RP[q];   // partner slot
    const int sid = args.blk_scene[i / lbs];
    if (sid < 0) return;
    const int s0 = __builtin_amdgcn_readfirstlane(sid);
    if (__ballot(sid != s0) == 0) {
        const SceneDev S = args.scenes[s0];
        bond_record<F64>(args, S, s0, q, i, j);
    } else {
        bond_record<F64>(args, args.scenes[sid], sid, q, i, j);
    }
}

// The bond pass of a launch of one f32 scene: the scene record and flags are read into
// registers once, and each record's loads come in two rounds (its slots and record
// word; then both endpoints' position and velocity records, the owner's flag word, the
// force record and the bond record), each issued together.
__global__ void __launch_bounds__(256) k_bond_forces1(StepArgs args, int nrec) {
    const DevArrays& a = args.a;
    const int sid = args.one_scene;
    const SceneDev S = args.scenes[sid];
    const int2 lfw = args.lf[sid];
    const uint32_t fl = (uint32_t)lfw.x;
    if (!(fl & SF_ACTIVE) || !(fl & SF_SIM)) return;
    const int q = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (q >= nrec) return;
    const int i = a.RS[q];   // owner slot (-1: no record)
    const int j = a.RP[q];   // partner slot
    const int2 rb = a.BRB[q];
    if (i < 0) return;
    if (settle_dead_record(a, q, rb)) return;
    const int cur = args.cur;
    const uint32_t mi = (uint32_t)a.M[i];
    const float4 p = a.P[cur][i];
    const float4 vi = a.V[i];
    RecIn in;
    in.rb = rb;
    in.bf = a.BF[q];
    in.pb = a.P[cur][j];
    in.vb = a.V[j];
    in.bl = a.BLS[S.b0 + min(max(rb.x, 0), max(S.nb - 1, 0))];
    asm volatile("" ::: "memory");   // the loads above go out together
    if (mi & PF_INACTIVE) return;   // an inactive owner evaluates nothing
    bond_pass_body<false>(args, S, q, i, j, lfw, p, vi, mi, in);
}

// Owner and partner slots of every record (after an upload and after each re-sort):
// RS[q] = slot of the owner RO[q] (-1: no record), RP[q] = slot of the record's partner.
__global__ void k_record_slots(StepArgs args, int nrec, int lbs) {
    const DevArrays& a = args.a;
    const int q = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (q >= nrec) return;
    const int g = a.RO[q];
    int i = -1, j = 0;
    if (g >= 0) {
        const int o = a.INV[g];
        const int sid = args.blk_scene[o / lbs];
        if (sid >= 0) {
            i = o;
            j = a.INV[args.scenes[sid].p0 + a.BR[q].x];
        }
    }
    a.RS[q] = i;
    a.RP[q] = j;
}

// Whether this block's scene measures distances in f64.  A launch that holds any
// f64 scene runs the F64 instantiation; its f32 scenes still take the f32 body, so a
// scene's arithmetic never depends on the other scenes of the launch.
__device__ __forceinline__ bool block_scene_f64(const StepArgs& args) {
    const int sid = args.blk_scene[xcd_block((int)blockIdx.x, (int)gridDim.x)];
    return sid >= 0 && args.scenes[sid].use_f64 != 0;
}

// Plain f32 scenes: 4 waves per SIMD, at which the list path (four candidates' position
// and velocity records in flight, a round's new list entries held until its end) fits
// its registers without spilling; launches with a large body take
// k_step_plain_large, whose stencil fallback also covers them.  Launches holding an
// f64 scene (k_step_plain_f64) and bonded scenes carry more live state and keep the
// default register budget; both handle large bodies.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) k_step_plain(StepArgs args) {
    step_slot<false, false, false>(args);
}
// 4 waves per SIMD: 97 VGPRs without spills (the budget of 5 waves spills the
// large body's hit-list walk).
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) k_step_plain_large(StepArgs args) {
    step_slot<false, false, true>(args);
}
__global__ void __launch_bounds__(256) k_step_plain_f64(StepArgs args) {
    if (block_scene_f64(args)) step_slot<true, false, true>(args);
    else step_slot<false, false, true>(args);
}
// Bonded launches: the f32 variant fits 106 VGPRs without spills at 4 waves per SIMD;
// the f64 one keeps the default budget.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) k_step_bonds_f32(StepArgs args) {
    step_slot<false, true, true>(args);
}
__global__ void __launch_bounds__(256) k_step_bonds_f64(StepArgs args) {
    if (block_scene_f64(args)) step_slot<true, true, true>(args);
    else step_slot<false, true, true>(args);
}

// Binning of the current positions only (parity views and the bond builder).
__global__ void k_bin(StepArgs args) {
    const DevArrays& a = args.a;
    const int s = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int sid = args.blk_scene[blockIdx.x];
    if (sid < 0) return;
    const SceneDev& S = args.scenes[sid];
    if (!(args.lf[sid].x & SF_ACTIVE)) return;
    if (s < S.p0 || s >= S.p0 + S.np) return;
    const int2 m = make_int2(a.M[s], a.ID[s]);
    int key;
    if (m.x & PF_INACTIVE) {
        key = S.c0 + S.w * S.h;
    } else {
        const float4 p = a.P[args.cur][s];
        int cx, cy;
        key = S.c0 + cell_of(S, p.x, p.y, cx, cy);
    }
    bin_slot(a, s, key);
}

// ---- cell scan: exclusive prefix of cnt[0..n) into start[0..n], zeroing cnt ----
// Reduce-then-scan: the binning has added every cell count to its tile's total;
// k_scan_tiles adds the totals of the tiles before its own (read from L2) and scans
// its tile.  No block waits on another.


__device__ __forceinline__ int wave_incl_scan(int x) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    return x;
}

__device__ __forceinline__ int wave_sum(int x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

__device__ __forceinline__ void load_tile(const int32_t* cnt, int64_t base, int n, int (&v)[SCAN_ITEMS]) {
    if (base + SCAN_ITEMS <= n) {
        const int4* src = reinterpret_cast<const int4*>(cnt + base);
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS / 4; k++) {
            const int4 q = src[k];
            v[4 * k] = q.x; v[4 * k + 1] = q.y; v[4 * k + 2] = q.z; v[4 * k + 3] = q.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; k++) v[k] = base + k < n ? cnt[base + k] : 0;
    }
}

// Exclusive scan of tile `tile` (of nt) into start, zeroing its counts; the last tile
// also writes start[n] = total.
__device__ __forceinline__ void scan_tile_body(int32_t* cnt, int32_t* start, int n, const int32_t* tsum, int tile,
                                               int nt) {
    __shared__ int s_wave[SCAN_THREADS / 64];
    __shared__ int s_pre[SCAN_THREADS / 64];
    // prefix of the tiles before this one
    int pre = 0;
    for (int q = threadIdx.x; q < tile; q += SCAN_THREADS) pre += tsum[q];
    pre = wave_sum(pre);
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) s_pre[wv] = pre;
    const int64_t base = (int64_t)tile * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    int v[SCAN_ITEMS];
    load_tile(cnt, base, n, v);
    int sum = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) sum += v[k];
    const int incl = wave_incl_scan(sum);
    if (lane == 63) s_wave[wv] = incl;
    __syncthreads();
    int off = 0, tile_pre = 0, block_total = 0;
#pragma unroll
    for (int q = 0; q < SCAN_THREADS / 64; q++) {
        if (q < wv) off += s_wave[q];
        tile_pre += s_pre[q];
        block_total += s_wave[q];
    }
    int run = tile_pre + off + incl - sum;
    if (base + SCAN_ITEMS <= n) {
        int4* dst = reinterpret_cast<int4*>(start + base);
        int4* zc = reinterpret_cast<int4*>(cnt + base);
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS / 4; k++) {
            int4 q;
            q.x = run; run += v[4 * k];
            q.y = run; run += v[4 * k + 1];
            q.z = run; run += v[4 * k + 2];
            q.w = run; run += v[4 * k + 3];
            dst[k] = q;
            zc[k] = make_int4(0, 0, 0, 0);
        }
    } else {
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; k++) {
            const int64_t i = base + k;
            if (i < n) {
                start[i] = run;
                cnt[i] = 0;
            }
            run += v[k];
        }
    }
    if (tile == nt - 1 && threadIdx.x == SCAN_THREADS - 1) start[n] = tile_pre + block_total;
    __syncthreads();   // s_wave / s_pre are reused by the block's next tile
}

__global__ void __launch_bounds__(SCAN_THREADS) k_scan_tiles(int32_t* cnt, int32_t* start, int n,
                                                             const int32_t* tsum, const int32_t* gate) {
    if (gate && gate[0] == 0) return;
    scan_tile_body(cnt, start, n, tsum, (int)blockIdx.x, (int)gridDim.x);
}

// ---- k_scatter: cell-ordered slot list (positions relative to the scene) ----
__global__ void k_scatter(StepArgs args) {
    const DevArrays& a = args.a;
    clear_tile_sums(a);
    const int s = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int sid = args.blk_scene[blockIdx.x];
    if (sid < 0) return;
    const SceneDev& S = args.scenes[sid];
    const uint32_t fl = (uint32_t)args.lf[sid].x;
    if (!(fl & SF_ACTIVE) || !(fl & SF_BIN)) return;
    if (s < S.p0 || s >= S.p0 + S.np) return;
    const int pos = a.START[a.KEY[s]] - a.START[S.c0] + S.p0 + a.RANK[s];
    if (pos >= S.p0 && pos < S.p0 + S.np) a.SIDX[pos] = s;
    else atomicAdd(&a.STAT[3], 1ull);
}

// ---- neighbour-list rebuild ----
// Every kernel returns at once unless RB[cur] is set.  The grid strides over logical
// blocks of blockDim.x slots (the scene table's blocks), so a launch with the flag
// clear costs a few workgroups.  The rebuild bins every live scene.

__global__ void k_rb_store(int32_t* rb, int v) {
    if (blockIdx.x == 0 && threadIdx.x == 0) rb[0] = v;
}

// The per-slot stages of a rebuild; lbs is the scene table's block of slots.
__device__ __forceinline__ void rb_bin_slot(const StepArgs& args, int s, int lbs) {
    const DevArrays& a = args.a;
    const int sid = args.blk_scene[s / lbs];
    if (sid < 0) return;   // whole waves: lbs is a multiple of 64
    const SceneDev& S = args.scenes[sid];
    const int2 m = make_int2(a.M[s], a.ID[s]);
    int key = S.c0 + S.w * S.h;
    if (!(m.x & PF_INACTIVE)) {
        const float4 p = a.P[args.cur][s];
        int cx, cy;
        key = S.c0 + cell_of(S, p.x, p.y, cx, cy);
    }
    bin_slot(a, s, key);
}

__device__ __forceinline__ void rb_scatter_slot(const StepArgs& args, int s, int lbs) {
    const DevArrays& a = args.a;
    const int sid = args.blk_scene[s / lbs];
    if (sid < 0) return;
    const SceneDev& S = args.scenes[sid];
    const int pos = a.START[a.KEY[s]] - a.START[S.c0] + S.p0 + a.RANK[s];
    if (pos >= S.p0 && pos < S.p0 + S.np) a.SIDX[pos] = s;
    else atomicAdd(&a.STAT[3], 1ull);
}

// Neighbour list of slot s (a slot of the scene whose record is S): the slots within
// r_a + r_b + skin (x 1.0001) in the 3x3 stencil of its cell, sorted by partner id; more
// than NLMAX marks the particle for the stencil fallback (count NL_OVERFLOW).  Records
// the build position.  Every load comes in branch-free rounds (indices clamped into the
// scene's own ranges, entries past a row's end dropped by the test), so that a round's
// loads are in flight together.
constexpr int NL_ROW_STEP = 2;   // cell-list entries per stencil row and round
__device__ __forceinline__ void build_nl_slot(const StepArgs& args, const SceneDev& S, int pos0, int s, int2* nl_lds) {
    const DevArrays& a = args.a;
    const int cap = (int)a.cap;
    const int cur = args.cur;
    const int mw = a.M[s];
    const float4 p = a.P[cur][s];
    const int2 brs = a.BRS[s];
    const float4 vs = a.V[s];
    int cx, cy;
    cell_of(S, p.x, p.y, cx, cy);
    const int xa = max(cx - 1, 0), xb = min(cx + 1, S.w - 1);
    // the three stencil rows' ranges of the cell list, rows off the grid empty
    int e[3], e1[3];
#pragma unroll
    for (int r = 0; r < 3; r++) {
        const int ry = cy - 1 + r;
        const int ryc = min(max(ry, 0), S.h - 1);
        e[r] = a.START[S.c0 + ryc * S.w + xa];
        e1[r] = a.START[S.c0 + ryc * S.w + xb + 1];
    }
    // contacts kept beside the old list (PF_ALIGNED): moved beside the new one below, or
    // left as a partner-id list where that cannot hold them
    const bool al = ((uint32_t)mw & PF_ALIGNED) != 0;
    const uint32_t om = al ? contact_mask(vs) : 0u;
    const int w0old = a.NL[s];
    auto leave_aligned = [&]() {
        contacts_from_aligned(a, s, om, a.CP, a.CF, s);
        a.M[s] = (int)((uint32_t)mw & ~PF_ALIGNED);
    };
    a.XR[s] = make_float2(p.x, p.y);
    const uint32_t vw0 = __float_as_uint(vs.w) & VW_EXPLICIT;   // path since the build: 0
    if (mw & (PF_INACTIVE | PF_LARGE)) {
        if (al) leave_aligned();
        a.V[s].w = __uint_as_float(vw0);
        // a large body searches the stencil of the cell list every step
        a.NL[s] = (mw & PF_INACTIVE) ? nl_word(0, 0) : nl_word(-1, 0);
        if (!(mw & PF_INACTIVE)) atomicAdd(&a.STAT[5], 1ull);
        return;
    }
#pragma unroll
    for (int r = 0; r < 3; r++) {
        const int ry = cy - 1 + r;
        const bool row = ry >= 0 && ry < S.h;
        e[r] = row ? max(e[r] + pos0, S.p0) : 0;
        e1[r] = row ? min(e1[r] + pos0, S.p0 + S.np) : 0;
    }
    // live bond partners stay out of the list: their records carry the pair until the
    // record dies, when the step kernel adds the partner to the list
    auto bonded = [&](int oj) {
        if (brs.y <= brs.x) return false;
        for (int q = brs.x; q < brs.y; q++) {
            const int4 br = a.BR[q];
            if (br.w != 0 && br.x == oj) return true;
        }
        return false;
    };
    // The list is sorted by partner id (the step kernel forces its hits in that order):
    // candidates are appended to this thread's column of the block's LDS area (entry k
    // of thread t at k * blockDim.x + t, so that a wave's lanes hit distinct banks) and
    // sorted once at the end.  Appending costs one LDS store, so a wave's candidate
    // rounds, which run whenever any lane still has candidates, stay cheap.
    int2* const lst = nl_lds + threadIdx.x;   // {partner id, slot}, stride blockDim.x
    const int lstride = (int)blockDim.x;
    int cnt = 0;
    auto add = [&](int j, int id) {
        if (cnt < NLMAX) lst[cnt * lstride] = make_int2(id, j);
        cnt++;
    };
    const float lr = p.z + S.skin;
    const int lo = S.p0, hi = S.p0 + S.np - 1;   // the scene's cell-list positions
    while (e[0] < e1[0] || e[1] < e1[1] || e[2] < e1[2]) {
        // NL_ROW_STEP entries of each row: their slots, then their positions (ids in the
        // last words), each in one round of loads
        constexpr int K = 3 * NL_ROW_STEP;
        int j[K];
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
            for (int u = 0; u < NL_ROW_STEP; u++) j[r * NL_ROW_STEP + u] = a.SIDX[min(max(e[r] + u, lo), hi)];
        asm volatile("" ::: "memory");
        float4 q[K];
#pragma unroll
        for (int k = 0; k < K; k++) q[k] = a.P[cur][j[k]];
        asm volatile("" ::: "memory");
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int r = k / NL_ROW_STEP, u = k % NL_ROW_STEP;
            if (e[r] + u >= e1[r] || j[k] == s) continue;
            const float dx = q[k].x - p.x, dy = q[k].y - p.y;
            const float xx = dx * dx;
            const float yy = dy * dy;
            const float L = (lr + q[k].z) * 1.0001f;
            const int idk = __float_as_int(q[k].w);
            if (xx + yy < L * L && (S.nlb == 0 || q[k].z <= S.rsmax) && !bonded(idk)) add(j[k], idk);
        }
#pragma unroll
        for (int r = 0; r < 3; r++) e[r] += NL_ROW_STEP;
    }
    // the scene's large bodies, from their list (never through the cells)
    for (int k = 0; k < S.nlb; k++) {
        const int jl = a.INV[S.p0 + a.LB[S.lb0 + k]];
        const float4 q = a.P[cur][jl];
        const float dx = q.x - p.x, dy = q.y - p.y;
        const float xx = dx * dx;
        const float yy = dy * dy;
        const float L = (lr + q.z) * 1.0001f;
        if (xx + yy < L * L) {
            const int idl = __float_as_int(q.w);
            if (!bonded(idl)) add(jl, idl);
        }
    }
    int first = 0;
    uint32_t nm = 0u;   // contacts moved beside the new list: bits of their candidates
    if (cnt <= NLMAX) {
        // the old contacts' candidate slots and forces (the first OLD_AHEAD of them),
        // loaded now so that they arrive during the sort
        constexpr int OLD_AHEAD = 6;
        int ojs[OLD_AHEAD];
        float oft[OLD_AHEAD];
        {
            uint32_t mk = (al && cnt <= KMAX) ? om : 0u;
#pragma unroll
            for (int u = 0; u < OLD_AHEAD; u++) {
                ojs[u] = -1;
                oft[u] = 0.0f;
                if (mk != 0u) {
                    const int k = __ffs(mk) - 1;
                    ojs[u] = k == 0 ? nl_first(w0old) : a.NL[k * cap + s];
                    oft[u] = a.FT[k * cap + s];
                    mk &= mk - 1u;
                }
            }
        }
        // insertion sort of this thread's entries by partner id (ids are unique)
        for (int k = 1; k < cnt; k++) {
            const int2 x = lst[k * lstride];
            int m = k - 1;
            for (; m >= 0; m--) {
                const int2 y = lst[m * lstride];
                if (y.x < x.x) break;
                lst[(m + 1) * lstride] = y;
            }
            lst[(m + 1) * lstride] = x;
        }
        // Old and new candidates are both in partner-id order, so one forward walk over
        // the new list finds each old contact's slot; its tangential force goes into
        // that entry's id word (ids are not needed once sorted).  Every old row is read
        // before any row is written.  An old contact that is no longer a candidate, or a
        // list longer than the mask, leaves the contacts as a partner-id list.
        bool keep = al && cnt <= KMAX;
        if (keep) {
            int kp = 0, u = 0;
            for (uint32_t mk = om; mk != 0u; mk &= mk - 1u, u++) {
                int js;
                float f;
                if (u < OLD_AHEAD) {
                    // selects, not an indexed register array (which would live in scratch)
                    js = ojs[0];
                    f = oft[0];
#pragma unroll
                    for (int v = 1; v < OLD_AHEAD; v++) {
                        if (u == v) {
                            js = ojs[v];
                            f = oft[v];
                        }
                    }
                } else {
                    const int k = __ffs(mk) - 1;
                    js = k == 0 ? nl_first(w0old) : a.NL[k * cap + s];
                    f = a.FT[k * cap + s];
                }
                while (kp < cnt && lst[kp * lstride].y != js) kp++;
                if (kp >= cnt) {
                    keep = false;
                    break;
                }
                lst[kp * lstride].x = __float_as_int(f);
                nm |= 1u << kp;
                kp++;
            }
        }
        if (al && !keep) {
            leave_aligned();
            nm = 0u;
        }
        for (int k = 1; k < cnt; k++) a.NL[k * cap + s] = lst[k * lstride].y;
        for (uint32_t mk = nm; mk != 0u; mk &= mk - 1u) {
            const int k = __ffs(mk) - 1;
            a.FT[k * cap + s] = __int_as_float(lst[k * lstride].x);
        }
        if (cnt > 0) first = lst[0].y;
    } else if (al) {
        leave_aligned();
    }
    a.V[s].w = __uint_as_float(vw0 | (nm << VW_MASK_SHIFT));
    a.NL[s] = nl_word(cnt <= NLMAX ? cnt : -1, first);
    if (cnt > NLMAX) atomicAdd(&a.STAT[5], 1ull);   // particles on the stencil fallback
}

// The rebuild flag is double-buffered by position buffer: a drifting launch sets
// RB[nxt] for the rebuild that follows it, and that rebuild's first kernel clears
// the other flag, whose rebuild has run (no kernel of this chain reads it).
__global__ void k_rb_bin(StepArgs args, int nlb) {
    if (blockIdx.x == 0 && threadIdx.x == 0) args.a.RB[args.cur ^ 1] = 0;
    if (args.a.RB[args.cur] == 0) return;
    const int lbs = (int)blockDim.x;
    for (int lb = (int)blockIdx.x; lb < nlb; lb += (int)gridDim.x) rb_bin_slot(args, lb * lbs + (int)threadIdx.x, lbs);
}

__global__ void k_rb_scatter(StepArgs args, int nlb) {
    if (args.a.RB[args.cur] == 0) return;
    clear_tile_sums(args.a);
    const int lbs = (int)blockDim.x;
    for (int lb = (int)blockIdx.x; lb < nlb; lb += (int)gridDim.x)
        rb_scatter_slot(args, lb * lbs + (int)threadIdx.x, lbs);
}

// The lists are built in cell-list order (thread e takes the slot at list position e),
// so that a wave's stencils overlap whatever the storage order: the re-sort moves
// candidates to the front of their units, out of cell order.
// The list build's LDS: NLMAX {id, slot} entries per thread (dynamic, sized by the
// launch: 8 KB at 64 threads to 32 KB at the largest block, 256).
constexpr int NL_LDS_THREADS = 256;

__global__ void __launch_bounds__(NL_LDS_THREADS) k_build_nl(StepArgs args, int nlb) {
    if (args.a.RB[args.cur] == 0) return;
    extern __shared__ int2 nl_lds[];
    const int lbs = (int)blockDim.x;
    // the scene's list is a permutation of its slots: list position and slot share the
    // scene.  One live scene: its record is read into registers before any store (read
    // after one, every field would be loaded again).
    auto block = [&](const SceneDev& S, int lb) {
        const int e = lb * lbs + (int)threadIdx.x;   // the scene's list positions are its slot range
        if (e < S.p0 || e >= S.p0 + S.np) return;
        const int pos0 = S.p0 - args.a.START[S.c0];   // cell-list position -> slot-list index
        const int s = args.a.SIDX[e];
        if (s >= S.p0 && s < S.p0 + S.np) build_nl_slot(args, S, pos0, s, nl_lds);
    };
    if (args.one_scene >= 0) {
        const SceneDev S = args.scenes[args.one_scene];
        for (int lb = (int)blockIdx.x; lb < nlb; lb += (int)gridDim.x) block(S, lb);
    } else {
        for (int lb = (int)blockIdx.x; lb < nlb; lb += (int)gridDim.x) {
            const int sid = __builtin_amdgcn_readfirstlane(args.blk_scene[lb]);
            if (sid < 0) continue;
            const SceneDev S = args.scenes[sid];
            block(S, lb);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&args.a.STAT[4], 1ull);   // rebuilds
}

// ---- (cell, id) order of the cell list (the debug view of stages 1-3) ----
// Each cell's entries sorted by particle id, one thread per cell.
__global__ void k_cell_sort(StepArgs args, int cell_lo, int cell_hi, int p0) {
    const DevArrays& a = args.a;
    const int c = cell_lo + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (c >= cell_hi) return;
    // cells [cell_lo, cell_hi) of one scene whose slots start at p0
    const int pos0 = p0 - a.START[cell_lo];
    const int e0 = a.START[c] + pos0, e1 = a.START[c + 1] + pos0;
    const int plim = p0 + (a.START[cell_hi] - a.START[cell_lo]);  // end of the scene's list
    if (e0 < p0 || e1 > plim || e1 < e0) return;
    for (int e = e0 + 1; e < e1; e++) {
        const int v = a.SIDX[e];
        const int key = a.ID[v];
        int f = e - 1;
        while (f >= e0 && a.ID[a.SIDX[f]] > key) {
            a.SIDX[f + 1] = a.SIDX[f];
            f--;
        }
        a.SIDX[f + 1] = v;
    }
}

// ---- storage re-sort ----
// 1. move every per-particle array to its destination: slot j takes the particle
//    order[j] (the scene's cells, sentinel included, hold exactly its np slots, so the
//    list, and with it the order, is a permutation of the scene's slots).
__global__ void k_reorder_move(ReorderArgs r) {
    const int sid = r.blk_scene[blockIdx.x];
    if (sid < 0) return;
    const SceneDev& S = r.scenes[sid];
    const int j = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (j < S.p0 || j >= S.p0 + S.np) return;
    const int s = r.order[j];
    const int cur = r.cur;
    r.dst.P[cur][j] = r.src.P[cur][s];
    const float4 v = r.src.V[s];
    r.dst.V[j] = v;
    if (vword(v) & VW_EXPLICIT) r.dst.DX[j] = r.src.DX[s];
    const int2 m = make_int2(r.src.M[s], r.src.ID[s]);
    r.dst.M[j] = (int)((uint32_t)m.x & ~PF_ALIGNED);
    r.dst.ID[j] = m.y;
    r.dst.A[j] = r.src.A[s];
    r.dst.TH[j] = r.src.TH[s];
    if (r.with_ext) r.dst.EXT[j] = r.src.EXT[s];
    if (r.with_data) {
        r.dst.DATA0[j] = r.src.DATA0[s];
        r.dst.DATA1[j] = r.src.DATA1[s];
    }
    const int c = cc_of((uint32_t)m.x);
    const int cap = (int)r.src.cap;
    // the lists arrive as partner-id lists (the re-sort follows a list build); one kept
    // beside the neighbour list would leave as one
    if ((uint32_t)m.x & PF_ALIGNED) {
        contacts_from_aligned(r.src, s, contact_mask(v), r.dst.CP, r.dst.CF, j);
    } else {
        for (int q = 0; q < c && q < KMAX; q++) {
            r.dst.CP[q * cap + j] = r.src.CP[q * cap + s];
            r.dst.CF[q * cap + j] = r.src.CF[q * cap + s];
        }
    }
    // the neighbour list, its candidates renamed to their destination slots
    const int w0 = r.src.NL[s];
    const int nc = nl_count(w0);
    r.dst.NL[j] = nl_word(nc, nc > 0 ? r.newslot[nl_first(w0)] : 0);
    r.dst.XR[j] = r.src.XR[s];
    r.dst.BRS[j] = r.src.BRS[s];
    for (int q = 1; q < nc; q++) r.dst.NL[q * cap + j] = r.newslot[r.src.NL[q * cap + s]];
    r.dst.INV[S.p0 + m.y] = j;
}

// 0. the destination of every slot.  Units of PART_BLOCKS consecutive blocks of list
//    positions (cut at scene boundaries: the scene table's blocks never straddle
//    scenes) keep their positions, partitioned stably into three classes: particles
//    with more neighbour-list candidates than one round of the step kernel's contact
//    loop takes (PART_HEAVY or more, or the stencil fallback), then those with fewer,
//    then those with none.  Waves then hold particles that take the same number of
//    candidate rounds (a wave runs as many as its busiest lane), waves without
//    candidates skip the narrow phase, and particles stay within their unit (16384
//    slots at blocks of 256), near their cell-order neighbours.  The cell list is
//    renamed to the destination slots in place.  One workgroup per unit,
//    blockDim.x = the logical block size.
constexpr int PART_BLOCKS = 64;
constexpr int PART_HEAVY = 5;   // the step kernel's list path tests candidates four at a time
static_assert(PART_BLOCKS <= 64, "one flag bit per block in a 64-bit mask");

// class of the particle at list position j: 0 heavy, 1 light, 2 no candidates
__device__ __forceinline__ int part_class(const ReorderArgs& r, const SceneDev& S, int j, int& s) {
    s = r.src.SIDX[j];
    if (s < S.p0 || s >= S.p0 + S.np) s = j;     // a damaged list keeps the slot in place
    const int nc = nl_count(r.src.NL[s]);
    return nc == 0 ? 2 : ((nc < 0 || nc >= PART_HEAVY) ? 0 : 1);
}

__global__ void __launch_bounds__(256) k_partition(ReorderArgs r, int nlb) {
    const int lbs = (int)blockDim.x;
    const int b0 = (int)blockIdx.x * PART_BLOCKS;
    const int b1 = min(b0 + PART_BLOCKS, nlb);
    const int tid = (int)threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6, nw = lbs >> 6;
    const uint64_t below = (1ull << lane) - 1ull;
    __shared__ uint64_t s_bal[2][PART_BLOCKS][4];   // lanes of class 0 / 1 of each wave of each block
    __shared__ int s_pre[2][PART_BLOCKS];           // class 0 / 1 slots of the segment before each block
    __shared__ int s_tot[2];
    int bs = b0;
    while (bs < b1) {                               // segments: runs of blocks of one scene
        const int sid = r.blk_scene[bs];
        int be = bs + 1;
        while (be < b1 && r.blk_scene[be] == sid) be++;
        if (sid >= 0) {
            const SceneDev& S = r.scenes[sid];
            // this thread's slot of every block of the segment: two class bits per block,
            // the loads of several blocks in flight together
            uint64_t m0 = 0, m1 = 0;
#pragma unroll 8
            for (int b = bs; b < be; b++) {
                int s;
                const int c = part_class(r, S, b * lbs + tid, s);
                if (c == 0) m0 |= 1ull << (b - bs);
                if (c == 1) m1 |= 1ull << (b - bs);
            }
            for (int b = bs; b < be; b++) {
                const uint64_t f0 = __ballot((m0 >> (b - bs)) & 1ull);
                const uint64_t f1 = __ballot((m1 >> (b - bs)) & 1ull);
                if (lane == 0) {
                    s_bal[0][b - bs][wv] = f0;
                    s_bal[1][b - bs][wv] = f1;
                }
            }
            __syncthreads();
            if (tid < 2) {
                int run = 0;
                for (int q = 0; q < be - bs; q++) {
                    s_pre[tid][q] = run;
                    for (int w = 0; w < nw; w++) run += __popcll(s_bal[tid][q][w]);
                }
                s_tot[tid] = run;
            }
            __syncthreads();
            const int seg0 = bs * lbs;
#pragma unroll 4
            for (int b = bs; b < be; b++) {
                const int k = b - bs;
                const int j = b * lbs + tid;
                int s = r.src.SIDX[j];
                if (s < S.p0 || s >= S.p0 + S.np) s = j;
                int r0 = __popcll(s_bal[0][k][wv] & below);    // lanes of each class before this one
                int r1 = __popcll(s_bal[1][k][wv] & below);
                for (int w = 0; w < wv; w++) {
                    r0 += __popcll(s_bal[0][k][w]);
                    r1 += __popcll(s_bal[1][k][w]);
                }
                const int c0 = s_pre[0][k], c1 = s_pre[1][k];   // class 0 / 1 slots before block b
                int dst;
                if ((m0 >> k) & 1ull) dst = seg0 + c0 + r0;
                else if ((m1 >> k) & 1ull) dst = seg0 + s_tot[0] + c1 + r1;
                else dst = seg0 + s_tot[0] + s_tot[1] + (k * lbs - c0 - c1) + (tid - r0 - r1);
                r.order[dst] = s;
                r.newslot[s] = dst;
                r.src.SIDX[j] = dst;
            }
        }
        bs = be;
        __syncthreads();   // s_bal / s_pre / s_tot are rewritten by the next segment
    }
}

// ---- observables ----
__global__ void k_kinetic(DevArrays a, int cur, int m0, int nm, int p0, int np, double* out_by_id) {
    const int s = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (s >= np) return;
    const int slot = p0 + s;
    const int2 m = make_int2(a.M[slot], a.ID[slot]);
    double e = 0.0;
    if (!(m.x & PF_INACTIVE)) {
        const int mat = m.x & PF_MAT_MASK;
        const double rr = a.P[cur][slot].z;
        const double mass = (double)a.MAT[m0 + (mat < nm ? mat : 0)].density * 3.141592653589793 * rr * rr;
        const double inertia = 0.5 * mass * rr * rr;
        const float4 v = a.V[slot];
        const double vx = v.x, vy = v.y, w = v.z;
        e = 0.5 * mass * (vx * vx + vy * vy) + 0.5 * inertia * w * w;
    }
    if (m.y >= 0 && m.y < np) out_by_id[m.y] = e;
}

// Group sums (grip force and displacement of a tensile test): per listed id, the f64
// terms {m ax, m ay, x, y, vx, vy, |data0|} with m the step's f32 mass; each block
// reduces its ids in a fixed tree order, the partials are added in block order.
constexpr int GROUP_BLOCKS = 64;
constexpr int GROUP_TERMS = 7;
__global__ void __launch_bounds__(256) k_group_partials(DevArrays a, int cur, int m0, int nm, int p0,
                                                         const int32_t* ids, int n_ids, int with_data,
                                                         double* partials) {
    __shared__ double s_t[GROUP_TERMS][256];
    double t[GROUP_TERMS] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int k = (int)(blockIdx.x * 256 + threadIdx.x); k < n_ids; k += (int)(gridDim.x * 256)) {
        const int slot = a.INV[p0 + ids[k]];
        const int mat = a.M[slot] & PF_MAT_MASK;
        const float4 p = a.P[cur][slot];
        const float density = a.MAT[m0 + (mat < nm ? mat : 0)].density;
        const float rr = p.z;
        const float mass = density * PI_F * rr * rr;
        const float4 acc = a.A[slot];
        const float4 v = a.V[slot];
        t[0] += (double)mass * (double)acc.x;
        t[1] += (double)mass * (double)acc.y;
        t[2] += (double)p.x;
        t[3] += (double)p.y;
        t[4] += (double)v.x;
        t[5] += (double)v.y;
        if (with_data) t[6] += fabs((double)a.DATA0[slot].x);
    }
#pragma unroll
    for (int c = 0; c < GROUP_TERMS; c++) s_t[c][threadIdx.x] = t[c];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int c = 0; c < GROUP_TERMS; c++) s_t[c][threadIdx.x] += s_t[c][threadIdx.x + w];
        }
        __syncthreads();
    }
    if ((int)threadIdx.x < GROUP_TERMS) partials[blockIdx.x * GROUP_TERMS + threadIdx.x] = s_t[threadIdx.x][0];
}

__global__ void k_group_total(const double* partials, int nblocks, double* out) {
    const int c = (int)threadIdx.x;
    if (c >= GROUP_TERMS) return;
    double s = 0.0;
    for (int b = 0; b < nblocks; b++) s += partials[b * GROUP_TERMS + c];
    out[c] = s;
}

// Sum of the contact counts of a scene (mean contacts per particle) - integer,
// order independent.  A grid of COUNT_BLOCKS strides over the scene, so only that
// many same-address atomics meet in L2.
constexpr int COUNT_BLOCKS = 256;
// Slots [p0, p0 + np): every contact list kept beside its neighbour list becomes a
// partner-id list in CP / CF (for reads from the host; the next step kernel takes
// either form).
__global__ void __launch_bounds__(256) k_contacts_by_id(StepArgs args, int p0, int np) {
    const DevArrays& a = args.a;
    const int s = p0 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (s >= p0 + np) return;
    const uint32_t mw = (uint32_t)a.M[s];
    if (!(mw & PF_ALIGNED)) return;
    contacts_from_aligned(a, s, contact_mask(a.V[s]), a.CP, a.CF, s);
    a.M[s] = (int)(mw & ~PF_ALIGNED);
}

__global__ void __launch_bounds__(256) k_count_contacts(const int32_t* m, int p0, int np, unsigned long long* out) {
    __shared__ int s_w[4];
    int c = 0;
    for (int s = (int)(blockIdx.x * blockDim.x + threadIdx.x); s < np; s += (int)(gridDim.x * blockDim.x))
        c += cc_of((uint32_t)m[p0 + s]);
    c = wave_sum(c);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        const int tot = s_w[0] + s_w[1] + s_w[2] + s_w[3];
        if (tot) atomicAdd(out, (unsigned long long)tot);
    }
}

// ---- large-body search: one wave per large body ----
// Runs before the step kernel of every launch with a force half in a scene with large
// bodies.  The small discs that overlap large body g (the contact search's predicate
// over the stencil of the last cell list around its build position, with the reach
// ceil((R + r_small + skin) / cs) + 1 of the step kernel's own stencil walk), found by the
// wave's 64 lanes together, in any order, into LBH[g]; their count into LBHC[g] (-1
// beyond LBH_MAX).  The large body's thread then tests these instead of walking the
// stencil alone: the same candidates pass the same predicate, and its contact search
// sorts what it admits, so the result does not depend on the order found here.
__global__ void __launch_bounds__(64) k_large_search(StepArgs args) {
    const DevArrays& a = args.a;
    const int g = (int)blockIdx.x;
    const int sid = g / LB_MAX, k = g % LB_MAX;
    const SceneDev& S = args.scenes[sid];
    const uint32_t fl = (uint32_t)args.lf[sid].x;
    if (!(fl & SF_ACTIVE) || !(fl & SF_SIM) || !(fl & SF_NARROW) || k >= S.nlb) return;
    __shared__ int cnt;
    __shared__ int hits[LBH_MAX];
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    const int s = a.INV[S.p0 + a.LB[S.lb0 + k]];
    const float4 p = a.P[args.cur][s];
    const float2 xr0 = a.XR[s];
    int cx, cy;
    cell_of(S, xr0.x, xr0.y, cx, cy);
    const int pos0 = S.p0 - a.START[S.c0];
    const int reach = max(1, (int)ceilf((p.z + S.rsmax + S.skin) * S.inv_cs) + 1);
    const int xa = max(cx - reach, 0), xb = min(cx + reach, S.w - 1);
    const int ry0 = max(cy - reach, 0), nrows = min(cy + reach, S.h - 1) - ry0 + 1;
    const int lane = (int)threadIdx.x;
    // Up to 64 stencil rows at a time: each lane takes one row's range of the cell list,
    // a prefix over the lanes lays the rows' entries end to end, and every lane then takes
    // LS_K entries per round (its row found by a binary search over the prefix), their
    // slots and then their positions each in one round of loads.
    constexpr int LS_K = 4;
    for (int rb = 0; rb < nrows; rb += 64) {
        int e0 = 0, len = 0;
        if (rb + lane < nrows) {
            const int ry = ry0 + rb + lane;
            const int s0 = a.START[S.c0 + ry * S.w + xa];
            const int s1 = a.START[S.c0 + ry * S.w + xb + 1];
            e0 = max(s0 + pos0, S.p0);
            len = max(min(s1 + pos0, S.p0 + S.np) - e0, 0);
        }
        const int incl = wave_incl_scan(len);
        const int total = __shfl(incl, 63, 64);
        for (int t0 = 0; t0 < total; t0 += 64 * LS_K) {
            int jj[LS_K];
#pragma unroll
            for (int u = 0; u < LS_K; u++) {
                const int t = t0 + u * 64 + lane;
                const int tc = min(t, total - 1);
                // the first row whose inclusive prefix passes tc (all 64 lanes take part)
                int lo = 0;
#pragma unroll
                for (int step = 32; step > 0; step >>= 1) {
                    const int probe = __shfl(incl, lo + step - 1, 64);
                    if (probe <= tc) lo += step;
                }
                const int row_incl = __shfl(incl, lo, 64);
                const int row_len = __shfl(len, lo, 64);
                const int row_e0 = __shfl(e0, lo, 64);
                const int e = row_e0 + tc - (row_incl - row_len);   // in range for every lane
                const int je = a.SIDX[e];
                jj[u] = t < total ? je : -1;
            }
            asm volatile("" ::: "memory");
            float4 q[LS_K];
#pragma unroll
            for (int u = 0; u < LS_K; u++) q[u] = a.P[args.cur][jj[u] >= 0 ? jj[u] : s];
            asm volatile("" ::: "memory");
#pragma unroll
            for (int u = 0; u < LS_K; u++) {
                const int j = jj[u];
                if (j >= 0 && j != s && q[u].z <= S.rsmax && overlaps(q[u], p)) {
                    const int w = atomicAdd(&cnt, 1);
                    if (w < LBH_MAX) hits[w] = j;
                }
            }
        }
    }
    __syncthreads();
    const int c = cnt;
    for (int i = (int)threadIdx.x; i < min(c, LBH_MAX); i += 64) a.LBH[g * LBH_MAX + i] = hits[i];
    if (threadIdx.x == 0) a.LBHC[g] = c <= LBH_MAX ? c : -1;
}

// ---- bond builder (State::regen_bonds rule on the cell list) ----
// For particle id i: every j > i with sqrt((xj-xi)^2 + (yj-yi)^2) < (ri + rj) * factor + gap,
// the `maxb` smallest j kept.  Stencil of (2*reach+1)^2 cells.
__global__ void k_bond_pairs(StepArgs args, int reach, int maxb, float factor, float gap, int32_t* out_cnt,
                             int32_t* out_j) {
    const DevArrays& a = args.a;
    const int s = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int sid = args.blk_scene[blockIdx.x];
    if (sid < 0) return;
    const SceneDev& S = args.scenes[sid];
    if (s < S.p0 || s >= S.p0 + S.np) return;
    const int2 m = make_int2(a.M[s], a.ID[s]);
    const int i = m.y;
    if (m.x & PF_INACTIVE) {
        if (i < S.n_real) out_cnt[i] = 0;
        return;
    }
    const float4 p = a.P[args.cur][s];
    int cx, cy;
    cell_of(S, p.x, p.y, cx, cy);
    const int pos0 = S.p0 - a.START[S.c0];
    int best[16];
    int n = 0;
    for (int ry = max(cy - reach, 0); ry <= min(cy + reach, S.h - 1); ry++) {
        const int e0 = max(a.START[S.c0 + ry * S.w + max(cx - reach, 0)] + pos0, S.p0);
        const int e1 = min(a.START[S.c0 + ry * S.w + min(cx + reach, S.w - 1) + 1] + pos0, S.p0 + S.np);
        for (int e = e0; e < e1; e++) {
            const int js = a.SIDX[e];
            const int j = a.ID[js];
            if (j <= i) continue;
            const float4 q = a.P[args.cur][js];
            const float dx = q.x - p.x, dy = q.y - p.y;
            const float dist = sqrtf(dx * dx + dy * dy);
            if (!(dist < (p.z + q.z) * factor + gap)) continue;
            // keep the maxb smallest j, ascending
            int pos = 0;
            for (int k = 0; k < n; k++) pos += best[k] < j ? 1 : 0;
            if (pos >= maxb) continue;
            for (int k = min(n, maxb - 1); k > pos; k--) best[k] = best[k - 1];
            best[pos] = j;
            n = min(n + 1, maxb);
        }
    }
    out_cnt[i] = n;
    for (int k = 0; k < n; k++) out_j[(int64_t)i * maxb + k] = best[k];
}

// ---- launchers (host context) ----

// With events (timing only, never under graph capture) hipExtLaunchKernelGGL stamps
// them with the kernel's own start and end.
template <class K, class... A>
static void launch_k(K kernel, dim3 g, dim3 b, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1, A... args) {
    if (ev0) hipExtLaunchKernelGGL(kernel, g, b, 0, st, ev0, ev1, 0, args...);
    else hipLaunchKernelGGL(kernel, g, b, 0, st, args...);
}
// the same with `lds` bytes of dynamic LDS per workgroup
template <class K, class... A>
static void launch_k_lds(K kernel, dim3 g, dim3 b, size_t lds, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1,
                         A... args) {
    if (ev0) hipExtLaunchKernelGGL(kernel, g, b, lds, st, ev0, ev1, 0, args...);
    else hipLaunchKernelGGL(kernel, g, b, lds, st, args...);
}

template <class K>
static void launch_step_kernel(K kernel, dim3 g, dim3 b, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1,
                               const StepArgs& a) {
    launch_k(kernel, g, b, st, ev0, ev1, a);
}

hipError_t launch_step(const StepArgs& args, int nblocks, int block, bool f64, bool bonds, bool large,
                       hipStream_t st, hipEvent_t ev0, hipEvent_t ev1) {
    dim3 g(nblocks), b(block);
    StepArgs a = args;
    a.lbs = block;
    if (f64) {
        if (bonds) launch_step_kernel(k_step_bonds_f64, g, b, st, ev0, ev1, a);
        else launch_step_kernel(k_step_plain_f64, g, b, st, ev0, ev1, a);
    } else {
        if (bonds) launch_step_kernel(k_step_bonds_f32, g, b, st, ev0, ev1, a);
        else if (large) launch_step_kernel(k_step_plain_large, g, b, st, ev0, ev1, a);
        else launch_step_kernel(k_step_plain, g, b, st, ev0, ev1, a);
    }
    return hipGetLastError();
}

hipError_t launch_bond_forces(const StepArgs& args, int nrec, int lbs, bool f64, hipStream_t st,
                              hipEvent_t ev0, hipEvent_t ev1) {
    const int b = 256;
    const dim3 g((unsigned)((nrec + b - 1) / b));
    if (f64) launch_k(k_bond_forces<true>, g, dim3(b), st, ev0, ev1, args, nrec, lbs);
    else if (args.one_scene >= 0) launch_k(k_bond_forces1, g, dim3(b), st, ev0, ev1, args, nrec);
    else launch_k(k_bond_forces<false>, g, dim3(b), st, ev0, ev1, args, nrec, lbs);
    return hipGetLastError();
}

hipError_t launch_record_slots(const StepArgs& args, int nrec, int lbs, hipStream_t st, hipEvent_t ev0,
                               hipEvent_t ev1) {
    const int b = 256;
    launch_k(k_record_slots, dim3((unsigned)((nrec + b - 1) / b)), dim3(b), st, ev0, ev1, args, nrec, lbs);
    return hipGetLastError();
}

hipError_t launch_bin(const StepArgs& args, int nblocks, int block, hipStream_t st) {
    hipLaunchKernelGGL(k_bin, dim3(nblocks), dim3(block), 0, st, args);
    return hipGetLastError();
}


int scan_tiles(int n) { return tile_count(n); }

// gate: NULL to scan unconditionally, else the rebuild flag
hipError_t launch_scan(const DevArrays& a, int ncells, const int32_t* gate, hipStream_t st, hipEvent_t ev0,
                       hipEvent_t ev1) {
    const int nt = scan_tiles(ncells);
    launch_k(k_scan_tiles, dim3(nt), dim3(SCAN_THREADS), st, ev0, ev1, a.CNT.p, a.START.p, ncells, a.TSUM.p,
             gate);
    return hipGetLastError();
}

hipError_t launch_rb_bin(const StepArgs& args, int nlb, int grid, int block, hipStream_t st, hipEvent_t ev0,
                         hipEvent_t ev1) {
    launch_k(k_rb_bin, dim3(grid), dim3(block), st, ev0, ev1, args, nlb);
    return hipGetLastError();
}

hipError_t launch_rb_scatter(const StepArgs& args, int nlb, int grid, int block, hipStream_t st, hipEvent_t ev0,
                             hipEvent_t ev1) {
    launch_k(k_rb_scatter, dim3(grid), dim3(block), st, ev0, ev1, args, nlb);
    return hipGetLastError();
}

hipError_t launch_build_nl(const StepArgs& args, int nlb, int grid, int block, hipStream_t st, hipEvent_t ev0,
                           hipEvent_t ev1) {
    if (block > NL_LDS_THREADS) return hipErrorInvalidValue;
    launch_k_lds(k_build_nl, dim3(grid), dim3(block), sizeof(int2) * NLMAX * (size_t)block, st, ev0, ev1, args, nlb);
    return hipGetLastError();
}

hipError_t launch_rb_store(int32_t* rb, int v, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1) {
    launch_k(k_rb_store, dim3(1), dim3(64), st, ev0, ev1, rb, v);
    return hipGetLastError();
}

hipError_t launch_scatter(const StepArgs& args, int nblocks, int block, hipStream_t st) {
    hipLaunchKernelGGL(k_scatter, dim3(nblocks), dim3(block), 0, st, args);
    return hipGetLastError();
}

hipError_t launch_cell_sort(const StepArgs& args, int cell_lo, int cell_hi, int p0, hipStream_t st) {
    const int b = 256;
    const int n = cell_hi - cell_lo;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_cell_sort, dim3((n + b - 1) / b), dim3(b), 0, st, args, cell_lo, cell_hi, p0);
    return hipGetLastError();
}

hipError_t launch_reorder_move(const ReorderArgs& r, int nblocks, int block, hipStream_t st, hipEvent_t ev0,
                               hipEvent_t ev1) {
    launch_k(k_reorder_move, dim3(nblocks), dim3(block), st, ev0, ev1, r);
    return hipGetLastError();
}

hipError_t launch_partition(const ReorderArgs& r, int nblocks, int block, hipStream_t st, hipEvent_t ev0,
                            hipEvent_t ev1) {
    launch_k(k_partition, dim3((nblocks + PART_BLOCKS - 1) / PART_BLOCKS), dim3(block), st, ev0, ev1, r, nblocks);
    return hipGetLastError();
}

hipError_t launch_kinetic(const DevArrays& a, int cur, int m0, int nm, int p0, int np, double* out, hipStream_t st) {
    const int b = 256;
    hipLaunchKernelGGL(k_kinetic, dim3((np + b - 1) / b), dim3(b), 0, st, a, cur, m0, nm, p0, np, out);
    return hipGetLastError();
}

hipError_t launch_contacts_by_id(const StepArgs& args, int p0, int np, hipStream_t st) {
    const int b = 256;
    if (np > 0) hipLaunchKernelGGL(k_contacts_by_id, dim3((np + b - 1) / b), dim3(b), 0, st, args, p0, np);
    return hipGetLastError();
}

hipError_t launch_count_contacts(const int32_t* m, int p0, int np, unsigned long long* out, hipStream_t st) {
    const int b = 256;
    const int nb = (np + b - 1) / b;
    hipLaunchKernelGGL(k_count_contacts, dim3(nb < COUNT_BLOCKS ? nb : COUNT_BLOCKS), dim3(b), 0, st, m, p0, np, out);
    return hipGetLastError();
}

hipError_t launch_large_search(const StepArgs& args, int nblocks, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1) {
    launch_k(k_large_search, dim3(nblocks), dim3(64), st, ev0, ev1, args);
    return hipGetLastError();
}

hipError_t launch_bond_pairs(const StepArgs& args, int nblocks, int block, int reach, int maxb, float factor,
                             float gap, int32_t* out_cnt, int32_t* out_j, hipStream_t st) {
    hipLaunchKernelGGL(k_bond_pairs, dim3(nblocks), dim3(block), 0, st, args, reach, maxb, factor, gap, out_cnt,
                       out_j);
    return hipGetLastError();
}

// partials: GROUP_BLOCKS * GROUP_TERMS doubles; out: GROUP_TERMS doubles (device)
hipError_t launch_group_sums(const DevArrays& a, int cur, int m0, int nm, int p0, const int32_t* ids, int n_ids,
                             int with_data, double* partials, double* out, hipStream_t st) {
    int nb = (n_ids + 255) / 256;
    nb = nb < 1 ? 1 : (nb > GROUP_BLOCKS ? GROUP_BLOCKS : nb);
    hipLaunchKernelGGL(k_group_partials, dim3(nb), dim3(256), 0, st, a, cur, m0, nm, p0, ids, n_ids, with_data,
                       partials);
    hipLaunchKernelGGL(k_group_total, dim3(1), dim3(64), 0, st, (const double*)partials, nb, out);
    return hipGetLastError();
}

}  // namespace psim
