// dem_oracle.cpp - CPU oracle for the DEM time step.  TEST INFRASTRUCTURE ONLY.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// this library, and only as a checker or as the timed CPU baseline.  The product
// path (physics-sim_amd/) never links or calls it.
//
// What it restates (reference files, read as text; the reference itself is Rust +
// WGSL and cannot be built or run here):
//   * LOM pass      src/shaders/physics/2D_LOM.wgsl:80-201
//   * Sim pass      src/shaders/physics/2D_Simulation.wgsl:108-478
//   * step driver   src/wgpu_prog.rs:1420-1473 (LOM then Sim, gen_per_frame times)
//   * grid sizing   src/wgpu_prog.rs:698-717 (= settings.rs:1955-1974)
//   * bond builder  src/state.rs:291-367 (State::regen_bonds)
//
// Invocations of each pass run one after another in id order.  That turns the
// races of the WGSL kernels on the cell grid and the coll_cont words into one
// defined behaviour.  The one exception is bonds[].index, written by the endpoint
// that breaks a bond and read by its partner in the same pass
// (2D_Simulation.wgsl:236-239, :324): a break becomes visible to the partner from
// the next pass.  That is what the reference's GPU does when both endpoints sit in
// one wave (the partner's load runs in lockstep with, not after, the store); its
// recorded trace data/Chain_T_7.csv shows it (row 80: the partner's data words are
// 0, as with the intact-bond law, not the contact force a same-pass read gives).
//
// Two slot policies:
//   ORC_MODE_REFERENCE  the WGSL's own rules: 14-slot table per particle, candidates
//                       from the fixed-capacity AABB bucket grid in grid order, new
//                       contacts in the last empty slot, forces summed in slot order.
//   ORC_MODE_CANONICAL  the semantics the HIP path implements: the candidate set is
//                       every overlapping pair (no bucket capacity), each particle's
//                       occupied slots are kept sorted by partner id b (forces summed in
//                       that order) and deleted slots are cleared to -1 words.
//   Both modes give the same pair sets, the same per-contact forces and the same bond
//   histories; they differ only in FP summation order (and in the reference's
//   overflow behaviour, which canonical mode lacks).
//
// FP semantics: every expression keeps the WGSL operand order; build with
// -ffp-contract=off so no multiply-add is fused; division and sqrt are IEEE.
// length(v) = sqrt(x*x + y*y); normalize(v) = v / length(v).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <vector>
#include <chrono>
#include <omp.h>

#include "../include/dem.h"

namespace {

constexpr float PI_F = 3.14159265358979323846f;  // WGSL `const PI` converted to f32
constexpr float G_F = 9.81f;
constexpr int DATA_SIZE = 7;
constexpr int REF_MAX_CONTACTS = 14;              // 2D_Simulation.wgsl:125

inline int32_t fbits(float f) { int32_t i; std::memcpy(&i, &f, 4); return i; }
inline float bitsf(int32_t i) { float f; std::memcpy(&f, &i, 4); return f; }

// WGSL i32(f32): truncation toward zero, saturating, NaN -> 0.
inline int32_t f2i(float v) {
    if (!(v == v)) return 0;
    if (v >= 2147483520.0f) return INT32_MAX;
    if (v <= -2147483648.0f) return INT32_MIN;
    return (int32_t)v;
}
inline float lenf(float x, float y) { float xx = x * x; float yy = y * y; return std::sqrt(xx + yy); }
inline double lend(double x, double y) { double xx = x * x; double yy = y * y; return std::sqrt(xx + yy); }
inline int sign_i(float x) { return x > 0.0f ? 1 : (x < 0.0f ? -1 : 0); }
inline float wmin(float a, float b) { return std::fmin(a, b); }
inline float wmax(float a, float b) { return std::fmax(a, b); }

struct Slot {  // struct Contact, 2D_Simulation.wgsl:19-26
    int32_t a, b;
    float ft, fbt, tb;
    int32_t bonded;
};
static_assert(sizeof(Slot) == 24, "Contact is 6 words");

struct V3 { float x, y, z; };

}  // namespace

extern "C" {

// Host-side arrays of one scene in the reference layouts (state.rs:69-93 plus the
// device-only buffers del_pos, del_rot, data, coll_cont and grid).
typedef struct orc_state {
    int32_t n;
    float* pos; float* vel; float* acc; float* rot; float* rot_vel; float* rot_acc;
    float* forces; float* radii; int32_t* fixity; int32_t* material_pointers;
    int32_t* bonds; int32_t n_bonds;
    const float* bond_strengths;      // optional {normal, shear} per bond (extension)
    float* contacts; int32_t slots;   // slots * 6 words per particle
    float* data;                      // 7N
    float* del_pos; float* del_rot;   // 2N, N
    int32_t* coll_cont;               // 4
    int32_t* grid; int64_t grid_len;  // reference mode only: w*h*cap
} orc_state;

}  // extern "C"

namespace {

// Every overlapping pair (the narrow-phase predicate of 2D_Simulation.wgsl:159,
// sqrt(dx*dx + dy*dy) < r_b + r_a) among the first n discs with r != 0 and finite
// positions, found over a cell list of side >= 2 r_max built here (3x3 stencil), the
// few discs much larger than the rest by brute force: found[a] = ascending partner
// ids.  O(N); the same set as the O(N^2) brute force.
void cell_list_pairs(int n, const float* pos, const float* radii, std::vector<std::vector<int32_t>>& found) {
    found.assign(n, {});
    auto px = [&](int i) { return pos[2 * i]; };
    auto py = [&](int i) { return pos[2 * i + 1]; };
    auto touch = [&](int a, int b) { return lenf(px(b) - px(a), py(b) - py(a)) < (radii[b] + radii[a]); };
    float xmin = INFINITY, xmax = -INFINITY, ymin = INFINITY, ymax = -INFINITY;
    std::vector<int> act;
    for (int i = 0; i < n; i++) {
        if (radii[i] == 0.0f) continue;
        if (!(px(i) == px(i)) || !(py(i) == py(i))) continue;
        act.push_back(i);
        xmin = std::min(xmin, px(i)); xmax = std::max(xmax, px(i));
        ymin = std::min(ymin, py(i)); ymax = std::max(ymax, py(i));
    }
    if (act.empty()) return;
    // A few discs far larger than the rest (a striker among grains) would size the cells
    // for everyone: with the radii in decreasing order, the first k <= 64 with
    // r_{k-1} > 2 r_k are paired by brute force, the others over cells of 2 r_k.
    std::vector<float> rs;
    rs.reserve(act.size());
    for (int i : act) rs.push_back(radii[i]);
    const size_t top = std::min<size_t>(rs.size(), 65);
    std::partial_sort(rs.begin(), rs.begin() + top, rs.end(), std::greater<float>());
    float rmax = rs[0];
    for (size_t k = 1; k < top; k++)
        if (rs[k - 1] > 2.0f * rs[k]) { rmax = rs[k]; break; }
    std::vector<int> big, small;
    for (int i : act) (radii[i] > rmax ? big : small).push_back(i);
    // large bodies against every disc (both directions)
    std::vector<std::vector<int32_t>> big_of(n);
    for (int a : big) {
        for (int b : act)
            if (b != a && touch(a, b)) {
                found[a].push_back(b);
                if (radii[b] <= rmax) big_of[b].push_back(a);
            }
    }
    double cs = std::max(2.0 * rmax, 1e-30);
    int64_t gw = std::min<int64_t>(std::max<int64_t>((int64_t)((xmax - xmin) / cs) + 1, 1), 1 << 14);
    int64_t gh = std::min<int64_t>(std::max<int64_t>((int64_t)((ymax - ymin) / cs) + 1, 1), 1 << 14);
    cs = std::max(cs, std::max((double)(xmax - xmin) / (double)gw, (double)(ymax - ymin) / (double)gh) * 1.000001);
    auto cell_of = [&](int i, int64_t& cx, int64_t& cy) {
        cx = std::min<int64_t>(std::max<int64_t>((int64_t)(((double)px(i) - xmin) / cs), 0), gw - 1);
        cy = std::min<int64_t>(std::max<int64_t>((int64_t)(((double)py(i) - ymin) / cs), 0), gh - 1);
    };
    std::vector<int64_t> start(gw * gh + 1, 0);
    std::vector<int32_t> order(small.size());
    for (int i : small) { int64_t cx, cy; cell_of(i, cx, cy); start[cy * gw + cx + 1]++; }
    for (int64_t c = 0; c < gw * gh; c++) start[c + 1] += start[c];
    std::vector<int64_t> fill(start.begin(), start.end() - 1);
    for (int i : small) { int64_t cx, cy; cell_of(i, cx, cy); order[fill[cy * gw + cx]++] = i; }
#pragma omp parallel for schedule(dynamic, 256)
    for (size_t k = 0; k < small.size(); k++) {
        int a = small[k];
        int64_t cx, cy;
        cell_of(a, cx, cy);
        std::vector<int32_t>& out = found[a];
        for (int64_t yy = std::max<int64_t>(cy - 1, 0); yy <= std::min<int64_t>(cy + 1, gh - 1); yy++)
            for (int64_t xx = std::max<int64_t>(cx - 1, 0); xx <= std::min<int64_t>(cx + 1, gw - 1); xx++) {
                int64_t c = yy * gw + xx;
                for (int64_t e = start[c]; e < start[c + 1]; e++) {
                    int b = order[e];
                    if (b != a && touch(a, b)) out.push_back(b);
                }
            }
        out.insert(out.end(), big_of[a].begin(), big_of[a].end());
        std::sort(out.begin(), out.end());
    }
    for (int a : big) std::sort(found[a].begin(), found[a].end());
}

struct Ctx {
    const dem_uniform& u;
    const dem_material* mats;
    int n_mats;
    const dem_grid& g;
    orc_state& s;
    bool f64;
    bool canonical;
    std::vector<int32_t> bond_snapshot;  // bond index words at the start of the Sim pass

    Slot* slot(int32_t j) { return reinterpret_cast<Slot*>(s.contacts) + j; }
    float px(int i) const { return s.pos[2 * i]; }
    float py(int i) const { return s.pos[2 * i + 1]; }
    float r(int i) const { return s.radii[i]; }
    const dem_material& mat(int i) const {
        int m = s.material_pointers[i];
        if (m < 0 || m >= n_mats) m = 0;
        return mats[m];
    }
    int32_t bond_index(int32_t k) const { return (k >= 0 && k < s.n_bonds) ? s.bonds[3 * k] : 0; }
    int32_t bond_index_seen(int32_t k) const {
        return (k >= 0 && k < (int32_t)bond_snapshot.size()) ? bond_snapshot[k] : 0;
    }
    float bond_length(int32_t k) const { return (k >= 0 && k < s.n_bonds) ? bitsf(s.bonds[3 * k + 2]) : 0.0f; }
    float normal_limit(int32_t k) const {
        if (s.bond_strengths && k >= 0 && k < s.n_bonds) return s.bond_strengths[2 * k];
        return u.bond_normal_strength;
    }
    float shear_limit(int32_t k) const {
        if (s.bond_strengths && k >= 0 && k < s.n_bonds) return s.bond_strengths[2 * k + 1];
        return u.bond_shear_strength;
    }
    // With `defer_breaks` (the canonical Sim pass run in parallel) a break is queued on
    // the breaking particle and written after the pass in particle-id order, the order
    // in which the serial pass writes them.
    bool defer_breaks = false;
    std::vector<std::vector<std::pair<int32_t, int32_t>>> breaks;   // per particle {bond, code}
    static int32_t& owner() { static thread_local int32_t id = -1; return id; }
    void break_bond(int32_t k, int32_t code) {
        if (!(k >= 0 && k < s.n_bonds)) return;
        if (defer_breaks) breaks[owner()].push_back({k, code});
        else s.bonds[3 * k] = code;
    }

    // distance / distance2, 2D_Simulation.wgsl:264-284 ($ F32 and $ F64 variants)
    float distance(int a, int b) const {
        if (f64) {
            double dx = (double)px(a) - (double)px(b), dy = (double)py(a) - (double)py(b);
            return (float)(lend(dx, dy) - ((double)r(a) + (double)r(b)));
        }
        return lenf(px(a) - px(b), py(a) - py(b)) - (r(a) + r(b));
    }
    float distance2(int a, int b, int32_t bonded) const {
        float L = r(a) + r(b);
        if (bonded >= 0) L = bond_length(bonded);
        if (f64) {
            double dx = (double)px(a) - (double)px(b), dy = (double)py(a) - (double)py(b);
            return (float)(lend(dx, dy) - (double)L);
        }
        return lenf(px(a) - px(b), py(a) - py(b)) - L;
    }
    void add_data(int a, float v0, float v1, float v2) {
        s.data[DATA_SIZE * a + 0] += v0;
        s.data[DATA_SIZE * a + 1] += v1;
        s.data[DATA_SIZE * a + 2] += v2;
    }
    // relative tangential displacement, 2D_Simulation.wgsl:352-360
    float rel_tangent(int a, int b, float tx, float ty) const {
        float dax = s.del_pos[2 * a], day = s.del_pos[2 * a + 1];
        float dbx = s.del_pos[2 * b], dby = s.del_pos[2 * b + 1];
        float dra = s.del_rot[a] * r(a);
        float drb = s.del_rot[b] * r(b);
        float rtx = dbx - dax, rty = dby - day;
        float rr = drb + dra;
        return (rtx * tx + rty * ty) + rr;
    }

    // linear_model, 2D_Simulation.wgsl:338-373
    // With cn2 != 0: the normal dashpot of the configurations' spring-dashpot law, an
    // extension (the reference has none): 2 zeta sqrt(m_eff kn) times the approach speed,
    // the normal component of del_pos[b] - del_pos[a] over dT.
    V3 linear_model(int a, int b, Slot& sl, int32_t bonded) {
        float nd = wmin(0.0f, distance2(a, b, bonded));
        if (nd == 0.0f) return {0.0f, 0.0f, 0.0f};
        const dem_material& ma = mat(a);
        const dem_material& mb = mat(b);
        float kn = 1.0f / (1.0f / ma.normal_stiffness + 1.0f / mb.normal_stiffness);
        float ks = 1.0f / (1.0f / ma.shear_stiffness + 1.0f / mb.shear_stiffness);
        float nf = -nd * kn;
        float dx = px(a) - px(b), dy = py(a) - py(b);
        float l = lenf(dx, dy);
        float nx = dx / l, ny = dy / l;
        if (cn2 != 0.0f) {
            float m_a = ma.density * PI_F * r(a) * r(a);
            float m_b = mb.density * PI_F * r(b) * r(b);
            float me = m_a * m_b / (m_a + m_b);
            float cn = cn2 * std::sqrt(me * kn);
            float ax = s.del_pos[2 * b] - s.del_pos[2 * a], ay = s.del_pos[2 * b + 1] - s.del_pos[2 * a + 1];
            float va = (ax * nx + ay * ny) / u.dT;
            nf = nf + cn * va;
        }
        float tx = -ny, ty = nx;
        float rt = rel_tangent(a, b, tx, ty);
        float lim = std::fabs(nf) * u.friction_coefficient;
        // WGSL clamp (min(max(.))) written with comparisons, so that the sign of a zero
        // result (mu = 0) is defined
        const float t = sl.ft + rt * ks;
        sl.ft = t < -lim ? -lim : (t > lim ? lim : t);
        float m = -r(a) * sl.ft;
        V3 f{nx * nf + tx * sl.ft, ny * nf + ty * sl.ft, m};
        add_data(a, nf, sl.ft, m);
        return f;
    }

    // linear_parallel_bonds, 2D_Simulation.wgsl:286-336
    V3 linear_parallel_bonds(int a, int b, Slot& sl, int32_t bonded) {
        float kn = u.bond_normal_stiffness;
        float ks = u.bond_shear_stiffness;
        float R = 0.5f * wmin(r(a), r(b));
        float A = 2.0f * R * 1.0f;
        float I = 2.0f / 3.0f * R * R * R;
        float nd = distance2(a, b, bonded);
        float nf = -kn * nd * A;
        float dx = px(a) - px(b), dy = py(a) - py(b);
        float l = lenf(dx, dy);
        float nx = dx / l, ny = dy / l;
        float tx = -ny, ty = nx;
        float rt = rel_tangent(a, b, tx, ty);
        sl.fbt = sl.fbt + rt * ks * A;
        float dtb = s.del_rot[a] - s.del_rot[b];
        sl.tb = sl.tb + dtb;
        float m = -kn * I * sl.tb - sl.fbt * r(a);
        float fx = nx * nf + tx * sl.fbt, fy = ny * nf + ty * sl.fbt;
        float nl = normal_limit(bonded), sh = shear_limit(bonded);
        if (u.bonds_tear == 1 &&
            (nf / A - u.moment_contribution_factor * (std::fabs(-kn * I * sl.tb) * R) / I < -nl ||
             std::fabs(sl.fbt) / A > sh)) {
            int32_t code = -10 + sign_i(nd);
            break_bond(bonded, code);
            sl.bonded = code;
            nf = 0.0f;
            sl.fbt = 0.0f;
            fx = 0.0f; fy = 0.0f;
            m = 0.0f;
        }
        add_data(a, -nf, sl.fbt, m);
        V3 lm = linear_model(a, b, sl, bonded);
        return {fx + lm.x, fy + lm.y, m + lm.z};
    }

    // linear_contact_bonds, 2D_Simulation.wgsl:375-417
    V3 linear_contact_bonds(int a, int b, Slot& sl, int32_t bonded) {
        float nd = distance(a, b);
        float kn = 1.0f / (1.0f / u.bond_normal_stiffness + 1.0f / u.bond_normal_stiffness);
        float ks = 1.0f / (1.0f / u.bond_shear_stiffness + 1.0f / u.bond_shear_stiffness);
        float nf = -nd * kn;
        float dx = px(a) - px(b), dy = py(a) - py(b);
        float l = lenf(dx, dy);
        float nx = dx / l, ny = dy / l;
        float tx = -ny, ty = nx;
        float rt = rel_tangent(a, b, tx, ty);
        sl.ft = sl.ft + rt * ks;
        float cd = u.contact_damping;
        float fx = cd * (nx * nf + tx * sl.ft), fy = cd * (ny * nf + ty * sl.ft);
        float m = -r(a) * sl.ft;
        float sh = shear_limit(bonded), nl = normal_limit(bonded);
        if (u.bonds_tear == 1 && (nf < -nl || std::fabs(sl.ft) > sh)) {
            int32_t code = -10 + sign_i(nd);
            break_bond(bonded, code);
            sl.bonded = code;
            nf = 0.0f;
            sl.fbt = 0.0f;
            fx = 0.0f; fy = 0.0f;
            m = 0.0f;
        }
        s.data[DATA_SIZE * a + 0] = -nf;
        s.data[DATA_SIZE * a + 1] = sl.ft;
        s.data[DATA_SIZE * a + 2] = m;
        return {fx, fy, m};
    }

    // normal_bonds, 2D_Simulation.wgsl:419-432
    V3 normal_bonds(int a, int b, Slot& sl, int32_t bonded) {
        float disp = distance(a, b);
        float sf = u.bond_normal_stiffness * disp;
        float dx = px(b) - px(a), dy = py(b) - py(a);
        float l = lenf(dx, dy);
        float fx = 0.0f + sf * (dx / l), fy = 0.0f + sf * (dy / l);
        if (u.bonds_tear == 1 && sf >= normal_limit(bonded)) {
            int32_t code = -10 + sign_i(disp);
            break_bond(bonded, code);
            sl.bonded = code;
            fx = 0.0f; fy = 0.0f;
        }
        V3 lm = linear_model(a, b, sl, bonded);
        return {lm.x + fx, lm.y + fy, lm.z + 0.0f};
    }

    // One slot of the force loop, 2D_Simulation.wgsl:230-255.
    V3 slot_force(Slot& sl) {
        int a = sl.a, b = sl.b;
        int32_t bonded = sl.bonded;
        if (bonded >= 0) {
            int32_t idx = bond_index_seen(bonded);
            if (idx < 0) { sl.bonded = idx; bonded = idx; }
        }
        if (bonded < 0 || u.bonds == 0) return linear_model(a, b, sl, -1);
        if (u.bonds == 1) return normal_bonds(a, b, sl, bonded);
        if (u.bonds == 2) return linear_contact_bonds(a, b, sl, bonded);
        if (u.bonds == 3) return linear_parallel_bonds(a, b, sl, bonded);
        return {0.0f, 0.0f, 0.0f};
    }

    // store_forces, 2D_Simulation.wgsl:434-478
    void store_forces(int id, int mat_id, float nfx, float nfy, float nm) {
        if (mat_id < 0 || mat_id >= n_mats) mat_id = 0;
        float density = mats[mat_id].density;
        float rr = r(id);
        float mass = density * PI_F * rr * rr;
        float inertia = 0.5f * mass * rr * rr;
        float fx = nfx + s.forces[6 * id], fy = nfy + s.forces[6 * id + 1];
        float m = nm + s.forces[6 * id + 2];
        if (u.gravity == 1 && u.planet_mode == 1) {
            float cx = 0.0f, cy = 0.0f;
            if (u.mouse_gravity == 1) { cx = u.gravity_x; cy = u.gravity_y; }
            float ddx = cx - px(id), ddy = cy - py(id);
            if (ddx != 0.0f || ddy != 0.0f) {
                float l = lenf(ddx, ddy);
                fx = fx + ddx / l * G_F * u.gravity_acc * mass;
                fy = fy + ddy / l * G_F * u.gravity_acc * mass;
            }
        } else if (u.gravity == 1) {
            float ga = G_F * u.gravity_acc;
            fx = fx + 0.0f;
            fy = fy + -ga * mass;
        }
        float ax = fx / mass, ay = fy / mass;
        s.acc[2 * id] = ax;
        s.acc[2 * id + 1] = ay;
        s.rot_acc[id] = m / inertia;
        const int32_t* fix = s.fixity + 6 * id;
        float dt = u.dT;
        if (fix[0] == 0) s.vel[2 * id] = s.vel[2 * id] + 0.5f * s.acc[2 * id] * dt;
        if (fix[1] == 0) s.vel[2 * id + 1] = s.vel[2 * id + 1] + 0.5f * s.acc[2 * id + 1] * dt;
        if (fix[2] == 0) s.rot_vel[id] = s.rot_vel[id] + 0.5f * s.rot_acc[id] * dt;
    }

    // ---- LOM pass: 2D_LOM.wgsl:80-201 ----
    void lom(int id) {
        if (s.radii[id] == 0.0f || id >= u.particles) return;
        float dt = u.dT;
        float ivx = s.vel[2 * id], ivy = s.vel[2 * id + 1];
        float iw = s.rot_vel[id];
        const int32_t* fix = s.fixity + 6 * id;
        if (fix[0] == 0) ivx += s.acc[2 * id] * dt * 0.5f;
        if (fix[1] == 0) ivy += s.acc[2 * id + 1] * dt * 0.5f;
        if (fix[2] == 0) iw += s.rot_acc[id] * dt * 0.5f;
        s.data[DATA_SIZE * id + 4] = ivx;
        s.data[DATA_SIZE * id + 5] = ivy;
        s.data[DATA_SIZE * id + 6] = iw;
        if (s.coll_cont[3] == 0) {
            s.del_pos[2 * id] = ivx * dt;
            s.del_pos[2 * id + 1] = ivy * dt;
            s.del_rot[id] = iw * dt;
        }
        float dpx = s.del_pos[2 * id], dpy = s.del_pos[2 * id + 1];
        float npx = s.pos[2 * id] + dpx, npy = s.pos[2 * id + 1] + dpy;
        float rr = s.radii[id];
        if (u.round_bounds == 0) {
            float yH = u.vert_bound / 2.0f;
            float xW = u.hor_bound / 2.0f;
            if (fix[1] != 1) {
                if (npy - rr < -yH) {
                    ivy = -ivy * 0.5f;
                    s.pos[2 * id + 1] += -yH - (npy - rr);
                } else if (npy + rr > yH) {
                    ivy = -ivy * 0.5f;
                    s.pos[2 * id + 1] -= (npy + rr) - yH;
                }
            }
            if (fix[0] != 1) {
                if (npx - rr < -xW) {
                    ivx = -ivx * 0.5f;
                    s.pos[2 * id] += -xW - (npx - rr);
                } else if (npx + rr > xW) {
                    ivx = -ivx * 0.5f;
                    s.pos[2 * id] -= (npx + rr) - xW;
                }
            }
        } else if (lenf(npx, npy) + rr > u.bound_radius) {
            float l = lenf(npx, npy);
            float nx = npx / l, ny = npy / l;
            float dd = dpx * nx + dpy * ny;
            float dcx = dd * nx, dcy = dd * ny;
            float vv = ivx * nx + ivy * ny;
            float cvx = vv * nx, cvy = vv * ny;
            if (fix[0] != 1) {
                s.pos[2 * id] = nx * (u.bound_radius - rr) - dcx;
                ivx = -cvx * 0.5f;
            }
            if (fix[1] != 1) {
                s.pos[2 * id + 1] = ny * (u.bound_radius - rr) - dcy;
                ivy = -cvy * 0.5f;
            }
        }
        s.pos[2 * id] += dpx;
        s.pos[2 * id + 1] += dpy;
        s.rot[id] += s.del_rot[id];
        s.vel[2 * id] = ivx;
        s.vel[2 * id + 1] = ivy;
        s.rot_vel[id] = iw;

        if (id == 0) {
            s.coll_cont[1] += 1;
            if (s.coll_cont[1] >= u.collision_interval) {
                s.coll_cont[0] = 1;
                s.coll_cont[1] = 0;
            } else {
                s.coll_cont[0] = 0;
            }
        }
        if (!canonical && s.coll_cont[0] == 1) bin(id);
    }

    void aabb_cells(int id, int& minx, int& maxx, int& miny, int& maxy) const {
        float cs = g.cell_size;
        float base_x = -cs * (float)g.w * 0.5f;
        float base_y = cs * (float)g.h * 0.5f;
        float l = px(id) - r(id), rgt = px(id) + r(id);
        float bot = py(id) - r(id), top = py(id) + r(id);
        minx = std::max(f2i((l - base_x) / cs), 0);
        maxx = std::min(f2i((rgt - base_x) / cs), g.w - 1);
        miny = std::max(f2i((base_y - top) / cs), 0);
        maxy = std::min(f2i((base_y - bot) / cs), g.h - 1);
    }

    // Broad phase binning, 2D_LOM.wgsl:165-200 (sequential: first toucher resets).
    void bin(int id) {
        int minx, maxx, miny, maxy;
        aabb_cells(id, minx, maxx, miny, maxy);
        const int cap = g.cell_cap;
        for (int cy = miny; cy <= maxy; cy++)
            for (int cx = minx; cx <= maxx; cx++) {
                int64_t base = ((int64_t)cy * g.w + cx) * cap;
                if (base + 1 >= s.grid_len) continue;
                int32_t old = s.grid[base + 1];
                s.grid[base + 1] = s.coll_cont[2];
                if (old != s.coll_cont[2]) s.grid[base] = 0;
            }
        for (int cy = miny; cy <= maxy; cy++)
            for (int cx = minx; cx <= maxx; cx++) {
                int64_t base = ((int64_t)cy * g.w + cx) * cap;
                if (base + 1 >= s.grid_len) continue;
                int32_t pc = s.grid[base] + 1;
                s.grid[base] = pc;
                if (pc < cap - 1 && base + 1 + pc < s.grid_len) s.grid[base + 1 + pc] = id;
            }
    }

    // ---- Sim pass, reference slot rules: 2D_Simulation.wgsl:108-262 ----
    void sim_reference(int id) {
        s.data[DATA_SIZE * id + 0] = 0.0f;
        s.data[DATA_SIZE * id + 1] = 0.0f;
        s.data[DATA_SIZE * id + 2] = 0.0f;
        if (s.radii[id] == 0.0f || id >= u.particles) return;
        int mat_id = s.material_pointers[id];
        float nfx = 0.0f, nfy = 0.0f, nm = 0.0f;
        const int K = s.slots;
        Slot* tab = slot(id * K);
        if (u.collisions == 1 && s.coll_cont[0] == 1) {
            int32_t coll[REF_MAX_CONTACTS];
            unsigned count = 0;
            int minx, maxx, miny, maxy;
            aabb_cells(id, minx, maxx, miny, maxy);
            for (int cy = miny; cy <= maxy && count < REF_MAX_CONTACTS; cy++) {
                for (int cx = minx; cx <= maxx && count < REF_MAX_CONTACTS; cx++) {
                    int64_t base = ((int64_t)cy * g.w + cx) * g.cell_cap;
                    if (base >= s.grid_len) continue;
                    int32_t nb = s.grid[base];
                    if (nb > 1) {
                        for (int i = 2; i < nb + 2; i++) {
                            if (base + i >= s.grid_len) break;
                            int32_t b = s.grid[base + i];
                            if (b < 0 || b >= s.n) continue;  // stale / overflow word
                            if ((uint32_t)b != (uint32_t)id) {
                                bool already = false;
                                for (unsigned j = 0; j < count; j++)
                                    if (coll[j] == b) already = true;
                                if (!already && lenf(px(b) - px(id), py(b) - py(id)) < (r(b) + r(id))) {
                                    coll[count] = b;
                                    count += 1;
                                    if (count == REF_MAX_CONTACTS) break;
                                }
                            }
                        }
                    }
                }
            }
            // delete contacts that no longer exist, :185-204
            for (int j = 0; j < K; j++) {
                Slot& sl = tab[j];
                if (sl.b == -1) continue;
                bool found = false;
                for (unsigned i = 0; i < count; i++)
                    if (sl.b == coll[i]) found = true;
                if (!found && (sl.bonded < 0 || u.bonds == 0)) {
                    sl.a = -1;
                    sl.b = -1;
                } else if (!found && sl.bonded > 0 && u.bonds <= 1) {
                    sl.ft = 0.0f;
                }
            }
            // create new contacts in the last empty slot, :206-227
            for (unsigned i = 0; i < count; i++) {
                int existing = -1, empty = -1;
                for (int j = 0; j < K; j++) {
                    if (tab[j].b == coll[i]) { existing = j; break; }
                    else if (tab[j].b == -1) empty = j;
                }
                if (existing == -1 && empty == -1) continue;
                if (existing == -1) {
                    tab[empty].a = id;
                    tab[empty].b = coll[i];
                    tab[empty].ft = 0.0f;
                }
            }
        }
        for (int j = 0; j < K; j++) {
            if (tab[j].b == -1) continue;
            V3 f = slot_force(tab[j]);
            nfx += f.x; nfy += f.y; nm += f.z;
        }
        store_forces(id, mat_id, nfx, nfy, nm);
        if (id == 0) {
            s.coll_cont[2] += 1;
            s.coll_cont[3] = 0;
        }
    }

    // ---- Sim pass, canonical slot order (the HIP path's semantics) ----
    std::vector<std::vector<int32_t>> found;  // per particle, ascending partner ids

    void find_all_pairs() { cell_list_pairs(u.particles, s.pos, s.radii, found); }

    void sim_canonical(int id) {
        owner() = id;
        s.data[DATA_SIZE * id + 0] = 0.0f;
        s.data[DATA_SIZE * id + 1] = 0.0f;
        s.data[DATA_SIZE * id + 2] = 0.0f;
        if (s.radii[id] == 0.0f || id >= u.particles) return;
        int mat_id = s.material_pointers[id];
        float nfx = 0.0f, nfy = 0.0f, nm = 0.0f;
        const int K = s.slots;
        Slot* tab = slot(id * K);
        if (u.collisions == 1 && s.coll_cont[0] == 1) {
            const std::vector<int32_t>& f = found[id];
            for (int j = 0; j < K; j++) {
                Slot& sl = tab[j];
                if (sl.b == -1) continue;
                bool fnd = std::binary_search(f.begin(), f.end(), sl.b);
                if (!fnd && (sl.bonded < 0 || u.bonds == 0)) {
                    std::memset(&sl, 0xff, sizeof(Slot));
                } else if (!fnd && sl.bonded > 0 && u.bonds <= 1) {
                    sl.ft = 0.0f;
                }
            }
            // new partners fill free slots in ascending id while the particle holds fewer
            // than 14 (2D_Simulation.wgsl:125: contacts and bonds share the 14 slots;
            // existing slots are never evicted)
            int occupied = 0;
            for (int j = 0; j < K; j++) occupied += tab[j].b != -1;
            for (int32_t b : f) {
                bool present = false;
                int empty = -1;
                for (int j = 0; j < K; j++) {
                    if (tab[j].b == b) { present = true; break; }
                    if (tab[j].b == -1 && empty < 0) empty = j;
                }
                if (present || empty < 0 || occupied >= REF_MAX_CONTACTS) continue;
                occupied++;
                tab[empty].a = id;
                tab[empty].b = b;
                tab[empty].ft = 0.0f;
                tab[empty].fbt = bitsf(-1);
                tab[empty].tb = bitsf(-1);
                tab[empty].bonded = -1;
            }
        }
        canonicalize(tab, K);
        for (int j = 0; j < K; j++) {
            if (tab[j].b == -1) continue;
            V3 fc = slot_force(tab[j]);
            nfx += fc.x; nfy += fc.y; nm += fc.z;
        }
        store_forces(id, mat_id, nfx, nfy, nm);
        if (id == 0) {
            s.coll_cont[2] += 1;
            s.coll_cont[3] = 0;
        }
    }

    float cn2 = 0.0f;   // dashpot factor 2 zeta (0: the reference's law)

    static void canonicalize(Slot* tab, int K) {
        std::stable_sort(tab, tab + K, [](const Slot& x, const Slot& y) {
            bool ex = x.b == -1, ey = y.b == -1;
            if (ex != ey) return ey;  // occupied first
            if (ex) return false;
            return x.b < y.b;
        });
        for (int j = 0; j < K; j++)
            if (tab[j].b == -1) std::memset(&tab[j], 0xff, sizeof(Slot));
    }

    void step() {
        const int n = u.particles;
        if (!canonical) {
            for (int id = 0; id < n; id++) lom(id);
        } else {
            // invocation 0 first (it owns the coll_cont counters the others read), then
            // the rest, each of which touches only its own particle
            if (n > 0) lom(0);
#pragma omp parallel for schedule(static)
            for (int id = 1; id < n; id++) lom(id);
        }
        bond_snapshot.resize(s.n_bonds);
        for (int k = 0; k < s.n_bonds; k++) bond_snapshot[k] = s.bonds[3 * k];
        if (canonical) {
            if (u.collisions == 1 && s.coll_cont[0] == 1) find_all_pairs();
            // every invocation writes only its own particle's slots, data and
            // accelerations (and invocation 0 the coll_cont words no other one reads in
            // this pass); the shared bond words are queued and written in id order
            breaks.resize(n);
            defer_breaks = true;
#pragma omp parallel for schedule(dynamic, 256)
            for (int id = 0; id < n; id++) sim_canonical(id);
            defer_breaks = false;
            for (int id = 0; id < n; id++) {
                for (const auto& b : breaks[id]) s.bonds[3 * b.first] = b.second;
                breaks[id].clear();
            }
        } else {
            for (int id = 0; id < n; id++) sim_reference(id);
        }
    }
};

// 2 zeta for restitution e (zeta = -ln e / sqrt(pi^2 + ln^2 e), f64, rounded once); 0
// outside (0, 1): the linear spring-dashpot whose binary collisions rebound at e.
float dashpot_factor(float e) {
    if (!(e > 0.0f && e < 1.0f)) return 0.0f;
    const double le = std::log((double)e);
    return (float)(2.0 * (-le / std::sqrt(M_PI * M_PI + le * le)));
}

}  // namespace

extern "C" {

enum { ORC_MODE_REFERENCE = 0, ORC_MODE_CANONICAL = 1 };

// Runs `nsteps` time steps on the arrays in `s` (in place).
// restitution: the dashpot extension (0: none, the reference's law).
int orc_run(const dem_uniform* u, const dem_material* mats, int32_t n_mats, const dem_grid* g,
            orc_state* s, int32_t mode, int32_t use_f64, int32_t nsteps, float restitution) {
    if (!u || !mats || n_mats <= 0 || !g || !s || nsteps < 0) return -1;
    if (u->particles > s->n || u->particles < 0) return -1;
    if (mode == ORC_MODE_REFERENCE && (!s->grid || s->slots < 1)) return -1;
    Ctx c{*u, mats, n_mats, *g, *s, use_f64 != 0, mode == ORC_MODE_CANONICAL, {}, {}};
    c.cn2 = dashpot_factor(restitution);
    if (c.canonical)
        for (int i = 0; i < u->particles; i++) Ctx::canonicalize(c.slot(i * s->slots), s->slots);
    for (int k = 0; k < nsteps; k++) c.step();
    return 0;
}

// Sorts every particle's occupied slots by partner id and clears empty slots to
// -1 words: the canonical table layout the HIP path downloads.
// Threads of the oracle's OpenMP loops (the canonical mode's LOM, pair search and Sim
// pass); results do not depend on it.  Returns the previous count.
int orc_set_threads(int32_t threads) {
    const int prev = omp_get_max_threads();
    if (threads > 0) omp_set_num_threads(threads);
    return prev;
}

int orc_canonicalize(float* contacts, int32_t n, int32_t slots) {
    Slot* t = reinterpret_cast<Slot*>(contacts);
    for (int i = 0; i < n; i++) Ctx::canonicalize(t + (int64_t)i * slots, slots);
    return 0;
}

// Reference grid: grid_capacity, wgpu_prog.rs:698-717 (width/height = 2x the box).
int orc_grid_capacity(float hor_bound, float vert_bound, float max_radius, float min_radius,
                      int32_t variable_rad, int32_t particles, dem_grid* out) {
    float width = hor_bound * 2.0f;
    float height = vert_bound * 2.0f;
    float max_rad = max_radius * 2.0f;
    float min_rad = variable_rad ? min_radius : max_radius;
    out->w = (int32_t)std::ceil(width / max_rad);
    out->h = (int32_t)std::ceil(height / max_rad);
    int32_t cap = (int32_t)std::ceil(std::pow(max_rad / min_rad + 1.0f, 2.0f));
    out->cell_cap = std::min(cap, particles) + 2;
    out->cell_size = max_rad;
    return 0;
}

// Stage 1 closed form: centre cell of each particle in the dem_grid convention
// (grid centred on the origin, rows from the top, as GridInfo in 2D_LOM.wgsl:167-178):
// cx = trunc((x - base_x) * (1/cs)) and cy = trunc((base_y - y) * (1/cs)), clamped to
// the grid, with 1/cs rounded to f32.  Inactive particles (r == 0 or id >= n_active)
// get key w*h.
int orc_cell_keys(int32_t n, int32_t n_active, const float* pos, const float* radii, const dem_grid* g,
                  int32_t* keys) {
    float cs = g->cell_size;
    float base_x = -cs * (float)g->w * 0.5f;
    float base_y = cs * (float)g->h * 0.5f;
    float inv = 1.0f / cs;
    for (int i = 0; i < n; i++) {
        if (i >= n_active || radii[i] == 0.0f) { keys[i] = g->w * g->h; continue; }
        int cx = std::min(std::max(f2i((pos[2 * i] - base_x) * inv), 0), g->w - 1);
        int cy = std::min(std::max(f2i((base_y - pos[2 * i + 1]) * inv), 0), g->h - 1);
        keys[i] = cy * g->w + cx;
    }
    return 0;
}

// Stages 2-3 closed form: stable sort of particle ids by key (ties by id) and the
// exclusive prefix sum of the per-cell counts (ncells + 1 entries).
int orc_sort_cells(int32_t n, const int32_t* keys, int32_t ncells, int32_t* sorted_ids, int32_t* cell_start) {
    std::vector<int64_t> cnt(ncells + 2, 0);
    for (int i = 0; i < n; i++) {
        int k = keys[i];
        if (k < 0 || k > ncells) return -1;
        cnt[k + 1]++;
    }
    for (int c = 0; c <= ncells; c++) cnt[c + 1] += cnt[c];
    if (cell_start)
        for (int c = 0; c <= ncells; c++) cell_start[c] = (int32_t)cnt[c];
    std::vector<int64_t> fill(cnt.begin(), cnt.end());
    if (sorted_ids)
        for (int i = 0; i < n; i++) sorted_ids[fill[keys[i]]++] = i;
    return 0;
}

// Stage 4 brute force: for every active particle a, every active b != a with
// sqrt(dx*dx + dy*dy) < r_b + r_a (the narrow-phase predicate of
// 2D_Simulation.wgsl:159), ascending.  O(N^2).  Returns the total pair count or
// -1 if some particle has more than max_per partners.
int64_t orc_contact_pairs(int32_t n, const float* pos, const float* radii, int32_t max_per, int32_t* counts,
                          int32_t* partners) {
    int64_t total = 0;
    bool over = false;
#pragma omp parallel for reduction(+ : total) schedule(dynamic, 64)
    for (int a = 0; a < n; a++) {
        int c = 0;
        if (radii[a] != 0.0f) {
            for (int b = 0; b < n; b++) {
                if (b == a || radii[b] == 0.0f) continue;
                float dx = pos[2 * b] - pos[2 * a], dy = pos[2 * b + 1] - pos[2 * a + 1];
                if (lenf(dx, dy) < radii[b] + radii[a]) {
                    if (c < max_per) partners[(int64_t)a * max_per + c] = b;
                    else over = true;
                    c++;
                }
            }
        }
        for (int k = c; k < max_per; k++) partners[(int64_t)a * max_per + k] = -1;
        counts[a] = c;
        total += c;
    }
    return over ? -1 : total;
}

// Stage 4 on a cell list: the pair set of orc_contact_pairs in O(N) (cell_list_pairs),
// for inputs too large for the brute force.  Same outputs and return value.
int64_t orc_contact_pairs_grid(int32_t n, const float* pos, const float* radii, int32_t max_per, int32_t* counts,
                               int32_t* partners) {
    std::vector<std::vector<int32_t>> found;
    cell_list_pairs(n, pos, radii, found);
    int64_t total = 0;
    bool over = false;
    for (int a = 0; a < n; a++) {
        const int c = (int)found[a].size();
        for (int k = 0; k < max_per; k++) partners[(int64_t)a * max_per + k] = k < c ? found[a][k] : -1;
        if (c > max_per) over = true;
        counts[a] = c;
        total += c;
    }
    return over ? -1 : total;
}

// State::regen_bonds restated (state.rs:291-367).  O(N^2), in (i, j) order.  The
// bond test is dist < (r_i + r_j) * factor + gap in f32: the reference's is factor
// 1.05, gap 0 (state.rs:302); a bonded assembly bonds every pair closer than a gap
// (factor 1).  bonds_out: 3 words per bond; contacts_out: slots*6 words per particle
// (all -1 words on entry is not required: the function initialises them).  Returns
// the bond count, or -1 if bond_cap is too small.
int64_t orc_regen_bonds(int32_t n, const float* pos, const float* radii, int32_t max_bonds, int32_t slots,
                        int32_t* bonds_out, int64_t bond_cap, float* contacts_out, float factor, float gap) {
    Slot* tab = reinterpret_cast<Slot*>(contacts_out);
    std::memset(contacts_out, 0xff, sizeof(Slot) * (size_t)n * slots);
    int64_t bond_index = 0;
    for (int i = 0; i < n; i++) {
        int col_num = 0;
        for (int j = i; j < n; j++) {
            if (i == j) continue;
            float dx = pos[2 * j] - pos[2 * i], dy = pos[2 * j + 1] - pos[2 * i + 1];
            float dist = std::sqrt(dx * dx + dy * dy);
            float sum = radii[i] + radii[j];
            if (dist < sum * factor + gap) {
                if (col_num < max_bonds) {
                    if (bond_index >= bond_cap) return -1;
                    float mag = std::sqrt(dx * dx + dy * dy);
                    float nx = dx / mag, ny = dy / mag;
                    float angle = std::atan2(nx, ny);
                    bonds_out[3 * bond_index] = 1;
                    bonds_out[3 * bond_index + 1] = fbits(angle);
                    bonds_out[3 * bond_index + 2] = fbits(mag);
                    for (int k = 0; k < slots; k++) {
                        Slot& sl = tab[(int64_t)i * slots + k];
                        if (sl.a == -1) { sl = Slot{i, j, 0.0f, 0.0f, 0.0f, (int32_t)bond_index}; break; }
                    }
                    for (int k = 0; k < slots; k++) {
                        Slot& sl = tab[(int64_t)j * slots + k];
                        if (sl.a == -1) { sl = Slot{j, i, 0.0f, 0.0f, 0.0f, (int32_t)bond_index}; break; }
                    }
                    col_num++;
                    bond_index++;
                } else if (col_num == max_bonds) {
                    break;
                }
            }
        }
    }
    return bond_index;
}

// Kinetic energy Σ ½ m |v|² + ½ I ω² (m = ρπr², I = ½ m r²), f64 accumulation in id order.
double orc_kinetic_energy(int32_t n, const float* vel, const float* rot_vel, const float* radii,
                          const int32_t* material_pointers, const dem_material* mats, int32_t n_mats) {
    double e = 0.0;
    for (int i = 0; i < n; i++) {
        if (radii[i] == 0.0f) continue;
        int m = material_pointers[i];
        if (m < 0 || m >= n_mats) m = 0;
        double rr = radii[i];
        double mass = (double)mats[m].density * M_PI * rr * rr;
        double inertia = 0.5 * mass * rr * rr;
        double vx = vel[2 * i], vy = vel[2 * i + 1], w = rot_vel[i];
        e += 0.5 * mass * (vx * vx + vy * vy) + 0.5 * inertia * w * w;
    }
    return e;
}

// Reference-algorithm CPU timing: runs `nsteps` reference-mode steps with the
// per-particle loops split over `threads` OpenMP threads (threads == 1: the
// sequential restatement).  The parallel version follows the WGSL's own
// concurrency: LOM drift in parallel, then the AABB binning with atomic cell
// counters, then the Sim pass in parallel.  Returns seconds.
double orc_time_reference(const dem_uniform* u, const dem_material* mats, int32_t n_mats, const dem_grid* g,
                          orc_state* s, int32_t nsteps, int32_t threads, float restitution) {
    Ctx c{*u, mats, n_mats, *g, *s, false, false, {}, {}};
    c.cn2 = dashpot_factor(restitution);
    if (threads <= 1) {
        auto t0 = std::chrono::steady_clock::now();
        for (int k = 0; k < nsteps; k++) c.step();
        auto t1 = std::chrono::steady_clock::now();
        return std::chrono::duration<double>(t1 - t0).count();
    }
    omp_set_num_threads(threads);
    const int n = u->particles;
    auto t0 = std::chrono::steady_clock::now();
    for (int k = 0; k < nsteps; k++) {
        // LOM: invocation 0 first (it owns the coll_cont counters), the rest in parallel.
        Ctx c0{*u, mats, n_mats, *g, *s, false, true, {}, {}};  // canonical flag: skip binning inside lom()
        c0.cn2 = c.cn2;
        if (n > 0) c0.lom(0);
#pragma omp parallel for schedule(static)
        for (int id = 1; id < n; id++) c0.lom(id);
        if (s->coll_cont[0] == 1) {
            const int cap = g->cell_cap;
#pragma omp parallel for schedule(static)
            for (int id = 0; id < n; id++) {
                if (s->radii[id] == 0.0f) continue;
                int minx, maxx, miny, maxy;
                c.aabb_cells(id, minx, maxx, miny, maxy);
                for (int cy = miny; cy <= maxy; cy++)
                    for (int cx = minx; cx <= maxx; cx++) {
                        int64_t base = ((int64_t)cy * g->w + cx) * cap;
                        if (base + 1 >= s->grid_len) continue;
                        int32_t old = __atomic_exchange_n(&s->grid[base + 1], s->coll_cont[2], __ATOMIC_RELAXED);
                        if (old != s->coll_cont[2]) __atomic_store_n(&s->grid[base], 0, __ATOMIC_RELAXED);
                    }
            }
#pragma omp parallel for schedule(static)
            for (int id = 0; id < n; id++) {
                if (s->radii[id] == 0.0f) continue;
                int minx, maxx, miny, maxy;
                c.aabb_cells(id, minx, maxx, miny, maxy);
                for (int cy = miny; cy <= maxy; cy++)
                    for (int cx = minx; cx <= maxx; cx++) {
                        int64_t base = ((int64_t)cy * g->w + cx) * cap;
                        if (base + 1 >= s->grid_len) continue;
                        int32_t pc = __atomic_add_fetch(&s->grid[base], 1, __ATOMIC_RELAXED);
                        if (pc < cap - 1 && base + 1 + pc < s->grid_len) s->grid[base + 1 + pc] = id;
                    }
            }
        }
        c.bond_snapshot.resize(s->n_bonds);
        for (int b = 0; b < s->n_bonds; b++) c.bond_snapshot[b] = s->bonds[3 * b];
#pragma omp parallel for schedule(static)
        for (int id = 0; id < n; id++) c.sim_reference(id);
    }
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

}  // extern "C"
