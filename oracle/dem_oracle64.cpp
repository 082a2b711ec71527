// dem_oracle64.cpp - the oracle's time step in f64 throughout ("ref64").  TEST
// INFRASTRUCTURE ONLY: loaded by tests/ as a checker; the product path never links it.
//
// The same step as the f32 oracle's canonical mode (dem_oracle.cpp), restated with every
// quantity in double: positions, velocities, displacements, forces, tangential and
// bond histories, constants (pi, 9.81).  It follows the same reference files:
//   LOM pass      src/shaders/physics/2D_LOM.wgsl:80-163 (kick, drift, rect and round
//                 walls, the collision-interval counter)
//   Sim pass      src/shaders/physics/2D_Simulation.wgsl:108-478 (narrow phase with the
//                 14-slot table, :125-227; linear_model :338-373; bond laws :286-432;
//                 store_forces :434-478)
// with the canonical contact order (every overlapping pair, slots sorted by partner id,
// new partners into free slots in ascending id while fewer than 14 are held) and the
// contact dashpot extension.  Its purpose is to state the f32 path's error against
// arithmetic without rounding to speak of: one-step positions and forces, and the
// kinetic-energy trace of a multi-step run.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../include/dem.h"

namespace {

constexpr int SLOTS = 14;   // 2D_Simulation.wgsl:125
constexpr double G = 9.81;

struct Slot {
    int32_t b = -1;
    double ft = 0.0, fbt = 0.0, tb = 0.0;
    int32_t bonded = -1;
};

struct Sim {
    const dem_uniform& u;
    const dem_material* mats;
    int nm;
    int n;
    double dt, cn2;
    std::vector<double> x, y, vx, vy, ax, ay, th, w, al, r, ex, ey, em, dpx, dpy, dth, data;
    std::vector<int32_t> fix, mat;
    std::vector<int32_t> bidx;       // bond index words
    std::vector<double> blen, bsn, bss;
    std::vector<Slot> tab;           // SLOTS per particle
    int32_t cc[4] = {0, 0, 0, 0};
    std::vector<int32_t> snap;
    std::vector<std::vector<int32_t>> found;

    Sim(const dem_uniform& u_, const dem_material* m_, int nm_, int n_, double dt_)
        : u(u_), mats(m_), nm(nm_), n(n_), dt(dt_), cn2(0.0) {}

    const dem_material& M(int i) const {
        int m = mat[i];
        if (m < 0 || m >= nm) m = 0;
        return mats[m];
    }
    double mass(int i) const { return (double)M(i).density * M_PI * r[i] * r[i]; }

    void lom(int i) {
        if (r[i] == 0.0 || i >= u.particles) return;
        const int32_t* f = &fix[6 * i];
        double ivx = vx[i], ivy = vy[i], iw = w[i];
        if (f[0] == 0) ivx += ax[i] * dt * 0.5;
        if (f[1] == 0) ivy += ay[i] * dt * 0.5;
        if (f[2] == 0) iw += al[i] * dt * 0.5;
        data[7 * i + 4] = ivx;
        data[7 * i + 5] = ivy;
        data[7 * i + 6] = iw;
        if (cc[3] == 0) {
            dpx[i] = ivx * dt;
            dpy[i] = ivy * dt;
            dth[i] = iw * dt;
        }
        const double npx = x[i] + dpx[i], npy = y[i] + dpy[i], rr = r[i];
        if (u.round_bounds == 0) {
            const double yH = u.vert_bound / 2.0, xW = u.hor_bound / 2.0;
            if (f[1] != 1) {
                if (npy - rr < -yH) { ivy = -ivy * 0.5; y[i] += -yH - (npy - rr); }
                else if (npy + rr > yH) { ivy = -ivy * 0.5; y[i] -= (npy + rr) - yH; }
            }
            if (f[0] != 1) {
                if (npx - rr < -xW) { ivx = -ivx * 0.5; x[i] += -xW - (npx - rr); }
                else if (npx + rr > xW) { ivx = -ivx * 0.5; x[i] -= (npx + rr) - xW; }
            }
        } else if (std::hypot(npx, npy) + rr > u.bound_radius) {
            const double l = std::hypot(npx, npy), nx = npx / l, ny = npy / l;
            const double d = dpx[i] * nx + dpy[i] * ny, v = ivx * nx + ivy * ny;
            if (f[0] != 1) { x[i] = nx * (u.bound_radius - rr) - d * nx; ivx = -v * nx * 0.5; }
            if (f[1] != 1) { y[i] = ny * (u.bound_radius - rr) - d * ny; ivy = -v * ny * 0.5; }
        }
        x[i] += dpx[i];
        y[i] += dpy[i];
        th[i] += dth[i];
        vx[i] = ivx;
        vy[i] = ivy;
        w[i] = iw;
        if (i == 0) {
            cc[1] += 1;
            if (cc[1] >= u.collision_interval) { cc[0] = 1; cc[1] = 0; }
            else cc[0] = 0;
        }
    }

    void pairs() {
        // every overlapping pair over a cell list of side 2 r_max
        found.assign(n, {});
        double rmax = 0.0, x0 = INFINITY, y0 = INFINITY, x1 = -INFINITY, y1 = -INFINITY;
        std::vector<int> act;
        for (int i = 0; i < u.particles; i++)
            if (r[i] != 0.0 && std::isfinite(x[i]) && std::isfinite(y[i])) {
                act.push_back(i);
                rmax = std::max(rmax, r[i]);
                x0 = std::min(x0, x[i]); x1 = std::max(x1, x[i]);
                y0 = std::min(y0, y[i]); y1 = std::max(y1, y[i]);
            }
        if (act.empty()) return;
        const double cs = std::max(2.0 * rmax, 1e-30) * 1.000001;
        const int64_t gw = std::min<int64_t>((int64_t)((x1 - x0) / cs) + 1, 1 << 14);
        const int64_t gh = std::min<int64_t>((int64_t)((y1 - y0) / cs) + 1, 1 << 14);
        auto cell = [&](int i, int64_t& cx, int64_t& cy) {
            cx = std::min<int64_t>(std::max<int64_t>((int64_t)((x[i] - x0) / cs), 0), gw - 1);
            cy = std::min<int64_t>(std::max<int64_t>((int64_t)((y[i] - y0) / cs), 0), gh - 1);
        };
        std::vector<std::vector<int>> bin(gw * gh);
        for (int i : act) { int64_t cx, cy; cell(i, cx, cy); bin[cy * gw + cx].push_back(i); }
        for (int a : act) {
            int64_t cx, cy;
            cell(a, cx, cy);
            for (int64_t yy = std::max<int64_t>(cy - 1, 0); yy <= std::min<int64_t>(cy + 1, gh - 1); yy++)
                for (int64_t xx = std::max<int64_t>(cx - 1, 0); xx <= std::min<int64_t>(cx + 1, gw - 1); xx++)
                    for (int b : bin[yy * gw + xx])
                        if (b != a && std::hypot(x[b] - x[a], y[b] - y[a]) < r[b] + r[a]) found[a].push_back(b);
            std::sort(found[a].begin(), found[a].end());
        }
    }

    // relative tangential displacement (2D_Simulation.wgsl:352-360)
    double rel_t(int a, int b, double tx, double ty) const {
        return ((dpx[b] - dpx[a]) * tx + (dpy[b] - dpy[a]) * ty) + (dth[b] * r[b] + dth[a] * r[a]);
    }
    void add(int a, double p, double q, double s) { data[7 * a] += p; data[7 * a + 1] += q; data[7 * a + 2] += s; }

    void linear_model(int a, int b, Slot& sl, int32_t bonded, double& fx, double& fy, double& m) {
        const double L = bonded >= 0 ? blen[bonded] : r[a] + r[b];
        const double dx = x[a] - x[b], dy = y[a] - y[b], l = std::hypot(dx, dy);
        const double nd = std::min(0.0, l - L);
        if (nd == 0.0) return;
        const double kn = 1.0 / (1.0 / M(a).normal_stiffness + 1.0 / M(b).normal_stiffness);
        const double ks = 1.0 / (1.0 / M(a).shear_stiffness + 1.0 / M(b).shear_stiffness);
        double nf = -nd * kn;
        const double nx = dx / l, ny = dy / l, tx = -ny, ty = nx;
        if (cn2 != 0.0) {
            const double ma = mass(a), mb = mass(b);
            nf += cn2 * std::sqrt(ma * mb / (ma + mb) * kn) * (((dpx[b] - dpx[a]) * nx + (dpy[b] - dpy[a]) * ny) / dt);
        }
        const double lim = std::fabs(nf) * u.friction_coefficient;
        const double t = sl.ft + rel_t(a, b, tx, ty) * ks;
        sl.ft = t < -lim ? -lim : (t > lim ? lim : t);
        const double mm = -r[a] * sl.ft;
        fx += nx * nf + tx * sl.ft;
        fy += ny * nf + ty * sl.ft;
        m += mm;
        add(a, nf, sl.ft, mm);
    }

    double dist(int a, int b) const { return std::hypot(x[a] - x[b], y[a] - y[b]) - (r[a] + r[b]); }
    static int sgn(double v) { return v > 0 ? 1 : (v < 0 ? -1 : 0); }

    void bond(int a, int b, Slot& sl, int32_t k, double& fx, double& fy, double& m) {
        if (u.bonds == 3) {
            const double kn = u.bond_normal_stiffness, ks = u.bond_shear_stiffness;
            const double R = 0.5 * std::min(r[a], r[b]), A = 2.0 * R, I = 2.0 / 3.0 * R * R * R;
            const double dx = x[a] - x[b], dy = y[a] - y[b], l = std::hypot(dx, dy);
            const double nd = l - blen[k];
            double nf = -kn * nd * A;
            const double nx = dx / l, ny = dy / l, tx = -ny, ty = nx;
            sl.fbt += rel_t(a, b, tx, ty) * ks * A;
            sl.tb += dth[a] - dth[b];
            double mm = -kn * I * sl.tb - sl.fbt * r[a];
            double bx = nx * nf + tx * sl.fbt, by = ny * nf + ty * sl.fbt;
            if (u.bonds_tear == 1 &&
                (nf / A - u.moment_contribution_factor * (std::fabs(-kn * I * sl.tb) * R) / I < -bsn[k] ||
                 std::fabs(sl.fbt) / A > bss[k])) {
                bidx[k] = -10 + sgn(nd);
                sl.bonded = bidx[k];
                nf = 0.0; sl.fbt = 0.0; bx = by = mm = 0.0;
            }
            add(a, -nf, sl.fbt, mm);
            fx += bx; fy += by; m += mm;
            linear_model(a, b, sl, k, fx, fy, m);
        } else if (u.bonds == 2) {
            const double nd = dist(a, b);
            const double kn = 1.0 / (2.0 / u.bond_normal_stiffness), ks = 1.0 / (2.0 / u.bond_shear_stiffness);
            double nf = -nd * kn;
            const double dx = x[a] - x[b], dy = y[a] - y[b], l = std::hypot(dx, dy);
            const double nx = dx / l, ny = dy / l, tx = -ny, ty = nx;
            sl.ft += rel_t(a, b, tx, ty) * ks;
            const double cd = u.contact_damping;
            double bx = cd * (nx * nf + tx * sl.ft), by = cd * (ny * nf + ty * sl.ft), mm = -r[a] * sl.ft;
            if (u.bonds_tear == 1 && (nf < -bsn[k] || std::fabs(sl.ft) > bss[k])) {
                bidx[k] = -10 + sgn(nd);
                sl.bonded = bidx[k];
                nf = 0.0; sl.fbt = 0.0; bx = by = mm = 0.0;
            }
            data[7 * a] = -nf; data[7 * a + 1] = sl.ft; data[7 * a + 2] = mm;
            fx += bx; fy += by; m += mm;
        } else if (u.bonds == 1) {
            const double disp = dist(a, b), sf = u.bond_normal_stiffness * disp;
            const double dx = x[b] - x[a], dy = y[b] - y[a], l = std::hypot(dx, dy);
            double bx = sf * (dx / l), by = sf * (dy / l);
            if (u.bonds_tear == 1 && sf >= bsn[k]) {
                bidx[k] = -10 + sgn(disp);
                sl.bonded = bidx[k];
                bx = by = 0.0;
            }
            double lx = 0, ly = 0, lm = 0;
            linear_model(a, b, sl, k, lx, ly, lm);
            fx += lx + bx; fy += ly + by; m += lm;
        }
    }

    void sim(int id) {
        data[7 * id] = data[7 * id + 1] = data[7 * id + 2] = 0.0;
        if (r[id] == 0.0 || id >= u.particles) return;
        Slot* t = &tab[(size_t)id * SLOTS];
        if (u.collisions == 1 && cc[0] == 1) {
            const std::vector<int32_t>& f = found[id];
            for (int j = 0; j < SLOTS; j++) {
                Slot& sl = t[j];
                if (sl.b == -1) continue;
                const bool fnd = std::binary_search(f.begin(), f.end(), sl.b);
                if (!fnd && (sl.bonded < 0 || u.bonds == 0)) sl = Slot();
                else if (!fnd && sl.bonded > 0 && u.bonds <= 1) sl.ft = 0.0;
            }
            int occ = 0;
            for (int j = 0; j < SLOTS; j++) occ += t[j].b != -1;
            for (int32_t b : f) {
                bool present = false;
                int empty = -1;
                for (int j = 0; j < SLOTS; j++) {
                    if (t[j].b == b) { present = true; break; }
                    if (t[j].b == -1 && empty < 0) empty = j;
                }
                if (present || empty < 0 || occ >= SLOTS) continue;
                occ++;
                t[empty] = Slot();
                t[empty].b = b;
            }
        }
        std::stable_sort(t, t + SLOTS, [](const Slot& p, const Slot& q) {
            if ((p.b == -1) != (q.b == -1)) return q.b == -1;
            return p.b != -1 && p.b < q.b;
        });
        double fx = 0.0, fy = 0.0, m = 0.0;
        for (int j = 0; j < SLOTS; j++) {
            Slot& sl = t[j];
            if (sl.b == -1) continue;
            int32_t k = sl.bonded;
            if (k >= 0 && k < (int32_t)snap.size() && snap[k] < 0) { sl.bonded = snap[k]; k = snap[k]; }
            if (k < 0 || u.bonds == 0) linear_model(id, sl.b, sl, -1, fx, fy, m);
            else bond(id, sl.b, sl, k, fx, fy, m);
        }
        const double ms = mass(id), inertia = 0.5 * ms * r[id] * r[id];
        fx += ex[id]; fy += ey[id]; m += em[id];
        if (u.gravity == 1 && u.planet_mode == 1) {
            double cx = 0.0, cy = 0.0;
            if (u.mouse_gravity == 1) { cx = u.gravity_x; cy = u.gravity_y; }
            const double ddx = cx - x[id], ddy = cy - y[id];
            if (ddx != 0.0 || ddy != 0.0) {
                const double l = std::hypot(ddx, ddy);
                fx += ddx / l * G * u.gravity_acc * ms;
                fy += ddy / l * G * u.gravity_acc * ms;
            }
        } else if (u.gravity == 1) {
            fy += -G * u.gravity_acc * ms;
        }
        ax[id] = fx / ms;
        ay[id] = fy / ms;
        al[id] = m / inertia;
        const int32_t* f = &fix[6 * id];
        if (f[0] == 0) vx[id] += 0.5 * ax[id] * dt;
        if (f[1] == 0) vy[id] += 0.5 * ay[id] * dt;
        if (f[2] == 0) w[id] += 0.5 * al[id] * dt;
    }

    void step() {
        for (int i = 0; i < u.particles; i++) lom(i);
        snap = bidx;
        if (u.collisions == 1 && cc[0] == 1) pairs();
        for (int i = 0; i < u.particles; i++) sim(i);
        cc[2] += 1;
        cc[3] = 0;
    }

    double kinetic() const {
        double e = 0.0;
        for (int i = 0; i < u.particles; i++) {
            if (r[i] == 0.0) continue;
            const double ms = mass(i), inertia = 0.5 * ms * r[i] * r[i];
            e += 0.5 * ms * (vx[i] * vx[i] + vy[i] * vy[i]) + 0.5 * inertia * w[i] * w[i];
        }
        return e;
    }
};

inline float bitsf(int32_t i) { float f; std::memcpy(&f, &i, 4); return f; }

}  // namespace

extern "C" {

// Runs `nsteps` steps in f64 from the f32 state given in the reference layouts (state.rs:69-93;
// contacts: `slots` words x 6 per particle, bonds {index, angle bits, length bits},
// optional per-bond strengths), restitution as dem_params.restitution.  Writes the final
// state in f64 (pos/vel/acc 2n, rot/rot_vel/rot_acc n, data 7n; any may be NULL), the
// bond index words (n_bonds, may be NULL) and the kinetic energy after every step
// (nsteps entries, may be NULL).  Returns 0, or -1 on bad arguments.
int orc64_run(const dem_uniform* u, const dem_material* mats, int32_t n_mats, int32_t n, const float* pos,
              const float* vel, const float* acc, const float* rot, const float* rot_vel, const float* rot_acc,
              const float* forces, const float* radii, const int32_t* fixity, const int32_t* material_pointers,
              const int32_t* bonds, int32_t n_bonds, const float* bond_strengths, const float* contacts, int32_t slots,
              int32_t nsteps, float restitution, double* pos_out, double* vel_out, double* acc_out, double* rot_out,
              double* rot_vel_out, double* rot_acc_out, double* data_out, int32_t* bonds_out, double* ke_out) {
    if (!u || !mats || n_mats <= 0 || n < 0 || nsteps < 0 || u->particles > n || slots < 1) return -1;
    Sim s(*u, mats, n_mats, n, (double)u->dT);
    if (restitution > 0.0f && restitution < 1.0f) {
        const double le = std::log((double)restitution);
        s.cn2 = 2.0 * (-le / std::sqrt(M_PI * M_PI + le * le));
    }
    auto col = [&](std::vector<double>& v, const float* src, int stride, int k) {
        v.assign(n, 0.0);
        for (int i = 0; i < n; i++) v[i] = src[(size_t)stride * i + k];
    };
    col(s.x, pos, 2, 0); col(s.y, pos, 2, 1); col(s.vx, vel, 2, 0); col(s.vy, vel, 2, 1);
    col(s.ax, acc, 2, 0); col(s.ay, acc, 2, 1); col(s.th, rot, 1, 0); col(s.w, rot_vel, 1, 0);
    col(s.al, rot_acc, 1, 0); col(s.r, radii, 1, 0);
    col(s.ex, forces, 6, 0); col(s.ey, forces, 6, 1); col(s.em, forces, 6, 2);
    // WGPUComputeProg::restore uploads vel into del_pos and rot_vel into del_rot
    s.dpx = s.vx; s.dpy = s.vy; s.dth = s.w;
    s.data.assign((size_t)7 * n, 0.0);
    s.fix.assign(fixity, fixity + 6 * (size_t)n);
    s.mat.assign(material_pointers, material_pointers + n);
    s.bidx.assign(n_bonds, 0);
    s.blen.assign(n_bonds, 0.0);
    s.bsn.assign(n_bonds, u->bond_normal_strength);
    s.bss.assign(n_bonds, u->bond_shear_strength);
    for (int k = 0; k < n_bonds; k++) {
        s.bidx[k] = bonds[3 * k];
        s.blen[k] = bitsf(bonds[3 * k + 2]);
        if (bond_strengths) { s.bsn[k] = bond_strengths[2 * k]; s.bss[k] = bond_strengths[2 * k + 1]; }
    }
    s.tab.assign((size_t)n * SLOTS, Slot());
    if (contacts) {
        const int32_t* c = reinterpret_cast<const int32_t*>(contacts);
        for (int i = 0; i < n; i++) {
            int q = 0;
            for (int k = 0; k < slots && q < SLOTS; k++) {
                const int32_t* w = c + ((size_t)i * slots + k) * 6;
                if (w[1] == -1) continue;
                Slot& sl = s.tab[(size_t)i * SLOTS + q++];
                sl.b = w[1];
                sl.ft = bitsf(w[2]);
                sl.fbt = w[5] >= 0 ? (double)bitsf(w[3]) : 0.0;
                sl.tb = w[5] >= 0 ? (double)bitsf(w[4]) : 0.0;
                sl.bonded = w[5];
            }
        }
    }
    for (int k = 0; k < nsteps; k++) {
        s.step();
        if (ke_out) ke_out[k] = s.kinetic();
    }
    for (int i = 0; i < n; i++) {
        if (pos_out) { pos_out[2 * i] = s.x[i]; pos_out[2 * i + 1] = s.y[i]; }
        if (vel_out) { vel_out[2 * i] = s.vx[i]; vel_out[2 * i + 1] = s.vy[i]; }
        if (acc_out) { acc_out[2 * i] = s.ax[i]; acc_out[2 * i + 1] = s.ay[i]; }
        if (rot_out) rot_out[i] = s.th[i];
        if (rot_vel_out) rot_vel_out[i] = s.w[i];
        if (rot_acc_out) rot_acc_out[i] = s.al[i];
        if (data_out)
            for (int q = 0; q < 7; q++) data_out[7 * i + q] = s.data[7 * i + q];
    }
    if (bonds_out)
        for (int k = 0; k < n_bonds; k++) bonds_out[k] = s.bidx[k];
    return 0;
}

}  // extern "C"
