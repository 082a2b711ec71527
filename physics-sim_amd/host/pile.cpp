// pile.cpp - packed piles of polydisperse discs for the benchmark scenes (host only).
//
// Drop-and-roll deposition: discs are placed one after another, each at the lowest
// position near its drop line where it rests on two discs already placed (or on one
// disc and the floor, or on the floor), without overlapping any placed disc by more
// than the designed contact overlap.  The drop lines sweep across the column from left
// to right, row after row, so the pile grows in layers and every disc sits in a valley
// of the layer below: a pile that can stand under gravity with friction, touching
// about four neighbours per disc, instead of discs balanced on a single support.
//
// No counterpart in the reference: its scenes come from setup::grid (src/setup.rs:24-50),
// a lattice, which for polydisperse discs leaves them apart or overlapping.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace {

struct SplitMix {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

struct Pile {
    double left, width, floor, ovl, rmax;
    const float* r;
    const double* delta;   // optional per-disc contact overlap (absolute)
    std::vector<double> x, y;
    // cell grid of placed discs (cells of side >= 2 rmax, rows grow upward)
    double cs;
    int gw;
    std::vector<int> head;   // per cell, first disc; next[] links
    std::vector<int> next;
    // height map: the disc whose top surface is highest over each bin
    double hb;
    std::vector<int> top;
    std::vector<double> toph;

    int cell_col(double px) const { return std::min(std::max((int)((px - left) / cs), 0), gw - 1); }
    int cell_row(double py) const { return std::max((int)((py - floor) / cs), 0); }
    void insert(int i) {
        const int cr = cell_row(y[i]), cc = cell_col(x[i]);
        const size_t c = (size_t)cr * gw + cc;
        if (head.size() <= c) head.resize(((size_t)cr + 8) * gw, -1);
        next[i] = head[c];
        head[c] = i;
        const int b0 = std::max(0, (int)std::floor((x[i] - r[i] - left) / hb));
        const int b1 = std::min((int)top.size() - 1, (int)std::floor((x[i] + r[i] - left) / hb));
        for (int b = b0; b <= b1; b++) {
            const double xb = left + (b + 0.5) * hb;
            const double dx = xb - x[i];
            if (std::fabs(dx) >= r[i]) continue;
            const double h = y[i] + std::sqrt((double)r[i] * r[i] - dx * dx);
            if (h > toph[b]) { toph[b] = h; top[b] = i; }
        }
    }
    // designed centre distance of disc i (radius ri) to disc j: (ri + rj) less the larger
    // of ovl (ri + rj) and disc i's own overlap
    double sep(int i, double ri, double rj) const {
        const double s = ri + rj;
        return s - std::max(s * ovl, delta ? delta[i] : 0.0);
    }
    // whether disc i (radius ri) at (px, py) keeps every placed disc at its designed
    // distance or more, and stays on the floor and inside the column
    bool free_at(int i, double ri, double px, double py) const {
        if (py < floor + ri * (1.0 - 1e-9)) return false;
        if (px < left + ri || px > left + width - ri) return false;
        const int cr = cell_row(py), cc = cell_col(px);
        for (int rr = std::max(cr - 1, 0); rr <= cr + 1; rr++)
            for (int c = std::max(cc - 1, 0); c <= std::min(cc + 1, gw - 1); c++) {
                const size_t k = (size_t)rr * gw + c;
                if (k >= head.size()) continue;
                for (int j = head[k]; j >= 0; j = next[j]) {
                    const double s = sep(i, ri, r[j]) * (1.0 - 1e-12);
                    const double dx = px - x[j], dy = py - y[j];
                    if (dx * dx + dy * dy < s * s) return false;
                }
            }
        return true;
    }
};

}  // namespace

extern "C" {

// Deposits n discs of radii r[0..n) into a column [left, left + width] standing on the
// floor y = floor; disc i is dropped at the sweep position and rests where described
// above, touching its supports at centre distance r_i + r_j - max(overlap (r_i + r_j),
// delta[i]) (delta may be NULL: pressing a disc into its supports by the load it will
// carry lets a pile start near its static compression).  Writes f64 centres.  Returns
// 0, or -1 on bad arguments.
int psim_deposit_pile(int32_t n, const float* r, double left, double width, double floor, double overlap,
                      const double* delta, uint64_t seed, double* x_out, double* y_out) {
    if (n < 0 || !r || !x_out || !y_out || !(width > 0.0)) return -1;
    Pile p;
    p.left = left;
    p.width = width;
    p.floor = floor;
    p.ovl = overlap;
    p.r = r;
    p.delta = delta;
    p.rmax = 0.0;
    for (int32_t i = 0; i < n; i++) p.rmax = std::max(p.rmax, (double)r[i]);
    if (!(p.rmax > 0.0)) return -1;
    for (int32_t i = 0; i < n; i++)
        if (!(r[i] > 0.0f)) return -1;
    p.x.assign(n, 0.0);
    p.y.assign(n, 0.0);
    p.cs = 2.0 * p.rmax * 1.000001;
    p.gw = std::max(1, (int)std::ceil(width / p.cs));
    p.next.assign(n, -1);
    p.hb = p.rmax / 10.0;
    const int nb = std::max(1, (int)std::ceil(width / p.hb));
    p.top.assign(nb, -1);
    p.toph.assign(nb, -INFINITY);
    SplitMix rng{seed};
    double cursor = left;
    std::vector<int> sup;
    struct Cand { double y, x; };
    std::vector<Cand> cand;
    for (int32_t i = 0; i < n; i++) {
        const double ri = r[i];
        // drop line: the sweep advances by the disc's diameter and wraps at the right edge
        if (cursor + 2.0 * ri > left + width) cursor = left + rng.uniform() * p.rmax;
        const double x0 = std::min(cursor + ri, left + width - ri);
        cursor += 2.0 * ri;
        // surface discs over the window that a disc dropped at x0 can roll across
        const double reach = ri + p.rmax;
        const int b0 = std::max(0, (int)std::floor((x0 - reach - left) / p.hb));
        const int b1 = std::min(nb - 1, (int)std::floor((x0 + reach - left) / p.hb));
        sup.clear();
        for (int b = b0; b <= b1; b++)
            if (p.top[b] >= 0 && (sup.empty() || sup.back() != p.top[b])) sup.push_back(p.top[b]);
        std::sort(sup.begin(), sup.end());
        sup.erase(std::unique(sup.begin(), sup.end()), sup.end());
        cand.clear();
        const double yf = floor + ri * (1.0 + 1e-4);
        cand.push_back({yf, x0});
        const double win = p.rmax;
        for (size_t a = 0; a < sup.size(); a++) {
            const int A = sup[a];
            const double sa = p.sep(i, ri, r[A]);
            // on A alone, on the drop line
            const double dx0 = x0 - p.x[A];
            if (std::fabs(dx0) < sa) cand.push_back({p.y[A] + std::sqrt(sa * sa - dx0 * dx0), x0});
            // on A and the floor
            const double dyf = yf - p.y[A];
            if (std::fabs(dyf) < sa) {
                const double h = std::sqrt(sa * sa - dyf * dyf);
                for (double xs : {p.x[A] - h, p.x[A] + h})
                    if (std::fabs(xs - x0) <= win) cand.push_back({yf, xs});
            }
            // on A and B: the upper intersection of the two contact circles
            for (size_t b = a + 1; b < sup.size(); b++) {
                const int B = sup[b];
                const double sb = p.sep(i, ri, r[B]);
                const double ux = p.x[B] - p.x[A], uy = p.y[B] - p.y[A];
                const double d2 = ux * ux + uy * uy, d = std::sqrt(d2);
                if (!(d > 0.0) || d >= sa + sb || d <= std::fabs(sa - sb)) continue;
                const double t = (d2 + sa * sa - sb * sb) / (2.0 * d);
                const double h = std::sqrt(std::max(sa * sa - t * t, 0.0));
                const double ex = ux / d, ey = uy / d;
                const double mx = p.x[A] + t * ex, my = p.y[A] + t * ey;
                // the two points mx,my +- h * (-ey, ex): keep the upper one
                double px = mx - h * ey, py = my + h * ex;
                const double qx = mx + h * ey, qy = my - h * ex;
                if (qy > py) { px = qx; py = qy; }
                if (std::fabs(px - x0) <= win) cand.push_back({py, px});
            }
        }
        std::sort(cand.begin(), cand.end(), [](const Cand& u, const Cand& v) { return u.y < v.y || (u.y == v.y && u.x < v.x); });
        bool placed = false;
        for (const Cand& c : cand)
            if (p.free_at(i, ri, c.x, c.y)) {
                p.x[i] = c.x;
                p.y[i] = c.y;
                placed = true;
                break;
            }
        if (!placed) {
            // above everything in the window, on the drop line
            double yt = yf;
            for (int A : sup) yt = std::max(yt, p.y[A] + r[A] + ri);
            while (!p.free_at(i, ri, x0, yt)) yt += ri;
            p.x[i] = x0;
            p.y[i] = yt;
        }
        p.insert(i);
    }
    std::copy(p.x.begin(), p.x.end(), x_out);
    std::copy(p.y.begin(), p.y.end(), y_out);
    return 0;
}

}  // extern "C"
