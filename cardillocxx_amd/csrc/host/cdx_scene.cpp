// cdx_scene.cpp -- CPU contact generation and scene builders (see cdx_scene.h).
//
// Host-only C++ (g++ -fopenmp).  Contact generation runs on the CPU, as in the reference,
// and produces the contact buffer that cdx_set_contacts uploads to the device.
#include "cdx_scene.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

namespace {

struct V3 {
    double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

// R^T v for the rotation of unit quaternion (qx, qy, qz, qw) (Eigen coefficient order).
inline V3 rot_t(const double* q, V3 v) {
    const double x = q[0], y = q[1], z = q[2], w = q[3];
    const double r00 = 1 - 2 * (y * y + z * z), r01 = 2 * (x * y - z * w), r02 = 2 * (x * z + y * w);
    const double r10 = 2 * (x * y + z * w), r11 = 1 - 2 * (x * x + z * z), r12 = 2 * (y * z - x * w);
    const double r20 = 2 * (x * z - y * w), r21 = 2 * (y * z + x * w), r22 = 1 - 2 * (x * x + y * y);
    return {r00 * v.x + r10 * v.y + r20 * v.z, r01 * v.x + r11 * v.y + r21 * v.z, r02 * v.x + r12 * v.y + r22 * v.z};
}

// Spatial hash: bodies bucketed by cell via a counting sort.
struct Grid {
    double lo[3], cell;
    int n[3];
    std::vector<int> start, items;
    void build(int nb, const double* pos, int stride, double cell_size) {
        cell = cell_size;
        double hi[3];
        for (int k = 0; k < 3; ++k) {
            lo[k] = 1e300;
            hi[k] = -1e300;
        }
        for (int i = 0; i < nb; ++i)
            for (int k = 0; k < 3; ++k) {
                lo[k] = std::min(lo[k], pos[(size_t)stride * i + k]);
                hi[k] = std::max(hi[k], pos[(size_t)stride * i + k]);
            }
        if (nb == 0)
            for (int k = 0; k < 3; ++k) lo[k] = hi[k] = 0;
        // keep the table bounded: widen the cells if the box would need more than ~8M of them
        for (;;) {
            for (int k = 0; k < 3; ++k) n[k] = std::max(1, std::min(1 << 12, (int)std::floor((hi[k] - lo[k]) / cell) + 1));
            if ((double)n[0] * n[1] * n[2] <= 8e6) break;
            cell *= 2;
        }
        const size_t nc = (size_t)n[0] * n[1] * n[2];
        start.assign(nc + 1, 0);
        std::vector<int> cid(nb);
        for (int i = 0; i < nb; ++i) {
            cid[i] = cell_of(pos + (size_t)stride * i);
            start[cid[i] + 1]++;
        }
        for (size_t c = 0; c < nc; ++c) start[c + 1] += start[c];
        items.resize(nb);
        std::vector<int> fill(start.begin(), start.end() - 1);
        for (int i = 0; i < nb; ++i) items[fill[cid[i]]++] = i;
    }
    int coord(double v, int k) const { return std::max(0, std::min(n[k] - 1, (int)std::floor((v - lo[k]) / cell))); }
    int cell_of(const double* p) const { return (coord(p[2], 2) * n[1] + coord(p[1], 1)) * n[0] + coord(p[0], 0); }
    // Calls f(j) for every body in the 3x3x3 cells around p.
    template <typename F>
    void around(const double* p, F&& f) const {
        const int cx = coord(p[0], 0), cy = coord(p[1], 1), cz = coord(p[2], 2);
        for (int z = std::max(0, cz - 1); z <= std::min(n[2] - 1, cz + 1); ++z)
            for (int y = std::max(0, cy - 1); y <= std::min(n[1] - 1, cy + 1); ++y)
                for (int x = std::max(0, cx - 1); x <= std::min(n[0] - 1, cx + 1); ++x) {
                    const int c = (z * n[1] + y) * n[0] + x;
                    for (int k = start[c]; k < start[c + 1]; ++k) f(items[k]);
                }
    }
};

struct RawContact {
    int a, b;  // body ids, -1 static
    int wall;  // wall index for sphere-wall contacts, -1 otherwise
    V3 n, p;
    double depth;
};

void fill_contact(const RawContact& rc, const double* q, const double* mu_body, const double* mu_wall, int c,
                  int* body_a, int* body_b, double* mu, double* frame, double* angA, double* angB, double* point,
                  double* depth) {
    body_a[c] = rc.a;
    body_b[c] = rc.b;
    const double ma = rc.a >= 0 ? mu_body[rc.a] : mu_wall[rc.wall];
    const double mb = rc.b >= 0 ? mu_body[rc.b] : mu_wall[rc.wall];
    mu[c] = std::min(std::max(0.0, ma), std::max(0.0, mb));  // combineFrictionMu "min" (:94-106)
    double nn[3] = {rc.n.x, rc.n.y, rc.n.z}, t1[3], t2[3];
    cdx_tangent_frame(nn, t1, t2);
    const V3 d[3] = {{nn[0], nn[1], nn[2]}, {t1[0], t1[1], t1[2]}, {t2[0], t2[1], t2[2]}};
    for (int k = 0; k < 3; ++k) {
        frame[9 * (size_t)c + 3 * k + 0] = d[k].x;
        frame[9 * (size_t)c + 3 * k + 1] = d[k].y;
        frame[9 * (size_t)c + 3 * k + 2] = d[k].z;
    }
    // body-frame rows r_body x d_body (makeContact :151-160 + buildContactRowByDof)
    for (int side = 0; side < 2; ++side) {
        const int body = side == 0 ? rc.a : rc.b;
        double* ang = (side == 0 ? angA : angB) + 9 * (size_t)c;
        if (body < 0) {
            for (int k = 0; k < 9; ++k) ang[k] = 0;
            continue;
        }
        const double* qb = q + 7 * (size_t)body;
        const V3 r = rot_t(qb + 3, rc.p - V3{qb[0], qb[1], qb[2]});
        for (int k = 0; k < 3; ++k) {
            const V3 a = cross(r, rot_t(qb + 3, d[k]));
            ang[3 * k + 0] = a.x;
            ang[3 * k + 1] = a.y;
            ang[3 * k + 2] = a.z;
        }
    }
    point[3 * (size_t)c + 0] = rc.p.x;
    point[3 * (size_t)c + 1] = rc.p.y;
    point[3 * (size_t)c + 2] = rc.p.z;
    depth[c] = rc.depth;
}

}  // namespace

extern "C" {

void cdx_tangent_frame(const double n_in[3], double t1[3], double t2[3]) {
    double n[3] = {n_in[0], n_in[1], n_in[2]};
    const double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (len <= 2.220446049250313e-16) {  // degenerate: world axes (tangent_frame.hpp:26-31)
        t1[0] = 1, t1[1] = 0, t1[2] = 0;
        t2[0] = 0, t2[1] = 1, t2[2] = 0;
        return;
    }
    for (double& c : n) c /= len;
    V3 t;
    if (n[2] < -0.9999999) {  // south pole: fixed basis re-orthogonalised (:34-41)
        t = {0, -1, 0};
        const double dd = n[0] * t.x + n[1] * t.y + n[2] * t.z;
        t = t - dd * V3{n[0], n[1], n[2]};
    } else {  // Frisvad (:43-49)
        const double a = 1.0 / (1.0 + n[2]);
        const double b = -n[0] * n[1] * a;
        t = {1.0 - n[0] * n[0] * a, b, -n[0]};
    }
    const double tl = std::sqrt(dot(t, t));
    if (tl > 0) t = {t.x / tl, t.y / tl, t.z / tl};  // normalize(): divide by the norm
    t1[0] = t.x, t1[1] = t.y, t1[2] = t.z;
    const V3 c = cross({n[0], n[1], n[2]}, t);
    t2[0] = c.x, t2[1] = c.y, t2[2] = c.z;
}

int cdx_sphere_contacts(int nb, const double* q, const double* radius, const double* mu_body, int nw, const double* wall_n,
                        const double* wall_d, const double* mu_wall, double margin, int cap, int* body_a, int* body_b,
                        double* mu, double* frame, double* angA, double* angB, double* point, double* depth) {
    if (nb < 0 || nw < 0 || (nb > 0 && (!q || !radius || !mu_body))) return -1;
    double rmax = 0;
    for (int i = 0; i < nb; ++i) rmax = std::max(rmax, radius[i]);
    Grid g;
    g.build(nb, q, 7, std::max(2 * rmax + margin, 1e-9));
    std::vector<std::vector<RawContact>> per(nb);
#pragma omp parallel for schedule(dynamic, 256)
    for (int a = 0; a < nb; ++a) {
        const double* xa = q + 7 * (size_t)a;
        const V3 pa{xa[0], xa[1], xa[2]};
        auto& out = per[a];
        for (int w = 0; w < nw; ++w) {
            const V3 n{wall_n[3 * w], wall_n[3 * w + 1], wall_n[3 * w + 2]};
            const double h = dot(n, pa) - wall_d[w];  // signed distance of the centre
            const double gap = h - radius[a];
            if (gap > margin) continue;
            const V3 p1 = pa - h * n;            // witness on the wall
            const V3 p2 = pa - radius[a] * n;    // witness on the sphere
            out.push_back(RawContact{-1, a, w, n, 0.5 * (p1 + p2), std::max(0.0, -gap)});
        }
        std::vector<int> nbrs;
        g.around(xa, [&](int b) {
            if (b > a) nbrs.push_back(b);
        });
        std::sort(nbrs.begin(), nbrs.end());
        for (int b : nbrs) {
            const double* xb = q + 7 * (size_t)b;
            const V3 pb{xb[0], xb[1], xb[2]};
            const V3 dlt = pb - pa;
            const double dist = std::sqrt(dot(dlt, dlt));
            const double gap = dist - radius[a] - radius[b];
            if (gap > margin) continue;
            const V3 n = dist > 0 ? (1.0 / dist) * dlt : V3{0, 0, 1};
            const V3 p1 = pa + radius[a] * n, p2 = pb - radius[b] * n;
            out.push_back(RawContact{a, b, -1, n, 0.5 * (p1 + p2), std::max(0.0, -gap)});
        }
    }
    std::vector<int> off(nb + 1, 0);
    for (int a = 0; a < nb; ++a) off[a + 1] = off[a] + (int)per[a].size();
    const int total = off[nb];
    if (total > cap) return total;
#pragma omp parallel for schedule(dynamic, 256)
    for (int a = 0; a < nb; ++a)
        for (size_t k = 0; k < per[a].size(); ++k)
            fill_contact(per[a][k], q, mu_body, mu_wall, off[a] + (int)k, body_a, body_b, mu, frame, angA, angB, point, depth);
    return total;
}

int cdx_match_warmstart(int nprev, const int* prev_a, const int* prev_b, const double* prev_point, const double* prev_imp,
                        int ncur, const int* cur_a, const int* cur_b, const double* cur_point, double max_dist,
                        double* warm_out) {
    if (nprev < 0 || ncur < 0) return -1;
    auto key = [](int a, int b) { return ((int64_t)a << 32) ^ (int64_t)(uint32_t)b; };
    std::vector<int> pi(nprev), ci(ncur);
    std::iota(pi.begin(), pi.end(), 0);
    std::iota(ci.begin(), ci.end(), 0);
    std::stable_sort(pi.begin(), pi.end(), [&](int x, int y) { return key(prev_a[x], prev_b[x]) < key(prev_a[y], prev_b[y]); });
    std::stable_sort(ci.begin(), ci.end(), [&](int x, int y) { return key(cur_a[x], cur_b[x]) < key(cur_a[y], cur_b[y]); });
    std::memset(warm_out, 0, sizeof(double) * 3 * (size_t)ncur);
    const double max2 = max_dist * max_dist;
    size_t p = 0;
    for (size_t i = 0; i < ci.size();) {
        const int64_t k = key(cur_a[ci[i]], cur_b[ci[i]]);
        size_t i1 = i;
        while (i1 < ci.size() && key(cur_a[ci[i1]], cur_b[ci[i1]]) == k) ++i1;
        while (p < pi.size() && key(prev_a[pi[p]], prev_b[pi[p]]) < k) ++p;
        size_t p1 = p;
        while (p1 < pi.size() && key(prev_a[pi[p1]], prev_b[pi[p1]]) == k) ++p1;
        if (p1 > p) {
            // greedy nearest-neighbour matching, closest pairs first, each previous used once
            std::vector<std::pair<double, std::pair<size_t, size_t>>> cand;
            for (size_t u = i; u < i1; ++u)
                for (size_t v = p; v < p1; ++v) {
                    const double* a = cur_point + 3 * (size_t)ci[u];
                    const double* b = prev_point + 3 * (size_t)pi[v];
                    const double d2 = (a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]);
                    if (d2 <= max2) cand.push_back({d2, {u - i, v - p}});
                }
            std::stable_sort(cand.begin(), cand.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
            std::vector<char> cu(i1 - i, 0), pu(p1 - p, 0);
            for (const auto& cd : cand) {
                const size_t u = cd.second.first, v = cd.second.second;
                if (cu[u] || pu[v]) continue;
                cu[u] = pu[v] = 1;
                for (int m = 0; m < 3; ++m) warm_out[3 * (size_t)ci[i + u] + m] = prev_imp[3 * (size_t)pi[p + v] + m];
            }
        }
        i = i1;
    }
    return 0;
}

static void random_quat(std::mt19937_64& rng, double* q) {
    // Shoemake's uniform random rotation
    std::uniform_real_distribution<double> U(0.0, 1.0);
    const double u1 = U(rng), u2 = U(rng), u3 = U(rng);
    const double s1 = std::sqrt(1 - u1), s2 = std::sqrt(u1);
    q[0] = s1 * std::sin(2 * M_PI * u2);
    q[1] = s1 * std::cos(2 * M_PI * u2);
    q[2] = s2 * std::sin(2 * M_PI * u3);
    q[3] = s2 * std::cos(2 * M_PI * u3);
}

int cdx_scene_sphere_grid(int n, int nx, int ny, double pitch, double z0, double jitter, uint64_t seed, double* q) {
    if (n < 0 || nx <= 0 || ny <= 0) return -1;
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> J(-jitter, jitter);
    const double ox = -0.5 * (nx - 1) * pitch, oy = -0.5 * (ny - 1) * pitch;
    for (int i = 0; i < n; ++i) {
        const int ix = i % nx, iy = (i / nx) % ny, iz = i / (nx * ny);
        double* b = q + 7 * (size_t)i;
        b[0] = ox + ix * pitch + J(rng);
        b[1] = oy + iy * pitch + J(rng);
        b[2] = z0 + iz * pitch;
        random_quat(rng, b + 3);
    }
    return n;
}

static uint64_t spread3(uint64_t v) {
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
}

int cdx_scene_sphere_pile(int n, double rmin, double rmax, double Lx, double Ly, double Lz, uint64_t seed, int settle_steps,
                          double* q, double* radius) {
    if (n <= 0 || !(rmin > 0) || rmax < rmin) return -1;
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    std::vector<double> r(n), x(3 * (size_t)n);
    const double lr = std::log(rmax / rmin);
    for (int i = 0; i < n; ++i) r[i] = rmin * std::exp(U(rng) * lr);  // log-uniform
    // Random sequential placement, largest first, rejection against a cell grid.
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return r[a] > r[b]; });
    const double cell = 2 * rmax;
    const int gx = std::max(1, (int)(Lx / cell)), gy = std::max(1, (int)(Ly / cell));
    double zmax = Lz;
    std::vector<std::vector<int>> cells;
    int gz = 0;
    auto ensure_gz = [&](double top) {
        const int need = (int)(top / cell) + 2;
        if (need > gz) {
            gz = need;
            cells.resize((size_t)gx * gy * gz);
        }
    };
    ensure_gz(zmax);
    auto cidx = [&](double px, double py, double pz, int& cx, int& cy, int& cz) {
        cx = std::min(gx - 1, std::max(0, (int)(px / (Lx / gx))));
        cy = std::min(gy - 1, std::max(0, (int)(py / (Ly / gy))));
        cz = std::max(0, (int)(pz / cell));
    };
    for (int oi = 0; oi < n; ++oi) {
        const int i = order[oi];
        int tries = 0;
        for (;;) {
            const double px = r[i] + U(rng) * (Lx - 2 * r[i]);
            const double py = r[i] + U(rng) * (Ly - 2 * r[i]);
            const double pz = r[i] + U(rng) * (zmax - 2 * r[i]);
            int cx, cy, cz;
            cidx(px, py, pz, cx, cy, cz);
            ensure_gz(pz + cell);
            bool ok = true;
            for (int z = std::max(0, cz - 1); ok && z <= std::min(gz - 1, cz + 1); ++z)
                for (int y = std::max(0, cy - 1); ok && y <= std::min(gy - 1, cy + 1); ++y)
                    for (int xx = std::max(0, cx - 1); ok && xx <= std::min(gx - 1, cx + 1); ++xx)
                        for (int j : cells[((size_t)z * gy + y) * gx + xx]) {
                            const double dx = px - x[3 * j], dy = py - x[3 * j + 1], dz = pz - x[3 * j + 2];
                            const double rr = r[i] + r[j];
                            if (dx * dx + dy * dy + dz * dz < rr * rr) {
                                ok = false;
                                break;
                            }
                        }
            if (ok) {
                x[3 * i] = px;
                x[3 * i + 1] = py;
                x[3 * i + 2] = pz;
                cells[((size_t)cz * gy + cy) * gx + cx].push_back(i);
                break;
            }
            if (++tries % 200 == 0) zmax *= 1.02;  // column grows when the box is too full
        }
    }
    // Position-based gravity settle: predict, Jacobi overlap projection, velocity update.
    std::vector<double> v(3 * (size_t)n, 0.0), xo(3 * (size_t)n), dx(3 * (size_t)n), m(n);
    for (int i = 0; i < n; ++i) m[i] = r[i] * r[i] * r[i];
    const double dt = 0.01, g = 9.81, dmax = 0.5 * rmin, skin = 0.6 * rmin;
    std::vector<std::vector<int>> nb(n);
    // the last steps run without gravity so that the Jacobi projection removes residual overlap
    const int relax_steps = settle_steps / 6;
    for (int s = 0; s < settle_steps; ++s) {
        const bool relax = s >= settle_steps - relax_steps;
        xo = x;
        for (int i = 0; i < n; ++i) {
            if (relax) {
                v[3 * i] = v[3 * i + 1] = v[3 * i + 2] = 0.0;
                continue;
            }
            v[3 * i + 2] -= g * dt;
            for (int k = 0; k < 3; ++k) {
                const double d = std::max(-dmax, std::min(dmax, v[3 * i + k] * dt));
                x[3 * i + k] += d;
            }
        }
        Grid grid;
        grid.build(n, x.data(), 3, 2 * rmax + skin);
#pragma omp parallel for schedule(dynamic, 512)
        for (int i = 0; i < n; ++i) {
            nb[i].clear();
            grid.around(&x[3 * i], [&](int j) {
                if (j == i) return;
                const double ddx = x[3 * i] - x[3 * j], ddy = x[3 * i + 1] - x[3 * j + 1], ddz = x[3 * i + 2] - x[3 * j + 2];
                const double cut = r[i] + r[j] + skin;
                if (ddx * ddx + ddy * ddy + ddz * ddz < cut * cut) nb[i].push_back(j);
            });
            std::sort(nb[i].begin(), nb[i].end());
        }
        for (int pass = 0; pass < 8; ++pass) {
#pragma omp parallel for schedule(dynamic, 512)
            for (int i = 0; i < n; ++i) {
                double d[3] = {0, 0, 0};
                for (int j : nb[i]) {
                    double e[3] = {x[3 * i] - x[3 * j], x[3 * i + 1] - x[3 * j + 1], x[3 * i + 2] - x[3 * j + 2]};
                    const double dist = std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
                    const double ov = r[i] + r[j] - dist;
                    if (ov <= 0 || dist <= 0) continue;
                    const double w = ov * m[j] / (m[i] + m[j]) / dist;
                    for (int k = 0; k < 3; ++k) d[k] += w * e[k];
                }
                for (int k = 0; k < 3; ++k) dx[3 * i + k] = 0.7 * d[k];
            }
            for (int i = 0; i < n; ++i) {
                for (int k = 0; k < 3; ++k) x[3 * i + k] += dx[3 * i + k];
                x[3 * i] = std::min(Lx - r[i], std::max(r[i], x[3 * i]));
                x[3 * i + 1] = std::min(Ly - r[i], std::max(r[i], x[3 * i + 1]));
                x[3 * i + 2] = std::max(r[i], x[3 * i + 2]);
            }
        }
        for (size_t k = 0; k < v.size(); ++k) v[k] = 0.9 * (x[k] - xo[k]) / dt;
    }
    // Morton order for locality of the body arrays (and so of the contact list).
    std::vector<uint64_t> code(n);
    const double cs = rmin;
    for (int i = 0; i < n; ++i)
        code[i] = spread3((uint64_t)(x[3 * i] / cs)) | spread3((uint64_t)(x[3 * i + 1] / cs)) << 1 |
                  spread3((uint64_t)(std::max(0.0, x[3 * i + 2]) / cs)) << 2;
    std::vector<int> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return code[a] < code[b]; });
    for (int k = 0; k < n; ++k) {
        const int i = perm[k];
        double* b = q + 7 * (size_t)k;
        b[0] = x[3 * i];
        b[1] = x[3 * i + 1];
        b[2] = x[3 * i + 2];
        random_quat(rng, b + 3);
        radius[k] = r[i];
    }
    return n;
}

}  // extern "C"
