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
    // the normal row is the unit normal the frame was built from (n / |n|, as cdx_tangent_frame
    // normalises it; at most an ulp from the narrowphase's), so that t2 = n x t1 holds exactly
    const double len = std::sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
    if (len > 2.220446049250313e-16)
        for (double& x : nn) x /= len;
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


}  // extern "C"

namespace {

struct Body {
    int type;  // 0 sphere, 1 box, 2 capsule
    V3 c;
    double R[3][3];  // columns are the body axes in world coordinates
    V3 h;            // sphere: h.x = radius; box: half extents; capsule: h.x = radius, h.y = half length (body z)
    double bound;    // bounding-sphere radius
};

inline V3 col(const double R[3][3], int k) { return {R[0][k], R[1][k], R[2][k]}; }

void rot_of(const double* q, double R[3][3]) {
    const double x = q[0], y = q[1], z = q[2], w = q[3];
    R[0][0] = 1 - 2 * (y * y + z * z), R[0][1] = 2 * (x * y - z * w), R[0][2] = 2 * (x * z + y * w);
    R[1][0] = 2 * (x * y + z * w), R[1][1] = 1 - 2 * (x * x + z * z), R[1][2] = 2 * (y * z - x * w);
    R[2][0] = 2 * (x * z - y * w), R[2][1] = 2 * (y * z + x * w), R[2][2] = 1 - 2 * (x * x + y * y);
}

Body make_body(const double* q, int type, const double* dims) {
    Body b;
    b.type = type;
    b.c = {q[0], q[1], q[2]};
    rot_of(q + 3, b.R);
    if (type == 0) {
        b.h = {dims[0], 0, 0};
        b.bound = dims[0];
    } else if (type == 2) {
        b.h = {dims[0], dims[1], 0};
        b.bound = dims[0] + dims[1];
    } else {
        b.h = {dims[0], dims[1], dims[2]};
        b.bound = std::sqrt(dims[0] * dims[0] + dims[1] * dims[1] + dims[2] * dims[2]);
    }
    return b;
}

double hcomp(const V3& h, int k) { return k == 0 ? h.x : (k == 1 ? h.y : h.z); }

// Capsule segment end points (the axis is the body z axis).
inline void capsule_ends(const Body& C, V3& p0, V3& p1) {
    const V3 ax = col(C.R, 2);
    p0 = C.c - C.h.y * ax;
    p1 = C.c + C.h.y * ax;
}


// Body (dynamic, index `bi`) against static half-space n.x >= d; the wall is side A.
void wall_contacts(int bi, int w, const Body& B, V3 n, double d, double margin, std::vector<RawContact>& out) {
    if (B.type == 2) {  // capsule: its two segment ends
        V3 ends[2];
        capsule_ends(B, ends[0], ends[1]);
        for (const V3& P : ends) {
            const double hgt = dot(n, P) - d;
            const double gap = hgt - B.h.x;
            if (gap > margin) continue;
            out.push_back(RawContact{-1, bi, w, n, 0.5 * ((P - hgt * n) + (P - B.h.x * n)), std::max(0.0, -gap)});
        }
        return;
    }
    if (B.type == 0) {
        const double hgt = dot(n, B.c) - d;
        const double gap = hgt - B.h.x;
        if (gap > margin) return;
        const V3 p1 = B.c - hgt * n, p2 = B.c - B.h.x * n;
        out.push_back(RawContact{-1, bi, w, n, 0.5 * (p1 + p2), std::max(0.0, -gap)});
        return;
    }
    for (int k = 0; k < 8; ++k) {
        const double sx = (k & 1) ? 1 : -1, sy = (k & 2) ? 1 : -1, sz = (k & 4) ? 1 : -1;
        const V3 corner = B.c + (sx * B.h.x) * col(B.R, 0) + (sy * B.h.y) * col(B.R, 1) + (sz * B.h.z) * col(B.R, 2);
        const double dist = dot(n, corner) - d;
        if (dist > margin) continue;
        out.push_back(RawContact{-1, bi, w, n, corner - (0.5 * dist) * n, std::max(0.0, -dist)});
    }
}

// Sphere S against box X; returns normal from the box to the sphere, gap and the two witnesses.
bool sphere_box(const Body& S, const Body& X, double margin, V3& n, V3& pbox, V3& psph, double& gap) {
    const V3 dl = S.c - X.c;
    double loc[3] = {dot(dl, col(X.R, 0)), dot(dl, col(X.R, 1)), dot(dl, col(X.R, 2))};
    double cp[3];
    bool inside = true;
    for (int k = 0; k < 3; ++k) {
        const double hk = hcomp(X.h, k);
        cp[k] = std::max(-hk, std::min(hk, loc[k]));
        if (cp[k] != loc[k]) inside = false;
    }
    V3 nl;
    if (!inside) {
        V3 diff{loc[0] - cp[0], loc[1] - cp[1], loc[2] - cp[2]};
        const double dist = std::sqrt(dot(diff, diff));
        gap = dist - S.h.x;
        if (gap > margin) return false;
        nl = (1.0 / dist) * diff;
    } else {
        int kb = 0;
        double pen = 1e300;
        for (int k = 0; k < 3; ++k) {
            const double pk = hcomp(X.h, k) - std::fabs(loc[k]);
            if (pk < pen) pen = pk, kb = k;
        }
        const double sg = loc[kb] >= 0 ? 1.0 : -1.0;
        nl = {kb == 0 ? sg : 0.0, kb == 1 ? sg : 0.0, kb == 2 ? sg : 0.0};
        cp[kb] = sg * hcomp(X.h, kb);
        gap = -(pen + S.h.x);
    }
    n = nl.x * col(X.R, 0) + nl.y * col(X.R, 1) + nl.z * col(X.R, 2);
    pbox = X.c + cp[0] * col(X.R, 0) + cp[1] * col(X.R, 1) + cp[2] * col(X.R, 2);
    psph = S.c - S.h.x * n;
    return true;
}

// Box-box by the separating-axis test; contacts from clipping the incident face against the
// reference face (face axes) or one closest-point contact (edge-edge axes).  Normal from A to B.
void box_box(int ia, int ib, const Body& A, const Body& B, double margin, std::vector<RawContact>& out) {
    const V3 d = B.c - A.c;
    double best = -1e300;
    int kind = -1, ai = -1, bj = -1;
    V3 axis_best{0, 0, 1};
    auto proj = [](const Body& X, V3 ax) {
        return X.h.x * std::fabs(dot(col(X.R, 0), ax)) + X.h.y * std::fabs(dot(col(X.R, 1), ax)) +
               X.h.z * std::fabs(dot(col(X.R, 2), ax));
    };
    auto test = [&](V3 ax, int k, int i, int j, double bias) {
        const double sep = std::fabs(dot(d, ax)) - proj(A, ax) - proj(B, ax);
        if (sep > margin) return false;
        if (sep > best + bias) {
            best = sep;
            kind = k;
            ai = i;
            bj = j;
            axis_best = ax;
        }
        return true;
    };
    for (int i = 0; i < 3; ++i)
        if (!test(col(A.R, i), 0, i, -1, 0.0)) return;
    for (int j = 0; j < 3; ++j)
        if (!test(col(B.R, j), 1, -1, j, 0.0)) return;
    const double edge_bias = 1e-9 + 0.05 * std::fabs(best);  // prefer face contacts when nearly tied
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            V3 ax = cross(col(A.R, i), col(B.R, j));
            const double len = std::sqrt(dot(ax, ax));
            if (len < 1e-9) continue;
            if (!test((1.0 / len) * ax, 2, i, j, edge_bias)) return;
        }
    V3 n = dot(d, axis_best) < 0 ? -1.0 * axis_best : axis_best;  // A -> B
    if (kind == 2) {
        // edges of A along axis ai and of B along bj closest to each other
        V3 ea = A.c, eb = B.c;
        for (int k = 0; k < 3; ++k) {
            if (k != ai) ea = ea + ((dot(col(A.R, k), n) > 0 ? 1.0 : -1.0) * hcomp(A.h, k)) * col(A.R, k);
            if (k != bj) eb = eb - ((dot(col(B.R, k), n) > 0 ? 1.0 : -1.0) * hcomp(B.h, k)) * col(B.R, k);
        }
        const V3 ua = col(A.R, ai), ub = col(B.R, bj);
        const V3 r = ea - eb;
        const double b = dot(ua, ub), dd = dot(ua, r), e = dot(ub, r);
        const double den = 1 - b * b;
        double s = den > 1e-12 ? (b * e - dd) / den : 0.0;
        s = std::max(-hcomp(A.h, ai), std::min(hcomp(A.h, ai), s));
        double t = std::max(-hcomp(B.h, bj), std::min(hcomp(B.h, bj), b * s + e));
        const V3 pa = ea + s * ua, pb = eb + t * ub;
        const double sep = dot(pb - pa, n);
        if (sep > margin) return;
        out.push_back(RawContact{ia, ib, -1, n, 0.5 * (pa + pb), std::max(0.0, -sep)});
        return;
    }
    // face contact: reference face on R (outward normal nr), incident face on I
    const Body& Rb = kind == 0 ? A : B;
    const Body& Ib = kind == 0 ? B : A;
    const int rf = kind == 0 ? ai : bj;
    const V3 nr = kind == 0 ? n : -1.0 * n;
    int jf = 0;
    double bd = -1;
    for (int k = 0; k < 3; ++k) {
        const double v = std::fabs(dot(col(Ib.R, k), nr));
        if (v > bd) bd = v, jf = k;
    }
    const double sg = dot(col(Ib.R, jf), nr) > 0 ? -1.0 : 1.0;
    const V3 fc = Ib.c + (sg * hcomp(Ib.h, jf)) * col(Ib.R, jf);
    const int u1 = (jf + 1) % 3, u2 = (jf + 2) % 3;
    std::vector<V3> poly = {fc + hcomp(Ib.h, u1) * col(Ib.R, u1) + hcomp(Ib.h, u2) * col(Ib.R, u2),
                            fc - hcomp(Ib.h, u1) * col(Ib.R, u1) + hcomp(Ib.h, u2) * col(Ib.R, u2),
                            fc - hcomp(Ib.h, u1) * col(Ib.R, u1) - hcomp(Ib.h, u2) * col(Ib.R, u2),
                            fc + hcomp(Ib.h, u1) * col(Ib.R, u1) - hcomp(Ib.h, u2) * col(Ib.R, u2)};
    // Sutherland-Hodgman against the four side planes of the reference face
    const int r1 = (rf + 1) % 3, r2 = (rf + 2) % 3;
    const int sides[4][2] = {{r1, 1}, {r1, -1}, {r2, 1}, {r2, -1}};
    for (auto& sd : sides) {
        const V3 ax = ((double)sd[1]) * col(Rb.R, sd[0]);
        const double lim = dot(ax, Rb.c) + hcomp(Rb.h, sd[0]);
        std::vector<V3> np;
        for (size_t k = 0; k < poly.size(); ++k) {
            const V3 P = poly[k], Q = poly[(k + 1) % poly.size()];
            const double dp = dot(ax, P) - lim, dq = dot(ax, Q) - lim;
            if (dp <= 0) np.push_back(P);
            if ((dp <= 0) != (dq <= 0)) np.push_back(P + (dp / (dp - dq)) * (Q - P));
        }
        poly.swap(np);
        if (poly.empty()) return;
    }
    const double plane = dot(nr, Rb.c) + hcomp(Rb.h, rf);
    std::vector<V3> pts;
    std::vector<double> seps;
    for (const V3& P : poly) {
        const double sep = dot(nr, P) - plane;
        if (sep > margin) continue;
        pts.push_back(P);
        seps.push_back(sep);
    }
    // At most four points per face contact: the deepest, the point farthest from it, the one
    // spanning the largest triangle with those two, and the one adding the largest triangle on
    // an edge of that triangle.  A clipped polygon can carry up to eight vertices; more rows on
    // one body pair tighten the coupling the Jacobi relaxation alpha = 0.3 has to absorb.
    std::vector<int> keep;
    if (pts.size() <= 4) {
        for (size_t k = 0; k < pts.size(); ++k) keep.push_back((int)k);
    } else {
        auto tri = [&](int i, int j, int k) {
            const V3 c = cross(pts[j] - pts[i], pts[k] - pts[i]);
            return dot(c, c);
        };
        auto chosen = [&](int k) { return std::find(keep.begin(), keep.end(), k) != keep.end(); };
        int s0 = 0;
        for (int k = 1; k < (int)pts.size(); ++k)
            if (seps[k] < seps[s0]) s0 = k;
        keep.push_back(s0);
        int s1 = -1;
        double best1 = -1;
        for (int k = 0; k < (int)pts.size(); ++k) {
            const V3 e = pts[k] - pts[s0];
            if (!chosen(k) && dot(e, e) > best1) best1 = dot(e, e), s1 = k;
        }
        keep.push_back(s1);
        int s2 = -1;
        double best2 = -1;
        for (int k = 0; k < (int)pts.size(); ++k)
            if (!chosen(k) && tri(s0, s1, k) > best2) best2 = tri(s0, s1, k), s2 = k;
        keep.push_back(s2);
        int s3 = -1;
        double best3 = -1;
        for (int k = 0; k < (int)pts.size(); ++k) {
            if (chosen(k)) continue;
            const double a = std::max(tri(s0, s1, k), std::max(tri(s1, s2, k), tri(s2, s0, k)));
            if (a > best3) best3 = a, s3 = k;
        }
        keep.push_back(s3);
        std::sort(keep.begin(), keep.end());  // polygon order
    }
    for (int k : keep)
        out.push_back(RawContact{ia, ib, -1, n, pts[k] - (0.5 * seps[k]) * nr, std::max(0.0, -seps[k])});
}

// Closest point on segment [p0, p1] to x.
inline V3 closest_on_segment(V3 p0, V3 p1, V3 x) {
    const V3 e = p1 - p0;
    const double ee = dot(e, e);
    const double t = ee > 0 ? std::max(0.0, std::min(1.0, dot(x - p0, e) / ee)) : 0.0;
    return p0 + t * e;
}

// A sphere of radius ra at pa against one of radius rb at pb: normal from a to b.
inline void ball_ball(int ia, int ib, V3 pa, double ra, V3 pb, double rb, double margin, std::vector<RawContact>& out) {
    const V3 d = pb - pa;
    const double dist = std::sqrt(dot(d, d));
    const double gap = dist - ra - rb;
    if (gap > margin) return;
    const V3 n = dist > 0 ? (1.0 / dist) * d : V3{0, 0, 1};
    out.push_back(RawContact{ia, ib, -1, n, 0.5 * ((pa + ra * n) + (pb - rb * n)), std::max(0.0, -gap)});
}

// Capsule-capsule: closest points of the two segments (Ericson, Real-Time Collision Detection,
// 5.1.9); for nearly parallel overlapping segments the two ends of the overlap, so a capsule
// lying on another gets a two-point support.
void capsule_capsule(int ia, int ib, const Body& A, const Body& B, double margin, std::vector<RawContact>& out) {
    V3 a0, a1, b0, b1;
    capsule_ends(A, a0, a1);
    capsule_ends(B, b0, b1);
    const V3 d1 = a1 - a0, d2 = b1 - b0, r = a0 - b0;
    const double aa = dot(d1, d1), ee = dot(d2, d2), ff = dot(d2, r);
    const double cc = dot(d1, r), bb = dot(d1, d2);
    const double den = aa * ee - bb * bb;
    if (aa > 0 && ee > 0 && den <= 1e-6 * aa * ee) {
        // parallel: project B's ends onto A's axis and keep the overlap
        const double t0 = std::max(0.0, std::min(1.0, dot(b0 - a0, d1) / aa));
        const double t1 = std::max(0.0, std::min(1.0, dot(b1 - a0, d1) / aa));
        const double lo = std::min(t0, t1), hi = std::max(t0, t1);
        const V3 pa_lo = a0 + lo * d1, pa_hi = a0 + hi * d1;
        ball_ball(ia, ib, pa_lo, A.h.x, closest_on_segment(b0, b1, pa_lo), B.h.x, margin, out);
        if (hi - lo > 1e-9) ball_ball(ia, ib, pa_hi, A.h.x, closest_on_segment(b0, b1, pa_hi), B.h.x, margin, out);
        return;
    }
    double sa = 0, tb = 0;
    if (aa <= 0 && ee <= 0) {
        sa = tb = 0;
    } else if (aa <= 0) {
        tb = std::max(0.0, std::min(1.0, ff / ee));
    } else if (ee <= 0) {
        sa = std::max(0.0, std::min(1.0, -cc / aa));
    } else {
        sa = std::max(0.0, std::min(1.0, (bb * ff - cc * ee) / den));
        tb = (bb * sa + ff) / ee;
        if (tb < 0) {
            tb = 0;
            sa = std::max(0.0, std::min(1.0, -cc / aa));
        } else if (tb > 1) {
            tb = 1;
            sa = std::max(0.0, std::min(1.0, (bb - cc) / aa));
        }
    }
    ball_ball(ia, ib, a0 + sa * d1, A.h.x, b0 + tb * d2, B.h.x, margin, out);
}

// Capsule-box by sphere probes: the two segment ends and the point of the segment nearest the
// box centre, each against the box (closest point); probes closer than a radius to an earlier
// one are skipped.
void capsule_box(int ia, int ib, const Body& A, const Body& B, double margin, std::vector<RawContact>& out) {
    const bool a_caps = A.type == 2;
    const Body& C = a_caps ? A : B;
    const Body& X = a_caps ? B : A;
    V3 p0, p1;
    capsule_ends(C, p0, p1);
    const V3 probes[3] = {p0, p1, closest_on_segment(p0, p1, X.c)};
    for (int k = 0; k < 3; ++k) {
        bool dup = false;
        for (int m = 0; m < k; ++m) {
            const V3 d = probes[k] - probes[m];
            if (dot(d, d) < C.h.x * C.h.x) dup = true;
        }
        if (dup) continue;
        Body S;
        S.type = 0;
        S.c = probes[k];
        S.h = {C.h.x, 0, 0};
        S.bound = C.h.x;
        V3 n, pbox, psph;
        double gap;
        if (!sphere_box(S, X, margin, n, pbox, psph, gap)) continue;
        const V3 nab = a_caps ? -1.0 * n : n;  // n points from the box to the probe
        out.push_back(RawContact{ia, ib, -1, nab, 0.5 * (pbox + psph), std::max(0.0, -gap)});
    }
}

// ---- static heightfield: samples h[i + nx j] at (x0 + i dx, y0 + j dy); cell (i, j) is split
// along the diagonal (i+1, j)-(i, j+1) into two triangles.
struct HField {
    int nx = 0, ny = 0;
    double x0 = 0, y0 = 0, dx = 1, dy = 1;
    const double* h = nullptr;
    int wall = -1;  // friction slot
    V3 vert(int i, int j) const { return {x0 + i * dx, y0 + j * dy, h[i + (size_t)nx * j]}; }
};

// Closest point of triangle (a, b, c) to p (Ericson, Real-Time Collision Detection, 5.1.5).
V3 closest_on_triangle(V3 p, V3 a, V3 b, V3 c) {
    const V3 ab = b - a, ac = c - a, ap = p - a;
    const double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0 && d2 <= 0) return a;
    const V3 bp = p - b;
    const double d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0 && d4 <= d3) return b;
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) return a + (d1 / (d1 - d3)) * ab;
    const V3 cp = p - c;
    const double d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0 && d5 <= d6) return c;
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) return a + (d2 / (d2 - d6)) * ac;
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) return b + ((d4 - d3) / ((d4 - d3) + (d5 - d6))) * (c - b);
    const double den = 1.0 / (va + vb + vc);
    return a + (vb * den) * ab + (vc * den) * ac;
}

// One probe (a sphere of radius r at P; r = 0 for a point) against the heightfield.  Above the
// surface: the closest point over the triangles under the probe's footprint; below it: the
// triangle under P, its upward normal and the signed distance to its plane.  Normal from the
// terrain (side A, static) to the body.  Outside the sampled rectangle there is no terrain.
bool hf_probe(const HField& hf, V3 P, double r, double margin, V3& n, V3& q, double& gap) {
    const double fx = (P.x - hf.x0) / hf.dx, fy = (P.y - hf.y0) / hf.dy;
    if (!(fx >= 0 && fy >= 0 && fx <= hf.nx - 1 && fy <= hf.ny - 1)) return false;
    const int ci = std::min(hf.nx - 2, (int)fx), cj = std::min(hf.ny - 2, (int)fy);
    V3 t[3];
    const bool upper = (fx - ci) + (fy - cj) > 1.0;
    if (!upper) t[0] = hf.vert(ci, cj), t[1] = hf.vert(ci + 1, cj), t[2] = hf.vert(ci, cj + 1);
    else t[0] = hf.vert(ci + 1, cj + 1), t[1] = hf.vert(ci, cj + 1), t[2] = hf.vert(ci + 1, cj);
    V3 tn = cross(t[1] - t[0], t[2] - t[0]);
    if (tn.z < 0) tn = -1.0 * tn;
    tn = (1.0 / std::sqrt(dot(tn, tn))) * tn;
    const double plane = dot(P - t[0], tn);
    if (plane < 0) {  // below the surface
        gap = plane - r;
        if (gap > margin) return false;
        n = tn;
        q = P - plane * tn;
        return true;
    }
    const double reach = r + margin;
    const int i0 = std::max(0, (int)std::floor((P.x - reach - hf.x0) / hf.dx));
    const int i1 = std::min(hf.nx - 2, (int)std::floor((P.x + reach - hf.x0) / hf.dx));
    const int j0 = std::max(0, (int)std::floor((P.y - reach - hf.y0) / hf.dy));
    const int j1 = std::min(hf.ny - 2, (int)std::floor((P.y + reach - hf.y0) / hf.dy));
    double best = 1e300;
    V3 bq{0, 0, 0};
    for (int j = j0; j <= j1; ++j)
        for (int i = i0; i <= i1; ++i) {
            const V3 v00 = hf.vert(i, j), v10 = hf.vert(i + 1, j), v01 = hf.vert(i, j + 1), v11 = hf.vert(i + 1, j + 1);
            const V3 c0 = closest_on_triangle(P, v00, v10, v01), c1 = closest_on_triangle(P, v11, v01, v10);
            const V3 e0 = P - c0, e1 = P - c1;
            const double q0 = dot(e0, e0), q1 = dot(e1, e1);
            if (q0 < best) best = q0, bq = c0;
            if (q1 < best) best = q1, bq = c1;
        }
    const double dist = std::sqrt(best);
    gap = dist - r;
    if (gap > margin) return false;
    n = dist > 1e-12 ? (1.0 / dist) * (P - bq) : tn;
    q = bq;
    return true;
}

// Probes of a body against the heightfield: sphere centre; capsule ends and midpoint; box corners.
void hf_contacts(int bi, const Body& B, const HField& hf, double margin, std::vector<RawContact>& out) {
    V3 probes[8];
    int np = 0;
    double r = 0;
    if (B.type == 0) {
        probes[np++] = B.c;
        r = B.h.x;
    } else if (B.type == 2) {
        capsule_ends(B, probes[0], probes[1]);
        probes[2] = B.c;
        np = 3;
        r = B.h.x;
    } else {
        for (int k = 0; k < 8; ++k)
            probes[np++] = B.c + (((k & 1) ? 1.0 : -1.0) * B.h.x) * col(B.R, 0) + (((k & 2) ? 1.0 : -1.0) * B.h.y) * col(B.R, 1) +
                           (((k & 4) ? 1.0 : -1.0) * B.h.z) * col(B.R, 2);
    }
    for (int k = 0; k < np; ++k) {
        V3 n, q;
        double gap;
        if (!hf_probe(hf, probes[k], r, margin, n, q, gap)) continue;
        out.push_back(RawContact{-1, bi, hf.wall, n, 0.5 * (q + (probes[k] - r * n)), std::max(0.0, -gap)});
    }
}

void pair_contacts(int ia, int ib, const Body& A, const Body& B, double margin, std::vector<RawContact>& out) {
    const V3 d = B.c - A.c;
    const double reach = A.bound + B.bound + margin;
    if (dot(d, d) > reach * reach) return;
    if (A.type == 2 || B.type == 2) {
        if (A.type == 2 && B.type == 2) return capsule_capsule(ia, ib, A, B, margin, out);
        if (A.type == 1 || B.type == 1) return capsule_box(ia, ib, A, B, margin, out);
        // capsule-sphere: the sphere against the nearest point of the segment
        V3 p0, p1;
        const Body& C = A.type == 2 ? A : B;
        const Body& S = A.type == 2 ? B : A;
        capsule_ends(C, p0, p1);
        const V3 pc = closest_on_segment(p0, p1, S.c);
        if (A.type == 2) return ball_ball(ia, ib, pc, C.h.x, S.c, S.h.x, margin, out);
        return ball_ball(ia, ib, S.c, S.h.x, pc, C.h.x, margin, out);
    }
    if (A.type == 0 && B.type == 0) {
        const double dist = std::sqrt(dot(d, d));
        const double gap = dist - A.h.x - B.h.x;
        if (gap > margin) return;
        const V3 n = dist > 0 ? (1.0 / dist) * d : V3{0, 0, 1};
        out.push_back(RawContact{ia, ib, -1, n, 0.5 * ((A.c + A.h.x * n) + (B.c - B.h.x * n)), std::max(0.0, -gap)});
    } else if (A.type == 1 && B.type == 1) {
        box_box(ia, ib, A, B, margin, out);
    } else {
        const bool a_sphere = A.type == 0;
        V3 n, pbox, psph;
        double gap;
        if (!sphere_box(a_sphere ? A : B, a_sphere ? B : A, margin, n, pbox, psph, gap)) return;
        // n points from the box to the sphere; contact normals run from A to B
        const V3 nab = a_sphere ? -1.0 * n : n;
        out.push_back(RawContact{ia, ib, -1, nab, 0.5 * (pbox + psph), std::max(0.0, -gap)});
    }
}

}  // namespace

extern "C" {

int cdx_shape_contacts(int nscenes, const int* body_off, const double* q, const int* shape, const double* dims,
                       const double* mu_body, int nw, const double* wall_n, const double* wall_d, const double* mu_wall,
                       double margin, int cap, int* body_a, int* body_b, double* mu, double* frame, double* angA,
                       double* angB, double* point, double* depth, int* contact_off) {
    return cdx_shape_contacts_hf(nscenes, body_off, q, shape, dims, mu_body, nw, wall_n, wall_d, mu_wall, nullptr, 0, nullptr,
                                 nullptr, margin, cap, body_a, body_b, mu, frame, angA, angB, point, depth, contact_off);
}

int cdx_shape_contacts_hf(int nscenes, const int* body_off, const double* q, const int* shape, const double* dims,
                          const double* mu_body, int nw, const double* wall_n, const double* wall_d, const double* mu_wall,
                          const cdx_heightfield* terrain, int ndisabled, const int* dis_a, const int* dis_b, double margin,
                          int cap, int* body_a, int* body_b, double* mu, double* frame, double* angA, double* angB,
                          double* point, double* depth, int* contact_off) {
    if (nscenes < 1 || !body_off || nw < 0 || ndisabled < 0 || (ndisabled > 0 && (!dis_a || !dis_b))) return -1;
    // pairs that never collide (CollisionCoal::disablePair / isPairDisabled, collision_coal.cpp:558-570)
    std::vector<uint64_t> disabled(ndisabled);
    for (int k = 0; k < ndisabled; ++k) {
        const uint64_t lo = (uint32_t)std::min(dis_a[k], dis_b[k]), hi = (uint32_t)std::max(dis_a[k], dis_b[k]);
        disabled[k] = (lo << 32) | hi;
    }
    std::sort(disabled.begin(), disabled.end());
    auto is_disabled = [&](int a, int b) {
        return !disabled.empty() &&
               std::binary_search(disabled.begin(), disabled.end(), ((uint64_t)(uint32_t)a << 32) | (uint32_t)b);
    };
    const int nb = body_off[nscenes];
    for (int i = 0; i < nb; ++i)
        if (shape[i] != 0 && shape[i] != 1 && shape[i] != 2) return -1;
    HField hf;
    std::vector<double> mu_static(mu_wall, mu_wall + nw);
    if (terrain) {
        if (terrain->nx < 2 || terrain->ny < 2 || !(terrain->dx > 0) || !(terrain->dy > 0) || !terrain->h) return -1;
        hf.nx = terrain->nx, hf.ny = terrain->ny, hf.x0 = terrain->x0, hf.y0 = terrain->y0;
        hf.dx = terrain->dx, hf.dy = terrain->dy, hf.h = terrain->h;
        hf.wall = nw;
        mu_static.push_back(terrain->mu);
    }
    std::vector<Body> bodies(nb);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nb; ++i) bodies[i] = make_body(q + 7 * (size_t)i, shape[i], dims + 3 * (size_t)i);
    // contacts of scene s, in body order (walls of body a, then its pairs (a, b > a) by b)
    std::vector<std::vector<RawContact>> per(nscenes);
#pragma omp parallel
    {
        std::vector<int> nbrs;
        std::vector<double> cen;
#pragma omp for schedule(dynamic, 1)
        for (int s = 0; s < nscenes; ++s) {
            const int b0 = body_off[s], b1 = body_off[s + 1];
            if (b1 <= b0) continue;
            double bmax = 0;
            cen.resize(3 * (size_t)(b1 - b0));
            for (int i = b0; i < b1; ++i) {
                bmax = std::max(bmax, bodies[i].bound);
                cen[3 * (i - b0)] = bodies[i].c.x, cen[3 * (i - b0) + 1] = bodies[i].c.y, cen[3 * (i - b0) + 2] = bodies[i].c.z;
            }
            Grid g;
            g.build(b1 - b0, cen.data(), 3, std::max(2 * bmax + margin, 1e-9));
            auto& out = per[s];
            for (int a = b0; a < b1; ++a) {
                for (int w = 0; w < nw; ++w)
                    wall_contacts(a, w, bodies[a], V3{wall_n[3 * w], wall_n[3 * w + 1], wall_n[3 * w + 2]}, wall_d[w], margin,
                                  out);
                if (hf.h) hf_contacts(a, bodies[a], hf, margin, out);
                nbrs.clear();
                g.around(&cen[3 * (a - b0)], [&](int j) {
                    if (j + b0 > a) nbrs.push_back(j + b0);
                });
                std::sort(nbrs.begin(), nbrs.end());
                for (int b : nbrs)
                    if (!is_disabled(a, b)) pair_contacts(a, b, bodies[a], bodies[b], margin, out);
            }
        }
    }
    std::vector<int> off(nscenes + 1, 0);
    for (int s = 0; s < nscenes; ++s) off[s + 1] = off[s] + (int)per[s].size();
    const int total = off[nscenes];
    if (contact_off)
        for (int s = 0; s <= nscenes; ++s) contact_off[s] = off[s];
    if (total > cap) return total;
#pragma omp parallel for schedule(dynamic, 16)
    for (int s = 0; s < nscenes; ++s)
        for (size_t k = 0; k < per[s].size(); ++k)
            fill_contact(per[s][k], q, mu_body, mu_static.data(), off[s] + (int)k, body_a, body_b, mu, frame, angA, angB, point,
                         depth);
    return total;
}


int cdx_scene_rsa(int nscenes, const int* body_off, const double* bound, const double* lo, const double* hi, uint64_t seed,
                  double* pos) {
    if (nscenes < 1 || !body_off) return -1;
#pragma omp parallel for schedule(dynamic, 1)
    for (int s = 0; s < nscenes; ++s) {
        std::mt19937_64 rng(seed * 1000003ULL + (uint64_t)s);
        std::uniform_real_distribution<double> U(0.0, 1.0);
        const int b0 = body_off[s], b1 = body_off[s + 1];
        std::vector<int> order(b1 - b0);
        std::iota(order.begin(), order.end(), b0);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return bound[a] > bound[b]; });
        std::vector<int> placed;
        double top = hi[2];
        for (int i : order) {
            const double r = bound[i];
            int tries = 0;
            for (;;) {
                const double p[3] = {lo[0] + r + U(rng) * std::max(0.0, hi[0] - lo[0] - 2 * r),
                                     lo[1] + r + U(rng) * std::max(0.0, hi[1] - lo[1] - 2 * r),
                                     lo[2] + r + U(rng) * std::max(0.0, top - lo[2] - 2 * r)};
                bool ok = true;
                for (int j : placed) {
                    const double dx = p[0] - pos[3 * j], dy = p[1] - pos[3 * j + 1], dz = p[2] - pos[3 * j + 2];
                    const double rr = r + bound[j];
                    if (dx * dx + dy * dy + dz * dz < rr * rr) {
                        ok = false;
                        break;
                    }
                }
                if (ok) {
                    for (int k = 0; k < 3; ++k) pos[3 * i + k] = p[k];
                    placed.push_back(i);
                    break;
                }
                if (++tries % 200 == 0) top *= 1.1;  // column grows when the box is too full
            }
        }
    }
    return 0;
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
