// gnp_geom.cpp -- nanoplatelet geometry for the workload generator (host C++).
//
// Closest points on a plate, plate-plate distance, the crossings of a plate with the
// window faces and the in-plane Delaunay triangulation of a plate's points.  The
// triangulation follows the reference's Bowyer-Watson steps (Triangulation.cpp:139-475)
// so that the edge list comes out in the same order: the assembly accumulates the
// diagonal edge by edge, and only the same order gives the same bits.
#include "geom3.h"

#include <cmath>

namespace cicy_geom {

namespace {

constexpr double kZero = 1.0e-10;  // the reference's Zero (Hns.h)

// closest point of the mid-plane square of g to p (local clamp with w = 0)
double point_square(const Plate &g, const double *p, double *q) {
    double loc[3];
    to_local(g, p, loc);
    const double h = 0.5 * g.l;
    const double ql[3] = {std::clamp(loc[0], -h, h), std::clamp(loc[1], -h, h), 0.0};
    to_global(g, ql, q);
    const double d[3] = {p[0] - q[0], p[1] - q[1], p[2] - q[2]};
    return dot3(d, d);
}

void square_corner(const Plate &g, int k, double *p) {
    // corners in cyclic order so that corner k and k+1 (mod 4) share an edge
    static const double su[4] = {-1, 1, 1, -1}, sv[4] = {-1, -1, 1, 1};
    const double q[3] = {0.5 * g.l * su[k], 0.5 * g.l * sv[k], 0.0};
    to_global(g, q, p);
}

// does segment p0->p1 cross the mid-plane square of g?  x receives the crossing point
bool segment_crosses_square(const Plate &g, const double *p0, const double *p1, double *x) {
    const double d0[3] = {p0[0] - g.c[0], p0[1] - g.c[1], p0[2] - g.c[2]};
    const double d1[3] = {p1[0] - g.c[0], p1[1] - g.c[1], p1[2] - g.c[2]};
    const double w0 = dot3(d0, g.ax[2]), w1 = dot3(d1, g.ax[2]);
    if ((w0 > 0 && w1 > 0) || (w0 < 0 && w1 < 0) || w0 == w1) return false;
    const double s = w0 / (w0 - w1);
    for (int a = 0; a < 3; a++) x[a] = p0[a] + s * (p1[a] - p0[a]);
    double loc[3];
    to_local(g, x, loc);
    const double h = 0.5 * g.l;
    return std::fabs(loc[0]) <= h && std::fabs(loc[1]) <= h;
}

}  // namespace

bool plate_closest_point(const Plate &g, const double *p, double *q, double *dist) {
    double loc[3];
    to_local(g, p, loc);
    const double h[3] = {0.5 * g.l, 0.5 * g.l, 0.5 * g.t};
    bool inside = true;
    double ql[3];
    for (int k = 0; k < 3; k++) {
        if (!(std::fabs(loc[k]) < h[k])) inside = false;
        ql[k] = std::clamp(loc[k], -h[k], h[k]);
    }
    to_global(g, ql, q);
    const double d[3] = {p[0] - q[0], p[1] - q[1], p[2] - q[2]};
    *dist = std::sqrt(dot3(d, d));
    return !inside;
}

double midplane_distance(const Plate &a, const Plate &b, double *pa, double *pb) {
    double ca[4][3], cb[4][3];
    for (int k = 0; k < 4; k++) {
        square_corner(a, k, ca[k]);
        square_corner(b, k, cb[k]);
    }
    // crossing squares: an edge of one passes through the other
    for (int k = 0; k < 4; k++) {
        double x[3];
        if (segment_crosses_square(b, ca[k], ca[(k + 1) & 3], x) ||
            segment_crosses_square(a, cb[k], cb[(k + 1) & 3], x)) {
            for (int q = 0; q < 3; q++) pa[q] = pb[q] = x[q];
            return 0.0;
        }
    }
    double best = INFINITY;
    double q[3];
    for (int k = 0; k < 4; k++) {
        double d2 = point_square(b, ca[k], q);
        if (d2 < best) {
            best = d2;
            for (int r = 0; r < 3; r++) { pa[r] = ca[k][r]; pb[r] = q[r]; }
        }
        d2 = point_square(a, cb[k], q);
        if (d2 < best) {
            best = d2;
            for (int r = 0; r < 3; r++) { pa[r] = q[r]; pb[r] = cb[k][r]; }
        }
    }
    for (int i = 0; i < 4; i++) {
        double e1[3];
        for (int r = 0; r < 3; r++) e1[r] = ca[(i + 1) & 3][r] - ca[i][r];
        for (int j = 0; j < 4; j++) {
            double e2[3];
            for (int r = 0; r < 3; r++) e2[r] = cb[(j + 1) & 3][r] - cb[j][r];
            double s, t;
            const double d2 = seg_seg(ca[i], e1, cb[j], e2, s, t);
            if (d2 < best) {
                best = d2;
                for (int r = 0; r < 3; r++) {
                    pa[r] = ca[i][r] + s * e1[r];
                    pb[r] = cb[j][r] + t * e2[r];
                }
            }
        }
    }
    return std::sqrt(best);
}

void plate_window_crossings(const Plate &g, double box, std::vector<double> acc[6]) {
    // window faces by (axis, side): x_min 4, x_max 2, y_min 5, y_max 3, z_min 1, z_max 0
    static const int face[3][2] = {{4, 2}, {5, 3}, {1, 0}};
    double v[8][3];
    for (int k = 0; k < 8; k++) plate_vertex(g, k, v[k]);
    for (int k = 0; k < 8; k++)
        for (int bit = 1; bit < 8; bit <<= 1) {
            if (k & bit) continue;
            const double *p0 = v[k], *p1 = v[k | bit];
            // clip the edge to the cube (slab method), remembering which plane cut it
            double t0 = 0.0, t1 = 1.0;
            int ent_ax = -1, ent_side = 0, ext_ax = -1, ext_side = 0;
            bool hit = true;
            for (int a = 0; a < 3 && hit; a++) {
                const double d = p1[a] - p0[a];
                if (std::fabs(d) < 1e-300) {
                    if (p0[a] < 0.0 || p0[a] > box) hit = false;
                    continue;
                }
                double ta = (0.0 - p0[a]) / d, tb = (box - p0[a]) / d;
                int sa = 0, sb = 1;
                if (ta > tb) {
                    std::swap(ta, tb);
                    std::swap(sa, sb);
                }
                if (ta > t0) { t0 = ta; ent_ax = a; ent_side = sa; }
                if (tb < t1) { t1 = tb; ext_ax = a; ext_side = sb; }
                if (t0 > t1) hit = false;
            }
            if (!hit) continue;
            const int axs[2] = {ent_ax, ext_ax}, sides[2] = {ent_side, ext_side};
            const double ts[2] = {t0, t1};
            for (int e = 0; e < 2; e++) {
                if (axs[e] < 0) continue;  // this end of the edge lies inside the cube
                double x[3];
                for (int a = 0; a < 3; a++) x[a] = p0[a] + ts[e] * (p1[a] - p0[a]);
                x[axs[e]] = sides[e] ? box : 0.0;
                std::vector<double> &f = acc[face[axs[e]][sides[e]]];
                f.insert(f.end(), x, x + 3);
            }
        }
}

void plate_triangulation(double l, const std::vector<double> &uv,
                         std::vector<std::pair<int, int>> &edges) {
    edges.clear();
    const int n = (int)(uv.size() / 2);
    if (n <= 4) {
        for (int i = 0; i + 1 < n; i++)
            for (int j = i + 1; j < n; j++) edges.emplace_back(i, j);
        return;
    }
    // super triangle of Triangulation::Generate_supertriangle (:199-236), plate frame
    const double s3 = std::sqrt(3.0);
    const double sx = (0.75 + 0.5 * s3) * l;
    const double sup[3][2] = {{0.0, (0.75 * s3 + 0.5) * l}, {sx, -l}, {-sx, -l}};
    // vertex ids: >= 0 the points, -1..-3 the super triangle
    auto pt = [&](int id, double &x, double &y) {
        if (id >= 0) {
            x = uv[2 * (size_t)id];
            y = uv[2 * (size_t)id + 1];
        } else {
            x = sup[-1 - id][0];
            y = sup[-1 - id][1];
        }
    };
    struct Tri { int a, b, c; };
    struct Ed { int a, b; };
    auto same = [](const Ed &x, const Ed &y) {
        return (x.a == y.a && x.b == y.b) || (x.a == y.b && x.b == y.a);
    };
    std::vector<Tri> tris{{-1, -2, -3}};
    std::vector<Ed> bad;
    std::vector<int> rep;
    for (int i = 0; i < n; i++) {
        double px, py;
        pt(i, px, py);
        bad.clear();
        // triangles whose circumcircle holds the point, scanned from the back and removed
        for (int j = (int)tris.size() - 1; j >= 0; j--) {
            double x1, y1, x2, y2, x3, y3;
            pt(tris[j].a, x1, y1);
            pt(tris[j].b, x2, y2);
            pt(tris[j].c, x3, y3);
            const double ax = x1 - x3, ay = y1 - y3, bx = x2 - x3, by = y2 - y3;
            const double an = ax * ax + ay * ay, bn = bx * bx + by * by;
            const double dd = 2.0 * (ax * by - ay * bx);
            const double cx = x3 + (an * by - bn * ay) / dd;
            const double cy = y3 + (bn * ax - an * bx) / dd;
            const double r2 = (x1 - cx) * (x1 - cx) + (y1 - cy) * (y1 - cy);
            const double d2 = (px - cx) * (px - cx) + (py - cy) * (py - cy);
            if (d2 - r2 <= kZero) {
                bad.push_back({tris[j].a, tris[j].b});
                bad.push_back({tris[j].b, tris[j].c});
                bad.push_back({tris[j].c, tris[j].a});
                tris.erase(tris.begin() + j);
            }
        }
        // the boundary of the cavity: edges of the removed triangles seen only once
        rep.assign(bad.size(), 0);
        for (size_t j = 0; j + 1 < bad.size(); j++)
            for (size_t k = j + 1; k < bad.size(); k++)
                if (same(bad[j], bad[k])) {
                    rep[j]++;
                    rep[k]++;
                }
        for (size_t j = 0; j < bad.size(); j++)
            if (!rep[j]) tris.push_back({bad[j].a, bad[j].b, i});
    }
    // edges not touching the super triangle, each kept at its first appearance
    std::vector<Ed> all;
    for (const Tri &t : tris) {
        if (t.a >= 0 && t.b >= 0) all.push_back({t.a, t.b});
        if (t.b >= 0 && t.c >= 0) all.push_back({t.b, t.c});
        if (t.c >= 0 && t.a >= 0) all.push_back({t.c, t.a});
    }
    rep.assign(all.size(), 0);
    for (size_t j = 0; j + 1 < all.size(); j++)
        for (size_t k = j + 1; k < all.size(); k++)
            if (same(all[j], all[k])) rep[k]++;
    for (size_t j = 0; j < all.size(); j++)
        if (!rep[j]) edges.emplace_back(all[j].a, all[j].b);
}

}  // namespace cicy_geom
