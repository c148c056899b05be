// geom3.h -- small 3-D geometry helpers shared by the workload generator (host C++).
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <utility>
#include <vector>

namespace cicy_geom {

inline double dot3(const double *a, const double *b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// Closest points of segments P0+s*d1 (s in [0,1]) and Q0+t*d2 (t in [0,1]); returns the
// squared distance.  Derived from minimising |P(s)-Q(t)|^2 over the unit square.
inline double seg_seg(const double *p0, const double *d1, const double *q0, const double *d2,
                      double &s, double &t) {
    double r[3] = {p0[0] - q0[0], p0[1] - q0[1], p0[2] - q0[2]};
    double a = dot3(d1, d1);
    double e = dot3(d2, d2);
    double f = dot3(d2, r);
    double c = dot3(d1, r);
    double bb = dot3(d1, d2);
    const double eps = 1e-30;
    if (a <= eps && e <= eps) {
        s = t = 0.0;
    } else if (a <= eps) {
        s = 0.0;
        t = std::clamp(f / e, 0.0, 1.0);
    } else if (e <= eps) {
        t = 0.0;
        s = std::clamp(-c / a, 0.0, 1.0);
    } else {
        double denom = a * e - bb * bb;
        s = denom > 1e-14 * a * e ? std::clamp((bb * f - c * e) / denom, 0.0, 1.0) : 0.0;
        t = (bb * s + f) / e;
        if (t < 0.0) {
            t = 0.0;
            s = std::clamp(-c / a, 0.0, 1.0);
        } else if (t > 1.0) {
            t = 1.0;
            s = std::clamp((bb - c) / a, 0.0, 1.0);
        }
    }
    double dx = p0[0] + d1[0] * s - (q0[0] + d2[0] * t);
    double dy = p0[1] + d1[1] * s - (q0[1] + d2[1] * t);
    double dz = p0[2] + d1[2] * s - (q0[2] + d2[2] * t);
    return dx * dx + dy * dy + dz * dz;
}

// A graphene nanoplatelet: a thin square cuboid, side l, thickness t.  ax[0], ax[1] are
// the unit in-plane axes and ax[2] the unit normal of its large face (the reference's
// face 0, GNP struct in Geometry_3D.h).
struct Plate {
    double c[3];
    double ax[3][3];
    double l, t;
};

inline void to_local(const Plate &g, const double *p, double *q) {
    const double d[3] = {p[0] - g.c[0], p[1] - g.c[1], p[2] - g.c[2]};
    for (int k = 0; k < 3; k++) q[k] = dot3(d, g.ax[k]);
}

inline void to_global(const Plate &g, const double *q, double *p) {
    for (int a = 0; a < 3; a++)
        p[a] = g.c[a] + q[0] * g.ax[0][a] + q[1] * g.ax[1][a] + q[2] * g.ax[2][a];
}

// Vertex k (0..7) of the cuboid; bit 0/1/2 of k selects the sign along ax[0]/ax[1]/ax[2].
inline void plate_vertex(const Plate &g, int k, double *p) {
    const double q[3] = {(k & 1) ? 0.5 * g.l : -0.5 * g.l, (k & 2) ? 0.5 * g.l : -0.5 * g.l,
                         (k & 4) ? 0.5 * g.t : -0.5 * g.t};
    to_global(g, q, p);
}

// Closest point of the cuboid to p (clamping in the plate frame) and its distance.
// Returns false when p lies strictly inside the cuboid (the reference's flag -1 case of
// Generate_Network::Get_gnp_point_closest_to_point, Generate_Network.cpp:5577).
bool plate_closest_point(const Plate &g, const double *p, double *q, double *dist);

// Distance between the mid-plane squares of two plates, with the closest points on each
// (0 when the squares cross).
double midplane_distance(const Plate &a, const Plate &b, double *pa, double *pb);

// Points where the cuboid's twelve edges cross the faces of the cube [0,box]^3, grouped
// by face (window face numbering of Cutoff_Wins::Where_is_with_boundary); acc[f] holds
// the crossings on face f.
void plate_window_crossings(const Plate &g, double box, std::vector<double> acc[6]);

// Delaunay edges of points given in the plate's in-plane coordinates (u, v), built with
// the Bowyer-Watson steps of Triangulation.cpp:139-475 (super triangle around the plate,
// points inserted in the given order, circumcircle test with the reference's 1e-10
// tolerance, edges touching the super triangle dropped, repeated edges kept once).
// Up to four points give every pair (Generate_trivial_triangulation, :116-137).  Edges are
// returned as local index pairs in the reference's final order.
void plate_triangulation(double l, const std::vector<double> &uv,
                         std::vector<std::pair<int, int>> &edges);

}  // namespace cicy_geom
