// percolation.cpp -- clusters of tubes, their percolation families and junction groups, from
// the labels and junctions of the contact search (host only: O(tubes + junctions) integer
// bookkeeping).
//
// Follows Hoshen_Kopelman.cpp:
//   Make_particle_clusters                   :2331-2364  tubes of each label, window order;
//                                                        unlabelled tubes are isolated
//   Find_percolated_clusters                 :2366-2396  boundary pairs {2,4} x, {3,5} y, {0,1} z
//   Find_clusters_connected_to_boundaries    :2398-2468
//   Add_clusters_in_boundary                 :2470-2495
//   Add_percolated_direction                 :2497-2522  direction d adds 10 d + 1
//   Determine_family_of_percolated_clusters  :2524-2589  percolated clusters numbered in label order
//   Family_map                               :2591-2622  1 x, 11 y, 21 z, 12 xy, 22 xz, 32 yz, 33 xyz
//   Group_junctions_same_particle            :2672-2705  junction -> its percolated cluster
#include "cicy_clusters.h"

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <vector>

namespace {

int fail(char *err, int64_t len, const char *fmt, ...) {
    if (err && len > 0) {
        va_list ap;
        va_start(ap, fmt);
        std::vsnprintf(err, (size_t)len, fmt, ap);
        va_end(ap);
    }
    return 0;
}

int family_of(int sum) {
    switch (sum) {
        case 1: return 0;
        case 11: return 1;
        case 21: return 2;
        case 12: return 3;
        case 22: return 4;
        case 32: return 5;
        case 33: return 6;
        default: return -1;
    }
}

}  // namespace

extern "C" int cicy_cnt_percolation(int64_t n_cnts, const int32_t *labels, int64_t n_labels,
                                    int64_t n_inside, const int64_t *cnts_inside,
                                    const int64_t *boundary_off, const int64_t *boundary_cnt,
                                    int64_t n_junc, const int64_t *junc, int32_t *family,
                                    int32_t *perc_index, int64_t *n_perc, int64_t *cluster_off,
                                    int64_t *cluster_cnts, int64_t *n_isolated,
                                    int64_t *isolated, int32_t *junc_cluster, char *err,
                                    int64_t err_len) {
    if (n_cnts < 0 || n_labels < 0 || n_inside < 0 || n_junc < 0 || (n_cnts && !labels) ||
        (n_inside && !cnts_inside) || !boundary_off || (n_labels && (!family || !perc_index)) ||
        !n_perc || !cluster_off || !n_isolated || (n_junc && (!junc || !junc_cluster)))
        return fail(err, err_len, "cicy_cnt_percolation: invalid arguments");
    for (int64_t i = 0; i < n_cnts; i++)
        if (labels[i] < -1 || labels[i] >= n_labels)
            return fail(err, err_len, "cicy_cnt_percolation: label %d of tube %lld out of range",
                        labels[i], (long long)i);
    if (boundary_off[0] != 0)
        return fail(err, err_len, "cicy_cnt_percolation: boundary offsets must start at 0");
    for (int b = 0; b < 6; b++)
        if (boundary_off[b + 1] < boundary_off[b])
            return fail(err, err_len, "cicy_cnt_percolation: boundary offsets decrease");
    for (int64_t q = 0; q < boundary_off[6]; q++)
        if (boundary_cnt[q] < 0 || boundary_cnt[q] >= n_cnts)
            return fail(err, err_len, "cicy_cnt_percolation: boundary tube %lld out of range",
                        (long long)boundary_cnt[q]);
    // Make_particle_clusters: tubes of each label in window order; the rest are isolated
    std::vector<int64_t> cnt((size_t)n_labels + 1, 0);
    int64_t niso = 0;
    for (int64_t k = 0; k < n_inside; k++) {
        const int64_t t = cnts_inside[k];
        if (t < 0 || t >= n_cnts)
            return fail(err, err_len, "cicy_cnt_percolation: cnts_inside[%lld] out of range",
                        (long long)k);
        if (labels[t] == -1)
            niso++;
        else
            cnt[(size_t)labels[t] + 1]++;
    }
    cluster_off[0] = 0;
    for (int64_t L = 0; L < n_labels; L++) cluster_off[L + 1] = cluster_off[L] + cnt[(size_t)L + 1];
    if (cluster_cnts || isolated) {
        std::vector<int64_t> fill(cluster_off, cluster_off + n_labels);
        int64_t w = 0;
        for (int64_t k = 0; k < n_inside; k++) {
            const int64_t t = cnts_inside[k];
            if (labels[t] == -1) {
                if (isolated) isolated[w] = t;
                w++;
            } else if (cluster_cnts) {
                cluster_cnts[fill[(size_t)labels[t]]++] = t;
            }
        }
    }
    *n_isolated = niso;
    // Find_clusters_connected_to_boundaries: direction d percolates a cluster that has tubes
    // on both faces of d
    static const int pairs[3][2] = {{2, 4}, {3, 5}, {0, 1}};
    std::vector<int> dirs((size_t)n_labels, 0);  // bit d set: 10 d + 1 in the reference's set
    std::vector<char> on1((size_t)n_labels);
    for (int d = 0; d < 3; d++) {
        std::fill(on1.begin(), on1.end(), 0);
        bool any = false;
        const int b1 = pairs[d][0], b2 = pairs[d][1];
        for (int64_t q = boundary_off[b1]; q < boundary_off[b1 + 1]; q++) {
            const int L = labels[boundary_cnt[q]];
            if (L != -1) {
                on1[(size_t)L] = 1;
                any = true;
            }
        }
        if (!any) continue;
        for (int64_t q = boundary_off[b2]; q < boundary_off[b2 + 1]; q++) {
            const int L = labels[boundary_cnt[q]];
            if (L != -1 && on1[(size_t)L]) dirs[(size_t)L] |= 1 << d;
        }
    }
    // Determine_family_of_percolated_clusters + Family_map
    int64_t np = 0;
    for (int64_t L = 0; L < n_labels; L++) {
        if (!dirs[(size_t)L]) {
            family[L] = -1;
            perc_index[L] = -1;
            continue;
        }
        int sum = 0;
        for (int d = 0; d < 3; d++)
            if (dirs[(size_t)L] >> d & 1) sum += 10 * d + 1;
        family[L] = family_of(sum);
        perc_index[L] = (int32_t)np++;
    }
    *n_perc = np;
    // Group_junctions_same_particle: every junction joins tubes of one cluster
    for (int64_t q = 0; q < n_junc; q++) {
        const int64_t a = junc[4 * q + 1], b = junc[4 * q + 3];
        if (a < 0 || a >= n_cnts || b < 0 || b >= n_cnts)
            return fail(err, err_len, "cicy_cnt_percolation: junction %lld joins a tube out of "
                        "range", (long long)q);
        const int L1 = labels[a], L2 = labels[b];
        if (L1 != L2)
            return fail(err, err_len, "cicy_cnt_percolation: junction %lld joins tubes of "
                        "different clusters", (long long)q);
        junc_cluster[q] = L1 >= 0 ? perc_index[L1] : -1;
    }
    return 1;
}
