/* cicy_clusters.h -- tube clusters, percolation families and junction groups (libcicy_net.so,
 * host only).  Returns 1 on success, 0 with a message in err[err_len].
 *
 * cicy_cnt_percolation replaces the tube part of what Hoshen_Kopelman::
 * Determine_clusters_and_percolation does after labelling (Hoshen_Kopelman.cpp:124-202):
 * Make_particle_clusters (:2331-2364), Find_percolated_clusters (:2366-2396) with its helpers
 * (:2398-2522), Determine_family_of_percolated_clusters (:2524-2589), Family_map (:2591-2622)
 * and Group_junctions_same_particle (:2672-2705).
 *
 * In:  labels[n_cnts] and n_labels as cicy_cnt_contacts returns them (labels_cnt);
 *      cnts_inside[n_inside], the window's tubes in order;
 *      the six boundary lists (the reference's boundary_cnt, boundaries numbered as
 *      Cutoff_Wins does: x pairs 4 and 2, y pairs 5 and 3, z pairs 1 and 0) as
 *      boundary_cnt[boundary_off[b] .. boundary_off[b + 1]], b = 0..5;
 *      junctions junc[4 q..] = (Pi, tube i, Pj, tube j), q < n_junc (junctions_cnt).
 * Out: family[n_labels]: 0 x, 1 y, 2 z, 3 xy, 4 xz, 5 yz, 6 xyz, -1 not percolated;
 *      perc_index[n_labels]: index of the label among percolated clusters (label order), -1;
 *      *n_perc; cluster_off[n_labels + 1] into cluster_cnts (tubes of each label in window
 *      order, clusters_cnt_tmp; may be null); *n_isolated and isolated (window tubes without
 *      label, the first vector of isolated_cnt; may be null); junc_cluster[n_junc]: the
 *      percolated cluster of each junction or -1 (cluster_cnt_junctions).  A junction between
 *      tubes of different labels is an error, as in the reference. */
#ifndef CICY_CLUSTERS_H
#define CICY_CLUSTERS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

int cicy_cnt_percolation(int64_t n_cnts, const int32_t *labels, int64_t n_labels,
                         int64_t n_inside, const int64_t *cnts_inside,
                         const int64_t *boundary_off, const int64_t *boundary_cnt,
                         int64_t n_junc, const int64_t *junc, int32_t *family,
                         int32_t *perc_index, int64_t *n_perc, int64_t *cluster_off,
                         int64_t *cluster_cnts, int64_t *n_isolated, int64_t *isolated,
                         int32_t *junc_cluster, char *err, int64_t err_len);

#ifdef __cplusplus
}
#endif

#endif
