/* cicy_netio.h -- the reference's network files (libcicy_net.so, host only).
 *
 * Each call replaces one reader of Read_Network (Read_Network.cpp) or writer of Printer
 * (Printer.cpp) and reads or writes the same files, in `dir` instead of the current
 * directory.  Return 1 on success, 0 on failure with a message in err[err_len].  Where the
 * reference would read a short or malformed file into zeros or stale values these calls fail
 * with the file and line.  fmt: 0 = csv, 1 = dat (binary). */
#ifndef CICY_NETIO_H
#define CICY_NETIO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Read_Network::Read_sample_geometry (Read_Network.cpp:153-199): sample_geom.csv.
 * out[0..2] = corner (poi_min), out[3..5] = len_x, wid_y, hei_z. */
int cicy_read_sample_geometry(const char *dir, double *out, char *err, int64_t err_len);

/* Read_Network::Read_cnt_data_from_csv (:201-344) / Read_cnt_data_from_dat (:346-462):
 * cnt_struct.* and cnt_coordinates.*.  With npts, radii and xyz null it sets only n_cnts
 * and n_points; otherwise it fills npts[n_cnts] (points per tube), radii[n_cnts] and
 * xyz[3 n_points], tube by tube (the reference's structure[i] is the next npts[i] points). */
int cicy_read_cnts(const char *dir, int fmt, int64_t *n_cnts, int64_t *n_points, int64_t *npts,
                   double *radii, double *xyz, char *err, int64_t err_len);

/* Read_Network::Read_gnp_data_from_csv (:464-571) / _from_dat (:573-678): gnp_data.*.
 * *n_gnps is the room in geo/rot on entry and the plate count on return; with geo and rot
 * null only the count is returned.  geo[7 i..] = l, t, theta, phi, cx, cy, cz;
 * rot[9 i..] = Generate_Network::Get_transformation_matrix(theta, phi), row major
 * (Generate_Network.cpp:3466-3490). */
int cicy_read_gnps(const char *dir, int fmt, int64_t *n_gnps, double *geo, double *rot,
                   char *err, int64_t err_len);

/* Generate_Network's file output (Generate_Network.cpp:6118-6180) with
 * Printer::Print_cnt_points_and_structure[_binary] (Printer.cpp:273-380) and
 * Printer::Print_gnp_data[_binary] (:140-243): sample[6] as read above; tubes as read above;
 * plates as lt[2 i..] = (side, thickness), rot[9 i..], centre[3 i..].  The angles written
 * are recovered from rot by Printer::Recover_angle (:170-202). */
int cicy_write_network_files(const char *dir, int fmt, int prec_cnt, int prec_gnp,
                             const double *sample, int64_t n_cnts, const int64_t *npts,
                             const double *radii, const double *xyz, int64_t n_gnps,
                             const double *lt, const double *rot, const double *centre,
                             char *err, int64_t err_len);

#ifdef __cplusplus
}
#endif

#endif
