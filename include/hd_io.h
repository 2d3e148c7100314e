/* Point-set loaders of the reference's input formats (host-side, no GPU).
 *
 * Replaces src/loaders/loader.h:26-234 (WKT via Boost.Geometry, OFF, the
 * `.bin` point cache), src/loaders/ply_loader.h:37-101 (PLY via rply) and the
 * input-type switch of src/run_hausdorff_distance.cu:91-112. Points are
 * returned as AoS FP32 (n_dims floats per point), the layout every entry
 * point of hd.h takes. */
#ifndef HD_IO_H
#define HD_IO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Input types (src/input_type.h:4, plus the cache file itself). */
enum {
  HD_INPUT_WKT = 0, /* one geometry per line: POINT, LINESTRING, POLYGON
                       (outer ring), MULTILINESTRING, MULTIPOLYGON (outer
                       rings); GEOMETRYCOLLECTION lines are skipped */
  HD_INPUT_OFF = 1, /* "OFF", counts, then n_dims numbers per vertex */
  HD_INPUT_PLY = 2, /* ascii / binary_little_endian / binary_big_endian;
                       vertex x, y (, z); each axis shifted so that its
                       minimum is 0, in FP64, then rounded to FP32 */
  HD_INPUT_BIN = 3  /* uint64 count, then count * n_dims FP32 */
};

/* Loads at most `limit` points of `path`. With a non-empty
 * serialize_folder (WKT and OFF only, as the reference), the points are
 * cached there as "points_<path with '/' -> '_'>_limit_<limit>.bin" and a
 * later call reads the cache instead of parsing (loader.h:176-234); the
 * folder is created if missing. *out receives a malloc'd array of
 * n_dims * *n_out floats, released with hd_free_points. Returns 0 or a
 * negative HD_E* code (hd_last_error() has the message); a file with no
 * points is an error (run_hausdorff_distance.cu:114-115). */
int hd_load_points(const char* path, int input_type, int n_dims,
                   uint64_t limit, const char* serialize_folder, float** out,
                   uint64_t* n_out);
void hd_free_points(float* points);

/* In-place preprocessing of src/loaders/translate_points.h: every
 * coordinate mapped to (x - lower) / max_extent (:42-63, a zero extent
 * counts as 1), and dimension 0 shifted by extent * ratio (:65-80). */
int hd_normalize_points(float* points, uint64_t n, int n_dims);
int hd_translate_points(float* points, uint64_t n, int n_dims, int dim,
                        double ratio);

#ifdef __cplusplus
}
#endif

#endif /* HD_IO_H */
