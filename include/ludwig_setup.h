/*
 * ludwig_setup.h -- C ABI of the host-side table producers (libludwig_setup.so).
 *
 * These replace the CPU setup loops of OpenLUDWIG's Julia host that produce
 * the time step's inputs.  Each function cites the reference routine whose
 * results it reproduces bit for bit.  All arrays are caller-allocated unless
 * the function returns malloc'd buffers (release them with lbs_free).
 *
 * Conventions: block coordinates are 1-based (bx,by,bz) rows of an (nb,3)
 * int32 array in the level's sorted block order; `grid` is the level's dense
 * block pointer of dims (gdx,gdy,gdz), column-major, 1-based ids, 0 = none;
 * cell arrays are (8,8,8,nb) column-major as in src/geometry/blocks.jl.
 * Status: 0 = OK, negative = error (message in lbs_last_error()).
 */
#ifndef LUDWIG_SETUP_H
#define LUDWIG_SETUP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

const char* lbs_last_error(void);
void lbs_free(void* p);

/* Geometry.load_mesh / parse_binary_stl / parse_ascii_stl
 * (src/geometry/geometry.jl:385-520): triangles (n,3,3), winding-corrected. */
int lbs_stl_read(const char* path, double scale, double** tris_out, int64_t* n_out);

/* compute_geometry_properties (src/geometry/geometry.jl:75-103). */
int lbs_geometry_properties(const double* tris, int64_t n, double* normals, double* areas, double* centers);

/* subdivide_mesh / subdivide_triangle (src/geometry/geometry.jl:120-252):
 * longest-edge bisection; parent_out[i] = source triangle of output i.     */
int lbs_subdivide(const double* tris, int64_t n, double min_size, double** tris_out, int64_t** parent_out,
                  int64_t* n_out);

/* voxelize_blocks! + triangle_intersects_aabb (src/geometry/domain_generation.jl:10-112). */
int lbs_voxelize(int64_t nb, const int32_t* coords, const int32_t* grid, int gdx, int gdy, int gdz,
                 const double* tris, int64_t ntri, double dx, const double* offset, uint8_t* obst);

/* perform_flood_fill! (src/geometry/domain_generation.jl:114-203); returns the
 * number of interior cells filled. */
int64_t lbs_flood_fill(int64_t nb, const int32_t* coords, const int32_t* grid, int gdx, int gdy, int gdz,
                       uint8_t* obst);

/* apply_sponge! (src/geometry/domain_generation.jl:205-289). sponge_thickness is
 * the Float64 value of the Float32 config value, as the reference converts it. */
int lbs_sponge(int64_t nb, const int32_t* coords, double dx, double Lx, double Ly, double Lz, double sponge_thickness,
               int symmetric, float* sponge);

/* compute_wall_distances! (src/geometry/domain_generation.jl:371-434); returns
 * the number of near-wall cells. */
int64_t lbs_wall_distances(int64_t nb, const int32_t* coords, const int32_t* grid, int gdx, int gdy, int gdz,
                           const uint8_t* obst, double dx, float* wall_dist);

/* compute_bouzidi_qmap_sparse (src/geometry/bouzidi_setup.jl:64-167) with
 * compute_q_for_cell / ray_triangle_intersection (src/geometry/bouzidi_math.jl).
 * Boundary cells in canonical (block, lz, ly, lx) order: cells (n,4) =
 * (b,lx,ly,lz) 1-based; q_bits (n,27) FP16; tri (n,27) 1-based, 0 = none.  */
int lbs_bouzidi_qmap(int64_t nb, const int32_t* coords, const int32_t* grid, int gdx, int gdy, int gdz,
                     const double* tris, int64_t ntri, double dx, const double* offset, int32_t** cells_out,
                     uint16_t** q_out, int32_t** tri_out, int64_t* n_out);

/* Julia Float16(::Float64): round-to-nearest-even. */
uint16_t lbs_double_to_half(double d);

#ifdef __cplusplus
}
#endif
#endif
