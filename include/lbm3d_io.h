/*
 * lbm3d_io.h -- geometry input of liblbm3d.so: the reference's voxel-file
 * reader and an STL voxelizer.
 */
#ifndef LBM3D_IO_H
#define LBM3D_IO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

const char *lbm3d_io_last_error(void);

/* init_geo (Single_phase/LBM_3D_SinglePhase_Solver.py:173-177): read the
 * ASCII voxel file `path` -- whitespace-separated numbers, x fastest, then y,
 * then z -- into the HOST int8 array solid[nx][ny][nz] (C order).  As the
 * reference does with np.loadtxt, values > 0 become 1 and the others are cast
 * to int8 (truncated toward zero).  Returns the number of values read, or a
 * negative LBM3D_E* code (also when the count is not nx*ny*nz). */
int64_t lbm3d_read_geo_ascii(const char *path, int32_t nx, int32_t ny, int32_t nz, int8_t *solid);

/* Voxelize a closed triangle mesh, the job of
 * Single_phase/Convert_stl_to_binary.cpp:29-157: node (i, j, k) at
 * (x0 + i*dx, y0 + j*dx, z0 + k*dx) gets solid[(i*ny + j)*nz + k] = 1 when its
 * generalized winding number with respect to the mesh is >= 1/2 (inside or on
 * the surface), else 0.  tris: DEVICE float [n_tris][3][3] (corner, xyz);
 * solid: DEVICE int8.  The reference decides inside/outside by the sign of
 * libigl's fast-winding-number signed distance; this sums the exact winding
 * number. */
int lbm3d_voxelize(void *stream, const float *tris, int32_t n_tris, double x0, double y0, double z0, double dx,
                   int32_t nx, int32_t ny, int32_t nz, int8_t *solid);

#ifdef __cplusplus
}
#endif

#endif /* LBM3D_IO_H */
