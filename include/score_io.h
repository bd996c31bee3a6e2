/*
 * score_io.h -- grain-map files in SCORE's formats (host only, no GPU).
 *
 * The growth machine hands back the final grain map in memory
 * (score_gm_get_grainmap with SCORE_GM_MAP_RAW2GID: 1 = infected, else 2+gid,
 * x fastest, then y, then z).  These functions write and read the files SCORE
 * makes of such a map:
 *   - the headerless uint32 stack of ompcrit_write_voxeldata_coloring_grainids
 *     (SCORE_Io_Characterization.cpp:583-676) and its file name;
 *   - the HDF5 dataset of hdf5_write_voxelgrid_grainid (SCORE_Io_HDF5.cpp:126-262)
 *     under the group and name ompcrit_write_voxeldata_h5 gives it
 *     (SCORE_Io_HDF5.cpp:862-935, called with "/GROWTH3DGrainID" at the end of
 *     solve_RXGROWTH, SCORE_Kernel.cpp:6756-6771);
 *   - the XDMF file hdf5_write_xmffile_3dcorectmesh writes beside it
 *     (SCORE_Io_HDF5.cpp:715-797).
 * All functions return 0 on success and a negative SCORE_IO_ERR_* otherwise.
 */
#ifndef SCORE_IO_H
#define SCORE_IO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SCORE_IO_OK 0
#define SCORE_IO_ERR_ARG -1   /* bad argument, or a buffer too small */
#define SCORE_IO_ERR_FILE -2  /* the file could not be created, written or read */
#define SCORE_IO_ERR_HDF5 -3  /* an HDF5 call failed, or the dataset is not a 3D uint32 one */
#define SCORE_IO_ERR_SIZE -4  /* more than HDF5_MAXIMUM_SINGLE_WRITE (1600 MiB) in one write */

/* "SCORE.<simid>.JobID.<jobid>.Iter.<step>.X.<XXXX>.<postfix>.raw" with XXXX =
 * (uint32)(X*1000) written with four digits (SCORE_Io_Characterization.cpp:601-609). */
int score_io_raw_name(char *buf, size_t cap, uint32_t simid, uint32_t jobid, uint32_t step, double X,
                      const char *postfix);

/* The map as SCORE writes it: nx*ny*nz native (little-endian) uint32, one xy
 * layer after the other along +z. */
int score_io_write_raw(const char *path, const uint32_t *map, uint32_t nx, uint32_t ny, uint32_t nz);

/* "SCORE.<simid>.VoxelData.h5" and "SCORE.<simid>.VoxelData.xdmf" */
int score_io_h5_name(char *buf, size_t cap, uint32_t simid);
int score_io_xdmf_name(char *buf, size_t cap, uint32_t simid);

/* The map, or its part inside the inclusive render window win = {xmin, xmax,
 * ymin, ymax, zmin, zmax} (NULL: the whole domain), as dataset
 * "<group>/grainid_<step>" of type H5T_STD_U32LE and dimensions
 * {wx, wy, wz} holding the values x fastest, as SCORE stores them.  The file is
 * opened for writing, or created when absent; the group is created when absent. */
int score_io_write_h5_grainid(const char *h5path, const char *group, uint32_t step, const uint32_t *map,
                              uint32_t nx, uint32_t ny, uint32_t nz, const uint32_t *win);

/* Dimensions of a 3D dataset; then its values (cap entries at least). */
int score_io_read_h5_dims(const char *h5path, const char *dataset, uint64_t dims[3]);
int score_io_read_h5_u32(const char *h5path, const char *dataset, uint32_t *out, size_t cap);

/* The XDMF time series of nsnap grain-ID snapshots, snapshot i given as
 * snap[4*i .. 4*i+3] = {nx, ny, nz, step}, linking to h5name:<group>/<step> with
 * the text SCORE writes. */
int score_io_write_xdmf(const char *xmfpath, const char *h5name, const char *group, const uint32_t *snap,
                        uint32_t nsnap);

#ifdef __cplusplus
}
#endif

#endif
