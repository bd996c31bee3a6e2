/*
 * score_io.c -- grain-map files in SCORE's formats (see include/score_io.h).
 * Host code; links against the HDF5 C library.
 */
#include "../../include/score_io.h"

#include <hdf5.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* HDF5_MAXIMUM_SINGLE_WRITE (SCORE_Defs_Functionality.h:367) */
#define SCORE_IO_H5_MAX_WRITE ((uint64_t)1600 * 1024 * 1024)

int score_io_raw_name(char *buf, size_t cap, uint32_t simid, uint32_t jobid, uint32_t step, double X,
                      const char *postfix) {
    if (!buf || !postfix) return SCORE_IO_ERR_ARG;
    /* uint32_t XX = (X * 1000.0); then four digits, zero-padded below 1000 */
    const uint32_t XX = (uint32_t)(X * 1000.0);
    int n = snprintf(buf, cap, "SCORE.%u.JobID.%u.Iter.%u.X.%04u.%s.raw", simid, jobid, step, XX, postfix);
    return (n < 0 || (size_t)n >= cap) ? SCORE_IO_ERR_ARG : SCORE_IO_OK;
}

int score_io_write_raw(const char *path, const uint32_t *map, uint32_t nx, uint32_t ny, uint32_t nz) {
    if (!path || !map || !nx || !ny || !nz) return SCORE_IO_ERR_ARG;
    FILE *f = fopen(path, "wb");
    if (!f) return SCORE_IO_ERR_FILE;
    const size_t nxy = (size_t)nx * ny;
    int rc = SCORE_IO_OK;
    /* one xy layer per write, along +z */
    for (uint32_t z = 0; z < nz && rc == SCORE_IO_OK; z++)
        if (fwrite(map + nxy * z, sizeof(uint32_t), nxy, f) != nxy) rc = SCORE_IO_ERR_FILE;
    if (fclose(f) != 0 && rc == SCORE_IO_OK) rc = SCORE_IO_ERR_FILE;
    return rc;
}

int score_io_h5_name(char *buf, size_t cap, uint32_t simid) {
    if (!buf) return SCORE_IO_ERR_ARG;
    int n = snprintf(buf, cap, "SCORE.%u.VoxelData.h5", simid);
    return (n < 0 || (size_t)n >= cap) ? SCORE_IO_ERR_ARG : SCORE_IO_OK;
}

int score_io_xdmf_name(char *buf, size_t cap, uint32_t simid) {
    if (!buf) return SCORE_IO_ERR_ARG;
    int n = snprintf(buf, cap, "SCORE.%u.VoxelData.xdmf", simid);
    return (n < 0 || (size_t)n >= cap) ? SCORE_IO_ERR_ARG : SCORE_IO_OK;
}

int score_io_write_h5_grainid(const char *h5path, const char *group, uint32_t step, const uint32_t *map,
                              uint32_t nx, uint32_t ny, uint32_t nz, const uint32_t *win) {
    if (!h5path || !group || !map || !nx || !ny || !nz) return SCORE_IO_ERR_ARG;
    uint32_t w[6] = {0, nx - 1, 0, ny - 1, 0, nz - 1};
    if (win) {
        memcpy(w, win, sizeof(w));
        if (w[0] > w[1] || w[1] >= nx || w[2] > w[3] || w[3] >= ny || w[4] > w[5] || w[5] >= nz) return SCORE_IO_ERR_ARG;
    }
    const uint64_t wx = w[1] - w[0] + 1, wy = w[3] - w[2] + 1, wz = w[5] - w[4] + 1;
    if (wx * wy * wz * sizeof(uint32_t) > SCORE_IO_H5_MAX_WRITE) return SCORE_IO_ERR_SIZE;
    char dname[512];
    int n = snprintf(dname, sizeof(dname), "%s/grainid_%u", group, step);
    if (n < 0 || (size_t)n >= sizeof(dname)) return SCORE_IO_ERR_ARG;
    uint32_t *buf = (uint32_t *)malloc(wx * wy * wz * sizeof(uint32_t));
    if (!buf) return SCORE_IO_ERR_ARG;
    /* the window's cells, x fastest, then y, then z */
    size_t c = 0;
    for (uint64_t z = w[4]; z <= w[5]; z++)
        for (uint64_t y = w[2]; y <= w[3]; y++) {
            memcpy(buf + c, map + w[0] + y * nx + z * (uint64_t)nx * ny, wx * sizeof(uint32_t));
            c += wx;
        }
    int rc = SCORE_IO_ERR_HDF5;
    H5Eset_auto2(H5E_DEFAULT, NULL, NULL);
    hid_t file = H5Fopen(h5path, H5F_ACC_RDWR, H5P_DEFAULT);
    if (file < 0) file = H5Fcreate(h5path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (file >= 0) {
        if (H5Lexists(file, group, H5P_DEFAULT) <= 0) {
            hid_t g = H5Gcreate2(file, group, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
            if (g >= 0) H5Gclose(g);
        }
        const hsize_t dims[3] = {wx, wy, wz};
        hid_t space = H5Screate_simple(3, dims, NULL);
        if (space >= 0) {
            hid_t dset = H5Dcreate2(file, dname, H5T_STD_U32LE, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
            if (dset >= 0) {
                if (H5Dwrite(dset, H5T_NATIVE_UINT, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf) >= 0) rc = SCORE_IO_OK;
                H5Dclose(dset);
            }
            H5Sclose(space);
        }
        if (H5Fclose(file) < 0) rc = SCORE_IO_ERR_HDF5;
    } else {
        rc = SCORE_IO_ERR_FILE;
    }
    free(buf);
    return rc;
}

/* a 3D dataset of 32-bit unsigned integers, its dimensions in dims */
static hid_t open_u32_3d(hid_t file, const char *dataset, hsize_t dims[3]) {
    hid_t dset = H5Dopen2(file, dataset, H5P_DEFAULT);
    if (dset < 0) return -1;
    hid_t space = H5Dget_space(dset);
    hid_t type = H5Dget_type(dset);
    int ok = space >= 0 && type >= 0 && H5Sget_simple_extent_ndims(space) == 3 &&
             H5Sget_simple_extent_dims(space, dims, NULL) == 3 && H5Tget_class(type) == H5T_INTEGER &&
             H5Tget_size(type) == 4 && H5Tget_sign(type) == H5T_SGN_NONE;
    if (space >= 0) H5Sclose(space);
    if (type >= 0) H5Tclose(type);
    if (!ok) {
        H5Dclose(dset);
        return -1;
    }
    return dset;
}

int score_io_read_h5_dims(const char *h5path, const char *dataset, uint64_t dims[3]) {
    if (!h5path || !dataset || !dims) return SCORE_IO_ERR_ARG;
    H5Eset_auto2(H5E_DEFAULT, NULL, NULL);
    hid_t file = H5Fopen(h5path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (file < 0) return SCORE_IO_ERR_FILE;
    hsize_t d[3];
    hid_t dset = open_u32_3d(file, dataset, d);
    int rc = SCORE_IO_ERR_HDF5;
    if (dset >= 0) {
        dims[0] = d[0];
        dims[1] = d[1];
        dims[2] = d[2];
        rc = SCORE_IO_OK;
        H5Dclose(dset);
    }
    H5Fclose(file);
    return rc;
}

int score_io_read_h5_u32(const char *h5path, const char *dataset, uint32_t *out, size_t cap) {
    if (!h5path || !dataset || !out) return SCORE_IO_ERR_ARG;
    H5Eset_auto2(H5E_DEFAULT, NULL, NULL);
    hid_t file = H5Fopen(h5path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (file < 0) return SCORE_IO_ERR_FILE;
    hsize_t d[3];
    hid_t dset = open_u32_3d(file, dataset, d);
    int rc = SCORE_IO_ERR_HDF5;
    if (dset >= 0) {
        if ((uint64_t)d[0] * d[1] * d[2] > cap) rc = SCORE_IO_ERR_ARG;
        else if (H5Dread(dset, H5T_NATIVE_UINT, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) >= 0) rc = SCORE_IO_OK;
        H5Dclose(dset);
    }
    H5Fclose(file);
    return rc;
}

int score_io_write_xdmf(const char *xmfpath, const char *h5name, const char *group, const uint32_t *snap,
                        uint32_t nsnap) {
    if (!xmfpath || !h5name || !group || (nsnap && !snap)) return SCORE_IO_ERR_ARG;
    FILE *f = fopen(xmfpath, "w");
    if (!f) return SCORE_IO_ERR_FILE;
    /* program name SCORE_v<VERSION_MAJOR>.<VERSION_MINOR>.<VERSION_BUILD> (SCORE_Defs_Functionality.h:61-64) */
    const char *prg = "SCORE_v1.2.0";
    fprintf(f, "<?xml version=\"1.0\" ?>\n");
    fprintf(f, "<!DOCTYPE Xdmf SYSTEM \"Xdmf.dtd\" []>\n");
    fprintf(f, "<Xdmf xmlns:xi=\"http://www.w3.org/2003/XInclude\" Version=\"2.2\">\n");
    fprintf(f, "\t<Domain>\n");
    fprintf(f, "\t\t<Grid Name=\"%s_TimeSeries\" GridType=\"Collection\" CollectionType=\"Temporal\">\n", prg);
    for (uint32_t i = 0; i < nsnap; i++) {
        const uint32_t nx = snap[4 * i], ny = snap[4 * i + 1], nz = snap[4 * i + 2], step = snap[4 * i + 3];
        fprintf(f, "\t\t\t<Grid Name=\"%s_Result\" GridType=\"Uniform\">\n", prg);
        /* the snapshot's ordinal, as (double) snp+1 with six significant digits */
        fprintf(f, "\t\t\t\t<Time Type=\"Single\" Value=\"   %.6g\" />\n", (double)i + 1.0);
        fprintf(f, "\t\t\t\t<Topology TopologyType=\"3DCoRectMesh\" Dimensions=\"%u %u %u\"></Topology>\n", nx, ny, nz);
        fprintf(f, "\t\t\t\t<Geometry Type=\"Origin_DXDYDZ\">\n");
        fprintf(f, "\t\t\t\t\t<DataItem Format=\"XML\" Dimensions=\"3\" NumberType=\"Float\" Precision=\"3\">0.0 0.0 0.0</DataItem>\n");
        fprintf(f, "\t\t\t\t\t<DataItem Format=\"XML\" Dimensions=\"3\" NumberType=\"Float\" Precision=\"3\">1.0 1.0 1.0</DataItem>\n");
        fprintf(f, "\t\t\t\t</Geometry>\n");
        fprintf(f, "\t\t\t\t<Attribute Name=\"GrainID+2\" AttributeType=\"Scalar\" Center=\"Cell\">\n");
        fprintf(f, "\t\t\t\t\t<DataItem Dimensions=\"%u %u %u\" Format=\"HDF\" NumberType=\"UInt\" Precision=\"4\">\n", nx, ny, nz);
        /* SCORE links <group>/<step>; its dataset is named <group>/grainid_<step> */
        fprintf(f, "\t\t\t\t\t%s:%s/%u</DataItem>\n", h5name, group, step);
        fprintf(f, "\t\t\t\t</Attribute>\n");
        fprintf(f, "\t\t\t</Grid>\n");
    }
    fprintf(f, "\t\t</Grid>\n");
    fprintf(f, "\t</Domain>\n");
    fprintf(f, "</Xdmf>");
    return fclose(f) == 0 ? SCORE_IO_OK : SCORE_IO_ERR_FILE;
}
