// hdf5_io.cpp -- HDF5 field I/O in the reference's file format
// (zmodel/utils.cu:88-281), for the whole virtual grid held by one process.
//
// Per rank: the packed slab is vector-unpacked to block-cyclic order
// (zmodel/utils.h:99-119) and written to, or read from, that rank's
// block-cyclic hyperslab of the global dataset.  Serial HDF5: the ranks are
// visited in turn instead of writing collectively.
#include <hdf5.h>

#include <cstdint>
#include <vector>

#include "../../include/stencilfft_hdf5.h"

namespace {

struct C128 {
    double re, im;
};

// {"real": double, "imag": double}, the reference's compound type
hid_t complex_type() {
    hid_t t = H5Tcreate(H5T_COMPOUND, sizeof(C128));
    if (t < 0) return t;
    if (H5Tinsert(t, "real", 0, H5T_NATIVE_DOUBLE) < 0 || H5Tinsert(t, "imag", sizeof(double), H5T_NATIVE_DOUBLE) < 0) {
        H5Tclose(t);
        return -1;
    }
    return t;
}

// rank id -> coordinates along (depth, row, col) in file-axis order
void rank_coords(const zm_shape& s, int id, int* coord) {
    const int p = s.p;
    const int cid = id % p, rid = (id / p) % p, did = id / (p * p);
    if (s.dims == 2) {
        coord[0] = rid;
        coord[1] = cid;
    } else {
        coord[0] = did;
        coord[1] = rid;
        coord[2] = cid;
    }
}

// packed slab <-> block-cyclic slab: each row of L values, packed position
// j*VEC + jj <-> block-cyclic position j + jj*B
void unpack_slab(const C128* packed, C128* bc, int64_t rows, int l, int b, int vec) {
    for (int64_t r = 0; r < rows; r++) {
        const C128* pr = packed + r * l;
        C128* br = bc + r * l;
        for (int j = 0; j < b; j++)
            for (int jj = 0; jj < vec; jj++) br[j + jj * b] = pr[j * vec + jj];
    }
}

void pack_slab(const C128* bc, C128* packed, int64_t rows, int l, int b, int vec) {
    for (int64_t r = 0; r < rows; r++) {
        const C128* br = bc + r * l;
        C128* pr = packed + r * l;
        for (int j = 0; j < b; j++)
            for (int jj = 0; jj < vec; jj++) pr[j * vec + jj] = br[j + jj * b];
    }
}

// the rank's hyperslab: start = coord*B, stride = B*P, count = L/B, block = B
herr_t select_rank(hid_t space, const zm_shape& s, const zm_dims& d, int id) {
    int coord[3];
    rank_coords(s, id, coord);
    hsize_t start[3], stride[3], count[3], block[3];
    for (int a = 0; a < s.dims; a++) {
        start[a] = (hsize_t)coord[a] * (hsize_t)s.b;
        stride[a] = (hsize_t)s.b * (hsize_t)s.p;
        count[a] = (hsize_t)(d.l / s.b);
        block[a] = (hsize_t)s.b;
    }
    return H5Sselect_hyperslab(space, H5S_SELECT_SET, start, stride, count, block);
}

struct Handles {
    hid_t file = -1, space = -1, mem = -1, type = -1, dset = -1;
    ~Handles() {
        if (dset >= 0) H5Dclose(dset);
        if (type >= 0) H5Tclose(type);
        if (mem >= 0) H5Sclose(mem);
        if (space >= 0) H5Sclose(space);
        if (file >= 0) H5Fclose(file);
    }
};

zm_status check_args(const zm_shape* s, const void* grid, const char* f, const char* ds, zm_dims* d) {
    if (!s || !grid || !f || !ds) return ZM_ERR_INVALID_ARG;
    return zm_shape_check(s, d);
}

}  // namespace

extern "C" zm_status zm_store_hdf5(const zm_shape* s, const double* host_grid, const char* filename,
                                   const char* dataset_name) {
    zm_dims d;
    zm_status st = check_args(s, host_grid, filename, dataset_name, &d);
    if (st != ZM_OK) return st;
    Handles h;
    h.file = H5Fcreate(filename, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (h.file < 0) return ZM_ERR_IO;
    hsize_t dims[3] = {(hsize_t)s->n, (hsize_t)s->n, (hsize_t)s->n};
    hsize_t ldims[3] = {(hsize_t)d.l, (hsize_t)d.l, (hsize_t)d.l};
    h.space = H5Screate_simple(s->dims, dims, nullptr);
    h.mem = H5Screate_simple(s->dims, ldims, nullptr);
    h.type = complex_type();
    if (h.space < 0 || h.mem < 0 || h.type < 0) return ZM_ERR_IO;
    h.dset = H5Dcreate2(h.file, dataset_name, h.type, h.space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (h.dset < 0) return ZM_ERR_IO;
    const C128* grid = reinterpret_cast<const C128*>(host_grid);
    std::vector<C128> bc((size_t)d.local);
    const int64_t rows = d.local / d.l;
    for (int id = 0; id < d.nranks; id++) {
        unpack_slab(grid + (int64_t)id * d.local, bc.data(), rows, d.l, s->b, d.vec);
        if (select_rank(h.space, *s, d, id) < 0) return ZM_ERR_IO;
        if (H5Dwrite(h.dset, h.type, h.mem, h.space, H5P_DEFAULT, bc.data()) < 0) return ZM_ERR_IO;
    }
    return ZM_OK;
}

extern "C" zm_status zm_load_hdf5(const zm_shape* s, double* host_grid, const char* filename,
                                  const char* dataset_name) {
    zm_dims d;
    zm_status st = check_args(s, host_grid, filename, dataset_name, &d);
    if (st != ZM_OK) return st;
    Handles h;
    h.file = H5Fopen(filename, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (h.file < 0) return ZM_ERR_IO;
    h.dset = H5Dopen2(h.file, dataset_name, H5P_DEFAULT);
    if (h.dset < 0) return ZM_ERR_IO;
    h.space = H5Dget_space(h.dset);
    if (h.space < 0) return ZM_ERR_IO;
    // the dataset must be N^d
    hsize_t fdims[3] = {0, 0, 0};
    if (H5Sget_simple_extent_ndims(h.space) != s->dims) return ZM_ERR_SHAPE;
    H5Sget_simple_extent_dims(h.space, fdims, nullptr);
    for (int a = 0; a < s->dims; a++)
        if (fdims[a] != (hsize_t)s->n) return ZM_ERR_SHAPE;
    hsize_t ldims[3] = {(hsize_t)d.l, (hsize_t)d.l, (hsize_t)d.l};
    h.mem = H5Screate_simple(s->dims, ldims, nullptr);
    h.type = complex_type();
    if (h.mem < 0 || h.type < 0) return ZM_ERR_IO;
    C128* grid = reinterpret_cast<C128*>(host_grid);
    std::vector<C128> bc((size_t)d.local);
    const int64_t rows = d.local / d.l;
    for (int id = 0; id < d.nranks; id++) {
        if (select_rank(h.space, *s, d, id) < 0) return ZM_ERR_IO;
        if (H5Dread(h.dset, h.type, h.mem, h.space, H5P_DEFAULT, bc.data()) < 0) return ZM_ERR_IO;
        pack_slab(bc.data(), grid + (int64_t)id * d.local, rows, d.l, s->b, d.vec);
    }
    return ZM_OK;
}
