/*
 * stencilfft_hdf5.h -- optional HDF5 field I/O (libstencilfft_hdf5.so, linked
 * against the serial HDF5 C library).
 *
 * File format: the reference's (zmodel/utils.cu:88-93, 103-281).  A field
 * is one N x N (N x N x N) dataset of the compound type {"real": double,
 * "imag": double} in global natural order, axis 0 = rows gr (depth gd in 3D),
 * last axis = columns gc.  Each rank's part is the block-cyclic hyperslab
 * start = rank coordinates * B, stride = B * P, count = L / B, block = B per
 * axis, vector-packed in memory (zmodel/utils.h:99-128).
 *
 * Buffers are host memory holding the whole virtual grid: the P^d rank slabs
 * of L^d complex128 values, (re, im) interleaved, in world-rank order.  The
 * reference writes and reads with parallel HDF5, one rank per process; here one
 * process writes or reads every rank's hyperslab in turn.
 */
#ifndef STENCILFFT_HDF5_H
#define STENCILFFT_HDF5_H

#include "stencilfft.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces store_host_hdf5 / store_host_hdf5_3d (zmodel/utils.cu:150-196, 239-281):
 * creates (truncates) `filename` and writes dataset `dataset_name`. */
zm_status zm_store_hdf5(const zm_shape* shape, const double* host_grid, const char* filename,
                        const char* dataset_name);

/* Replaces init_host_hdf5 / init_host_hdf5_3d (zmodel/utils.cu:103-148, 198-237):
 * reads dataset `dataset_name` (N^d of the compound type above) into the packed
 * virtual grid. */
zm_status zm_load_hdf5(const zm_shape* shape, double* host_grid, const char* filename, const char* dataset_name);

#ifdef __cplusplus
}
#endif

#endif
