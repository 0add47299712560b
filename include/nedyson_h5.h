/* GREEN files in the reference's HDF5 layout: green_func::print_to_file and
 * read_from_file (src/greens.cpp:199-306).  A function is a group with the
 * flat complex datasets GR, GL ((nt+1)(nt+2)/2 blocks each), GTV
 * ((nt+1)(ntau+1) blocks), GM (ntau+1 blocks) of nao x nao complex128 blocks
 * in green_func's packing, and the int scalars nt, ntau, sig, nao.  Complex
 * values are the compound {r, i} of two doubles.  The reference writes the
 * Matsubara dataset as "M" but reads "GM"; these functions write GM.
 *
 * Library: nedyson_amd/libnedyson_h5.so (links libnedyson_amd.so and HDF5).
 * Status codes are those of nedyson.h; the message is ned_h5_last_error(). */
#ifndef NEDYSON_H5_H
#define NEDYSON_H5_H

#include "nedyson.h"

#ifdef __cplusplus
extern "C" {
#endif

const char *ned_h5_last_error(void);

/* green_func::read_from_file's size fields of the group (group "" = root) */
ned_status ned_h5_green_dims(const char *file, const char *group, int *nt, int *ntau, int *nao, int *sig);

/* host buffers (any may be NULL to skip that component); nt, ntau, nao must
 * equal the file's */
ned_status ned_h5_green_read_host(const char *file, const char *group, int nt, int ntau, int nao, void *ret,
                                  void *les, void *tv, void *mat);
/* green_func::print_to_file from host buffers; creates the file or updates
 * the group of an existing one */
ned_status ned_h5_green_write_host(const char *file, const char *group, int nt, int ntau, int nao, int sig,
                                   const void *ret, const void *les, const void *tv, const void *mat);

/* read_from_file into G (which = NED_G) or Sigma (NED_SIGMA) of a context:
 * the file's nao and ntau must equal the context's and its nt must not
 * exceed it (later slices are left as they are) */
ned_status ned_h5_green_read(ned_ctx *ctx, int which, const char *file, const char *group);
/* print_to_file of slices 0..nt of G or Sigma of a context (nt < 0: all) */
ned_status ned_h5_green_write(ned_ctx *ctx, int which, const char *file, const char *group, int nt);

#ifdef __cplusplus
}
#endif

#endif
