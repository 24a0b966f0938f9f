/*
 * modle_cool.h: C-ABI of the .cool writer for the banded contact matrices the simulate path
 * produces (the Cooler<u32> writer of libmodle_io, src/libmodle_io/cooler_impl.hpp and
 * cooler_write_impl.hpp). Host code on the HDF5 C library; no device code.
 *
 * A file holds, under its root group, the Cooler format version 3 layout the reference writes:
 *   chroms/name (fixed-length strings), chroms/length (int64);
 *   bins/chrom, bins/start, bins/end (int64; fixed-size bins of every chromosome, in file order);
 *   pixels/bin1_id, pixels/bin2_id (int64), pixels/count (int32): the nonzero pixels of the upper
 *     triangle within the band, bin1-major, bin2 ascending;
 *   indexes/chrom_offset (nchroms + 1), indexes/bin1_offset (one per bin, plus the final nnz);
 *   root attributes format, format-version, bin-type, bin-size, storage-mode, assembly-name,
 *   generated-by, creation-date, metadata (JSON), nchroms, nbins, nnz, sum.
 *
 * Every function returns 0 on success or a negative MCOOL_E* code; mcool_last_error() gives the
 * message of the last failure on this thread. A writer is not thread-safe.
 */
#ifndef MODLE_COOL_H
#define MODLE_COOL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MCOOL_OK 0
#define MCOOL_EINVAL -1 /* invalid argument */
#define MCOOL_EIO -2    /* HDF5 error */

typedef struct mcool_writer mcool_writer;

/* Cooler<u32>(path, WRITE_ONLY, bin_size, max_str_length) (cooler_impl.hpp:22-93): creates (or
 * truncates) the file, its groups and datasets, and writes the format attributes
 * (Cooler::write_metadata, cooler_write_impl.hpp:42-97). assembly_name may be NULL or empty
 * (no assembly-name attribute). Chromosome names are stored as max_str_length-byte strings. */
int mcool_open(const char* path, uint64_t bin_size, uint64_t max_str_length,
               const char* assembly_name, mcool_writer** out);
/* Cooler::write_metadata_attribute (cooler_write_impl.hpp:99-121): the "metadata" attribute, the
 * JSON of the command line the simulate CLI stores there. */
int mcool_write_metadata(mcool_writer* w, const char* json);
/* Cooler::write_or_append_cmatrix_to_file (cooler_write_impl.hpp:139-371) for one chromosome of
 * `length` bp simulated over [start, end): appends its name and length, its bins, its nonzero
 * pixels and its index entries. band is the chromosome's ContactMatrixDense<u32> data (nrows *
 * ncols + 1 values, data[col * nrows + (col - row)]); NULL writes the chromosome's bins and index
 * only (an empty matrix). Counts above INT32_MAX are MCOOL_EINVAL (the pixel count is int32). */
int mcool_write_chrom(mcool_writer* w, const char* name, uint64_t name_len, uint64_t start,
                      uint64_t end, uint64_t length, const uint32_t* band, uint64_t nrows,
                      uint64_t ncols);
/* ~Cooler (cooler_impl.hpp:95-125): appends the closing entries of both indexes, writes nnz and
 * sum, and closes the file. The writer is freed whatever the result. */
int mcool_close(mcool_writer* w);
const char* mcool_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
