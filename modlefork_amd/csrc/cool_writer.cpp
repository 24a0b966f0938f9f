// .cool writer of the simulate path's banded contact matrices (include/modle_cool.h): the layout
// libmodle_io's Cooler<u32> writes (src/libmodle_io/cooler_impl.hpp,
// src/libmodle_io/cooler_write_impl.hpp), on the HDF5 C library. Host code only.
#include <hdf5.h>

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "../../include/modle_cool.h"

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

// datasets in the order the reference creates them (cooler_impl.hpp:572-598)
enum Dset { kChromLen, kChromName, kBinChrom, kBinStart, kBinEnd, kB1, kB2, kCount, kIdxBin1,
            kIdxChrom, kNumDsets };
const char* const kDsetPath[kNumDsets] = {
    "chroms/length", "chroms/name",     "bins/chrom",          "bins/start",
    "bins/end",      "pixels/bin1_id", "pixels/bin2_id",      "pixels/count",
    "indexes/bin1_offset", "indexes/chrom_offset"};

constexpr unsigned kDeflateLevel = 6;          // Cooler::DEFAULT_COMPRESSION_LEVEL
constexpr hsize_t kChunkBytes = 1024 * 1024;   // Cooler::DEFAULT_HDF5_CHUNK_SIZE
constexpr size_t kFlushPixels = 1u << 20;      // pixels buffered before a write

}  // namespace

struct mcool_writer {
  hid_t file = -1;
  hid_t dset[kNumDsets];
  hid_t str_type = -1;
  uint64_t bin_size = 0, max_str = 0;
  int64_t nchroms = 0, nbins = 0, nnz = 0, sum = 0;
  hsize_t len[kNumDsets] = {};
  std::vector<int64_t> b1, b2;
  std::vector<int32_t> count;
  mcool_writer() { std::fill(dset, dset + kNumDsets, (hid_t)-1); }
};

namespace {

// A chunked, deflated, extensible 1-D dataset (generate_creat_prop_list, hdf5_impl.hpp:619-631).
hid_t create_dataset(hid_t file, const char* path, hid_t type, size_t elem_bytes) {
  const hsize_t dims = 0, maxdims = H5S_UNLIMITED;
  const hsize_t chunk = std::max<hsize_t>(1, kChunkBytes / elem_bytes);
  hid_t space = H5Screate_simple(1, &dims, &maxdims);
  hid_t dcpl = H5Pcreate(H5P_DATASET_CREATE);
  hid_t lcpl = H5Pcreate(H5P_LINK_CREATE);
  H5Pset_chunk(dcpl, 1, &chunk);
  H5Pset_deflate(dcpl, kDeflateLevel);
  H5Pset_create_intermediate_group(lcpl, 1);
  const hid_t d = H5Dcreate2(file, path, type, space, lcpl, dcpl, H5P_DEFAULT);
  H5Pclose(lcpl);
  H5Pclose(dcpl);
  H5Sclose(space);
  return d;
}

// Appends n elements of memory type mtype to dataset k.
bool append(mcool_writer* w, int k, hid_t mtype, const void* buf, hsize_t n) {
  if (n == 0) return true;
  const hsize_t off = w->len[k], newlen = off + n;
  if (H5Dset_extent(w->dset[k], &newlen) < 0) return false;
  hid_t fspace = H5Dget_space(w->dset[k]);
  hid_t mspace = H5Screate_simple(1, &n, nullptr);
  bool ok = H5Sselect_hyperslab(fspace, H5S_SELECT_SET, &off, nullptr, &n, nullptr) >= 0 &&
            H5Dwrite(w->dset[k], mtype, mspace, fspace, H5P_DEFAULT, buf) >= 0;
  H5Sclose(mspace);
  H5Sclose(fspace);
  if (ok) w->len[k] = newlen;
  return ok;
}

bool append_i64(mcool_writer* w, int k, const std::vector<int64_t>& v) {
  return append(w, k, H5T_NATIVE_INT64, v.data(), v.size());
}

// Appends `count` copies of value to the bin1 index.
bool fill_idx_bin1(mcool_writer* w, uint64_t count, int64_t value) {
  if (count == 0) return true;
  std::vector<int64_t> v(count, value);
  return append_i64(w, kIdxBin1, v);
}

// A scalar root attribute (hdf5::write_or_create_attribute, hdf5_impl.hpp:408-437): strings as
// variable-length UTF-8 (METADATA_STR_TYPE), integers as int64.
bool write_attr_str(hid_t file, const char* name, const char* value) {
  hid_t t = H5Tcopy(H5T_C_S1);
  H5Tset_size(t, H5T_VARIABLE);
  H5Tset_strpad(t, H5T_STR_NULLTERM);
  H5Tset_cset(t, H5T_CSET_UTF8);
  if (H5Aexists(file, name) > 0) H5Adelete(file, name);
  hid_t space = H5Screate(H5S_SCALAR);
  hid_t a = H5Acreate2(file, name, t, space, H5P_DEFAULT, H5P_DEFAULT);
  bool ok = a >= 0 && H5Awrite(a, t, &value) >= 0;
  if (a >= 0) H5Aclose(a);
  H5Sclose(space);
  H5Tclose(t);
  return ok;
}

bool write_attr_i64(hid_t file, const char* name, int64_t value) {
  hid_t a;
  if (H5Aexists(file, name) > 0) {
    a = H5Aopen(file, name, H5P_DEFAULT);
  } else {
    hid_t space = H5Screate(H5S_SCALAR);
    a = H5Acreate2(file, name, H5T_NATIVE_INT64, space, H5P_DEFAULT, H5P_DEFAULT);
    H5Sclose(space);
  }
  bool ok = a >= 0 && H5Awrite(a, H5T_NATIVE_INT64, &value) >= 0;
  if (a >= 0) H5Aclose(a);
  return ok;
}

// The creation date as an RFC 3339 UTC timestamp with microseconds (the reference formats
// absl::Now() in UTC).
std::string utc_now() {
  const auto now = std::chrono::system_clock::now();
  const std::time_t t = std::chrono::system_clock::to_time_t(now);
  const long us = (long)(std::chrono::duration_cast<std::chrono::microseconds>(
                             now.time_since_epoch()).count() % 1000000);
  std::tm tm{};
  gmtime_r(&t, &tm);
  char buf[64];
  std::strftime(buf, sizeof(buf), "%Y-%m-%dT%H:%M:%S", &tm);
  char out[96];
  std::snprintf(out, sizeof(out), "%s.%06ld+00:00", buf, us);
  return out;
}

bool flush_pixels(mcool_writer* w) {
  bool ok = append_i64(w, kB1, w->b1) && append_i64(w, kB2, w->b2) &&
            append(w, kCount, H5T_NATIVE_INT32, w->count.data(), w->count.size());
  w->b1.clear();
  w->b2.clear();
  w->count.clear();
  return ok;
}

void close_handles(mcool_writer* w) {
  for (auto& d : w->dset)
    if (d >= 0) H5Dclose(d), d = -1;
  if (w->str_type >= 0) H5Tclose(w->str_type), w->str_type = -1;
  if (w->file >= 0) H5Fclose(w->file), w->file = -1;
}

}  // namespace

extern "C" {

const char* mcool_last_error(void) { return g_err.c_str(); }

int mcool_open(const char* path, uint64_t bin_size, uint64_t max_str_length,
               const char* assembly_name, mcool_writer** out) {
  if (!path || !out) return fail(MCOOL_EINVAL, "null argument");
  if (bin_size == 0) return fail(MCOOL_EINVAL, "bin size must be > 0");
  if (max_str_length == 0) return fail(MCOOL_EINVAL, "chromosome names need a length > 0");
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);  // errors are reported through the return codes
  auto* w = new mcool_writer();
  w->bin_size = bin_size;
  w->max_str = max_str_length;
  w->file = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (w->file < 0) {
    delete w;
    return fail(MCOOL_EIO, "cannot create %s", path);
  }
  // chroms/name: fixed-length C strings of the longest name (generate_default_str_type)
  w->str_type = H5Tcopy(H5T_C_S1);
  H5Tset_size(w->str_type, max_str_length);
  const hid_t types[kNumDsets] = {H5T_NATIVE_INT64, w->str_type,     H5T_NATIVE_INT64,
                                  H5T_NATIVE_INT64, H5T_NATIVE_INT64, H5T_NATIVE_INT64,
                                  H5T_NATIVE_INT64, H5T_NATIVE_INT32, H5T_NATIVE_INT64,
                                  H5T_NATIVE_INT64};
  const size_t sizes[kNumDsets] = {8, max_str_length, 8, 8, 8, 8, 8, 4, 8, 8};
  bool ok = true;
  for (int k = 0; k < kNumDsets && ok; ++k) {
    w->dset[k] = create_dataset(w->file, kDsetPath[k], types[k], sizes[k]);
    ok = w->dset[k] >= 0;
  }
  // Cooler::write_metadata (cooler_write_impl.hpp:42-97)
  ok = ok && write_attr_str(w->file, "format", "HDF5::Cooler") &&
       write_attr_i64(w->file, "format-version", 3) &&
       write_attr_str(w->file, "bin-type", "fixed") &&
       write_attr_i64(w->file, "bin-size", (int64_t)bin_size) &&
       write_attr_str(w->file, "storage-mode", "symmetric-upper");
  if (ok && assembly_name && assembly_name[0])
    ok = write_attr_str(w->file, "assembly-name", assembly_name);
  ok = ok && write_attr_str(w->file, "generated-by", "modle-mi355x 0.1.0") &&
       write_attr_str(w->file, "creation-date", utc_now().c_str());
  if (!ok) {
    close_handles(w);
    delete w;
    return fail(MCOOL_EIO, "cannot lay out the Cooler groups, datasets and attributes in %s", path);
  }
  *out = w;
  return MCOOL_OK;
}

int mcool_write_metadata(mcool_writer* w, const char* json) {
  if (!w || !json || !json[0]) return fail(MCOOL_EINVAL, "null or empty metadata");
  if (!write_attr_str(w->file, "metadata", json)) return fail(MCOOL_EIO, "cannot write metadata");
  return MCOOL_OK;
}

int mcool_write_chrom(mcool_writer* w, const char* name, uint64_t name_len, uint64_t start,
                      uint64_t end, uint64_t length, const uint32_t* band, uint64_t nrows,
                      uint64_t ncols) {
  if (!w || !name) return fail(MCOOL_EINVAL, "null argument");
  if (name_len == 0 || name_len > w->max_str)
    return fail(MCOOL_EINVAL, "chromosome name of %llu bytes: the file stores up to %llu",
                (unsigned long long)name_len, (unsigned long long)w->max_str);
  if (start > end || end > length) return fail(MCOOL_EINVAL, "invalid chromosome range");
  const uint64_t bs = w->bin_size;
  if (band && ncols != (end - start + bs - 1) / bs)
    return fail(MCOOL_EINVAL, "band of %llu columns for a %llu bp range at %llu bp bins",
                (unsigned long long)ncols, (unsigned long long)(end - start),
                (unsigned long long)bs);
  const int64_t chrom_id = w->nchroms++;
  // chroms/name, chroms/length
  std::vector<char> nm(w->max_str, '\0');
  std::memcpy(nm.data(), name, name_len);
  const int64_t len64 = (int64_t)length;
  bool ok = append(w, kChromName, w->str_type, nm.data(), 1) &&
            append(w, kChromLen, H5T_NATIVE_INT64, &len64, 1);
  // the index entry of the chromosome's first bin, appended at the end of the call
  const int64_t first_bin = w->nbins;
  // bins of the whole chromosome (Cooler::write_bins, cooler_write_impl.hpp:374-415)
  const uint64_t nb = (length + bs - 1) / bs;
  {
    std::vector<int64_t> chrom(nb, chrom_id), s(nb), e(nb);
    for (uint64_t b = 0; b < nb; ++b) {
      s[b] = (int64_t)(b * bs);
      e[b] = (int64_t)std::min<uint64_t>((b + 1) * bs, length);
    }
    ok = ok && append_i64(w, kBinChrom, chrom) && append_i64(w, kBinStart, s) &&
         append_i64(w, kBinEnd, e);
  }
  w->nbins += (int64_t)nb;
  // index of the bins before the simulated range
  ok = ok && fill_idx_bin1(w, (start + bs - 1) / bs, w->nnz);
  if (!band || ncols == 0) ok = ok && fill_idx_bin1(w, nb, w->nnz);
  // nonzero pixels, bin1-major: for bin1 column i, bin2 j in [i, i + nrows) within the matrix,
  // pixel (i, j) at data[j * nrows + (j - i)] (ContactMatrixDense::unsafe_get)
  const int64_t pxl_offset = first_bin + (int64_t)((start + bs - 1) / bs);
  if (band && ok) {
    std::vector<int64_t> idx;
    idx.reserve(std::min<uint64_t>(ncols, kFlushPixels));
    for (uint64_t i = 0; i < ncols && ok; ++i) {
      idx.push_back(w->nnz);
      for (uint64_t j = i; j < i + nrows && j < ncols; ++j) {
        const uint32_t m = band[j * nrows + (j - i)];
        if (m == 0) continue;
        if (m > (uint32_t)INT32_MAX) {
          ok = false;
          g_err = "a pixel count exceeds the int32 range of pixels/count";
          break;
        }
        w->b1.push_back(pxl_offset + (int64_t)i);
        w->b2.push_back(pxl_offset + (int64_t)j);
        w->count.push_back((int32_t)m);
        w->sum += m;
        ++w->nnz;
        if (w->b1.size() == kFlushPixels) ok = flush_pixels(w);
      }
      if (idx.size() == kFlushPixels) {
        ok = ok && append_i64(w, kIdxBin1, idx);
        idx.clear();
      }
    }
    if (ok && !idx.empty()) ok = append_i64(w, kIdxBin1, idx);
  }
  // index of the bins after the simulated range
  ok = ok && fill_idx_bin1(w, (length - end + bs - 1) / bs, w->nnz);
  ok = ok && flush_pixels(w);
  ok = ok && append(w, kIdxChrom, H5T_NATIVE_INT64, &first_bin, 1);
  ok = ok && write_attr_i64(w->file, "nchroms", w->nchroms) &&
       write_attr_i64(w->file, "nbins", w->nbins) && write_attr_i64(w->file, "nnz", w->nnz) &&
       write_attr_i64(w->file, "sum", w->sum);
  if (!ok) return fail(MCOOL_EIO, "writing %.*s failed%s%s", (int)name_len, name,
                       g_err.empty() ? "" : ": ", g_err.c_str());
  return MCOOL_OK;
}

int mcool_close(mcool_writer* w) {
  if (!w) return fail(MCOOL_EINVAL, "null writer");
  bool ok = true;
  if (w->nchroms != 0) {
    // the closing entries: chrom_offset ends with nbins, bin1_offset with nnz. The reference
    // appends each only when it differs from the last entry (cooler_impl.hpp:95-125), which drops
    // bin1_offset's closing entry when the last bin holds no pixel; it is always appended here,
    // as the Cooler format has it (nbins + 1 entries).
    ok = append(w, kIdxChrom, H5T_NATIVE_INT64, &w->nbins, 1) &&
         append(w, kIdxBin1, H5T_NATIVE_INT64, &w->nnz, 1) &&
         write_attr_i64(w->file, "nnz", w->nnz) && write_attr_i64(w->file, "sum", w->sum);
  }
  close_handles(w);
  delete w;
  if (!ok) return fail(MCOOL_EIO, "finalising the file failed");
  return MCOOL_OK;
}

}  // extern "C"
