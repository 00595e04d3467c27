// io.cpp -- the raw-binary CSR format of the reference.
//
// Reference: loadCSRSparseMatrixBin (als/src/cuda/als.cu:914-933) reads
// three raw files: int32 row_ptr[m+1], int32 col[nnz], fp32 val[nnz], in
// host byte order, no header.  The reference opens all three but checks
// only two (it tests dFile twice, :920), ignores fread's result and prints
// on failure.  Here every file is checked, its size must match, and a
// status is returned; the C++ symbol of the reference (void return) is kept
// for binary callers.
#include <cstdio>
#include <cstdint>
#include <string>

#include "../../include/als.h"
#include "cumf_internal.h"

namespace {

int read_exact(const char* path, void* dst, size_t bytes) {
  if (!path) {
    cumf::set_last_error("csr file path is NULL");
    return CUMF_ERR_INVALID;
  }
  FILE* fp = std::fopen(path, "rb");
  if (!fp) {
    cumf::set_last_error(std::string("cannot open ") + path);
    return CUMF_ERR_INVALID;
  }
  int rc = CUMF_OK;
  if (std::fseek(fp, 0, SEEK_END) != 0 || std::ftell(fp) != (long)bytes) {
    cumf::set_last_error(std::string(path) + ": size is not " + std::to_string(bytes) + " bytes");
    rc = CUMF_ERR_INVALID;
  } else {
    std::rewind(fp);
    if (bytes && std::fread(dst, 1, bytes, fp) != bytes) {
      cumf::set_last_error(std::string("short read from ") + path);
      rc = CUMF_ERR_INVALID;
    }
  }
  std::fclose(fp);
  return rc;
}

int write_exact(const char* path, const void* src, size_t bytes) {
  FILE* fp = path ? std::fopen(path, "wb") : nullptr;
  if (!fp) {
    cumf::set_last_error(std::string("cannot create ") + (path ? path : "(null)"));
    return CUMF_ERR_INVALID;
  }
  const bool ok = !bytes || std::fwrite(src, 1, bytes, fp) == bytes;
  const bool closed = std::fclose(fp) == 0;
  if (!ok || !closed) {
    cumf::set_last_error(std::string("short write to ") + path);
    return CUMF_ERR_INVALID;
  }
  return CUMF_OK;
}

}  // namespace

extern "C" int cumf_load_csr_bin(const char* dataFile, const char* rowFile, const char* colFile,
                                 float* data, int* row, int* col, int m, long nnz) {
  cumf::clear_last_error();
  if (m < 0 || nnz < 0 || !data || !row || !col) {
    cumf::set_last_error("cumf_load_csr_bin: bad arguments");
    return CUMF_ERR_INVALID;
  }
  int rc = read_exact(rowFile, row, sizeof(int) * ((size_t)m + 1));
  if (rc == CUMF_OK) rc = read_exact(colFile, col, sizeof(int) * (size_t)nnz);
  if (rc == CUMF_OK) rc = read_exact(dataFile, data, sizeof(float) * (size_t)nnz);
  if (rc == CUMF_OK && (row[0] != 0 || row[m] != nnz)) {
    cumf::set_last_error("cumf_load_csr_bin: row pointers must start at 0 and end at nnz");
    rc = CUMF_ERR_INVALID;
  }
  return rc;
}

extern "C" int cumf_save_csr_bin(const char* dataFile, const char* rowFile, const char* colFile,
                                 const float* data, const int* row, const int* col, int m, long nnz) {
  cumf::clear_last_error();
  if (m < 0 || nnz < 0 || !row || (nnz > 0 && (!data || !col))) {
    cumf::set_last_error("cumf_save_csr_bin: bad arguments");
    return CUMF_ERR_INVALID;
  }
  int rc = write_exact(rowFile, row, sizeof(int) * ((size_t)m + 1));
  if (rc == CUMF_OK) rc = write_exact(colFile, col, sizeof(int) * (size_t)nnz);
  if (rc == CUMF_OK) rc = write_exact(dataFile, data, sizeof(float) * (size_t)nnz);
  return rc;
}

// Same name, C++ linkage and parameters as the reference
// (als/src/cuda/als.cu:914-916); failures are reported through
// cumf_last_error() instead of a printf.
void loadCSRSparseMatrixBin(const char* dataFile, const char* rowFile, const char* colFile, float* data,
                            int* row, int* col, const int m, const long nnz) {
  (void)cumf_load_csr_bin(dataFile, rowFile, colFile, data, row, col, m, nnz);
}
