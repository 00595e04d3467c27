// io.cpp -- the raw-binary CSR format of the reference.
//
// Reference: loadCSRSparseMatrixBin (als/src/cuda/als.cu:914-933) reads
// three raw files: int32 row_ptr[m+1], int32 col[nnz], fp32 val[nnz], in
// host byte order, no header.  The reference opens all three but checks
// only two (it tests dFile twice, :920), ignores fread's result and prints
// on failure.  Here every file is checked, its size must match, and a
// status is returned; the C++ symbol of the reference (void return) is kept
// for binary callers.
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

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

// ---------------------------------------------------------------------------
// Held-out split of an ingested rating matrix.  The reference's files carry
// no test set (its Spark caller splits the data before the GPU sees it), so
// a run from raw CSR files needs one made here: every rating is held out
// with probability `fraction`, decided by a hash of (seed, rating index) so
// the split does not depend on thread count or platform; then every row and
// every column with ratings keeps at least one training rating (of its
// held-out ratings, the one with the smallest hash goes back to training).
// ---------------------------------------------------------------------------
namespace {

uint64_t mix64(uint64_t x) {  // splitmix64 finaliser
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

}  // namespace

extern "C" int cumf_split_holdout(int64_t m, int64_t n, int64_t nnz, const int32_t* row_ptr, const int32_t* col,
                                  const float* val, double fraction, uint64_t seed, int32_t* train_row_ptr,
                                  int32_t* train_col, float* train_val, int64_t* train_nnz, int32_t* test_u,
                                  int32_t* test_v, float* test_r, int64_t* test_nnz) {
  cumf::clear_last_error();
  if (m < 0 || n < 0 || nnz < 0 || nnz > INT32_MAX || m > INT32_MAX || n > INT32_MAX || !row_ptr ||
      !train_row_ptr || !train_nnz || !test_nnz || !(fraction >= 0.0 && fraction <= 1.0) ||
      (nnz > 0 && (!col || !val || !train_col || !train_val || !test_u || !test_v || !test_r))) {
    cumf::set_last_error("cumf_split_holdout: bad arguments");
    return CUMF_ERR_INVALID;
  }
  if (row_ptr[0] != 0 || row_ptr[m] != nnz) {
    cumf::set_last_error("cumf_split_holdout: row pointers must start at 0 and end at nnz");
    return CUMF_ERR_INVALID;
  }
  for (int64_t u = 0; u < m; ++u)
    if (row_ptr[u + 1] < row_ptr[u]) {
      cumf::set_last_error("cumf_split_holdout: row pointers decrease at row " + std::to_string(u));
      return CUMF_ERR_INVALID;
    }
  for (int64_t k = 0; k < nnz; ++k)
    if (col[k] < 0 || col[k] >= n) {
      cumf::set_last_error("cumf_split_holdout: column index out of range at rating " + std::to_string(k));
      return CUMF_ERR_INVALID;
    }
  const uint64_t key = mix64(seed);
  auto hash = [&](int64_t k) { return mix64(key ^ (uint64_t)k); };
  // 2^-53 * (53-bit hash) < fraction
  std::vector<uint8_t> held((size_t)nnz);
  for (int64_t k = 0; k < nnz; ++k) held[k] = (double)(hash(k) >> 11) * 0x1.0p-53 < fraction;
  // every row with ratings keeps one
  for (int64_t u = 0; u < m; ++u) {
    const int64_t s = row_ptr[u], e = row_ptr[u + 1];
    if (s == e) continue;
    bool any = false;
    int64_t best = -1;
    uint64_t bh = 0;
    for (int64_t k = s; k < e; ++k) {
      if (!held[k]) {
        any = true;
        break;
      }
      const uint64_t hk = hash(k);
      if (best < 0 || hk < bh) {
        best = k;
        bh = hk;
      }
    }
    if (!any) held[best] = 0;
  }
  // every column with ratings keeps one
  std::vector<int64_t> ntrain((size_t)n, 0), best((size_t)n, -1);
  std::vector<uint64_t> bh((size_t)n, 0);
  for (int64_t k = 0; k < nnz; ++k) {
    const int32_t c = col[k];
    if (!held[k]) {
      ++ntrain[c];
    } else {
      const uint64_t hk = hash(k);
      if (best[c] < 0 || hk < bh[c]) {
        best[c] = k;
        bh[c] = hk;
      }
    }
  }
  for (int64_t c = 0; c < n; ++c)
    if (ntrain[c] == 0 && best[c] >= 0) held[best[c]] = 0;
  int64_t nt = 0, nh = 0;
  train_row_ptr[0] = 0;
  for (int64_t u = 0; u < m; ++u) {
    for (int64_t k = row_ptr[u]; k < row_ptr[u + 1]; ++k) {
      if (held[k]) {
        test_u[nh] = (int32_t)u;
        test_v[nh] = col[k];
        test_r[nh] = val[k];
        ++nh;
      } else {
        train_col[nt] = col[k];
        train_val[nt] = val[k];
        ++nt;
      }
    }
    train_row_ptr[u + 1] = (int32_t)nt;
  }
  *train_nnz = nt;
  *test_nnz = nh;
  return CUMF_OK;
}
