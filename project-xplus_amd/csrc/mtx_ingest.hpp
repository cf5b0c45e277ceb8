// mtx_ingest.hpp -- MatrixMarket ingestion and the binary CSR cache.
//
// read_mtx_matrix / read_mtx_vector / reference_solution follow the dataset
// converter scripts/download_suitesparse_data.py:
//   parse_coordinate_matrix (:224-284)  coordinate storage, real|integer|pattern
//                                       field, general|symmetric symmetry;
//                                       explicit zeros skipped, duplicates summed
//                                       in file order, the mirror (j,i) of every
//                                       off-diagonal symmetric entry added, the
//                                       leading subset_n x subset_n principal
//                                       submatrix kept, columns ascending, sums
//                                       that cancel to 0 dropped;
//   parse_vector            (:287-336)  array or coordinate right-hand sides;
//   reference_solution      (:339-349)  the smooth x_ref with b = A x_ref when
//                                       the collection has no b;
//   convert_dataset         (:418-444)  x0 = 0.
// Same error messages.  One deliberate difference: a 0 row or column index
// (MatrixMarket indices start at 1) is an error here; the converter would
// wrap it to the last row.
//
// The binary cache is this library's own format (the reference reparses the
// text files on every run): a 128-byte header, then row_ptr (int32, n+1),
// col_idx (int32, nnz), values (f64, nnz), b (f64, n), x0 (f64, n), all
// little-endian, with a 64-bit checksum of the payload and the sizes and
// modification times of the five text files it was made from.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "generators.hpp"

namespace xp {

struct MtxMatrix {
    int n = 0, m = 0;                           // shape after the leading-principal subset
    std::vector<int> row_ptr, col_idx;
    std::vector<double> values;
    long long nrows = 0, ncols = 0, nnz_reported = 0;   // the size line
    int field = 0;                              // 0 real, 1 integer, 2 pattern
    int symmetry = 0;                           // 0 general, 1 symmetric
    long long entries = 0;                      // entry lines read
    bool has_diag = false;                      // any nonzero diagonal entry kept
    double diag_min_abs = 0.0;
};

MtxMatrix read_mtx_matrix(const std::string& path, long long subset_n);
std::vector<double> read_mtx_vector(const std::string& path, int expected_n, long long subset_n);
std::vector<double> reference_solution(int n);
// A dataset from a MatrixMarket matrix (and optional right-hand side file):
// b from b_path, else b = A x_ref; x0 = 0.  Throws when not square.
HostCsr mtx_dataset(const std::string& mtx_path, const std::string& b_path, long long subset_n,
                    MtxMatrix* info);

// y = A x over CSR rows in storage order, FP64, one sequential sum per row:
// the converter's csr_spmv (download_suitesparse_data.py:352-358), so that b
// is bit-identical to the one it writes.  Data preparation only; solves run
// their SpMV on the GPU.
void csr_spmv_host_sequential(const HostCsr& a, const std::vector<double>& x, std::vector<double>& y);

// Binary cache.  `stamp` carries the sizes and modification times of the text
// files (zeros when the dataset has no text source).
struct SourceStamp {
    uint64_t size[5] = {0, 0, 0, 0, 0};
    int64_t mtime_ns[5] = {0, 0, 0, 0, 0};
    bool operator==(const SourceStamp& o) const;
};
SourceStamp stamp_dataset_dir(const std::string& dir);   // throws if a file is missing
void save_binary(const HostCsr& a, const std::string& path, const SourceStamp& stamp);
HostCsr load_binary(const std::string& path, SourceStamp* stamp);
// Text dataset through the cache: a cache whose stamp matches the directory's
// files is read; otherwise the text files are parsed (CsrDataset::load) and
// the cache is rewritten when the location is writable.
HostCsr load_cached(const std::string& dir, const std::string& cache_path, bool* hit);
uint64_t payload_checksum(const void* data, size_t bytes, uint64_t h);

}  // namespace xp
