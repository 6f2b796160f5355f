// HDF5 access for the drop-in host: the orbital file the reference reads and
// the QMCPACK THC file it writes (src/h5helper.cpp, src/distributed_matrix.h:
// 115-197 and 315-330).  libhdf5 (C API, 1.10) is opened at run time with
// dlopen, from THC_HDF5_LIB or the usual install locations, so the library
// does not link against it.
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace thc {
namespace h5 {

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// True when a usable libhdf5 was found (does not throw).
bool available();

class File {
 public:
  enum Mode { ReadOnly, ReadWrite, Truncate };
  File(const std::string& path, Mode mode);
  ~File();
  File(const File&) = delete;
  File& operator=(const File&) = delete;

  bool exists(const std::string& name) const;
  std::vector<uint64_t> dims(const std::string& name) const;
  // Reads a real dataset (f64 or int) as double / int.
  std::vector<double> read_f64(const std::string& name, std::vector<uint64_t>* dims = nullptr) const;
  std::vector<int> read_int(const std::string& name, std::vector<uint64_t>* dims = nullptr) const;
  // Reads a complex dataset stored as h5py's compound {r, i}, or a real f64
  // dataset (imaginary parts zero); returns interleaved (re, im).
  std::vector<double> read_complex(const std::string& name, std::vector<uint64_t>* dims = nullptr) const;

  void ensure_group(const std::string& path);
  void write_f64(const std::string& name, const double* data, const std::vector<uint64_t>& dims,
                 bool gzip = false);
  void write_int(const std::string& name, const int* data, const std::vector<uint64_t>& dims);
  // Complex as QMCPACK expects: f64 with a trailing dimension of 2
  // (src/h5helper.cpp:63-76); gzip level 6, chunks of 20 (src/h5helper.cpp:28-35).
  void write_complex_as_f64(const std::string& name, const double* interleaved,
                            const std::vector<uint64_t>& dims, bool gzip);
  // Writes a complex dataset as h5py's compound {r, i} (for orbital files).
  void write_compound_complex(const std::string& name, const double* interleaved,
                              const std::vector<uint64_t>& dims);

 private:
  int64_t id_ = -1;
};

}  // namespace h5
}  // namespace thc
