// HDF5 I/O through a run-time-loaded libhdf5 (C API of HDF5 1.10).
#include "h5io.h"

#include <dlfcn.h>

#include <cstdlib>
#include <cstring>
#include <mutex>

namespace thc {
namespace h5 {
namespace {

typedef int64_t hid_t;
typedef int herr_t;
typedef unsigned long long hsize_t;
typedef int htri_t;

// HDF5 1.10 constants used here (H5Fpublic.h, H5Ppublic.h, H5Spublic.h, H5Tpublic.h).
constexpr unsigned F_ACC_RDONLY = 0x0000u, F_ACC_RDWR = 0x0001u, F_ACC_TRUNC = 0x0002u;
constexpr hid_t P_DEFAULT = 0, S_ALL = 0;
constexpr int T_COMPOUND = 6;  // H5T_COMPOUND in H5T_class_t
constexpr int T_INTEGER = 0, T_FLOAT = 1;

struct Api {
  void* h = nullptr;
  herr_t (*open)();
  hid_t (*fopen)(const char*, unsigned, hid_t);
  hid_t (*fcreate)(const char*, unsigned, hid_t, hid_t);
  herr_t (*fclose)(hid_t);
  hid_t (*dopen)(hid_t, const char*, hid_t);
  hid_t (*dcreate)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t, hid_t);
  herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void*);
  herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void*);
  herr_t (*dclose)(hid_t);
  hid_t (*dget_space)(hid_t);
  hid_t (*dget_type)(hid_t);
  int (*s_ndims)(hid_t);
  int (*s_dims)(hid_t, hsize_t*, hsize_t*);
  hid_t (*s_create)(int, const hsize_t*, const hsize_t*);
  herr_t (*sclose)(hid_t);
  hid_t (*tcreate)(int, size_t);
  herr_t (*tinsert)(hid_t, const char*, size_t, hid_t);
  herr_t (*tclose)(hid_t);
  int (*tclass)(hid_t);
  hid_t (*gcreate)(hid_t, const char*, hid_t, hid_t, hid_t);
  hid_t (*gopen)(hid_t, const char*, hid_t);
  herr_t (*gclose)(hid_t);
  hid_t (*pcreate)(hid_t);
  herr_t (*pset_chunk)(hid_t, int, const hsize_t*);
  herr_t (*pset_deflate)(hid_t, unsigned);
  herr_t (*pclose)(hid_t);
  htri_t (*lexists)(hid_t, const char*, hid_t);
  herr_t (*eset_auto)(hid_t, void*, void*);
  hid_t* native_double;
  hid_t* native_int;
  hid_t* cls_dataset_create;
  bool ok = false;
};

template <class F>
bool sym(void* h, const char* name, F& f) {
  f = reinterpret_cast<F>(dlsym(h, name));
  return f != nullptr;
}

Api& api() {
  static Api a;
  static std::once_flag once;
  std::call_once(once, [] {
    const char* cands[] = {std::getenv("THC_HDF5_LIB"), "/opt/conda/lib/libhdf5.so", "libhdf5.so",
                           "libhdf5_serial.so", "/usr/lib/x86_64-linux-gnu/hdf5/serial/libhdf5.so"};
    for (const char* c : cands) {
      if (!c) continue;
      a.h = dlopen(c, RTLD_NOW | RTLD_LOCAL);
      if (a.h) break;
    }
    if (!a.h) return;
    bool ok = sym(a.h, "H5open", a.open) && sym(a.h, "H5Fopen", a.fopen) && sym(a.h, "H5Fcreate", a.fcreate) &&
              sym(a.h, "H5Fclose", a.fclose) && sym(a.h, "H5Dopen2", a.dopen) && sym(a.h, "H5Dcreate2", a.dcreate) &&
              sym(a.h, "H5Dread", a.dread) && sym(a.h, "H5Dwrite", a.dwrite) && sym(a.h, "H5Dclose", a.dclose) &&
              sym(a.h, "H5Dget_space", a.dget_space) && sym(a.h, "H5Dget_type", a.dget_type) &&
              sym(a.h, "H5Sget_simple_extent_ndims", a.s_ndims) &&
              sym(a.h, "H5Sget_simple_extent_dims", a.s_dims) && sym(a.h, "H5Screate_simple", a.s_create) &&
              sym(a.h, "H5Sclose", a.sclose) && sym(a.h, "H5Tcreate", a.tcreate) &&
              sym(a.h, "H5Tinsert", a.tinsert) && sym(a.h, "H5Tclose", a.tclose) &&
              sym(a.h, "H5Tget_class", a.tclass) && sym(a.h, "H5Gcreate2", a.gcreate) &&
              sym(a.h, "H5Gopen2", a.gopen) && sym(a.h, "H5Gclose", a.gclose) && sym(a.h, "H5Pcreate", a.pcreate) &&
              sym(a.h, "H5Pset_chunk", a.pset_chunk) && sym(a.h, "H5Pset_deflate", a.pset_deflate) &&
              sym(a.h, "H5Pclose", a.pclose) && sym(a.h, "H5Lexists", a.lexists) &&
              sym(a.h, "H5Eset_auto2", a.eset_auto) && sym(a.h, "H5T_NATIVE_DOUBLE_g", a.native_double) &&
              sym(a.h, "H5T_NATIVE_INT_g", a.native_int) &&
              sym(a.h, "H5P_CLS_DATASET_CREATE_ID_g", a.cls_dataset_create);
    if (ok && a.open() >= 0) {
      a.eset_auto(0, nullptr, nullptr);  // errors are reported through return codes
      a.ok = true;
    }
  });
  if (!a.ok) throw Error("libhdf5 could not be loaded (set THC_HDF5_LIB to its path)");
  return a;
}

void check(bool good, const std::string& what) {
  if (!good) throw Error("HDF5: " + what);
}

struct Dset {
  hid_t id;
  explicit Dset(hid_t file, const std::string& name) {
    id = api().dopen(file, name.c_str(), P_DEFAULT);
    check(id >= 0, "cannot open dataset " + name);
  }
  ~Dset() { api().dclose(id); }
  std::vector<uint64_t> dims() const {
    hid_t sp = api().dget_space(id);
    const int nd = api().s_ndims(sp);
    std::vector<hsize_t> d(nd > 0 ? nd : 0);
    if (nd > 0) api().s_dims(sp, d.data(), nullptr);
    api().sclose(sp);
    return std::vector<uint64_t>(d.begin(), d.end());
  }
  int type_class() const {
    hid_t t = api().dget_type(id);
    const int c = api().tclass(t);
    api().tclose(t);
    return c;
  }
};

size_t count(const std::vector<uint64_t>& d) {
  size_t n = 1;
  for (auto v : d) n *= v;
  return n;
}

hid_t complex_type() {
  hid_t t = api().tcreate(T_COMPOUND, 2 * sizeof(double));
  api().tinsert(t, "r", 0, *api().native_double);
  api().tinsert(t, "i", sizeof(double), *api().native_double);
  return t;
}

}  // namespace

bool available() {
  try {
    api();
    return true;
  } catch (const Error&) {
    return false;
  }
}

File::File(const std::string& path, Mode mode) {
  Api& a = api();
  if (mode == Truncate) id_ = a.fcreate(path.c_str(), F_ACC_TRUNC, P_DEFAULT, P_DEFAULT);
  else id_ = a.fopen(path.c_str(), mode == ReadOnly ? F_ACC_RDONLY : F_ACC_RDWR, P_DEFAULT);
  check(id_ >= 0, "cannot open " + path);
}

File::~File() {
  if (id_ >= 0) api().fclose(id_);
}

bool File::exists(const std::string& name) const {
  // walk the path so missing intermediate groups report false
  std::string cur;
  size_t pos = 0;
  while (pos != std::string::npos) {
    size_t next = name.find('/', pos + 1);
    cur = name.substr(0, next);
    if (!cur.empty() && cur != "/" && api().lexists(id_, cur.c_str(), P_DEFAULT) <= 0) return false;
    pos = next;
  }
  return true;
}

std::vector<uint64_t> File::dims(const std::string& name) const { return Dset(id_, name).dims(); }

std::vector<double> File::read_f64(const std::string& name, std::vector<uint64_t>* dims) const {
  Dset d(id_, name);
  auto dm = d.dims();
  std::vector<double> out(count(dm));
  check(api().dread(d.id, *api().native_double, S_ALL, S_ALL, P_DEFAULT, out.data()) >= 0, "read " + name);
  if (dims) *dims = dm;
  return out;
}

std::vector<int> File::read_int(const std::string& name, std::vector<uint64_t>* dims) const {
  Dset d(id_, name);
  auto dm = d.dims();
  std::vector<int> out(count(dm));
  check(api().dread(d.id, *api().native_int, S_ALL, S_ALL, P_DEFAULT, out.data()) >= 0, "read " + name);
  if (dims) *dims = dm;
  return out;
}

std::vector<double> File::read_complex(const std::string& name, std::vector<uint64_t>* dims) const {
  Dset d(id_, name);
  auto dm = d.dims();
  const size_t n = count(dm);
  std::vector<double> out(2 * n, 0.0);
  if (d.type_class() == T_COMPOUND) {
    hid_t t = complex_type();
    const herr_t e = api().dread(d.id, t, S_ALL, S_ALL, P_DEFAULT, out.data());
    api().tclose(t);
    check(e >= 0, "read complex " + name);
  } else {
    check(d.type_class() == T_FLOAT || d.type_class() == T_INTEGER, "unsupported type for " + name);
    std::vector<double> re(n);
    check(api().dread(d.id, *api().native_double, S_ALL, S_ALL, P_DEFAULT, re.data()) >= 0, "read " + name);
    for (size_t i = 0; i < n; ++i) out[2 * i] = re[i];
  }
  if (dims) *dims = dm;
  return out;
}

void File::ensure_group(const std::string& path) {
  size_t pos = 0;
  while (pos != std::string::npos) {
    size_t next = path.find('/', pos + 1);
    std::string cur = path.substr(0, next);
    if (!cur.empty() && cur != "/") {
      if (api().lexists(id_, cur.c_str(), P_DEFAULT) <= 0) {
        hid_t g = api().gcreate(id_, cur.c_str(), P_DEFAULT, P_DEFAULT, P_DEFAULT);
        check(g >= 0, "create group " + cur);
        api().gclose(g);
      }
    }
    pos = next;
  }
}

namespace {
void write_raw(hid_t file, const std::string& name, hid_t type, const void* data, const std::vector<uint64_t>& dims,
               bool gzip) {
  std::vector<hsize_t> d(dims.begin(), dims.end());
  hid_t sp = api().s_create((int)d.size(), d.data(), nullptr);
  hid_t pl = P_DEFAULT;
  if (gzip) {
    pl = api().pcreate(*api().cls_dataset_create);
    std::vector<hsize_t> ch(d.size());
    for (size_t i = 0; i < d.size(); ++i) ch[i] = d[i] < 20 ? d[i] : 20;
    api().pset_chunk(pl, (int)ch.size(), ch.data());
    api().pset_deflate(pl, 6);
  }
  hid_t ds = api().dcreate(file, name.c_str(), type, sp, P_DEFAULT, pl, P_DEFAULT);
  const bool ok = ds >= 0 && api().dwrite(ds, type, S_ALL, S_ALL, P_DEFAULT, data) >= 0;
  if (ds >= 0) api().dclose(ds);
  if (gzip) api().pclose(pl);
  api().sclose(sp);
  check(ok, "write " + name);
}
}  // namespace

void File::write_f64(const std::string& name, const double* data, const std::vector<uint64_t>& dims, bool gzip) {
  write_raw(id_, name, *api().native_double, data, dims, gzip);
}

void File::write_int(const std::string& name, const int* data, const std::vector<uint64_t>& dims) {
  write_raw(id_, name, *api().native_int, data, dims, false);
}

void File::write_complex_as_f64(const std::string& name, const double* interleaved, const std::vector<uint64_t>& dims,
                                bool gzip) {
  std::vector<uint64_t> d = dims;
  d.push_back(2);
  write_raw(id_, name, *api().native_double, interleaved, d, gzip);
}

void File::write_compound_complex(const std::string& name, const double* interleaved,
                                  const std::vector<uint64_t>& dims) {
  hid_t t = complex_type();
  try {
    write_raw(id_, name, t, interleaved, dims, false);
  } catch (...) {
    api().tclose(t);
    throw;
  }
  api().tclose(t);
}

}  // namespace h5
}  // namespace thc
