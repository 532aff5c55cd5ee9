// pandas-HDF5 tables: the on-disk layout of every continuous-energy and
// thermal-scattering table of the reference (HDF5DataSet<D>,
// src/HDF5DataSet.hpp:18-167).
//
// A file holds one pandas DataFrame with a single column, written with key
// "pandas", whose index is a D-level MultiIndex over the full product of its
// sorted levels (level 0 slowest, level D-1 fastest):
//   /pandas                  group
//   /pandas@axis1_nlevels    attribute, D
//   /pandas/axis1_level<i>   dataset, the values of level i (doubles)
//   /pandas/block0_values    dataset, the column (doubles, product of the
//                            level sizes, last level fastest)
// Only these are read (HDF5DataSet::ReadPandasAxis / ReadPandasValues,
// :87-130).  The writer puts the same items into a new file, with the column
// as an (n x 1) dataset as pandas stores it.
//
// Host code linked against the HDF5 C library; it has no GPU part.
#include <hdf5.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int Fail(const std::string& msg) {
  g_err = msg;
  return 1;
}

struct Quiet {  // no HDF5 error stack printed on failures we report ourselves
  H5E_auto2_t fn = nullptr;
  void* data = nullptr;
  Quiet() {
    H5Eget_auto2(H5E_DEFAULT, &fn, &data);
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  }
  ~Quiet() { H5Eset_auto2(H5E_DEFAULT, fn, data); }
};

struct Handle {  // closes an HDF5 id with the matching call
  hid_t id;
  herr_t (*close)(hid_t);
  Handle(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  ~Handle() {
    if (id >= 0) close(id);
  }
  Handle(const Handle&) = delete;
  Handle& operator=(const Handle&) = delete;
};

bool FileExists(const char* path) {
  FILE* f = fopen(path, "rb");
  if (!f) return false;
  fclose(f);
  return true;
}

int ReadDoubles(hid_t group, const std::string& name, std::vector<double>* out,
                uint64_t* n_out, const char* path) {
  Handle ds(H5Dopen2(group, name.c_str(), H5P_DEFAULT), H5Dclose);
  if (ds.id < 0) return Fail(std::string(path) + ": no dataset /pandas/" + name);
  Handle sp(H5Dget_space(ds.id), H5Sclose);
  const hssize_t n = H5Sget_simple_extent_npoints(sp.id);
  if (n < 0) return Fail(std::string(path) + ": unreadable dataset /pandas/" + name);
  *n_out = uint64_t(n);
  if (out) {
    out->resize(size_t(n));
    if (n > 0 && H5Dread(ds.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT,
                         out->data()) < 0)
      return Fail(std::string(path) + ": cannot read /pandas/" + name + " as double");
  }
  return 0;
}

}  // namespace

extern "C" {

const char* mmc_h5_last_error(void) { return g_err.c_str(); }

// Level count, level sizes (up to max_levels) and value count of a file.
// Messages follow HDF5DataSet::ReadPandasAxis: "File not found: <path>".
int mmc_h5_info(const char* path, int32_t* levels, uint64_t* sizes, int32_t max_levels,
                uint64_t* n_values) {
  if (!path || !levels || !n_values) return Fail("null argument");
  if (!FileExists(path)) return Fail(std::string("File not found: ") + path);
  Quiet q;
  Handle f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (f.id < 0) return Fail(std::string(path) + ": not an HDF5 file");
  Handle g(H5Gopen2(f.id, "/pandas", H5P_DEFAULT), H5Gclose);
  if (g.id < 0) return Fail(std::string(path) + ": no group /pandas");
  Handle a(H5Aopen(g.id, "axis1_nlevels", H5P_DEFAULT), H5Aclose);
  unsigned long long d = 0;
  if (a.id < 0 || H5Aread(a.id, H5T_NATIVE_ULLONG, &d) < 0)
    return Fail(std::string(path) + ": no attribute axis1_nlevels");
  *levels = int32_t(d);
  for (unsigned long long i = 0; i < d && int32_t(i) < max_levels; i++) {
    uint64_t n = 0;
    if (ReadDoubles(g.id, "axis1_level" + std::to_string(i), nullptr, &n, path)) return 1;
    if (sizes) sizes[i] = n;
  }
  return ReadDoubles(g.id, "block0_values", nullptr, n_values, path);
}

// Reads the levels into axes[i] (sized by mmc_h5_info) and the values.
int mmc_h5_read(const char* path, int32_t levels, double* const* axes, double* values) {
  if (!path || levels < 1 || !axes || !values) return Fail("null argument");
  if (!FileExists(path)) return Fail(std::string("File not found: ") + path);
  Quiet q;
  Handle f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (f.id < 0) return Fail(std::string(path) + ": not an HDF5 file");
  Handle g(H5Gopen2(f.id, "/pandas", H5P_DEFAULT), H5Gclose);
  if (g.id < 0) return Fail(std::string(path) + ": no group /pandas");
  std::vector<double> buf;
  uint64_t n = 0;
  for (int32_t i = 0; i < levels; i++) {
    if (ReadDoubles(g.id, "axis1_level" + std::to_string(i), &buf, &n, path)) return 1;
    for (uint64_t k = 0; k < n; k++) axes[i][k] = buf[k];
  }
  if (ReadDoubles(g.id, "block0_values", &buf, &n, path)) return 1;
  for (uint64_t k = 0; k < n; k++) values[k] = buf[k];
  return 0;
}

// Writes a new file (replacing any) with `levels` levels of sizes[i] values
// axes[i] and prod(sizes) values, last level fastest.
int mmc_h5_write(const char* path, int32_t levels, const uint64_t* sizes,
                 const double* const* axes, const double* values) {
  if (!path || levels < 1 || !sizes || !axes || !values) return Fail("null argument");
  uint64_t total = 1;
  for (int32_t i = 0; i < levels; i++) total *= sizes[i];
  Quiet q;
  Handle f(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
  if (f.id < 0) return Fail(std::string(path) + ": cannot create");
  Handle g(H5Gcreate2(f.id, "/pandas", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
  if (g.id < 0) return Fail(std::string(path) + ": cannot create /pandas");
  {
    Handle sp(H5Screate(H5S_SCALAR), H5Sclose);
    Handle a(H5Acreate2(g.id, "axis1_nlevels", H5T_STD_I64LE, sp.id, H5P_DEFAULT,
                        H5P_DEFAULT), H5Aclose);
    const long long d = levels;
    if (a.id < 0 || H5Awrite(a.id, H5T_NATIVE_LLONG, &d) < 0)
      return Fail(std::string(path) + ": cannot write axis1_nlevels");
  }
  auto put = [&](const std::string& name, int rank, const hsize_t* dims,
                 const double* data) -> int {
    Handle sp(H5Screate_simple(rank, dims, nullptr), H5Sclose);
    Handle ds(H5Dcreate2(g.id, name.c_str(), H5T_IEEE_F64LE, sp.id, H5P_DEFAULT,
                         H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    if (ds.id < 0 ||
        H5Dwrite(ds.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
      return Fail(std::string(path) + ": cannot write /pandas/" + name);
    return 0;
  };
  for (int32_t i = 0; i < levels; i++) {
    const hsize_t dims[1] = {hsize_t(sizes[i])};
    if (put("axis1_level" + std::to_string(i), 1, dims, axes[i])) return 1;
  }
  const hsize_t dims[2] = {hsize_t(total), 1};
  return put("block0_values", 2, dims, values);
}

}  // extern "C"
