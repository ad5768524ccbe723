// libfopis_io.so: plain HDF5 datasets (include/fopis_io.h) over the HDF5 C
// library. The reference stores its resume and capacity-matrix files through
// HighFive, which maps double / size_t / int to the native HDF5 types used
// here and a single value to a scalar dataspace (src/Misc/IO.cpp:28-120,
// src/Grid/RootGrid/RootGridIO.cpp:25-290, src/Misc/Environment.cpp:379-438).
#include "fopis_io.h"

#include <hdf5.h>

#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

struct fopis_h5 {
  hid_t file = -1;
};

namespace {

thread_local std::string g_err;

struct H5Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

template <class F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

void quiet() {
  static bool done = false;
  if (!done) {
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);  // errors come back as codes
    done = true;
  }
}

hid_t native(int type) {
  switch (type) {
    case FOPIS_IO_F64:
      return H5T_NATIVE_DOUBLE;
    case FOPIS_IO_U64:
      return H5T_NATIVE_UINT64;
    case FOPIS_IO_I32:
      return H5T_NATIVE_INT32;
    case FOPIS_IO_F32:
      return H5T_NATIVE_FLOAT;
    case FOPIS_IO_U8:
      return H5T_NATIVE_UINT8;
    default:
      throw H5Error("unknown element type");
  }
}

// Closes an HDF5 id at scope exit.
struct Id {
  hid_t id;
  herr_t (*close)(hid_t);
  Id(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  ~Id() {
    if (id >= 0) close(id);
  }
  Id(const Id&) = delete;
  Id& operator=(const Id&) = delete;
};

// Opens (creating when asked) every group on the way to the last path element.
hid_t parent_group(hid_t file, const std::string& name, bool create, std::string* leaf) {
  std::vector<std::string> parts;
  size_t b = 0;
  while (b <= name.size()) {
    size_t e = name.find('/', b);
    if (e == std::string::npos) e = name.size();
    if (e > b) parts.push_back(name.substr(b, e - b));
    b = e + 1;
  }
  if (parts.empty()) throw H5Error("empty dataset name");
  *leaf = parts.back();
  hid_t g = H5Gopen2(file, "/", H5P_DEFAULT);
  for (size_t q = 0; q + 1 < parts.size(); ++q) {
    hid_t next;
    if (H5Lexists(g, parts[q].c_str(), H5P_DEFAULT) > 0) {
      next = H5Gopen2(g, parts[q].c_str(), H5P_DEFAULT);
    } else if (create) {
      next = H5Gcreate2(g, parts[q].c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    } else {
      H5Gclose(g);
      throw H5Error("no group " + parts[q] + " on the way to " + name);
    }
    H5Gclose(g);
    if (next < 0) throw H5Error("cannot open group " + parts[q]);
    g = next;
  }
  return g;
}

hid_t open_dataset(hid_t file, const std::string& name) {
  std::string leaf;
  Id g(parent_group(file, name, false, &leaf), H5Gclose);
  if (H5Lexists(g.id, leaf.c_str(), H5P_DEFAULT) <= 0) throw H5Error("no dataset " + name);
  hid_t d = H5Dopen2(g.id, leaf.c_str(), H5P_DEFAULT);
  if (d < 0) throw H5Error("cannot open dataset " + name);
  return d;
}

}  // namespace

extern "C" {

const char* fopis_io_last_error(void) { return g_err.c_str(); }

int fopis_io_open(const char* path, int mode, fopis_h5** out) {
  return guarded([&] {
    if (!path || !out) throw H5Error("null argument");
    quiet();
    hid_t f = mode == 1 ? H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT)
                        : H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) throw H5Error(std::string(mode == 1 ? "cannot create " : "cannot open ") + path);
    *out = new fopis_h5{f};
  });
}

int fopis_io_close(fopis_h5* f) {
  return guarded([&] {
    if (!f) return;
    const herr_t r = f->file >= 0 ? H5Fclose(f->file) : 0;
    delete f;
    if (r < 0) throw H5Error("close failed");
  });
}

int fopis_io_write(fopis_h5* f, const char* name, int type, int rank, const int64_t* dims, const void* data) {
  return guarded([&] {
    if (!f || !name || (rank > 0 && !dims)) throw H5Error("null argument");
    if (rank < 0 || rank > 3) throw H5Error("rank must be 0..3");
    int64_t count = 1;
    hsize_t hd[3] = {0, 0, 0};
    for (int q = 0; q < rank; ++q) {
      if (dims[q] < 0) throw H5Error("negative dimension");
      hd[q] = (hsize_t)dims[q];
      count *= dims[q];
    }
    if (count > 0 && !data) throw H5Error("null data");
    std::string leaf;
    Id g(parent_group(f->file, name, true, &leaf), H5Gclose);
    if (H5Lexists(g.id, leaf.c_str(), H5P_DEFAULT) > 0) throw H5Error(std::string("dataset exists: ") + name);
    Id space(rank == 0 ? H5Screate(H5S_SCALAR) : H5Screate_simple(rank, hd, nullptr), H5Sclose);
    const hid_t t = native(type);
    Id d(H5Dcreate2(g.id, leaf.c_str(), t, space.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    if (d.id < 0) throw H5Error(std::string("cannot create dataset ") + name);
    if (count > 0 && H5Dwrite(d.id, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
      throw H5Error(std::string("write failed: ") + name);
  });
}

int fopis_io_exists(fopis_h5* f, const char* name) {
  if (!f || !name) return 0;
  try {
    Id d(open_dataset(f->file, name), H5Dclose);
    return 1;
  } catch (const std::exception&) {
    return 0;
  }
}

int fopis_io_info(fopis_h5* f, const char* name, int* rank, int64_t* dims) {
  return guarded([&] {
    if (!f || !name || !rank || !dims) throw H5Error("null argument");
    Id d(open_dataset(f->file, name), H5Dclose);
    Id s(H5Dget_space(d.id), H5Sclose);
    const int r = H5Sget_simple_extent_ndims(s.id);
    if (r < 0 || r > 8) throw H5Error(std::string("bad dataspace: ") + name);
    hsize_t hd[8];
    H5Sget_simple_extent_dims(s.id, hd, nullptr);
    *rank = r;
    for (int q = 0; q < r; ++q) dims[q] = (int64_t)hd[q];
  });
}

int fopis_io_read(fopis_h5* f, const char* name, int type, int64_t count, void* data) {
  return guarded([&] {
    if (!f || !name || (count > 0 && !data)) throw H5Error("null argument");
    Id d(open_dataset(f->file, name), H5Dclose);
    Id s(H5Dget_space(d.id), H5Sclose);
    const hssize_t n = H5Sget_simple_extent_npoints(s.id);
    if (n != (hssize_t)count)
      throw H5Error(std::string(name) + ": holds " + std::to_string((long long)n) + " elements, expected " +
                    std::to_string((long long)count));
    if (count > 0 && H5Dread(d.id, native(type), H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
      throw H5Error(std::string("read failed: ") + name);
  });
}

}  // extern "C"
