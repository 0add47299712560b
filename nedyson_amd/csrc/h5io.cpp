// GREEN reader/writer in the reference's HDF5 layout (green_func::
// print_to_file / read_from_file, src/greens.cpp:199-306): one group per
// function holding the flat complex datasets GR ((nt+1)(nt+2)/2 blocks), GL
// (same), GTV ((nt+1)(ntau+1) blocks), GM (ntau+1 blocks) and the int scalars
// nt, ntau, sig, nao.  Complex numbers are the compound {r, i} of two doubles
// (the layout of std::complex<double>).  The reference's writer stores the
// Matsubara block as "M" while its reader expects "GM"; this writer stores GM,
// so that its files read back in both.
//
// Built as its own library (libnedyson_h5.so) on top of the C ABI, so that the
// compute library does not depend on HDF5.  Device contexts are read and
// written in chunks of blocks through ned_green_read / ned_green_write.
#include "../../include/nedyson.h"
#include "../../include/nedyson_h5.h"

#include <hdf5.h>

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

struct H5Err : std::runtime_error {
  explicit H5Err(const std::string &m) : std::runtime_error(m) {}
};

// RAII for hid_t handles
struct Hid {
  hid_t id = -1;
  herr_t (*close)(hid_t) = nullptr;
  Hid(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  Hid(const Hid &) = delete;
  Hid &operator=(const Hid &) = delete;
  ~Hid() {
    if (id >= 0 && close) close(id);
  }
  operator hid_t() const { return id; }
};

hid_t complex_type() {
  hid_t t = H5Tcreate(H5T_COMPOUND, 16);
  H5Tinsert(t, "r", 0, H5T_NATIVE_DOUBLE);
  H5Tinsert(t, "i", 8, H5T_NATIVE_DOUBLE);
  return t;
}

std::string join(const char *group, const char *name) {
  std::string g = group ? group : "";
  while (!g.empty() && g.back() == '/') g.pop_back();
  return g.empty() ? std::string("/") + name : (g[0] == '/' ? g : "/" + g) + "/" + name;
}

bool exists(hid_t f, const std::string &path) {
  // every prefix must exist for H5Lexists to be safe
  size_t pos = 1;
  while (true) {
    size_t nxt = path.find('/', pos);
    std::string pre = path.substr(0, nxt);
    if (H5Lexists(f, pre.c_str(), H5P_DEFAULT) <= 0) return false;
    if (nxt == std::string::npos) return true;
    pos = nxt + 1;
  }
}

int read_int(hid_t f, const std::string &path) {
  Hid d(H5Dopen2(f, path.c_str(), H5P_DEFAULT), H5Dclose);
  if (d < 0) throw H5Err("missing dataset " + path);
  int v = 0;
  if (H5Dread(d, H5T_NATIVE_INT, H5S_ALL, H5S_ALL, H5P_DEFAULT, &v) < 0) throw H5Err("cannot read " + path);
  return v;
}

void write_int(hid_t f, const std::string &path, int v) {
  if (exists(f, path)) H5Ldelete(f, path.c_str(), H5P_DEFAULT);
  Hid sp(H5Screate(H5S_SCALAR), H5Sclose);
  Hid lcpl(H5Pcreate(H5P_LINK_CREATE), H5Pclose);
  H5Pset_create_intermediate_group(lcpl, 1);
  Hid d(H5Dcreate2(f, path.c_str(), H5T_NATIVE_INT, sp, lcpl, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
  if (d < 0 || H5Dwrite(d, H5T_NATIVE_INT, H5S_ALL, H5S_ALL, H5P_DEFAULT, &v) < 0) throw H5Err("cannot write " + path);
}

// a flat complex dataset of nblk blocks of es complex numbers
struct CDataset {
  Hid d, ct;
  hsize_t len = 0;
  CDataset(hid_t f, const std::string &path, bool create, hsize_t n)
      : d(-1, H5Dclose), ct(complex_type(), H5Tclose) {
    if (create) {
      if (exists(f, path)) H5Ldelete(f, path.c_str(), H5P_DEFAULT);
      Hid sp(H5Screate_simple(1, &n, nullptr), H5Sclose);
      Hid lcpl(H5Pcreate(H5P_LINK_CREATE), H5Pclose);
      H5Pset_create_intermediate_group(lcpl, 1);
      d.id = H5Dcreate2(f, path.c_str(), ct, sp, lcpl, H5P_DEFAULT, H5P_DEFAULT);
      if (d < 0) throw H5Err("cannot create " + path);
      len = n;
    } else {
      d.id = H5Dopen2(f, path.c_str(), H5P_DEFAULT);
      if (d < 0) throw H5Err("missing dataset " + path);
      Hid sp(H5Dget_space(d), H5Sclose);
      hssize_t np = H5Sget_simple_extent_npoints(sp);
      if (np < 0) throw H5Err("bad dataspace " + path);
      len = (hsize_t)np;
    }
  }
  // elements [first, first+count) <-> buf
  void io(bool write, hsize_t first, hsize_t count, void *buf) {
    if (count == 0) return;
    if (first + count > len) throw H5Err("dataset shorter than its GREEN sizes");
    Hid fs(H5Dget_space(d), H5Sclose);
    H5Sselect_hyperslab(fs, H5S_SELECT_SET, &first, nullptr, &count, nullptr);
    Hid ms(H5Screate_simple(1, &count, nullptr), H5Sclose);
    herr_t r = write ? H5Dwrite(d, ct, ms, fs, H5P_DEFAULT, buf) : H5Dread(d, ct, ms, fs, H5P_DEFAULT, buf);
    if (r < 0) throw H5Err("hdf5 transfer failed");
  }
};

hid_t open_file(const char *file, bool write) {
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  if (!write) {
    hid_t f = H5Fopen(file, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) throw H5Err(std::string("cannot open ") + file);
    return f;
  }
  hid_t f = -1;
  if (H5Fis_hdf5(file) > 0) f = H5Fopen(file, H5F_ACC_RDWR, H5P_DEFAULT);
  if (f < 0) f = H5Fcreate(file, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (f < 0) throw H5Err(std::string("cannot create ") + file);
  return f;
}

long long tri(int nt) { return (long long)(nt + 1) * (nt + 2) / 2; }

template <class F>
ned_status guard(F &&f) {
  try {
    f();
    return NED_OK;
  } catch (const H5Err &e) {
    g_err = e.what();
    return NED_ERR_ARG;
  } catch (const std::bad_alloc &) {
    g_err = "out of host memory";
    return NED_ERR_NOMEM;
  } catch (const std::exception &e) {
    g_err = e.what();
    return NED_ERR_DEVICE;
  }
}

void check_ned(ned_status s) {
  if (s != NED_OK) throw std::runtime_error(std::string("nedyson_amd: ") + ned_last_error());
}

constexpr long long CHUNK = 4096;  // blocks per transfer through the host

}  // namespace

extern "C" {

const char *ned_h5_last_error(void) { return g_err.c_str(); }

ned_status ned_h5_green_dims(const char *file, const char *group, int *nt, int *ntau, int *nao, int *sig) {
  if (!file || !nt || !ntau || !nao || !sig) { g_err = "null argument"; return NED_ERR_ARG; }
  return guard([&] {
    Hid f(open_file(file, false), H5Fclose);
    *nt = read_int(f, join(group, "nt"));
    *ntau = read_int(f, join(group, "ntau"));
    *nao = read_int(f, join(group, "nao"));
    *sig = read_int(f, join(group, "sig"));
  });
}

ned_status ned_h5_green_read_host(const char *file, const char *group, int nt, int ntau, int nao, void *ret,
                                  void *les, void *tv, void *mat) {
  if (!file) { g_err = "null argument"; return NED_ERR_ARG; }
  return guard([&] {
    Hid f(open_file(file, false), H5Fclose);
    const int fnt = read_int(f, join(group, "nt")), fntau = read_int(f, join(group, "ntau")),
              fnao = read_int(f, join(group, "nao"));
    if (fnt != nt || fntau != ntau || fnao != nao) throw H5Err("GREEN sizes differ from the file's");
    const hsize_t es = (hsize_t)nao * nao;
    if (ret) CDataset(f, join(group, "GR"), false, 0).io(false, 0, tri(nt) * es, ret);
    if (les) CDataset(f, join(group, "GL"), false, 0).io(false, 0, tri(nt) * es, les);
    if (tv) CDataset(f, join(group, "GTV"), false, 0).io(false, 0, (hsize_t)(nt + 1) * (ntau + 1) * es, tv);
    if (mat) CDataset(f, join(group, "GM"), false, 0).io(false, 0, (hsize_t)(ntau + 1) * es, mat);
  });
}

ned_status ned_h5_green_write_host(const char *file, const char *group, int nt, int ntau, int nao, int sig,
                                   const void *ret, const void *les, const void *tv, const void *mat) {
  if (!file || !ret || !les || !tv || !mat) { g_err = "null argument"; return NED_ERR_ARG; }
  if (nt < 0 || ntau < 0 || nao <= 0) { g_err = "bad sizes"; return NED_ERR_ARG; }
  return guard([&] {
    Hid f(open_file(file, true), H5Fclose);
    const hsize_t es = (hsize_t)nao * nao;
    CDataset(f, join(group, "GM"), true, (ntau + 1) * es).io(true, 0, (ntau + 1) * es, const_cast<void *>(mat));
    CDataset(f, join(group, "GR"), true, tri(nt) * es).io(true, 0, tri(nt) * es, const_cast<void *>(ret));
    CDataset(f, join(group, "GL"), true, tri(nt) * es).io(true, 0, tri(nt) * es, const_cast<void *>(les));
    CDataset(f, join(group, "GTV"), true, (hsize_t)(nt + 1) * (ntau + 1) * es)
        .io(true, 0, (hsize_t)(nt + 1) * (ntau + 1) * es, const_cast<void *>(tv));
    write_int(f, join(group, "nt"), nt);
    write_int(f, join(group, "ntau"), ntau);
    write_int(f, join(group, "sig"), sig);
    write_int(f, join(group, "nao"), nao);
    H5Fflush(f, H5F_SCOPE_LOCAL);
  });
}

ned_status ned_h5_green_read(ned_ctx *ctx, int which, const char *file, const char *group) {
  if (!ctx || !file) { g_err = "null argument"; return NED_ERR_ARG; }
  return guard([&] {
    ned_dims dm;
    check_ned(ned_ctx_get_dims(ctx, &dm));
    const int nt_ctx = dm.nt, ntau_ctx = dm.ntau, nao_ctx = dm.nao;
    Hid f(open_file(file, false), H5Fclose);
    const int nt = read_int(f, join(group, "nt")), ntau = read_int(f, join(group, "ntau")),
              nao = read_int(f, join(group, "nao"));
    if (nao != nao_ctx || ntau != ntau_ctx) throw H5Err("file nao/ntau differ from the context's");
    if (nt > nt_ctx) throw H5Err("file nt exceeds the context's nt");
    const long long es = (long long)nao * nao;
    std::vector<char> buf((size_t)CHUNK * es * 16);
    struct Comp { const char *name; int comp; long long nblk; };
    const Comp comps[] = {{"GR", NED_RET, tri(nt)}, {"GL", NED_LES, tri(nt)},
                          {"GTV", NED_TV, (long long)(nt + 1) * (ntau + 1)}, {"GM", NED_MAT, ntau + 1}};
    for (const Comp &c : comps) {
      CDataset d(f, join(group, c.name), false, 0);
      for (long long b = 0; b < c.nblk; b += CHUNK) {
        const long long nb = std::min(CHUNK, c.nblk - b);
        d.io(false, (hsize_t)(b * es), (hsize_t)(nb * es), buf.data());
        check_ned(ned_green_write(ctx, which, c.comp, b, nb, buf.data()));
      }
    }
  });
}

ned_status ned_h5_green_write(ned_ctx *ctx, int which, const char *file, const char *group, int nt) {
  if (!ctx || !file) { g_err = "null argument"; return NED_ERR_ARG; }
  return guard([&] {
    ned_dims dm;
    check_ned(ned_ctx_get_dims(ctx, &dm));
    if (nt < 0) nt = dm.nt;
    if (nt > dm.nt) throw H5Err("nt exceeds the context's nt");
    const int ntau = dm.ntau, nao = dm.nao, sig = dm.sig;
    Hid f(open_file(file, true), H5Fclose);
    const long long es = (long long)nao * nao;
    std::vector<char> buf((size_t)CHUNK * es * 16);
    struct Comp { const char *name; int comp; long long nblk; };
    const Comp comps[] = {{"GM", NED_MAT, ntau + 1}, {"GR", NED_RET, tri(nt)}, {"GL", NED_LES, tri(nt)},
                          {"GTV", NED_TV, (long long)(nt + 1) * (ntau + 1)}};
    for (const Comp &c : comps) {
      CDataset d(f, join(group, c.name), true, (hsize_t)(c.nblk * es));
      for (long long b = 0; b < c.nblk; b += CHUNK) {
        const long long nb = std::min(CHUNK, c.nblk - b);
        check_ned(ned_green_read(ctx, which, c.comp, b, nb, buf.data()));
        d.io(true, (hsize_t)(b * es), (hsize_t)(nb * es), buf.data());
      }
    }
    write_int(f, join(group, "nt"), nt);
    write_int(f, join(group, "ntau"), ntau);
    write_int(f, join(group, "sig"), sig);
    write_int(f, join(group, "nao"), nao);
    H5Fflush(f, H5F_SCOPE_LOCAL);
  });
}

}  // extern "C"
