// HDF5 output in the reference's layout, written through the headers and
// read back through the HDF5 C API.
//
// usage: hdf5_check write <dir> <initfile>
//          writes <dir>/pd.h5 (PD::writeToHDF5), <dir>/pd2.h5
//          (PDDouble::writeToHDF5), <dir>/chain.h5 (the chain layout of
//          affineChainSet::writeToHDF5Handle from a known 3-step,
//          4-walker, 2-parameter chain: theta = 100 s + 10 w + p,
//          logL = -(10 s + w)), <dir>/meta.h5 (initFileKnots and specFile
//          into groups) and tries an unwritable path
//        hdf5_check dump <file>
//          prints every group ("G <path>"), attribute
//          ("A <path>@<name> <kind><bytes> <n> <values>") and dataset
//          ("D <path> <kind><bytes> <rank> <dims> : <values>"), kind i/u/f/s
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/pofd_affine_amd_io.hpp"

#ifndef POFD_AFFINE_HAVE_HDF5
#error "hdf5_check needs the HDF5 header on the include path"
#endif

namespace {

std::string kind(hid_t ty) {
  const H5T_class_t c = H5Tget_class(ty);
  const size_t sz = H5Tget_size(ty);
  if (c == H5T_STRING) return "s" + std::to_string(H5Tis_variable_str(ty) > 0 ? 0 : sz);
  if (c == H5T_FLOAT) return "f" + std::to_string(sz);
  if (c == H5T_INTEGER) return std::string(H5Tget_sign(ty) == H5T_SGN_NONE ? "u" : "i") +
                               std::to_string(sz);
  return "?" + std::to_string(sz);
}

// reads n values of an attribute (attr) or dataset (!attr) and prints them
void print_values(hid_t id, bool attr, hid_t ty, size_t n) {
  const H5T_class_t c = H5Tget_class(ty);
  if (c == H5T_STRING) {
    hid_t mt = H5Tcopy(H5T_C_S1);
    H5Tset_size(mt, H5T_VARIABLE);
    std::vector<char*> buf(n, nullptr);
    const herr_t e = attr ? H5Aread(id, mt, buf.data())
                          : H5Dread(id, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data());
    for (size_t i = 0; i < n; ++i) {
      std::printf(" [%s]", (e >= 0 && buf[i]) ? buf[i] : "");
      if (buf[i]) H5free_memory(buf[i]);
    }
    H5Tclose(mt);
    return;
  }
  if (c == H5T_FLOAT) {
    std::vector<double> v(n);
    if (attr)
      H5Aread(id, H5T_NATIVE_DOUBLE, v.data());
    else
      H5Dread(id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data());
    for (double x : v) std::printf(" %.17g", x);
    return;
  }
  std::vector<long long> v(n);
  if (attr)
    H5Aread(id, H5T_NATIVE_LLONG, v.data());
  else
    H5Dread(id, H5T_NATIVE_LLONG, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data());
  for (long long x : v) std::printf(" %lld", x);
}

herr_t on_attr(hid_t loc, const char* name, const H5A_info_t*, void* path) {
  hid_t a = H5Aopen(loc, name, H5P_DEFAULT);
  hid_t ty = H5Aget_type(a), sp = H5Aget_space(a);
  const hssize_t n = H5Sget_simple_extent_npoints(sp);
  std::printf("A %s@%s %s %lld", static_cast<const std::string*>(path)->c_str(), name,
              kind(ty).c_str(), static_cast<long long>(n));
  print_values(a, true, ty, static_cast<size_t>(n));
  std::printf("\n");
  H5Sclose(sp);
  H5Tclose(ty);
  H5Aclose(a);
  return 0;
}

void dump_object(hid_t loc, const std::string& path);

herr_t on_link(hid_t grp, const char* name, const H5L_info_t*, void* path) {
  const std::string& base = *static_cast<const std::string*>(path);
  dump_object(grp, (base == "/" ? "/" : base + "/") + name);
  return 0;
}

void dump_object(hid_t file, const std::string& path) {
  hid_t o = H5Oopen(file, path.c_str(), H5P_DEFAULT);
  if (o < 0) return;
  std::string p = path;
  if (H5Iget_type(o) == H5I_GROUP) {
    std::printf("G %s\n", path.c_str());
    H5Aiterate2(o, H5_INDEX_NAME, H5_ITER_INC, nullptr, on_attr, &p);
    H5Literate(o, H5_INDEX_NAME, H5_ITER_INC, nullptr, on_link, &p);
  } else if (H5Iget_type(o) == H5I_DATASET) {
    hid_t ty = H5Dget_type(o), sp = H5Dget_space(o);
    const int rank = H5Sget_simple_extent_ndims(sp);
    std::vector<hsize_t> dims(rank > 0 ? rank : 1, 1);
    if (rank > 0) H5Sget_simple_extent_dims(sp, dims.data(), nullptr);
    const hssize_t n = H5Sget_simple_extent_npoints(sp);
    std::printf("D %s %s %d", path.c_str(), kind(ty).c_str(), rank);
    for (int i = 0; i < rank; ++i) std::printf(" %llu", static_cast<unsigned long long>(dims[i]));
    std::printf(" :");
    print_values(o, false, ty, static_cast<size_t>(n));
    std::printf("\n");
    H5Sclose(sp);
    H5Tclose(ty);
    H5Aiterate2(o, H5_INDEX_NAME, H5_ITER_INC, nullptr, on_attr, &p);
  }
  H5Oclose(o);
}

int write_all(const std::string& dir, const std::string& initfile) {
  PD pd;
  const double v[5] = {0.5, 1.5, -2.0, 3.25, 4.0};
  pd.fill(5, -0.25, 0.125, v, true);
  pd.writeToHDF5(dir + "/pd.h5");

  PDDouble pd2(3, -1.0, 0.5, 4, 2.0, 0.25);
  for (unsigned i = 0; i < 12; ++i) pd2.pd_[i] = 0.1 * i;
  pd2.writeToHDF5(dir + "/pd2.h5");

  const unsigned ns = 3, nw = 4, np = 2;
  std::vector<float> th(ns * nw * np);
  std::vector<double> ll(ns * nw);
  for (unsigned s = 0; s < ns; ++s)
    for (unsigned w = 0; w < nw; ++w) {
      for (unsigned p = 0; p < np; ++p) th[(s * nw + w) * np + p] = 100.0f * s + 10.0f * w + p;
      ll[s * nw + w] = -(10.0 * s + w);
    }
  pofd_hdf5::withNewFile(dir + "/chain.h5", "test", "chain", [&](hid_t f) {
    pofd_hdf5::writeChainSet(f, th.data(), ll.data(), ns, nw, np);
  });

  initFileKnots ifk(initfile, true, true);
  specFile sp;
  sp.datafiles = {"map_a.fits", "map_b.fits"};
  sp.psffiles = {"beam.fits", "beam.fits"};
  sp.sigmas = {0.002, 0.003};
  sp.like_norm = {1.0, 0.5};
  pofd_hdf5::withNewFile(dir + "/meta.h5", "test", "meta", [&](hid_t f) {
    pofd_hdf5::Id g(H5Gcreate2(f, "ParamInfo", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    ifk.writeToHDF5Handle(g.get());
    pofd_hdf5::Id g2(H5Gcreate2(f, "LikelihoodParams", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                     H5Gclose);
    sp.writeToHDF5Handle(g2.get());
    pofd_hdf5::writeAttStrings(g2.get(), "Names", {"alpha", "", "gamma"});
  });

  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);  // the failure below is expected
  try {
    pd.writeToHDF5(dir + "/no_such_dir/pd.h5");
    std::printf("unwritable none\n");
  } catch (const affineExcept& e) {
    std::printf("unwritable %s %s\n", e.getErrClass().c_str(), e.getErrMethod().c_str());
  }
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc >= 4 && std::string(argv[1]) == "write") return write_all(argv[2], argv[3]);
  if (argc == 3 && std::string(argv[1]) == "dump") {
    hid_t f = H5Fopen(argv[2], H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return 3;
    dump_object(f, "/");
    H5Fclose(f);
    return 0;
  }
  std::fprintf(stderr, "usage: hdf5_check write <dir> <initfile> | dump <file>\n");
  return 2;
}
