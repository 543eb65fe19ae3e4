/*
 * pofd_affine_amd_hdf5.hpp -- HDF5 output in the reference's file layout.
 *
 * Writers for attributes and datasets with the reference's names, types and
 * shapes (src/hdf5utils.cc), and the chain layout of
 * affineChainSet::writeToHDF5Handle (src/affineChainSet.cc:1306-1450): a
 * "Chain" float dataset [nwalkers][nsteps][nparams] with NWalkers, NSteps
 * and NParams attributes, and a "Likelihood" double dataset
 * [nwalkers][nsteps] with NWalkers and NSteps.  The classes that write
 * files with these (PD, PDDouble, affineEnsemble, pofdMCMC) have their
 * writeToHDF5 / writeToHDF5Handle members defined where the HDF5 header is
 * on the include path (for this image: -idirafter /opt/conda/include, so
 * that the system headers still come first, and -L/opt/conda/lib -lhdf5);
 * define POFD_AFFINE_NO_HDF5 to leave them out.
 *
 * Included by pofd_affine_amd.hpp after affineExcept; not meant to be
 * included on its own.  Host code only.  Boolean attributes and datasets
 * are written as H5T_NATIVE_HBOOL from hbool_t values.
 */
#ifndef POFD_AFFINE_AMD_HDF5_HPP
#define POFD_AFFINE_AMD_HDF5_HPP

#include <hdf5.h>

#include <memory>
#include <string>
#include <vector>

namespace pofd_hdf5 {

/* an HDF5 identifier closed by its close function when the holder ends */
class Id {
 public:
  Id(hid_t id, herr_t (*close)(hid_t)) : id_(id), close_(close) {}
  ~Id() {
    if (id_ >= 0) close_(id_);
  }
  Id(const Id&) = delete;
  Id& operator=(const Id&) = delete;
  hid_t get() const { return id_; }
  bool ok() const { return id_ >= 0; }

 private:
  hid_t id_;
  herr_t (*close_)(hid_t);
};

inline void need(bool ok, const char* cls, const char* meth, const std::string& what) {
  if (!ok) throw affineExcept(cls, meth, what);
}
inline void need_handle(hid_t obj, const char* cls, const char* meth, const std::string& name) {
  need(obj >= 0 && H5Iget_ref(obj) >= 0, cls, meth,
       "Input handle is not valid when writing " + name);
}

/* a one-dimensional attribute of n values of type ftype (file) / mtype
   (memory) */
inline void writeAttArray(hid_t obj, const std::string& name, hid_t ftype, hid_t mtype,
                          hsize_t n, const void* v, const char* meth) {
  need_handle(obj, "hdf5utils", meth, name);
  need(n > 0, "hdf5utils", meth, "Invalid number of elements in value when writing " + name);
  Id sp(H5Screate_simple(1, &n, nullptr), H5Sclose);
  need(sp.ok(), "hdf5utils", meth, "Failed to create dataspace for " + name);
  Id at(H5Acreate2(obj, name.c_str(), ftype, sp.get(), H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  need(at.ok(), "hdf5utils", meth, "Failed to create attribute " + name);
  need(H5Awrite(at.get(), mtype, v) >= 0, "hdf5utils", meth, "Failed to write attribute " + name);
}
inline void writeAttUnsignedInt(hid_t obj, const std::string& name, unsigned int v) {
  writeAttArray(obj, name, H5T_NATIVE_UINT, H5T_NATIVE_UINT, 1, &v, "writeAttUnsignedInt");
}
inline void writeAttFloat(hid_t obj, const std::string& name, float v) {
  writeAttArray(obj, name, H5T_NATIVE_FLOAT, H5T_NATIVE_FLOAT, 1, &v, "writeAttFloat");
}
inline void writeAttDouble(hid_t obj, const std::string& name, double v) {
  writeAttArray(obj, name, H5T_NATIVE_DOUBLE, H5T_NATIVE_DOUBLE, 1, &v, "writeAttDouble");
}
inline void writeAttBool(hid_t obj, const std::string& name, bool v) {
  const hbool_t b = static_cast<hbool_t>(v);
  writeAttArray(obj, name, H5T_NATIVE_HBOOL, H5T_NATIVE_HBOOL, 1, &b, "writeAttBool");
}
/* hbool_t copies of v (hbool_t may be bool, whose vector has no data()) */
inline std::unique_ptr<hbool_t[]> hbools(const std::vector<bool>& v) {
  std::unique_ptr<hbool_t[]> b(new hbool_t[v.size() ? v.size() : 1]);
  for (size_t i = 0; i < v.size(); ++i) b[i] = static_cast<hbool_t>(v[i]);
  return b;
}
inline void writeAttBools(hid_t obj, const std::string& name, const std::vector<bool>& v) {
  writeAttArray(obj, name, H5T_NATIVE_HBOOL, H5T_NATIVE_HBOOL, v.size(), hbools(v).get(),
                "writeAttBools");
}
/* variable-length C strings (hdf5utils::writeAttStrings) */
inline void writeAttStrings(hid_t obj, const std::string& name, const std::vector<std::string>& v) {
  Id ty(H5Tcopy(H5T_C_S1), H5Tclose);
  need(ty.ok() && H5Tset_size(ty.get(), H5T_VARIABLE) >= 0, "hdf5utils", "writeAttStrings",
       "Failed to make the string type for " + name);
  std::vector<const char*> p;
  for (const std::string& s : v) p.push_back(s.c_str());
  writeAttArray(obj, name, ty.get(), ty.get(), p.size(), p.data(), "writeAttStrings");
}

/* a dataset of rank dims.size() */
inline void writeData(hid_t obj, const std::string& name, hid_t ftype, hid_t mtype,
                      const std::vector<hsize_t>& dims, const void* v, const char* meth) {
  need_handle(obj, "hdf5utils", meth, name);
  for (hsize_t d : dims)
    need(d > 0, "hdf5utils", meth, "Invalid number of elements in value when writing " + name);
  Id sp(H5Screate_simple(static_cast<int>(dims.size()), dims.data(), nullptr), H5Sclose);
  need(sp.ok(), "hdf5utils", meth, "Failed to create dataspace for " + name);
  Id ds(H5Dcreate2(obj, name.c_str(), ftype, sp.get(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
        H5Dclose);
  need(ds.ok(), "hdf5utils", meth, "Failed to create dataset " + name);
  need(H5Dwrite(ds.get(), mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, v) >= 0, "hdf5utils", meth,
       "Failed to write dataset " + name);
}
inline void writeDataDoubles(hid_t obj, const std::string& name, const std::vector<double>& v) {
  writeData(obj, name, H5T_NATIVE_DOUBLE, H5T_NATIVE_DOUBLE, {v.size()}, v.data(),
            "writeDataDoubles");
}
inline void writeDataFloats(hid_t obj, const std::string& name, const std::vector<float>& v) {
  writeData(obj, name, H5T_NATIVE_FLOAT, H5T_NATIVE_FLOAT, {v.size()}, v.data(),
            "writeDataFloats");
}
inline void writeDataUnsignedInts(hid_t obj, const std::string& name,
                                  const std::vector<unsigned int>& v) {
  writeData(obj, name, H5T_NATIVE_UINT, H5T_NATIVE_UINT, {v.size()}, v.data(),
            "writeDataUnsignedInts");
}
inline void writeDataBools(hid_t obj, const std::string& name, const std::vector<bool>& v) {
  writeData(obj, name, H5T_NATIVE_HBOOL, H5T_NATIVE_HBOOL, {v.size()}, hbools(v).get(),
            "writeDataBools");
}
inline void writeDataStrings(hid_t obj, const std::string& name, const std::vector<std::string>& v) {
  Id ty(H5Tcopy(H5T_C_S1), H5Tclose);
  need(ty.ok() && H5Tset_size(ty.get(), H5T_VARIABLE) >= 0, "hdf5utils", "writeDataStrings",
       "Failed to make the string type for " + name);
  std::vector<const char*> p;
  for (const std::string& s : v) p.push_back(s.c_str());
  writeData(obj, name, ty.get(), ty.get(), {p.size()}, p.data(), "writeDataStrings");
}
/* one double as a one-element dataset (calcLikeSingle::writeToHDF5Handle's
   data and R ranges) */
inline void writeDataDouble(hid_t obj, const std::string& name, double v) {
  writeData(obj, name, H5T_NATIVE_DOUBLE, H5T_NATIVE_DOUBLE, {1}, &v, "writeDataDouble");
}

/* a group created in obj, closed when the holder ends */
inline hid_t createGroup(hid_t obj, const std::string& name, const char* cls, const char* meth) {
  const hid_t g = H5Gcreate2(obj, name.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  need(g >= 0, cls, meth, "Failed to create " + name + " HDF5 group");
  return g;
}

/* the chain as affineChainSet::writeToHDF5Handle lays it out, from the
   sampler's [nsteps][nwalkers][nparams] / [nsteps][nwalkers] arrays */
inline void writeChainSet(hid_t obj, const float* theta, const double* loglike,
                          unsigned int nsteps, unsigned int nwalkers, unsigned int nparams) {
  if (nsteps == 0) return;  // an empty chain writes nothing
  need_handle(obj, "affineChainSet", "writeToHDF5Handle", "Chain");
  std::vector<float> ch(static_cast<size_t>(nwalkers) * nsteps * nparams);
  std::vector<double> ll(static_cast<size_t>(nwalkers) * nsteps);
  for (unsigned w = 0; w < nwalkers; ++w)
    for (unsigned s = 0; s < nsteps; ++s) {
      const size_t src = static_cast<size_t>(s) * nwalkers + w;
      const size_t dst = static_cast<size_t>(w) * nsteps + s;
      for (unsigned p = 0; p < nparams; ++p) ch[dst * nparams + p] = theta[src * nparams + p];
      ll[dst] = loglike[src];
    }
  {
    const std::vector<hsize_t> dims = {nwalkers, nsteps, nparams};
    writeData(obj, "Chain", H5T_NATIVE_FLOAT, H5T_NATIVE_FLOAT, dims, ch.data(),
              "writeToHDF5Handle");
    Id ds(H5Dopen2(obj, "Chain", H5P_DEFAULT), H5Dclose);
    need(ds.ok(), "affineChainSet", "writeToHDF5Handle", "Failed to reopen Chain");
    writeAttUnsignedInt(ds.get(), "NWalkers", nwalkers);
    writeAttUnsignedInt(ds.get(), "NSteps", nsteps);
    writeAttUnsignedInt(ds.get(), "NParams", nparams);
  }
  {
    const std::vector<hsize_t> dims = {nwalkers, nsteps};
    writeData(obj, "Likelihood", H5T_NATIVE_DOUBLE, H5T_NATIVE_DOUBLE, dims, ll.data(),
              "writeToHDF5Handle");
    Id ds(H5Dopen2(obj, "Likelihood", H5P_DEFAULT), H5Dclose);
    need(ds.ok(), "affineChainSet", "writeToHDF5Handle", "Failed to reopen Likelihood");
    writeAttUnsignedInt(ds.get(), "NWalkers", nwalkers);
    writeAttUnsignedInt(ds.get(), "NSteps", nsteps);
  }
}

/* a new file (truncating), handed to fn, closed afterwards */
template <typename F>
inline void withNewFile(const std::string& path, const char* cls, const char* meth, F fn) {
  Id f(H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
  need(f.ok(), cls, meth, "Failed to open HDF5 file to write: " + path);
  fn(f.get());
}

}  // namespace pofd_hdf5

#endif
