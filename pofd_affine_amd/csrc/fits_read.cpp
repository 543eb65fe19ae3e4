// Minimal FITS image reader (host ingestion for beams and maps).
//
// Covers what the reference reads through cfitsio (src/beam.cc:295-380,
// src/fitsData.cc:56-315): 2880-byte blocks of 80-column cards, IMAGE HDUs
// with BITPIX 8/16/32/64/-32/-64, BZERO/BSCALE, the PIXSIZE/PIXSCALE key,
// extension selection "file.fits[k]" and the image/IMAGE, mask/MASK
// extension names.  Pixels are returned in FITS storage order (NAXIS1
// fastest), the order the reference fills its arrays in.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "internal.h"

namespace {

struct Hdu {
  std::string extname;
  bool has_pixsize = false;
  double pixsize = 0.0;
  int bitpix = 0;
  std::vector<long> dims;
  double bscale = 1.0, bzero = 0.0;
  bool image = true;
  std::vector<double> data;
};

std::string trim(const std::string& s) {
  size_t a = s.find_first_not_of(' '), b = s.find_last_not_of(' ');
  return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

std::string card_value(const std::string& card) {
  std::string v = card.substr(10);
  std::string t = trim(v);
  if (!t.empty() && t[0] == '\'') {
    std::string out;
    for (size_t i = 1; i < t.size(); ++i) {
      if (t[i] == '\'') {
        if (i + 1 < t.size() && t[i + 1] == '\'') {
          out += '\'';
          ++i;
        } else {
          break;
        }
      } else {
        out += t[i];
      }
    }
    return trim(out);
  }
  size_t slash = t.find('/');
  if (slash != std::string::npos) t = t.substr(0, slash);
  return trim(t);
}

double to_double(const std::string& s) {
  std::string t = s;
  for (char& ch : t)
    if (ch == 'D') ch = 'E';
  return std::strtod(t.c_str(), nullptr);
}

int read_file(const std::string& path, std::vector<Hdu>& hdus,
              std::string& err) {
  std::ifstream in(path, std::ios::binary);
  if (!in) {
    err = "Error opening input file: " + path;
    return POFD_E_ARG;
  }
  std::vector<char> buf((std::istreambuf_iterator<char>(in)),
                        std::istreambuf_iterator<char>());
  const size_t B = 2880;
  size_t pos = 0;
  while (pos + B <= buf.size()) {
    Hdu h;
    bool done = false;
    bool first = true;
    while (!done) {
      if (pos + B > buf.size()) {
        err = "truncated FITS header in " + path;
        return POFD_E_ARG;
      }
      for (size_t i = 0; i < B && !done; i += 80) {
        std::string card(&buf[pos + i], 80);
        std::string key = trim(card.substr(0, 8));
        if (key == "END") {
          done = true;
          break;
        }
        if (first && key == "XTENSION") {
          h.image = (card_value(card) == "IMAGE");
        }
        if (card.compare(8, 2, "= ") != 0) continue;
        std::string val = card_value(card);
        if (key == "BITPIX") h.bitpix = std::atoi(val.c_str());
        else if (key == "NAXIS") h.dims.assign(std::atoi(val.c_str()), 0);
        else if (key.compare(0, 5, "NAXIS") == 0 && key.size() > 5) {
          int k = std::atoi(key.c_str() + 5);
          if (k >= 1 && static_cast<size_t>(k) <= h.dims.size())
            h.dims[k - 1] = std::atol(val.c_str());
        } else if (key == "BSCALE") h.bscale = to_double(val);
        else if (key == "BZERO") h.bzero = to_double(val);
        else if (key == "EXTNAME") h.extname = val;
        else if (key == "PIXSIZE") {
          h.has_pixsize = true;
          h.pixsize = to_double(val);
        } else if (key == "PIXSCALE" && !h.has_pixsize) {
          h.has_pixsize = true;
          h.pixsize = to_double(val);
        }
      }
      first = false;
      pos += B;
    }
    size_t nelem = h.dims.empty() ? 0 : 1;
    for (long d : h.dims) nelem *= static_cast<size_t>(d);
    const size_t bpe = static_cast<size_t>(std::abs(h.bitpix)) / 8;
    const size_t nbytes = nelem * bpe;
    if (h.image && nelem && pos + nbytes <= buf.size()) {
      h.data.resize(nelem);
      const unsigned char* p = reinterpret_cast<const unsigned char*>(&buf[pos]);
      for (size_t i = 0; i < nelem; ++i) {
        const unsigned char* q = p + i * bpe;
        uint64_t u = 0;
        for (size_t b = 0; b < bpe; ++b) u = (u << 8) | q[b];  // big endian
        double v;
        switch (h.bitpix) {
          case 8: v = static_cast<double>(static_cast<uint8_t>(u)); break;
          case 16: v = static_cast<double>(static_cast<int16_t>(u)); break;
          case 32: v = static_cast<double>(static_cast<int32_t>(u)); break;
          case 64: v = static_cast<double>(static_cast<int64_t>(u)); break;
          case -32: {
            uint32_t w = static_cast<uint32_t>(u);
            float f;
            std::memcpy(&f, &w, 4);
            v = f;
            break;
          }
          case -64: {
            double d;
            std::memcpy(&d, &u, 8);
            v = d;
            break;
          }
          default:
            err = "unsupported BITPIX in " + path;
            return POFD_E_ARG;
        }
        if (h.bscale != 1.0 || h.bzero != 0.0) v = v * h.bscale + h.bzero;
        h.data[i] = v;
      }
    }
    pos += (nbytes + B - 1) / B * B;
    hdus.push_back(std::move(h));
  }
  if (hdus.empty()) {
    err = "no HDU in " + path;
    return POFD_E_ARG;
  }
  return POFD_OK;
}

void split_ext(const std::string& path, std::string& base, int& ext) {
  ext = -1;
  base = path;
  if (!path.empty() && path.back() == ']') {
    size_t lb = path.rfind('[');
    if (lb != std::string::npos) {
      base = path.substr(0, lb);
      ext = std::atoi(path.c_str() + lb + 1);
    }
  }
}

bool name_is(const std::string& a, const char* b) {
  if (a.size() != std::strlen(b)) return false;
  for (size_t i = 0; i < a.size(); ++i)
    if (std::tolower(static_cast<unsigned char>(a[i])) != b[i]) return false;
  return true;
}

thread_local std::string g_fits_err;

}  // namespace

extern "C" const char* pofd_fits_last_error(void) { return g_fits_err.c_str(); }

extern "C" int pofd_fits_read_beam(const char* path, double* pix,
                                   uint64_t capacity, uint64_t* npix,
                                   double* pixsize) {
  if (!path || !npix) return POFD_E_ARG;
  std::string base;
  int ext;
  split_ext(path, base, ext);
  std::vector<Hdu> hdus;
  int rc = read_file(base, hdus, g_fits_err);
  if (rc) return rc;
  const Hdu& h = hdus[ext >= 0 ? static_cast<size_t>(ext) : 0];
  if (ext >= static_cast<int>(hdus.size())) {
    g_fits_err = "beam::readFile: no such extension";
    return POFD_E_ARG;
  }
  if (h.dims.size() != 2) {
    g_fits_err = "beam::readFile: Input BEAM is not 2D";
    return POFD_E_ARG;
  }
  if (!h.has_pixsize) {
    g_fits_err = std::string("beam::readFile: Unable to find pixel scale "
                             "information in ") + path;
    return POFD_E_ARG;
  }
  *npix = h.data.size();
  if (pixsize) *pixsize = h.pixsize;
  if (pix) {
    if (capacity < h.data.size()) return POFD_E_ARG;
    std::memcpy(pix, h.data.data(), h.data.size() * sizeof(double));
  }
  return POFD_OK;
}

namespace {
// image HDU ("file[k]", else the image/IMAGE extension, else the primary)
// and, unless ignored, the mask/MASK extension of a map file
int load_map(const char* path, bool ignore_mask, std::vector<double>& img_out,
             std::vector<double>& mask_out, const char* who) {
  std::string base;
  int ext;
  split_ext(path, base, ext);
  std::vector<Hdu> hdus;
  int rc = read_file(base, hdus, g_fits_err);
  if (rc) return rc;
  const Hdu* img = nullptr;
  if (ext >= 0) {
    if (ext >= static_cast<int>(hdus.size())) {
      g_fits_err = std::string(who) + ": no such extension";
      return POFD_E_ARG;
    }
    img = &hdus[ext];
  } else {
    for (const Hdu& h : hdus)
      if (name_is(h.extname, "image")) {
        img = &h;
        break;
      }
    if (!img) img = &hdus[0];
  }
  if (img->dims.size() != 2 || img->data.empty()) {
    g_fits_err = std::string(who) + ": Image is not 2D";
    return POFD_E_ARG;
  }
  mask_out.clear();
  if (!ignore_mask)
    for (const Hdu& h : hdus)
      if (name_is(h.extname, "mask")) {
        if (h.dims != img->dims) {
          g_fits_err = std::string(who) + ": Mask does not match image extent";
          return POFD_E_ARG;
        }
        mask_out = h.data;
        break;
      }
  img_out = img->data;
  return POFD_OK;
}
}  // namespace

// Pixel pairs of two maps as fitsDataDouble::readData (src/fitsDataDouble.cc:
// 305-406): masks are unsigned integers; with both present a pixel is kept
// when (mask1 & mask2) == 0, with one when it is 0.
extern "C" int pofd_fits_read_map_pair(const char* path1, const char* path2,
                                       int ignore_mask, double* d1, double* d2,
                                       uint64_t capacity, uint64_t* n) {
  if (!path1 || !path2 || !n) return POFD_E_ARG;
  std::vector<double> i1, m1, i2, m2;
  int rc = load_map(path1, ignore_mask != 0, i1, m1, "fitsDataDouble::readData");
  if (rc) return rc;
  rc = load_map(path2, ignore_mask != 0, i2, m2, "fitsDataDouble::readData");
  if (rc) return rc;
  if (i1.size() != i2.size()) {
    g_fits_err = "fitsDataDouble::readData: Data from file1 and file2 not the same size";
    return POFD_E_ARG;
  }
  auto mval = [](double v) { return static_cast<uint32_t>(static_cast<int64_t>(v)); };
  std::vector<size_t> keep;
  keep.reserve(i1.size());
  for (size_t i = 0; i < i1.size(); ++i) {
    bool k = true;
    if (!m1.empty() && !m2.empty())
      k = (mval(m1[i]) & mval(m2[i])) == 0;
    else if (!m1.empty())
      k = mval(m1[i]) == 0;
    else if (!m2.empty())
      k = mval(m2[i]) == 0;
    if (k) keep.push_back(i);
  }
  if (keep.empty()) {
    g_fits_err = "fitsDataDouble::readData: All data masked";
    return POFD_E_ARG;
  }
  *n = keep.size();
  if (d1 || d2) {
    if (!d1 || !d2 || capacity < keep.size()) return POFD_E_ARG;
    for (size_t k = 0; k < keep.size(); ++k) {
      d1[k] = i1[keep[k]];
      d2[k] = i2[keep[k]];
    }
  }
  return POFD_OK;
}

extern "C" int pofd_fits_read_map(const char* path, int ignore_mask,
                                  double* data, uint64_t capacity,
                                  uint64_t* n) {
  if (!path || !n) return POFD_E_ARG;
  std::string base;
  int ext;
  split_ext(path, base, ext);
  std::vector<Hdu> hdus;
  int rc = read_file(base, hdus, g_fits_err);
  if (rc) return rc;
  const Hdu* img = nullptr;
  if (ext >= 0) {
    if (ext >= static_cast<int>(hdus.size())) {
      g_fits_err = "fitsData::readData: no such extension";
      return POFD_E_ARG;
    }
    img = &hdus[ext];
  } else {
    for (const Hdu& h : hdus)
      if (name_is(h.extname, "image")) {
        img = &h;
        break;
      }
    if (!img) img = &hdus[0];
  }
  if (img->dims.size() != 2 || img->data.empty()) {
    g_fits_err = "fitsData::readData: Image is not 2D";
    return POFD_E_ARG;
  }
  std::vector<double> out;
  const Hdu* mask = nullptr;
  if (!ignore_mask)
    for (const Hdu& h : hdus)
      if (name_is(h.extname, "mask")) {
        mask = &h;
        break;
      }
  if (mask) {
    if (mask->dims != img->dims) {
      g_fits_err = "fitsData::readFile: Mask does not match image extent";
      return POFD_E_ARG;
    }
    for (size_t i = 0; i < img->data.size(); ++i)
      if (mask->data[i] == 0.0) out.push_back(img->data[i]);
    if (out.empty()) {
      g_fits_err = "fitsData::readData: All data masked";
      return POFD_E_ARG;
    }
  } else {
    out = img->data;
  }
  *n = out.size();
  if (data) {
    if (capacity < out.size()) return POFD_E_ARG;
    std::memcpy(data, out.data(), out.size() * sizeof(double));
  }
  return POFD_OK;
}
