// WAMIT numeric files .1, .2, .3 of a solve (wamit_files.rs:794-867 build, :534-550 force_file,
// :92-97 complex_cols, :1718-1757 fortran_e / fortran_i, :1822-1867 write_wamit; normalization
// docs/reference/file-formats.md section 1):
//   .1  period, I, J, A/(rho L^(3+ri+rj)), B/(w rho L^(3+ri+rj))            per frequency, I, J;
//       the radiation limits as rows of period -1 (omega = 0) and 0 (omega = infinity) with the
//       added mass only (wamit_files.rs:794-833; file-formats.md), in the order of the frequencies
//   .2  period, beta, I, |X|, phase, Re, Im of the Haskind force / (rho g L^(2+ri))
//   .2fk, .2sc  the same for its Froude-Krylov part and Haskind - Froude-Krylov (:836-853)
//   .3  the same columns for the pressure-integrated force; .3fk, .3sc its Froude-Krylov and
//       scattering parts; .4 the motions (RAO) xi L^ri (:855-866)
//   .6p field-point elevations (radiation block, then diffraction block per frequency), .fpt the
//       field points (:938-984)
// (each written when its arrays are given; the limits have no rows in them)
// Period 2 pi / w in E14.6 0P form, values E14.6 1P, integers I6, CRLF line ends, a header line
// " HydroBEM <version>   <file>   <stamp>"; the files are written to temporary names and renamed
// once all exist (all or none).
#include <cerrno>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <string>
#include <sys/stat.h>
#include <vector>

namespace hbg {

void set_error(int status, const std::string& msg, const char* where);

namespace {

const double PI_W = 3.14159265358979323846;

bool exists(const std::string& p) {
  struct stat st;
  return ::stat(p.c_str(), &st) == 0;
}

// E<width>.<digits> of x, 1P (d.dddddd E+xx) or 0P (0.dddddd E+xx), right-aligned
bool fortran_e(double x, bool one_p, int width, int digits, std::string& out) {
  if (!std::isfinite(x)) return false;
  if (x == 0.0) x = 0.0;
  std::string body;
  if (x == 0.0) {
    body = "0." + std::string(digits, '0') + "E+00";
  } else {
    const int p = one_p ? digits : digits - 1;
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.*e", p, std::fabs(x));
    std::string s(buf);
    const size_t epos = s.find('e');
    std::string mant = s.substr(0, epos);
    int e = std::atoi(s.c_str() + epos + 1);
    if (!one_p) {
      std::string m;
      for (char c : mant)
        if (c != '.') m += c;
      mant = "0." + m;
      e += 1;
    }
    if (std::abs(e) > 99) return false;
    char eb[8];
    std::snprintf(eb, sizeof eb, "E%c%02d", e < 0 ? '-' : '+', std::abs(e));
    body = mant + eb;
  }
  if (x < 0.0) body = "-" + body;
  if ((int)body.size() > width) return false;
  out += std::string(width - body.size(), ' ') + body;
  return true;
}

bool fortran_i(long n, int width, std::string& out) {
  std::string s = std::to_string(n);
  if ((int)s.size() > width) return false;
  out += std::string(width - s.size(), ' ') + s;
  return true;
}

double degrees(std::complex<double> z) {
  return (z == std::complex<double>(0.0, 0.0)) ? 0.0 : std::atan2(z.imag(), z.real()) * (180.0 / PI_W);
}

bool complex_cols(std::complex<double> z, std::string& out) {
  return fortran_e(std::hypot(z.real(), z.imag()), true, 14, 6, out) && fortran_e(degrees(z), true, 14, 6, out) &&
         fortran_e(z.real(), true, 14, 6, out) && fortran_e(z.imag(), true, 14, 6, out);
}

}  // namespace

// A, B [F][D][D]; HK [F][D][H] complex (Output.HK order); X [F][H][D] complex. Field points
// (P > 0): pts [P][3] global, frad [F][P][D] and fdif [F][P][H] complex, the elevations (p / (rho g)
// below z = 0) per unit displacement of the radiation columns and of the total diffraction wave.
int write_wamit(const char* stem, int overwrite, double rho, double g, double L, const double* omegas, int F,
                const double* heads, int H, int nb, const double* A, const double* B, const double* HK,
                const double* X, const double* FK, const double* SC, const double* RAO, const char* stamp,
                std::string& err, int P = 0, const double* pts = nullptr, const double* frad = nullptr,
                const double* fdif = nullptr) {
  const int D = 6 * nb, K = D;
  if (!(rho > 0.0) || !(g > 0.0) || !(L > 0.0)) {
    err = "density, gravity and reference_length must be > 0";
    return 1;
  }
  std::string st = stamp && *stamp ? stamp : "";
  if (st.empty()) {
    char buf[64];
    std::time_t t = std::time(nullptr);
    std::strftime(buf, sizeof buf, "%Y-%m-%d %H:%M:%S", std::localtime(&t));
    st = buf;
  }
  if (st.find('\r') != std::string::npos || st.find('\n') != std::string::npos) {
    err = "the time stamp must be a single line";
    return 1;
  }
  auto r = [](int i) { return (i % 6) >= 3 ? 1 : 0; };
  auto limit = [&](int f) { return omegas[f] == 0.0 || std::isinf(omegas[f]); };
  std::vector<std::string> per(F), bcol(H);
  for (int f = 0; f < F; ++f) {
    if (!(omegas[f] >= 0.0)) {
      err = "frequencies must be > 0, or 0 or +infinity (the limits)";
      return 1;
    }
    const double p = limit(f) ? (omegas[f] == 0.0 ? -1.0 : 0.0) : 2.0 * PI_W / omegas[f];
    if (!fortran_e(p, false, 14, 6, per[f])) {
      err = "frequency does not fit E14.6";
      return 1;
    }
  }
  for (int h = 0; h < H; ++h)
    if (!fortran_e(heads[h], false, 14, 6, bcol[h])) {
      err = "heading does not fit E14.6";
      return 1;
    }
  for (int a = 0; a < F; ++a)
    for (int b = a + 1; b < F; ++b)
      if (per[a] == per[b]) {
        err = "distinct frequencies collide at the printed precision of the numeric files";
        return 1;
      }
  std::vector<std::pair<std::string, std::vector<std::string>>> files;
  {
    std::vector<std::string> lines;
    for (int f = 0; f < F; ++f)
      for (int i = 0; i < D; ++i)
        for (int k = 0; k < K; ++k) {
          const double s = rho * std::pow(L, 3 + r(i) + r(k));
          const size_t o = ((size_t)f * D + i) * K + k;
          // a limit's entry that is not defined (NaN: at omega = 0 in finite depth the modes with a
          // net volume flux) has no row, as a mode the reference refuses there
          if (limit(f) && std::isnan(A[o])) continue;
          std::string l = per[f];
          if (!fortran_i(i + 1, 6, l) || !fortran_i(k + 1, 6, l) || !fortran_e(A[o] / s, true, 14, 6, l) ||
              (!limit(f) && !fortran_e(B[o] / (omegas[f] * s), true, 14, 6, l))) {
            err = "value does not fit E14.6";
            return 1;
          }
          lines.push_back(l);
        }
    files.push_back({".1", lines});
  }
  // a force-style file: for each ordinary frequency and heading, row i of value(f, h, i) / scale(i)
  auto force_file = [&](const char* ext, auto value, auto scale) {
    std::vector<std::string> lines;
    for (int f = 0; f < F; ++f)
      for (int h = 0; h < H && !limit(f); ++h)
        for (int i = 0; i < D; ++i) {
          std::string l = per[f] + bcol[h];
          if (!fortran_i(i + 1, 6, l) || !complex_cols(value(f, h, i) / scale(i), l)) {
            err = "value does not fit E14.6";
            return false;
          }
          lines.push_back(l);
        }
    files.push_back({ext, lines});
    return true;
  };
  typedef std::complex<double> C;
  auto at = [](const double* p, size_t o) { return C(p[2 * o], p[2 * o + 1]); };
  auto hk = [&](int f, int h, int n) { return at(HK, ((size_t)f * K + n) * H + h); };  // [F][K][H]
  auto fhd = [&](const double* p) { return [=](int f, int h, int i) { return at(p, ((size_t)f * H + h) * D + i); }; };
  auto force_scale = [&](int i) { return rho * g * std::pow(L, 2 + r(i)); };
  auto rao_scale = [&](int i) { return std::pow(L, -r(i)); };
  if (H > 0) {
    // the Haskind force by solved column (every mode of every body: column n is mode n)
    if (!force_file(".2", hk, force_scale) || !force_file(".3", fhd(X), force_scale)) return 1;
    if (FK) {
      auto fk = fhd(FK);
      if (!force_file(".2fk", fk, force_scale) ||
          !force_file(".2sc", [&](int f, int h, int n) { return hk(f, h, n) - fk(f, h, n); }, force_scale) ||
          !force_file(".3fk", fk, force_scale))
        return 1;
    }
    if (SC && !force_file(".3sc", fhd(SC), force_scale)) return 1;
    if (RAO && !force_file(".4", fhd(RAO), rao_scale)) return 1;
  }
  if (P > 0 && pts) {
    // .6p (wamit_files.rs:938-974, file-formats.md section 2): per ordinary frequency the
    // radiation block (the points in order, (Re, Im) of eta / L^r_k per column on each row), then
    // the diffraction block (per heading the points, Re, Im of eta / A); .fpt the points once
    auto e1p = [&](double x, std::string& l) { return fortran_e(x, true, 14, 6, l); };
    if (frad || fdif) {
      std::vector<std::string> lines;
      for (int f = 0; f < F; ++f) {
        if (limit(f)) continue;
        for (int q = 0; frad && K > 0 && q < P; ++q) {
          std::string l = per[f];
          bool ok = fortran_i(q + 1, 10, l);
          for (int k = 0; k < K && ok; ++k) {
            const C z = at(frad, ((size_t)f * P + q) * K + k) * std::pow(L, -r(k));
            ok = e1p(z.real(), l) && e1p(z.imag(), l);
          }
          if (!ok) {
            err = "field value does not fit E14.6";
            return 1;
          }
          lines.push_back(l);
        }
        for (int h = 0; fdif && h < H; ++h)
          for (int q = 0; q < P; ++q) {
            std::string l = per[f] + bcol[h];
            const C z = at(fdif, ((size_t)f * P + q) * H + h);
            if (!fortran_i(q + 1, 10, l) || !e1p(z.real(), l) || !e1p(z.imag(), l)) {
              err = "field value does not fit E14.6";
              return 1;
            }
            lines.push_back(l);
          }
      }
      files.push_back({".6p", lines});
    }
    std::vector<std::string> lines;
    for (int q = 0; q < P; ++q) {
      std::string l;
      if (!fortran_i(q + 1, 10, l) || !e1p(pts[3 * q], l) || !e1p(pts[3 * q + 1], l) || !e1p(pts[3 * q + 2], l)) {
        err = "field point does not fit E14.6";
        return 1;
      }
      lines.push_back(l);
    }
    files.push_back({".fpt", lines});
  }
  const std::string sstem(stem);
  if (!overwrite)
    for (auto& f : files)
      if (exists(sstem + f.first)) {
        err = sstem + f.first + " exists";
        return 10;
      }
  std::string base = sstem;
  const size_t slash = base.find_last_of("/\\");
  if (slash != std::string::npos) base = base.substr(slash + 1);
  std::vector<std::string> temps;
  auto cleanup = [&]() {
    for (auto& t : temps) std::remove(t.c_str());
  };
  for (size_t q = 0; q < files.size(); ++q) {
    const std::string tmp = sstem + ".hbg-" + std::to_string((long)std::time(nullptr)) + "-" + std::to_string(q) + ".tmp";
    FILE* fp = std::fopen(tmp.c_str(), "wb");
    if (!fp) {
      cleanup();
      err = "cannot write " + tmp + ": " + std::strerror(errno);
      return 9;
    }
    temps.push_back(tmp);
    std::string text = " HydroBEM 0.1.0-amd   " + base + files[q].first + "   " + st + "\r\n";
    for (auto& l : files[q].second) text += l + "\r\n";
    const bool ok = std::fwrite(text.data(), 1, text.size(), fp) == text.size();
    if (std::fclose(fp) != 0 || !ok) {
      cleanup();
      err = "cannot write " + tmp;
      return 9;
    }
  }
  for (size_t q = 0; q < files.size(); ++q)
    if (std::rename(temps[q].c_str(), (sstem + files[q].first).c_str()) != 0) {
      cleanup();
      err = "cannot rename to " + sstem + files[q].first;
      return 9;
    }
  return 0;
}

}  // namespace hbg

// ---- the number fields of the files at the C ABI (include/hydrobem_amd.h hbg_fortran_e /
// hbg_fortran_i), for the reference's formatter table (crates/hydrobem/tests/io.rs:37-65)
#include "../../include/hydrobem_amd.h"

namespace {

HbgStatus field_out(bool ok, const std::string& s, char* out, int32_t capacity, const char* what) {
  if (!ok) {
    hbg::set_error(HBG_ERR_VALUE, std::string(what) + ": the number does not fit the field", what);
    return HBG_ERR_VALUE;
  }
  if ((int64_t)s.size() + 1 > capacity) {
    hbg::set_error(HBG_ERR_ARGUMENT, std::string(what) + ": output capacity too small", what);
    return HBG_ERR_ARGUMENT;
  }
  std::memcpy(out, s.c_str(), s.size() + 1);
  return HBG_OK;
}

}  // namespace

extern "C" HbgStatus hbg_fortran_e(double x, int32_t one_p, int32_t width, int32_t digits, char* out,
                                   int32_t capacity) {
  if (!out || width < 1 || width > 64 || digits < 1 || digits > 30) {
    hbg::set_error(HBG_ERR_ARGUMENT, "null output or width / digits out of range", "hbg_fortran_e");
    return HBG_ERR_ARGUMENT;
  }
  std::string s;
  const bool ok = hbg::fortran_e(x, one_p != 0, width, digits, s);
  return field_out(ok, s, out, capacity, "hbg_fortran_e");
}

extern "C" HbgStatus hbg_fortran_i(int64_t n, int32_t width, char* out, int32_t capacity) {
  if (!out || width < 1 || width > 64) {
    hbg::set_error(HBG_ERR_ARGUMENT, "null output or width out of range", "hbg_fortran_i");
    return HBG_ERR_ARGUMENT;
  }
  std::string s;
  const bool ok = hbg::fortran_i((long)n, width, s);
  return field_out(ok, s, out, capacity, "hbg_fortran_i");
}
