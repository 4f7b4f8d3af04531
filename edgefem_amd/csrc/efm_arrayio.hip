// Array-package writers (host code): export_array_package_json, export_patterns_csv,
// export_coupling_csv and export_scan_csv (src/io/array_export.cpp:33-191), with the
// coupling list of extract_all_couplings (src/array/coupling_extract.cpp:21-57) and the
// VSWR of active_vswr (src/array/active_impedance.cpp:117-123) that they call.  The layouts
// are the reference's file formats; numbers are written as iostream writes them at
// precision 12 (printf's %.12g), here with std::to_chars(general, 12).  |s| and arg s are
// the C library's hypot and atan2 (what std::abs / std::arg of a complex<double> call under
// g++), degrees are radians * 180 / pi in that order, and a position distance is the square
// root of ((dx^2 + dy^2) + dz^2).
#include "efm_internal.h"

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <string>

namespace {

struct Out {
  std::string s;
  void num(double v) {
    char b[40];
    const auto r = std::to_chars(b, b + sizeof(b), v, std::chars_format::general, 12);
    s.append(b, r.ptr);
  }
  void integer(int64_t v) {
    char b[24];
    const auto r = std::to_chars(b, b + sizeof(b), v);
    s.append(b, r.ptr);
  }
  void lit(const char *t) { s.append(t); }
  void ch(char c) { s.push_back(c); }
};

double deg(double rad) { return rad * 180.0 / M_PI; }
double cabs_(double re, double im) { return std::hypot(re, im); }
double to_db(double lin) { return 20.0 * std::log10(std::max(lin, 1e-20)); }

int write_file(const char *path, const std::string &text) {
  FILE *f = std::fopen(path, "wb");
  if (!f) {
    efm::set_error(std::string("Cannot open file for writing: ") + path);
    return EFM_ERR_ARG;
  }
  const size_t w = std::fwrite(text.data(), 1, text.size(), f);
  const int c = std::fclose(f);
  if (w != text.size() || c != 0) {
    efm::set_error(std::string("Write failed: ") + path);
    return EFM_ERR_ARG;
  }
  return EFM_OK;
}

void vec3(Out &o, const double *p) {
  o.ch('[');
  o.num(p[0]);
  o.lit(", ");
  o.num(p[1]);
  o.lit(", ");
  o.num(p[2]);
  o.ch(']');
}

}  // namespace

extern "C" int efm_array_package_json(const char *path, const char *model_name, double reference_impedance,
                                      int64_t n_elements, const double *h_positions, int64_t n_freq,
                                      const double *h_freqs, int64_t n_smat, int nports, const double *h_S,
                                      int64_t n_pattern_freqs, int64_t n_scan, const double *h_scan_theta,
                                      const double *h_scan_phi, const double *h_scan_loss_db) {
  EFM_REQUIRE(path && model_name, "efm_array_package_json: NULL path or model name");
  EFM_REQUIRE(n_elements >= 0 && n_freq >= 0 && n_smat >= 0 && nports >= 0 && n_pattern_freqs >= 0 && n_scan >= 0,
              "efm_array_package_json: negative size");
  EFM_REQUIRE(n_smat <= n_freq, "efm_array_package_json: more S matrices than frequencies");
  EFM_REQUIRE(n_elements == 0 || h_positions, "efm_array_package_json: positions missing");
  EFM_REQUIRE(n_freq == 0 || h_freqs, "efm_array_package_json: frequencies missing");
  EFM_REQUIRE(n_smat == 0 || nports == 0 || h_S, "efm_array_package_json: S matrices missing");
  EFM_REQUIRE(n_scan == 0 || (h_scan_theta && h_scan_phi && h_scan_loss_db), "efm_array_package_json: scan data missing");
  Out o;
  o.lit("{\n");
  o.lit("  \"format\": \"EdgeFEM_ArrayPackage\",\n");
  o.lit("  \"version\": \"1.0\",\n");
  o.lit("  \"model_name\": \"");
  o.lit(model_name);
  o.lit("\",\n");
  o.lit("  \"reference_impedance\": ");
  o.num(reference_impedance);
  o.lit(",\n");
  o.lit("  \"num_elements\": ");
  o.integer(n_elements);
  o.lit(",\n");
  o.lit("  \"element_positions\": [\n");
  for (int64_t i = 0; i < n_elements; ++i) {
    o.lit("    ");
    vec3(o, h_positions + 3 * i);
    if (i < n_elements - 1)
      o.ch(',');
    o.ch('\n');
  }
  o.lit("  ],\n");
  o.lit("  \"num_frequencies\": ");
  o.integer(n_freq);
  o.lit(",\n");
  o.lit("  \"frequencies\": [");
  for (int64_t i = 0; i < n_freq; ++i) {
    o.num(h_freqs[i]);
    if (i < n_freq - 1)
      o.lit(", ");
  }
  o.lit("],\n");
  o.lit("  \"S_matrices\": [\n");
  for (int64_t f = 0; f < n_smat; ++f) {
    const double *S = h_S + (size_t)f * nports * nports * 2;
    o.lit("    {\n");
    o.lit("      \"frequency\": ");
    o.num(h_freqs[f]);
    o.lit(",\n");
    o.lit("      \"data\": [\n");
    for (int i = 0; i < nports; ++i) {
      o.lit("        [");
      for (int j = 0; j < nports; ++j) {
        const double *v = S + 2 * ((size_t)i * nports + j);
        o.ch('[');
        o.num(v[0]);
        o.lit(", ");
        o.num(v[1]);
        o.ch(']');
        if (j < nports - 1)
          o.lit(", ");
      }
      o.ch(']');
      if (i < nports - 1)
        o.ch(',');
      o.ch('\n');
    }
    o.lit("      ]\n");
    o.lit("    }");
    if (f < n_smat - 1)
      o.ch(',');
    o.ch('\n');
  }
  o.lit("  ],\n");
  o.lit("  \"has_embedded_patterns\": ");
  o.lit(n_pattern_freqs == 0 ? "false" : "true");
  o.lit(",\n");
  o.lit("  \"num_pattern_frequencies\": ");
  o.integer(n_pattern_freqs);
  o.lit(",\n");
  o.lit("  \"has_scan_data\": ");
  o.lit(n_scan == 0 ? "false" : "true");
  o.lit(",\n");
  o.lit("  \"num_scan_points\": ");
  o.integer(n_scan);
  if (n_scan > 0) {
    o.lit(",\n");
    o.lit("  \"scan_angles\": [\n");
    for (int64_t i = 0; i < n_scan; ++i) {
      o.lit("    {\"theta\": ");
      o.num(deg(h_scan_theta[i]));
      o.lit(", \"phi\": ");
      o.num(deg(h_scan_phi[i]));
      o.lit(", \"scan_loss_db\": ");
      o.num(h_scan_loss_db[i]);
      o.ch('}');
      if (i < n_scan - 1)
        o.ch(',');
      o.ch('\n');
    }
    o.lit("  ]\n");
  } else {
    o.ch('\n');
  }
  o.lit("}\n");
  return write_file(path, o.s);
}

extern "C" int efm_patterns_csv(const char *path, int64_t n_patterns, const int32_t *h_port_index,
                                const int64_t *h_npts, const double *h_rows) {
  EFM_REQUIRE(path, "efm_patterns_csv: NULL path");
  EFM_REQUIRE(n_patterns >= 0, "efm_patterns_csv: negative size");
  EFM_REQUIRE(n_patterns == 0 || (h_port_index && h_npts), "efm_patterns_csv: NULL argument");
  int64_t total = 0;
  for (int64_t p = 0; p < n_patterns; ++p) {
    EFM_REQUIRE(h_npts[p] >= 0, "efm_patterns_csv: negative point count");
    total += h_npts[p];
  }
  EFM_REQUIRE(total == 0 || h_rows, "efm_patterns_csv: rows missing");
  Out o;
  o.lit("port_index,theta_deg,phi0_mag,phi0_phase_deg,phi90_mag,phi90_phase_deg\n");
  const double *r = h_rows;
  for (int64_t p = 0; p < n_patterns; ++p) {
    for (int64_t k = 0; k < h_npts[p]; ++k, r += 5) {
      o.integer(h_port_index[p]);
      o.ch(',');
      o.num(deg(r[0]));
      o.ch(',');
      o.num(r[1]);
      o.ch(',');
      o.num(deg(r[2]));
      o.ch(',');
      o.num(r[3]);
      o.ch(',');
      o.num(deg(r[4]));
      o.ch('\n');
    }
  }
  return write_file(path, o.s);
}

extern "C" int efm_coupling_csv(const char *path, int nports, const double *h_S, const double *h_positions) {
  EFM_REQUIRE(path, "efm_coupling_csv: NULL path");
  EFM_REQUIRE(nports >= 0, "efm_coupling_csv: negative port count");
  EFM_REQUIRE(nports == 0 || h_S, "efm_coupling_csv: S missing");
  Out o;
  o.lit("port_i,port_j,S_real,S_imag,magnitude_db,phase_deg,distance_m\n");
  for (int i = 0; i < nports; ++i)
    for (int j = 0; j < nports; ++j) {
      if (i == j)
        continue;
      const double re = h_S[2 * ((size_t)i * nports + j)], im = h_S[2 * ((size_t)i * nports + j) + 1];
      double dist = 0.0;
      if (h_positions) {
        const double dx = h_positions[3 * i] - h_positions[3 * j];
        const double dy = h_positions[3 * i + 1] - h_positions[3 * j + 1];
        const double dz = h_positions[3 * i + 2] - h_positions[3 * j + 2];
        dist = std::sqrt((dx * dx + dy * dy) + dz * dz);
      }
      o.integer(i);
      o.ch(',');
      o.integer(j);
      o.ch(',');
      o.num(re);
      o.ch(',');
      o.num(im);
      o.ch(',');
      o.num(to_db(cabs_(re, im)));
      o.ch(',');
      o.num(deg(std::atan2(im, re)));
      o.ch(',');
      o.num(dist);
      o.ch('\n');
    }
  return write_file(path, o.s);
}

extern "C" int efm_scan_csv(const char *path, int64_t n_scan, const double *h_theta, const double *h_phi,
                            const int64_t *h_nelem, const double *h_gamma, const double *h_Z) {
  EFM_REQUIRE(path, "efm_scan_csv: NULL path");
  EFM_REQUIRE(n_scan >= 0, "efm_scan_csv: negative size");
  EFM_REQUIRE(n_scan == 0 || (h_theta && h_phi && h_nelem), "efm_scan_csv: NULL argument");
  int64_t total = 0;
  for (int64_t s = 0; s < n_scan; ++s) {
    EFM_REQUIRE(h_nelem[s] >= 0, "efm_scan_csv: negative element count");
    total += h_nelem[s];
  }
  EFM_REQUIRE(total == 0 || (h_gamma && h_Z), "efm_scan_csv: Gamma or Z missing");
  Out o;
  o.lit("theta_deg,phi_deg,element_idx,Gamma_real,Gamma_imag,Z_real,Z_imag,VSWR\n");
  const double *g = h_gamma, *z = h_Z;
  for (int64_t s = 0; s < n_scan; ++s) {
    for (int64_t i = 0; i < h_nelem[s]; ++i, g += 2, z += 2) {
      const double rho = cabs_(g[0], g[1]);
      const double vswr = rho >= 1.0 - 1e-12 ? 1e12 : (1.0 + rho) / (1.0 - rho);
      o.num(deg(h_theta[s]));
      o.ch(',');
      o.num(deg(h_phi[s]));
      o.ch(',');
      o.integer(i);
      o.ch(',');
      o.num(g[0]);
      o.ch(',');
      o.num(g[1]);
      o.ch(',');
      o.num(z[0]);
      o.ch(',');
      o.num(z[1]);
      o.ch(',');
      o.num(vswr);
      o.ch('\n');
    }
  }
  return write_file(path, o.s);
}
