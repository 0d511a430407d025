// formats.cpp — host-side readers of the reference's input files and the ELL adapter.
//
// File formats (utils.h:1851-1887, 2334-2484; main.cpp:210-289): whitespace-separated
// numbers, read in order; <job>.data1 = nodesNo stepsNo bcNodesNo bunnyNodesNo (read as
// doubles, main.cpp:83,210), <job>.data2 = 51 values, <job>.datac<k> = 7 values per species,
// <job>.datac<S+1> = 4 values for phi, <job>.nodes = n lines "x y z", <job>.bctypes = n ints,
// <job>.bcvalues<f> / <job>.icvalues<f> = n doubles per field.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/nnpp.h"

namespace {

bool read_doubles(const std::string& path, int64_t n, double* out) {
  std::ifstream f(path.c_str());
  if (!f) return false;
  for (int64_t i = 0; i < n; i++) {
    if (!(f >> out[i])) return false;
  }
  return true;
}

std::vector<std::string> column_names(int S) {
  static const char* ref[10] = {"Mg++", "H+", "OH-", "HCO3-", "CO32-",
                                "HPO42-", "PO43-", "Ca++", "Na+", "Cl-"};
  std::vector<std::string> c;
  for (int k = 0; k < S; k++) c.push_back(S == 10 ? ref[k] : "U" + std::to_string(k + 1));
  c.push_back("Phi");
  return c;
}

}  // namespace

extern "C" {

int nnpp_write_output_csv(const char* path, int32_t n, int32_t S, const double* x,
                          const double* y, const double* z, const double* U,
                          const int32_t* phase) {
  if (!path || n < 0 || S < 1 || S > NNPP_MAX_SPECIES || (n && (!x || !y || !z || !U || !phase)))
    return NNPP_ERR_ARG;
  FILE* f = std::fopen(path, "w");
  if (!f) return NNPP_ERR_IO;
  const std::vector<std::string> names = column_names(S);
  std::fputs("X,Y,Z", f);
  for (const std::string& c : names) std::fprintf(f, ",%s", c.c_str());
  std::fputs(",Phase\n", f);
  for (int32_t i = 0; i < n; i++) {  // precision 16, default float format (utils.h:2088-2096)
    std::fprintf(f, "%.16g,%.16g,%.16g", x[i], y[i], z[i]);
    for (int fld = 0; fld <= S; fld++) std::fprintf(f, ",%.16g", U[(int64_t)fld * n + i]);
    std::fprintf(f, ",%d\n", phase[i]);
  }
  return std::fclose(f) == 0 ? NNPP_OK : NNPP_ERR_IO;
}

int nnpp_read_start_csv(const char* path, int32_t n, int32_t S, double* U, int32_t* phase) {
  if (!path || n < 0 || S < 1 || S > NNPP_MAX_SPECIES || (n && (!U || !phase)))
    return NNPP_ERR_ARG;
  std::ifstream file(path);
  if (!file) return NNPP_ERR_IO;
  for (int64_t k = 0; k < (int64_t)(S + 1) * n; k++) U[k] = 0.0;
  for (int32_t i = 0; i < n; i++) phase[i] = 0;
  std::string line, col;
  if (!std::getline(file, line)) return NNPP_ERR_IO;
  std::map<std::string, size_t> pos;
  {
    std::istringstream h(line);
    size_t idx = 0;
    while (std::getline(h, col, ',')) pos[col] = idx++;
  }
  const std::vector<std::string> names = column_names(S);
  int32_t row = 0;
  std::vector<std::string> data;
  while (row < n && std::getline(file, line)) {
    std::istringstream ss(line);
    data.clear();
    while (std::getline(ss, col, ',')) data.push_back(col);
    for (int fld = 0; fld <= S; fld++) {
      auto it = pos.find(names[fld]);
      if (it != pos.end() && it->second < data.size())
        U[(int64_t)fld * n + row] = std::stod(data[it->second]);
    }
    auto it = pos.find("Phase");
    if (it != pos.end() && it->second < data.size()) phase[row] = std::stoi(data[it->second]);
    row++;
  }
  return row == n ? NNPP_OK : NNPP_ERR_IO;
}

int64_t nnpp_abi_params_size(void) { return (int64_t)sizeof(nnpp_params); }
int64_t nnpp_abi_stats_size(void) { return (int64_t)sizeof(nnpp_step_stats); }

int nnpp_params_defaults(nnpp_params* p, int32_t n_species, int32_t chemistry) {
  if (!p || n_species < 1 || n_species > NNPP_MAX_SPECIES) return NNPP_ERR_ARG;
  if (chemistry != NNPP_CHEM_GENERIC && chemistry != NNPP_CHEM_REFERENCE) return NNPP_ERR_ARG;
  if (chemistry == NNPP_CHEM_REFERENCE && n_species != 10) return NNPP_ERR_UNSUPPORTED;
  p->n_species = n_species;
  p->chemistry = chemistry;
  for (int k = 0; k < NNPP_MAX_SPECIES; k++) {
    p->solid_bonds[k] = (k == 0) ? 1 : 0;  // Mg only: utils.h:100-159 vs 211-213
    // main.cpp:805-837: U1..U9 always, U10 only in liquid
    p->clamp[k] = (k >= n_species) ? NNPP_CLAMP_NONE
                                   : (k == 9 ? NNPP_CLAMP_LIQUID : NNPP_CLAMP_ALWAYS);
  }
  p->clamp_floor = 1.0e-8 / p->CSolid;
  p->front_threshold = p->CSat / p->CSolid;  // main.cpp:916
  p->phase_eps = 10E-6;                      // EPS, main.cpp:35,414
  p->sub_steps = 8;                          // sub_step_no, main.cpp:39
  return NNPP_OK;
}

int nnpp_read_vector(const char* path, int64_t n, double* out) {
  if (!path || n < 0 || (n && !out)) return NNPP_ERR_ARG;
  return read_doubles(path, n, out) ? NNPP_OK : NNPP_ERR_IO;
}

int nnpp_read_ivector(const char* path, int64_t n, int32_t* out) {
  if (!path || n < 0 || (n && !out)) return NNPP_ERR_ARG;
  std::ifstream f(path);
  if (!f) return NNPP_ERR_IO;
  for (int64_t i = 0; i < n; i++) {
    double v;  // the reference reads into Eigen::VectorXi with >>; accept "0" or "0.0"
    if (!(f >> v)) return NNPP_ERR_IO;
    out[i] = (int32_t)v;
  }
  return NNPP_OK;
}

int nnpp_read_nodes(const char* path, int64_t n, double* x, double* y, double* z) {
  if (!path || n < 0 || (n && (!x || !y || !z))) return NNPP_ERR_ARG;
  std::ifstream f(path);
  if (!f) return NNPP_ERR_IO;
  for (int64_t i = 0; i < n; i++)
    if (!(f >> x[i] >> y[i] >> z[i])) return NNPP_ERR_IO;
  return NNPP_OK;
}

int nnpp_read_params(const char* job, int32_t n_species, int32_t chemistry, nnpp_params* p) {
  if (!job || !p || n_species < 1 || n_species > NNPP_MAX_SPECIES) return NNPP_ERR_ARG;
  const std::string J(job);
  double d1[4], d2[NNPP_DATA2_LEN], dc[NNPP_DATAC_LEN], dp[NNPP_DATAC_PHI_LEN];
  std::memset(p, 0, sizeof(*p));
  if (!read_doubles(J + ".data1", 4, d1)) return NNPP_ERR_IO;
  if (!read_doubles(J + ".data2", NNPP_DATA2_LEN, d2)) return NNPP_ERR_IO;
  p->n_nodes = (int32_t)d1[0];  // utils.h:2336-2339
  p->n_steps = (int32_t)d1[1];
  double* f = &p->t0;  // the 51 .data2 doubles are laid out contiguously in declaration order
  for (int i = 0; i < NNPP_DATA2_LEN; i++) f[i] = d2[i];
  for (int k = 0; k < n_species; k++) {
    if (!read_doubles(J + ".datac" + std::to_string(k + 1), NNPP_DATAC_LEN, dc))
      return NNPP_ERR_IO;
    p->z[k] = dc[0];  // utils.h:2400-2406
    p->D_solid[k] = dc[1];
    p->D_liquid[k] = dc[2];
    p->s1[k] = dc[3];
    p->s2[k] = dc[4];
    p->K1[k] = dc[5];
    p->K2[k] = dc[6];
  }
  if (!read_doubles(J + ".datac" + std::to_string(n_species + 1), NNPP_DATAC_PHI_LEN, dp))
    return NNPP_ERR_IO;
  p->sigma0 = dp[0];  // utils.h:2480-2483
  p->sigma1 = dp[1];
  p->s_phi = dp[2];
  p->K_phi = dp[3];
  int st = nnpp_params_defaults(p, n_species, chemistry);
  if (st) return st;
  std::ifstream ext((J + ".nnpp").c_str());
  if (!ext) return NNPP_OK;
  std::string line;
  while (std::getline(ext, line)) {
    const size_t hash = line.find('#');
    if (hash != std::string::npos) line.resize(hash);
    const size_t eq = line.find('=');
    if (eq == std::string::npos) continue;
    std::string key = line.substr(0, eq);
    key.erase(0, key.find_first_not_of(" \t"));
    key.erase(key.find_last_not_of(" \t") + 1);
    std::istringstream val(line.substr(eq + 1));
    if (key == "r_diss") val >> p->r_diss;
    else if (key == "front_threshold") val >> p->front_threshold;
    else if (key == "clamp_floor") val >> p->clamp_floor;
    else if (key == "phase_eps") val >> p->phase_eps;
    else if (key == "sub_steps") val >> p->sub_steps;
    else if (key == "solid_bonds" || key == "clamp") {
      int32_t* dst = key == "clamp" ? p->clamp : p->solid_bonds;
      for (int k = 0; k < n_species; k++)
        if (!(val >> dst[k])) return NNPP_ERR_IO;
    } else if (!key.empty()) {
      return NNPP_ERR_IO;  // unknown key: refuse rather than silently ignore
    }
    if (val.fail() && key != "solid_bonds" && key != "clamp") return NNPP_ERR_IO;
  }
  return NNPP_OK;
}

int nnpp_neighbors_to_ell(int32_t n, const int64_t* row_ptr, const int32_t* col,
                          int32_t max_len, int32_t* ell, int32_t* lengths) {
  if (n < 0 || max_len < 1 || (n && (!row_ptr || !ell || !lengths))) return NNPP_ERR_ARG;
  for (int32_t i = 0; i < n; i++) {
    const int64_t len = row_ptr[i + 1] - row_ptr[i];
    if (len + 1 > max_len) return NNPP_ERR_OVERFLOW;
  }
  // column-major n x max_len (Eigen::MatrixXi default, main.cpp:138; grid.h:145-149)
  for (int64_t k = 0; k < (int64_t)n * max_len; k++) ell[k] = -1;
  for (int32_t i = 0; i < n; i++) {
    const int64_t b0 = row_ptr[i], len = row_ptr[i + 1] - b0;
    ell[i] = i;  // clouds(i, 0) = i, grid.h:171
    for (int64_t t = 0; t < len; t++) ell[(int64_t)(t + 1) * n + i] = col[b0 + t];
    lengths[i] = (int32_t)(len + 1);  // grid.h:195
  }
  return NNPP_OK;
}

int nnpp_neighbors_from_ell(int32_t n, int32_t max_len, const int32_t* ell,
                            const int32_t* lengths, int64_t* row_ptr, int32_t* col) {
  if (n < 0 || max_len < 1 || (n && (!ell || !lengths || !row_ptr))) return NNPP_ERR_ARG;
  if (!row_ptr) return NNPP_ERR_ARG;
  row_ptr[0] = 0;
  for (int32_t i = 0; i < n; i++) {
    if (lengths[i] < 1 || lengths[i] > max_len) return NNPP_ERR_ARG;
    row_ptr[i + 1] = row_ptr[i] + (lengths[i] - 1);
  }
  if (!col) return NNPP_OK;
  std::vector<int32_t> row;
  for (int32_t i = 0; i < n; i++) {
    row.clear();
    for (int32_t t = 0; t < lengths[i]; t++) {  // column-major, self at column 0 (grid.h:171)
      const int32_t j = ell[(int64_t)t * n + i];
      if (j < 0 || j >= n) return NNPP_ERR_ARG;
      if (t == 0) {
        if (j != i) return NNPP_ERR_ARG;
        continue;
      }
      row.push_back(j);
    }
    std::sort(row.begin(), row.end());
    for (size_t t = 0; t < row.size(); t++) {
      if (row[t] == i || (t && row[t] == row[t - 1])) return NNPP_ERR_ARG;
      col[row_ptr[i] + (int64_t)t] = row[t];
    }
  }
  return NNPP_OK;
}

}  // extern "C"
