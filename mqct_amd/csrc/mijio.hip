// mijio.hip -- the parallel-IO matrix directory (host code).
//
// A split build of the reference writes its matrix not as one MTRX file but
// as a directory (MATRIX_TRUNCATED, iotest.f:427): per writer rank one pair
// of files MIJ_<k_st>_<k_fn>.DAT (the elements of its range whose |M| at the
// two cutoff R points exceeds MIJ_CUTOFF) and Ind_<k_st>_<k_fn>.DAT (their
// ind_mat pairs), the counts in MATRIX_NONZERO_INDEX.DAT, the ranges in
// MATRIX_FILE_INDEX.DAT, and the state tables in MATRIX_INFO.DAT
// (bk_print_matrix_info, bk_print_matrix, matrix_ini.f:1545-1575,
// 6836-7244; bk_matrix_combination, 9053-9145).  The reader is
// bk_read_matrix (matrix_ini.f:7246-7525) for one reading rank: every
// element in file order, then (in the caller) MIJ_SHIFT and the spline.
// Formatted (UNFORMAT=NO) and Fortran sequential unformatted (UNFORMAT=YES)
// forms; the two index files are formatted in both.
#include "mqct_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

constexpr double kEvToWn = 8065.5448, kAuToEv = 27.21113845;  // iotest.f:20-24

// Fortran Ew.d with the default scale factor: [-]0.d1..dd E+xx (a three-
// digit exponent drops the E), right-justified in w columns.
std::string fort_e(double x, int w, int d) {
  char m[64];
  std::string body;
  if (x == 0.0) {
    body = "0." + std::string(d, '0') + "E+00";
  } else {
    std::snprintf(m, sizeof m, "%.*e", d - 1, std::fabs(x));
    // m = "D.DDDDe+XX"
    std::string s(m);
    size_t e = s.find('e');
    std::string digits = s.substr(0, 1) + s.substr(2, e - 2);
    int ex = std::atoi(s.c_str() + e + 1) + 1;
    char t[16];
    if (std::abs(ex) <= 99) std::snprintf(t, sizeof t, "E%c%02d", ex < 0 ? '-' : '+', std::abs(ex));
    else std::snprintf(t, sizeof t, "%c%03d", ex < 0 ? '-' : '+', std::abs(ex));
    body = "0." + digits + t;
  }
  if (x < 0.0) body = "-" + body;
  if ((int)body.size() < w) body = std::string(w - body.size(), ' ') + body;
  return body;
}

// A<w> of a shorter string: right-justified.
std::string fa(const char *s, int w) {
  std::string t(s);
  if ((int)t.size() < w) t = std::string(w - t.size(), ' ') + t;
  return t;
}

struct Rec {  // Fortran sequential unformatted record
  std::vector<unsigned char> b;
  void i32(int v) { put(&v, 4); }
  void put(const void *p, size_t n) {
    const unsigned char *c = (const unsigned char *)p;
    b.insert(b.end(), c, c + n);
  }
  bool write(FILE *f) {
    uint32_t n = (uint32_t)b.size();
    bool ok = fwrite(&n, 4, 1, f) == 1;
    if (n) ok = ok && fwrite(b.data(), 1, n, f) == n;
    ok = ok && fwrite(&n, 4, 1, f) == 1;
    b.clear();
    return ok;
  }
};

bool read_rec(FILE *f, std::vector<unsigned char> &b) {
  uint32_t n, n2;
  if (fread(&n, 4, 1, f) != 1) return false;
  b.resize(n);
  if (n && fread(b.data(), 1, n, f) != n) return false;
  return fread(&n2, 4, 1, f) == 1 && n2 == n;
}

// Element range of writer rank i (1-based) of n over T elements
// (bk_matrix_combination's split, matrix_ini.f:9073-9087).
void rank_range(int64_t T, int n, int i, int64_t &k_st, int64_t &k_fn) {
  const int64_t chunk = T / n, rem = T - chunk * n;
  k_st = (int64_t)(i - 1) * chunk + 1;
  k_fn = (int64_t)i * chunk;
  if (i - 1 < rem) {
    k_st += i - 1;
    k_fn += i;
  } else {
    k_st += rem;
    k_fn += rem;
  }
}

std::string mij_name(const char *dir, const char *pre, int64_t a, int64_t b) {
  return std::string(dir) + "/" + pre + std::to_string(a) + "_" + std::to_string(b) + ".DAT";
}

int qn_cols(int coll_type) { return coll_type == 1 ? 1 : coll_type == 4 ? 3 : 6; }

}  // namespace

extern "C" int mqct_hip_write_mij_info(const char *dir, int unformatted, int coll_type,
                                       int n_channels, int n_states, int64_t total_size,
                                       int n_r, const int *indx_chann, const int *j12,
                                       const int *m12, const int *chan_qn,
                                       const int *parity_state, const int *ind_mat,
                                       const double *R_com) {
  if (!dir || !indx_chann || !j12 || !m12 || !chan_qn || !ind_mat || !R_com)
    return mqct_fail(MQCT_ERR_ARG, "write_mij_info: null argument");
  if (coll_type != 1 && coll_type != 4 && coll_type != 0)
    return mqct_fail(MQCT_ERR_ARG, "write_mij_info: SYS_TYPE 1, 4, 0 only");
  if (parity_state && coll_type != 0)
    return mqct_fail(MQCT_ERR_ARG, "write_mij_info: exchange parities belong to SYS_TYPE 0");
  const std::string path = std::string(dir) + "/MATRIX_INFO.DAT";
  FILE *f = std::fopen(path.c_str(), unformatted ? "wb" : "w");
  if (!f) return mqct_fail(MQCT_ERR_ARG, "write_mij_info: cannot open " + path);
  const int nq = qn_cols(coll_type);
  bool ok = true;
  if (!unformatted) {
    // bk_print_matrix_info, formatted branch (matrix_ini.f:6846-7021)
    ok &= std::fprintf(f, "COLLISION_TYPE= %1d\n", coll_type) > 0;
    ok &= std::fprintf(f, "NUMBER_OF_CHANLS= %4d\n", n_channels) > 0;
    ok &= std::fprintf(f, "NUMBER_OF_STATES= %6d\n", n_states) > 0;
    ok &= std::fprintf(f, "MATRIX_SIZE= %9lld\n", (long long)total_size) > 0;
    ok &= std::fprintf(f, "R_GRID_SIZE= %4d\n", n_r) > 0;
    static const char *h1[] = {"J"}, *h4[] = {"J", "KA", "KC"},
                      *h0[] = {"J1", "KA1", "KC1", "J2", "KA2", "KC2"};
    const char **hq = coll_type == 1 ? h1 : coll_type == 4 ? h4 : h0;
    std::string hd = fa("#STATE", 8) + " " + fa("#CHANNEL", 8) + " " + fa("J12", 8) + " " + fa("M12", 8);
    for (int q = 0; q < nq; q++) hd += " " + fa(hq[q], 8);
    if (parity_state) hd += " " + fa("P", 2);
    ok &= std::fprintf(f, "%s\n", hd.c_str()) > 0;
    for (int s = 0; s < n_states && ok; s++) {
      const int c = indx_chann[s];
      ok &= std::fprintf(f, "%8d %8d %8d %8d", s + 1, c, j12[s], m12[s]) > 0;
      for (int q = 0; q < nq; q++) ok &= std::fprintf(f, " %8d", chan_qn[(size_t)(c - 1) * nq + q]) > 0;
      if (parity_state) ok &= std::fprintf(f, " %2d", parity_state[s]) > 0;
      ok &= std::fputc('\n', f) != EOF;
    }
    ok &= std::fprintf(f, "%s %s %s\n", fa("MIJ_INDEX", 16).c_str(), fa("ST_1", 8).c_str(), fa("ST_2", 8).c_str()) > 0;
    for (int64_t k = 0; k < total_size && ok; k++)
      ok &= std::fprintf(f, "%16lld %8d %8d\n", (long long)(k + 1), ind_mat[2 * k], ind_mat[2 * k + 1]) > 0;
    ok &= std::fprintf(f, "%s %s\n", fa("#iR", 5).c_str(), fa("R_COM(iR)", 19).c_str()) > 0;
    for (int i = 0; i < n_r && ok; i++) ok &= std::fprintf(f, "%5d %s\n", i + 1, fort_e(R_com[i], 19, 12).c_str()) > 0;
    ok &= std::fprintf(f, " \n") > 0;  // the closing empty list-directed WRITE
  } else {
    // unformatted branch (matrix_ini.f:7030-7135): no pair list
    Rec r;
    r.i32(coll_type); ok &= r.write(f);
    r.i32(n_channels); ok &= r.write(f);
    r.i32(n_states); ok &= r.write(f);
    r.i32((int)total_size); ok &= r.write(f);
    r.i32(n_r); ok &= r.write(f);
    for (int s = 0; s < n_states && ok; s++) {
      const int c = indx_chann[s];
      r.i32(s + 1); r.i32(c); r.i32(j12[s]); r.i32(m12[s]);
      for (int q = 0; q < nq; q++) r.i32(chan_qn[(size_t)(c - 1) * nq + q]);
      if (parity_state) r.i32(parity_state[s]);
      ok &= r.write(f);
    }
    for (int i = 0; i < n_r && ok; i++) {
      r.i32(i + 1);
      r.put(&R_com[i], 8);
      ok &= r.write(f);
    }
  }
  ok &= std::fclose(f) == 0;
  if (!ok) return mqct_fail(MQCT_ERR_ARG, "write_mij_info: write failed");
  return MQCT_OK;
}

extern "C" int mqct_hip_write_mij_dir(const char *dir, int unformatted, int n_files,
                                      int64_t total_size, int n_r, const int *ind_mat,
                                      const double *mat_el, double mij_zero_cut,
                                      int ir_cut1, int ir_cut2) {
  if (!dir || !ind_mat || !mat_el) return mqct_fail(MQCT_ERR_ARG, "write_mij_dir: null argument");
  if (n_files < 1 || total_size < 1 || n_r < 1)
    return mqct_fail(MQCT_ERR_ARG, "write_mij_dir: bad sizes");
  if (n_files > total_size)
    return mqct_fail(MQCT_ERR_ARG, "The #procs is larger than #elements in matrix.");
  if (ir_cut1 < 1 || ir_cut1 > n_r || ir_cut2 < 1 || ir_cut2 > n_r)
    return mqct_fail(MQCT_ERR_ARG, "write_mij_dir: cutoff R index out of range");
  std::vector<int64_t> nz(n_files, 0);
  bool ok = true;
  for (int i = 1; i <= n_files && ok; i++) {
    int64_t k_st, k_fn;
    rank_range(total_size, n_files, i, k_st, k_fn);
    const std::string pm = mij_name(dir, "MIJ_", k_st, k_fn), pi = mij_name(dir, "Ind_", k_st, k_fn);
    FILE *fm = std::fopen(pm.c_str(), unformatted ? "wb" : "w");
    FILE *fi = std::fopen(pi.c_str(), unformatted ? "wb" : "w");
    if (!fm || !fi) {
      if (fm) std::fclose(fm);
      if (fi) std::fclose(fi);
      return mqct_fail(MQCT_ERR_ARG, "write_mij_dir: cannot open " + pm);
    }
    Rec r;
    for (int64_t k = k_st; k <= k_fn && ok; k++) {
      const double *col = mat_el + (size_t)(k - 1) * n_r;
      // bk_print_matrix writes the elements above the cut
      // (matrix_ini.f:7195-7205)
      if (!(std::max(std::fabs(col[ir_cut1 - 1]), std::fabs(col[ir_cut2 - 1])) > mij_zero_cut)) continue;
      nz[i - 1]++;
      const int a = ind_mat[2 * (k - 1)], b = ind_mat[2 * (k - 1) + 1];
      if (!unformatted) {
        for (int ir = 0; ir < n_r && ok; ir++)
          ok &= std::fprintf(fm, "%16lld %8d %s\n", (long long)k, ir + 1, fort_e(col[ir], 19, 12).c_str()) > 0;
        ok &= std::fprintf(fi, "%12d%12d\n", a, b) > 0;  // list-directed
      } else {
        r.i32((int)k); ok &= r.write(fm);
        r.put(col, (size_t)n_r * 8); ok &= r.write(fm);
        r.i32(a); r.i32(b); ok &= r.write(fi);
      }
    }
    ok &= std::fclose(fm) == 0;
    ok &= std::fclose(fi) == 0;
  }
  int64_t total_nz = 0;
  for (int64_t v : nz) total_nz += v;
  // MATRIX_NONZERO_INDEX.DAT (matrix_ini.f:1558-1570): the cut in cm^-1
  {
    const std::string p = std::string(dir) + "/MATRIX_NONZERO_INDEX.DAT";
    FILE *f = std::fopen(p.c_str(), "w");
    if (!f) return mqct_fail(MQCT_ERR_ARG, "write_mij_dir: cannot open " + p);
    ok &= std::fprintf(f, "#Files = %16d\n", n_files) > 0;
    ok &= std::fprintf(f, "Non-Zero #Matrix Elements = %16lld\n", (long long)total_nz) > 0;
    ok &= std::fprintf(f, "The Matrix Cut-Off = %s cm^-1\n", fort_e(mij_zero_cut * kEvToWn * kAuToEv, 19, 12).c_str()) > 0;
    ok &= std::fprintf(f, "%s  %s\n", fa("file_#", 16).c_str(), fa("#non_zero_Mij", 16).c_str()) > 0;
    for (int i = 0; i < n_files && ok; i++) ok &= std::fprintf(f, "%16d  %16lld\n", i + 1, (long long)nz[i]) > 0;
    ok &= std::fclose(f) == 0;
  }
  // MATRIX_FILE_INDEX.DAT (bk_matrix_combination, matrix_ini.f:9064-9092)
  {
    const std::string p = std::string(dir) + "/MATRIX_FILE_INDEX.DAT";
    FILE *f = std::fopen(p.c_str(), "w");
    if (!f) return mqct_fail(MQCT_ERR_ARG, "write_mij_dir: cannot open " + p);
    ok &= std::fprintf(f, "Total #files = %10d\n", n_files) > 0;
    ok &= std::fprintf(f, "%s  %s  %s\n", fa("#Mij", 16).c_str(), fa("#Mij_begin", 16).c_str(),
                       fa("#Mij_end", 16).c_str()) > 0;
    for (int i = 1; i <= n_files && ok; i++) {
      int64_t k_st, k_fn;
      rank_range(total_size, n_files, i, k_st, k_fn);
      ok &= std::fprintf(f, "%16d  %16lld  %16lld\n", i, (long long)k_st, (long long)k_fn) > 0;
    }
    ok &= std::fclose(f) == 0;
  }
  if (!ok) return mqct_fail(MQCT_ERR_ARG, "write_mij_dir: write failed");
  return MQCT_OK;
}

extern "C" int mqct_hip_read_mij_dir(const char *dir, int unformatted, int n_r, int64_t cap,
                                     int64_t *n_nonzero, double *cutoff_cm, int *ind_mat,
                                     double *mat_el) {
  if (!dir || !n_nonzero) return mqct_fail(MQCT_ERR_ARG, "read_mij_dir: null argument");
  // MATRIX_NONZERO_INDEX.DAT: files, non-zero count, cut, per-file counts
  std::string p = std::string(dir) + "/MATRIX_NONZERO_INDEX.DAT";
  FILE *f = std::fopen(p.c_str(), "r");
  if (!f) return mqct_fail(MQCT_ERR_ARG, "read_mij_dir: cannot open " + p);
  char line[512];
  auto tail_after = [&](int col) -> const char * { return std::strlen(line) > (size_t)col ? line + col : ""; };
  int n_files = 0;
  long long nnz = 0;
  double cut = 0.0;
  bool ok = std::fgets(line, sizeof line, f) && std::sscanf(tail_after(9), "%d", &n_files) == 1;
  ok = ok && std::fgets(line, sizeof line, f) && std::sscanf(tail_after(28), "%lld", &nnz) == 1;
  ok = ok && std::fgets(line, sizeof line, f);
  if (ok) {
    std::string v(tail_after(21));
    v = v.substr(0, 19);
    for (char &c : v)
      if (c == 'D' || c == 'd') c = 'E';
    cut = std::strtod(v.c_str(), nullptr);
  }
  ok = ok && std::fgets(line, sizeof line, f);
  std::vector<long long> cnt(n_files > 0 ? n_files : 0);
  for (int i = 0; ok && i < n_files; i++) {
    int idx;
    ok = std::fgets(line, sizeof line, f) && std::sscanf(line, "%d %lld", &idx, &cnt[i]) == 2 && idx == i + 1;
  }
  std::fclose(f);
  if (!ok || n_files < 1) return mqct_fail(MQCT_ERR_ARG, "Error reading MATRIX_NONZERO_INDEX.DAT");
  // MATRIX_FILE_INDEX.DAT: the element range of each file
  p = std::string(dir) + "/MATRIX_FILE_INDEX.DAT";
  f = std::fopen(p.c_str(), "r");
  if (!f) return mqct_fail(MQCT_ERR_ARG, "read_mij_dir: cannot open " + p);
  int fc = 0;
  ok = std::fgets(line, sizeof line, f) && std::sscanf(tail_after(15), "%d", &fc) == 1;
  ok = ok && std::fgets(line, sizeof line, f);
  std::vector<long long> kb(n_files), ke(n_files);
  for (int i = 0; ok && i < n_files; i++) {
    int idx;
    ok = std::fgets(line, sizeof line, f) && std::sscanf(line, "%d %lld %lld", &idx, &kb[i], &ke[i]) == 3 &&
         idx == i + 1;
  }
  std::fclose(f);
  if (!ok) return mqct_fail(MQCT_ERR_ARG, "read_mij_dir: cannot read MATRIX_FILE_INDEX.DAT");
  if (fc != n_files) return mqct_fail(MQCT_ERR_ARG, "#Files do not match in matrix directories");
  *n_nonzero = nnz;
  if (cutoff_cm) *cutoff_cm = cut;
  if (!mat_el && !ind_mat) return MQCT_OK;  // sizes only
  if (nnz > cap) return mqct_fail(MQCT_ERR_ARG, "read_mij_dir: more non-zero elements than cap");
  int64_t st = 0;
  std::vector<unsigned char> b;
  for (int i = 0; i < n_files; i++) {
    const std::string pm = mij_name(dir, "MIJ_", kb[i], ke[i]), pi = mij_name(dir, "Ind_", kb[i], ke[i]);
    FILE *fm = std::fopen(pm.c_str(), unformatted ? "rb" : "r");
    FILE *fi = std::fopen(pi.c_str(), unformatted ? "rb" : "r");
    if (!fm || !fi) {
      if (fm) std::fclose(fm);
      if (fi) std::fclose(fi);
      return mqct_fail(MQCT_ERR_ARG, "ERROR: MATRIX FILES NOT FOUND " + pm);
    }
    for (long long e = 0; ok && e < cnt[i]; e++, st++) {
      if (st >= nnz) { ok = false; break; }
      if (!unformatted) {
        for (int ir = 0; ok && ir < n_r; ir++) {
          ok = std::fgets(line, sizeof line, fm) != nullptr && std::strlen(line) >= 45;
          if (!ok) break;
          std::string v(line + 26, 19);
          for (char &c : v)
            if (c == 'D' || c == 'd') c = 'E';
          if (v.find('E') == std::string::npos) {  // E19.12 with a three-digit exponent
            size_t s = v.find_last_of("+-");
            if (s != std::string::npos && s > 0) v.insert(s, "E");
          }
          if (mat_el) mat_el[(size_t)st * n_r + ir] = std::strtod(v.c_str(), nullptr);
        }
        int a = 0, c = 0;
        ok = ok && std::fgets(line, sizeof line, fi) && std::sscanf(line, "%d %d", &a, &c) == 2;
        if (ind_mat) {
          ind_mat[2 * st] = a;
          ind_mat[2 * st + 1] = c;
        }
      } else {
        ok = read_rec(fm, b) && b.size() == 4;
        ok = ok && read_rec(fm, b) && b.size() == (size_t)n_r * 8;
        if (ok && mat_el) std::memcpy(mat_el + (size_t)st * n_r, b.data(), (size_t)n_r * 8);
        ok = ok && read_rec(fi, b) && b.size() == 8;
        if (ok && ind_mat) std::memcpy(ind_mat + 2 * st, b.data(), 8);
      }
    }
    std::fclose(fm);
    std::fclose(fi);
    if (!ok) return mqct_fail(MQCT_ERR_ARG, "read_mij_dir: short or malformed " + pm);
  }
  if (st != nnz) return mqct_fail(MQCT_ERR_ARG, "read_mij_dir: element count differs from MATRIX_NONZERO_INDEX.DAT");
  return MQCT_OK;
}
