// caseio.cpp -- case-directory IO, text-to-bin and the compare-x metric,
// with the reference's formats and message texts (see include/hbj_io.h).
#include <sys/stat.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "hbj_internal.hpp"
#include "hbj_io.h"

using hbj::set_error;
using hbj::clear_error;

struct hbj_case {
  uint32_t n = 0, n_cols = 0, nnz = 0;
  std::vector<uint32_t> row_ptr, col_idx, block_starts;
  std::vector<float> values, b, x0;
};

namespace {

constexpr uint32_t kMagicCsr1 = 0x43535231u;

bool read_file(const std::string& path, std::vector<uint8_t>& out) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) {
    set_error("Failed to read %s: %s", path.c_str(), std::strerror(errno));
    return false;
  }
  std::fseek(f, 0, SEEK_END);
  const long len = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  out.resize(len > 0 ? (size_t)len : 0);
  const size_t got = out.empty() ? 0 : std::fread(out.data(), 1, out.size(), f);
  std::fclose(f);
  if (got != out.size()) {
    set_error("Failed to read %s: short read", path.c_str());
    return false;
  }
  return true;
}

// loaders.rs:186-237: little-endian cursor with labelled EOF errors
struct Cursor {
  const std::vector<uint8_t>& b;
  size_t pos = 0;
  size_t remaining() const { return pos < b.size() ? b.size() - pos : 0; }
  bool take(size_t n, const std::string& label, const uint8_t** p) {
    if (n > remaining()) {
      set_error("Unexpected EOF while reading %s: need %zu bytes, have %zu", label.c_str(), n,
                remaining());
      return false;
    }
    *p = b.data() + pos;
    pos += n;
    return true;
  }
  bool u32(const std::string& label, uint32_t* v) {
    const uint8_t* p;
    if (!take(4, label, &p)) return false;
    *v = (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
    return true;
  }
  bool u32s(size_t n, const char* label, std::vector<uint32_t>& out) {
    if (n * 4 > remaining()) {  // report the first missing element, as the per-element reader does
      const size_t idx = remaining() / 4;
      set_error("Unexpected EOF while reading %s[%zu]: need 4 bytes, have %zu", label, idx,
                remaining() % 4);
      return false;
    }
    out.resize(n);
    if (n) std::memcpy(out.data(), b.data() + pos, n * 4);
    pos += n * 4;
    return true;
  }
  bool f32s(size_t n, const char* label, std::vector<float>& out) {
    if (n * 4 > remaining()) {
      const size_t idx = remaining() / 4;
      set_error("Unexpected EOF while reading %s[%zu]: need 4 bytes, have %zu", label, idx,
                remaining() % 4);
      return false;
    }
    out.resize(n);
    if (n) std::memcpy(out.data(), b.data() + pos, n * 4);
    pos += n * 4;
    return true;
  }
};

std::string join(const std::string& dir, const char* name) {
  if (dir.empty()) return name;
  return dir.back() == '/' ? dir + name : dir + "/" + name;
}

bool load_vector(const std::string& path, std::vector<float>& v, uint32_t* n) {
  std::vector<uint8_t> bytes;
  if (!read_file(path, bytes)) return false;
  Cursor c{bytes};
  if (!c.u32("n", n)) return false;
  if (!c.f32s(*n, "values", v)) return false;
  if (c.remaining() != 0) {
    set_error("%s has %zu trailing bytes after expected payload", path.c_str(), c.remaining());
    return false;
  }
  return true;
}

}  // namespace

extern "C" {

hbj_status hbj_case_load(const char* case_dir, hbj_case** out) {
  clear_error();
  if (!case_dir || !out) {
    set_error("hbj_case_load: bad argument");
    return HBJ_E_INVALID;
  }
  *out = nullptr;
  const std::string dir = case_dir;
  auto* cs = new hbj_case();
  auto fail = [&]() {
    delete cs;
    return HBJ_E_INVALID;
  };
  {  // loaders.rs:44-124
    const std::string path = join(dir, "matrix.csr.bin");
    std::vector<uint8_t> bytes;
    if (!read_file(path, bytes)) return fail();
    Cursor c{bytes};
    uint32_t magic, version, r0, r1, r2;
    if (!c.u32("magic", &magic)) return fail();
    if (magic != kMagicCsr1) {
      set_error("Bad matrix magic in %s: expected 0x%08x, got 0x%08x", path.c_str(), kMagicCsr1, magic);
      return fail();
    }
    if (!c.u32("version", &version)) return fail();
    if (version != 1) {
      set_error("Unsupported matrix version in %s: expected 1, got %u", path.c_str(), version);
      return fail();
    }
    if (!c.u32("n_rows", &cs->n) || !c.u32("n_cols", &cs->n_cols) || !c.u32("nnz", &cs->nnz) ||
        !c.u32("reserved0", &r0) || !c.u32("reserved1", &r1) || !c.u32("reserved2", &r2))
      return fail();
    if (!c.u32s((size_t)cs->n + 1, "row_ptr", cs->row_ptr)) return fail();
    if (!c.u32s(cs->nnz, "col_idx", cs->col_idx)) return fail();
    if (!c.f32s(cs->nnz, "values", cs->values)) return fail();
    if (cs->row_ptr[0] != 0) {
      set_error("CSR row_ptr must start with 0");
      return fail();
    }
    if (cs->row_ptr.back() != cs->nnz) {
      set_error("CSR row_ptr must end with nnz (%u), got Some(%u)", cs->nnz, cs->row_ptr.back());
      return fail();
    }
    for (size_t i = 0; i + 1 < cs->row_ptr.size(); ++i)
      if (cs->row_ptr[i + 1] < cs->row_ptr[i]) {
        set_error("CSR row_ptr must be non-decreasing");
        return fail();
      }
    uint32_t max_col = 0;
    for (uint32_t v : cs->col_idx) max_col = std::max(max_col, v);
    if (cs->n_cols > 0 && max_col >= cs->n_cols) {
      set_error("CSR col_idx contains out-of-range column %u for n_cols=%u", max_col, cs->n_cols);
      return fail();
    }
    if (c.remaining() != 0) {
      set_error("matrix.csr.bin has %zu trailing bytes after expected payload", c.remaining());
      return fail();
    }
  }
  uint32_t nb_ = 0, nx0 = 0;
  if (!load_vector(join(dir, "rhs.bin"), cs->b, &nb_)) return fail();
  if (!load_vector(join(dir, "x0.bin"), cs->x0, &nx0)) return fail();
  {  // loaders.rs:146-162
    const std::string path = join(dir, "block_starts.bin");
    std::vector<uint8_t> bytes;
    if (!read_file(path, bytes)) return fail();
    Cursor c{bytes};
    uint32_t len;
    if (!c.u32("len", &len) || !c.u32s(len, "block_starts", cs->block_starts)) return fail();
    if (c.remaining() != 0) {
      set_error("%s has %zu trailing bytes after expected payload", path.c_str(), c.remaining());
      return fail();
    }
  }
  // loaders.rs:14-34
  if (cs->n != cs->n_cols) {
    set_error("matrix must be square for PCG: got %ux%u", cs->n, cs->n_cols);
    return fail();
  }
  if (nb_ != cs->n) {
    set_error("rhs length mismatch: b.n=%u but matrix n_rows=%u", nb_, cs->n);
    return fail();
  }
  if (nx0 != cs->n) {
    set_error("x0 length mismatch: x0.n=%u but matrix n_rows=%u", nx0, cs->n);
    return fail();
  }
  const auto& s = cs->block_starts;
  if (s.empty()) {
    set_error("block_starts must not be empty");
    return fail();
  }
  if (s[0] != 0) {
    set_error("block_starts must start with 0, got %u", s[0]);
    return fail();
  }
  if (s.back() != cs->n) {
    set_error("block_starts must end with n (%u), got %u", cs->n, s.back());
    return fail();
  }
  for (size_t i = 0; i + 1 < s.size(); ++i)
    if (s[i + 1] <= s[i]) {
      set_error("block_starts must be strictly increasing");
      return fail();
    }
  *out = cs;
  return HBJ_OK;
}

void hbj_case_free(hbj_case* c) { delete c; }

hbj_status hbj_case_view(const hbj_case* c, uint32_t* n, uint32_t* nnz, uint32_t* n_starts,
                         const uint32_t** row_ptr, const uint32_t** col_idx, const float** values,
                         const float** b, const float** x0, const uint32_t** block_starts) {
  if (!c) {
    set_error("hbj_case_view: null case");
    return HBJ_E_INVALID;
  }
  if (n) *n = c->n;
  if (nnz) *nnz = c->nnz;
  if (n_starts) *n_starts = (uint32_t)c->block_starts.size();
  if (row_ptr) *row_ptr = c->row_ptr.data();
  if (col_idx) *col_idx = c->col_idx.data();
  if (values) *values = c->values.data();
  if (b) *b = c->b.data();
  if (x0) *x0 = c->x0.data();
  if (block_starts) *block_starts = c->block_starts.data();
  return HBJ_OK;
}

}  // extern "C"

// ------------------------------------------------------------ text-to-bin

namespace {

struct LineReader {
  std::ifstream in;
  std::string line;
  // next non-empty, non-comment line, trimmed (convert main.rs:118-131)
  bool next(std::string& out) {
    while (std::getline(in, line)) {
      size_t a = line.find_first_not_of(" \t\r\n");
      if (a == std::string::npos) continue;
      size_t b = line.find_last_not_of(" \t\r\n");
      std::string t = line.substr(a, b - a + 1);
      if (t[0] == '#') continue;
      out = t;
      return true;
    }
    return false;
  }
};

bool ends_with(const std::string& s, const char* suf) {
  const size_t k = std::strlen(suf);
  return s.size() >= k && s.compare(s.size() - k, k, suf) == 0;
}
bool header_like(const std::string& l) {  // convert main.rs:141-146
  return l.rfind("n ", 0) == 0 || l.rfind("nnz ", 0) == 0 || ends_with(l, "_u32") ||
         ends_with(l, "_f32");
}

template <typename T>
bool read_tokens(LineReader& r, const std::string& section, size_t need, std::vector<T>& out,
                 bool is_f32) {
  out.clear();
  out.reserve(need);
  std::string l;
  while (out.size() < need) {
    if (!r.next(l)) {
      set_error("Section %s: EOF while reading values (got %zu, need %zu)", section.c_str(),
                out.size(), need);
      return false;
    }
    if (header_like(l)) {
      set_error("Section %s: hit unexpected header '%s' before collecting len=%zu values",
                section.c_str(), l.c_str(), need);
      return false;
    }
    std::istringstream ss(l);
    std::string t;
    while (out.size() < need && ss >> t) {
      char* end = nullptr;
      errno = 0;
      if (is_f32) {
        const float v = std::strtof(t.c_str(), &end);
        if (end == t.c_str() || *end) {
          set_error("Section %s: failed to parse f32 token '%s': invalid float literal",
                    section.c_str(), t.c_str());
          return false;
        }
        out.push_back((T)v);
      } else {
        const unsigned long long v = std::strtoull(t.c_str(), &end, 10);
        if (end == t.c_str() || *end || t[0] == '-' || v > 0xffffffffull) {
          set_error("Section %s: failed to parse u32 token '%s': invalid digit found in string",
                    section.c_str(), t.c_str());
          return false;
        }
        out.push_back((T)v);
      }
    }
  }
  return true;
}

bool write_bytes(const std::string& path, const std::vector<uint8_t>& data) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) {
    set_error("Failed to create %s: %s", path.c_str(), std::strerror(errno));
    return false;
  }
  const size_t w = data.empty() ? 0 : std::fwrite(data.data(), 1, data.size(), f);
  std::fclose(f);
  if (w != data.size()) {
    set_error("write %s: short write", path.c_str());
    return false;
  }
  return true;
}

void put_u32(std::vector<uint8_t>& o, uint32_t v) {
  for (int i = 0; i < 4; ++i) o.push_back((uint8_t)(v >> (8 * i)));
}
template <typename T>
void put_arr(std::vector<uint8_t>& o, const std::vector<T>& a) {
  const size_t at = o.size();
  o.resize(at + a.size() * 4);
  if (!a.empty()) std::memcpy(o.data() + at, a.data(), a.size() * 4);
}

bool rm_rf(const std::string& dir) {
  const std::string cmd = "rm -rf -- '" + dir + "'";
  return std::system(cmd.c_str()) == 0;
}

}  // namespace

extern "C" {

hbj_status hbj_text_to_bin(const char* input_txt, const char* out_dir, int force) {
  clear_error();
  if (!input_txt || !out_dir) {
    set_error("hbj_text_to_bin: bad argument");
    return HBJ_E_INVALID;
  }
  LineReader r;
  r.in.open(input_txt);
  if (!r.in) {
    set_error("Failed to open %s: %s", input_txt, std::strerror(errno));
    return HBJ_E_INVALID;
  }
  long long n = -1, nnz = -1;
  std::vector<uint32_t> row_ptr, col_idx, bs;
  std::vector<float> values, b, x0, x_ref;
  bool has[7] = {false, false, false, false, false, false, false};
  std::string l;
  while (r.next(l)) {  // convert main.rs:238-325
    if (l.rfind("n ", 0) == 0) {
      n = std::atoll(l.c_str() + 2);
      continue;
    }
    if (l.rfind("nnz ", 0) == 0) {
      nnz = std::atoll(l.c_str() + 4);
      continue;
    }
    const std::string section = l;
    std::string len_line;
    if (!r.next(len_line)) {
      set_error("Section %s: missing len line (EOF)", section.c_str());
      return HBJ_E_INVALID;
    }
    if (len_line.rfind("len ", 0) != 0) {
      set_error("Section %s: expected 'len N', got '%s'", section.c_str(), len_line.c_str());
      return HBJ_E_INVALID;
    }
    const size_t len = (size_t)std::strtoull(len_line.c_str() + 4, nullptr, 10);
    bool ok = true;
    if (section == "row_ptr_u32") ok = read_tokens(r, section, len, row_ptr, false), has[0] = true;
    else if (section == "col_idx_u32") ok = read_tokens(r, section, len, col_idx, false), has[1] = true;
    else if (section == "values_f32") ok = read_tokens(r, section, len, values, true), has[2] = true;
    else if (section == "b_f32") ok = read_tokens(r, section, len, b, true), has[3] = true;
    else if (section == "x0_f32") ok = read_tokens(r, section, len, x0, true), has[4] = true;
    else if (section == "block_starts_u32") ok = read_tokens(r, section, len, bs, false), has[5] = true;
    else if (section == "x_ref_f32") ok = read_tokens(r, section, len, x_ref, true), has[6] = true;
    else {
      std::fprintf(stderr, "Warning: skipping unknown section: %s (len %zu)\n", section.c_str(), len);
      std::vector<float> skip;  // token count only; values are not parsed
      size_t skipped = 0;
      std::string sl;
      while (skipped < len) {
        if (!r.next(sl)) {
          set_error("Section %s: EOF while skipping values (skipped %zu, need %zu)", section.c_str(),
                    skipped, len);
          return HBJ_E_INVALID;
        }
        if (header_like(sl)) {
          set_error("Section %s: hit unexpected header '%s' before skipping len=%zu tokens",
                    section.c_str(), sl.c_str(), len);
          return HBJ_E_INVALID;
        }
        std::istringstream ss(sl);
        std::string t;
        while (ss >> t) ++skipped;
      }
    }
    if (!ok) return HBJ_E_INVALID;
  }
  // convert main.rs:327-399
  if (n < 0) { set_error("Missing 'n'"); return HBJ_E_INVALID; }
  if (nnz < 0) { set_error("Missing 'nnz'"); return HBJ_E_INVALID; }
  const char* missing[6] = {"Missing row_ptr_u32", "Missing col_idx_u32", "Missing values_f32",
                            "Missing b_f32", "Missing x0_f32", "Missing block_starts_u32"};
  for (int i = 0; i < 6; ++i)
    if (!has[i]) { set_error("%s", missing[i]); return HBJ_E_INVALID; }
  if ((long long)row_ptr.size() != n + 1) {
    set_error("row_ptr_u32 len must be n+1 (%lld), got %zu", n + 1, row_ptr.size());
    return HBJ_E_INVALID;
  }
  if ((long long)col_idx.size() != nnz || (long long)values.size() != nnz) {
    set_error("col_idx/values len must be nnz (%lld), got col_idx=%zu, values=%zu", nnz,
              col_idx.size(), values.size());
    return HBJ_E_INVALID;
  }
  if ((long long)b.size() != n || (long long)x0.size() != n) {
    set_error("b/x0 len must be n (%lld), got b=%zu, x0=%zu", n, b.size(), x0.size());
    return HBJ_E_INVALID;
  }
  if (has[6] && (long long)x_ref.size() != n) {
    set_error("x_ref_f32 len must be n (%lld), got %zu", n, x_ref.size());
    return HBJ_E_INVALID;
  }
  if (bs.empty()) { set_error("block_starts_u32 must not be empty"); return HBJ_E_INVALID; }
  if (bs[0] != 0) { set_error("block_starts_u32 must start with 0, got %u", bs[0]); return HBJ_E_INVALID; }
  if ((long long)bs.back() != n) {
    set_error("block_starts_u32 must end with n (%lld), got %u", n, bs.back());
    return HBJ_E_INVALID;
  }
  for (size_t i = 0; i + 1 < bs.size(); ++i)
    if (bs[i + 1] <= bs[i]) { set_error("block_starts_u32 must be strictly increasing"); return HBJ_E_INVALID; }
  // convert main.rs:407-494
  const std::string dir = out_dir;
  struct stat stt;
  if (stat(out_dir, &stt) == 0) {
    if (!force) {
      set_error("Output dir %s already exists. Use --force to overwrite.", out_dir);
      return HBJ_E_INVALID;
    }
    if (!rm_rf(dir)) { set_error("Failed to remove %s", out_dir); return HBJ_E_INVALID; }
  }
  if (std::system(("mkdir -p -- '" + dir + "'").c_str()) != 0) {
    set_error("Failed to create %s", out_dir);
    return HBJ_E_INVALID;
  }
  std::vector<uint8_t> o;
  put_u32(o, kMagicCsr1);
  put_u32(o, 1);
  put_u32(o, (uint32_t)n);
  put_u32(o, (uint32_t)n);
  put_u32(o, (uint32_t)nnz);
  put_u32(o, 0);
  put_u32(o, 0);
  put_u32(o, 0);
  put_arr(o, row_ptr);
  put_arr(o, col_idx);
  put_arr(o, values);
  if (!write_bytes(join(dir, "matrix.csr.bin"), o)) return HBJ_E_INVALID;
  auto vec = [&](const char* name, const std::vector<float>& v) {
    std::vector<uint8_t> w;
    put_u32(w, (uint32_t)v.size());
    put_arr(w, v);
    return write_bytes(join(dir, name), w);
  };
  if (!vec("rhs.bin", b) || !vec("x0.bin", x0)) return HBJ_E_INVALID;
  {
    std::vector<uint8_t> w;
    put_u32(w, (uint32_t)bs.size());
    put_arr(w, bs);
    if (!write_bytes(join(dir, "block_starts.bin"), w)) return HBJ_E_INVALID;
  }
  if (has[6] && !vec("x_ref.bin", x_ref)) return HBJ_E_INVALID;
  char meta[1024];
  std::snprintf(meta, sizeof meta,
                "{\n  \"format\": \"wgpu-solver-backend-text-v0\",\n  \"n\": %lld,\n  \"nnz\": %lld,\n"
                "  \"row_ptr_len\": %zu,\n  \"col_idx_len\": %zu,\n  \"values_len\": %zu,\n"
                "  \"rhs_len\": %zu,\n  \"x0_len\": %zu,\n  \"block_starts_len\": %zu,\n"
                "  \"x_ref_present\": %s\n}",
                n, nnz, row_ptr.size(), col_idx.size(), values.size(), b.size(), x0.size(),
                bs.size(), has[6] ? "true" : "false");
  std::vector<uint8_t> mb(meta, meta + std::strlen(meta));
  if (!write_bytes(join(dir, "meta.json"), mb)) return HBJ_E_INVALID;
  return HBJ_OK;
}

// ---------------------------------------------------------- x.bin, compare

hbj_status hbj_vec_bin_write(const char* path, const float* x, uint32_t n) {
  clear_error();
  if (!path || (n && !x)) {
    set_error("hbj_vec_bin_write: bad argument");
    return HBJ_E_INVALID;
  }
  std::string p = path;
  const size_t slash = p.find_last_of('/');
  if (slash != std::string::npos && slash > 0)  // cli main.rs:450-455 create_dir_all(parent)
    if (std::system(("mkdir -p -- '" + p.substr(0, slash) + "'").c_str()) != 0) {
      set_error("create_dir_all %s: failed", p.substr(0, slash).c_str());
      return HBJ_E_INVALID;
    }
  std::vector<uint8_t> w;
  put_u32(w, n);
  w.resize(4 + 4ull * n);
  if (n) std::memcpy(w.data() + 4, x, 4ull * n);
  return write_bytes(p, w) ? HBJ_OK : HBJ_E_INVALID;
}

hbj_status hbj_vec_bin_read(const char* path, float* x, uint32_t cap, uint32_t* n) {
  clear_error();
  if (!path || !n) {
    set_error("hbj_vec_bin_read: bad argument");
    return HBJ_E_INVALID;
  }
  std::vector<uint8_t> bytes;
  if (!read_file(path, bytes)) return HBJ_E_INVALID;
  if (bytes.size() < 4) {
    set_error("read len header %s: failed to fill whole buffer", path);
    return HBJ_E_INVALID;
  }
  uint32_t len;
  std::memcpy(&len, bytes.data(), 4);
  *n = len;
  if (!x) return HBJ_OK;
  if (len > cap) {
    set_error("hbj_vec_bin_read: %s holds %u values, capacity %u", path, len, cap);
    return HBJ_E_INVALID;
  }
  if (bytes.size() < 4 + 4ull * len) {  // cli main.rs:571-575
    set_error("read f32[%llu] %s: failed to fill whole buffer",
              (unsigned long long)((bytes.size() - 4) / 4), path);
    return HBJ_E_INVALID;
  }
  if (len) std::memcpy(x, bytes.data() + 4, 4ull * len);
  return HBJ_OK;
}

hbj_status hbj_compare_x(const float* x_ref, const float* x, uint32_t n, uint32_t top_k,
                         hbj_compare_stats* out, uint32_t* worst_idx) {
  clear_error();
  if (!out || (n && (!x_ref || !x)) || (top_k && !worst_idx)) {
    set_error("hbj_compare_x: bad argument");
    return HBJ_E_INVALID;
  }
  double se = 0.0, sr = 0.0;
  float max_abs = 0.0f;
  std::vector<std::pair<float, uint32_t>> worst;  // (abs_err, i), largest first
  for (uint32_t i = 0; i < n; ++i) {
    const float a = x_ref[i], b = x[i];
    const float ae = std::fabs(b - a);
    max_abs = std::max(max_abs, ae);
    se += (double)ae * (double)ae;
    sr += (double)a * (double)a;
    if (top_k) {  // cli main.rs:623-641, stable descending order
      if (worst.size() < top_k || ae > worst.back().first) {
        if (worst.size() == top_k) worst.pop_back();
        worst.push_back({ae, i});
        std::stable_sort(worst.begin(), worst.end(),
                         [](const auto& p, const auto& q) { return p.first > q.first; });
      }
    }
  }
  std::memset(out, 0, sizeof *out);
  out->n = n;
  out->l2_err = std::sqrt(se);
  out->l2_ref = std::sqrt(sr);
  out->rel_l2 = out->l2_err / std::max(out->l2_ref, 1e-30);
  out->rmse = std::sqrt(se / (double)std::max<uint32_t>(n, 1));
  out->max_abs_err = max_abs;
  out->n_worst = (uint32_t)worst.size();
  for (size_t k = 0; k < worst.size(); ++k) worst_idx[k] = worst[k].second;
  return HBJ_OK;
}

}  // extern "C"
