// rnpx_tune.cpp -- the persisted autotuning table (the device counterpart of
// RNP's QR::Tuning hooks, RNP/LA/QR.hpp:41-51, which fix nb = 32 and the
// crossover for every shape).  rnp2_amd/autotune.py times the candidate
// configurations on the device and writes a JSON table:
//
//   {"version": 1, "device": "...", "entries": [
//     {"kind": "factor", "dtype": "d", "m": 8192, "n": 8192, "nb": 128, "ksplit": 0, "ms": 28.2},
//     {"kind": "batched", "dtype": "d", "lds": 40960, "blocked_min": 33, "ms": 8.7}]}
//
// Factor entries give the outer block width (nb 32, 64, 128 or 256: the width of
// the compact-WY block the trailing matrix sees), the split of W = V^H C and
// optionally the CUs reserved for the look-ahead panel stream (panel_cus);
// a factorization whose handle tuning is automatic (nb = 0) takes the entry
// of its element type nearest in (log2 m, log2 n), within a factor of two in
// each.  Batched entries give the LDS bytes per one-wave team of the small
// tier (which also sets the small/large threshold) and the column count from
// which real problems take the blocked large tier.
//
// The table is read once per process, on first use: from RNPX_TUNING_FILE if
// set, else rnpx_tuning.json beside librnpx.so; rnpx_tuning_load replaces it.
// A missing file means defaults; a malformed one is reported by
// rnpx_tuning_load and otherwise ignored.
#include "rnpx_internal.hpp"

#include <dlfcn.h>

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

namespace rnpx {

namespace {

struct Entry {
  std::string kind;  // "factor" or "batched"
  char dtype = 'd';
  double m = 0, n = 0;
  int nb = 0, ksplit = 0, lds = 0, blocked_min = 0;
  int panel_cus = 0;  // factor: CUs of the look-ahead panel stream (0 automatic, -1 none)
};

std::mutex g_mu;
std::vector<Entry> g_table;
bool g_loaded = false;

// A flat reader for the table's shape: an object holding "entries", an
// array of flat objects whose values are numbers or strings.  Returns false
// on anything else.
struct Reader {
  const std::string& s;
  size_t i = 0;
  explicit Reader(const std::string& t) : s(t) {}
  void ws() {
    while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
  }
  bool eat(char c) {
    ws();
    if (i < s.size() && s[i] == c) {
      ++i;
      return true;
    }
    return false;
  }
  bool str(std::string* out) {
    ws();
    if (i >= s.size() || s[i] != '"') return false;
    size_t j = s.find('"', i + 1);
    if (j == std::string::npos) return false;
    *out = s.substr(i + 1, j - i - 1);
    i = j + 1;
    return true;
  }
  bool num(double* out) {
    ws();
    const char* b = s.c_str() + i;
    char* e = nullptr;
    const double v = std::strtod(b, &e);
    if (e == b) return false;
    i += (size_t)(e - b);
    *out = v;
    return true;
  }
  // any scalar, nested object or array (skipped)
  bool skip() {
    ws();
    if (i >= s.size()) return false;
    if (s[i] == '"') {
      std::string t;
      return str(&t);
    }
    if (s[i] == '{' || s[i] == '[') {
      const char open = s[i], close = (open == '{') ? '}' : ']';
      int depth = 0;
      for (; i < s.size(); ++i) {
        if (s[i] == '"') {
          const size_t j = s.find('"', i + 1);
          if (j == std::string::npos) return false;
          i = j;
        } else if (s[i] == open) {
          ++depth;
        } else if (s[i] == close && --depth == 0) {
          ++i;
          return true;
        }
      }
      return false;
    }
    double v;
    if (num(&v)) return true;
    for (const char* w : {"true", "false", "null"})
      if (s.compare(i, std::strlen(w), w) == 0) {
        i += std::strlen(w);
        return true;
      }
    return false;
  }
  bool entry(Entry* e) {
    if (!eat('{')) return false;
    if (eat('}')) return true;
    do {
      std::string k;
      if (!str(&k) || !eat(':')) return false;
      ws();
      if (i < s.size() && s[i] == '"') {
        std::string v;
        if (!str(&v)) return false;
        if (k == "kind") e->kind = v;
        if (k == "dtype") e->dtype = v.empty() ? '?' : v[0];
      } else {
        double v;
        if (!num(&v)) return false;
        if (k == "m") e->m = v;
        else if (k == "n") e->n = v;
        else if (k == "nb") e->nb = (int)v;
        else if (k == "ksplit") e->ksplit = (int)v;
        else if (k == "lds") e->lds = (int)v;
        else if (k == "blocked_min") e->blocked_min = (int)v;
        else if (k == "panel_cus") e->panel_cus = (int)v;
      }
    } while (eat(','));
    return eat('}');
  }
  bool table(std::vector<Entry>* out) {
    if (!eat('{')) return false;
    if (eat('}')) return true;
    do {
      std::string k;
      if (!str(&k) || !eat(':')) return false;
      if (k != "entries") {
        if (!skip()) return false;
        continue;
      }
      if (!eat('[')) return false;
      if (eat(']')) continue;
      do {
        Entry e;
        if (!entry(&e)) return false;
        out->push_back(e);
      } while (eat(','));
      if (!eat(']')) return false;
    } while (eat(','));
    return eat('}');
  }
};

bool valid(const Entry& e) {
  if (e.dtype != 'd' && e.dtype != 'z') return false;
  if (e.kind == "factor")
    return e.m >= 1 && e.n >= 1 && (e.nb == 32 || e.nb == 64 || e.nb == 128 || e.nb == 256) && e.ksplit >= 0 &&
           e.ksplit <= 64 && e.panel_cus >= -1 && e.panel_cus <= 1024;
  if (e.kind == "batched") return e.lds >= 4096 && e.lds <= 160 * 1024 && e.blocked_min >= 1;
  return false;
}

// -1: unreadable file (defaults), -2: malformed, else the entry count.
int read_file(const std::string& path, std::vector<Entry>* out) {
  std::ifstream f(path);
  if (!f) return -1;
  std::stringstream ss;
  ss << f.rdbuf();
  std::vector<Entry> t;
  const std::string text = ss.str();  // the reader keeps a reference: text must outlive it
  Reader r(text);
  if (!r.table(&t)) return -2;
  r.ws();
  if (r.i != r.s.size()) return -2;
  for (const Entry& e : t)
    if (!valid(e)) return -2;
  *out = t;
  return (int)t.size();
}

std::string default_path() {
  if (const char* e = getenv("RNPX_TUNING_FILE")) return e;
  Dl_info info;
  if (dladdr(reinterpret_cast<void*>(&default_path), &info) && info.dli_fname) {
    std::string p(info.dli_fname);
    const size_t slash = p.rfind('/');
    return (slash == std::string::npos ? std::string(".") : p.substr(0, slash)) + "/rnpx_tuning.json";
  }
  return "rnpx_tuning.json";
}

void ensure_loaded() {
  if (g_loaded) return;
  g_loaded = true;
  std::vector<Entry> t;
  if (read_file(default_path(), &t) > 0) g_table = t;
}

}  // namespace

int tuning_load(const char* path) {
  std::lock_guard<std::mutex> lk(g_mu);
  std::vector<Entry> t;
  const int r = read_file(path ? std::string(path) : default_path(), &t);
  g_loaded = true;
  if (r == -1 && !path) {
    g_table.clear();  // no table at the default location: defaults
    return 0;
  }
  if (r < 0) return r == -1 ? RNPX_ERR_UNSUPPORTED : RNPX_ERR_ARG;
  g_table = t;
  return r;
}

bool tuning_factor(char dtype, int m, int n, int* nb, int* ksplit, int* panel_cus) {
  std::lock_guard<std::mutex> lk(g_mu);
  ensure_loaded();
  const Entry* best = nullptr;
  double bd = 0;
  for (const Entry& e : g_table) {
    if (e.kind != "factor" || e.dtype != dtype) continue;
    const double dm = std::fabs(std::log2((double)m / e.m)), dn = std::fabs(std::log2((double)n / e.n));
    if (dm > 1.0 || dn > 1.0) continue;
    if (!best || dm + dn < bd) {
      best = &e;
      bd = dm + dn;
    }
  }
  if (!best) return false;
  *nb = best->nb;
  *ksplit = best->ksplit;
  if (panel_cus) *panel_cus = best->panel_cus;
  return true;
}

bool tuning_batched(char dtype, int* lds, int* blocked_min) {
  std::lock_guard<std::mutex> lk(g_mu);
  ensure_loaded();
  for (const Entry& e : g_table) {
    if (e.kind == "batched" && e.dtype == dtype) {
      *lds = e.lds;
      *blocked_min = e.blocked_min;
      return true;
    }
  }
  return false;
}

}  // namespace rnpx

extern "C" {

int rnpx_tuning_load(const char* path) { return rnpx::tuning_load(path); }

int rnpx_tuning_lookup_factor(char dtype, size_t m, size_t n, int* nb, int* ksplit) {
  if (!nb || !ksplit) return RNPX_ERR_ARG;
  if (m > (size_t)0x7fffffff || n > (size_t)0x7fffffff) return 0;
  return rnpx::tuning_factor(dtype, (int)m, (int)n, nb, ksplit) ? 1 : 0;
}

int rnpx_tuning_lookup_panel_cus(char dtype, size_t m, size_t n, int* panel_cus) {
  if (!panel_cus) return RNPX_ERR_ARG;
  if (m > (size_t)0x7fffffff || n > (size_t)0x7fffffff) return 0;
  int nb = 0, ks = 0;
  return rnpx::tuning_factor(dtype, (int)m, (int)n, &nb, &ks, panel_cus) ? 1 : 0;
}

int rnpx_tuning_lookup_batched(char dtype, int* lds_bytes, int* blocked_min) {
  if (!lds_bytes || !blocked_min) return RNPX_ERR_ARG;
  return rnpx::tuning_batched(dtype, lds_bytes, blocked_min) ? 1 : 0;
}

}  // extern "C"
