// Reader of an OpenFOAM polyMesh directory in ASCII format (points, faces,
// owner, neighbour, boundary), the on-disk input of the reference's Readit
// (GridAndGeometry/ReadFoamMesh.f90:31-230).  Host code only.
//
// Conventions kept from the reference's reader: the count of each list is
// the last token before the line that opens the list; a point is a
// parenthesised triple (possibly spread over lines); a face is
// "n(i0 i1 ...)" with 0-based point indices; owner and neighbour hold one
// 0-based cell index per entry; a boundary entry is "name { ... nFaces N;
// startFace S; }".  The FoamFile header and comments are skipped.
#include <hip/hip_runtime.h>

#include <cctype>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "fsamd_simple.h"

namespace fs {
void set_error(const std::string &msg);
}

struct fs_polymesh {
  std::vector<double> x, y, z;
  std::vector<int> face_ptr, face_nodes, owner, neighbour;
  std::vector<std::string> patch_names;
  std::vector<int> patch_start, patch_nfaces;
};

namespace {

// The file's contents with /* */ and // comments removed.
std::string slurp(const std::string &path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string s = ss.str();
  std::string out;
  out.reserve(s.size());
  for (size_t i = 0; i < s.size(); ++i) {
    if (s[i] == '/' && i + 1 < s.size() && s[i + 1] == '/') {
      while (i < s.size() && s[i] != '\n') ++i;
      out.push_back('\n');
    } else if (s[i] == '/' && i + 1 < s.size() && s[i + 1] == '*') {
      i += 2;
      while (i + 1 < s.size() && !(s[i] == '*' && s[i + 1] == '/')) ++i;
      ++i;
    } else if (s[i] != '\r') {
      out.push_back(s[i]);
    }
  }
  return out;
}

struct Cursor {
  const std::string &s;
  size_t p = 0;
  explicit Cursor(const std::string &str) : s(str) {}
  void skip_ws() {
    while (p < s.size() && std::isspace((unsigned char)s[p])) ++p;
  }
  // skips a "FoamFile { ... }" header when present
  void skip_header() {
    skip_ws();
    if (s.compare(p, 8, "FoamFile") == 0) {
      const size_t b = s.find('{', p), e = s.find('}', p);
      if (b == std::string::npos || e == std::string::npos) throw std::runtime_error("broken FoamFile header");
      p = e + 1;
    }
  }
  long long integer() {
    skip_ws();
    size_t q = p;
    if (q < s.size() && (s[q] == '-' || s[q] == '+')) ++q;
    while (q < s.size() && std::isdigit((unsigned char)s[q])) ++q;
    if (q == p) throw std::runtime_error("integer expected at offset " + std::to_string(p));
    const long long v = std::stoll(s.substr(p, q - p));
    p = q;
    return v;
  }
  double real() {
    skip_ws();
    const char *b = s.c_str() + p;
    char *e = nullptr;
    const double v = std::strtod(b, &e);
    if (e == b) throw std::runtime_error("number expected at offset " + std::to_string(p));
    p += (size_t)(e - b);
    return v;
  }
  void expect(char c) {
    skip_ws();
    if (p >= s.size() || s[p] != c)
      throw std::runtime_error(std::string("'") + c + "' expected at offset " + std::to_string(p));
    ++p;
  }
  std::string word() {
    skip_ws();
    size_t q = p;
    while (q < s.size() && !std::isspace((unsigned char)s[q]) && s[q] != '{' && s[q] != '}' && s[q] != ';' &&
           s[q] != '(' && s[q] != ')')
      ++q;
    std::string w = s.substr(p, q - p);
    p = q;
    return w;
  }
};

// count and opening parenthesis of a list
long long list_head(Cursor &c) {
  c.skip_header();
  const long long n = c.integer();
  if (n < 0) throw std::runtime_error("negative list size");
  c.expect('(');
  return n;
}

} // namespace

extern "C" {

int fs_polymesh_read(const char *dir, int *counts, fs_polymesh **out) {
  try {
    if (!dir || !counts || !out) {
      fs::set_error("fs_polymesh_read: NULL argument");
      return -1;
    }
    *out = nullptr;
    auto m = std::make_unique<fs_polymesh>();
    const std::string d(dir);
    {
      const std::string s = slurp(d + "/points");
      Cursor c(s);
      const long long n = list_head(c);
      m->x.resize(n);
      m->y.resize(n);
      m->z.resize(n);
      for (long long i = 0; i < n; ++i) {
        c.expect('(');
        m->x[i] = c.real();
        m->y[i] = c.real();
        m->z[i] = c.real();
        c.expect(')');
      }
      c.expect(')');
    }
    {
      const std::string s = slurp(d + "/faces");
      Cursor c(s);
      const long long n = list_head(c);
      m->face_ptr.assign(1, 0);
      for (long long f = 0; f < n; ++f) {
        const long long k = c.integer();
        if (k < 3) throw std::runtime_error("face " + std::to_string(f) + " has fewer than 3 points");
        c.expect('(');
        for (long long q = 0; q < k; ++q) {
          const long long v = c.integer();
          if (v < 0 || v >= (long long)m->x.size())
            throw std::runtime_error("face " + std::to_string(f) + " names a point out of range");
          m->face_nodes.push_back((int)v);
        }
        c.expect(')');
        m->face_ptr.push_back((int)m->face_nodes.size());
      }
      c.expect(')');
    }
    const int nfaces = (int)m->face_ptr.size() - 1;
    auto read_cells = [&](const char *name, std::vector<int> &v) {
      const std::string s = slurp(d + "/" + name);
      Cursor c(s);
      const long long n = list_head(c);
      v.resize(n);
      for (long long i = 0; i < n; ++i) {
        const long long k = c.integer();
        if (k < 0) throw std::runtime_error(std::string(name) + ": negative cell index");
        v[i] = (int)k;
      }
      c.expect(')');
    };
    read_cells("owner", m->owner);
    read_cells("neighbour", m->neighbour);
    if ((int)m->owner.size() != nfaces) throw std::runtime_error("owner and faces differ in length");
    if (m->neighbour.size() > m->owner.size()) throw std::runtime_error("more neighbours than faces");
    {
      const std::string s = slurp(d + "/boundary");
      Cursor c(s);
      const long long n = list_head(c);
      for (long long b = 0; b < n; ++b) {
        const std::string name = c.word();
        c.expect('{');
        long long nf = -1, sf = -1;
        while (true) {
          c.skip_ws();
          if (c.p < s.size() && s[c.p] == '}') {
            ++c.p;
            break;
          }
          const std::string key = c.word();
          if (key.empty()) throw std::runtime_error("broken boundary entry " + name);
          if (key == "nFaces") nf = c.integer();
          else if (key == "startFace") sf = c.integer();
          // skip the rest of the entry (values may be words or lists)
          int depth = 0;
          while (c.p < s.size()) {
            const char ch = s[c.p];
            if (ch == '(') ++depth;
            if (ch == ')') --depth;
            if (ch == ';' && depth == 0) {
              ++c.p;
              break;
            }
            ++c.p;
          }
        }
        if (nf < 0 || sf < 0) throw std::runtime_error("boundary " + name + " lacks nFaces or startFace");
        m->patch_names.push_back(name);
        m->patch_nfaces.push_back((int)nf);
        m->patch_start.push_back((int)sf);
      }
    }
    // the patches must cover the boundary faces in order (the reference's
    // face numbering: interior faces, then the patches one after another)
    int next = (int)m->neighbour.size();
    for (size_t b = 0; b < m->patch_start.size(); ++b) {
      if (m->patch_start[b] != next) throw std::runtime_error("patch " + m->patch_names[b] + " does not start where the previous one ends");
      next += m->patch_nfaces[b];
    }
    if (next != nfaces) throw std::runtime_error("patches do not cover the boundary faces");
    counts[0] = (int)m->x.size();
    counts[1] = nfaces;
    counts[2] = (int)m->face_nodes.size();
    counts[3] = (int)m->neighbour.size();
    counts[4] = (int)m->patch_names.size();
    *out = m.release();
    return 0;
  } catch (const std::exception &e) {
    fs::set_error(std::string("fs_polymesh_read: ") + e.what());
    return -1;
  }
}

int fs_polymesh_get(fs_polymesh *m, double *x, double *y, double *z, int *face_ptr, int *face_nodes, int *owner,
                    int *neighbour, int *patch_start, int *patch_nfaces, char *patch_names, int name_len) {
  if (!m) {
    fs::set_error("fs_polymesh_get: NULL mesh");
    return -1;
  }
  auto put = [](const auto &v, auto *dst) {
    if (dst) std::copy(v.begin(), v.end(), dst);
  };
  put(m->x, x);
  put(m->y, y);
  put(m->z, z);
  put(m->face_ptr, face_ptr);
  put(m->face_nodes, face_nodes);
  put(m->owner, owner);
  put(m->neighbour, neighbour);
  put(m->patch_start, patch_start);
  put(m->patch_nfaces, patch_nfaces);
  if (patch_names && name_len > 0)
    for (size_t b = 0; b < m->patch_names.size(); ++b) {
      char *dst = patch_names + b * name_len;
      std::memset(dst, 0, name_len);
      std::strncpy(dst, m->patch_names[b].c_str(), name_len - 1);
    }
  return 0;
}

int fs_polymesh_destroy(fs_polymesh *m) {
  delete m;
  return 0;
}

} // extern "C"
