// OpenFOAM polyMesh reader for tetrahedral meshes.
// Input contract of gpu/serial_cuda/foam_mesh_g2.hpp:25-57.
#include "mesh.hpp"

#include <algorithm>
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string_view>

#include "error.hpp"

namespace nodals {

namespace {

// Tokenizer over a whole file: skips // and /* */ comments, returns the
// punctuation ( ) { } ; as single-character tokens, quoted strings without
// quotes, and maximal runs of other characters.
class Scanner {
 public:
  explicit Scanner(const std::string& path) : path_(path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) fail("cannot open " + path);
    std::ostringstream ss;
    ss << in.rdbuf();
    buf_ = ss.str();
  }

  bool next(std::string_view& tok) {
    const char* s = buf_.data();
    const std::size_t n = buf_.size();
    while (pos_ < n) {
      const char c = s[pos_];
      if (c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\f' || c == '\v') {
        ++pos_;
        continue;
      }
      if (c == '/' && pos_ + 1 < n && s[pos_ + 1] == '/') {
        while (pos_ < n && s[pos_] != '\n' && s[pos_] != '\r') ++pos_;
        continue;
      }
      if (c == '/' && pos_ + 1 < n && s[pos_ + 1] == '*') {
        const std::size_t e = buf_.find("*/", pos_ + 2);
        if (e == std::string::npos) fail("unterminated block comment in " + path_);
        pos_ = e + 2;
        continue;
      }
      if (c == '(' || c == ')' || c == '{' || c == '}' || c == ';') {
        tok = std::string_view(s + pos_, 1);
        ++pos_;
        return true;
      }
      if (c == '"') {
        const std::size_t b = pos_ + 1;
        std::size_t e = b;
        while (e < n && s[e] != '"') e += (s[e] == '\\') ? 2 : 1;
        if (e >= n) fail("unterminated quote in " + path_);
        tok = std::string_view(s + b, e - b);
        pos_ = e + 1;
        return true;
      }
      const std::size_t b = pos_;
      while (pos_ < n) {
        const char d = s[pos_];
        if (d == ' ' || d == '\t' || d == '\n' || d == '\r' || d == '\f' || d == '\v' || d == '(' ||
            d == ')' || d == '{' || d == '}' || d == ';')
          break;
        if (d == '/' && pos_ + 1 < n && (s[pos_ + 1] == '/' || s[pos_ + 1] == '*')) break;
        ++pos_;
      }
      tok = std::string_view(s + b, pos_ - b);
      return true;
    }
    return false;
  }

  std::string_view need() {
    std::string_view t;
    if (!next(t)) fail("unexpected end of file in " + path_);
    return t;
  }

  void expect(char c) {
    std::string_view t = need();
    if (t.size() != 1 || t[0] != c) fail(std::string("expected '") + c + "' in " + path_);
  }

  static bool toInt(std::string_view t, long long& v) {
    if (t.empty()) return false;
    std::size_t i = 0;
    bool neg = false;
    if (t[0] == '-' || t[0] == '+') {
      neg = t[0] == '-';
      i = 1;
      if (t.size() == 1) return false;
    }
    long long x = 0;
    for (; i < t.size(); ++i) {
      if (t[i] < '0' || t[i] > '9') return false;
      x = x * 10 + (t[i] - '0');
      if (x > (long long)INT_MAX + 1) return false;
    }
    v = neg ? -x : x;
    return true;
  }

  long long needInt(const char* what) {
    long long v = 0;
    if (!toInt(need(), v)) fail(std::string("invalid ") + what + " in " + path_);
    return v;
  }

  double needDouble(const char* what) {
    std::string_view t = need();
    std::string s(t);
    char* e = nullptr;
    errno = 0;
    const double v = std::strtod(s.c_str(), &e);
    if (errno || !e || *e != '\0' || s.empty()) fail(std::string("invalid ") + what + " in " + path_);
    return v;
  }

  // Position after "<count> (": the first non-negative integer token that is
  // directly followed by an opening parenthesis.
  int listStart() {
    std::string_view t;
    while (next(t)) {
      long long v = 0;
      if (!toInt(t, v) || v < 0 || v > INT_MAX) continue;
      std::string_view u;
      if (!next(u)) break;
      if (u == "(") return (int)v;
    }
    fail("cannot find OpenFOAM list count in " + path_);
  }

  const std::string& path() const { return path_; }

 private:
  std::string path_;
  std::string buf_;
  std::size_t pos_ = 0;
};

std::vector<std::array<double, 3>> readPoints(const std::string& p) {
  Scanner s(p);
  const int n = s.listStart();
  std::vector<std::array<double, 3>> out((std::size_t)n);
  for (int i = 0; i < n; ++i) {
    s.expect('(');
    out[(std::size_t)i] = {s.needDouble("x"), s.needDouble("y"), s.needDouble("z")};
    s.expect(')');
  }
  s.expect(')');
  return out;
}

std::vector<std::array<std::int32_t, 3>> readFaces(const std::string& p) {
  Scanner s(p);
  const int n = s.listStart();
  std::vector<std::array<std::int32_t, 3>> out((std::size_t)n);
  for (int i = 0; i < n; ++i) {
    if (s.needInt("face arity") != 3) fail("non-triangular face in " + p);
    s.expect('(');
    for (int j = 0; j < 3; ++j) {
      const long long v = s.needInt("face vertex");
      if (v < 0 || v > INT_MAX) fail("face vertex out of range in " + p);
      out[(std::size_t)i][(std::size_t)j] = (std::int32_t)v;
    }
    s.expect(')');
  }
  s.expect(')');
  return out;
}

std::vector<std::int32_t> readLabels(const std::string& p) {
  Scanner s(p);
  const int n = s.listStart();
  std::vector<std::int32_t> out((std::size_t)n);
  for (int i = 0; i < n; ++i) {
    const long long v = s.needInt("label");
    if (v < INT_MIN || v > INT_MAX) fail("label out of range in " + p);
    out[(std::size_t)i] = (std::int32_t)v;
  }
  s.expect(')');
  return out;
}

std::vector<Patch> readBoundary(const std::string& p) {
  Scanner s(p);
  const int n = s.listStart();
  std::vector<Patch> out;
  for (int i = 0; i < n; ++i) {
    std::string name(s.need());
    if (name.empty() || name == ")") fail("missing patch name in " + p);
    s.expect('{');
    long long nf = -1, sf = -1;
    int depth = 1;
    while (depth) {
      std::string_view t = s.need();
      if (t == "{") {
        ++depth;
        continue;
      }
      if (t == "}") {
        --depth;
        continue;
      }
      if (depth != 1) continue;
      if (t == "nFaces") {
        nf = s.needInt("nFaces");
        s.expect(';');
      } else if (t == "startFace") {
        sf = s.needInt("startFace");
        s.expect(';');
      }
    }
    if (nf < 0 || sf < 0) fail("invalid patch range " + name);
    out.push_back({name, (std::int32_t)sf, (std::int32_t)nf});
  }
  s.expect(')');
  return out;
}

}  // namespace

double det3(const double J[3][3]) {
  return J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
         J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
         J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
}

void inv3(const double J[3][3], double I[3][3]) {
  const double d = det3(J);
  if (std::abs(d) < 1e-30) fail("singular tetrahedron");
  I[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / d;
  I[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / d;
  I[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / d;
  I[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / d;
  I[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / d;
  I[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / d;
  I[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / d;
  I[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / d;
  I[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / d;
}

int TetMesh::patchIndex(const std::string& name) const {
  for (std::size_t i = 0; i < patches.size(); ++i)
    if (patches[i].name == name) return (int)i;
  return -1;
}

TetMesh readPolyMesh(const std::string& dir) {
  TetMesh M;
  M.points = readPoints(dir + "/points");
  M.faces = readFaces(dir + "/faces");
  M.owner = readLabels(dir + "/owner");
  M.neighbour = readLabels(dir + "/neighbour");
  M.patches = readBoundary(dir + "/boundary");
  require(M.owner.size() == M.faces.size(), "owner size != faces size in " + dir);
  require(M.neighbour.size() <= M.faces.size(), "neighbour longer than faces in " + dir);
  M.facePatch.assign(M.faces.size(), -1);
  for (std::size_t p = 0; p < M.patches.size(); ++p) {
    const Patch& P = M.patches[p];
    require(P.start >= 0 && (std::size_t)P.start + (std::size_t)P.count <= M.faces.size(),
            "patch range out of bounds: " + P.name);
    for (int f = P.start; f < P.start + P.count; ++f) M.facePatch[(std::size_t)f] = (std::int32_t)p;
  }
  int maxc = -1;
  for (auto x : M.owner) maxc = std::max(maxc, (int)x);
  for (auto x : M.neighbour) maxc = std::max(maxc, (int)x);
  const int nc = maxc + 1;
  require(nc > 0, "mesh has no cells: " + dir);
  for (auto x : M.owner) require(x >= 0, "negative owner label");
  for (auto x : M.neighbour) require(x >= 0, "negative neighbour label");
  for (const auto& f : M.faces)
    for (auto v : f) require(v < (std::int32_t)M.points.size(), "face vertex beyond points");

  // Faces of each cell: owner faces in face order, then neighbour faces in
  // face order (the cell-face list order of the reference reader).
  std::vector<std::int32_t> cnt((std::size_t)nc + 1, 0);
  for (auto c : M.owner) ++cnt[(std::size_t)c + 1];
  for (auto c : M.neighbour) ++cnt[(std::size_t)c + 1];
  for (int c = 0; c < nc; ++c) {
    require(cnt[(std::size_t)c + 1] == 4, "non-tet cell " + std::to_string(c));
    cnt[(std::size_t)c + 1] += cnt[(std::size_t)c];
  }
  std::vector<std::int32_t> cellFaces((std::size_t)cnt[(std::size_t)nc]);
  std::vector<std::int32_t> fill(cnt.begin(), cnt.end() - 1);
  for (std::size_t f = 0; f < M.owner.size(); ++f)
    cellFaces[(std::size_t)fill[(std::size_t)M.owner[f]]++] = (std::int32_t)f;
  for (std::size_t f = 0; f < M.neighbour.size(); ++f)
    cellFaces[(std::size_t)fill[(std::size_t)M.neighbour[f]]++] = (std::int32_t)f;

  M.tets.resize((std::size_t)nc);
  M.oppFace.resize((std::size_t)nc);
  for (int c = 0; c < nc; ++c) {
    const std::int32_t* cf = &cellFaces[(std::size_t)cnt[(std::size_t)c]];
    std::int32_t v[12];
    for (int k = 0; k < 4; ++k)
      for (int j = 0; j < 3; ++j) v[3 * k + j] = M.faces[(std::size_t)cf[k]][(std::size_t)j];
    std::sort(v, v + 12);
    std::int32_t* e = std::unique(v, v + 12);
    require(e - v == 4, "tet vertex count mismatch at cell " + std::to_string(c));
    std::array<std::int32_t, 4> t = {v[0], v[1], v[2], v[3]};
    const auto& X0 = M.points[(std::size_t)t[0]];
    const auto& X1 = M.points[(std::size_t)t[1]];
    const auto& X2 = M.points[(std::size_t)t[2]];
    const auto& X3 = M.points[(std::size_t)t[3]];
    const double J[3][3] = {{X1[0] - X0[0], X2[0] - X0[0], X3[0] - X0[0]},
                            {X1[1] - X0[1], X2[1] - X0[1], X3[1] - X0[1]},
                            {X1[2] - X0[2], X2[2] - X0[2], X3[2] - X0[2]}};
    if (det3(J) < 0) std::swap(t[1], t[2]);
    M.tets[(std::size_t)c] = t;
    for (int i = 0; i < 4; ++i) {
      int found = -1;
      for (int k = 0; k < 4; ++k) {
        const auto& F = M.faces[(std::size_t)cf[k]];
        if (F[0] != t[(std::size_t)i] && F[1] != t[(std::size_t)i] && F[2] != t[(std::size_t)i]) {
          require(found < 0, "multiple opposite faces");
          found = cf[k];
        }
      }
      require(found >= 0, "missing opposite face");
      M.oppFace[(std::size_t)c][(std::size_t)i] = found;
    }
  }
  return M;
}

}  // namespace nodals
