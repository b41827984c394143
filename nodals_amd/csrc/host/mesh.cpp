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

namespace {

// Cells of a face list (owner/neighbour/patches set): the tetrahedra as the
// sorted vertex set of their four faces with vertices 1 and 2 swapped for a
// negative Jacobian, and oppFace.
void finishTetMesh(TetMesh& M, const std::string& dir);

}  // namespace

TetMesh readPolyMesh(const std::string& dir) {
  TetMesh M;
  M.points = readPoints(dir + "/points");
  M.faces = readFaces(dir + "/faces");
  M.owner = readLabels(dir + "/owner");
  M.neighbour = readLabels(dir + "/neighbour");
  M.patches = readBoundary(dir + "/boundary");
  finishTetMesh(M, dir);
  return M;
}

namespace {

void finishTetMesh(TetMesh& M, const std::string& dir) {
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
}

template <class T>
void readRaw(std::ifstream& f, T* dst, std::size_t n, const std::string& what) {
  if (n == 0) return;
  f.read(reinterpret_cast<char*>(dst), (std::streamsize)(n * sizeof(T)));
  require((bool)f, "HXT1: short read (" + what + ")");
}

}  // namespace

bool isHxt1File(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  char magic[8] = {};
  f.read(magic, 8);
  return (bool)f && std::string(magic, 8) == "HXT1BIN1";
}

// HXT1 binary (turbulence/sst_gpu_v2/source/hybrid_hxt1/hxt1_mesh.hpp:
// 20-131, written by tools/prepare_hxt1.py): magic "HXT1BIN1", uint32
// version 1 and a reserved word, thirteen uint64 counts (nv nhex ntet
// nhexface ntetface nvel npressure ninterface nwallhex ninhex nintet nouthex
// nouttet), then points [nv][3] f64, hexes [nhex][8] i32, tets [ntet][4] i32,
// hex / tet velocity DOFs [nhex][14] / [ntet][8] i32, hex / tet pressure ids,
// hex / tet geometry [nhex][82] / [ntet][12] f64, interface records [][9],
// and the boundary records (cell, local face, bubble) [][3] i32 of the hex
// wall, hex and tet inlet, hex and tet outlet, with nothing after them.
// Tetrahedral meshes only.  Local face lf of a tet is the one opposite
// vertex lf.  Faces become an OpenFOAM-ordered face list: internal faces by
// (owner, neighbour), then the boundary faces as patches patch_0_0 (wall:
// every boundary face that is neither inlet nor outlet), patch_2_0 (inlet)
// and patch_1_0 (outlet), the H8 duct patch names.
TetMesh readHxt1(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  require((bool)f, "cannot open HXT1 mesh " + path);
  char magic[8] = {};
  f.read(magic, 8);
  require((bool)f && std::string(magic, 8) == "HXT1BIN1", "bad HXT1 magic in " + path);
  std::uint32_t version = 0, reserved = 0;
  readRaw(f, &version, 1, "version");
  readRaw(f, &reserved, 1, "reserved");
  require(version == 1, "unsupported HXT1 version in " + path);
  std::uint64_t h[13] = {};
  readRaw(f, h, 13, "header");
  const std::uint64_t nv = h[0], nhex = h[1], ntet = h[2], ninter = h[7], nwallhex = h[8], ninhex = h[9],
                      nintet = h[10], nouthex = h[11], nouttet = h[12];
  require(nhex == 0 && ninter == 0 && nwallhex == 0 && ninhex == 0 && nouthex == 0,
          "HXT1 mesh with hexahedra: only tetrahedral meshes are supported (" + path + ")");
  require(nv > 0 && ntet > 0 && nv < (1ull << 31) && ntet < (1ull << 31), "HXT1 sizes out of range in " + path);
  std::vector<double> pts((std::size_t)nv * 3);
  readRaw(f, pts.data(), pts.size(), "points");
  std::vector<std::int32_t> tets((std::size_t)ntet * 4), tvel((std::size_t)ntet * 8), tpres((std::size_t)ntet);
  readRaw(f, tets.data(), tets.size(), "tets");
  readRaw(f, tvel.data(), tvel.size(), "tet velocity dofs");
  readRaw(f, tpres.data(), tpres.size(), "tet pressure ids");
  std::vector<double> tgeom((std::size_t)ntet * 12);
  readRaw(f, tgeom.data(), tgeom.size(), "tet geometry");
  std::vector<std::int32_t> inTet((std::size_t)nintet * 3), outTet((std::size_t)nouttet * 3);
  readRaw(f, inTet.data(), inTet.size(), "inlet tet records");
  readRaw(f, outTet.data(), outTet.size(), "outlet tet records");
  char extra = 0;
  f.read(&extra, 1);
  require(f.gcount() == 0, "trailing bytes in HXT1 mesh " + path);
  for (std::int32_t v : tets) require(v >= 0 && (std::uint64_t)v < nv, "HXT1 tet vertex out of range");

  static const int kFace[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};
  struct Side {
    std::array<std::int32_t, 3> key;
    std::int32_t cell, lf;
  };
  std::vector<Side> sides;
  sides.reserve((std::size_t)ntet * 4);
  for (std::int32_t c = 0; c < (std::int32_t)ntet; ++c)
    for (int lf = 0; lf < 4; ++lf) {
      std::array<std::int32_t, 3> k;
      for (int j = 0; j < 3; ++j) k[(std::size_t)j] = tets[(std::size_t)c * 4 + (std::size_t)kFace[lf][j]];
      std::sort(k.begin(), k.end());
      sides.push_back({k, c, lf});
    }
  std::sort(sides.begin(), sides.end(), [](const Side& a, const Side& b) {
    return a.key != b.key ? a.key < b.key : a.cell < b.cell;
  });
  // boundary kind of (cell, local face): 1 inlet, 2 outlet
  std::vector<std::uint8_t> kind((std::size_t)ntet * 4, 0);
  for (int pass = 0; pass < 2; ++pass) {
    const std::vector<std::int32_t>& R = pass ? outTet : inTet;
    for (std::size_t r = 0; r < R.size(); r += 3) {
      const std::int32_t c = R[r], lf = R[r + 1];
      require(c >= 0 && (std::uint64_t)c < ntet && lf >= 0 && lf < 4, "HXT1 boundary record out of range");
      kind[(std::size_t)c * 4 + (std::size_t)lf] = (std::uint8_t)(pass + 1);
    }
  }
  struct IFace {
    std::int32_t own, nei;
    std::array<std::int32_t, 3> v;
  };
  std::vector<IFace> inner;
  std::vector<std::array<std::int32_t, 3>> bnd[3];
  std::vector<std::int32_t> bown[3];
  for (std::size_t i = 0; i < sides.size();) {
    std::size_t j = i + 1;
    while (j < sides.size() && sides[j].key == sides[i].key) ++j;
    require(j - i <= 2, "non-manifold face in HXT1 mesh " + path);
    const Side& a = sides[i];
    std::array<std::int32_t, 3> fv;
    for (int q = 0; q < 3; ++q) fv[(std::size_t)q] = tets[(std::size_t)a.cell * 4 + (std::size_t)kFace[a.lf][q]];
    if (j - i == 2) {
      require(kind[(std::size_t)a.cell * 4 + (std::size_t)a.lf] == 0 &&
                  kind[(std::size_t)sides[i + 1].cell * 4 + (std::size_t)sides[i + 1].lf] == 0,
              "HXT1 boundary record on an internal face");
      inner.push_back({a.cell, sides[i + 1].cell, fv});
    } else {
      const int k = kind[(std::size_t)a.cell * 4 + (std::size_t)a.lf];  // 0 wall, 1 inlet, 2 outlet
      bnd[k].push_back(fv);
      bown[k].push_back(a.cell);
    }
    i = j;
  }
  std::sort(inner.begin(), inner.end(),
            [](const IFace& a, const IFace& b) { return a.own != b.own ? a.own < b.own : a.nei < b.nei; });
  TetMesh M;
  M.points.resize((std::size_t)nv);
  for (std::size_t v = 0; v < (std::size_t)nv; ++v) M.points[v] = {pts[3 * v], pts[3 * v + 1], pts[3 * v + 2]};
  for (const IFace& q : inner) {
    M.faces.push_back(q.v);
    M.owner.push_back(q.own);
    M.neighbour.push_back(q.nei);
  }
  static const char* kNames[3] = {"patch_0_0", "patch_2_0", "patch_1_0"};
  for (int k = 0; k < 3; ++k) {
    Patch P;
    P.name = kNames[k];
    P.start = (std::int32_t)M.faces.size();
    P.count = (std::int32_t)bnd[k].size();
    for (std::size_t r = 0; r < bnd[k].size(); ++r) {
      M.faces.push_back(bnd[k][r]);
      M.owner.push_back(bown[k][r]);
    }
    M.patches.push_back(P);
  }
  require(M.patches[1].count == (std::int32_t)nintet && M.patches[2].count == (std::int32_t)nouttet,
          "HXT1 inlet/outlet records do not match the mesh's boundary faces");
  finishTetMesh(M, path);
  // the file's velocity DOFs in this mesh's (cell, basis) order
  const std::uint64_t nvel = h[5];
  for (std::size_t c = 0; c < (std::size_t)ntet; ++c)
    require(tpres[c] == (std::int32_t)c, "HXT1 pressure ids other than the cell index are not supported (" + path + ")");
  M.fileDof.resize((std::size_t)ntet);
  for (std::size_t c = 0; c < (std::size_t)ntet; ++c)
    for (int a = 0; a < 4; ++a) {
      int j = -1;
      for (int q = 0; q < 4; ++q)
        if (tets[c * 4 + (std::size_t)q] == M.tets[c][(std::size_t)a]) j = q;
      require(j >= 0, "HXT1 cell vertices changed while reading");
      const std::int32_t dv = tvel[c * 8 + (std::size_t)j], df = tvel[c * 8 + 4 + (std::size_t)j];
      require(dv >= 0 && (std::uint64_t)dv < nvel && df >= 0 && (std::uint64_t)df < nvel, "HXT1 velocity DOF out of range");
      M.fileDof[c][(std::size_t)a] = dv;
      M.fileDof[c][(std::size_t)a + 4] = df;
    }
  M.fileNvel = (std::int64_t)nvel;
  M.fileNvert = (std::int64_t)nv;
  return M;
}

TetMesh readMesh(const std::string& path) {
  if (isHxt1File(path)) return readHxt1(path);
  TetMesh M = readPolyMesh(path);
  M.fileDof.resize(M.tets.size());
  for (std::size_t c = 0; c < M.tets.size(); ++c)
    for (int a = 0; a < 4; ++a) {
      M.fileDof[c][(std::size_t)a] = M.tets[c][(std::size_t)a];
      M.fileDof[c][(std::size_t)a + 4] = M.nPoints() + M.oppFace[c][(std::size_t)a];
    }
  M.fileNvel = (std::int64_t)M.nPoints() + M.nFaces();
  M.fileNvert = M.nPoints();
  return M;
}

}  // namespace nodals
