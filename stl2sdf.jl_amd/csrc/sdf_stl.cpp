// STL import with the reference's vertex merge (host code, no device work).
//
// Reference: src/DataImport/ImportStlData.jl.  import_stl (:10-27) reads the
// file (MeshIO, or the package's own readers at :102-196: a file whose first
// five bytes are "solid" is ASCII, anything else binary) into one vertex per
// triangle corner, then process_stl_mesh (:198-309) merges vertices: each
// corner joins the first earlier unique vertex within 1e-10, searched first
// in its cell of a 1e-10 hash grid and then in the 26 neighbouring cells,
// in the order of Julia's `for dx = -1:1, dy = -1:1, dz = -1:1` loop (dz
// fastest), each cell's vertices in the order they became unique.
#include <ctype.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <unordered_map>
#include <vector>

namespace {

struct CellKey {
  int64_t x, y, z;
  bool operator==(const CellKey &o) const { return x == o.x && y == o.y && z == o.z; }
};
struct CellHash {
  size_t operator()(const CellKey &k) const {
    uint64_t h = (uint64_t)k.x * 0x9E3779B97F4A7C15ull;
    h ^= (uint64_t)k.y * 0xC2B2AE3D27D4EB4Full + (h << 6) + (h >> 2);
    h ^= (uint64_t)k.z * 0x165667B19E3779F9ull + (h << 6) + (h >> 2);
    return (size_t)h;
  }
};

// read_binary_stl (:127-161): 80-byte header, uint32 count, then per
// triangle a normal (skipped), three float32 vertices and a 2-byte attribute.
bool read_binary(FILE *f, std::vector<double> &v, std::string &err) {
  unsigned char hdr[80];
  uint32_t nt = 0;
  if (fread(hdr, 1, 80, f) != 80 || fread(&nt, 4, 1, f) != 1) {
    err = "binary STL: file shorter than its header";
    return false;
  }
  v.resize((size_t)nt * 9);
  unsigned char rec[50];
  for (uint32_t t = 0; t < nt; t++) {
    if (fread(rec, 1, 50, f) != 50) {
      err = "binary STL: file ends before triangle " + std::to_string(t + 1) + " of " +
            std::to_string(nt);
      return false;
    }
    float xyz[9];
    memcpy(xyz, rec + 12, sizeof(xyz));
    for (int c = 0; c < 9; c++) v[(size_t)t * 9 + c] = (double)xyz[c];
  }
  return true;
}

// read_ascii_stl (:163-196): every line starting with "vertex" (after
// stripping) counts; one with at least four fields also adds a vertex
// parsed as Float64, and when the count is then a multiple of three the
// last three vertices added form a triangle.  So a short "vertex" line
// shifts the grouping, as in the reference; a triangle that would reach
// before the first vertex is an error there (and here).
bool read_ascii(FILE *f, std::vector<double> &v, std::string &err) {
  std::vector<double> verts;
  int64_t count = 0;
  bool bad = false;
  std::string bad_number;
  std::string line;
  auto flush_line = [&]() {
    size_t a = 0;
    while (a < line.size() && isspace((unsigned char)line[a])) a++;
    if (line.compare(a, 6, "vertex") == 0) {
      count++;
      std::string copy(line.c_str() + a);
      char *tok[4];
      int n = 0;
      char *save = nullptr;
      for (char *t = strtok_r(&copy[0], " \t\r\n\v\f", &save); t && n < 4;
           t = strtok_r(nullptr, " \t\r\n\v\f", &save))
        tok[n++] = t;
      if (n >= 4) {
        for (int c = 1; c <= 3; c++) {
          char *end = nullptr;
          verts.push_back(strtod(tok[c], &end));
          if (end == tok[c] || *end != 0) bad_number = tok[c];  // parse(Float64, ...) throws
        }
        if (count % 3 == 0) {
          if (verts.size() < 9)
            bad = true;
          else
            v.insert(v.end(), verts.end() - 9, verts.end());
        }
      }
    }
    line.clear();
  };
  int ch;
  while ((ch = fgetc(f)) != EOF) {
    if (ch == '\n')
      flush_line();
    else
      line.push_back((char)ch);
  }
  flush_line();
  if (!bad_number.empty()) {
    err = "ASCII STL: cannot parse \"" + bad_number + "\" as a number";
    return false;
  }
  if (bad) {
    err = "ASCII STL: a triangle refers to a vertex before the first one";
    return false;
  }
  return true;
}

// process_stl_mesh (:198-309) over corners v (3 doubles each, three per
// triangle): unique vertices into X (3 per vertex), 1-based IEN.
void merge_vertices(const std::vector<double> &v, std::vector<double> &X,
                    std::vector<int64_t> &IEN) {
  const double tol = 1e-10;
  const double scale = floor(1.0 / tol);  // floor(Int, 1 / tol)
  const double edge = tol * scale;
  std::unordered_map<CellKey, std::vector<int64_t>, CellHash> cells;
  const size_t nc = v.size() / 3;
  IEN.resize(nc);
  X.clear();
  for (size_t q = 0; q < nc; q++) {
    const double *p = &v[3 * q];
    const CellKey g{(int64_t)floor(p[0] * scale), (int64_t)floor(p[1] * scale),
                    (int64_t)floor(p[2] * scale)};
    auto within = [&](int64_t u) {
      const double dx = p[0] - X[3 * u], dy = p[1] - X[3 * u + 1], dz = p[2] - X[3 * u + 2];
      return sqrt(dx * dx + dy * dy + dz * dz) < tol;
    };
    int64_t found = -1;
    auto search = [&](const CellKey &k) {
      auto it = cells.find(k);
      if (it == cells.end()) return;
      for (int64_t u : it->second)
        if (within(u)) {
          found = u;
          return;
        }
    };
    search(g);
    if (found < 0) {
      double pic[3];
      for (int d = 0; d < 3; d++) pic[d] = p[d] * scale - floor(p[d] * scale);
      for (int dx = -1; dx <= 1 && found < 0; dx++)
        for (int dy = -1; dy <= 1 && found < 0; dy++)
          for (int dz = -1; dz <= 1 && found < 0; dz++) {
            if (dx == 0 && dy == 0 && dz == 0) continue;
            const bool near = (dx == -1 && pic[0] < edge) || (dx == 1 && pic[0] > 1 - edge) ||
                              (dy == -1 && pic[1] < edge) || (dy == 1 && pic[1] > 1 - edge) ||
                              (dz == -1 && pic[2] < edge) || (dz == 1 && pic[2] > 1 - edge);
            if (near) search(CellKey{g.x + dx, g.y + dy, g.z + dz});
          }
    }
    if (found < 0) {
      found = (int64_t)(X.size() / 3);
      X.insert(X.end(), p, p + 3);
      cells[g].push_back(found);
    }
    IEN[q] = found + 1;
  }
}

bool import_file(const char *path, std::vector<double> &X, std::vector<int64_t> &IEN,
                 std::string &err) {
  FILE *f = fopen(path, "rb");
  if (!f) {
    err = std::string("cannot open ") + path;
    return false;
  }
  char head[5] = {0, 0, 0, 0, 0};
  const size_t got = fread(head, 1, 5, f);
  const bool ascii = got == 5 && memcmp(head, "solid", 5) == 0;  // detect_stl_type
  rewind(f);
  std::vector<double> v;
  const bool ok = ascii ? read_ascii(f, v, err) : read_binary(f, v, err);
  fclose(f);
  if (!ok) return false;
  merge_vertices(v, X, IEN);
  return true;
}

}  // namespace

namespace sdf {

// The mesh of an STL file as import_stl returns it: X (3 per vertex),
// 1-based IEN (3 per triangle).  False with err set on a read failure.
bool import_stl_file(const char *path, std::vector<double> &X, std::vector<int64_t> &IEN,
                     std::string &err) {
  return import_file(path, X, IEN, err);
}

}  // namespace sdf
