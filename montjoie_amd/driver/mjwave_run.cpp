// mjwave_run -- runs a Montjoie explicit wave deck (.ini data file) on the
// MI355X through the C ABI of mjwave.h.
//
//   mjwave_run deck.ini [--precision fp64|fp32|fp32_g64] [--device N]
//                       [--dump-mesh FILE]
//
// Replaces, for the explicit leapfrog path, Montjoie's driver chain
// (Program/Beginner/acous3D.cc -> HyperbolicProblem::RunAll,
// Instationary/VarInstationary.cxx:1629-1831): read the data file
// (Output/OutputHarmonic.cxx keyword parsing), build the mesh (REGULAR,
// Mesh/Mesh2D.cxx:5095-5135, Mesh/Mesh3D.cxx:10468-10514, or a Medit .mesh
// file, Mesh/Mesh3D.cxx:1027-1060,3734-3770), number the Q_r nodes
// (Mesh/NumberMap.cxx), apply the media, Dirichlet references, Dirac source
// with a Ricker pulse and SismoPoint receivers, run RunSecondOrderScheme
// (ModifiedEquation.cxx:16-53) and write the point seismograms in
// Montjoie's text format (one row per output time: t then the values, %.15e,
// tab-separated; Output/GridInterpolation.cxx:3150-3152 sets precision 15).
//
// Keywords read: TypeElement (QUADRANGLE_LOBATTO, HEXAHEDRON_LOBATTO),
// TypeEquation (ACOUSTIC, TIME_ELASTIC), MeshPath, FileMesh, TypeMesh,
// OrderDiscretization, ConditionReference (DIRICHLET), MateriauDielec
// (ISOTROPE, elastic also YOUNG_POISSON; damping sigma), PhysicalMedia
// (rho / mu / sigma, SINUS or MESH / SAME_MESH data files on the
// computational mesh, acoustic), TypeSource (SRC_DIRAC, SRC_VOLUME GAUSSIAN / UNIFORM), OriginePhase,
// Polarization, Frequency,
// TemporalSource (RICKER, DERIVATIVE_RICKER, MODIFIED_RICKER, GAUSSIAN, HARMONIC,
// SINUS_GAUSSIAN), TimeStep (a value or AUTO), TimeInterval, OrderTimeScheme (2 or 4
// LEAP_FROG, 2 LEAP_FROG_PML), SismoPoint, FileOutputPoint, ParametersOutputPoint, SismoCircle,
// FileOutputCircle, ParametersOutputCircle, SismoPlane, FileOutputPlane, ParametersOutputPlane,
// SismoLine, FileOutputLine, ParametersOutputLine, OutputFormat (BINARY or ASCII, DOUBLE or FLOAT),
// DisplayStress, SaveReprise, PathReprise.  Other keywords are reported and ignored; a value
// this path does not implement (another scheme, source or condition) stops
// the run with a message, as Montjoie's own parser does (cout + abort).
//
// Elastic decks (quadrilaterals or hexahedra): Montjoie's rows hold the
// displacement of every point (d values per point, point-major) followed by
// the stress tensor of every point (d*d values, row-major; DisplayStress =
// YES, its default), or the raw gradient of u (DisplayStress = NO,
// ModifyOutputUnknown returns early, Elliptic/Elastic/VarElastic.cxx:
// 257-258); this driver writes both.
//
// --dump-mesh writes the assembled mesh (sizes, node positions, element ->
// node map in local order q = i + n j + n^2 k, Dirichlet flags, media) as raw
// little-endian arrays and exits without touching the GPU.
#include <algorithm>
#include <cctype>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "mjwave.h"

namespace {

[[noreturn]] void die(const std::string &m) {
  std::fprintf(stderr, "mjwave_run: %s\n", m.c_str());
  std::exit(1);
}

void check(int rc, const mjw_problem *p, const char *what) {
  if (rc != MJW_OK) die(std::string(what) + ": " + mjw_last_error(p));
}

std::vector<std::string> split(const std::string &s) {
  std::istringstream is(s);
  std::vector<std::string> v;
  std::string w;
  while (is >> w) v.push_back(w);
  return v;
}

double num(const std::string &s) {
  char *end = nullptr;
  const double v = std::strtod(s.c_str(), &end);
  if (end == s.c_str()) die("not a number: " + s);
  return v;
}

// ----------------------------------------------------------------- deck
struct Deck {
  std::multimap<std::string, std::vector<std::string>> kv;
  std::string dir;
  const std::vector<std::string> *get(const std::string &k) const {
    auto it = kv.find(k);
    return it == kv.end() ? nullptr : &it->second;
  }
  std::vector<std::vector<std::string>> all(const std::string &k) const {
    std::vector<std::vector<std::string>> out;
    auto r = kv.equal_range(k);
    for (auto it = r.first; it != r.second; ++it) out.push_back(it->second);
    return out;
  }
  const std::vector<std::string> &need(const std::string &k) const {
    const auto *v = get(k);
    if (!v || v->empty()) die("keyword " + k + " is missing");
    return *v;
  }
};

Deck read_deck(const std::string &path) {
  std::ifstream in(path);
  if (!in) die("cannot open " + path);
  Deck d;
  const size_t sl = path.find_last_of('/');
  d.dir = sl == std::string::npos ? "." : path.substr(0, sl);
  std::string line;
  while (std::getline(in, line)) {
    const size_t h = line.find('#');
    if (h != std::string::npos) line = line.substr(0, h);
    const size_t eq = line.find('=');
    if (eq == std::string::npos) continue;
    std::vector<std::string> key = split(line.substr(0, eq));
    if (key.size() != 1) continue;
    d.kv.emplace(key[0], split(line.substr(eq + 1)));
  }
  return d;
}

// ----------------------------------------------------------------- mesh
// vertices + Q1 elements (quadrilaterals / hexahedra in Medit corner order)
// + boundary faces (edges / quadrilaterals) with references
struct Q1Mesh {
  int dim = 3;
  std::vector<double> vert;               // nv * dim
  std::vector<int64_t> elem;              // E * 2^dim vertex ids
  std::vector<int> elem_ref;
  std::vector<std::vector<int64_t>> bnd;  // boundary faces (vertex ids)
  std::vector<int> bnd_ref;
};

Q1Mesh regular_mesh(int dim, const std::vector<std::string> &p) {
  // FileMesh = REGULAR nbx [xmin xmax ymin ymax [zmin zmax]] [ref_domain ref_bnd...]
  Q1Mesh m;
  m.dim = dim;
  int nbx = (int)num(p.at(1));
  if (nbx <= 1) nbx = 2;
  double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
  const size_t nb = p.size();
  const size_t need_box = dim == 3 ? 8 : 6;
  if (nb >= need_box)
    for (int a = 0; a < dim; ++a) {
      lo[a] = num(p[2 + 2 * a]);
      hi[a] = num(p[3 + 2 * a]);
    }
  int npt[3] = {nbx, nbx, nbx};
  for (int a = 1; a < dim; ++a) {
    const double ratio = (hi[a] - lo[a]) / (hi[0] - lo[0]) * nbx;
    // 2-D rounds up (Mesh2D.cxx:5109-5110), 3-D to the nearest (Mesh3D.cxx:10486-10487)
    npt[a] = dim == 2 ? (int)std::ceil(ratio) : (int)std::lround(ratio);
    if (npt[a] <= 1) npt[a] = 2;
  }
  int ref_dom = 1;
  std::vector<int> rb(2 * dim, 3);
  const size_t first_ref = need_box;
  if (nb >= first_ref + 2) {
    if (dim == 2) ref_dom = (int)num(p[first_ref]);  // 3-D always uses domain 1
    const int r0 = (int)num(p[first_ref + 1]);
    std::fill(rb.begin(), rb.end(), r0);
    if (nb >= first_ref + 1 + 2 * dim)
      for (int k = 1; k < 2 * dim; ++k) rb[k] = (int)num(p[first_ref + 1 + k]);
  }
  const int nx = npt[0] - 1, ny = npt[1] - 1, nz = dim == 3 ? npt[2] - 1 : 0;
  auto coord = [&](int a, int i) {
    return i == npt[a] - 1 ? hi[a] : lo[a] + i * (hi[a] - lo[a]) / (npt[a] - 1);
  };
  const int pz = dim == 3 ? npt[2] : 1;
  for (int k = 0; k < pz; ++k)
    for (int j = 0; j < npt[1]; ++j)
      for (int i = 0; i < npt[0]; ++i) {
        m.vert.push_back(coord(0, i));
        m.vert.push_back(coord(1, j));
        if (dim == 3) m.vert.push_back(coord(2, k));
      }
  auto vid = [&](int i, int j, int k) {
    return (int64_t)i + (int64_t)npt[0] * (j + (int64_t)npt[1] * k);
  };
  if (dim == 2) {
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        for (int64_t v : {vid(i, j, 0), vid(i + 1, j, 0), vid(i + 1, j + 1, 0), vid(i, j + 1, 0)})
          m.elem.push_back(v);
        m.elem_ref.push_back(ref_dom);
      }
    // boundary references: y = ymin, x = xmax, y = ymax, x = xmin (Mesh2D.cxx:4856-4873)
    for (int i = 0; i < nx; ++i) {
      m.bnd.push_back({vid(i, 0, 0), vid(i + 1, 0, 0)});
      m.bnd_ref.push_back(rb[0]);
      m.bnd.push_back({vid(i, ny, 0), vid(i + 1, ny, 0)});
      m.bnd_ref.push_back(rb[2]);
    }
    for (int j = 0; j < ny; ++j) {
      m.bnd.push_back({vid(nx, j, 0), vid(nx, j + 1, 0)});
      m.bnd_ref.push_back(rb[1]);
      m.bnd.push_back({vid(0, j, 0), vid(0, j + 1, 0)});
      m.bnd_ref.push_back(rb[3]);
    }
  } else {
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          for (int64_t v : {vid(i, j, k), vid(i + 1, j, k), vid(i + 1, j + 1, k), vid(i, j + 1, k),
                            vid(i, j, k + 1), vid(i + 1, j, k + 1), vid(i + 1, j + 1, k + 1),
                            vid(i, j + 1, k + 1)})
            m.elem.push_back(v);
          m.elem_ref.push_back(1);
        }
    // boundary references: x = xmin, y = ymin, z = zmin, z = zmax, y = ymax,
    // x = xmax (Mesh3D.cxx:8700-8840)
    auto quad = [&](int64_t a, int64_t b, int64_t c, int64_t d, int ref) {
      m.bnd.push_back({a, b, c, d});
      m.bnd_ref.push_back(ref);
    };
    for (int b = 0; b < std::max(ny, nz); ++b)
      for (int a = 0; a < std::max(nx, ny); ++a) {
        if (a < ny && b < nz) {
          quad(vid(0, a, b), vid(0, a + 1, b), vid(0, a + 1, b + 1), vid(0, a, b + 1), rb[0]);
          quad(vid(nx, a, b), vid(nx, a + 1, b), vid(nx, a + 1, b + 1), vid(nx, a, b + 1), rb[5]);
        }
        if (a < nx && b < nz) {
          quad(vid(a, 0, b), vid(a + 1, 0, b), vid(a + 1, 0, b + 1), vid(a, 0, b + 1), rb[1]);
          quad(vid(a, ny, b), vid(a + 1, ny, b), vid(a + 1, ny, b + 1), vid(a, ny, b + 1), rb[4]);
        }
      }
    for (int b = 0; b < ny; ++b)
      for (int a = 0; a < nx; ++a) {
        quad(vid(a, b, 0), vid(a + 1, b, 0), vid(a + 1, b + 1, 0), vid(a, b + 1, 0), rb[2]);
        quad(vid(a, b, nz), vid(a + 1, b, nz), vid(a + 1, b + 1, nz), vid(a, b + 1, nz), rb[3]);
      }
  }
  return m;
}

Q1Mesh medit_mesh(int dim, const std::string &path) {
  std::ifstream in(path);
  if (!in) die("cannot open mesh file " + path);
  Q1Mesh m;
  m.dim = dim;
  std::string w;
  auto read_int = [&]() -> int64_t {
    int64_t v;
    if (!(in >> v)) die("bad Medit file " + path);
    return v;
  };
  while (in >> w) {
    if (w == "Dimension") {
      const int d = (int)read_int();
      if (d != dim) die("mesh dimension does not match TypeElement");
    } else if (w == "Vertices") {
      const int64_t n = read_int();
      m.vert.resize(n * dim);
      for (int64_t i = 0; i < n; ++i) {
        for (int a = 0; a < dim; ++a) in >> m.vert[i * dim + a];
        read_int();  // reference
      }
    } else if (w == "Hexahedra" || (w == "Quadrilaterals" && dim == 2)) {
      const int64_t n = read_int();
      const int nv = dim == 3 ? 8 : 4;
      for (int64_t e = 0; e < n; ++e) {
        for (int c = 0; c < nv; ++c) m.elem.push_back(read_int() - 1);
        m.elem_ref.push_back((int)read_int());
      }
    } else if ((w == "Quadrilaterals" && dim == 3) || (w == "Edges" && dim == 2)) {
      const int64_t n = read_int();
      const int nv = dim == 3 ? 4 : 2;
      for (int64_t f = 0; f < n; ++f) {
        std::vector<int64_t> v(nv);
        for (int c = 0; c < nv; ++c) v[c] = read_int() - 1;
        m.bnd.push_back(v);
        m.bnd_ref.push_back((int)read_int());
      }
    } else if (w == "Triangles" || w == "Tetrahedra" || w == "Prisms" || w == "Pyramids") {
      die("Medit section " + w + ": only quadrilateral / hexahedral meshes are supported");
    } else if (w == "End") {
      break;
    }
  }
  if (m.elem.empty()) die("no elements in " + path);
  return m;
}

// ------------------------------------------------------------ PML layers
// AddPML = YES <axes> thickness layers|AUTO (PmlRegionParameter_Base::
// SetParameters, Mesh/MeshBoundaries.cxx:32-102): X, Y, Z extrude both sides of
// the box along that axis, X- / X+ (Y-, Z-, ...) one side.
struct PmlSpec {
  bool on = false;
  int side[3] = {0, 0, 0};  // bit 0: negative side, bit 1: positive side
  double thickness = 0.0;
  int layers = 0;           // 0: AUTO
};

PmlSpec parse_pml(const Deck &d, int dim) {
  PmlSpec s;
  const auto lines = d.all("AddPML");
  if (lines.empty()) return s;
  if (lines.size() > 1) die("AddPML: one PML region per deck");
  const auto &v = lines[0];
  if (v.size() < 4) die("AddPML needs YES axes thickness layers");
  if (v[0] != "YES") die("AddPML: only AddPML = YES <axes> thickness layers (no reference-based layers)");
  const std::string &ax = v[1];
  if (ax.find('R') != std::string::npos) die("AddPML: radial layers are not supported");
  for (int a = 0; a < dim; ++a) {
    const char c = "XYZ"[a];
    if (ax.find(c) == std::string::npos) continue;
    if (ax.find(std::string(1, c) + "-") != std::string::npos) s.side[a] = 1;
    else if (ax.find(std::string(1, c) + "+") != std::string::npos) s.side[a] = 2;
    else s.side[a] = 3;
  }
  s.thickness = num(v[2]);
  if (!(s.thickness > 0)) die("AddPML: the thickness must be positive");
  s.layers = v[3] == "AUTO" ? 0 : (int)num(v[3]);
  if (v[3] != "AUTO" && s.layers < 1) die("AddPML: the number of layers must be positive");
  s.on = s.side[0] || s.side[1] || s.side[2];
  return s;
}

// vertex-set key of a face (sorted vertex ids)
std::vector<int64_t> face_key(std::vector<int64_t> f) {
  std::sort(f.begin(), f.end());
  return f;
}

// faces of a Q1 element in Medit corner order (quadrilateral: 4 edges,
// hexahedron: 6 faces)
std::vector<std::vector<int64_t>> elem_faces(const Q1Mesh &m, int64_t e) {
  const int nc = 1 << m.dim;
  const int64_t *c = &m.elem[(size_t)e * nc];
  if (m.dim == 2) return {{c[0], c[1]}, {c[1], c[2]}, {c[2], c[3]}, {c[3], c[0]}};
  return {{c[0], c[1], c[2], c[3]}, {c[4], c[5], c[6], c[7]}, {c[0], c[1], c[5], c[4]},
          {c[1], c[2], c[6], c[5]}, {c[2], c[3], c[7], c[6]}, {c[3], c[0], c[4], c[7]}};
}

// Layers along axis a beyond the plane x_a = pos (delta < 0: below it), the
// way Mesh<>::ExtrudeCoordinate does (Mesh/Mesh2D.cxx:4901-5073,
// Mesh/Mesh3D.cxx:9414-9660): the vertices within 1e-10 of the plane (R_N::
// threshold, Mesh/NumberMesh.cxx:1621,2727) are copied nl times with x_a
// moved to the layer planes; every referenced boundary face lying on the plane
// grows one element per layer, of the medium (reference) of the element it
// bounds, and is itself moved to the outer plane with its reference; the
// rim of those faces gets lateral faces, referenced like the non-extruded
// boundary face they continue (the extruded face's reference otherwise).
void extrude(Q1Mesh &m, int a, int nl, double pos, double delta) {
  const int dim = m.dim, nc = 1 << dim;
  const int64_t nv0 = (int64_t)(m.vert.size() / dim);
  std::vector<int64_t> to(nv0, -1);
  int64_t nn = 0;
  for (int64_t i = 0; i < nv0; ++i)
    if (std::fabs(pos - m.vert[(size_t)i * dim + a]) <= 1e-10) to[i] = nv0 + nn++;
  if (!nn) return;
  m.vert.resize((size_t)(nv0 + nn * nl) * dim);
  for (int k = 0; k < nl; ++k)
    for (int64_t i = 0; i < nv0; ++i)
      if (to[i] >= 0) {
        double *x = &m.vert[(size_t)(to[i] + k * nn) * dim];
        const double *x0 = &m.vert[(size_t)i * dim];
        for (int b = 0; b < dim; ++b) x[b] = x0[b];
        // 2-D: pos + (k+1) delta / nl; 3-D: x_a + delta (k+1) / nl
        x[a] = dim == 2 ? pos + (double)(k + 1) * delta / nl : x0[a] + delta * (double)(k + 1) / nl;
      }
  auto layer = [&](int64_t v, int k) { return k == 0 ? v : to[v] + (int64_t)(k - 1) * nn; };
  const size_t nb0 = m.bnd.size();
  std::vector<char> top(nb0, 0);
  std::map<std::vector<int64_t>, int64_t> owner;  // top face -> element it bounds
  for (size_t f = 0; f < nb0; ++f) {
    bool on = true;
    for (int64_t v : m.bnd[f]) on = on && to[v] >= 0;
    top[f] = on;
    if (on) owner[face_key(m.bnd[f])] = -1;
  }
  const int64_t E0 = (int64_t)m.elem_ref.size();
  for (int64_t e = 0; e < E0; ++e)
    for (const auto &fc : elem_faces(m, e)) {
      auto it = owner.find(face_key(fc));
      if (it != owner.end()) it->second = e;
    }
  // rim of the extruded patch: vertices (2-D) / edges (3-D) on one top face only
  std::map<std::vector<int64_t>, std::pair<int, size_t>> rim;  // key -> (count, face)
  for (size_t f = 0; f < nb0; ++f) {
    if (!top[f]) continue;
    const auto &fv = m.bnd[f];
    const int n = (int)fv.size();
    if (dim == 2) {
      for (int64_t v : fv) {
        auto &r = rim[{v}];
        r.first++;
        r.second = f;
      }
    } else {
      for (int j = 0; j < n; ++j) {
        auto &r = rim[face_key({fv[j], fv[(j + 1) % n]})];
        r.first++;
        r.second = f;
      }
    }
  }
  std::vector<std::vector<int64_t>> new_bnd;
  std::vector<int> new_ref;
  for (const auto &kv : rim) {
    if (kv.second.first != 1) continue;
    const size_t f = kv.second.second;
    int ref = m.bnd_ref[f];
    if (dim == 2) {
      // another boundary edge at this vertex running along axis a
      // (GetPointWithOtherReference, Mesh2D.cxx:5043-5058)
      const int64_t v = kv.first[0];
      for (size_t g = 0; g < nb0; ++g) {
        if (top[g] || m.bnd_ref[g] == ref) continue;
        const auto &ev = m.bnd[g];
        if (ev[0] != v && ev[1] != v) continue;
        const int64_t w = ev[0] == v ? ev[1] : ev[0];
        if (std::fabs(m.vert[(size_t)v * 2 + (1 - a)] - m.vert[(size_t)w * 2 + (1 - a)]) <= 1e-9 &&
            m.bnd_ref[g] != 0)
          ref = m.bnd_ref[g];
      }
      for (int k = 0; k < nl; ++k) {
        new_bnd.push_back({layer(v, k), layer(v, k + 1)});
        new_ref.push_back(ref);
      }
    } else {
      // a non-extruded boundary face sharing the edge (Mesh3D.cxx:9606-9619)
      const int64_t v0 = kv.first[0], v1 = kv.first[1];
      for (size_t g = 0; g < nb0; ++g) {
        if (top[g]) continue;
        const auto &fv = m.bnd[g];
        const int n = (int)fv.size();
        bool has = false;
        for (int j = 0; j < n && !has; ++j)
          has = face_key({fv[j], fv[(j + 1) % n]}) == std::vector<int64_t>{v0, v1};
        if (has && m.bnd_ref[g] > 0) ref = m.bnd_ref[g];
      }
      // the edge in the order the top face runs it
      const auto &fv = m.bnd[f];
      const int n = (int)fv.size();
      int64_t p0 = v0, p1 = v1;
      for (int j = 0; j < n; ++j)
        if (face_key({fv[j], fv[(j + 1) % n]}) == std::vector<int64_t>{v0, v1}) {
          p0 = fv[j];
          p1 = fv[(j + 1) % n];
        }
      for (int k = 0; k < nl; ++k) {
        new_bnd.push_back({layer(p0, k), layer(p1, k), layer(p1, k + 1), layer(p0, k + 1)});
        new_ref.push_back(ref);
      }
    }
  }
  // the new elements, positively oriented
  const double *X = nullptr;
  auto sub = [&](int64_t i, int64_t j, double *o) {
    for (int b = 0; b < dim; ++b) o[b] = X[(size_t)j * dim + b] - X[(size_t)i * dim + b];
  };
  for (size_t f = 0; f < nb0; ++f) {
    if (!top[f]) continue;
    const int64_t owner_e = owner[face_key(m.bnd[f])];
    if (owner_e < 0) die("AddPML: a boundary face on the extruded plane bounds no element");
    std::vector<int64_t> fv = m.bnd[f];
    for (int k = 0; k < nl; ++k) {
      std::vector<int64_t> el(nc);
      for (int j = 0; j < nc / 2; ++j) {
        el[j] = layer(fv[j], k);
        el[dim == 2 ? 3 - j : j + 4] = layer(fv[j], k + 1);
      }
      X = m.vert.data();
      double u[3], w[3], z[3] = {0, 0, 0};
      sub(el[0], el[1], u);
      sub(el[0], el[3], w);
      double vol;
      if (dim == 2) {
        vol = u[0] * w[1] - u[1] * w[0];
      } else {
        sub(el[0], el[4], z);
        vol = (u[1] * w[2] - u[2] * w[1]) * z[0] + (u[2] * w[0] - u[0] * w[2]) * z[1] +
              (u[0] * w[1] - u[1] * w[0]) * z[2];
      }
      if (vol < 0) {
        if (dim == 2) {
          std::swap(el[0], el[1]);
          std::swap(el[2], el[3]);
        } else {
          std::swap(el[1], el[3]);
          std::swap(el[5], el[7]);
        }
      }
      m.elem.insert(m.elem.end(), el.begin(), el.end());
      m.elem_ref.push_back(m.elem_ref[owner_e]);
    }
    for (auto &v : m.bnd[f]) v = layer(v, nl);
  }
  m.bnd.insert(m.bnd.end(), new_bnd.begin(), new_bnd.end());
  m.bnd_ref.insert(m.bnd_ref.end(), new_ref.begin(), new_ref.end());
}

// PML region of a Q1 mesh: the layers extruded axis by axis, x- then x+,
// y-, y+ (z-, z+) (PmlRegionParameter<>::AddPMLElements, Mesh/
// Mesh2DBoundaries.cxx:104-192, Mesh3DBoundaries.cxx:59-140); returns the
// kind of every element: bit 2a set below the box along axis a, bit 2a+1
// above it, from its centre and the box of the mesh before extrusion (0 for
// the physical elements).  layers = AUTO: ceil(thickness / mean length of the
// boundary edges on the box), at least 1 (AddPML, MeshBoundaries.cxx:108-118,
// GetAverageLengthOnBoundary, Mesh2DBoundaries.cxx:16-95,
// Mesh3DBoundaries.cxx:14-50).
std::vector<int> add_pml(Q1Mesh &m, PmlSpec &s) {
  const int dim = m.dim, nc = 1 << dim;
  const int64_t nv = (int64_t)(m.vert.size() / dim);
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int64_t i = 0; i < nv; ++i)
    for (int a = 0; a < dim; ++a) {
      lo[a] = std::min(lo[a], m.vert[(size_t)i * dim + a]);
      hi[a] = std::max(hi[a], m.vert[(size_t)i * dim + a]);
    }
  if (s.layers == 0) {
    std::vector<char> onb(nv, 0);
    for (int64_t i = 0; i < nv; ++i)
      for (int a = 0; a < dim; ++a) {
        const double x = m.vert[(size_t)i * dim + a];
        if (std::fabs(x - lo[a]) <= 1e-10 || std::fabs(x - hi[a]) <= 1e-10) onb[i] = 1;
      }
    double sum = 0.0;
    int64_t cnt = 0;
    for (const auto &f : m.bnd) {
      const int n = (int)f.size();
      for (int j = 0; j < (dim == 2 ? 1 : n); ++j) {
        const int64_t v0 = f[j], v1 = f[(j + 1) % n];
        if (!onb[v0] || !onb[v1]) continue;
        double d2 = 0.0;
        for (int a = 0; a < dim; ++a) {
          const double dx = m.vert[(size_t)v0 * dim + a] - m.vert[(size_t)v1 * dim + a];
          d2 += dx * dx;
        }
        sum += std::sqrt(d2);
        ++cnt;
      }
    }
    const double h = cnt ? sum / cnt : 0.0;
    s.layers = h > 0 ? std::max(1, (int)std::ceil(s.thickness / h)) : 1;
  }
  const int64_t E0 = (int64_t)m.elem_ref.size();
  for (int a = 0; a < dim; ++a) {
    // the current extent along a (the earlier axes' layers do not move it)
    double mn = 1e300, mx = -1e300;
    for (size_t i = 0; i < m.vert.size() / dim; ++i) {
      mn = std::min(mn, m.vert[i * dim + a]);
      mx = std::max(mx, m.vert[i * dim + a]);
    }
    if (s.side[a] & 1) extrude(m, a, s.layers, mn, -s.thickness);
    if (s.side[a] & 2) extrude(m, a, s.layers, mx, s.thickness);
  }
  const int64_t E = (int64_t)m.elem_ref.size();
  std::vector<int> kind(E, 0);
  for (int64_t e = E0; e < E; ++e) {
    double c[3] = {0, 0, 0};
    for (int j = 0; j < nc; ++j)
      for (int a = 0; a < dim; ++a) c[a] += m.vert[(size_t)m.elem[(size_t)e * nc + j] * dim + a];
    for (int a = 0; a < dim; ++a) {
      c[a] *= 1.0 / nc;
      if (c[a] < lo[a]) kind[e] |= 1 << (2 * a);
      if (c[a] > hi[a]) kind[e] |= 2 << (2 * a);
    }
  }
  return kind;
}

// DampingPML = [PARABOLE|LINEAR|CONSTANT|SHIFTED_PARABOLE] vsigma [offset]
// (VarBoundaryCondition_Base::SetInputData, Harmonic/
// BoundaryConditionHarmonic.cxx:130-199) and TransverseDampingPML = coef
// (:117-129)
struct PmlDamping {
  int profile = 0;  // 0 parabola, 1 constant, 2 linear, 3 shifted parabola
  double vsigma = 0.0, offset = 0.0, transverse = 0.0;
};

PmlDamping parse_pml_damping(const Deck &d) {
  PmlDamping p;
  if (const auto *v = d.get("DampingPML")) {
    if (v->empty()) die("DampingPML needs vsigma");
    const std::string &w = (*v)[0];
    auto need = [&](size_t n) {
      if (v->size() < n) die("DampingPML = " + w + " needs more parameters");
    };
    if (w == "PARABOLE") { need(2); p.vsigma = num((*v)[1]); }
    else if (w == "CONSTANT") { need(2); p.profile = 1; p.vsigma = num((*v)[1]); }
    else if (w == "LINEAR") { need(2); p.profile = 2; p.vsigma = num((*v)[1]); }
    else if (w == "SHIFTED_PARABOLE") {
      need(3);
      p.profile = 3;
      p.vsigma = num((*v)[1]);
      p.offset = num((*v)[2]);
    } else {
      p.vsigma = num(w);
    }
  }
  if (const auto *v = d.get("TransverseDampingPML")) {
    if (v->empty()) die("TransverseDampingPML needs a coefficient");
    p.transverse = num((*v)[0]);
  }
  return p;
}

// ------------------------------------------------------- Q_r numbering
// Global GL nodes shared by neighbouring elements are identified
// topologically: a node on an edge by the edge's vertex ids and its index
// counted from the lower id; on a face by the face's lowest-id corner, its
// lower-id then higher-id neighbour corners, and its indices in that frame.
struct Key {
  uint64_t a, b;
  bool operator==(const Key &o) const { return a == o.a && b == o.b; }
};
struct KeyHash {
  size_t operator()(const Key &k) const {
    uint64_t h = k.a * 0x9E3779B97F4A7C15ULL ^ (k.b + 0x632BE59BD9B4E019ULL + (k.a << 6));
    return (size_t)(h ^ (h >> 31));
  }
};

struct QrMesh {
  int dim, order;
  int64_t N = 0, E = 0;
  std::vector<double> xyz;
  std::vector<int32_t> en;
  std::vector<uint8_t> dir;
};

QrMesh number_nodes(const Q1Mesh &m, int order, const std::set<int> &dirichlet_refs) {
  const int dim = m.dim, r = order, n = r + 1, nc = 1 << dim;
  const int nd = dim == 3 ? n * n * n : n * n;
  double xg[32], wg[32];
  check(mjw_gauss_lobatto(order, xg, wg), nullptr, "mjw_gauss_lobatto");
  QrMesh q;
  q.dim = dim;
  q.order = order;
  q.E = (int64_t)m.elem_ref.size();
  q.en.assign(q.E * nd, -1);
  std::unordered_map<Key, int32_t, KeyHash> ids;
  ids.reserve(q.E * nd / 2);
  // corner index of (a, b, c) in {0,1}^dim, Medit order
  auto corner = [&](int a, int b, int c) {
    static const int c2[2][2] = {{0, 3}, {1, 2}};
    return c2[a][b] + 4 * c;
  };
  std::set<std::vector<int64_t>> dir_faces;
  for (size_t f = 0; f < m.bnd.size(); ++f)
    if (dirichlet_refs.count(m.bnd_ref[f])) {
      std::vector<int64_t> v = m.bnd[f];
      std::sort(v.begin(), v.end());
      dir_faces.insert(v);
    }
  std::vector<uint8_t> dflag;
  for (int64_t e = 0; e < q.E; ++e) {
    const int64_t *V = &m.elem[e * nc];
    const int K = dim == 3 ? n : 1;
    for (int k = 0; k < K; ++k)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          const int idx[3] = {i, j, k};
          int side[3] = {-1, -1, -1}, nb = 0;
          for (int a = 0; a < dim; ++a)
            if (idx[a] == 0 || idx[a] == r) { side[a] = idx[a] == r; ++nb; }
          Key key{0, 0};
          bool shared = true;
          if (nb == dim) {  // vertex
            const int64_t g = V[corner(side[0], side[1], dim == 3 ? side[2] : 0)];
            key = {(1ull << 62) | (uint64_t)g, 0};
          } else if (nb == dim - 1) {  // edge
            int d = 0;
            while (side[d] >= 0) ++d;
            int s0[3] = {side[0], side[1], side[2]}, s1[3] = {side[0], side[1], side[2]};
            s0[d] = 0;
            s1[d] = 1;
            int64_t ga = V[corner(s0[0], s0[1], dim == 3 ? s0[2] : 0)];
            int64_t gb = V[corner(s1[0], s1[1], dim == 3 ? s1[2] : 0)];
            int t = idx[d];
            if (ga > gb) { std::swap(ga, gb); t = r - t; }
            key = {(2ull << 62) | ((uint64_t)ga << 31) | (uint64_t)gb, (uint64_t)t};
          } else if (dim == 3 && nb == 1) {  // face
            int d0 = 0;
            while (side[d0] < 0) ++d0;
            int du = -1, dv = -1;
            for (int a = 0; a < 3; ++a)
              if (a != d0) { if (du < 0) du = a; else dv = a; }
            int64_t g[2][2];
            for (int a = 0; a < 2; ++a)
              for (int b = 0; b < 2; ++b) {
                int s[3];
                s[d0] = side[d0];
                s[du] = a;
                s[dv] = b;
                g[a][b] = V[corner(s[0], s[1], s[2])];
              }
            int ou = 0, ov = 0;
            for (int a = 0; a < 2; ++a)
              for (int b = 0; b < 2; ++b)
                if (g[a][b] < g[ou][ov]) { ou = a; ov = b; }
            const int su = ou == 0 ? idx[du] : r - idx[du];
            const int sv = ov == 0 ? idx[dv] : r - idx[dv];
            const int64_t nu = g[1 - ou][ov], nv = g[ou][1 - ov];
            int s1, s2;
            int64_t first, second;
            if (nu < nv) { s1 = su; s2 = sv; first = nu; second = nv; }
            else { s1 = sv; s2 = su; first = nv; second = nu; }
            key = {(3ull << 62) | ((uint64_t)g[ou][ov] << 31) | (uint64_t)first,
                   ((uint64_t)second << 16) | ((uint64_t)s1 << 8) | (uint64_t)s2};
          } else {
            shared = false;
          }
          int32_t gid;
          bool fresh = true;
          if (shared) {
            auto it = ids.find(key);
            if (it != ids.end()) { gid = it->second; fresh = false; }
            else { gid = (int32_t)q.N; ids.emplace(key, gid); }
          } else {
            gid = (int32_t)q.N;
          }
          if (fresh) {
            ++q.N;
            // position: multilinear map of the element's vertices at the GL point
            double x[3] = {0, 0, 0};
            for (int c = 0; c < nc; ++c) {
              static const int ca[8] = {0, 1, 1, 0, 0, 1, 1, 0}, cb[8] = {0, 0, 1, 1, 0, 0, 1, 1},
                               cc[8] = {0, 0, 0, 0, 1, 1, 1, 1};
              const double wa = ca[c] ? xg[i] : 1 - xg[i], wb = cb[c] ? xg[j] : 1 - xg[j],
                           wc = dim == 3 ? (cc[c] ? xg[k] : 1 - xg[k]) : 1.0;
              for (int a = 0; a < dim; ++a) x[a] += wa * wb * wc * m.vert[V[c] * dim + a];
            }
            for (int a = 0; a < dim; ++a) q.xyz.push_back(x[a]);
            dflag.push_back(0);
          }
          q.en[e * nd + i + n * j + n * n * k] = gid;
        }
    // Dirichlet: every node of an element face listed as a Dirichlet boundary
    for (int d0 = 0; d0 < dim; ++d0)
      for (int sd = 0; sd < 2; ++sd) {
        std::vector<int64_t> fv;
        for (int c = 0; c < nc; ++c) {
          static const int ca[8] = {0, 1, 1, 0, 0, 1, 1, 0}, cb[8] = {0, 0, 1, 1, 0, 0, 1, 1},
                           cc[8] = {0, 0, 0, 0, 1, 1, 1, 1};
          const int s[3] = {ca[c], cb[c], cc[c]};
          if (s[d0] == sd) fv.push_back(V[c]);
        }
        std::sort(fv.begin(), fv.end());
        if (!dir_faces.count(fv)) continue;
        const int K = dim == 3 ? n : 1;
        for (int k = 0; k < K; ++k)
          for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
              const int idx[3] = {i, j, k};
              if (idx[d0] == (sd ? r : 0)) dflag[q.en[e * nd + i + n * j + n * n * k]] = 1;
            }
      }
  }
  q.dir = dflag;
  return q;
}

// ------------------------------------------------- Montjoie DOF numbering
// Montjoie's global number of every node (the order of its solution vectors,
// hence of SaveReprise files):
//   - referenced boundaries sorted first: 2-D edges as (min, max)
//     (Mesh/Mesh2DBoundaries.cxx:885-910); 3-D faces rotated to their lowest
//     vertex, reversed when the next vertex exceeds the previous one
//     (Mesh/Mesh3DBoundaries.cxx:650-700);
//   - edges numbered as first met on the referenced boundaries (local edges
//     (v_k, v_k+1)), then on the elements (square (0,1) (1,2) (2,3) (3,0);
//     cube: Mesh/Mesh3D.cxx:5201-5300 with the edge table of
//     Quadrature/HexahedronQuadrature.cxx:41-47); faces (3-D)
//     from the referenced faces, then each element's six faces rotated to their
//     lowest vertex (Mesh/Mesh3D.cxx:4928-5170);
//   - DOFs: vertices, then r-1 per edge from its lower to its higher vertex,
//     then (r-1)^2 per face, (s-1)(r-1) + (t-1) with s along v0 -> v1 and
//     t along v0 -> v3 of the stored face, then (r-1)^dim per element with the
//     first index outermost (Mesh/NumberMesh.cxx:695-900,
//     face table Mesh/NumberMesh.cxx:2609).
// Returns dof[g] for the driver's node g.
std::vector<int64_t> montjoie_dofs(const Q1Mesh &m, const QrMesh &q) {
  const int dim = m.dim, r = q.order, n = r + 1, nc = 1 << dim;
  const int nd = dim == 3 ? n * n * n : n * n;
  int64_t nv = 0;
  for (int64_t v : m.elem) nv = std::max(nv, v + 1);
  std::unordered_map<Key, int64_t, KeyHash> edges;
  auto add_edge = [&](int64_t a, int64_t b) {
    if (a > b) std::swap(a, b);
    const Key k{(uint64_t)a, (uint64_t)b};
    if (!edges.count(k)) edges.emplace(k, (int64_t)edges.size());
  };
  // faces: key = sorted corners; value = number and stored corner order
  struct Face {
    int64_t num;
    int64_t v[4];
  };
  auto face_key = [](const int64_t *f) {
    int64_t s[4] = {f[0], f[1], f[2], f[3]};
    std::sort(s, s + 4);
    return Key{((uint64_t)s[0] << 32) | (uint64_t)s[1], ((uint64_t)s[2] << 32) | (uint64_t)s[3]};
  };
  std::unordered_map<Key, Face, KeyHash> faces;
  auto add_face = [&](const int64_t *f) {
    const Key k = face_key(f);
    if (faces.count(k)) return;
    Face F;
    F.num = (int64_t)faces.size();
    std::copy(f, f + 4, F.v);
    faces.emplace(k, F);
  };
  // Medit corner c -> (a, b, c) in {0,1}^dim
  static const int ca[8] = {0, 1, 1, 0, 0, 1, 1, 0}, cb[8] = {0, 0, 1, 1, 0, 0, 1, 1},
                   cz[8] = {0, 0, 0, 0, 1, 1, 1, 1};
  if (dim == 2) {
    for (const auto &b : m.bnd) add_edge(b[0], b[1]);
    for (size_t e = 0; e * nc < m.elem.size(); ++e)
      for (int k = 0; k < 4; ++k) add_edge(m.elem[e * nc + k], m.elem[e * nc + (k + 1) % 4]);
  } else {
    std::vector<std::array<int64_t, 4>> ref;
    for (const auto &b : m.bnd) {
      if (b.size() != 4) die("SaveReprise: boundary faces must be quadrilaterals");
      int k0 = 0;
      for (int k = 1; k < 4; ++k)
        if (b[k] < b[k0]) k0 = k;
      std::array<int64_t, 4> f;
      const bool rev = b[(k0 + 1) % 4] > b[(k0 + 3) % 4];
      for (int j = 0; j < 4; ++j) f[j] = rev ? b[(k0 + 4 - j) % 4] : b[(k0 + j) % 4];
      ref.push_back(f);
    }
    for (const auto &f : ref)
      for (int k = 0; k < 4; ++k) add_edge(f[k], f[(k + 1) % 4]);
    static const int ce[12][2] = {{0, 1}, {1, 2}, {3, 2}, {0, 3}, {0, 4}, {1, 5},
                                  {2, 6}, {3, 7}, {4, 5}, {5, 6}, {7, 6}, {4, 7}};
    for (size_t e = 0; e * nc < m.elem.size(); ++e)
      for (const auto &ed : ce) add_edge(m.elem[e * nc + ed[0]], m.elem[e * nc + ed[1]]);
    for (const auto &f : ref) add_face(f.data());
    static const int cf[6][4] = {{0, 3, 7, 4}, {0, 1, 5, 4}, {0, 1, 2, 3},
                                 {4, 5, 6, 7}, {3, 2, 6, 7}, {1, 2, 6, 5}};
    for (size_t e = 0; e * nc < m.elem.size(); ++e)
      for (const auto &fl : cf) {
        int64_t v[4];
        for (int j = 0; j < 4; ++j) v[j] = m.elem[e * nc + fl[j]];
        int k0 = 0;
        for (int k = 1; k < 4; ++k)
          if (v[k] < v[k0]) k0 = k;
        int64_t f[4];
        for (int j = 0; j < 4; ++j) f[j] = v[(k0 + j) % 4];
        add_face(f);
      }
  }
  const int64_t ne = (int64_t)edges.size(), nf = (int64_t)faces.size();
  const int64_t off_e = nv, off_f = nv + ne * (r - 1), off_i = off_f + nf * (r - 1) * (r - 1);
  const int64_t ni = (int64_t)(dim == 3 ? (r - 1) * (r - 1) * (r - 1) : (r - 1) * (r - 1));
  std::vector<int64_t> dof(q.N, -1);
  for (int64_t e = 0; e < q.E; ++e) {
    const int64_t *V = &m.elem[e * nc];
    auto corner_of = [&](const int *cc) {  // coordinates in {0,1}^dim -> Medit corner
      for (int c = 0; c < nc; ++c)
        if (ca[c] == cc[0] && cb[c] == cc[1] && (dim == 2 || cz[c] == cc[2])) return c;
      return -1;
    };
    const int K = dim == 3 ? n : 1;
    for (int k = 0; k < K; ++k)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          const int idx[3] = {i, j, k};
          int nb = 0;
          for (int a = 0; a < dim; ++a) nb += idx[a] == 0 || idx[a] == r;
          int64_t d;
          if (nb == dim) {
            const int cc[3] = {i / r, j / r, k / r};
            d = V[corner_of(cc)];
          } else if (nb == dim - 1) {
            int ax = 0;
            while (idx[ax] == 0 || idx[ax] == r) ++ax;
            int s0[3], s1[3];
            for (int a = 0; a < 3; ++a) s0[a] = s1[a] = idx[a] / std::max(r, 1);
            s0[ax] = 0;
            s1[ax] = 1;
            int64_t a = V[corner_of(s0)], b = V[corner_of(s1)];
            int t = idx[ax];
            if (a > b) { std::swap(a, b); t = r - t; }
            d = off_e + edges.at(Key{(uint64_t)a, (uint64_t)b}) * (r - 1) + t - 1;
          } else if (dim == 3 && nb == 1) {
            int ax = 0;
            while (idx[ax] != 0 && idx[ax] != r) ++ax;
            int64_t fv[4];
            int pos[4][3], np = 0;
            for (int c = 0; c < 8; ++c) {
              const int cc[3] = {ca[c], cb[c], cz[c]};
              if (cc[ax] * r != idx[ax]) continue;
              fv[np] = V[c];
              std::copy(cc, cc + 3, pos[np]);
              ++np;
            }
            const Face &F = faces.at(face_key(fv));
            auto where = [&](int64_t v) {
              for (int c = 0; c < 4; ++c)
                if (fv[c] == v) return c;
              return -1;
            };
            const int *p0 = pos[where(F.v[0])], *p1 = pos[where(F.v[1])], *p3 = pos[where(F.v[3])];
            int s = 0, t = 0;
            for (int a = 0; a < 3; ++a) {
              s += (idx[a] - r * p0[a]) * (p1[a] - p0[a]);
              t += (idx[a] - r * p0[a]) * (p3[a] - p0[a]);
            }
            d = off_f + F.num * (r - 1) * (r - 1) + (int64_t)(s - 1) * (r - 1) + t - 1;
          } else {
            int64_t li = 0;
            for (int a = 0; a < dim; ++a) li = li * (r - 1) + idx[a] - 1;
            d = off_i + e * ni + li;
          }
          dof[q.en[e * nd + i + n * j + n * n * k]] = d;
        }
  }
  return dof;
}

// Seldon's Vector::Write: int32 size, then the values
template <class T>
void write_vector(const std::string &path, const std::vector<T> &v) {
  std::ofstream o(path, std::ios::binary);
  const int32_t n = (int32_t)v.size();
  o.write((const char *)&n, sizeof n);
  o.write((const char *)v.data(), (std::streamsize)(v.size() * sizeof(T)));
  if (!o) die("cannot write " + path);
}

// ------------------------------------------------------------------ media
// sigma_a at every quadrature point of the layers (GetDampingTauPML and
// EvaluateFunctionTauPML, Harmonic/BoundaryConditionHarmonic.cxx:1548-1570,
// 1639-1760; the first-order formulation keeps sigma itself, :1588-1636):
// dx = x_a minus the box face the element lies beyond, dsig = log(1000) 3
// vsigma vmax / (2 a^3), a the thickness, vmax = max sqrt(|mu / rho|) over the
// media of the layers (FindElementsInsidePML, :1424-1535); the box is the
// bounding box of the vertices of the physical elements.
std::vector<double> pml_tau(const Q1Mesh &m, const std::vector<int> &kind, const QrMesh &q,
                            const std::vector<double> &rho, const std::vector<double> &mu,
                            const PmlSpec &s, const PmlDamping &pd) {
  const int dim = m.dim, nc = 1 << dim, nd = (int)(q.en.size() / q.E);
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  double vmax = 0.0;
  for (int64_t e = 0; e < q.E; ++e) {
    if (kind[e]) {
      vmax = std::max(vmax, std::sqrt(std::fabs(mu[e] / rho[e])));
      continue;
    }
    for (int j = 0; j < nc; ++j)
      for (int a = 0; a < dim; ++a) {
        const double x = m.vert[(size_t)m.elem[(size_t)e * nc + j] * dim + a];
        lo[a] = std::min(lo[a], x);
        hi[a] = std::max(hi[a], x);
      }
  }
  const double A = s.thickness;
  const double dsig = std::log(1000e0) * 3e0 * pd.vsigma * vmax / (2 * std::pow(A, 3));
  auto f = [&](double dx) {
    switch (pd.profile) {
      case 1: return dsig * A * A;
      case 2: return dsig * std::fabs(dx) * A;
      case 3: return dsig * (dx * dx + pd.offset * A * A);
      default: return dsig * dx * dx;
    }
  };
  std::vector<double> tau((size_t)q.E * nd * dim, 0.0);
  for (int64_t e = 0; e < q.E; ++e) {
    if (!kind[e]) continue;
    for (int i = 0; i < nd; ++i) {
      const double *x = &q.xyz[(size_t)q.en[(size_t)e * nd + i] * dim];
      double z[3] = {0, 0, 0};
      for (int a = 0; a < dim; ++a) {
        if (kind[e] & (1 << (2 * a))) z[a] = f(x[a] - lo[a]);
        else if (kind[e] & (2 << (2 * a))) z[a] = f(x[a] - hi[a]);
      }
      if (pd.transverse != 0) {
        const double c = pd.transverse;
        double o[3];
        if (dim == 2) {
          o[0] = std::max(z[0], c * z[1]);
          o[1] = std::max(c * z[0], z[1]);
        } else {
          o[0] = std::max(z[0], c * std::max(z[1], z[2]));
          o[1] = std::max(c * std::max(z[0], z[2]), z[1]);
          o[2] = std::max(c * std::max(z[0], z[1]), z[2]);
        }
        for (int a = 0; a < dim; ++a) z[a] = o[a];
      }
      for (int a = 0; a < dim; ++a) tau[((size_t)e * nd + i) * dim + a] = z[a];
    }
  }
  return tau;
}

struct Media {
  std::vector<double> rho, mu, lam, sigma;  // sigma: damping (empty if none)
};

// MateriauDielec = ref ISOTROPE rho mu [sigma]          (acoustic)
//                  ref ISOTROPE rho lambda mu [sigma]   (elastic)
//                  ref YOUNG_POISSON rho E nu [sigma]   (elastic)
// read value after value as SetIndices does (VarHelmholtz / VarElastic,
// Elliptic/Elastic/VarElastic.cxx:1692-1708): rho, then the stiffness, then
// the damping sigma (rho u'' + sigma u' - div(...) = f), which is 0 when the
// line ends before it (GenericPhysicalIndice::SetInputVaryingMedia,
// Elliptic/PhysicalProperty.cxx:3614-3620).  YOUNG_POISSON gives mu =
// E / (2 (1 + nu)) and lambda = E nu / ((1 + nu)(1 - 2 nu)) in 3-D, but
// E nu / (1 - nu^2) in 2-D (ElasticPhysicalIndice::SetInputData,
// PhysicalProperty.cxx:3321-3345).
Media media(const Deck &d, const Q1Mesh &m, bool elastic) {
  std::map<int, std::vector<double>> mat;  // rho, lambda (elastic), mu, sigma
  for (const auto &v : d.all("MateriauDielec")) {
    const bool young = elastic && v.size() > 1 && v[1] == "YOUNG_POISSON";
    if (v.size() < 2 || (v[1] != "ISOTROPE" && !young))
      die(std::string("MateriauDielec: only ISOTROPE") + (elastic ? " and YOUNG_POISSON" : "") +
          " media are supported");
    const size_t nv = elastic ? 3 : 2;  // values before sigma
    if (v.size() < 2 + nv)
      die("MateriauDielec " + v[0] + ": " + (elastic ? (young ? "rho E nu" : "rho lambda mu")
                                                     : "rho mu") + " expected");
    std::vector<double> p;
    p.push_back(num(v[2]));
    if (young) {
      const double E = num(v[3]), nu = num(v[4]);
      p.push_back(m.dim == 2 ? E * nu / (1.0 - nu * nu) : E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)));
      p.push_back(0.5 * E / (1.0 + nu));
    } else {
      if (elastic) p.push_back(num(v[3]));
      p.push_back(num(v[2 + nv - 1]));
    }
    const double sg = v.size() > 2 + nv ? num(v[2 + nv]) : 0.0;
    if (!(sg >= 0.0)) die("MateriauDielec: the damping sigma must be >= 0");
    p.push_back(sg);
    mat[(int)num(v[0])] = p;
  }
  Media md;
  for (int ref : m.elem_ref) {
    auto it = mat.find(ref);
    if (it == mat.end()) die("no MateriauDielec for element reference " + std::to_string(ref));
    const auto &p = it->second;
    md.rho.push_back(p[0]);
    if (elastic) md.lam.push_back(p[1]);
    md.mu.push_back(p[elastic ? 2 : 1]);
    md.sigma.push_back(p.back());
  }
  bool damped = false;
  for (double v : md.sigma) damped = damped || v > 0.0;
  if (!damped) md.sigma.clear();
  return md;
}

// PhysicalMedia = rho|mu ref ISOTROPE SINUS xmin xmax ymin ymax [zmin zmax]
//                 c_x c_y [c_z] offset amplitude L_x L_y [L_z] [cte]
// read as GenericPhysicalIndice::SetInputVaryingMedia does
// (Elliptic/PhysicalProperty.cxx:3614-3672; the word after ref is skipped for
// the scalar rho, ScalarPhysicalIndice::SetInputData, and must be ISOTROPE
// for the tensor mu, PhysicalProperty.cxx:2421-2436).
struct Sinus {
  int dim = 3;
  double xmin[3] = {0, 0, 0}, xmax[3] = {0, 0, 0}, coef[3] = {0, 0, 0}, k[3] = {0, 0, 0};
  double offset = 0, amp = 0;
};

Sinus parse_sinus(const std::vector<std::string> &v, int dim) {
  const std::vector<std::string> p(v.begin() + 2, v.end());
  const int M = (int)p.size();
  int nb = 1;
  if (nb >= M || p[nb] != "SINUS")
    die("PhysicalMedia " + v[0] + ": only the SINUS field is on this path");
  if (v[0] == "mu" && p[0] != "ISOTROPE") die("PhysicalMedia mu: only ISOTROPE");
  ++nb;
  Sinus s;
  s.dim = dim;
  const double two_pi = 6.283185307179586476925286766559;
  if (nb < M - 3) {
    s.xmin[0] = num(p[nb++]); s.xmax[0] = num(p[nb++]);
    s.xmin[1] = num(p[nb++]); s.xmax[1] = num(p[nb++]);
  }
  if (dim == 3 && nb < M - 1) { s.xmin[2] = num(p[nb++]); s.xmax[2] = num(p[nb++]); }
  if (nb < M - 1) { s.coef[0] = num(p[nb++]); s.coef[1] = num(p[nb++]); }
  if (dim == 3 && nb < M) s.coef[2] = num(p[nb++]);
  if (nb < M - 3) {
    s.offset = num(p[nb++]);
    s.amp = num(p[nb++]);
    s.k[0] = two_pi / num(p[nb++]);
    s.k[1] = two_pi / num(p[nb++]);
  }
  if (dim == 3 && nb < M) s.k[2] = two_pi / num(p[nb++]);
  return s;
}

// the field at a point (PhysicalSinusoidalMedia::ComputeIndexAtPoint,
// PhysicalProperty.cxx:218-245; box test with the 1e-10 threshold
// Mesh/NumberMesh.cxx:1621,2727 sets; Gaussian cut-off of RectangleCutOff,
// PhysicalProperty.cxx:28-95)
double sinus_at(const Sinus &s, const double *x) {
  const double th = 1e-10;
  for (int a = 0; a < s.dim; ++a)
    if (x[a] < s.xmin[a] - th || x[a] > s.xmax[a] + th) return s.offset;
  double c = 1.0;
  for (int a = 0; a < s.dim; ++a) c *= std::sin(x[a] * s.k[a]);
  bool cut = false;
  for (int a = 0; a < s.dim; ++a) cut = cut || s.coef[a] != 0.0;
  if (cut)
    for (int a = 0; a < s.dim; ++a) {
      double f[2];
      const double b[2] = {s.xmin[a], s.xmax[a]};
      for (int q = 0; q < 2; ++q) {
        const double d = x[a] - b[q];
        f[q] = std::fabs(d) > th ? std::exp(-s.coef[a] / (d * d)) : (s.coef[a] != 0.0 ? 0.0 : 1.0);
      }
      c *= f[0] * f[1];
    }
  return s.offset + s.amp * c;
}

// Montjoie's local number of the nodal point (x_i, x_j[, x_k]) of a Q_r
// quadrilateral / hexahedron, stored at i + (r+1) (j + (r+1) k)
// (QuadrangleQuadrature.cxx:42-100, HexahedronQuadrature.cxx:43-182):
// vertices, edges (square: bottom, right, top and left reversed; cube: the 12
// edges of GetEdgeCube from first to second vertex), faces x=0, y=0, z=0,
// z=r, y=r, x=r (i outer, j inner), then the interior (i outer, k inner).
std::vector<int> nodal_numbering(int dim, int r) {
  const int n = r + 1;
  auto at = [&](int i, int j, int k) { return i + n * (j + n * k); };
  std::vector<int> num(dim == 3 ? n * n * n : n * n, -1);
  int nb = 0;
  if (dim == 2) {
    num[at(0, 0, 0)] = 0; num[at(r, 0, 0)] = 1; num[at(r, r, 0)] = 2; num[at(0, r, 0)] = 3;
    nb = 4;
    for (int i = 1; i < r; ++i) num[at(i, 0, 0)] = nb++;
    for (int i = 1; i < r; ++i) num[at(r, i, 0)] = nb++;
    for (int i = 1; i < r; ++i) num[at(r - i, r, 0)] = nb++;
    for (int i = 1; i < r; ++i) num[at(0, r - i, 0)] = nb++;
    for (int i = 1; i < r; ++i)
      for (int j = 1; j < r; ++j) num[at(i, j, 0)] = nb++;
    return num;
  }
  static const int cor[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  for (int c = 0; c < 8; ++c) num[at(cor[c][0] * r, cor[c][1] * r, cor[c][2] * r)] = nb++;
  static const int edge[12][2] = {{0, 1}, {1, 2}, {3, 2}, {0, 3}, {0, 4}, {1, 5},
                                  {2, 6}, {3, 7}, {4, 5}, {5, 6}, {7, 6}, {4, 7}};
  for (const auto &ed : edge)
    for (int l = 1; l < r; ++l) {
      int ijk[3];
      for (int a = 0; a < 3; ++a) ijk[a] = (l * cor[ed[1]][a] * r + (r - l) * cor[ed[0]][a] * r) / r;
      num[at(ijk[0], ijk[1], ijk[2])] = nb++;
    }
  for (int f = 0; f < 6; ++f)
    for (int i = 1; i < r; ++i)
      for (int j = 1; j < r; ++j) {
        const int ijk[6][3] = {{0, i, j}, {i, 0, j}, {i, j, 0}, {i, j, r}, {i, r, j}, {r, i, j}};
        num[at(ijk[f][0], ijk[f][1], ijk[f][2])] = nb++;
      }
  for (int i = 1; i < r; ++i)
    for (int j = 1; j < r; ++j)
      for (int k = 1; k < r; ++k) num[at(i, j, k)] = nb++;
  return num;
}

// PhysicalMedia = rho|mu ref ISOTROPE MESH|SAME_MESH mesh_file offset amplitude
//                 data_file order [cte] on the computational mesh
// (PhysicalProperty.cxx:3674-3704): the data file (ReadMeshData,
// Output/CommonInputOutput.cxx:325-352: int 0, int E, int offsets[E+1],
// doubles) gives every element's values at its order-`order` nodal points;
// offset + amplitude nu (:990-1000) is interpolated with the tensor Lagrange
// basis on the GL points of that order at the element's quadrature points
// (ProjectScalar, Harmonic/VarGeometryProblem.cxx:1428-1447).
void mesh_media(const std::vector<std::string> &v, const Deck &d, const Q1Mesh &m,
                const QrMesh &qm, int dim, int ref, std::vector<double> &dst) {
  const std::vector<std::string> p(v.begin() + 2, v.end());
  if (p.size() < 7) die("PhysicalMedia " + v[0] + " MESH: mesh_file offset amplitude data_file order");
  if (v[0] == "mu" && p[0] != "ISOTROPE") die("PhysicalMedia mu: only ISOTROPE");
  const std::string &mesh_file = p[2];
  const double offset = num(p[3]), amplitude = num(p[4]);
  std::string data_file = p[5];
  const int rm = (int)num(p[6]);
  const auto &fm = d.need("FileMesh");
  if (p[1] == "MESH" && mesh_file != fm[0])
    die("PhysicalMedia MESH: the field must live on the computational mesh (" + fm[0] + ")");
  if (rm < 1 || rm > 30) die("PhysicalMedia MESH: order out of range");
  std::ifstream in(data_file, std::ios::binary);
  if (!in && !data_file.empty() && data_file[0] != '/') in.open(d.dir + "/" + data_file, std::ios::binary);
  if (!in) die("cannot read " + data_file);
  int32_t head[2];
  in.read(reinterpret_cast<char *>(head), 8);
  if (!in || head[0] != 0) die(data_file + ": not a real-valued mesh data file");
  if (head[1] != qm.E) die(data_file + ": element count differs from the mesh");
  std::vector<int32_t> off(head[1] + 1);
  in.read(reinterpret_cast<char *>(off.data()), (std::streamsize)(off.size() * 4));
  const int npn = dim == 3 ? (rm + 1) * (rm + 1) * (rm + 1) : (rm + 1) * (rm + 1);
  for (int64_t e = 0; e < qm.E; ++e)
    if (off[e + 1] - off[e] != npn) die(data_file + ": nodal count differs from order " + p[6]);
  std::vector<double> nu((size_t)qm.E * npn);
  in.read(reinterpret_cast<char *>(nu.data()), (std::streamsize)(nu.size() * 8));
  if (!in) die(data_file + ": truncated");
  // L[q][a]: Lagrange basis of the order-rm GL points at the order-r points
  const int n = qm.order + 1, nm1 = rm + 1;
  double xm[32], wm[32], xq[32], wq[32];
  check(mjw_gauss_lobatto(rm, xm, wm), nullptr, "mjw_gauss_lobatto");
  check(mjw_gauss_lobatto(qm.order, xq, wq), nullptr, "mjw_gauss_lobatto");
  std::vector<double> L((size_t)n * nm1);
  for (int q = 0; q < n; ++q)
    for (int a = 0; a < nm1; ++a) {
      double v = 1.0;
      for (int b = 0; b < nm1; ++b)
        if (b != a) v *= (xq[q] - xm[b]) / (xm[a] - xm[b]);
      L[(size_t)q * nm1 + a] = v;
    }
  const std::vector<int> num_loc = nodal_numbering(dim, rm);
  const int K = dim == 3 ? nm1 : 1, Kq = dim == 3 ? n : 1;
  const int64_t nd = (int64_t)qm.en.size() / qm.E;
  std::vector<double> T((size_t)nm1 * nm1 * K), A1((size_t)n * nm1 * K), A2((size_t)n * n * K);
  for (int64_t e = 0; e < qm.E; ++e) {
    if (m.elem_ref[e] != ref) continue;
    for (int c = 0; c < K; ++c)
      for (int b = 0; b < nm1; ++b)
        for (int a = 0; a < nm1; ++a) {
          const size_t t = a + nm1 * (b + (size_t)nm1 * c);
          T[t] = offset + amplitude * nu[(size_t)e * npn + num_loc[t]];
        }
    // separable contractions: a -> i, b -> j, c -> k
    for (int c = 0; c < K; ++c)
      for (int b = 0; b < nm1; ++b)
        for (int i = 0; i < n; ++i) {
          double s = 0;
          for (int a = 0; a < nm1; ++a) s += L[(size_t)i * nm1 + a] * T[a + nm1 * (b + (size_t)nm1 * c)];
          A1[i + n * (b + (size_t)nm1 * c)] = s;
        }
    for (int c = 0; c < K; ++c)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          double s = 0;
          for (int b = 0; b < nm1; ++b) s += L[(size_t)j * nm1 + b] * A1[i + n * (b + (size_t)nm1 * c)];
          A2[i + n * (j + (size_t)n * c)] = s;
        }
    for (int k = 0; k < Kq; ++k)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          double s = 0;
          if (dim == 3) {
            for (int c = 0; c < K; ++c) s += L[(size_t)k * nm1 + c] * A2[i + n * (j + (size_t)n * c)];
          } else {
            s = A2[i + n * j];
          }
          dst[e * nd + i + n * (j + (int64_t)n * k)] = s;
        }
  }
}

// rho and mu at every quadrature point (the GL nodes of each element) when a
// PhysicalMedia line makes them vary; empty otherwise
struct NodalMedia {
  std::vector<double> rho, mu;  // empty unless a rho or mu field is given: E * nd each
  std::vector<double> sigma;    // empty unless a sigma (damping) field is given
};

NodalMedia nodal_media(const Deck &d, const Q1Mesh &m, const QrMesh &qm, const Media &md, int dim,
                       bool elastic) {
  NodalMedia nm;
  const auto lines = d.all("PhysicalMedia");
  if (lines.empty()) return nm;
  const int64_t nd = (int64_t)qm.en.size() / qm.E;
  bool rhomu = false, sig = false;
  for (const auto &v : lines) {
    if (v.size() < 4) die("PhysicalMedia = name ref ISOTROPE SINUS|MESH|SAME_MESH ...");
    // elastic media are rho, C and sigma (VarElastic.cxx:1752-1760); the
    // stiffness tensor C does not vary on this path
    if (elastic && v[0] == "C") die("PhysicalMedia C: a varying elastic stiffness is not supported");
    if (elastic ? (v[0] != "rho" && v[0] != "sigma")
                : (v[0] != "rho" && v[0] != "mu" && v[0] != "sigma"))
      die("PhysicalMedia " + v[0] + (elastic ? ": only rho and sigma" : ": only rho, mu and sigma"));
    (v[0] == "sigma" ? sig : rhomu) = true;
  }
  if (rhomu) {
    nm.rho.resize(qm.en.size());
    if (!elastic) nm.mu.resize(qm.en.size());
  }
  if (sig) nm.sigma.resize(qm.en.size());
  for (int64_t e = 0; e < qm.E; ++e)
    for (int64_t q = 0; q < nd; ++q) {
      if (rhomu) {
        nm.rho[e * nd + q] = md.rho[e];
        if (!elastic) nm.mu[e * nd + q] = md.mu[e];
      }
      if (sig) nm.sigma[e * nd + q] = md.sigma.empty() ? 0.0 : md.sigma[e];
    }
  for (const auto &v : lines) {
    const int ref = (int)num(v[1]);
    std::vector<double> &dst = v[0] == "rho" ? nm.rho : v[0] == "mu" ? nm.mu : nm.sigma;
    if (v[3] == "MESH" || v[3] == "SAME_MESH") {
      mesh_media(v, d, m, qm, dim, ref, dst);
      continue;
    }
    const Sinus s = parse_sinus(v, dim);
    for (int64_t e = 0; e < qm.E; ++e) {
      if (m.elem_ref[e] != ref) continue;
      for (int64_t q = 0; q < nd; ++q) {
        const int32_t g = qm.en[e * nd + q];
        dst[e * nd + q] = sinus_at(s, &qm.xyz[(size_t)g * dim]);
      }
    }
  }
  return nm;
}

// Snapshot grids (Output/GridInterpolation.cxx); SismoLine below.
// SismoPlane:
//   2-D: SismoPlane = xmin xmax ymin ymax Nx Ny | AUTO Nx Ny (:1780-1815; AUTO:
//        the mesh box shrunk by 5 machine epsilons of its size, :1244-1262),
//        points (xmin + i hx, ymin + j hy) (:2297-2345);
//   3-D: SismoPlane = xmin ymin zmin xmax ymax zmax x0 y0 z0 Nx Ny (:4732-4756),
//        points A + i/(Nx-1) (B - A) + j/(Ny-1) (C - A) (:5119-5166);
// value index j Nx + i; points outside the mesh hold 0.  Written at the
// ParametersOutputPlane times as <name>%04d.dat, numbered from 0, in
// Montjoie's binary double layout (WriteMatlab, :6060-6104: number of grids,
// dimension, data type, grid type, the grid's parameters, then the values
// with their count).
struct Plane {
  int dim = 2, nx = 0, ny = 0;
  bool line = false;                           // SismoLine: ny = 1
  bool circle = false;                         // SismoCircle: ny = 1, p = x0 y0 radius
  double p[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // plane 2-D: xmin xmax ymin ymax; 3-D: A, B, C
                                               // line: xmin xmax ymin ymax [zmin zmax]
  std::vector<double> pts;                     // nx*ny points, index j*nx + i
};

// SismoLine = xmin xmax ymin ymax [zmin zmax] N (GridInterpolation.cxx:1926-1943,
// :4782-4800): N points from (xmin, ymin[, zmin]) to (xmax, ymax[, zmax])
Plane line_grid(const std::vector<std::string> &v, int dim) {
  Plane g;
  g.dim = dim;
  g.line = true;
  const size_t nv = 2 * dim;
  if (v.size() < nv + 1) die(dim == 2 ? "SismoLine = xmin xmax ymin ymax N"
                                      : "SismoLine = xmin xmax ymin ymax zmin zmax N");
  for (size_t k = 0; k < nv; ++k) g.p[k] = num(v[k]);
  g.nx = (int)num(v[nv]);
  g.ny = 1;
  if (g.nx < 1) die("SismoLine: N must be positive");
  g.pts.resize((size_t)g.nx * dim);
  for (int i = 0; i < g.nx; ++i)
    for (int a = 0; a < dim; ++a) {
      const double h = g.nx > 1 ? (g.p[2 * a + 1] - g.p[2 * a]) / (g.nx - 1) : 0.0;
      g.pts[(size_t)i * dim + a] = g.p[2 * a] + i * h;
    }
  return g;
}

// SismoCircle = x0 y0 radius N (2-D; GridInterpolation.cxx:1966-1983): the box
// [x0 - r, x0 + r] x [y0 - r, y0 + r] is kept, and the points are
// c + rho (cos i dtheta, sin i dtheta) with c = the box centre, rho = half its
// width and dtheta = 2 pi / N (:2721-2740); points outside the mesh hold 0.
// Written at the ParametersOutputCircle times like the planes, grid type 3,
// header c, rho, N (:3076-3079, :3203-3205).
Plane circle_grid(const std::vector<std::string> &v, int dim) {
  if (dim != 2 || v.size() < 4) die("SismoCircle = x0 y0 radius N (2-D)");
  Plane g;
  g.dim = 2;
  g.circle = true;
  const double cx = num(v[0]), cy = num(v[1]), rad = num(v[2]);
  g.nx = (int)num(v[3]);
  g.ny = 1;
  if (g.nx < 1) die("SismoCircle: N must be positive");
  const double xmin = cx - rad, xmax = cx + rad, ymin = cy - rad, ymax = cy + rad;
  const double pi = 3.14159265358979323846;
  const double step = (2.0 * pi) / g.nx;
  const double c0 = (xmin + xmax) / 2.0, c1 = (ymin + ymax) / 2.0, r = (xmax - xmin) / 2.0;
  g.p[0] = 0.5 * (xmin + xmax);
  g.p[1] = 0.5 * (ymin + ymax);
  g.p[2] = 0.5 * (xmax - xmin);
  g.pts.resize((size_t)g.nx * 2);
  for (int i = 0; i < g.nx; ++i) {
    g.pts[(size_t)i * 2] = c0 + r * std::cos(step * i);
    g.pts[(size_t)i * 2 + 1] = c1 + r * std::sin(step * i);
  }
  return g;
}

Plane plane_grid(const std::vector<std::string> &v, int dim, const QrMesh &qm) {
  Plane g;
  g.dim = dim;
  if (dim == 2) {
    size_t nb = 0;
    if (!v.empty() && v[0] == "AUTO") {
      if (v.size() < 3) die("SismoPlane = AUTO Nx Ny");
      double lo[2] = {1e300, 1e300}, hi[2] = {-1e300, -1e300};
      for (int64_t q = 0; q < qm.N; ++q)
        for (int a = 0; a < 2; ++a) {
          lo[a] = std::min(lo[a], qm.xyz[q * 2 + a]);
          hi[a] = std::max(hi[a], qm.xyz[q * 2 + a]);
        }
      const double eps = 2.220446049250313e-16;
      for (int a = 0; a < 2; ++a) {
        const double w = hi[a] - lo[a];
        g.p[2 * a] = lo[a] + 5.0 * w * eps;
        g.p[2 * a + 1] = hi[a] - 5.0 * (hi[a] - g.p[2 * a]) * eps;
      }
      nb = 1;
    } else {
      if (v.size() < 6) die("SismoPlane = xmin xmax ymin ymax Nx Ny");
      for (int k = 0; k < 4; ++k) g.p[k] = num(v[k]);
      nb = 4;
    }
    g.nx = (int)num(v[nb]);
    g.ny = (int)num(v[nb + 1]);
    const double hx = g.nx > 1 ? (g.p[1] - g.p[0]) / (g.nx - 1) : 0.0;
    const double hy = g.ny > 1 ? (g.p[3] - g.p[2]) / (g.ny - 1) : 0.0;
    g.pts.resize((size_t)g.nx * g.ny * 2);
    for (int i = 0; i < g.nx; ++i)
      for (int j = 0; j < g.ny; ++j) {
        const size_t k = (size_t)j * g.nx + i;
        g.pts[k * 2] = g.p[0] + i * hx;
        g.pts[k * 2 + 1] = g.p[2] + j * hy;
      }
  } else {
    if (v.size() < 11) die("SismoPlane = xmin ymin zmin xmax ymax zmax x0 y0 z0 Nx Ny");
    double a[3], b[3], c[3];
    for (int k = 0; k < 3; ++k) {
      a[k] = num(v[k]);
      b[k] = num(v[3 + k]);
      c[k] = num(v[6 + k]);
    }
    // header order: xmin xmax ymin ymax zmin zmax x0 y0 z0
    for (int k = 0; k < 3; ++k) {
      g.p[2 * k] = a[k];
      g.p[2 * k + 1] = b[k];
      g.p[6 + k] = c[k];
    }
    g.nx = (int)num(v[9]);
    g.ny = (int)num(v[10]);
    const double sx = g.nx > 1 ? 1.0 / (g.nx - 1) : 0.0, sy = g.ny > 1 ? 1.0 / (g.ny - 1) : 0.0;
    g.pts.resize((size_t)g.nx * g.ny * 3);
    for (int i = 0; i < g.nx; ++i)
      for (int j = 0; j < g.ny; ++j) {
        const size_t k = (size_t)j * g.nx + i;
        for (int d = 0; d < 3; ++d)
          g.pts[k * 3 + d] = a[d] + i * sx * (b[d] - a[d]) + j * sy * (c[d] - a[d]);
      }
  }
  if (g.nx < 1 || g.ny < 1) die("SismoPlane: Nx and Ny must be positive");
  return g;
}

// OutputFormat = ASCII DOUBLE|FLOAT: the same fields as text, one per line
// (WriteMatlab's ascii branch, Output/GridInterpolation.cxx:6108-6130, with
// the grid header of WriteText, :3146-3175 (2-D) and :5597-5640 (3-D)):
// integers as such, reals at 15 (DOUBLE) or 7 (FLOAT) significant digits in
// the default stream notation (%g)
void write_plane_text(const std::string &name, const Plane &g, const std::vector<double> &vals,
                      bool dbl) {
  FILE *f = std::fopen(name.c_str(), "w");
  if (!f) die("cannot write " + name);
  const int prec = dbl ? 15 : 7;
  auto wr_d = [&](double x) { std::fprintf(f, "%.*g\n", prec, x); };
  std::fprintf(f, "1\n%d\n%d\n%d\n", g.dim, dbl ? 1 : 0,
               g.circle ? 3 : g.line ? (g.dim == 2 ? 1 : 3) : (g.dim == 2 ? 0 : 2));
  if (g.circle) {
    wr_d(g.p[0]); wr_d(g.p[1]); wr_d(g.p[2]); std::fprintf(f, "%d\n", g.nx);
  } else if (g.line) {
    for (int k = 0; k < 2 * g.dim; ++k) wr_d(g.p[k]);
    std::fprintf(f, "%d\n", g.nx);
  } else if (g.dim == 2) {
    wr_d(g.p[0]); wr_d(g.p[1]); std::fprintf(f, "%d\n", g.nx);
    wr_d(g.p[2]); wr_d(g.p[3]); std::fprintf(f, "%d\n", g.ny);
  } else {
    for (int k = 0; k < 9; ++k) wr_d(g.p[k]);
    std::fprintf(f, "%d\n%d\n", g.nx, g.ny);
  }
  for (double v : vals) wr_d(v);
  if (std::fclose(f) != 0) die("cannot write " + name);
}

// dbl: OutputFormat = BINARY DOUBLE (data type 1, grid parameters and values
// in double), else BINARY FLOAT (data type 0, floats)
void write_plane(const std::string &name, const Plane &g, const std::vector<double> &vals,
                 bool dbl, bool binary = true) {
  if (!binary) return write_plane_text(name, g, vals, dbl);
  std::ofstream o(name, std::ios::binary);
  if (!o) die("cannot write " + name);
  auto wi = [&](int32_t x) { o.write(reinterpret_cast<const char *>(&x), 4); };
  auto wd = [&](double x) {
    if (dbl) {
      o.write(reinterpret_cast<const char *>(&x), 8);
    } else {
      const float f = (float)x;
      o.write(reinterpret_cast<const char *>(&f), 4);
    }
  };
  wi(1);                        // one grid
  wi(g.dim);                    // dimension
  wi(dbl ? 1 : 0);              // real double / float values
  // grid type: plane 0 (2-D) / 2 (3-D), line 1 (2-D) / 3 (3-D), circle 3 (2-D)
  wi(g.circle ? 3 : g.line ? (g.dim == 2 ? 1 : 3) : (g.dim == 2 ? 0 : 2));
  if (g.circle) {
    wd(g.p[0]); wd(g.p[1]); wd(g.p[2]); wi(g.nx);
  } else if (g.line) {
    for (int k = 0; k < 2 * g.dim; ++k) wd(g.p[k]);
    wi(g.nx);
  } else if (g.dim == 2) {
    wd(g.p[0]); wd(g.p[1]); wi(g.nx);
    wd(g.p[2]); wd(g.p[3]); wi(g.ny);
  } else {
    for (int k = 0; k < 9; ++k) wd(g.p[k]);
    wi(g.nx); wi(g.ny);
  }
  wi((int32_t)vals.size());
  for (double v : vals) wd(v);
  if (!o) die("cannot write " + name);
}

template <class T>
void wr(std::ofstream &o, const std::vector<T> &v) {
  o.write(reinterpret_cast<const char *>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr,
                 "usage: mjwave_run deck.ini [--precision fp64|fp32|fp32_g64] [--device N] "
                 "[--dump-mesh FILE]\n");
    return 2;
  }
  std::string deck_path = argv[1], dump;
  int precision = MJW_FP64, device = 0;
  for (int a = 2; a < argc; ++a) {
    const std::string s = argv[a];
    if (s == "--precision" && a + 1 < argc) {
      const std::string p = argv[++a];
      precision = p == "fp32" ? MJW_FP32 : p == "fp32_g64" ? MJW_FP32_G64 : p == "fp64" ? MJW_FP64 : -1;
      if (precision < 0) die("unknown precision " + p);
    } else if (s == "--device" && a + 1 < argc) {
      device = std::atoi(argv[++a]);
    } else if (s == "--dump-mesh" && a + 1 < argc) {
      dump = argv[++a];
    } else {
      die("unknown option " + s);
    }
  }
  const Deck d = read_deck(deck_path);
  static const std::set<std::string> known = {
      "TypeElement", "TypeEquation", "MeshPath", "FileMesh", "TypeMesh", "OrderDiscretization",
      "ConditionReference", "MateriauDielec", "PhysicalMedia", "TypeSource", "OriginePhase", "Polarization",
      "Frequency", "TemporalSource", "TimeStep", "TimeInterval", "OrderTimeScheme", "SismoPoint",
      "SismoCircle", "FileOutputCircle", "ParametersOutputCircle", "FileOutputPoint",
      "ParametersOutputPoint", "SismoPlane",
      "FileOutputPlane", "ParametersOutputPlane", "SismoLine", "FileOutputLine",
      "ParametersOutputLine", "OutputFormat", "SaveReprise", "PathReprise", "LoadReprise", "DisplayStress",
      "ElectricOrMagnetic", "PrintLevel", "AddPML", "DampingPML", "TransverseDampingPML"};
  for (const auto &kv : d.kv)
    if (!known.count(kv.first))
      std::fprintf(stderr, "mjwave_run: keyword %s ignored\n", kv.first.c_str());

  const std::string te = d.need("TypeElement")[0];
  int dim;
  if (te == "QUADRANGLE_LOBATTO") dim = 2;
  else if (te == "HEXAHEDRON_LOBATTO") dim = 3;
  else die("TypeElement " + te + ": only QUADRANGLE_LOBATTO / HEXAHEDRON_LOBATTO");
  const std::string eq = d.need("TypeEquation")[0];
  const bool elastic = eq == "TIME_ELASTIC";
  if (!elastic && eq != "ACOUSTIC") die("TypeEquation " + eq + ": only ACOUSTIC / TIME_ELASTIC");
  const int ncomp = elastic ? dim : 1;
  const int order = (int)num(d.need("OrderDiscretization")[0]);
  // OrderTimeScheme = 2 LEAP_FROG or 4 LEAP_FROG (ModifiedEquationIterator of
  // that order, TimeSchemes.cxx:760-767) or 2 LEAP_FROG_PML (the first-order
  // split system, ModifiedEquationSystemIterator)
  bool first_order = false;
  int time_order = 2;
  if (const auto *o = d.get("OrderTimeScheme")) {
    const int r = o->empty() ? 0 : (int)num((*o)[0]);
    const bool lf = o->size() >= 2 && (*o)[1] == "LEAP_FROG";
    const bool lfp = o->size() >= 2 && (*o)[1] == "LEAP_FROG_PML";
    if (!((lf && (r == 2 || r == 4)) || (lfp && r == 2)))
      die("OrderTimeScheme: only 2 LEAP_FROG, 4 LEAP_FROG or 2 LEAP_FROG_PML");
    first_order = lfp;
    time_order = r;
  }
  PmlSpec pml_spec = parse_pml(d, dim);
  const PmlDamping pml_damp = parse_pml_damping(d);
  if (pml_spec.on && !first_order)
    die("AddPML: only with OrderTimeScheme = 2 LEAP_FROG_PML (the second-order PML formulation "
        "is not implemented)");
  if (first_order) {
    if (elastic) die("LEAP_FROG_PML: acoustic decks only");
    if (d.get("SaveReprise")) die("SaveReprise: not available with LEAP_FROG_PML");
    if (d.get("PhysicalMedia")) die("PhysicalMedia: not available with LEAP_FROG_PML (per-reference media only)");
    if (d.need("TimeStep")[0] == "AUTO") die("TimeStep = AUTO: not available with LEAP_FROG_PML");
  }
  // TypeSource = SRC_DIRAC (at OriginePhase), or SRC_VOLUME [references]
  // UNIFORM (f = 1, UniformSourceField, Source/SourceSpatiale.cxx:188-204) or
  // GAUSSIAN x y [z] radius [cut-off radius] (DefineSourceElliptic.cxx:
  // 470-560, InitGaussianParameter :169-195): f = beta exp(-alpha r^2),
  // alpha = -ln(1e-6) / radius^2, beta = sqrt(alpha / pi), 0 from the cut-off
  // radius on (the radius when not given; GaussianSource::GetAmplitude,
  // Source/SourceSpatiale.cxx:72-97), on the elements of the references
  // (every element when none is given), times the Polarization for elastic
  bool volume_source = false, vol_uniform = false;
  std::vector<int> vol_refs;
  double vol_x[3] = {0, 0, 0}, vol_r = 0, vol_rc = 0;
  if (const auto *s = d.get("TypeSource")) {
    const std::string k = s->empty() ? std::string() : (*s)[0];
    if (k == "SRC_VOLUME") {
      size_t w = 1;
      while (w < s->size() && !std::isalpha((unsigned char)(*s)[w][0]))
        vol_refs.push_back((int)num((*s)[w++]));
      if (w >= s->size() || ((*s)[w] != "GAUSSIAN" && (*s)[w] != "UNIFORM"))
        die("TypeSource = SRC_VOLUME: only the GAUSSIAN and UNIFORM volume sources");
      vol_uniform = (*s)[w] == "UNIFORM";
      ++w;
      if (!vol_uniform) {
        if (s->size() < w + dim + 1)
          die("TypeSource = SRC_VOLUME GAUSSIAN needs the centre and a radius");
        for (int a = 0; a < dim; ++a) vol_x[a] = num((*s)[w++]);
        vol_r = num((*s)[w++]);
        vol_rc = (w < s->size() && !std::isalpha((unsigned char)(*s)[w][0])) ? num((*s)[w]) : vol_r;
        if (!(vol_r > 0.0)) die("TypeSource = SRC_VOLUME GAUSSIAN: the radius must be positive");
      }
      volume_source = true;
    } else if (k != "SRC_DIRAC") {
      die("TypeSource " + k + ": only SRC_DIRAC and SRC_VOLUME GAUSSIAN");
    }
  }
  // TemporalSource = KIND [value] (GetNewTimeSource, VarInstationary.cxx:409-481)
  int tkind = MJW_TIME_RICKER;
  double tpar = 0.0;
  if (const auto *s = d.get("TemporalSource")) {
    const std::string k = s->empty() ? std::string() : (*s)[0];
    auto par = [&](const char *what) {
      if (s->size() < 2) die("TemporalSource = " + k + " " + what + ": the value is missing");
      return num((*s)[1]);
    };
    if (k == "RICKER") tkind = MJW_TIME_RICKER;
    else if (k == "DERIVATIVE_RICKER") tkind = MJW_TIME_DERIVATIVE_RICKER;
    else if (k == "MODIFIED_RICKER") { tkind = MJW_TIME_MODIFIED_RICKER; tpar = par("t0"); }
    else if (k == "GAUSSIAN") { tkind = MJW_TIME_GAUSSIAN; tpar = s->size() > 1 ? num((*s)[1]) : 0.0; }
    else if (k == "HARMONIC") { tkind = MJW_TIME_HARMONIC; tpar = par("tf"); }
    else if (k == "SINUS_GAUSSIAN") { tkind = MJW_TIME_SINUS_GAUSSIAN; tpar = par("alpha"); }
    else
      die("TemporalSource " + k + ": only RICKER, DERIVATIVE_RICKER, MODIFIED_RICKER, GAUSSIAN, "
          "HARMONIC and SINUS_GAUSSIAN");
  }
  // DisplayStress = NO: the elastic outputs keep grad u instead of C grad u
  // (VarElastic.cxx:134-139)
  bool display_stress = true;
  if (const auto *s = d.get("DisplayStress"))
    display_stress = s->empty() || (*s)[0] != "NO";
  if (const auto *s = d.get("LoadReprise"))
    if (!s->empty() && (*s)[0] == "YES") die("LoadReprise: restarting is not implemented");
  // SaveReprise = n: checkpoint every n steps (VarInstationary.cxx:321-333)
  int64_t save_every = 0;
  if (const auto *s = d.get("SaveReprise")) {
    if (s->empty()) die("SaveReprise needs a number of steps");
    save_every = (int64_t)num((*s)[0]);
    if (save_every <= 0) die("SaveReprise: the number of steps must be positive");
  }
  const std::string path_reprise = d.get("PathReprise") ? d.need("PathReprise")[0] : "./";
  // OutputFormat = BINARY|ASCII|... DOUBLE|FLOAT (Output/OutputHarmonic.cxx:321-380): the
  // precision sets the snapshot data type and the seismogram digits (15 or 7,
  // Output/GridInterpolation.cxx:3150-3152)
  bool out_binary = true, out_ascii = false, out_double = true;
  if (const auto *of = d.get("OutputFormat")) {
    if (of->size() < 2) die("OutputFormat needs a format and a precision");
    out_binary = (*of)[0] == "BINARY";
    out_ascii = (*of)[0] == "ASCII";
    if ((*of)[1] == "FLOAT") out_double = false;
    else if ((*of)[1] != "DOUBLE") die("OutputFormat: only DOUBLE or FLOAT precision");
  }

  // mesh
  const auto &fm = d.need("FileMesh");
  Q1Mesh q1;
  if (fm[0] == "REGULAR") {
    q1 = regular_mesh(dim, fm);
  } else {
    std::string mp = d.get("MeshPath") ? d.need("MeshPath")[0] : "./";
    if (!mp.empty() && mp[0] != '/') mp = d.dir + "/" + mp;
    q1 = medit_mesh(dim, mp + "/" + fm[0]);
  }
  std::set<int> drefs;
  for (const auto &v : d.all("ConditionReference")) {
    if (v.size() < 2) die("ConditionReference needs a reference and a condition");
    if (v.back() != "DIRICHLET") die("ConditionReference " + v.back() + ": only DIRICHLET");
    for (size_t k = 0; k + 1 < v.size(); ++k) drefs.insert((int)num(v[k]));
  }
  std::vector<int> pml_kind;
  if (pml_spec.on) pml_kind = add_pml(q1, pml_spec);
  pml_kind.resize(q1.elem_ref.size(), 0);
  const QrMesh qm = number_nodes(q1, order, drefs);
  const Media md = media(d, q1, elastic);
  if (first_order && !md.sigma.empty())
    for (double sg : md.sigma)
      if (sg != 0) die("MateriauDielec damping: not available with LEAP_FROG_PML");
  std::vector<uint8_t> pml_flag(pml_kind.size());
  for (size_t e = 0; e < pml_kind.size(); ++e) pml_flag[e] = pml_kind[e] != 0;
  std::vector<double> pml_tau_q;
  if (first_order) pml_tau_q = pml_tau(q1, pml_kind, qm, md.rho, md.mu, pml_spec, pml_damp);
  const NodalMedia nm = nodal_media(d, q1, qm, md, dim, elastic);
  if (!dump.empty()) {
    std::ofstream o(dump, std::ios::binary);
    const std::vector<int64_t> head = {dim, order, ncomp, qm.N, qm.E};
    wr(o, head);
    wr(o, qm.xyz);
    wr(o, qm.en);
    wr(o, qm.dir);
    wr(o, md.rho);
    wr(o, md.mu);
    wr(o, md.lam);
    // MateriauDielec's damping sigma per element (zeros when undamped)
    wr(o, md.sigma.empty() ? std::vector<double>((size_t)qm.E, 0.0) : md.sigma);
    wr(o, nm.rho);  // empty unless PhysicalMedia: E * nd each
    // (elastic: mu per element expanded to the quadrature points)
    std::vector<double> muq = nm.mu;
    if (muq.empty() && !nm.rho.empty()) {
      const int64_t nd = (int64_t)qm.en.size() / qm.E;
      muq.resize(nm.rho.size());
      for (int64_t e = 0; e < qm.E; ++e)
        for (int64_t q = 0; q < nd; ++q) muq[e * nd + q] = md.mu[e];
    }
    wr(o, muq);
    wr(o, montjoie_dofs(q1, qm));
    if (first_order) {
      // the layers: element flags, then sigma at every quadrature point
      std::ofstream op(dump + ".pml", std::ios::binary);
      wr(op, pml_flag);
      wr(op, pml_tau_q);
      if (!op) return 1;
    }
    return o ? 0 : 1;
  }

  const auto &tsv = d.need("TimeStep");
  const bool auto_dt = tsv[0] == "AUTO";
  const auto &ti = d.need("TimeInterval");
  const double t0 = ti.size() >= 2 ? num(ti[0]) : 0.0, t1 = num(ti.size() >= 2 ? ti[1] : ti[0]);
  const double f0 = num(d.need("Frequency")[0]);

  mjw_problem *p = nullptr;
  auto tc = std::chrono::steady_clock::now();
  check(mjw_create(&p, device, dim, order, ncomp, qm.N, qm.xyz.data(), qm.E, qm.en.data(),
                   md.rho.data(), md.mu.data(), elastic ? md.lam.data() : nullptr, qm.dir.data(),
                   1, precision),
        nullptr, "mjw_create");
  if (first_order)
    check(mjw_set_pml(p, pml_flag.data(), pml_tau_q.data()), p, "AddPML");
  if (!nm.rho.empty())
    check(mjw_set_media_nodal(p, nm.rho.data(), nm.mu.empty() ? nullptr : nm.mu.data()), p,
          "PhysicalMedia");
  if (!nm.sigma.empty())
    check(mjw_set_damping_nodal(p, nm.sigma.data()), p, "PhysicalMedia sigma");
  else if (!md.sigma.empty())
    check(mjw_set_damping(p, md.sigma.data()), p, "MateriauDielec sigma");
  if (time_order == 4) check(mjw_set_time_order(p, 4), p, "OrderTimeScheme");
  // time step: given, or TimeStep = AUTO: 0.999 times the leapfrog limit
  // c / sqrt(lambda_max(D^-1 K)) (VarInstationary.cxx:91-103,1229-1240,
  // EvaluateCFL_Generic, EvaluationCflWave.cxx:717-760), c = 2 for order 2
  // and 2 sqrt(3) for order 4 (ModifiedEquationIterator::GetCFL,
  // ModifiedEquation.cxx:217-225), lambda_max by Lanczos run until it
  // repeats to 1e-13 (ARPACK in the reference)
  double dt;
  if (auto_dt) {
    double lam = 0.0, prev = -1.0;
    const int64_t nfree = (int64_t)ncomp * qm.N;
    for (int it = 64; it <= 4096; it *= 2) {
      check(mjw_lambda_max(p, (int)std::min<int64_t>(it, std::max<int64_t>(nfree, 2)), 1, &lam), p,
            "TimeStep = AUTO");
      if (std::fabs(lam - prev) <= 1e-13 * lam || it >= nfree) break;
      prev = lam;
    }
    const double cfl = time_order == 4 ? 2.0 * std::sqrt(3.0) : 2.0;
    dt = 0.999 * cfl / std::sqrt(lam);
    std::printf("Maximal time step = %.15e\n", cfl / std::sqrt(lam));
  } else {
    dt = num(tsv[0]);
  }
  const int64_t nsteps = (int64_t)std::ceil((t1 - t0) / dt - 1e-9);  // RunSecondOrderScheme
  if (nsteps < 1) die("TimeInterval shorter than one TimeStep");
  // source
  if (volume_source) {
    std::vector<double> pol(ncomp, 1.0);
    if (elastic) {
      const auto &pv = d.need("Polarization");
      for (int a = 0; a < dim; ++a) pol[a] = num(pv.at(a));
    }
    const double pi = 3.14159265358979323846;
    const double alpha = vol_uniform ? 0.0 : -std::log(1.e-6) / (vol_r * vol_r);
    const double beta = std::sqrt(alpha / pi);
    const int64_t nd = qm.en.size() / qm.E;
    std::vector<double> fq((size_t)qm.E * nd, 0.0);
    for (int64_t e = 0; e < qm.E; ++e) {
      if (!vol_refs.empty() &&
          std::find(vol_refs.begin(), vol_refs.end(), q1.elem_ref[e]) == vol_refs.end())
        continue;
      for (int64_t q = 0; q < nd; ++q) {
        if (vol_uniform) {
          fq[e * nd + q] = 1.0;
          continue;
        }
        const double *xq = &qm.xyz[(size_t)qm.en[e * nd + q] * dim];
        double r2 = 0.0;
        for (int a = 0; a < dim; ++a) r2 += (xq[a] - vol_x[a]) * (xq[a] - vol_x[a]);
        const double r = std::sqrt(r2);
        if (r < vol_rc) fq[e * nd + q] = beta * std::exp(-alpha * (r * r));
      }
    }
    std::vector<double> b(qm.N), F((size_t)ncomp * qm.N);
    check(mjw_lumped_integral(p, fq.data(), b.data()), p, "SRC_VOLUME");
    for (int m = 0; m < ncomp; ++m)
      for (int64_t g = 0; g < qm.N; ++g) F[(size_t)m * qm.N + g] = pol[m] * b[g];
    check(mjw_set_source_vectors(p, F.data()), p, "SRC_VOLUME");
  } else {
    const auto &op = d.need("OriginePhase");
    if ((int)op.size() < dim) die("OriginePhase needs " + std::to_string(dim) + " coordinates");
    double x[3];
    for (int a = 0; a < dim; ++a) x[a] = num(op[a]);
    std::vector<double> pol(ncomp, 1.0);
    if (elastic) {
      const auto &pv = d.need("Polarization");
      for (int a = 0; a < dim; ++a) pol[a] = num(pv.at(a));
    }
    check(mjw_set_sources(p, 1, x, pol.data(), nullptr), p, "source");
  }
  std::vector<double> tab(nsteps);
  // the pulse at t_n + dt (ModifiedEquation.cxx:316); the first-order scheme
  // takes the primitive at t_n + dt/2 (ModifiedEquation.cxx:784-785)
  check(mjw_temporal_source_table(tkind, f0, tpar, t0, dt, (int)nsteps, first_order ? 0 : 1,
                                  first_order ? 0.5 * dt : dt, tab.data()),
        nullptr, "TemporalSource");
  check(mjw_set_pulse(p, (int)nsteps, tab.data()), p, "pulse");
  if (time_order == 4) {
    // f' and f'' at t_n + dt for the d^3U/dt^3 and d^4U/dt^4 terms
    std::vector<double> tab2(nsteps), tab3(nsteps);
    check(mjw_temporal_source_table(tkind, f0, tpar, t0, dt, (int)nsteps, 2, dt, tab2.data()),
          nullptr, "TemporalSource");
    check(mjw_temporal_source_table(tkind, f0, tpar, t0, dt, (int)nsteps, 3, dt, tab3.data()),
          nullptr, "TemporalSource");
    check(mjw_set_pulse_derivatives(p, (int)nsteps, tab2.data(), tab3.data()), p, "pulse");
  }
  // receivers: SismoPoint (SismoCircle is a snapshot grid, below)
  std::vector<double> pts;
  int npt = 0;
  for (const auto &v : d.all("SismoPoint")) {
    if ((int)v.size() < dim) die("SismoPoint needs " + std::to_string(dim) + " coordinates");
    for (int a = 0; a < dim; ++a) pts.push_back(num(v[a]));
    ++npt;
  }
  // a SismoPoint outside the mesh keeps its columns, written as 0 (the point
  // is never located, so its interpolation row stays empty:
  // Output/GridInterpolation.cxx TranslatePoint); only the points inside
  // become receivers, slot[r] = the point's index among them or -1
  std::vector<uint8_t> inside(npt, 1);
  std::vector<int> slot(npt, -1);
  if (npt) {
    check(mjw_set_probes(p, npt, pts.data(), inside.data()), p, "SismoPoint");
    check(mjw_set_probes(p, 0, nullptr, nullptr), p, "SismoPoint");
  }
  int nin = 0;
  for (int r = 0; r < npt; ++r)
    if (inside[r]) slot[r] = nin++;
  // receivers: every component of u at every point inside (component-major);
  // for elastic decks also d u_i / d x_a at every point (point-major, i
  // outer), from which the stress columns are formed
  std::vector<double> rpts;
  std::vector<int32_t> comp, deriv;
  for (int m = 0; m < ncomp; ++m)
    for (int r = 0; r < npt; ++r) {
      if (!inside[r]) continue;
      for (int a = 0; a < dim; ++a) rpts.push_back(pts[(size_t)r * dim + a]);
      comp.push_back(m);
      deriv.push_back(-1);
    }
  if (elastic)
    for (int r = 0; r < npt; ++r)
      if (inside[r])
        for (int m = 0; m < dim; ++m)
          for (int a = 0; a < dim; ++a) {
            for (int b = 0; b < dim; ++b) rpts.push_back(pts[(size_t)r * dim + b]);
            comp.push_back(m);
            deriv.push_back(a);
          }
  const int nrec = (int)comp.size();
  std::vector<int64_t> rec_elem(nrec, -1);
  if (nrec)
    check(mjw_set_receivers_grad(p, nrec, rpts.data(), comp.data(), deriv.data(),
                                 rec_elem.data()),
          p, "receivers");
  const double t_setup =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - tc).count();

  // snapshot grids (SismoPlane, SismoLine, SismoCircle): probe points, file names and the
  // steps each is written at (its own ParametersOutput* times and counter)
  struct Snap {
    Plane g;
    std::string name;
    std::vector<int64_t> steps;
    size_t off = 0;
    int count = 0;
  };
  std::vector<Snap> snaps;
  std::vector<double> probe_pts;
  std::vector<int32_t> probe_deriv;
  for (const char *kind : {"Plane", "Line", "Circle"}) {
    const std::string key = std::string("Sismo") + kind;
    const std::string k = kind;
    const auto lines = d.all(key);
    if (lines.empty()) continue;
    if (lines.size() > 1) die(key + ": one per deck");
    if (!out_binary && !out_ascii)
      die(key + ": only OutputFormat = BINARY or ASCII (DOUBLE or FLOAT) is written");
    Snap sn;
    sn.g = k == "Plane" ? plane_grid(lines[0], dim, qm)
           : k == "Line" ? line_grid(lines[0], dim) : circle_grid(lines[0], dim);
    const auto &fo = d.need(std::string("FileOutput") + kind);
    sn.name = fo.size() >= 2 ? fo[1] : fo[0];
    // the name's extension is dropped (GetBaseString, SELDON/share/Common.cxx:43-51,
    // applied by Output/OutputHarmonic.cxx:1254-1257); the step number and
    // ".dat" follow
    if (const size_t dot = sn.name.find_last_of('.'); dot != std::string::npos)
      sn.name = sn.name.substr(0, dot);
    double ps0 = t0, ps1 = t0 + nsteps * dt, psd = dt;
    if (const auto *po = d.get(std::string("ParametersOutput") + kind))
      if (po->size() >= 3) { ps0 = num((*po)[0]); ps1 = num((*po)[1]); psd = num((*po)[2]); }
    const int64_t every = std::max<int64_t>(1, std::llround(psd / dt));
    for (int64_t st = 0; st <= nsteps; ++st) {
      const double t = t0 + st * dt;
      if (t < ps0 - 1e-9 * dt || t > ps1 + 1e-9 * dt) continue;
      if (std::llround((t - ps0) / dt) % every) continue;
      sn.steps.push_back(st);
    }
    // elastic: the points again for d / d x_a, a < dim (stress)
    sn.off = probe_pts.size() / dim;
    for (int dv = -1; dv < (elastic ? dim : 0); ++dv) {
      probe_pts.insert(probe_pts.end(), sn.g.pts.begin(), sn.g.pts.end());
      probe_deriv.insert(probe_deriv.end(), sn.g.pts.size() / dim, dv);
    }
    snaps.push_back(sn);
  }
  std::set<int64_t> snap_steps;
  for (const auto &sn : snaps) snap_steps.insert(sn.steps.begin(), sn.steps.end());
  std::vector<int64_t> probe_elem(probe_deriv.size(), -1);
  if (!probe_pts.empty())
    check(mjw_set_probes_grad(p, (int)(probe_pts.size() / dim), probe_pts.data(),
                              probe_deriv.data(), nullptr, probe_elem.data()),
          p, "snapshot grid");
  std::vector<double> pv(probe_deriv.size() * ncomp);
  // acoustic: <name>%04d.dat; elastic (OutputFormat's multi-field layout):
  // <name>%04d_U<m>.dat per displacement component and <name>%04d_dU<k>.dat
  // per stress component sigma_ij, k = d i + j, with the medium of the
  // element holding the point (0 outside the mesh)
  auto snapshot = [&](int64_t st) {
    check(mjw_probe(p, pv.data()), p, "snapshot grid");
    for (auto &sn : snaps) {
      if (!std::binary_search(sn.steps.begin(), sn.steps.end(), st)) continue;
      const size_t n = (size_t)sn.g.nx * sn.g.ny;
      char num[16];
      std::snprintf(num, sizeof num, "%04d", sn.count++);
      std::vector<double> vals(n);
      if (!elastic) {
        for (size_t i = 0; i < n; ++i) vals[i] = pv[sn.off + i];
        write_plane(sn.name + num + ".dat", sn.g, vals, out_double, out_binary);
        continue;
      }
      for (int m = 0; m < dim; ++m) {
        for (size_t i = 0; i < n; ++i) vals[i] = pv[(sn.off + i) * dim + m];
        write_plane(sn.name + num + "_U" + std::to_string(m) + ".dat", sn.g, vals, out_double,
                    out_binary);
      }
      // gradient probes: sn.off + (1 + a) n + i holds d u_m / d x_a at point i
      auto grad = [&](size_t i, int m, int a) {
        return pv[(sn.off + (1 + a) * n + i) * dim + m];
      };
      for (int a = 0; a < dim; ++a)
        for (int b = 0; b < dim; ++b) {
          for (size_t i = 0; i < n; ++i) {
            const int64_t e = probe_elem[sn.off + n + i];
            if (e < 0) { vals[i] = 0.0; continue; }
            if (!display_stress) { vals[i] = grad(i, a, b); continue; }
            double div = 0.0;
            for (int c = 0; c < dim; ++c) div += grad(i, c, c);
            vals[i] = (a == b ? md.lam[e] * div : 0.0) + md.mu[e] * (grad(i, a, b) + grad(i, b, a));
          }
          write_plane(sn.name + num + "_dU" + std::to_string(dim * a + b) + ".dat", sn.g, vals,
                      out_double, out_binary);
        }
    }
  };

  // SaveReprise (SaveDataReprise, VarInstationary.cxx:917-972): at every step
  // nt with nt % n == 0, the final one included, U^nt in Montjoie's numbering
  // goes to SolAcous<b>_P0N1.dat, b alternating 0, 1, 0, ..., and
  // IterationAcousSaveNew0N1.dat = [nt, -1, b] after the previous one has been
  // copied to IterationAcousSave0N1.dat: a restart reads IterationAcousSave
  // (a complete SolAcous file) and the New file only when that one is damaged.
  std::vector<int64_t> mj_dof;
  if (save_every) mj_dof = montjoie_dofs(q1, qm);
  int bascule = -1;
  std::vector<double> un, ux;
  auto save_reprise = [&](int64_t st) {
    const std::string suffix = "0N1.dat";
    const std::vector<int32_t> it = {(int32_t)st, -1, bascule < 0 ? 0 : bascule};
    if (bascule < 0) {
      bascule = 0;
    } else {
      std::ifstream in(path_reprise + "IterationAcousSaveNew" + suffix, std::ios::binary);
      std::vector<char> old((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
      std::ofstream o(path_reprise + "IterationAcousSave" + suffix, std::ios::binary);
      o.write(old.data(), (std::streamsize)old.size());
      if (!o) die("cannot write " + path_reprise + "IterationAcousSave" + suffix);
    }
    write_vector(path_reprise + "IterationAcousSaveNew" + suffix, it);
    un.resize((size_t)ncomp * qm.N);
    ux.resize(un.size());
    check(mjw_get_state(p, un.data(), nullptr), p, "SaveReprise");
    for (int m = 0; m < ncomp; ++m)
      for (int64_t g = 0; g < qm.N; ++g) ux[(size_t)m * qm.N + mj_dof[g]] = un[(size_t)m * qm.N + g];
    write_vector(path_reprise + "SolAcous" + std::to_string(bascule) + "_P" + suffix, ux);
    bascule = 1 - bascule;
  };
  std::set<int64_t> stops = snap_steps;
  if (save_every)
    for (int64_t st = 0; st <= nsteps; st += save_every) stops.insert(st);

  // RunSecondOrderScheme: U^0 = U^1 = 0, rows n = 0..nsteps
  check(mjw_leapfrog_init(p, t0, dt, nullptr, nullptr), p, "leapfrog_init");
  std::vector<double> tr((size_t)(nsteps + 1) * (nrec ? nrec : 1));
  auto ts = std::chrono::steady_clock::now();
  int64_t done = 0;
  for (const int64_t s : stops) {
    if (s > done) {
      check(mjw_leapfrog_advance(p, (int)(s - done), nrec ? tr.data() + (size_t)done * nrec : nullptr,
                                 1),
            p, "advance");
      done = s;
    }
    if (save_every && s % save_every == 0) save_reprise(s);
    if (snap_steps.count(s)) snapshot(s);
  }
  if (nsteps > done)
    check(mjw_leapfrog_advance(p, (int)(nsteps - done),
                               nrec ? tr.data() + (size_t)done * nrec : nullptr, 1),
          p, "advance");
  if (nrec) check(mjw_record(p, tr.data() + (size_t)nsteps * nrec), p, "record");
  const double t_loop = std::chrono::duration<double>(std::chrono::steady_clock::now() - ts).count();

  // seismograms
  if (npt && d.get("FileOutputPoint")) {
    const auto &fo = d.need("FileOutputPoint");
    const std::string out = fo.size() >= 2 ? fo[1] : fo[0];
    double ts0 = t0, ts1 = t0 + nsteps * dt, tsd = dt;
    if (const auto *po = d.get("ParametersOutputPoint")) {
      if (po->size() >= 3) { ts0 = num((*po)[0]); ts1 = num((*po)[1]); tsd = num((*po)[2]); }
    }
    const int64_t every = std::max<int64_t>(1, std::llround(tsd / dt));
    FILE *f = std::fopen(out.c_str(), "w");
    if (!f) die("cannot write " + out);
    for (int64_t s = 0; s <= nsteps; ++s) {
      const double t = t0 + s * dt;
      if (t < ts0 - 1e-9 * dt || t > ts1 + 1e-9 * dt) continue;
      if (std::llround((t - ts0) / dt) % every) continue;
      const char *fmt = out_double ? "\t%.15e" : "\t%.7e";
      std::fprintf(f, fmt + 1, t);
      // point-major: the components of point 0, then of point 1, ...
      for (int r = 0; r < npt; ++r)
        for (int m = 0; m < ncomp; ++m)
          std::fprintf(f, fmt, slot[r] < 0 ? 0.0 : tr[(size_t)s * nrec + (size_t)m * nin + slot[r]]);
      // elastic: then the stress C grad u of each point, d*d values row-major
      // (sigma_ij = lambda tr(grad u) delta_ij + mu (du_i/dx_j + du_j/dx_i),
      // ModifyOutputUnknown, Elliptic/Elastic/VarElastic.cxx:257-282), with
      // the medium of the element holding the point
      if (elastic)
        for (int r = 0; r < npt; ++r) {
          const size_t dd = (size_t)dim * dim;
          if (slot[r] < 0) {
            for (size_t k = 0; k < dd; ++k) std::fprintf(f, fmt, 0.0);
            continue;
          }
          const double *g = &tr[(size_t)s * nrec + (size_t)dim * nin + (size_t)slot[r] * dd];
          const int64_t e = rec_elem[(size_t)dim * nin + (size_t)slot[r] * dd];
          const double lam = md.lam[e], mu = md.mu[e];
          double div = 0.0;
          for (int c = 0; c < dim; ++c) div += g[c * dim + c];
          for (int i = 0; i < dim; ++i)
            for (int j = 0; j < dim; ++j)
              std::fprintf(f, fmt,
                           !display_stress ? g[i * dim + j]
                                           : (i == j ? lam * div : 0.0) +
                                                 mu * (g[i * dim + j] + g[j * dim + i]));
        }
      std::fprintf(f, "\n");
    }
    std::fclose(f);
  }
  std::printf("mjwave_run: %s %dD Q%d, %lld nodes, %lld elements, %lld steps\n",
              elastic ? "elastic" : "acoustic", dim, order, (long long)qm.N, (long long)qm.E,
              (long long)nsteps);
  std::printf("Time to set up the problem : %.6f\n", t_setup);
  std::printf("Time to complete time iterations : %.6f\n", t_loop);
  mjw_destroy(p);
  return 0;
}
