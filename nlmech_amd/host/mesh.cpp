// Mesh reading for the finite-difference model (see mesh.h).
#include "mesh.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <unordered_map>

#include "vtu.h"

namespace nlm {
namespace fe {

namespace {

[[noreturn]] void error(const std::string& m) { throw inp::DeckError(m); }

// definitelyLessThan (src/util/compare.cpp:30-33)
bool less_than(double a, double b) {
  const double m = std::abs(a) < std::abs(b) ? std::abs(b) : std::abs(a);
  return (b - a) > m * 1e-5;
}

double dist(const double* a, const double* b) {
  const double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  return std::sqrt(dx * dx + dy * dy + dz * dz);
}

std::string ext(const std::string& f) {
  size_t p = f.find_last_of('.');
  return p == std::string::npos ? std::string() : f.substr(p + 1);
}

// rw::reader::readCsvFile (src/rw/reader.cpp:15-60): header with columns
// id, x[, y[, z]], volume (other columns ignored).
void read_csv(const std::string& file, size_t dim, std::vector<double>* x,
              std::vector<double>* vol) {
  std::ifstream f(file);
  if (!f) error("Error: cannot open mesh file " + file);
  std::string line;
  if (!std::getline(f, line)) error("Error: empty mesh file " + file);
  std::vector<std::string> head;
  {
    std::stringstream ss(line);
    std::string c;
    while (std::getline(ss, c, ',')) {
      c.erase(std::remove_if(c.begin(), c.end(), ::isspace), c.end());
      head.push_back(c);
    }
  }
  const char* need[4] = {"x", "y", "z", "volume"};
  int col[4] = {-1, -1, -1, -1};
  for (int k = 0; k < 4; k++)
    for (size_t j = 0; j < head.size(); j++)
      if (head[j] == need[k]) col[k] = int(j);
  if (col[0] < 0 || col[3] < 0 || (dim > 1 && col[1] < 0) ||
      (dim > 2 && col[2] < 0))
    error("Error: mesh file " + file + " lacks a required column");
  while (std::getline(f, line)) {
    if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
    std::vector<double> v;
    std::stringstream ss(line);
    std::string c;
    while (std::getline(ss, c, ',')) v.push_back(std::atof(c.c_str()));
    auto at = [&](int k) { return size_t(col[k]) < v.size() ? v[col[k]] : 0.; };
    x->push_back(at(0));
    x->push_back(dim > 1 ? at(1) : 0.);
    x->push_back(dim > 2 ? at(2) : 0.);
    vol->push_back(at(3));
  }
}

// rw::reader::MshReader::readMesh (src/rw/mshReader.cpp:20-114) reads a gmsh
// mesh through the gmsh library: every node at index tag - 1, and of the
// elements the 3-node triangles if the mesh has any, else the 4-node
// quadrangles (a mesh with both keeps the triangles after an error message;
// one with neither is an error), connectivity in the order the file lists
// them.  This reads the same from the ASCII formats 2.2 and 4.1 directly
// (binary files are refused).
void read_msh(const std::string& file, std::vector<double>* x,
              std::vector<int64_t>* enc, int* etype, size_t* nelems) {
  std::ifstream f(file);
  if (!f) error("Error: cannot open mesh file " + file);
  auto bad = [&](const std::string& what) {
    error("Error: mesh file " + file + ": " + what);
  };
  std::string line;
  auto next = [&]() -> std::string {
    if (!std::getline(f, line)) bad("unexpected end of file");
    if (!line.empty() && line.back() == '\r') line.pop_back();
    return line;
  };
  auto skip_to = [&](const std::string& end) {
    while (next() != end) {
    }
  };
  double version = 0.;
  bool have_nodes = false, have_elems = false;
  // counts read from the file are checked before they size anything
  const int64_t kMaxCount = int64_t(1) << 36;
  std::vector<int64_t> tri, quad;
  x->clear();
  auto put_node = [&](int64_t tag, double px, double py, double pz) {
    if (tag < 1) bad("node tag < 1");
    if (size_t(tag) * 3 > x->size()) bad("node tag beyond the node count");
    (*x)[3 * (tag - 1)] = px;
    (*x)[3 * (tag - 1) + 1] = py;
    (*x)[3 * (tag - 1) + 2] = pz;
  };
  // one element: gmsh type, node tags (1-based) from the stream
  auto put_elem = [&](int type, std::istringstream& ss) {
    std::vector<int64_t>* dst = type == 2 ? &tri : type == 3 ? &quad : nullptr;
    if (!dst) return;  // points, lines, other types: not used
    const int nv = type == 2 ? 3 : 4;
    for (int k = 0; k < nv; k++) {
      int64_t t = 0;
      if (!(ss >> t) || t < 1) bad("bad element connectivity");
      dst->push_back(t - 1);
    }
  };
  while (std::getline(f, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (line == "$MeshFormat") {
      std::istringstream ss(next());
      int ftype = -1, dsize = 0;
      ss >> version >> ftype >> dsize;
      if (ftype != 0) bad("binary gmsh files are not supported");
      if (!(version >= 2. && version < 3.) && !(version >= 4. && version < 5.))
        bad("gmsh format version " + std::to_string(version) +
            " is not supported (2.2 or 4.1)");
      skip_to("$EndMeshFormat");
    } else if (line == "$Nodes") {
      if (version == 0.) bad("$Nodes before $MeshFormat");
      if (version < 3.) {
        int64_t n = 0;
        std::istringstream(next()) >> n;
        if (n <= 0) bad("no nodes");
        if (n > kMaxCount) bad("node count out of range");
        x->assign(size_t(3 * n), 0.);
        for (int64_t i = 0; i < n; i++) {
          std::istringstream ss(next());
          int64_t tag;
          double px = 0, py = 0, pz = 0;
          if (!(ss >> tag >> px >> py >> pz)) bad("bad node line");
          put_node(tag, px, py, pz);
        }
      } else {
        int64_t nblk = 0, n = 0, tmin = 0, tmax = 0;
        std::istringstream(next()) >> nblk >> n >> tmin >> tmax;
        if (n <= 0) bad("no nodes");
        if (n > kMaxCount) bad("node count out of range");
        if (tmax != n) bad("node tags are not 1..N");
        x->assign(size_t(3 * n), 0.);
        for (int64_t b = 0; b < nblk; b++) {
          int edim = 0, etag = 0, param = 0;
          int64_t nb = 0;
          std::istringstream(next()) >> edim >> etag >> param >> nb;
          if (nb < 0 || nb > n) bad("bad node block size");
          std::vector<int64_t> tags(static_cast<size_t>(nb));
          for (int64_t k = 0; k < nb; k++)
            std::istringstream(next()) >> tags[size_t(k)];
          for (int64_t k = 0; k < nb; k++) {
            std::istringstream ss(next());
            double px = 0, py = 0, pz = 0;
            if (!(ss >> px >> py >> pz)) bad("bad node coordinates");
            put_node(tags[size_t(k)], px, py, pz);
          }
        }
      }
      skip_to("$EndNodes");
      have_nodes = true;
    } else if (line == "$Elements") {
      if (version == 0.) bad("$Elements before $MeshFormat");
      if (version < 3.) {
        int64_t n = 0;
        std::istringstream(next()) >> n;
        for (int64_t i = 0; i < n; i++) {
          std::istringstream ss(next());
          int64_t tag = 0;
          int type = 0, ntags = 0;
          if (!(ss >> tag >> type >> ntags)) bad("bad element line");
          if (ntags < 0 || ntags > 64) bad("bad element tag count");
          for (int k = 0; k < ntags; k++) {
            int64_t t;
            ss >> t;
          }
          put_elem(type, ss);
        }
      } else {
        int64_t nblk = 0, n = 0, tmin = 0, tmax = 0;
        std::istringstream(next()) >> nblk >> n >> tmin >> tmax;
        for (int64_t b = 0; b < nblk; b++) {
          int edim = 0, etag = 0, type = 0;
          int64_t nb = 0;
          std::istringstream(next()) >> edim >> etag >> type >> nb;
          for (int64_t k = 0; k < nb; k++) {
            std::istringstream ss(next());
            int64_t tag = 0;
            ss >> tag;
            put_elem(type, ss);
          }
        }
      }
      skip_to("$EndElements");
      have_elems = true;
    }
  }
  if (!have_nodes) bad("no $Nodes section");
  if (!have_elems) bad("no $Elements section");
  if (!tri.empty() && !quad.empty())
    std::fprintf(stderr,
                 "Error: Only mesh with one type of elements is supported!\n");
  if (tri.empty() && quad.empty())
    error("Error: Only 3-node triangle or 4-node square elements are "
          "supported!");
  const size_t nn = x->size() / 3;
  const bool use_tri = !tri.empty();
  *enc = use_tri ? std::move(tri) : std::move(quad);
  for (int64_t c : *enc)
    if (size_t(c) >= nn) bad("element refers to a missing node");
  *etype = use_tri ? rw::kVtkTriangle : rw::kVtkQuad;
  *nelems = enc->size() / (use_tri ? 3 : 4);
}

}  // namespace

Mesh::Mesh(const inp::MeshDeck& deck) : d_dim(deck.d_dim), d_h(deck.d_h) {
  const std::string& sd = deck.d_spatialDiscretization;
  if (sd != "finite_difference")
    error("Error: Spatial discretization type not known. Check input data.");
  if (d_dim < 1 || d_dim > 3) error("Error: Check Dimension in input data.");
  if (deck.d_filename.empty())
    error("Error: Filename for mesh data not specified.");
  const std::string e = ext(deck.d_filename);
  if (e == "csv") {
    read_csv(deck.d_filename, d_dim, &d_nodes, &d_vol);
  } else if (e == "vtu") {
    try {
      rw::VtuReader r(deck.d_filename);
      if (!r.readNodes(&d_nodes))
        error("Error: no points in mesh file " + deck.d_filename);
      bool found = r.readPointData("Node_Volume", &d_vol);
      if (!found) found = r.readPointData("Volume", &d_vol);
      if (!found || deck.d_keepElementConn) {
        std::vector<int64_t> off;
        std::vector<uint8_t> types;
        if (r.readCells(&d_enc, &off, &types) && !types.empty()) {
          d_eType = types[0];
          d_numElems = types.size();
          for (uint8_t t : types)
            if (t != types[0])
              error("Error: mesh file " + deck.d_filename +
                    " mixes element types");
        }
      }
      if (!found) d_vol.clear();
    } catch (const std::runtime_error& ex) {
      if (dynamic_cast<const inp::DeckError*>(&ex)) throw;
      error(std::string("Error: ") + ex.what());
    }
  } else if (e == "msh") {
    // the reference reads 2D triangle / quadrangle meshes; a 3D mesh would
    // yield its surface triangles
    if (d_dim != 2)
      error("Error: MshReader currently only supports reading of "
            "triangle/quadrangle elements in dimension 2.");
    read_msh(deck.d_filename, &d_nodes, &d_enc, &d_eType, &d_numElems);
    d_vol.clear();
  } else {
    error("Error: unknown mesh file type " + deck.d_filename);
  }
  if (d_vol.empty()) computeVol();
  if (deck.d_computeMeshSize) computeMeshSize();
  for (size_t i = 0; i < d_vol.size(); i++)
    if (d_vol[i] < 0.01 * std::pow(d_h, double(d_dim))) {
      std::ostringstream m;
      m << "Error: Check nodal volume " << d_vol[i] << " is less than "
        << 0.01 * std::pow(d_h, double(d_dim)) << ", Node = " << i;
      error(m.str());
    }
  // (the reference keeps a gmsh mesh's elements: fe::Mesh::createData,
  // mesh.cpp:127-129)
  if (!deck.d_keepElementConn && e != "msh") {
    d_enc.clear();
    d_numElems = 0;
  }
  computeBBox();
}

Mesh::Mesh(size_t dim, std::vector<double> nodes, std::vector<double> vol)
    : d_dim(dim), d_nodes(std::move(nodes)), d_vol(std::move(vol)) {
  if (d_nodes.size() != 3 * d_vol.size())
    error("Error: node and volume counts differ");
  computeBBox();
  computeMeshSize();
}

// fe::Mesh::computeVol (mesh.cpp:228-309): V_i = sum over the elements of
// node i of the integral of its shape function (triangles: area / 3,
// tetrahedra: volume / 4, quadrilaterals: 2x2 Gauss rule).
void Mesh::computeVol() {
  const int nv = rw::vtk_cell_vertices(d_eType);
  if (d_numElems == 0 || nv == 0)
    error("Error: Can not compute nodal volume for given finite element mesh "
          "as the element-node connectivity data is invalid.");
  d_vol.assign(d_nodes.size() / 3, 0.);
  const double* X = d_nodes.data();
  for (size_t e = 0; e < d_numElems; e++) {
    const int64_t* c = &d_enc[e * nv];
    if (d_eType == rw::kVtkTriangle) {
      const double* a = X + 3 * c[0];
      const double* b = X + 3 * c[1];
      const double* d = X + 3 * c[2];
      const double area =
          0.5 * std::abs((b[0] - a[0]) * (d[1] - a[1]) -
                         (d[0] - a[0]) * (b[1] - a[1]));
      for (int k = 0; k < 3; k++) d_vol[c[k]] += area / 3.;
    } else if (d_eType == rw::kVtkQuad) {
      const double g = 1. / std::sqrt(3.);
      const double gp[2] = {-g, g};
      const double xi[4] = {-1, 1, 1, -1}, eta[4] = {-1, -1, 1, 1};
      for (double s : gp)
        for (double t : gp) {
          double jx[2] = {0, 0}, jy[2] = {0, 0};
          for (int k = 0; k < 4; k++) {
            const double* p = X + 3 * c[k];
            const double dxi = 0.25 * xi[k] * (1 + eta[k] * t);
            const double deta = 0.25 * eta[k] * (1 + xi[k] * s);
            jx[0] += dxi * p[0];
            jx[1] += deta * p[0];
            jy[0] += dxi * p[1];
            jy[1] += deta * p[1];
          }
          const double det = std::abs(jx[0] * jy[1] - jx[1] * jy[0]);
          for (int k = 0; k < 4; k++)
            d_vol[c[k]] += 0.25 * (1 + xi[k] * s) * (1 + eta[k] * t) * det;
        }
    } else if (d_eType == rw::kVtkTetra) {
      const double* a = X + 3 * c[0];
      double m[3][3];
      for (int k = 0; k < 3; k++)
        for (int j = 0; j < 3; j++) m[k][j] = X[3 * c[k + 1] + j] - a[j];
      const double v =
          std::abs(m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
                   m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                   m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0])) /
          6.;
      for (int k = 0; k < 4; k++) d_vol[c[k]] += v / 4.;
    } else {
      error("Error: nodal volumes need triangle, quadrilateral or "
            "tetrahedral elements");
    }
  }
}

// fe::Mesh::computeMeshSize (mesh.cpp:326-350): starting from |x0 - x1| the
// guess is replaced by every pair distance that is definitelyLessThan it,
// pairs visited in (i, j) order.  Only pairs closer than the first guess can
// change it, so those are collected with a cell grid of that size and
// replayed in the same order.
void Mesh::computeMeshSize() {
  const size_t n = d_vol.size();
  if (n < 2) {
    d_h = 0.;
    return;
  }
  const double* X = d_nodes.data();
  double guess = dist(X, X + 3);
  const double cell = guess > 0. ? guess : 1.;
  double lo[3] = {X[0], X[1], X[2]};
  for (size_t i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) lo[d] = std::min(lo[d], X[3 * i + d]);
  auto key = [&](int64_t a, int64_t b, int64_t c) {
    return (a * 73856093) ^ (b * 19349663) ^ (c * 83492791);
  };
  auto cix = [&](size_t i, int d) {
    return int64_t(std::floor((X[3 * i + d] - lo[d]) / cell));
  };
  std::unordered_multimap<int64_t, size_t> grid;
  grid.reserve(n);
  for (size_t i = 0; i < n; i++)
    grid.emplace(key(cix(i, 0), cix(i, 1), cix(i, 2)), i);
  std::vector<std::pair<size_t, size_t>> cand;
  for (size_t i = 0; i < n; i++) {
    const int64_t a = cix(i, 0), b = cix(i, 1), c = cix(i, 2);
    for (int64_t da = -1; da <= 1; da++)
      for (int64_t db = -1; db <= 1; db++)
        for (int64_t dc = -1; dc <= 1; dc++) {
          auto r = grid.equal_range(key(a + da, b + db, c + dc));
          for (auto it = r.first; it != r.second; ++it) {
            const size_t j = it->second;
            if (j == i) continue;
            if (cix(j, 0) != a + da || cix(j, 1) != b + db ||
                cix(j, 2) != c + dc)
              continue;  // hash collision
            if (dist(X + 3 * i, X + 3 * j) < guess) cand.emplace_back(i, j);
          }
        }
  }
  std::sort(cand.begin(), cand.end());
  for (const auto& p : cand) {
    const double v = dist(X + 3 * p.first, X + 3 * p.second);
    if (less_than(v, guess)) guess = v;
  }
  d_h = guess;
}

void Mesh::computeBBox() {
  std::vector<double> p1(3, 0.), p2(3, 0.);
  for (size_t i = 0; i < d_vol.size(); i++)
    for (int d = 0; d < 3; d++) {
      const double x = d_nodes[3 * i + d];
      if (less_than(x, p1[d])) p1[d] = x;
      if (less_than(p2[d], x)) p2[d] = x;
    }
  d_bbox = std::make_pair(p1, p2);
}

}  // namespace fe
}  // namespace nlm
