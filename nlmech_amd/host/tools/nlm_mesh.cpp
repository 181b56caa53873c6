// nlm_mesh: uniform meshes for the finite-difference model, as the
// reference's mesh tool writes them (tools/mesh/mesh.cpp;
// tools/mesh/src/1d/fe1D.cpp:98-130; tools/mesh/src/2d/fe2D.cpp:115-576):
// "nlm_mesh -i input_mesh.yaml -d 2".  Node j (nx + 1) + i sits at
// (x0 + i h, y0 + j h); volumes h^dim, halved on each boundary.
#include <cmath>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "vtu.h"
#include "yaml.h"

namespace {

using nlm::yaml::Node;

bool less_than(double a, double b) {
  const double m = std::abs(a) < std::abs(b) ? std::abs(b) : std::abs(a);
  return (b - a) > m * 1e-5;
}

struct InpData {
  std::string mesh_file, out_format = "vtu", mesh_type = "uniform_tri",
                         compress;
  std::vector<double> domain;
  double horizon = 0., h = 0.;
  size_t r = 1;
  bool is_fd = false;
};

// readInputFile (fe2D.cpp:41-113, fe1D.cpp:20-96)
InpData read(const Node& c) {
  InpData d;
  std::string path = "./";
  if (c["Output"]) {
    if (c["Output"]["Path"]) path = c["Output"]["Path"].str() + "/";
    d.mesh_file = path + (c["Output"]["Mesh"] ? c["Output"]["Mesh"].str()
                                              : std::string("mesh"));
    if (c["Output"]["File_Format"])
      d.out_format = c["Output"]["File_Format"].str();
  } else {
    d.mesh_file = path + "mesh";
  }
  if (!c["Domain"])
    throw nlm::yaml::Error("Domain data is not provided in mesh input file.");
  d.domain = c["Domain"].as_doubles();
  if (!c["Horizon"])
    throw nlm::yaml::Error("Horizon is not provided in mesh input file.");
  d.horizon = c["Horizon"].as_double();
  if (c["Horizon_h_Ratio"]) {
    d.r = c["Horizon_h_Ratio"].as_size();
    d.h = d.horizon / double(d.r);
  }
  if (c["Mesh_Size"])
    d.h = c["Mesh_Size"].as_double();
  else if (!c["Horizon_h_Ratio"])
    throw nlm::yaml::Error(
        "Can not calculate mesh size. Either provide Horizon_h_Ratio or "
        "Mesh_Size.");
  if (c["Mesh_Type"]) d.mesh_type = c["Mesh_Type"].str();
  if (c["Is_FD"]) d.is_fd = c["Is_FD"].as_bool();
  if (c["Compress_Type"]) d.compress = c["Compress_Type"].str();
  if (d.out_format != "vtu")
    throw nlm::yaml::Error("only File_Format vtu is supported");
  return d;
}

void write(const InpData& d, const std::vector<double>& nodes,
           const std::vector<double>& vol, int type,
           const std::vector<int64_t>& conn) {
  nlm::rw::VtuWriter w(d.mesh_file, d.compress);
  w.appendNodes(nodes, nullptr);
  if (type > 0) w.appendCells(type, conn);
  w.appendPointData("Node_Volume", vol.data(), 1);
  w.addTimeStep(0.);
  w.close();
}

// domain end trimmed to whole cells (fe2D.cpp:124-140)
void trim(double first, double* second, double cell) {
  size_t ba_h = size_t((*second - first) / cell);
  double B = first + double(ba_h) * cell;
  if (less_than(B, *second)) *second = first + double(ba_h) * cell;
}

void fe1d(const InpData& d) {
  if (d.domain.size() < 2) throw nlm::yaml::Error("Domain needs 2 values");
  const size_t ne = size_t((d.domain[1] - d.domain[0]) / d.h);
  const size_t nn = ne + 1;
  std::vector<double> x(3 * nn, 0.), vol(nn, d.h);
  vol[0] *= 0.5;
  vol[nn - 1] *= 0.5;
  std::vector<int64_t> conn(2 * ne);
  for (size_t i = 0; i < nn; i++) x[3 * i] = d.domain[0] + double(i) * d.h;
  for (size_t i = 0; i < ne; i++) {
    conn[2 * i] = int64_t(i);
    conn[2 * i + 1] = int64_t(i + 1);
  }
  write(d, x, vol, nlm::rw::kVtkLine, conn);
}

void fe2d(const InpData& d0) {
  InpData d = d0;
  if (d.domain.size() < 4) throw nlm::yaml::Error("Domain needs 4 values");
  double x0 = d.domain[0], y0 = d.domain[1], x1 = d.domain[2],
         y1 = d.domain[3];
  const double h = d.h;
  if (d.mesh_type == "uniform_square" || d.mesh_type == "uniform_tri") {
    trim(x0, &x1, h);
    trim(y0, &y1, h);
    const size_t nx = size_t((x1 - x0) / h), ny = size_t((y1 - y0) / h);
    const size_t nn = (nx + 1) * (ny + 1);
    std::vector<double> x(3 * nn, 0.), vol(nn, h * h);
    for (size_t j = 0; j <= ny; j++)
      for (size_t i = 0; i <= nx; i++) {
        const size_t n = j * (nx + 1) + i;
        x[3 * n] = x0 + double(i) * h;
        x[3 * n + 1] = y0 + double(j) * h;
        if (i == 0 || i == nx) vol[n] *= 0.5;
        if (j == 0 || j == ny) vol[n] *= 0.5;
      }
    if (d.is_fd) return write(d, x, vol, 0, {});
    std::vector<int64_t> conn;
    if (d.mesh_type == "uniform_square") {
      conn.resize(4 * nx * ny);
      for (size_t j = 0; j < ny; j++)
        for (size_t i = 0; i < nx; i++) {
          const size_t n = j * nx + i;
          conn[4 * n + 0] = int64_t(j * (nx + 1) + i);
          conn[4 * n + 1] = int64_t(j * (nx + 1) + i + 1);
          conn[4 * n + 2] = int64_t((j + 1) * (nx + 1) + i + 1);
          conn[4 * n + 3] = int64_t((j + 1) * (nx + 1) + i);
        }
      return write(d, x, vol, nlm::rw::kVtkQuad, conn);
    }
    conn.resize(6 * nx * ny);
    for (size_t j = 0; j < ny; j++)
      for (size_t i = 0; i < nx; i++) {
        const int64_t n1 = int64_t(j * (nx + 1) + i), n2 = n1 + 1,
                      n3 = int64_t((j + 1) * (nx + 1) + i), n4 = n3 + 1;
        const size_t T1 = j * 2 * nx + 2 * i, T2 = T1 + 1;
        const int64_t a[6] = {n1, n4, n3, n1, n2, n4};
        const int64_t b[6] = {n1, n2, n3, n2, n4, n3};
        const int64_t* t = i % 2 == 0 ? a : b;
        for (int k = 0; k < 3; k++) {
          conn[3 * T1 + k] = t[k];
          conn[3 * T2 + k] = t[3 + k];
        }
      }
    return write(d, x, vol, nlm::rw::kVtkTriangle, conn);
  }
  if (d.mesh_type == "uniform_tri_sym") {
    // fe2D.cpp:340-557: cell corners plus one node at each cell centre
    const double hc = 2. * h;
    trim(x0, &x1, hc);
    trim(y0, &y1, hc);
    const size_t nx = size_t((x1 - x0) / hc), ny = size_t((y1 - y0) / hc);
    const size_t nn = (nx + 1) * (ny + 1) + nx * ny;
    std::vector<double> x(3 * nn, 0.), vol(nn, 2. * h * h);
    for (size_t j = 0; j <= ny; j++)
      for (size_t i = 0; i <= nx; i++) {
        size_t n = j * (nx + 1 + nx) + 2 * i;
        if (j == ny) n = j * (nx + 1 + nx) + i;
        x[3 * n] = x0 + double(i) * hc;
        x[3 * n + 1] = y0 + double(j) * hc;
        if (i == 0 || i == nx) vol[n] *= 0.5;
        if (j == 0 || j == ny) vol[n] *= 0.5;
        if (i < nx && j < ny) {
          n += 1;
          x[3 * n] = x0 + double(i) * hc + h;
          x[3 * n + 1] = y0 + double(j) * hc + h;
        }
      }
    if (d.is_fd) return write(d, x, vol, 0, {});
    std::vector<int64_t> conn(12 * nx * ny);
    for (size_t j = 0; j < ny; j++)
      for (size_t i = 0; i < nx; i++) {
        int64_t ns[5];
        ns[0] = int64_t(j * (nx + 1 + nx) + 2 * i);
        ns[1] = ns[0] + 1;
        ns[2] = ns[0] + 2;
        if (j + 1 < ny) {
          ns[4] = int64_t((j + 1) * (nx + 1 + nx) + 2 * i);
          ns[3] = ns[4] + 2;
        } else {
          ns[4] = int64_t((j + 1) * (nx + 1 + nx) + i);
          ns[3] = ns[4] + 1;
        }
        const int tri[4][3] = {{0, 1, 4}, {0, 2, 1}, {1, 2, 3}, {1, 3, 4}};
        for (int q = 0; q < 4; q++) {
          const size_t T = j * 4 * nx + 4 * i + q;
          for (int k = 0; k < 3; k++) conn[3 * T + k] = ns[tri[q][k]];
        }
      }
    return write(d, x, vol, nlm::rw::kVtkTriangle, conn);
  }
  throw nlm::yaml::Error(
      "Check Mesh_Type data. Currently only uniform_square, uniform_tri and "
      "uniform_tri_sym is implemented.");
}

}  // namespace

int main(int argc, char** argv) {
  std::string file;
  size_t dim = 0;
  int found = 0;
  for (int k = 1; k + 1 < argc; k++) {
    if (!std::strcmp(argv[k], "-i") || !std::strcmp(argv[k], "--input-file")) {
      file = argv[++k];
      found++;
    } else if (!std::strcmp(argv[k], "-d") ||
               !std::strcmp(argv[k], "--dimension")) {
      dim = size_t(std::atoi(argv[++k]));
      found++;
    }
  }
  if (found != 2) {
    std::cerr << argv[0] << " -i input.yaml -d 1" << std::endl;
    return 1;
  }
  try {
    const Node c = nlm::yaml::load_file(file);
    const InpData d = read(c);
    if (dim == 1) fe1d(d);
    else if (dim == 2) fe2d(d);
    else throw nlm::yaml::Error("dimension must be 1 or 2");
  } catch (const std::exception& e) {
    std::cerr << "Error: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
