// nlm_dc: compares two simulation outputs, as the reference's dc tool does
// for "-k fd_simple" (tools/dc/dc.cpp, tools/dc/src/fdSimple.cpp:76-399):
// per compare tag the L2 norm of the nodal differences and the sup norm
// (both over |.| of the per-node difference vector), written to Out_Filename;
// exit status 1 when an L2 error is not definitelyLessThan Tolerance.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "vtu.h"
#include "yaml.h"

namespace {

bool less_than(double a, double b) {
  const double m = std::abs(a) < std::abs(b) ? std::abs(b) : std::abs(a);
  return (b - a) > m * 1e-5;
}
bool greater_than(double a, double b) {
  const double m = std::abs(a) < std::abs(b) ? std::abs(b) : std::abs(a);
  return (a - b) > m * 1e-5;
}

[[noreturn]] void fail(const std::string& m) {
  std::cerr << m << std::endl;
  std::exit(1);
}

std::vector<double> read_tag(const nlm::rw::VtuReader& r,
                             const std::string& tag, const std::string& file,
                             int* nc) {
  std::vector<double> v;
  if (!r.readPointData(tag, &v, nc))
    fail("Error: " + tag + " data can not be found in the file = " + file);
  return v;
}

bool fd_simple(const nlm::yaml::Node& c) {
  const std::string f1 = c["Filename_1"].str(), f2 = c["Filename_2"].str();
  double tol = std::numeric_limits<double>::max();
  if (c["Tolerance"]) tol = c["Tolerance"].as_double();
  const std::string out = c["Out_Filename"].str();
  const bool screen = c["Print_Screen"] ? c["Print_Screen"].as_bool() : true;
  const std::vector<std::string> tags = c["Compare_Tags"].as_strings();
  if (tags.empty())
    fail("Error: Please specify at least one data which should be compared.");
  FILE* fo = std::fopen(out.c_str(), "w");
  if (!fo) fail("Error: cannot write " + out);
  std::ostringstream oss;
  for (size_t k = 0; k < tags.size(); k++) {
    oss << tags[k] << "_L2_Error, " << tags[k] << "_Sup_Error";
    if (k + 1 < tags.size()) oss << ", ";
  }
  oss << "\n";
  std::fprintf(fo, "%s", oss.str().c_str());
  if (screen) std::cout << "------------------------\n" << oss.str();
  oss.str("");
  oss.clear();
  nlm::rw::VtuReader r1(f1), r2(f2);
  std::vector<double> p1, p2;
  r1.readNodes(&p1);
  r2.readNodes(&p2);
  if (p1.size() != p2.size())
    fail("Error: Two files have different number of nodes.");
  bool error = false;
  for (size_t k = 0; k < tags.size(); k++) {
    const std::string& tag = tags[k];
    int n1 = 1, n2 = 1;
    const std::vector<double> a = read_tag(r1, tag, f1, &n1);
    const std::vector<double> b = read_tag(r2, tag, f2, &n2);
    if (n1 != n2 || a.size() != b.size())
      fail("Error: " + tag + " arrays differ in shape.");
    double l2 = 0., sup = 0.;
    const size_t nn = a.size() / size_t(n1);
    if (tag == "Displacement" || tag == "Velocity" || tag == "Force") {
      for (size_t i = 0; i < nn; i++) {
        const double dx = a[3 * i] - b[3 * i], dy = a[3 * i + 1] - b[3 * i + 1],
                     dz = a[3 * i + 2] - b[3 * i + 2];
        const double len = std::sqrt(dx * dx + dy * dy + dz * dz);
        l2 += len * len;
        if (greater_than(len, sup)) sup = len;
      }
    } else if (tag == "Strain_Energy") {
      for (size_t i = 0; i < nn; i++) {
        const double df = a[i] - b[i];
        l2 += df * df;
        if (greater_than(df, sup)) sup = df;
      }
    } else if (tag == "Strain_Tensor" || tag == "Stress_Tensor") {
      // symmetric tensors stored as xx, yy, zz, xy, xz, yz; the difference
      // summed over all nine entries row by row
      for (size_t i = 0; i < nn; i++) {
        double d[6];
        for (int q = 0; q < 6; q++) d[q] = a[6 * i + q] - b[6 * i + q];
        const double df =
            d[0] + d[3] + d[4] + d[3] + d[1] + d[5] + d[4] + d[5] + d[2];
        l2 += df * df;
        if (greater_than(df, sup)) sup = df;
      }
    } else {
      fail("Error: Comparison for tag = " + tag +
           " has not yet been implemented.");
    }
    l2 = std::sqrt(l2);
    if (!less_than(l2, tol)) error = true;
    oss << l2 << ", " << sup;
    if (k + 1 < tags.size()) oss << ", ";
  }
  oss << "\n";
  std::fprintf(fo, "%s", oss.str().c_str());
  if (screen) std::cout << oss.str();
  std::fclose(fo);
  return error;
}

}  // namespace

int main(int argc, char** argv) {
  std::string file, kind;
  for (int k = 1; k + 1 < argc; k++) {
    if (!std::strcmp(argv[k], "-i") || !std::strcmp(argv[k], "--input-file"))
      file = argv[++k];
    else if (!std::strcmp(argv[k], "-k") || !std::strcmp(argv[k], "--kind"))
      kind = argv[++k];
  }
  if (file.empty()) {
    std::cerr << argv[0] << " -i input.yaml -k type" << std::endl;
    return 1;
  }
  try {
    const nlm::yaml::Node c = nlm::yaml::load_file(file);
    if (kind != "fd_simple") {
      std::cerr << "Error: only -k fd_simple is supported" << std::endl;
      return 1;
    }
    return fd_simple(c) ? 1 : 0;
  } catch (const std::exception& e) {
    std::cerr << "Error: " << e.what() << std::endl;
    return 1;
  }
}
