// VTK XML UnstructuredGrid (.vtu) writer and reader.
//
// Writes what NLMech's rw::writer::VtkWriter writes through VTK
// (src/rw/vtkWriter.cpp:20-298): current positions x + u as "Points"
// (Float32, vtkPoints' default type), point data arrays as Float64, the time
// as the field array "TIME", optional cells, appended base64 data with a
// UInt32 header (UInt64 when an array exceeds 4 GiB) and zlib-compressed
// blocks of 32 KiB when the compression type is "zlib".
//
// Reads the same files and the other layouts VTK produces (appended raw or
// base64, inline binary or ascii, compressed or not, UInt32 or UInt64
// headers), which is what rw::reader::VtkReader uses VTK for
// (src/rw/reader.cpp:149-330, src/rw/vtkReader.cpp).
#pragma once

#include <cstdint>
#include <map>
#include <string>
#include <vector>

namespace nlm {
namespace rw {

// VTK cell types used by the reference's meshes (src/util/feElementDefs.h).
constexpr int kVtkLine = 3;
constexpr int kVtkTriangle = 5;
constexpr int kVtkQuad = 9;
constexpr int kVtkTetra = 10;
int vtk_cell_vertices(int type);

class VtuWriter {
 public:
  // filename without the ".vtu" extension (VtkWriter::VtkWriter appends it);
  // compress: "zlib" or anything else for no compression.  legacy: the ASCII
  // legacy format instead (File_Format: legacy_vtk, LegacyVtkWriter,
  // src/rw/legacyVtkWriter.cpp): "<filename>.vtk", points x + u, one vertex
  // cell per node, point data as SCALARS (TENSORS for the 6-component
  // symmetric tensors), numbers in the stream's default format; cells and
  // field data are left out with the reference's warning.
  VtuWriter(const std::string& filename, const std::string& compress,
            bool legacy = false);

  // Nodes (AoS [3n]) plus optional displacement (AoS [3n] or null):
  // VtkWriter::appendNodes.
  void appendNodes(const std::vector<double>& nodes, const double* u);
  // Cells of one type with connectivity [nverts * ncells]:
  // VtkWriter::appendMesh.
  void appendCells(int type, const std::vector<int64_t>& conn);
  // Point data with ncomp components per node (stored as Float64).
  void appendPointData(const std::string& name, const double* data,
                       int ncomp);
  void appendPointData(const std::string& name, const float* data, int ncomp);
  void appendPointData(const std::string& name, const uint8_t* data);
  // Field data: one Float64 tuple, or a UInt8 array.
  void appendFieldData(const std::string& name, double value);
  void appendFieldData(const std::string& name,
                       const std::vector<uint8_t>& bytes);
  // VtkWriter::addTimeStep: field array "TIME" (nothing in the legacy
  // format).
  void addTimeStep(double t) {
    if (!legacy_) appendFieldData("TIME", t);
  }
  // Writes the file; throws std::runtime_error on an I/O error.
  void close();

 private:
  struct Array {
    std::string name;
    std::string type;  // VTK type name
    int ncomp;
    int64_t ntuples;
    std::vector<uint8_t> bytes;
  };
  std::string file_;
  bool zlib_;
  bool legacy_ = false;
  std::string legacy_text_;  // the legacy file's text so far
  int64_t npoints_ = 0;
  std::vector<Array> field_, point_;
  Array points_;
  int64_t ncells_ = 0;
  Array conn_, offsets_, types_;
  bool closed_ = false;
};

class VtuReader {
 public:
  // Parses the file; throws std::runtime_error if it is not a readable
  // UnstructuredGrid file.
  explicit VtuReader(const std::string& filename);

  int64_t numPoints() const { return npoints_; }
  int64_t numCells() const { return ncells_; }
  // Point coordinates as AoS doubles [3n].
  bool readNodes(std::vector<double>* xyz) const;
  // A point data array converted to double; *ncomp receives its components.
  bool readPointData(const std::string& name, std::vector<double>* out,
                     int* ncomp = nullptr) const;
  bool hasPointData(const std::string& name) const;
  // A field data array converted to double / as raw bytes (UInt8 arrays).
  bool readFieldData(const std::string& name, std::vector<double>* out) const;
  bool readFieldBytes(const std::string& name,
                      std::vector<uint8_t>* out) const;
  // Cells: connectivity, offsets (end of each cell) and types.
  bool readCells(std::vector<int64_t>* conn, std::vector<int64_t>* offsets,
                 std::vector<uint8_t>* types) const;

 private:
  struct ArrayRef {
    std::string type;
    int ncomp = 1;
    std::string format;  // appended, binary, ascii
    int64_t offset = 0;
    std::string inline_text;
  };
  std::vector<uint8_t> decode(const ArrayRef& a) const;
  bool to_double(const ArrayRef& a, std::vector<double>* out) const;

  std::string file_;
  std::string text_;        // the XML part
  std::string appended_;    // appended data after '_' (base64 text or raw)
  bool appended_b64_ = true;
  bool zlib_ = false;
  bool header64_ = false;
  int64_t npoints_ = 0, ncells_ = 0;
  std::map<std::string, ArrayRef> field_, point_;
  ArrayRef points_, conn_, offsets_, types_;
  bool has_points_ = false, has_cells_ = false;
};

}  // namespace rw
}  // namespace nlm
