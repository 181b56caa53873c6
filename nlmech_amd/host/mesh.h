// Nodes, nodal volumes and (optionally) elements of a mesh file:
// fe::Mesh's finite-difference part (src/fe/mesh.cpp:24-330).
#pragma once

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "input.h"

namespace nlm {
namespace fe {

class Mesh {
 public:
  // fe::Mesh::Mesh(MeshDeck*) (mesh.cpp:37-94): reads .vtu or .csv, takes
  // volumes from "Node_Volume" / "Volume" or computes them from the
  // elements, computes the mesh size if the deck has none, checks every
  // volume against 0.01 h^dim.  Throws inp::DeckError.
  explicit Mesh(const inp::MeshDeck& deck);
  // fe::Mesh::setMeshData (mesh.cpp:352-375): nodes AoS [3n], volumes [n].
  Mesh(size_t dim, std::vector<double> nodes, std::vector<double> vol);

  size_t getDimension() const { return d_dim; }
  size_t getNumNodes() const { return d_vol.size(); }
  const std::vector<double>& getNodes() const { return d_nodes; }
  const std::vector<double>& getNodalVolumes() const { return d_vol; }
  double getMeshSize() const { return d_h; }
  // fe::Mesh::computeBBox (mesh.cpp:311-324): starts from the origin.
  const std::pair<std::vector<double>, std::vector<double>>& getBoundingBox()
      const {
    return d_bbox;
  }
  int getElementType() const { return d_eType; }
  size_t getNumElements() const { return d_numElems; }
  const std::vector<int64_t>& getElementConnectivities() const {
    return d_enc;
  }

 private:
  void computeVol();
  void computeMeshSize();
  void computeBBox();

  size_t d_dim = 0;
  double d_h = 0.;
  std::vector<double> d_nodes;
  std::vector<double> d_vol;
  int d_eType = 1;
  size_t d_numElems = 0;
  std::vector<int64_t> d_enc;
  std::pair<std::vector<double>, std::vector<double>> d_bbox;
};

}  // namespace fe
}  // namespace nlm
