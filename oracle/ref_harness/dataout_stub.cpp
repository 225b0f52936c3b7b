// Test infrastructure only.  Empty bodies for the reference's DataOut
// (Explicit3D/include/dataOut.h:44-97): its VTU writer needs VTK, which is
// shipped for Windows only, and output is not on the step path.
#include "dataOut.h"

namespace EnSC {
DataOut::DataOut(const Types::VerticesAll<3>& v, const std::vector<Element_HexN8>& e, Types::Real t)
    : opVerticesAll(v), opTriaHex(e), time(t) {}
void DataOut::add_node_data(DataOutType::DataType, const Eigen::Matrix<Types::Real, Eigen::Dynamic, 1>*, std::string) {}
void DataOut::add_ele_data(DataOutType::DataType, const Eigen::Matrix<Types::Real, Eigen::Dynamic, 1>*, std::string) {}
void DataOut::write_vtu(const std::string&) {}
void DataOut::write_vtu_virtualParticles(const unsigned int&, const FSI_share_data&) const {}
}  // namespace EnSC
