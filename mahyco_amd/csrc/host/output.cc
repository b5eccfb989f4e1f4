// VTP/PVD writer with the reference's text layout: default ostream
// formatting (6 significant digits), variables sorted by name, one PVD index
// rewritten after every snapshot.
#include "output.hpp"

#include <sys/stat.h>

#include <fstream>

#include "../../../include/mahyco.h"

namespace mhy {

OutputWriter::OutputWriter(const std::string& dir) : dir_(dir) {
  disabled_ = (dir == "none" || dir == "NONE");
  if (!disabled_) mkdir(dir_.c_str(), 0755);
}

int OutputWriter::write(mhy_ctx* x, int iteration, double time) {
  if (disabled_) return MHY_OK;
  long nx, ny;
  ctxMesh(x, nx, ny);
  const long N = (nx + 1) * (ny + 1), C = nx * ny;
  std::vector<double> X, Y;
  int st = ctxField(x, MHY_NODE_COORD_X, X);
  if (st == MHY_OK) st = ctxField(x, MHY_NODE_COORD_Y, Y);
  std::map<std::string, std::vector<double>> nodeVars, cellVars;
  if (st == MHY_OK) st = ctxField(x, MHY_NODE_OUT_VX, nodeVars["VitesseX"]);
  if (st == MHY_OK) st = ctxField(x, MHY_NODE_OUT_VY, nodeVars["VitesseY"]);
  if (st == MHY_OK) st = ctxField(x, MHY_CELL_PRESSURE, cellVars["Pressure"]);
  if (st == MHY_OK) st = ctxField(x, MHY_CELL_DENSITY, cellVars["Density"]);
  if (st == MHY_OK) st = ctxField(x, MHY_CELL_ENERGY, cellVars["Energy"]);
  const int m = ctxNbmat(x);
  if (st == MHY_OK && m > 1 && ctxProjection(x) == 1) st = ctxField(x, MHY_MAT_FRACVOL, cellVars["fracvol1"]);
  if (st == MHY_OK && m > 2 && ctxProjection(x) == 1)
    st = ctxField(x, MHY_MAT_FRACVOL + MHY_MAT_STRIDE, cellVars["fracvol2"]);
  if (st != MHY_OK) return st;
  const std::string name = "VnrRemap." + std::to_string(iteration) + ".vtp";
  std::ofstream w(dir_ + "/" + name);
  if (!w) return MHY_ERR_IO;
  w << "<?xml version=\"1.0\"?>" << std::endl;
  w << "<VTKFile type=\"PolyData\">" << std::endl;
  w << "\t<PolyData>" << std::endl;
  w << "\t\t<Piece NumberOfPoints=\"" << N << "\" NumberOfPolys=\"" << C << "\">" << std::endl;
  w << "\t\t\t<Points>" << std::endl;
  w << "\t\t\t\t<DataArray type=\"Float32\" NumberOfComponents=\"3\" format=\"ascii\">" << std::endl;
  for (long r = 0; r < N; ++r) w << " " << X[r] << " " << Y[r] << " " << 0.0;
  w << std::endl;
  w << "\t\t\t\t</DataArray>" << std::endl;
  w << "\t\t\t</Points>" << std::endl;
  w << "\t\t\t<Polys>" << std::endl;
  w << "\t\t\t\t<DataArray type=\"Int32\" Name=\"connectivity\" format=\"ascii\">" << std::endl;
  for (long j = 0; j < ny; ++j)
    for (long i = 0; i < nx; ++i) {
      const long bl = j * (nx + 1) + i;
      w << " " << " " << bl << " " << bl + 1 << " " << bl + nx + 2 << " " << bl + nx + 1;
    }
  w << std::endl;
  w << "\t\t\t\t</DataArray>" << std::endl;
  w << "\t\t\t\t<DataArray type=\"Int32\" Name=\"offsets\" format=\"ascii\">" << std::endl;
  for (long c = 1; c <= C; ++c) w << " " << c * 4;
  w << std::endl;
  w << "\t\t\t\t</DataArray>" << std::endl;
  w << "\t\t\t</Polys>" << std::endl;
  w << "\t\t\t<PointData Scalars=\"" << nodeVars.begin()->first << "\">" << std::endl;
  for (auto& kv : nodeVars) {
    w << "\t\t\t\t<DataArray Name=\"" << kv.first << "\" type=\"Float32\" format=\"ascii\">" << std::endl;
    for (long r = 0; r < N; ++r) w << " " << kv.second[r];
    w << std::endl;
    w << "\t\t\t\t</DataArray>" << std::endl;
  }
  w << "\t\t\t</PointData>" << std::endl;
  w << "\t\t\t<CellData Scalars=\"" << cellVars.begin()->first << "\">" << std::endl;
  for (auto& kv : cellVars) {
    w << "\t\t\t\t<DataArray Name=\"" << kv.first << "\" type=\"Float32\" format=\"ascii\">" << std::endl;
    for (long c = 0; c < C; ++c) w << " " << kv.second[c];
    w << std::endl;
    w << "\t\t\t\t</DataArray>" << std::endl;
  }
  w << "\t\t\t</CellData>" << std::endl;
  w << "\t\t</Piece>" << std::endl;
  w << "\t</PolyData>" << std::endl;
  w << "</VTKFile>" << std::endl;
  w.close();
  files_[time] = name;
  std::ofstream pvd(dir_ + "/VnrRemap.pvd");
  pvd << "<?xml version=\"1.0\"?>" << std::endl;
  pvd << "<VTKFile type=\"Collection\" version=\"0.1\">" << std::endl;
  pvd << "\t<Collection>" << std::endl;
  for (auto& kv : files_)
    pvd << "\t\t\t<DataSet timestep=\"" << kv.first << "\" group=\"\" part=\"0\" file=\"" << kv.second << "\"/>"
        << std::endl;
  pvd << "\t</Collection>" << std::endl;
  pvd << "</VTKFile>" << std::endl;
  return MHY_OK;
}

}  // namespace mhy
