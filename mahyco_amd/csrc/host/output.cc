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

namespace {

// One VTP snapshot of module <module> and its PVD index
// (PvdFileWriter2D.cc:26-114): point / cell data blocks only when they hold
// fields; nx * ny quads unless the piece is empty (C == 0).
int writeModule(const std::string& dir, const std::string& module, int iteration, double time, long N,
                const std::vector<double>& X, const std::vector<double>& Y, long nx, long ny, long C,
                const std::map<std::string, std::vector<double>>& nodeVars,
                const std::map<std::string, std::vector<double>>& cellVars, std::map<double, std::string>& files) {
  const std::string name = module + "." + std::to_string(iteration) + ".vtp";
  std::ofstream w(dir + "/" + name);
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
  if (C > 0)
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
  auto block = [&](const char* tag, const std::map<std::string, std::vector<double>>& vars, long n) {
    if (vars.empty()) return;
    w << "\t\t\t<" << tag << " Scalars=\"" << vars.begin()->first << "\">" << std::endl;
    for (auto& kv : vars) {
      w << "\t\t\t\t<DataArray Name=\"" << kv.first << "\" type=\"Float32\" format=\"ascii\">" << std::endl;
      for (long r = 0; r < n; ++r) w << " " << kv.second[r];
      w << std::endl;
      w << "\t\t\t\t</DataArray>" << std::endl;
    }
    w << "\t\t\t</" << tag << ">" << std::endl;
  };
  block("PointData", nodeVars, N);
  block("CellData", cellVars, C);
  w << "\t\t</Piece>" << std::endl;
  w << "\t</PolyData>" << std::endl;
  w << "</VTKFile>" << std::endl;
  w.close();
  files[time] = name;
  std::ofstream pvd(dir + "/" + module + ".pvd");
  pvd << "<?xml version=\"1.0\"?>" << std::endl;
  pvd << "<VTKFile type=\"Collection\" version=\"0.1\">" << std::endl;
  pvd << "\t<Collection>" << std::endl;
  for (auto& kv : files)
    pvd << "\t\t\t<DataSet timestep=\"" << kv.first << "\" group=\"\" part=\"0\" file=\"" << kv.second << "\"/>"
        << std::endl;
  pvd << "\t</Collection>" << std::endl;
  pvd << "</VTKFile>" << std::endl;
  return MHY_OK;
}

}  // namespace

// VNR: VnrRemap with the node velocities and the cell fields
// (vnr_remap/VnrRemap.cc:255-287).  Eucclhyd: EucclhydRemap with the cell
// fields, and the empty file of the disabled particle scheme
// (eucclhyd_remap/EucclhydRemap.cc:24-62).
int OutputWriter::write(mhy_ctx* x, int iteration, double time) {
  if (disabled_) return MHY_OK;
  long nx, ny;
  ctxMesh(x, nx, ny);
  const long N = (nx + 1) * (ny + 1), C = nx * ny;
  std::vector<double> X, Y;
  int st = ctxField(x, MHY_NODE_COORD_X, X);
  if (st == MHY_OK) st = ctxField(x, MHY_NODE_COORD_Y, Y);
  std::map<std::string, std::vector<double>> nodeVars, cellVars;
  const int m = ctxNbmat(x);
  if (ctxEuc(x)) {
    if (st == MHY_OK) st = ctxField(x, MHY_CELL_PRESSURE, cellVars["Pressure"]);
    if (st == MHY_OK) st = ctxField(x, MHY_CELL_DENSITY, cellVars["Density"]);
    if (st == MHY_OK) st = ctxField(x, MHY_CELL_ENERGY, cellVars["Energy"]);
    if (st == MHY_OK) st = ctxField(x, MHY_CELL_OUT_VX, cellVars["VelocityX"]);
    if (st == MHY_OK) st = ctxField(x, MHY_CELL_OUT_VY, cellVars["VelocityY"]);
    // the three material slots of the reference; absent materials print 0
    for (int k = 0; k < 3 && st == MHY_OK; ++k) {
      std::vector<double>& fk = cellVars["F" + std::to_string(k + 1)];
      if (k < m) st = ctxField(x, MHY_MAT_FRACVOL + k * MHY_MAT_STRIDE, fk);
      else fk.assign(C, 0.0);
    }
    if (st != MHY_OK) return st;
    st = writeModule(dir_, "EucclhydRemap", iteration, time, N, X, Y, nx, ny, C, nodeVars, cellVars, files_);
    if (st != MHY_OK) return st;
    std::map<std::string, std::vector<double>> partVars = {{"VolumePart", {}}, {"VxPart", {}}, {"VyPart", {}}};
    return writeModule(dir_, "Particules", iteration, time, 0, X, Y, nx, ny, 0, partVars, cellVars, particles_);
  }
  if (st == MHY_OK) st = ctxField(x, MHY_NODE_OUT_VX, nodeVars["VitesseX"]);
  if (st == MHY_OK) st = ctxField(x, MHY_NODE_OUT_VY, nodeVars["VitesseY"]);
  if (st == MHY_OK) st = ctxField(x, MHY_CELL_PRESSURE, cellVars["Pressure"]);
  if (st == MHY_OK) st = ctxField(x, MHY_CELL_DENSITY, cellVars["Density"]);
  if (st == MHY_OK) st = ctxField(x, MHY_CELL_ENERGY, cellVars["Energy"]);
  if (st == MHY_OK && m > 1 && ctxProjection(x) == 1) st = ctxField(x, MHY_MAT_FRACVOL, cellVars["fracvol1"]);
  if (st == MHY_OK && m > 2 && ctxProjection(x) == 1)
    st = ctxField(x, MHY_MAT_FRACVOL + MHY_MAT_STRIDE, cellVars["fracvol2"]);
  if (st != MHY_OK) return st;
  return writeModule(dir_, "VnrRemap", iteration, time, N, X, Y, nx, ny, C, nodeVars, cellVars, files_);
}

}  // namespace mhy
