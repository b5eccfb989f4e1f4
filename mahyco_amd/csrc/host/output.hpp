// VTP/PVD output of the time loop, text-identical to the reference writer
// (cppnabla/mesh/PvdFileWriter2D.cc:26-114, FileWriter.cc:20-35).
#pragma once

#include <map>
#include <string>
#include <vector>

struct mhy_ctx;

namespace mhy {

int ctxField(mhy_ctx* x, int field, std::vector<double>& out);
int ctxNbmat(const mhy_ctx* x);
int ctxProjection(const mhy_ctx* x);
int ctxEuc(const mhy_ctx* x);  // the Eucclhyd scheme
void ctxMesh(const mhy_ctx* x, long& nx, long& ny);

class OutputWriter {
 public:
  // "none" / "NONE" disables output; otherwise files go to <dir>.
  explicit OutputWriter(const std::string& dir);
  bool disabled() const { return disabled_; }
  int write(mhy_ctx* x, int iteration, double time);

 private:
  std::string dir_;
  bool disabled_ = false;
  std::map<double, std::string> files_, particles_;
};

}  // namespace mhy
