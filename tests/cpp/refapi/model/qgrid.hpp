// Test stand-in for the reference's include/model/qgrid.hpp: the QGrid singleton with the accessors
// the GPU classes read (qx, qy, qz_extended, nqy, nqz_extended), filled by the test as
// QGrid::create / create_qz_extended fill it in the reference.
#pragma once

#include <vector>

#include <common/typedefs.hpp>

namespace hig {
class QGrid {
 public:
  std::vector<real_t> qx_, qy_;
  std::vector<complex_t> qz_extended_;
  static QGrid& instance() {
    static QGrid g;
    return g;
  }
  int nqy() const { return (int)qy_.size(); }
  int nqz_extended() const { return (int)qz_extended_.size(); }
  real_t qx(unsigned int i) const { return qx_[i]; }
  real_t qy(unsigned int i) const { return qy_[i]; }
  complex_t qz_extended(unsigned int i) const { return qz_extended_[i]; }
};
}  // namespace hig
