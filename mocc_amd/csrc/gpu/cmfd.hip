#include "cmfd.hpp"

#include <string>

namespace mocc_amd {
namespace gpu {

void CmfdSolver::setup(const mocgpu_problem &, const mocgpu_options &, hipStream_t) {}
void CmfdSolver::initialize(hipStream_t) {}
int CmfdSolver::step(mocgpu_handle &, double &) { return MOCGPU_ERR_UNSUPPORTED; }

}  // namespace gpu
}  // namespace mocc_amd
