// Pin-cell CMFD on the device (FP64): homogenization, D-tilde/D-hat coupling,
// the coarse power iteration and its BiCGSTAB inner solves.
#pragma once

#include <hip/hip_runtime.h>

#include "../../../include/mocgpu.h"

struct mocgpu_handle;

namespace mocc_amd {
namespace gpu {

class CmfdSolver {
public:
    void setup(const mocgpu_problem &p, const mocgpu_options &opt, hipStream_t s);
    void initialize(hipStream_t s);
    // EigenSolver::do_cmfd: get_pin_flux, CMFD::solve(k), set_pin_flux.
    int step(mocgpu_handle &h, double &k);
};

}  // namespace gpu
}  // namespace mocc_amd
