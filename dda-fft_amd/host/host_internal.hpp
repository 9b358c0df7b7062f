// Host helpers shared between host_physics.cpp and the library's C ABI
// implementation; not part of the public interface.
#pragma once

#include <complex>

namespace ddafft_host {

double m00_grad(const float* kinc, const float* ksc, const float* normal, const float* einc0, const float* einc1,
                const float* esc0, const float* esc1, std::complex<double>* g0, std::complex<double>* g1);

}  // namespace ddafft_host
