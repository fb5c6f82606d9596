// Exports the reference's Hilbert curve functions (hilbert.h, compiled from
// the reference tree where it lies) through a C ABI so that tests can check
// the product's 2-D Hilbert tile order against them. TEST INFRASTRUCTURE ONLY.
// HIP's runtime header supplies the __host__/__device__ qualifiers that
// hilbert.h uses; nothing from the reference is copied here.
#include <hip/hip_runtime.h>
#include "hilbert.h"

extern "C" void ref_hilbert_i2c(unsigned n_dims, unsigned n_bits, unsigned long long index,
                                unsigned long long* coord) {
  hilbert_i2c(n_dims, n_bits, index, coord);
}

extern "C" unsigned long long ref_hilbert_c2i(unsigned n_dims, unsigned n_bits,
                                              const unsigned long long* coord) {
  return hilbert_c2i(n_dims, n_bits, coord);
}
