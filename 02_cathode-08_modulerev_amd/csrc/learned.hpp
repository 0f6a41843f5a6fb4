// learned.hpp -- the convolutional initial-guess network (learned.hip) as the
// NDD step uses it.
#pragma once

#include "common.hpp"

namespace cathode {

// out (FP64 plane, pitch nj+2, ghosts 0) = network(feat), feat = 3 FP64
// planes of (ni+2)(nj+2) values, on stream s.  net from cathode_cnn_create.
void cnn_infer(void* net, int ni, int nj, const double* feat, double* out, hipStream_t s);

}  // namespace cathode
