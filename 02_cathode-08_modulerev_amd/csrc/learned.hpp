// learned.hpp -- the convolutional initial-guess network (learned.hip) as the
// NDD step uses it.
#pragma once

#include "common.hpp"

namespace cathode {

// Activation planes of one inference.  Every NDD handle owns its own, so
// handles that share one network (independent instances stepped side by side
// on their own streams) never write the same buffers; the network object
// only holds the weights, which inference reads.
struct CnnScratch;
CnnScratch* cnn_scratch_new();
void cnn_scratch_free(CnnScratch* sc);

// out (FP64 plane, pitch nj+2, ghosts 0) = network(feat), feat = 3 FP64
// planes of (ni+2)(nj+2) values, on stream s, in the caller's scratch sc.
// net from cathode_cnn_create (its weights are packed when they are set).
void cnn_infer(void* net, int ni, int nj, const double* feat, double* out, CnnScratch* sc, hipStream_t s);

}  // namespace cathode
