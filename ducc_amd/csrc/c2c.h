// c2c.h -- complex FFT of an N-d device array along chosen axes (the drop-in
// for ducc0::c2c, src/ducc0/fft/fft.h:848-850).
#pragma once
#include <cstdint>
#include <memory>

#include "dmx.h"

namespace dmx {

// out = fct * FFT(in) along axes[0..naxes) (in any order; an axis named
// twice is transformed twice).  in: typecode DMX_C64/C128/F32/F64 (real
// input is read as complex with zero imaginary part); out: complex of the
// same precision.  Strides in elements, both arrays on the device, out may
// be the same array as in (same strides).  Enqueued on s; the returned
// object owns the buffers the enqueued kernels use and must outlive them.
std::shared_ptr<void> c2c_exec(const void *in, int in_tc, const int64_t *in_st,
                               void *out, const int64_t *out_st, int ndim,
                               const int64_t *shape, const int *axes,
                               int naxes, bool forward, double fct,
                               hipStream_t s);

// smallest 2,3,5,7,11-smooth integer >= n
int64_t smooth_length(int64_t n);

}  // namespace dmx
