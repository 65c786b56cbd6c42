// ars_decode.hpp -- launch interface of the error-decoding kernels
// (ars_decode.hip) for the C-ABI layer (ars_decode_api.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "ars_gf.hpp"
#include "ars_keyeq.hpp"

namespace ars {

// Where the syndrome kernel lists the columns whose syndrome is not zero.
struct SynOut {
    uint32_t *count;  // number of listed columns (zeroed before the launch)
    uint32_t *cols;   // column index per list entry (capacity L)
    uint8_t *syn;     // T syndrome bytes per list entry (capacity L * T)
};

hipError_t launch_syndrome(int tl, bool aligned, const uint8_t *rx, uint32_t L, uint32_t chunks,
                           const MulTab *tw, const SynOut &so, hipStream_t s);
hipError_t launch_decode_columns(uint8_t *rx, uint32_t L, uint32_t n, int tl, const uint32_t *count,
                                 const uint32_t *cols, const uint8_t *syn, bool formal, uint32_t grid,
                                 uint8_t *status, uint32_t *failed, const uint8_t *gtab, const MulTab *tw,
                                 const uint8_t *xt, hipStream_t s);
hipError_t launch_decode_tail(uint8_t *rx, uint32_t L, uint32_t n, int tl, const uint32_t *cols,
                              const DecPlan *plans, uint32_t count, uint8_t *status, uint32_t *failed,
                              const uint8_t *gtab, const MulTab *tw, hipStream_t s);

}  // namespace ars
