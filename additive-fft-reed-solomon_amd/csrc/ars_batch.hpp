// ars_batch.hpp -- launch interface of the batched persistent kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "ars_batch_desc.hpp"
#include "ars_gf.hpp"
#include "ars_kernels.hpp"

namespace ars {

// All three take the arena base (a BHeader followed by its regions).
hipError_t launch_batch_reset(uint8_t *arena, hipStream_t s);
hipError_t launch_batch_prep(uint8_t *arena, const uint8_t *gtab, uint32_t grid, hipStream_t s);
hipError_t launch_batch(uint8_t *arena, const MulTab *tw, uint32_t grid, hipStream_t s);
int batch_blocks_per_cu();

}  // namespace ars
