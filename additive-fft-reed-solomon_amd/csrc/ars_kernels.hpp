// ars_kernels.hpp -- launch interface between the C-ABI layer and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "ars_gf.hpp"

namespace ars {

// Per-recover-job multiplier tables, computed on the device (k_recover_prep).
struct RecPrep {
    MulTab tab[256];      // erased shard: 1 / Pi'(omega_i); other shards: Pi(omega_i)
    uint8_t erased[256];  // 1 for erased shards
};

struct PosArg {
    uint8_t p[256];
};

// Device copy of the field tables used by k_recover_prep:
// exp[0..511], log[512..767], omega[768..1023].
constexpr int kGtabBytes = 1024;

// Default tile variant (quads per workgroup, LDS buffers) of an operation
// (0 = encode, 1 = recover) on 2^tl points, used when the context holds no
// tuned choice.  Encode: 64 quads with one buffer, where a wave holds one
// shard group and its butterflies take the tables' high words by scalar
// loads (fastest or within 2% of it at every T from 32 to 128).  Recover:
// the streaming kernel (k_recover_stream, one buffer) with 32 quads at 256
// points and 64 below.  With 64 quads a wave skips erased rows with scalar
// branches; at 256 points the 32-quad tile, two workgroups per CU and no
// such branches, measured 3.5% faster once the butterflies' lookups went into
// priority bursts.  The byte path has only the two-buffer kernel.
constexpr int default_pq(int op, int tl) { return (op == 1 && tl >= 8) ? 32 : 64; }
constexpr int default_nbuf(int op, int tl) { return (void)op, (void)tl, 1; }

// pq / nbuf select a tile variant for transforms of 32+ points (0 = default):
// pq = column quads per workgroup (16, 32 or 64), nbuf = LDS exchange buffers (1 or 2).
// am: access mode of the buffers (BufIO in ars_tiles.hpp: 0 bytes, 1 aligned
// dwords, 2 dwords r bytes past alignment, 3 aligned dwords of a codeword of
// 4 GiB or more; io_mode() below picks it).  Mode 3 runs 64-quad tiles only
// (any pq / nbuf asked for is replaced); encode with T not a power of two and
// recompute take modes 0, 1 and 3 (2 runs as 0).
hipError_t launch_encode_ref(int tl, int am, bool accumulate, const uint8_t *msg, uint8_t *par,
                             uint32_t L, uint32_t chunks, const MulTab *tw, hipStream_t s,
                             int pq = 0, int nbuf = 0);
hipError_t launch_encode_kpow2(int kl, int am, const uint8_t *msg, uint8_t *par, uint32_t L,
                               uint32_t blocks, uint32_t msg_base, const MulTab *tw, hipStream_t s);
hipError_t launch_recover_prep(RecPrep *out, uint32_t n, uint32_t e, const PosArg &pos,
                               const uint8_t *gtab, hipStream_t s);
// Recover with a ready plan.  With nbuf == 1 and dword-addressable rows this
// is the streaming kernel, whose grid is sized from `cus` compute units.
hipError_t launch_recover(int nl, int am, uint8_t *data, uint32_t L, const RecPrep *prep,
                          const MulTab *tw, hipStream_t s, int pq = 0, int nbuf = 0, uint32_t cus = 0);
// Recover in one launch, each workgroup computing the plan itself (the
// default tile variant), when the launch has at most max_groups workgroups;
// hipErrorNotSupported (nothing launched) otherwise.
hipError_t launch_recover_fused(int nl, int am, uint8_t *data, uint32_t L, const PosArg &pos, uint32_t e,
                                const uint8_t *gtab, const MulTab *tw, hipStream_t s, uint32_t max_groups);
hipError_t launch_recompute(int tl, int am, uint8_t *data, uint32_t L, const MulTab *tw,
                            hipStream_t s);

// Access mode for buffers a and b (b = a for in-place work) of a codeword of
// `cw_bytes`: 1 when both are 4-byte aligned, 2 when both sit the same 1..3
// bytes past alignment, else 0; shard_len must be a multiple of 4 for 1 or 2.
// A codeword of 4 GiB or more takes mode 3 when mode 1's conditions hold and
// -1 (not supported) otherwise.
inline int io_mode(const void *a, const void *b, size_t L, uint64_t cw_bytes = 0) {
    const uintptr_t ra = reinterpret_cast<uintptr_t>(a) & 3u, rb = reinterpret_cast<uintptr_t>(b) & 3u;
    const int m = (L % 4 || ra != rb) ? 0 : ra ? 2 : 1;
    if (cw_bytes > 0xFFFFFFFFull) return m == 1 ? 3 : -1;
    return m;
}

}  // namespace ars
