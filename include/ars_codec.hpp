// ars_codec.hpp -- header-only C++ mirror of the reference's codec types over
// the C ABI in ars.h.
//
// The reference exposes `Codec<G, B, K, N, T>` with the aliases `Rs<N, T>` /
// `RsGfni<N, T>` / `RsLut<N, T>` (rs/src/codec.rs:10-27, rs/src/lib.rs:20-45)
// and two sharded hot-path methods:
//   encode_systematic_sharded(&self, message, parity, workspace, shard_len)
//       (rs/src/codec.rs:241-265)
//   recover_erasure_shards(&self, received, workspace, shard_len, positions) -> bool
//       (rs/src/codec.rs:479-569)
// plus recompute_data_from_parity_sharded (rs/src/codec.rs:321-340) and a
// sharded form of decode_systematic_scalar (rs/src/codec.rs:341-420).
// `Rs<N, T>` here keeps those names, argument meaning and return values on
// device buffers.  The reference panics on out-of-range positions; this
// mirror throws ars::Error, as it does for every negative ARS_E_* code.
// Like the reference's ZST codec, a Codec is cheap to copy; the per-device
// tables live in an ars_ctx shared by all codecs on that device.
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "ars.h"

namespace ars {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &what) : std::runtime_error(what), code(c) {}
};

inline int check(int rc) {
    if (rc < 0) throw Error(rc, ars_last_error());
    return rc;
}

// One context per device, created on first use and kept for the process.
inline ars_ctx *device_context(int device = 0) {
    static ars_ctx *ctxs[64] = {nullptr};
    if (device < 0 || device >= 64) throw Error(ARS_E_INVAL, "device index out of range");
    if (!ctxs[device]) check(ars_create(device, &ctxs[device]));
    return ctxs[device];
}

// N <= 256 a power of two; 1 <= T < N a power of two (rs/src/lib.rs:16-19).
template <std::size_t N, std::size_t T>
class Rs {
    static_assert(N >= 2 && N <= 256 && (N & (N - 1)) == 0, "N must be a power of two in [2, 256]");
    static_assert(T >= 1 && T < N && (T & (T - 1)) == 0, "T must be a power of two in [1, N)");

  public:
    explicit Rs(int device = 0, void *stream = nullptr) : ctx_(device_context(device)), stream_(stream) {}

    // Parity is XOR-accumulated into, as in the reference (callers pass
    // zeroed parity for a plain encode).  `workspace` is accepted for
    // signature parity and unused.  message: (N-T)*shard_len, parity:
    // T*shard_len, device pointers.
    void encode_systematic_sharded(const uint8_t *message, uint8_t *parity, uint8_t * /*workspace*/,
                                   std::size_t shard_len) const {
        check(ars_encode(ctx_, N, N - T, message, parity, shard_len, ARS_ENC_ACCUMULATE, stream_));
    }

    // received: N*shard_len device bytes, in place; positions on the host.
    bool recover_erasure_shards(uint8_t *received, uint8_t * /*workspace*/, std::size_t shard_len,
                                const uint8_t *positions, std::size_t e) const {
        return check(ars_recover(ctx_, N, N - T, received, shard_len, positions, (uint32_t)e, stream_)) ==
               ARS_TRUE;
    }

    // rs/src/codec.rs:321-340, N = 2T only.
    void recompute_data_from_parity_sharded(uint8_t *received, uint8_t * /*workspace*/,
                                            std::size_t shard_len) const {
        static_assert(N == 2 * T, "recompute_data_from_parity_sharded needs N = 2T");
        check(ars_recompute_data(ctx_, N, N - T, received, shard_len, stream_));
    }

    // Every byte column decoded as Codec::decode_systematic_scalar
    // (rs/src/codec.rs:341-420) decodes one received word, in place on the
    // device buffer `received` (N*shard_len bytes).  True iff no column
    // returned false.  formal_derivative: evaluate errors with the formal
    // derivative instead of the reference's poly::deriv_lnh.  Needs T >= 2.
    bool decode_systematic_sharded(uint8_t *received, std::size_t shard_len, bool formal_derivative = false,
                                   uint64_t *counts = nullptr) const {
        static_assert(T >= 2, "error decoding needs T >= 2");
        return check(ars_decode(ctx_, N, N - T, received, shard_len,
                                formal_derivative ? ARS_DEC_FORMAL_DERIVATIVE : 0u, nullptr, counts, stream_)) ==
               ARS_TRUE;
    }

  private:
    ars_ctx *ctx_;
    void *stream_;
};

}  // namespace ars
