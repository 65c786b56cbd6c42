// ars_serve_poll.hpp -- how the serving loop reads the completion event of a
// batch slot, kept free of HIP types so that it is unit-tested on the CPU.
#pragma once
#include <cstdint>

namespace ars {

enum class SlotPoll { kRunning, kDone, kFailed };

// The event query's result, given the runtime's codes for success and "not
// ready": anything else (a failed kernel or copy on the slot's stream) also
// ends the slot, so that its requests complete with an error instead of the
// slot staying busy forever and every waiter with it.
constexpr SlotPoll classify_slot_poll(int query_rc, int success_rc, int not_ready_rc) {
    return query_rc == not_ready_rc ? SlotPoll::kRunning : query_rc == success_rc ? SlotPoll::kDone : SlotPoll::kFailed;
}

// A request's final status when its slot retires: a failed slot turns every
// request that had not already failed into `hip_error`.
constexpr int32_t retired_status(SlotPoll p, int32_t status, int32_t hip_error) {
    return (p == SlotPoll::kFailed && status >= 0) ? hip_error : status;
}

// Jobs a slot must hold to run them in the batched kernel rather than as
// plain calls: the configured value, or the default for a loaded or an idle
// GPU, never more than the slot can hold.
constexpr uint64_t min_batch_jobs(uint64_t configured, bool loaded, uint64_t loaded_default, uint64_t idle_default,
                                  uint64_t max_batch_jobs) {
    const uint64_t m = configured ? configured : (loaded ? loaded_default : idle_default);
    return m < max_batch_jobs ? m : max_batch_jobs;
}

}  // namespace ars
