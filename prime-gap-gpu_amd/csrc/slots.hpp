// Work-counter slots of a runner's persistent launches (host side only).
#pragma once

#include <atomic>
#include <cstdint>

namespace pgx {

// Hands out slots 1..n-1 of a runner's counter array in turn (slot 0 belongs
// to the survivor pass of several-base batches).  One atomic per runner, so
// host threads sharing a runner get different slots: a slot comes round again
// only after n-2 other launches took theirs, which bounds how many persistent
// launches may be in flight on different streams at once.
class CounterSlots {
 public:
  explicit CounterSlots(uint32_t n) : n_(n) {}
  uint32_t next() { return 1u + next_.fetch_add(1u, std::memory_order_relaxed) % (n_ - 1u); }

 private:
  const uint32_t n_;
  std::atomic<uint32_t> next_{0};
};

}  // namespace pgx
