// Process-wide owner of the persistent (cooperative) Lanczos form.
//
// A cooperative launch is sized to the co-residency of the whole device
// (lanczos_plan_persistent).  Two such launches from different contexts, on
// different streams, can each get part of the device, and then every
// workgroup of both waits at a device-wide barrier for workgroups that
// cannot become resident.  Only one context of the process may therefore
// run a call that uses the persistent form at a time; a second one is
// refused with USBS_ERR_STATE before it launches anything, and concurrent
// solves use the per-op form.
//
// Header-only and free of HIP types so that a host-only build can check it.
#pragma once
#include <atomic>

namespace usbs {

inline std::atomic<const void*>& coop_owner_slot() {
  static std::atomic<const void*> owner{nullptr};
  return owner;
}

// Holds the persistent form for `owner` (a context) until destruction.  ok()
// is false when another owner holds it.  Nested guards of the same owner
// (a call inside a call of one context) succeed and leave the outer one in
// charge of the release.
class CoopGuard {
 public:
  // owner == nullptr: a call that does not use the persistent form (no-op).
  explicit CoopGuard(const void* owner) : owner_(owner) {
    if (!owner) {
      ok_ = true;
      return;
    }
    const void* expected = nullptr;
    if (coop_owner_slot().compare_exchange_strong(expected, owner, std::memory_order_acq_rel)) {
      held_ = true;
      ok_ = true;
    } else {
      ok_ = (expected == owner);
    }
  }
  ~CoopGuard() {
    if (held_) coop_owner_slot().store(nullptr, std::memory_order_release);
  }
  CoopGuard(const CoopGuard&) = delete;
  CoopGuard& operator=(const CoopGuard&) = delete;
  bool ok() const { return ok_; }

 private:
  const void* owner_;
  bool held_ = false;
  bool ok_ = false;
};

}  // namespace usbs
