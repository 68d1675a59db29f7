// lb_sync.h -- exclusion between an engine's graph capture and the GPU calls
// of other host threads (internal to libludwig_hip.so).
//
// A serving process builds new cases on set-up threads (allocations,
// synchronous copies, frees, the Q-map kernel on a private stream) while the
// main thread steps the engine.  Some of those calls invalidate a stream
// capture in progress on another thread, whatever the capture mode, so the
// capture of an engine step takes the lock exclusively and every other entry
// point that touches the GPU holds it shared for its duration.  Set-up
// threads still run concurrently with each other and with graph replays; only
// the capture itself (a few milliseconds of launch recording) waits for them.
#pragma once

#include <shared_mutex>

namespace lbsync {

std::shared_mutex& capture_mutex();

// Shared hold for one C-ABI call; re-entrant within a thread (an entry point
// that calls another takes the lock once).
class SetupGuard {
public:
    SetupGuard() {
        if (depth()++ == 0) capture_mutex().lock_shared();
    }
    ~SetupGuard() {
        if (--depth() == 0) capture_mutex().unlock_shared();
    }
    SetupGuard(const SetupGuard&) = delete;
    SetupGuard& operator=(const SetupGuard&) = delete;

private:
    static int& depth() {
        thread_local int d = 0;
        return d;
    }
};

}  // namespace lbsync
