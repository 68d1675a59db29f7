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
// The lock prefers the writer: once a capture waits for it, new shared
// holders wait too, so a stream of set-up calls (a serving run sets up dozens
// of cases back to back on several threads) cannot hold the capture off.  A
// reader-preferring std::shared_mutex did exactly that: the engine stood
// idle for 4 s of a 23 s run until every pending case was set up.
#pragma once

#include <condition_variable>
#include <mutex>

namespace lbsync {

class CaptureLock {
public:
    void lock() {
        std::unique_lock<std::mutex> l(m_);
        writers_waiting_++;
        cv_.wait(l, [this] { return !writer_ && readers_ == 0; });
        writers_waiting_--;
        writer_ = true;
    }
    void unlock() {
        {
            std::lock_guard<std::mutex> l(m_);
            writer_ = false;
        }
        cv_.notify_all();
    }
    void lock_shared() {
        std::unique_lock<std::mutex> l(m_);
        cv_.wait(l, [this] { return !writer_ && writers_waiting_ == 0; });
        readers_++;
    }
    void unlock_shared() {
        bool last;
        {
            std::lock_guard<std::mutex> l(m_);
            last = --readers_ == 0;
        }
        if (last) cv_.notify_all();
    }

private:
    std::mutex m_;
    std::condition_variable cv_;
    int readers_ = 0;
    int writers_waiting_ = 0;
    bool writer_ = false;
};

CaptureLock& capture_mutex();

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
