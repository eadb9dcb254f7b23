// fa_serve_wait.hpp -- the server's bounded wait for one launch.
//
// A class's completer waits for its launches in stream order.  Waiting with
// hipEventSynchronize has no limit: a launch that never completes (work that
// the device never schedules) would hold the completer, and with it every
// client of the class, forever.  The completer waits through wait_launch
// instead, which gives up after a limit and lets the server fail the class's
// requests with an error.  The function is free of HIP calls so that the CPU
// tests can drive it with a stub event.
#pragma once

#include <stdint.h>

namespace fa {

enum class WaitEnd { Done, Failed, TimedOut };

// query(): 0 once the launch has completed, 1 while it runs, any other value
// for a runtime error.  now(): seconds on a steady clock.  pause(n): idles
// before poll n + 1 (spins first, then sleeps).  limit: seconds.
template <class Query, class Now, class Pause>
WaitEnd wait_launch(Query query, Now now, Pause pause, double limit) {
    const double t0 = now();
    for (uint64_t n = 0;; ++n) {
        const int q = query();
        if (q == 0) return WaitEnd::Done;
        if (q != 1) return WaitEnd::Failed;
        if (now() - t0 > limit) return WaitEnd::TimedOut;
        pause(n);
    }
}

}  // namespace fa
