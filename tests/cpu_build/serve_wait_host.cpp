// Host build of the server's bounded launch wait
// (fast-atari-rs_amd/csrc/fa_serve_wait.hpp) driven by a stub event and a
// simulated clock, for the CPU tests only.  Not part of the library.
#include <stdint.h>

#include "../../fast-atari-rs_amd/csrc/fa_serve_wait.hpp"

// A stub launch that completes after `ready_after` polls (never when < 0), or
// reports a runtime error at poll `error_at` (none when < 0).  Each pause
// advances the simulated clock by `step` seconds.  Returns 0 Done, 1 Failed,
// 2 TimedOut; *polls and *elapsed report how long the wait took.
extern "C" int sw_run(int64_t ready_after, int64_t error_at, double step, double limit, uint64_t *polls,
                      double *elapsed) {
    double t = 0.0;
    uint64_t n = 0;
    const fa::WaitEnd w = fa::wait_launch(
        [&] {
            const int64_t i = (int64_t)n++;
            if (error_at >= 0 && i >= error_at) return 2;
            return (ready_after >= 0 && i >= ready_after) ? 0 : 1;
        },
        [&] { return t; }, [&](uint64_t) { t += step; }, limit);
    *polls = n;
    *elapsed = t;
    return w == fa::WaitEnd::Done ? 0 : (w == fa::WaitEnd::Failed ? 1 : 2);
}
