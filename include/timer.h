/* Host timers with the reference's names and interfaces (include/timer.h:21-71):
 *
 *   timer t;                    starts at construction
 *   t.milliseconds_elapsed()    time since construction, ms
 *   t.seconds_elapsed()         time since construction, s
 *   anonymouslib_timer u;       u.start(); ...; u.stop() returns ms since start()
 *
 * Both read the monotonic steady clock (the reference's `timer` counts cycles and divides by
 * a fixed CPU_MAX_FREQUENCY, which drifts with the core clock).  Calls into the drop-in
 * library return after their GPU work is complete, so host timing around them measures the
 * whole call, host<->device copies included. */
#ifndef SPOPS_DROPIN_TIMER_H
#define SPOPS_DROPIN_TIMER_H

#include <chrono>

#include "general_config.h"
#include "memopt.h"
#include "plat_config.h"

class timer {
  std::chrono::steady_clock::time_point t0_;

 public:
  timer() : t0_(std::chrono::steady_clock::now()) {}
  double milliseconds_elapsed() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0_).count();
  }
  double seconds_elapsed() { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count(); }
};

struct anonymouslib_timer {
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  void start() { t0 = std::chrono::steady_clock::now(); }
  double stop() { return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(); }
};

#endif
