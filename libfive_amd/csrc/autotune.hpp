// Launch-schedule autotuner.
//
// Several kernels can run the same work under more than one launch
// schedule: how many waves share a workgroup in the mesher's leaf passes,
// how many points each lane of the tape interpreter evaluates, the
// workgroup size of the specialised evaluation kernels.  Every schedule of
// a kernel computes the same results bit for bit; only the time differs,
// and which one is fastest depends on the tape (register banks, opcode mix)
// and on the device.
//
// The first time a kernel meets a key (its family plus the properties that
// select a different best schedule, such as the register bank count), the
// tuner launches every candidate on the caller's stream over the same
// slice of the real work, times each with HIP events (one untimed run
// each, then the minimum of `reps` timed runs) and keeps the fastest for
// the rest of the process.  Later calls with that key use the kept choice
// without timing anything.  Candidates must be idempotent on the slice
// (running one again, or another after it, leaves the same outputs).
//
// LIBFIVE_AMD_AUTOTUNE=0 turns tuning off (every kernel uses candidate 0,
// its default schedule); libfive_autotune_set() does the same at run time.
#pragma once

#include <hip/hip_runtime.h>

#include <functional>
#include <string>
#include <vector>

namespace libfive {
namespace amd {
namespace tune {

struct Decision {
    std::string key;                 // kernel family and shape class
    std::vector<std::string> names;  // candidate schedules
    std::vector<double> ms;          // minimum timed run of each candidate
    int choice = 0;                  // index of the schedule kept
};

// Returns the schedule to use for `key`: the kept choice, or, on the first
// call with this key, the fastest of the candidates after timing them with
// run(c) on stream s.  Returns 0 when tuning is off.  run(c) enqueues
// candidate c's work on s and returns its launch status.
int pick(const std::string& key, const std::vector<std::string>& names,
         const std::function<hipError_t(int)>& run, hipStream_t s, int reps = 2);

// The kept choice for key, or -1 when the key has not been tuned.
int known(const std::string& key);

void set_enabled(bool on);
bool enabled();

// Every decision made in this process, in the order they were made.
std::vector<Decision> decisions();

// Forgets every decision (the next call of each key tunes again).
void reset();

// Makes `choice` the kept decision for key without timing (for tests that
// run every schedule).
void force(const std::string& key, int choice);

}   // namespace tune
}   // namespace amd
}   // namespace libfive
