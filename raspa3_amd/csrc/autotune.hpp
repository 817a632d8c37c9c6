// Online autotuner for launch configurations.
//
// A tuned launch site offers a few candidate configurations that compute the same result (they differ in launch
// shape or work split only).  While a problem key is open, the site asks for the next candidate, launches the
// caller's real work with it between two HIP events on the caller's stream, and reports the elapsed time; after
// every candidate has been timed 'reps' times the fastest (by its best time) is fixed for that key for the rest
// of the process.  Tuning therefore costs no extra work and never changes what is computed.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace rg
{

class Autotuner
{
public:
  static Autotuner& instance();

  // Fixed choice for 'key', or -1 while it is still being tuned.
  int best(const std::string& key);
  // Candidate to time next for 'key' (registering the key with 'ncand' candidates), or -1 once tuning is done.
  int next(const std::string& key, int ncand, int reps);
  // Elapsed time of one timed launch of candidate 'cand'; fixes the winner when every candidate has 'reps' times.
  void report(const std::string& key, int cand, float ms);
  // One line per key: "key: best=i times=[...]".
  std::string summary();

private:
  struct Entry
  {
    int best = -1;
    int reps = 1;
    std::vector<float> bestMs;
    std::vector<int> runs;
  };
  std::mutex mu_;
  std::map<std::string, Entry> table_;
};

// Times one launch: records an event before and after 'launch' on 'stream' and waits for the second.
template <class F>
hipError_t timeLaunch(hipStream_t stream, F&& launch, float& ms)
{
  hipEvent_t a = nullptr, b = nullptr;
  hipError_t e = hipEventCreate(&a);
  if (e == hipSuccess) e = hipEventCreate(&b);
  if (e == hipSuccess) e = hipEventRecord(a, stream);
  if (e == hipSuccess) e = launch();
  if (e == hipSuccess) e = hipEventRecord(b, stream);
  if (e == hipSuccess) e = hipEventSynchronize(b);
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  return e;
}

}  // namespace rg
