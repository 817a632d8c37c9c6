// Online autotuner (see autotune.hpp) and its C-ABI report.
#include "autotune.hpp"

#include <algorithm>
#include <cstring>
#include <sstream>

#include "rg_internal.hpp"

namespace rg
{

Autotuner& Autotuner::instance()
{
  static Autotuner t;
  return t;
}

int Autotuner::best(const std::string& key)
{
  std::lock_guard<std::mutex> g(mu_);
  auto it = table_.find(key);
  return it == table_.end() ? -1 : it->second.best;
}

int Autotuner::next(const std::string& key, int ncand, int reps)
{
  std::lock_guard<std::mutex> g(mu_);
  Entry& e = table_[key];
  if (e.runs.empty())
  {
    e.runs.assign(ncand, 0);
    e.bestMs.assign(ncand, 0.0f);
    e.reps = std::max(1, reps);
  }
  if (e.best >= 0) return -1;
  // round robin: the candidate with the fewest timed runs
  int c = 0;
  for (int i = 1; i < static_cast<int>(e.runs.size()); ++i)
    if (e.runs[i] < e.runs[c]) c = i;
  return c;
}

void Autotuner::report(const std::string& key, int cand, float ms)
{
  std::lock_guard<std::mutex> g(mu_);
  auto it = table_.find(key);
  if (it == table_.end() || it->second.best >= 0 || cand < 0 || cand >= static_cast<int>(it->second.runs.size()))
    return;
  Entry& e = it->second;
  e.bestMs[cand] = e.runs[cand] == 0 ? ms : std::min(e.bestMs[cand], ms);
  e.runs[cand] += 1;
  if (*std::min_element(e.runs.begin(), e.runs.end()) >= e.reps)
    e.best = static_cast<int>(std::min_element(e.bestMs.begin(), e.bestMs.end()) - e.bestMs.begin());
}

std::string Autotuner::summary()
{
  std::lock_guard<std::mutex> g(mu_);
  std::ostringstream o;
  for (const auto& kv : table_)
  {
    o << kv.first << ": best=" << kv.second.best << " ms=[";
    for (size_t i = 0; i < kv.second.bestMs.size(); ++i) o << (i ? "," : "") << kv.second.bestMs[i];
    o << "]\n";
  }
  return o.str();
}

}  // namespace rg

extern "C"
{

size_t rg_autotune_summary(char* buf, size_t len)
{
  const std::string s = rg::Autotuner::instance().summary();
  if (buf && len > 0)
  {
    const size_t n = std::min(len - 1, s.size());
    std::memcpy(buf, s.data(), n);
    buf[n] = '\0';
  }
  return s.size();
}

}  // extern "C"
