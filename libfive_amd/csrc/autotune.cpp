#include "autotune.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>

#include "runtime.hpp"

namespace libfive {
namespace amd {
namespace tune {

namespace {

std::mutex g_mu;
std::map<std::string, int> g_choice;
std::vector<Decision> g_log;

bool initial_enabled() {
    const char* e = std::getenv("LIBFIVE_AMD_AUTOTUNE");
    return !(e && std::strcmp(e, "0") == 0);
}

bool g_enabled = initial_enabled();

// Events of one timing, destroyed on every exit path.
struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    EventPair() {
        hip_check(hipEventCreate(&a), "autotune event");
        hip_check(hipEventCreate(&b), "autotune event");
    }
    ~EventPair() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

}   // namespace

int pick(const std::string& key, const std::vector<std::string>& names,
         const std::function<hipError_t(int)>& run, hipStream_t s, int reps) {
    {
        std::lock_guard<std::mutex> lock(g_mu);
        if (!g_enabled || names.size() <= 1) return 0;
        auto it = g_choice.find(key);
        if (it != g_choice.end()) {
            return (it->second >= 0 && it->second < int(names.size())) ? it->second : 0;
        }
    }
    const int n = int(names.size());
    // one untimed run of every candidate (code loading, cold caches)
    for (int c = 0; c < n; ++c) hip_check(run(c), "autotune warm-up");
    Decision d;
    d.key = key;
    d.names = names;
    d.ms.assign(n, std::numeric_limits<double>::infinity());
    EventPair ev;
    for (int r = 0; r < std::max(1, reps); ++r) {
        for (int c = 0; c < n; ++c) {
            hip_check(hipEventRecord(ev.a, s), "autotune event");
            hip_check(run(c), "autotune run");
            hip_check(hipEventRecord(ev.b, s), "autotune event");
            hip_check(hipEventSynchronize(ev.b), "autotune wait");
            float ms = 0.0f;
            hip_check(hipEventElapsedTime(&ms, ev.a, ev.b), "autotune time");
            d.ms[c] = std::min(d.ms[c], double(ms));
        }
    }
    d.choice = int(std::min_element(d.ms.begin(), d.ms.end()) - d.ms.begin());
    std::lock_guard<std::mutex> lock(g_mu);
    auto it = g_choice.find(key);
    if (it != g_choice.end()) return it->second;     // tuned meanwhile by another thread
    g_choice[key] = d.choice;
    g_log.push_back(d);
    return d.choice;
}

int known(const std::string& key) {
    std::lock_guard<std::mutex> lock(g_mu);
    auto it = g_choice.find(key);
    return it == g_choice.end() ? -1 : it->second;
}

void set_enabled(bool on) {
    std::lock_guard<std::mutex> lock(g_mu);
    g_enabled = on;
}

bool enabled() {
    std::lock_guard<std::mutex> lock(g_mu);
    return g_enabled;
}

std::vector<Decision> decisions() {
    std::lock_guard<std::mutex> lock(g_mu);
    return g_log;
}

void force(const std::string& key, int choice) {
    std::lock_guard<std::mutex> lock(g_mu);
    g_choice[key] = choice;
}

void reset() {
    std::lock_guard<std::mutex> lock(g_mu);
    g_choice.clear();
    g_log.clear();
}

}   // namespace tune
}   // namespace amd
}   // namespace libfive
