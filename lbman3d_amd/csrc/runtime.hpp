// runtime.hpp — host-side helpers shared by capi.cpp, batch.cpp and the
// closure runtime: error reporting, configuration translation, the seeded
// counter-based RNG and per-launch event timing.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "../../include/lbman3d.h"
#include "common.hpp"

namespace lbm {

extern thread_local std::string g_last_error;
int fail(int code, const std::string& msg);

void anchored_q_host(const lbm_wall& w, double out[5]);
int validate_config(const lbm_config& c);
int build_slot_table(const lbm_config& c, std::vector<uint16_t>& table);
template <typename T>
BCDesc<T> make_bc(const lbm_config& c);
template <typename T>
Phys<T> make_phys(const lbm_config& c);
bool has_moving_wall(const lbm_config& c);
void feq_host(double rho, double ux, double uy, double uz, double out[kQ]);

// splitmix64 finaliser (Steele, Lea, Flood; "Fast splittable pseudorandom
// number generators", OOPSLA 2014).
inline uint64_t splitmix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// U[0,1) with 53 random bits for the key (seed, cell, k).
inline double uniform01(uint64_t seed, uint64_t cell, uint64_t k) {
    const uint64_t key = (seed * 0x100000001B3ull) ^ (cell * 8ull + k);
    return (double)(splitmix64(splitmix64(key)) >> 11) * (1.0 / 9007199254740992.0);
}

// Event pairs around individual launches on one stream. The events are for
// timing only (finish() synchronises the stream before reading them), so they
// are recorded without the system-scope fence that a default event performs:
// that fence writes back and invalidates the caches at every record, which
// would add its cost to each measured launch and to the launch after it.
class KernelRecorder {
public:
    explicit KernelRecorder(hipStream_t s) : stream_(s) {}
    ~KernelRecorder() {
        for (auto& r : recs_) {
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
    }
    void begin(const char* name) {
        Rec r;
        r.name = name;
        (void)hipEventCreateWithFlags(&r.a, hipEventDisableSystemFence);
        (void)hipEventCreateWithFlags(&r.b, hipEventDisableSystemFence);
        (void)hipEventRecord(r.a, stream_);
        recs_.push_back(r);
    }
    void end() { (void)hipEventRecord(recs_.back().b, stream_); }
    int finish() {
        if (hipStreamSynchronize(stream_) != hipSuccess) return fail(LBM_ERR_HIP, "profile synchronize failed");
        for (auto& r : recs_) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, r.a, r.b) != hipSuccess) return fail(LBM_ERR_HIP, "event timing failed");
            auto& t = totals_[r.name];
            t.first += ms;
            t.second += 1;
        }
        return LBM_OK;
    }
    const std::map<std::string, std::pair<double, int>>& totals() const { return totals_; }

private:
    struct Rec {
        std::string name;
        hipEvent_t a, b;
    };
    hipStream_t stream_;
    std::vector<Rec> recs_;
    std::map<std::string, std::pair<double, int>> totals_;
};

}  // namespace lbm
