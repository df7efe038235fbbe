// Optional per-kernel timing with HIP events recorded on the solver's stream
// around selected launches; used by the benchmark to turn algorithmic bytes
// into achieved bandwidth for the dominant kernels.  An event pair costs about
// 5 us of idle GPU time between the kernels it brackets, so a sampling period
// records only every period-th launch of each kernel.
#pragma once

#include <vector>

#include "dev_common.hpp"

namespace pdhcg {

enum TimedKernel {
    kTimedRx = 0,        // R x partial sums + D (dense R), ungated launches
    kTimedRtu = 1,       // R'u (+ diagonal Q) (dense R), ungated launches
    kTimedPrimalStep = 2,// A'y SpMV with the fused primal epilogue / store
    kTimedDualStep = 3,  // A x_bar SpMV with the fused dual epilogue
    kNumTimed = 4
};

class KernelTimer {
  public:
    bool enabled = false;
    int period = 1;  // record every period-th launch of each kernel


    ~KernelTimer() { reset(true); }

    // Returns a started pair slot; call stop() with the returned index.
    int start(int id, hipStream_t s) {
        if (!enabled) return -1;
        if (calls_[id]++ % period != 0) return -1;
        hipEvent_t a = take(), b = take();
        HIP_OK(hipEventRecord(a, s));
        rec_[id].push_back({a, b});
        return (int)rec_[id].size() - 1;
    }
    void stop(int id, int slot, hipStream_t s) {
        if (slot < 0) return;
        HIP_OK(hipEventRecord(rec_[id][slot].second, s));
    }
    // Sum of recorded durations (ms) and launch count; the stream must be idle.
    void totals(int id, double *ms, int *count) {
        double t = 0.0;
        for (auto &p : rec_[id]) {
            float e = 0.0f;
            HIP_OK(hipEventElapsedTime(&e, p.first, p.second));
            t += e;
        }
        *ms = t;
        *count = (int)rec_[id].size();
    }
    void reset(bool destroy = false) {
        for (auto &r : rec_)
            for (auto &p : r) {
                pool_.push_back(p.first);
                pool_.push_back(p.second);
            }
        for (auto &r : rec_) r.clear();
        for (int &c : calls_) c = 0;
        if (destroy) {
            for (hipEvent_t e : pool_) (void)hipEventDestroy(e);
            pool_.clear();
        }
    }

  private:
    hipEvent_t take() {
        if (!pool_.empty()) {
            hipEvent_t e = pool_.back();
            pool_.pop_back();
            return e;
        }
        hipEvent_t e;
        HIP_OK(hipEventCreate(&e));
        return e;
    }
    std::vector<hipEvent_t> pool_;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> rec_[kNumTimed];
    int calls_[kNumTimed] = {0, 0, 0, 0};
};

}  // namespace pdhcg
