/*
 * Per-stage timing report of the drainage program (utils/Timer.h,
 * utils/Logging.h:57-89 TIMING_* lines, utils/TimerTree.cpp:78-115 summary
 * table).  A Timer accumulates named intervals; its times come either from
 * the host clock around a blocking stage call or, for the fused device loop,
 * from hipEvent intervals added with add().
 */
#ifndef CAGEO_HOST_TIMER_H_
#define CAGEO_HOST_TIMER_H_

#include <algorithm>
#include <chrono>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

namespace cageo_host {

class Timer {
public:
    explicit Timer(const std::string& target = "") : target_(target) {}
    void startTimer(const std::string& target) {
        target_ = target;
        t0_ = std::chrono::steady_clock::now();
    }
    void endTimer() {
        add(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0_).count());
    }
    /* Add measured time (ms) covering `n` occurrences, e.g. the device
     * event time of one stage summed over n fused realizations. */
    void add(double ms, long n = 1) {
        last_ = n > 0 ? ms / n : ms;
        total_ += ms;
        count_ += n;
    }
    const std::string& getTimingTarget() const { return target_; }
    void setTimingTarget(const std::string& t) { target_ = t; }
    double getLastTime() const { return last_; }
    double getTotalTime() const { return total_; }
    double getAverageTime() const { return count_ ? total_ / count_ : 0.0; }
    long getCount() const { return count_; }

private:
    std::string target_;
    std::chrono::steady_clock::time_point t0_;
    double last_ = 0.0, total_ = 0.0;
    long count_ = 0;
};

/* Tree of timers printed as the reference's summary table: name, total ms,
 * average ms, and percentage of the parent's total. */
class TimerTree {
public:
    explicit TimerTree(const Timer* t) : timer_(t) {}
    ~TimerTree() {
        for (TimerTree* c : children_) delete c;
    }
    TimerTree* setChild(const Timer* t) {
        children_.push_back(new TimerTree(t));
        return children_.back();
    }
    void summary(std::ostream& os = std::cout) const {
        const int depth = maxDepth(), nameWidth = maxNameWidth();
        const int B = kIndent * depth + nameWidth;
        os << "Timer summary" << std::setw(B + kWidthTime + 3 - 13) << "Total time"
           << std::setw(kWidthTime + 3) << "Average time" << std::setw(kWidthPercent) << "% of total"
           << "\n";
        summary(os, 0, timer_->getTotalTime(), nameWidth, depth);
    }

private:
    static const int kIndent = 2, kWidthTime = 11, kWidthPercent = 12;
    void summary(std::ostream& os, int level, double parentTotal, int nameWidth, int depth) const {
        os << std::string((size_t)(kIndent * level), ' ') << std::left
           << std::setw(nameWidth + (depth - level) * kIndent) << timer_->getTimingTarget()
           << std::right << std::fixed << std::setprecision(1) << std::setw(kWidthTime)
           << timer_->getTotalTime() << " ms" << std::setw(kWidthTime) << timer_->getAverageTime()
           << " ms" << std::setw(kWidthPercent)
           << (parentTotal > 0 ? 100.0 * timer_->getTotalTime() / parentTotal : 0.0) << "\n";
        for (const TimerTree* c : children_)
            c->summary(os, level + 1, timer_->getTotalTime(), nameWidth, depth);
    }
    int maxDepth() const {
        int d = 0;
        for (const TimerTree* c : children_) d = std::max(d, 1 + c->maxDepth());
        return d;
    }
    int maxNameWidth() const {
        int w = (int)timer_->getTimingTarget().size();
        for (const TimerTree* c : children_) w = std::max(w, c->maxNameWidth());
        return w;
    }
    const Timer* timer_;
    std::vector<TimerTree*> children_;
};

}  // namespace cageo_host

#endif  // CAGEO_HOST_TIMER_H_
