// Vertex trajectories of the swept-CCD interface (reference: src/History.h, src/History.cpp).
// A history is a list of (time, position) entries per vertex, the first at t = 0 and, after
// finishHistory, the last at t = 1.  The device path consumes it as start/end arrays when every vertex
// has exactly those two entries, and as CSR arrays otherwise, and stitches on the device; the host
// methods getPosAtTime and stitchCommonHistory are here for narrowphase plugins that call them.
#pragma once

#include <algorithm>
#include <cstdint>
#include <limits>
#include <vector>

#include "EigenTypes.h"

struct HistoryEntry {
    double time;
    Eigen::Vector3d pos;
    bool operator==(const HistoryEntry &o) const { return time == o.time && pos == o.pos; }
};

// History.h:18-22: the positions of several vertices at one common breakpoint.
struct StitchedEntry {
    double time;
    std::vector<Eigen::Vector3d> pos;
};

class History {
public:
    explicit History(const Eigen::VectorXd &qstart)
    {
        int n = (int)(qstart.size() / 3);
        history_.resize(n);
        for (int i = 0; i < n; i++) history_[i].push_back(HistoryEntry{0.0, vec(qstart, 3 * i)});
    }
    void addHistory(int vert, double time, const Eigen::Vector3d &pos) { history_[vert].push_back(HistoryEntry{time, pos}); }
    void finishHistory(const Eigen::VectorXd &qend)
    {
        int n = (int)(qend.size() / 3);
        for (int i = 0; i < n; i++) history_[i].push_back(HistoryEntry{1.0, vec(qend, 3 * i)});
    }
    int countHistoryEntries() const
    {
        int total = 0;
        for (const auto &h : history_) total += (int)h.size();
        return total;
    }
    const std::vector<HistoryEntry> &getVertexHistory(int vert) const { return history_[vert]; }
    int numVertices() const { return (int)history_.size(); }

    // History.cpp (computeMinimumGap): the smallest time step between consecutive entries, at most 1.
    double computeMinimumGap() const
    {
        double result = 1.0;
        for (const auto &h : history_)
            for (size_t j = 1; j < h.size(); j++) result = std::min(h[j].time - h[j - 1].time, result);
        return result;
    }

    // History.cpp:42-76: vertex vert's position at `time` and the index of the entry that starts its
    // segment.  The reference bisects only to choose where a forward scan for the first entry later
    // than `time` starts; that entry is found here with one upper-bound search.  At or after the last
    // entry: the last entry.  (Times before the first entry, which the reference does not handle, give
    // the first entry.)
    void getPosAtTime(int vert, double time, Eigen::Vector3d &pos, int &idx) const
    {
        const std::vector<HistoryEntry> &h = history_[vert];
        const int next = (int)(std::upper_bound(h.begin(), h.end(), time,
                                                [](double t, const HistoryEntry &e) { return t < e.time; }) -
                               h.begin());
        if (next == (int)h.size() || next == 0) {
            idx = next == 0 ? 0 : next - 1;
            pos = h[idx].pos;
            return;
        }
        const int prev = next - 1;
        const double alpha = (time - h[prev].time) / (h[next].time - h[prev].time);
        pos = (1.0 - alpha) * h[prev].pos + alpha * h[next].pos;
        idx = prev;
    }

    // History.cpp:98-140: the trajectories of `verts` on the union of their breakpoints.  From t = 0 each
    // vertex's cursor moves to its last entry at or before the current time; the position there is the
    // cursor's segment interpolated as (b * old + a * new) / dt, with a and b the times since the
    // segment's start and until its end; the next time is the earliest segment end.  Every history
    // holds entries at t = 0 and t = 1 (the reference assumes the same).
    void stitchCommonHistory(const std::vector<int> &verts, std::vector<StitchedEntry> &stitchedHistory) const
    {
        stitchedHistory.clear();
        const size_t n = verts.size();
        std::vector<size_t> cur(n, 0);
        double curtime = 0;
        while (curtime <= 1.0) {
            double newtime = std::numeric_limits<double>::infinity();
            StitchedEntry e;
            e.time = curtime;
            e.pos.resize(n);
            for (size_t i = 0; i < n; i++) {
                const std::vector<HistoryEntry> &h = history_[verts[i]];
                while (cur[i] + 1 < h.size() && h[cur[i] + 1].time <= curtime) cur[i]++;
                const HistoryEntry &old_e = h[cur[i]];
                if (cur[i] + 1 == h.size()) {
                    e.pos[i] = old_e.pos;
                } else {
                    const HistoryEntry &new_e = h[cur[i] + 1];
                    const double dt = new_e.time - old_e.time;
                    const double a = curtime - old_e.time;
                    const double b = new_e.time - curtime;
                    e.pos[i] = (b * old_e.pos + a * new_e.pos) / dt;
                    newtime = std::min(newtime, new_e.time);
                }
            }
            stitchedHistory.push_back(e);
            curtime = newtime;
        }
    }

    // True when every vertex has exactly the entries at t = 0 and t = 1.
    bool twoEntry() const
    {
        for (const auto &h : history_)
            if (h.size() != 2 || h[0].time != 0.0 || h[1].time != 1.0) return false;
        return true;
    }
    void endpoints(std::vector<double> &qs, std::vector<double> &qe) const
    {
        qs.resize(3 * history_.size());
        qe.resize(3 * history_.size());
        for (size_t v = 0; v < history_.size(); v++)
            for (int k = 0; k < 3; k++) {
                qs[3 * v + k] = history_[v].front().pos[k];
                qe[3 * v + k] = history_[v].back().pos[k];
            }
    }
    void csr(std::vector<int64_t> &off, std::vector<double> &t, std::vector<double> &pos) const
    {
        off.assign(history_.size() + 1, 0);
        t.clear();
        pos.clear();
        for (size_t v = 0; v < history_.size(); v++) {
            for (const auto &e : history_[v]) {
                t.push_back(e.time);
                for (int k = 0; k < 3; k++) pos.push_back(e.pos[k]);
            }
            off[v + 1] = (int64_t)t.size();
        }
    }

private:
    static Eigen::Vector3d vec(const Eigen::VectorXd &q, int i) { return Eigen::Vector3d(q[i], q[i + 1], q[i + 2]); }
    std::vector<std::vector<HistoryEntry>> history_;
};
