// Vertex trajectories of the swept-CCD interface (reference: src/History.h, src/History.cpp:8-39).
// A history is a list of (time, position) entries per vertex, the first at t = 0 and, after
// finishHistory, the last at t = 1.  The device path consumes it as start/end arrays when every vertex
// has exactly those two entries, and as CSR arrays otherwise; stitching happens on the device.
#pragma once

#include <algorithm>
#include <cstdint>
#include <vector>

#include "EigenTypes.h"

struct HistoryEntry {
    double time;
    Eigen::Vector3d pos;
    bool operator==(const HistoryEntry &o) const { return time == o.time && pos == o.pos; }
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
