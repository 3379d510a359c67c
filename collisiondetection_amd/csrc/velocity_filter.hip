// Velocity filter: the layered penalty response of evouga/collisiondetection (include/VelocityFilter.h:7-45,
// src/VelocityFilter.cpp:16-87, src/ActiveLayers.cpp:23-256, src/PenaltyGroup.cpp:7-71,
// src/PenaltyPotential.cpp:9-61) driving this library's device collision queries.
//
// Division of labour.  The penalty integration is an event loop: groups of stencils ("layers") fire
// one at a time in the order of their next firing times, each firing reading and writing only the
// positions and velocities of its own vertices.  Every firing depends on the one before it, a group
// holds a few to a few hundred stencils, and a run makes 10^7..10^8 firings, so the loop stays on the
// host, where a firing costs well under a microsecond; a device round trip per firing would cost tens
// of microseconds.  What the reference spends its time on besides -- the O(V*F) meshSelfDistance,
// the k-DOP broadphase over the whole multi-entry history and the separating-plane narrowphase over
// every candidate -- runs on the device through this library's entry points, once per outer iteration.
//
// The loop is the reference's, operation for operation (the penalty arithmetic in the reference's
// expression order, the same std::priority_queue discipline for choosing the next group), but a firing
// touches only its group's vertices: the reference sizes and clears full-length vectors per firing
// (ActiveLayers.cpp:88-90, :111; PenaltyGroup.cpp:36-37), whose entries outside the group are never
// read.  A firing in which no penalty fires leaves the state as it was (every force entry is +0, so no
// vertex counts as touched, ActiveLayers.cpp:116-123) and is cut short after the penalty tests.
//
// Quiet groups.  Most firings apply no force: a penalty needs its stencil closer than the layer's
// thickness.  After a firing in which nothing fired, the group records a time before which none of its
// stencils can get that close while its vertices keep their velocities (the distance between a vertex
// and a triangle, or between two segments, changes no faster than the sum of the largest speeds on
// either side), with an absolute margin far above the rounding error of the distance evaluation.  Until
// then its firings only advance its step counter -- exactly what an evaluation would have done -- and
// any touch of one of its vertices by any group cancels the record.  The firing sequence, the priority
// queue's pushes and pops and every result stay those of the full evaluation.
#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <map>
#include <memory>
#include <queue>
#include <vector>

#include "ccd_core.h"
#include "ccd_internal.h"

namespace {

using ccd::V3;

struct Key4 {
    int32_t k[4];
};
struct KeyLess {   // Stencils.h:59-77, :119-137: lexicographic on the four indices
    bool operator()(const Key4 &a, const Key4 &b) const
    {
        for (int i = 0; i < 4; i++) {
            if (a.k[i] != b.k[i]) return a.k[i] < b.k[i];
        }
        return false;
    }
};

// One penalty group (PenaltyGroup.h:12-42): its stencils in insertion order, and the distinct vertices
// they touch in ascending order (the std::set<int> groupStencil_ of the reference).  slots map each
// stencil's 4 vertices to positions in verts.
struct Layer {
    double dt, outer, inner, k, cor;
    int next = 1;
    double t_next = 0;                 // next * dt (PenaltyGroup.cpp:59-62), kept with next
    double quiet_until = -INFINITY;    // no penalty of this group can fire before this time
    std::vector<Key4> vf, ee;
    std::vector<int32_t> verts, vf_slot, ee_slot;
    bool dirty = false;
    double fire_time() const { return t_next; }
    void set_next(int n)
    {
        next = n;
        t_next = next * dt;
    }

    void rebuild()
    {
        verts.clear();
        for (const Key4 &s : vf) verts.insert(verts.end(), s.k, s.k + 4);
        for (const Key4 &s : ee) verts.insert(verts.end(), s.k, s.k + 4);
        std::sort(verts.begin(), verts.end());
        verts.erase(std::unique(verts.begin(), verts.end()), verts.end());
        auto slot = [&](int32_t v) { return (int32_t)(std::lower_bound(verts.begin(), verts.end(), v) - verts.begin()); };
        vf_slot.resize(4 * vf.size());
        ee_slot.resize(4 * ee.size());
        for (size_t i = 0; i < vf.size(); i++)
            for (int j = 0; j < 4; j++) vf_slot[4 * i + j] = slot(vf[i].k[j]);
        for (size_t i = 0; i < ee.size(); i++)
            for (int j = 0; j < 4; j++) ee_slot[4 * i + j] = slot(ee[i].k[j]);
        dirty = false;
    }
};
struct LayerCmp {   // ActiveLayers.cpp:18-21: the group with the earliest next firing on top
    bool operator()(const Layer *a, const Layer *b) const { return b->t_next < a->t_next; }
};

inline V3 ld(const double *x, int32_t slot) { return ccd::mk(x[3 * slot], x[3 * slot + 1], x[3 * slot + 2]); }
inline void acc(double *g, int32_t slot, double s, V3 f)   // g.segment<3>(3*slot) += s * f
{
    g[3 * slot] = g[3 * slot] + s * f.x;
    g[3 * slot + 1] = g[3 * slot + 1] + s * f.y;
    g[3 * slot + 2] = g[3 * slot + 2] + s * f.z;
}
inline void dec(double *g, int32_t slot, double s, V3 f)   // g.segment<3>(3*slot) -= s * f
{
    g[3 * slot] = g[3 * slot] - s * f.x;
    g[3 * slot + 1] = g[3 * slot + 1] - s * f.y;
    g[3 * slot + 2] = g[3 * slot + 2] - s * f.z;
}

// VertexFacePenaltyPotential::addForce (PenaltyPotential.cpp:9-31) on slot-indexed group arrays.
inline bool vf_penalty(const double *q, const double *v, double *g, const int32_t *sl, double outer, double inner,
                       double k, double cor)
{
    V3 p = ld(q, sl[0]), a = ld(q, sl[1]), b = ld(q, sl[2]), c = ld(q, sl[3]);
    if (!ccd::vertex_plane_distance_less_than(p, a, b, c, outer)) return false;
    double b0, b1, b2;
    V3 cv = ccd::vertex_face_distance(p, a, b, c, b0, b1, b2);
    double dist = ccd::norm(cv);
    if (dist >= outer || dist < inner) return false;
    V3 vp = ld(v, sl[0]);
    V3 rel = ((ccd::mk(-vp.x, -vp.y, -vp.z) + b0 * ld(v, sl[1])) + b1 * ld(v, sl[2])) + b2 * ld(v, sl[3]);
    if (ccd::dot(rel, cv) > 0) k *= cor;
    double s = k * (outer - dist) / (outer - inner);
    V3 lf = ccd::mk((s * cv.x) / dist, (s * cv.y) / dist, (s * cv.z) / dist);
    g[3 * sl[0]] = g[3 * sl[0]] - lf.x;
    g[3 * sl[0] + 1] = g[3 * sl[0] + 1] - lf.y;
    g[3 * sl[0] + 2] = g[3 * sl[0] + 2] - lf.z;
    acc(g, sl[1], b0, lf);
    acc(g, sl[2], b1, lf);
    acc(g, sl[3], b2, lf);
    return true;
}

// EdgeEdgePenaltyPotential::addForce (PenaltyPotential.cpp:34-61).
inline bool ee_penalty(const double *q, const double *v, double *g, const int32_t *sl, double outer, double inner,
                       double k, double cor)
{
    V3 p0 = ld(q, sl[0]), p1 = ld(q, sl[1]), q0 = ld(q, sl[2]), q1 = ld(q, sl[3]);
    if (!ccd::line_line_distance_less_than(p0, p1, q0, q1, outer)) return false;
    double bp0, bp1, bq0, bq1;
    V3 cv = ccd::edge_edge_distance(p0, p1, q0, q1, bp0, bp1, bq0, bq1);
    double dist = ccd::norm(cv);
    if (dist >= outer || dist < inner) return false;
    V3 rel = (((-bp0) * ld(v, sl[0]) - bp1 * ld(v, sl[1])) + bq0 * ld(v, sl[2])) + bq1 * ld(v, sl[3]);
    if (ccd::dot(rel, cv) > 0) k *= cor;
    double s = k * (outer - dist) / (outer - inner);
    V3 lf = ccd::mk((s * cv.x) / dist, (s * cv.y) / dist, (s * cv.z) / dist);
    dec(g, sl[0], bp0, lf);
    dec(g, sl[1], bp1, lf);
    acc(g, sl[2], bq0, lf);
    acc(g, sl[3], bq1, lf);
    return true;
}

struct HEntry {
    double t, x, y, z;
};

// SimulationState.h: positions, velocities, inverse masses and per-DOF last update times.
struct State {
    std::vector<double> q, v, minv, last;
};

using Clock = std::chrono::steady_clock;
inline double secs(Clock::time_point a) { return std::chrono::duration<double>(Clock::now() - a).count(); }

// ActiveLayers (ActiveLayers.h:23-75) over the device queries.  Device arrays: faces (3F), fixed (V or null).
class Filter {
public:
    Filter(ccd_ctx *ctx, const int32_t *d_faces, int64_t F, int64_t V, const uint8_t *d_fixed, double outer,
           double inner, double dt, double k, double cor, bool verbose, ccd_vf_stats &st)
        : ctx_(ctx), faces_(d_faces), F_(F), V_(V), fixed_(d_fixed), outer_(outer), inner_(inner), dt_(dt), k_(k),
          cor_(cor), verbose_(verbose), st_(st), hist_(V)
    {
    }

    // ActiveLayers::runOneIteration (ActiveLayers.cpp:217-256).  *clean receives whether no collision was
    // found.  Returns a status code.
    int run_iteration(State &s, bool *clean)
    {
        if (verbose_) {
            printf("Taking an outer iteration, deepest layer is currently %d", deepest_);
            if (deepest_) printf(" with outer thickness %g and dt %g", layers_[deepest_ - 1]->outer, layers_[deepest_ - 1]->dt);
            printf("\n");
        }
        for (int64_t v = 0; v < V_; v++) {   // History(s.q), History.cpp:8-16
            hist_[v].clear();
            hist_[v].push_back(HEntry{0.0, s.q[3 * v], s.q[3 * v + 1], s.q[3 * v + 2]});
        }
        index_groups();
        Clock::time_point t0 = Clock::now();
        while (!step(s)) {
        }
        st_.host_seconds += secs(t0);
        int64_t entries = 0;
        for (const auto &h : hist_) entries += (int64_t)h.size();
        st_.history_entries = entries;
        if (verbose_) printf("Done simulating, accumulated %lld history entries\n", (long long)entries);
        std::vector<Key4> vfs, ees;
        t0 = Clock::now();
        int rc = detect(vfs, ees);
        st_.device_seconds += secs(t0);
        if (rc) return rc;
        if (verbose_)
            printf("Found %lld vertex-face and %lld edge-edge collisions\n", (long long)vfs.size(), (long long)ees.size());
        rollback();
        for (const Key4 &key : vfs) add_stencil(key, false);
        for (const Key4 &key : ees) add_stencil(key, true);
        st_.layers = deepest_;
        *clean = vfs.empty() && ees.empty();
        return CCD_OK;
    }

private:
    double layer_depth(int layer) const { return inner_ + (outer_ - inner_) / double(layer); }   // :167-170

    void add_groups(int maxdepth)
    {   // ActiveLayers.cpp:67-83
        while (deepest_ < maxdepth) {
            int depth = deepest_ + 1;
            const double fudge = 1e-4;   // keeps the layers' firing times from coinciding
            std::unique_ptr<Layer> g(new Layer);
            g->k = k_ * depth * depth * depth;
            g->outer = layer_depth(depth);
            g->inner = inner_;
            g->dt = dt_ / double(depth) / std::sqrt(double(depth) + fudge);
            g->cor = cor_;
            g->set_next(1);
            queue_.push(g.get());
            layers_.push_back(std::move(g));
            deepest_++;
        }
    }

    void add_stencil(const Key4 &key, bool ee)
    {   // ActiveLayers.cpp:41-65: the stencil joins the next 5 layers below its current depth
        int &depth = (ee ? eedepth_ : vfdepth_)[key];
        int old = depth;
        for (int i = 0; i < 5; i++) {
            add_groups(old + i + 1);
            Layer &g = *layers_[old + i];
            (ee ? g.ee : g.vf).push_back(key);
            g.dirty = true;
            depth++;
        }
    }

    void rollback()
    {   // ActiveLayers.cpp:152-165
        for (auto &g : layers_) g->set_next(1);
        std::priority_queue<Layer *, std::vector<Layer *>, LayerCmp> nq;
        while (!queue_.empty()) {
            nq.push(queue_.top());
            queue_.pop();
        }
        queue_ = nq;
    }

    double thickness(const Key4 &key, bool ee) const
    {   // ActiveLayers.cpp:172-186
        const auto &m = ee ? eedepth_ : vfdepth_;
        auto it = m.find(key);
        return it != m.end() ? layer_depth(it->second + 1) : layer_depth(1);
    }

    bool step(State &s)
    {   // ActiveLayers.cpp:85-150, with termination time 1 (VelocityFilter.cpp:49)
        if (!queue_.empty()) {
            Layer *g = queue_.top();
            double newtime = g->fire_time();
            if (newtime < 1.0) {
                queue_.pop();
                fire(*g, s, newtime);
                g->set_next(g->next + 1);
                queue_.push(g);
                return false;
            }
        }
        for (size_t i = 0; i < s.q.size(); i++) {
            s.q[i] = s.q[i] + (1.0 - s.last[i]) * s.v[i];
            s.last[i] = 1.0;
        }
        for (int64_t v = 0; v < V_; v++)   // History::finishHistory, History.cpp:31-39
            hist_[v].push_back(HEntry{1.0, s.q[3 * v], s.q[3 * v + 1], s.q[3 * v + 2]});
        return true;
    }

    // Vertex -> groups containing it (CSR over layer indices), and every group's quiet record cleared.
    void index_groups()
    {
        vg_off_.assign(V_ + 1, 0);
        for (auto &g : layers_) {
            if (g->dirty) g->rebuild();
            g->quiet_until = -INFINITY;
            for (int32_t v : g->verts) vg_off_[v + 1]++;
        }
        for (int64_t v = 0; v < V_; v++) vg_off_[v + 1] += vg_off_[v];
        vg_idx_.resize(vg_off_[V_]);
        std::vector<int64_t> pos(vg_off_.begin(), vg_off_.end() - 1);
        for (size_t i = 0; i < layers_.size(); i++)
            for (int32_t v : layers_[i]->verts) vg_idx_[pos[v]++] = (int32_t)i;
    }

    // A time before which no stencil of g can come within g.outer, from the group's positions nq_ at t0
    // and its vertices' velocities; -inf when some stencil is already that close (or not a number).
    double quiet_bound(const Layer &g, const State &s, double t0) const
    {
        auto speed = [&](int32_t slot) {
            const double *v = &s.v[3 * (size_t)g.verts[slot]];
            return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        };
        auto mag = [&](int32_t slot) {
            const double *x = &nq_[3 * (size_t)slot];
            return std::max(std::fabs(x[0]), std::max(std::fabs(x[1]), std::fabs(x[2])));
        };
        double lim = INFINITY;
        for (int pass = 0; pass < 2; pass++) {
            const bool ee = pass == 1;
            const std::vector<int32_t> &slots = ee ? g.ee_slot : g.vf_slot;
            for (size_t i = 0; i < slots.size(); i += 4) {
                const int32_t *sl = &slots[i];
                V3 a = ld(nq_.data(), sl[0]), b = ld(nq_.data(), sl[1]), c = ld(nq_.data(), sl[2]),
                   d = ld(nq_.data(), sl[3]);
                double w0, w1, w2, w3, S;
                double dist;
                if (ee) {
                    dist = ccd::norm(ccd::edge_edge_distance(a, b, c, d, w0, w1, w2, w3));
                    S = std::max(speed(sl[0]), speed(sl[1])) + std::max(speed(sl[2]), speed(sl[3]));
                } else {
                    dist = ccd::norm(ccd::vertex_face_distance(a, b, c, d, w0, w1, w2));
                    S = speed(sl[0]) + std::max(speed(sl[1]), std::max(speed(sl[2]), speed(sl[3])));
                }
                // Rounding of positions and distances is a few ulps of the coordinates' magnitude over
                // the rest of [0, 1]; the margin is 10^6 times that.
                double scale = std::max(std::max(mag(sl[0]), mag(sl[1])), std::max(mag(sl[2]), mag(sl[3]))) + S;
                double gap = dist - g.outer - 1e-10 * (1.0 + scale);
                if (!(gap > 0)) return -INFINITY;
                if (S > 0) lim = std::min(lim, gap / (S * (1.0 + 1e-9)));
            }
        }
        return t0 + 0.999999 * lim;
    }

    void fire(Layer &g, State &s, double newtime)
    {   // ActiveLayers.cpp:94-137 with PenaltyGroup::addForce (PenaltyGroup.cpp:35-53)
        st_.firings++;
        if (newtime < g.quiet_until) return;
        if (g.dirty) g.rebuild();
        const size_t m = g.verts.size();
        if (m == 0) return;
        nq_.resize(3 * m);
        nv_.resize(3 * m);
        gf_.assign(3 * m, 0.0);
        for (size_t i = 0; i < m; i++)
            for (int j = 0; j < 3; j++) {
                size_t idx = 3 * (size_t)g.verts[i] + j;
                double dt = newtime - s.last[idx];
                double x = s.q[idx] + (newtime - s.last[idx]) * s.v[idx];
                nq_[3 * i + j] = x;
                nv_[3 * i + j] = (x - s.q[idx]) / dt;
            }
        bool fired = false;
        for (size_t i = 0; i < g.vf.size(); i++)
            fired |= vf_penalty(nq_.data(), nv_.data(), gf_.data(), &g.vf_slot[4 * i], g.outer, g.inner, g.k, g.cor);
        for (size_t i = 0; i < g.ee.size(); i++)
            fired |= ee_penalty(nq_.data(), nv_.data(), gf_.data(), &g.ee_slot[4 * i], g.outer, g.inner, g.k, g.cor);
        if (!fired) {
            g.quiet_until = quiet_bound(g, s, newtime);
            return;
        }
        for (size_t i = 0; i < m; i++) {
            double F[3];
            bool touched = false;
            for (int j = 0; j < 3; j++) {
                F[j] = 0.0 + gf_[3 * i + j] * g.dt;
                if (F[j] != 0.0) touched = true;
            }
            if (!touched) continue;
            size_t b = 3 * (size_t)g.verts[i];
            for (int j = 0; j < 3; j++) {
                s.v[b + j] = s.v[b + j] + s.minv[b + j] * F[j];
                s.q[b + j] = nq_[3 * i + j];
                s.last[b + j] = newtime;
            }
            hist_[g.verts[i]].push_back(HEntry{newtime, s.q[b], s.q[b + 1], s.q[b + 2]});
            for (int64_t j = vg_off_[g.verts[i]]; j < vg_off_[g.verts[i] + 1]; j++)
                layers_[vg_idx_[j]]->quiet_until = -INFINITY;   // its velocity changed
        }
    }

    // ActiveLayers::collisionDetection (ActiveLayers.cpp:188-215): the k-DOP broadphase over the history
    // grown by outerEta, then the separating-plane narrowphase with each stencil's layer thickness.
    int detect(std::vector<Key4> &vfs, std::vector<Key4> &ees)
    {
        const int64_t E = st_.history_entries;
        off_.resize(V_ + 1);
        ht_.resize(E);
        hp_.resize(3 * E);
        off_[0] = 0;
        int64_t e = 0;
        for (int64_t v = 0; v < V_; v++) {
            for (const HEntry &h : hist_[v]) {
                ht_[e] = h.t;
                hp_[3 * e] = h.x;
                hp_[3 * e + 1] = h.y;
                hp_[3 * e + 2] = h.z;
                e++;
            }
            off_[v + 1] = e;
        }
        int64_t *d_off = (int64_t *)ccdi::buffer(ctx_, "vfl_off", 8 * (V_ + 1));
        double *d_t = (double *)ccdi::buffer(ctx_, "vfl_t", 8 * E);
        double *d_p = (double *)ccdi::buffer(ctx_, "vfl_p", 24 * E);
        if (!d_off || !d_t || !d_p) return CCD_E_OOM;
        int rc;
        if ((rc = ccd_copy_to_device(ctx_, d_off, off_.data(), 8 * (V_ + 1)))) return rc;
        if ((rc = ccd_copy_to_device(ctx_, d_t, ht_.data(), 8 * E))) return rc;
        if ((rc = ccd_copy_to_device(ctx_, d_p, hp_.data(), 24 * E))) return rc;

        int64_t nv = 0, ne = 0;
        int32_t *d_vk = nullptr, *d_ek = nullptr;
        for (int attempt = 0; attempt < 2; attempt++) {
            d_vk = (int32_t *)ccdi::buffer(ctx_, "vfl_vk", 16 * (cap_v_ + 1));
            d_ek = (int32_t *)ccdi::buffer(ctx_, "vfl_ek", 16 * (cap_e_ + 1));
            if (!d_vk || !d_ek) return CCD_E_OOM;
            rc = ccd_kdop_candidates_history(ctx_, faces_, F_, d_off, d_p, V_, E, outer_, fixed_, d_vk, cap_v_, &nv,
                                             d_ek, cap_e_, &ne, CCD_SORTED);
            if (rc == CCD_E_CAPACITY && attempt == 0) {
                cap_v_ = std::max(cap_v_, nv);
                cap_e_ = std::max(cap_e_, ne);
                continue;
            }
            break;
        }
        if (rc) return rc;
        st_.vf_candidates += nv;
        st_.ee_candidates += ne;
        if (verbose_)
            printf("Broad phase found %lld vertex-face and %lld edge-edge candidates\n", (long long)nv, (long long)ne);
        if (nv + ne == 0) return CCD_OK;

        keys_.resize(4 * (nv + ne));
        if ((rc = ccd_copy_to_host(ctx_, keys_.data(), d_vk, 16 * nv))) return rc;
        if ((rc = ccd_copy_to_host(ctx_, keys_.data() + 4 * nv, d_ek, 16 * ne))) return rc;
        eta_.resize(nv + ne);
        for (int64_t i = 0; i < nv + ne; i++) {
            Key4 key{{keys_[4 * i], keys_[4 * i + 1], keys_[4 * i + 2], keys_[4 * i + 3]}};
            eta_[i] = thickness(key, i >= nv);
        }
        double *d_eta = (double *)ccdi::buffer(ctx_, "vfl_eta", 8 * (nv + ne));
        uint8_t *d_st = (uint8_t *)ccdi::buffer(ctx_, "vfl_status", nv + ne + 2);
        if (!d_eta || !d_st) return CCD_E_OOM;
        if ((rc = ccd_copy_to_device(ctx_, d_eta, eta_.data(), 8 * (nv + ne)))) return rc;
        int64_t undecided = 0;
        rc = ccd_separating_plane_narrowphase(ctx_, d_off, d_t, d_p, V_, E, d_vk, d_eta, nv, d_ek, d_eta + nv, ne, 0.0,
                                              d_st, d_st + nv, &undecided);
        if (rc) return rc;
        if (undecided)
            return ccdi::set_error(ctx_, CCD_E_INVALID,
                                   "velocity filter: the separating-plane recursion does not terminate for %lld "
                                   "stencil(s) (the reference overflows its stack there)",
                                   (long long)undecided);
        status_.resize(nv + ne);
        if ((rc = ccd_copy_to_host(ctx_, status_.data(), d_st, nv + ne))) return rc;
        for (int64_t i = 0; i < nv + ne; i++) {
            if (status_[i] != 1) continue;
            Key4 key{{keys_[4 * i], keys_[4 * i + 1], keys_[4 * i + 2], keys_[4 * i + 3]}};
            (i < nv ? vfs : ees).push_back(key);
        }
        st_.vf_collisions += (int64_t)vfs.size();
        st_.ee_collisions += (int64_t)ees.size();
        return CCD_OK;
    }

    ccd_ctx *ctx_;
    const int32_t *faces_;
    int64_t F_, V_;
    const uint8_t *fixed_;
    double outer_, inner_, dt_, k_, cor_;
    bool verbose_;
    ccd_vf_stats &st_;
    int deepest_ = 0;
    std::vector<std::unique_ptr<Layer>> layers_;
    std::priority_queue<Layer *, std::vector<Layer *>, LayerCmp> queue_;
    std::map<Key4, int, KeyLess> vfdepth_, eedepth_;
    std::vector<std::vector<HEntry>> hist_;
    std::vector<double> nq_, nv_, gf_;
    std::vector<int64_t> vg_off_;
    std::vector<int32_t> vg_idx_;
    std::vector<int64_t> off_;
    std::vector<double> ht_, hp_, eta_;
    std::vector<int32_t> keys_;
    std::vector<uint8_t> status_;
    int64_t cap_v_ = 0, cap_e_ = 0;
};

}   // namespace

extern "C" int ccd_velocity_filter(ccd_ctx *ctx, const double *qstart, double *qend, int64_t nverts,
                                   const int32_t *faces, int64_t nfaces, const double *inv_masses, double outer_eta,
                                   double inner_eta, double base_stiffness, double base_substep, double cor,
                                   int max_rollbacks, int verbose, int *result, ccd_vf_stats *stats)
{
    if (!ctx) return CCD_E_INVALID;
    if (!result) return ccdi::set_error(ctx, CCD_E_INVALID, "null result");
    if (nverts < 0 || nfaces < 0) return ccdi::set_error(ctx, CCD_E_INVALID, "negative size");
    if (nfaces >= (int64_t)1 << 30 || nverts >= (int64_t)1 << 31)
        return ccdi::set_error(ctx, CCD_E_INVALID, "mesh too large for 32-bit indices");
    if ((nverts && (!qstart || !qend || !inv_masses)) || (nfaces && !faces))
        return ccdi::set_error(ctx, CCD_E_INVALID, "null mesh arrays");
    ccd_vf_stats st{};
    if (stats) *stats = st;

    // VelocityFilter.cpp:24-37: every vertex must belong to a face.
    std::vector<uint8_t> used(nverts, 0);
    for (int64_t i = 0; i < 3 * nfaces; i++) {
        if (faces[i] < 0 || faces[i] >= nverts)
            return ccdi::set_error(ctx, CCD_E_INVALID, "face index %d out of range [0, %lld)", faces[i],
                                   (long long)nverts);
        used[faces[i]] = 1;
    }
    for (int64_t i = 0; i < nverts; i++)
        if (!used[i]) {
            *result = -1;   // VelocityFilter::DANGLING_VERTICES
            return CCD_OK;
        }

    double maxmass = 0;   // VelocityFilter.cpp:39-44
    for (int64_t i = 0; i < 3 * nverts; i++) maxmass = ccd::smax(maxmass, inv_masses[i]);
    const double dt = base_substep;
    const double k = base_stiffness / dt * std::sqrt(maxmass);
    if (verbose)
        printf("For outer layer thickness %g and max inverse mass %g, using base timestep %g and stiffness %g\n",
               outer_eta, maxmass, dt, k);

    // Device copies of the mesh: faces and the fixed-vertex mask (a vertex with any zero inverse mass).
    std::vector<uint8_t> fixed(nverts, 0);
    bool any_fixed = false;
    for (int64_t i = 0; i < 3 * nverts; i++)
        if (inv_masses[i] == 0.0) fixed[i / 3] = 1, any_fixed = true;
    int32_t *d_faces = (int32_t *)ccdi::buffer(ctx, "vfl_faces", 12 * nfaces + 4);
    uint8_t *d_fixed = (uint8_t *)ccdi::buffer(ctx, "vfl_fixed", nverts + 1);
    double *d_q = (double *)ccdi::buffer(ctx, "vfl_q", 24 * nverts + 8);
    if (!d_faces || !d_fixed || !d_q) return CCD_E_OOM;
    int rc;
    if ((rc = ccd_copy_to_device(ctx, d_faces, faces, 12 * nfaces))) return rc;
    if ((rc = ccd_copy_to_device(ctx, d_fixed, fixed.data(), nverts))) return rc;
    if ((rc = ccd_copy_to_device(ctx, d_q, qstart, 24 * nverts))) return rc;

    // VelocityFilter.cpp:58-66: Distance::meshSelfDistance of the start state.
    Clock::time_point t0 = Clock::now();
    double closest = 0;
    int64_t nvf = 0, nee = 0;
    if ((rc = ccd_mesh_self_distance(ctx, d_q, nverts, d_faces, nfaces, any_fixed ? d_fixed : nullptr, &closest, &nvf,
                                     &nee)))
        return rc;
    st.device_seconds += secs(t0);
    st.closest = closest;
    if (stats) *stats = st;
    if (verbose) {
        printf("Checking %lld vertex-face and %lld edge-edge stencils\n", (long long)nvf, (long long)nee);
        printf("Closest primitive pair is distance %g apart at start\n", closest);
    }
    if (closest < inner_eta) {
        *result = -2;   // VelocityFilter::STARTING_STATE_INTERSECTS
        return CCD_OK;
    }
    if (verbose) {
        double needed = (outer_eta - inner_eta) / (closest - inner_eta);
        int n = (needed >= (double)INT_MIN && needed < (double)INT_MAX + 1.0) ? (int)needed : INT_MIN;
        printf("It will take at least %d layers to resolve this contact problem\n", n);
    }

    Filter al(ctx, d_faces, nfaces, nverts, any_fixed ? d_fixed : nullptr, outer_eta, inner_eta, dt, k, cor,
              verbose != 0, st);
    State s;
    for (int iter = 0; iter < max_rollbacks; iter++) {   // VelocityFilter.cpp:69-86
        s.q.assign(qstart, qstart + 3 * nverts);
        s.v.resize(3 * nverts);
        for (int64_t i = 0; i < 3 * nverts; i++) s.v[i] = qend[i] - qstart[i];
        s.minv.assign(inv_masses, inv_masses + 3 * nverts);
        s.last.assign(3 * nverts, 0.0);
        st.iterations = iter + 1;
        bool clean = false;
        rc = al.run_iteration(s, &clean);
        if (stats) *stats = st;
        if (rc) return rc;
        if (clean) {
            std::copy(s.q.begin(), s.q.end(), qend);
            *result = iter;
            if (verbose) fflush(stdout);
            return CCD_OK;
        }
    }
    *result = -3;   // VelocityFilter::MAXROLLBACKS_EXCEEDED
    if (verbose) fflush(stdout);
    return CCD_OK;
}
