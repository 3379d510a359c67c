// Load generator for the serving loop (ccd_server_*, include/ccd_amd.h): a Poisson stream of CCD
// requests against device-resident sessions, offered from C++ so that the client is not the bottleneck.
//
//   serve_loadgen SESSIONS RATE SECONDS [--inflight N] [--budget-gib G] [--threads T] [--seed S]
//                 [--sigma X] [--backlog-seconds B] [--small-tris N]
//
// SESSIONS is a binary file: int64 count, then per session int64 F, int64 V, double eta, int32 faces[3F],
// double positions[3V] (start = end = these).  Each request picks a session uniformly at random and
// supplies new end positions = the session's previous end + N(0, sigma^2) per coordinate, a fresh draw
// per request, so its tree is refit in place.  Requests are generated ahead of their arrival times by T
// worker threads (each owns the sessions s with s % T == its index, so a session's requests form one
// chain) and submitted by the main thread at their scheduled times.  Latency is measured from the
// scheduled arrival to the result on the host.  The offer stops early when more than RATE * B requests
// are waiting (overload).  Prints one JSON line.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ccd_amd.h"

namespace {

double now_s() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

struct Session {
    int64_t F = 0, V = 0;
    double eta = 0;
    std::vector<int32_t> faces;
    std::vector<double> q;
};

// xoshiro256** with Marsaglia's polar method for normals.
struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed)
    {
        for (int i = 0; i < 4; i++) {   // splitmix64 seeding
            seed += 0x9E3779B97F4A7C15ull;
            uint64_t z = seed;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            s[i] = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next()
    {
        const uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    // two independent N(0, 1) draws
    void normal2(double &a, double &b)
    {
        double u, v, w;
        do {
            u = 2.0 * uniform() - 1.0;
            v = 2.0 * uniform() - 1.0;
            w = u * u + v * v;
        } while (w >= 1.0 || w == 0.0);
        const double m = std::sqrt(-2.0 * std::log(w) / w);
        a = u * m;
        b = v * m;
    }
};

bool read_sessions(const char *path, std::vector<Session> &out)
{
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    int64_t n = 0;
    bool ok = std::fread(&n, 8, 1, f) == 1 && n >= 0;
    for (int64_t k = 0; ok && k < n; k++) {
        Session s;
        ok = std::fread(&s.F, 8, 1, f) == 1 && std::fread(&s.V, 8, 1, f) == 1 && std::fread(&s.eta, 8, 1, f) == 1 &&
             s.F >= 0 && s.V >= 0;
        if (!ok) break;
        s.faces.resize(3 * s.F);
        s.q.resize(3 * s.V);
        ok = std::fread(s.faces.data(), 4, s.faces.size(), f) == s.faces.size() &&
             std::fread(s.q.data(), 8, s.q.size(), f) == s.q.size();
        out.push_back(std::move(s));
    }
    std::fclose(f);
    return ok;
}

double percentile(std::vector<double> v, double p)
{
    if (v.empty()) return NAN;
    std::sort(v.begin(), v.end());
    const double r = p / 100.0 * (double)(v.size() - 1);
    const size_t lo = (size_t)r, hi = std::min(lo + 1, v.size() - 1);
    return v[lo] + (r - (double)lo) * (v[hi] - v[lo]);
}

}   // namespace

int main(int argc, char **argv)
{
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s SESSIONS RATE SECONDS [--inflight N] [--budget-gib G] [--threads T] [--seed S] "
                             "[--sigma X] [--backlog-seconds B] [--small-tris N]\n", argv[0]);
        return 2;
    }
    const char *path = argv[1];
    const double rate = std::atof(argv[2]), seconds = std::atof(argv[3]);
    int inflight = 16, threads = 8;
    double budget_gib = 32, sigma = 0.002, backlog_s = 2.0;
    uint64_t seed = 4;
    int64_t small_tris = 100000;
    for (int i = 4; i + 1 < argc; i += 2) {
        const std::string k = argv[i];
        if (k == "--inflight") inflight = std::atoi(argv[i + 1]);
        else if (k == "--budget-gib") budget_gib = std::atof(argv[i + 1]);
        else if (k == "--threads") threads = std::max(1, std::atoi(argv[i + 1]));
        else if (k == "--seed") seed = (uint64_t)std::atoll(argv[i + 1]);
        else if (k == "--sigma") sigma = std::atof(argv[i + 1]);
        else if (k == "--backlog-seconds") backlog_s = std::atof(argv[i + 1]);
        else if (k == "--small-tris") small_tris = std::atoll(argv[i + 1]);
        else {
            std::fprintf(stderr, "unknown option %s\n", argv[i]);
            return 2;
        }
    }
    std::vector<Session> sess;
    if (!read_sessions(path, sess) || sess.empty()) {
        std::fprintf(stderr, "cannot read sessions from %s\n", path);
        return 2;
    }
    ccd_server *srv = nullptr;
    int rc = ccd_server_create(0, (int64_t)(budget_gib * (double)(1ll << 30)), inflight, &srv);
    if (rc) {
        std::fprintf(stderr, "ccd_server_create failed: %d\n", rc);
        return 1;
    }
    const double t_open = now_s();
    std::vector<int> ids(sess.size());
    int64_t resident_tris = 0;
    for (size_t k = 0; k < sess.size(); k++) {
        const Session &s = sess[k];
        rc = ccd_server_open_session(srv, s.faces.data(), s.F, s.q.data(), s.q.data(), s.V, s.eta, s.eta, &ids[k]);
        if (rc) {
            std::fprintf(stderr, "ccd_server_open_session(%zu) failed: %d\n", k, rc);
            return 1;
        }
        resident_tris += s.F;
        if (k % 64 == 63) std::fprintf(stderr, "loadgen: opened %zu sessions\n", k + 1);
    }
    std::fprintf(stderr, "loadgen: %zu sessions, %lld resident triangles, opened in %.1f s\n", sess.size(),
                 (long long)resident_tris, now_s() - t_open);

    // The arrival process: Poisson times, uniform sessions.
    Rng arr(seed);
    std::vector<double> at;
    std::vector<int> who;
    for (double t = 0;;) {
        t += -std::log(1.0 - arr.uniform()) / rate;
        if (t >= seconds) break;
        at.push_back(t);
        who.push_back((int)(arr.next() % sess.size()));
    }
    const int64_t N = (int64_t)at.size();
    // Request buffers: a ring of slots filled by the worker threads ahead of the arrivals.
    const int64_t RING = 512;
    std::vector<std::vector<double>> slot(RING);
    std::vector<std::atomic<int64_t>> ready(RING);   // request id whose data the slot holds, or -1
    for (auto &r : ready) r.store(-1);
    std::atomic<int64_t> submitted_upto{0};            // requests before this index have been submitted
    std::atomic<bool> stop{false};
    std::vector<std::thread> pool;
    for (int w = 0; w < threads; w++) {
        pool.emplace_back([&, w]() {
            Rng rng(seed * 1000003ull + 17ull * (uint64_t)w + 1);
            std::vector<std::vector<double>> cur(sess.size());
            for (int64_t k = 0; k < N && !stop.load(); k++) {
                if (who[k] % threads != w) continue;
                while (k - submitted_upto.load() >= RING && !stop.load()) std::this_thread::yield();
                if (stop.load()) break;
                const Session &s = sess[who[k]];
                std::vector<double> &c = cur[who[k]];
                if (c.empty()) c = s.q;
                for (size_t i = 0; i + 1 < c.size(); i += 2) {
                    double a, b;
                    rng.normal2(a, b);
                    c[i] += sigma * a;
                    c[i + 1] += sigma * b;
                }
                if (c.size() % 2) {
                    double a, b;
                    rng.normal2(a, b);
                    c.back() += sigma * a;
                }
                slot[k % RING] = c;
                ready[k % RING].store(k);
            }
        });
    }
    std::vector<ccd_request_result> res;
    std::vector<ccd_request_result> buf(1024);
    auto poll = [&]() {
        int n = 0;
        do {
            ccd_server_poll(srv, buf.data(), (int)buf.size(), &n);
            res.insert(res.end(), buf.begin(), buf.begin() + n);
        } while (n == (int)buf.size());
    };
    int64_t stats[10];
    ccd_server_stats(srv, stats);
    const int64_t reruns0 = stats[6], waits0 = stats[8];
    const double start = now_s();
    std::vector<double> sched(N, 0.0);
    int64_t k = 0;
    bool overloaded = false;
    double lag_max = 0;
    for (; k < N; k++) {
        const double t_arr = start + at[k];
        while (now_s() < t_arr) {
            if (t_arr - now_s() > 0.0005) poll();
        }
        while (ready[k % RING].load() != k) std::this_thread::yield();
        const double lag = now_s() - t_arr;
        lag_max = std::max(lag_max, lag);
        sched[k] = t_arr;
        rc = ccd_server_submit(srv, ids[who[k]], slot[k % RING].data(), k);
        submitted_upto.store(k + 1);
        if (rc) {
            std::fprintf(stderr, "ccd_server_submit failed: %d\n", rc);
            break;
        }
        if (k % 64 == 0) {
            poll();
            ccd_server_stats(srv, stats);
            if ((double)stats[2] > rate * backlog_s) {
                overloaded = true;
                std::fprintf(stderr, "loadgen: overloaded at request %lld (%.2f s): %lld waiting\n", (long long)k, at[k],
                             (long long)stats[2]);
                k++;
                break;
            }
        }
        if (k % 2000 == 0)
            std::fprintf(stderr, "loadgen: %lld requests submitted at %.1f s, %zu done\n", (long long)k, at[k], res.size());
    }
    const int64_t offered = k;
    const double offered_wall = now_s() - start;
    stop.store(true);
    for (auto &t : pool) t.join();
    ccd_server_drain(srv);
    poll();
    const double end = now_s();
    ccd_server_stats(srv, stats);
    std::vector<double> lat, lat_small;
    double t_first = 1e300, t_last = -1e300, tris = 0;
    int64_t failed = 0, rerun_requests = 0, shown = 0;
    for (const auto &r : res) {
        if (r.status != CCD_OK) {
            failed++;
            continue;
        }
        if (r.reruns > 0) {
            rerun_requests++;
            if (shown++ < 8)
                std::fprintf(stderr, "loadgen: request %lld (session %d, %lld triangles) ran %d extra time(s): %lld face "
                             "pairs, %lld + %lld stencils\n", (long long)r.request_id, r.session,
                             (long long)sess[r.session].F, r.reruns, (long long)r.face_pairs,
                             (long long)r.vf_candidates, (long long)r.ee_candidates);
        }
        const double l = (r.t_done - sched[r.request_id]) * 1e3;
        lat.push_back(l);
        const int64_t F = sess[r.session].F;
        if (F <= small_tris) lat_small.push_back(l);
        t_first = std::min(t_first, sched[r.request_id]);
        t_last = std::max(t_last, r.t_done);
        tris += (double)F;
    }
    const double span = t_last > t_first ? t_last - t_first : 1.0;
    std::printf("{\"rate_offered\": %.3f, \"seconds\": %.3f, \"requests_scheduled\": %lld, \"submitted\": %lld, "
                "\"completed\": %zu, \"failed\": %lld, \"overloaded\": %s, \"offered_wall_s\": %.3f, "
                "\"drain_wall_s\": %.3f, \"offered_rate_achieved\": %.3f, \"requests_per_s\": %.3f, "
                "\"swept_tris_per_s\": %.1f, \"latency_ms\": {\"p50\": %.4f, \"p99\": %.4f, \"p50_le%lldk\": %.4f, "
                "\"p99_le%lldk\": %.4f, \"max\": %.4f}, \"reruns\": %lld, \"rerun_requests\": %lld, \"mem_waits\": %lld, \"client_max_lag_ms\": %.3f, "
                "\"sessions\": %zu, \"resident_triangles\": %lld, \"inflight\": %d, \"budget_gib\": %.1f, "
                "\"client\": \"C++ (%d generator threads), fresh N(0, %g^2) draws per request\"}\n",
                rate, seconds, (long long)N, (long long)offered, lat.size(), (long long)failed,
                overloaded ? "true" : "false", offered_wall, end - start - offered_wall,
                (double)offered / std::max(offered_wall, 1e-9), (double)lat.size() / span, tris / span,
                percentile(lat, 50), percentile(lat, 99), (long long)(small_tris / 1000), percentile(lat_small, 50),
                (long long)(small_tris / 1000), percentile(lat_small, 99),
                lat.empty() ? NAN : *std::max_element(lat.begin(), lat.end()), (long long)(stats[6] - reruns0),
                (long long)rerun_requests, (long long)(stats[8] - waits0), lag_max * 1e3, sess.size(), (long long)resident_tris, inflight,
                budget_gib, threads, sigma);
    ccd_server_destroy(srv);
    return 0;
}
