// Test-only: History::getPosAtTime and History::stitchCommonHistory of include/ccdamd/History.h on a
// CSR history read from a file, for the CPU tests to compare with the oracle.
//   in:  int64 nverts, nentries; int64 off[nverts + 1]; double t[nentries]; double pos[3 nentries];
//        int64 nq; nq x (int32 vert, double time)                       -- getPosAtTime queries
//        int64 ns; ns x (int32 k, int32 verts[4])                         -- stitchCommonHistory queries
//   out: text, one line per answer: "pos idx x y z" (%a hex floats), then per stitch query
//        "stitch n" followed by n lines "t x0 y0 z0 ... " (hex floats)
#include <cstdint>
#include <cstdio>
#include <vector>

#include "History.h"

template <class T>
static T rd(FILE *f)
{
    T v;
    if (std::fread(&v, sizeof(T), 1, f) != 1) throw 1;
    return v;
}

int main(int argc, char **argv)
{
    if (argc != 3) return 2;
    FILE *f = std::fopen(argv[1], "rb");
    FILE *o = std::fopen(argv[2], "w");
    if (!f || !o) return 2;
    const int64_t V = rd<int64_t>(f), E = rd<int64_t>(f);
    std::vector<int64_t> off(V + 1);
    for (auto &x : off) x = rd<int64_t>(f);
    std::vector<double> t(E), p(3 * E);
    for (auto &x : t) x = rd<double>(f);
    for (auto &x : p) x = rd<double>(f);
    Eigen::VectorXd q0(3 * V);
    for (int64_t v = 0; v < V; v++)
        for (int k = 0; k < 3; k++) q0[3 * v + k] = p[3 * off[v] + k];
    History h(q0);
    for (int64_t v = 0; v < V; v++)
        for (int64_t e = off[v] + 1; e < off[v + 1]; e++) h.addHistory((int)v, t[e], Eigen::Vector3d(p[3 * e], p[3 * e + 1], p[3 * e + 2]));
    const int64_t nq = rd<int64_t>(f);
    for (int64_t i = 0; i < nq; i++) {
        const int v = rd<int32_t>(f);
        const double tm = rd<double>(f);
        Eigen::Vector3d pos;
        int idx = -1;
        h.getPosAtTime(v, tm, pos, idx);
        std::fprintf(o, "pos %d %a %a %a\n", idx, pos[0], pos[1], pos[2]);
    }
    const int64_t ns = rd<int64_t>(f);
    for (int64_t i = 0; i < ns; i++) {
        const int k = rd<int32_t>(f);
        std::vector<int> verts(4);
        for (auto &x : verts) x = rd<int32_t>(f);
        verts.resize(k);
        std::vector<StitchedEntry> st;
        h.stitchCommonHistory(verts, st);
        std::fprintf(o, "stitch %zu\n", st.size());
        for (const auto &e : st) {
            std::fprintf(o, "%a", e.time);
            for (const auto &q : e.pos) std::fprintf(o, " %a %a %a", q[0], q[1], q[2]);
            std::fprintf(o, "\n");
        }
    }
    std::fclose(o);
    return 0;
}
