// Test-only: the product core (through core_harness.cpp) as a standalone program, for builds with
// AddressSanitizer + UndefinedBehaviorSanitizer (g++) or MemorySanitizer (clang), with every runtime
// index into the core's per-lane arrays bounds-checked (CCD_CHECK_INDEX).  It reads one case written by
// tests/test_core_sanitized.py:
//   int64 nverts, nentries, nkeys; int32 is_ee; double eta;
//   int64 off[nverts + 1]; double ht[nentries]; double hp[3 * nentries]; int32 keys[4 * nkeys]
// and runs the separating-plane narrowphase with pruned and with full cascades, then the CTCD
// narrowphase on the first and last history entries, and writes the statuses, flags, sub-tests and
// times of impact as raw arrays to the output file.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "core_harness.cpp"

template <class T>
static void rd(FILE *f, T *p, size_t n)
{
    if (n && std::fread(p, sizeof(T), n, f) != n) {
        std::fprintf(stderr, "short read\n");
        std::exit(2);
    }
}

int main(int argc, char **argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s case.bin out.bin\n", argv[0]);
        return 2;
    }
    FILE *f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    int64_t hdr[3];
    int32_t is_ee = 0;
    double eta = 0;
    rd(f, hdr, 3);
    rd(f, &is_ee, 1);
    rd(f, &eta, 1);
    const int64_t V = hdr[0], E = hdr[1], n = hdr[2];
    std::vector<int64_t> off(V + 1);
    std::vector<double> ht(E), hp(3 * E);
    std::vector<int32_t> keys(4 * n);
    rd(f, off.data(), off.size());
    rd(f, ht.data(), ht.size());
    rd(f, hp.data(), hp.size());
    rd(f, keys.data(), keys.size());
    std::fclose(f);

    std::vector<double> etas(n, eta);
    std::vector<int32_t> st_pruned(n), st_full(n), hit(n), sub(n);
    std::vector<double> t(n);
    core_set_sp_pruning(1);
    core_separating_plane(off.data(), ht.data(), hp.data(), V, keys.data(), etas.data(), n, is_ee, st_pruned.data());
    core_set_sp_pruning(0);
    core_separating_plane(off.data(), ht.data(), hp.data(), V, keys.data(), etas.data(), n, is_ee, st_full.data());
    // two-entry windows from each vertex's first and last entries
    std::vector<double> qs(3 * V), qe(3 * V);
    for (int64_t v = 0; v < V; v++)
        for (int c = 0; c < 3; c++) {
            qs[3 * v + c] = hp[3 * off[v] + c];
            qe[3 * v + c] = hp[3 * (off[v + 1] - 1) + c];
        }
    core_narrowphase(qs.data(), qe.data(), keys.data(), n, is_ee, eta, hit.data(), t.data(), sub.data());

    FILE *o = std::fopen(argv[2], "wb");
    if (!o) return 2;
    std::fwrite(st_pruned.data(), 4, n, o);
    std::fwrite(st_full.data(), 4, n, o);
    std::fwrite(hit.data(), 4, n, o);
    std::fwrite(sub.data(), 4, n, o);
    std::fwrite(t.data(), 8, n, o);
    std::fclose(o);
    int64_t diff = 0;
    for (int64_t i = 0; i < n; i++) diff += st_pruned[i] != st_full[i];
    std::printf("stencils %lld  pruned != full: %lld\n", (long long)n, (long long)diff);
    return 0;
}
