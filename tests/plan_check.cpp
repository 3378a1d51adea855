// CPU check of the assembly plan (minisod2d_amd/csrc/msod_plan.h), built by
// tests/test_plan.py with AddressSanitizer and UBSan.
//
// Input file (little endian): int64 nelem, npoin; int32 nn, B; int32 l2g[nn];
// int32 connec[nelem*nn] (reference layout connec(nelem,nnode), 1-based).
// Prints "OK ncolors=<c> nbatch=<b>" or the first violated invariant.
#include <cstdio>
#include <cstdlib>
#include <set>

#include "msod_plan.h"

static int die(const char *what, long long a = -1, long long b = -1)
{
    std::printf("FAIL %s (%lld, %lld)\n", what, a, b);
    return 1;
}

int main(int argc, char **argv)
{
    if (argc != 2) return die("usage");
    FILE *f = std::fopen(argv[1], "rb");
    if (!f) return die("open");
    int64_t nelem, npoin;
    int32_t nn, B;
    if (std::fread(&nelem, 8, 1, f) != 1 || std::fread(&npoin, 8, 1, f) != 1 ||
        std::fread(&nn, 4, 1, f) != 1 || std::fread(&B, 4, 1, f) != 1)
        return die("header");
    std::vector<int32_t> l2g(nn), raw(nelem * nn);
    if (std::fread(l2g.data(), 4, nn, f) != (size_t)nn ||
        std::fread(raw.data(), 4, raw.size(), f) != raw.size())
        return die("body");
    std::fclose(f);

    msod::HostPlan P;
    std::string err;
    if (!msod::build_host_plan(raw.data(), nelem, npoin, l2g, B, P, &err)) {
        std::printf("ERROR %s\n", err.c_str());
        return 2;
    }
    const int64_t NB = P.nbatch;
    if (NB != (nelem + B - 1) / B) return die("nbatch");
    if ((int64_t)P.uoff.size() != NB + 1 || P.unode.size() != P.uoff[NB]) return die("sizes");
    // connectivity and node positions
    for (int64_t e = 0; e < nelem; ++e)
        for (int l = 0; l < nn; ++l) {
            const int64_t k = e * nn + l;
            if (P.conn[k] != (uint32_t)(raw[e + nelem * l2g[l]] - 1)) return die("conn", e, l);
            const int64_t b = e / B;
            const uint32_t U = P.uoff[b + 1] - P.uoff[b];
            if (P.lidx[k] >= U) return die("lidx range", e, l);
            if ((P.unode[P.uoff[b] + P.lidx[k]] & 0x7fffffffu) != P.conn[k]) return die("lidx", e, l);
        }
    // each batch's list: sorted, distinct, exactly its elements' nodes
    for (int64_t b = 0; b < NB; ++b) {
        std::set<uint32_t> want;
        for (int64_t e = b * B; e < std::min<int64_t>(nelem, (b + 1) * B); ++e)
            for (int l = 0; l < nn; ++l) want.insert(P.conn[e * nn + l]);
        if (want.size() != P.uoff[b + 1] - P.uoff[b]) return die("batch count", b);
        uint32_t prev = 0;
        int i = 0;
        for (uint32_t v : want) {
            const uint32_t got = P.unode[P.uoff[b] + i] & 0x7fffffffu;
            if (got != v || (i > 0 && got <= prev)) return die("batch list", b, i);
            prev = got;
            ++i;
        }
    }
    if (!P.colored) {
        // atomic assembly only: no first-writer flags, order is the mesh
        // order, unref exactly the nodes no batch lists
        std::vector<char> refd(npoin, 0), unref(npoin, 0);
        for (uint32_t v : P.unode) {
            if (v >> 31) return die("first-writer flag without colouring", v);
            refd[v] = 1;
        }
        for (uint32_t v : P.unref) {
            if (v >= npoin || unref[v]) return die("unref entry", v);
            unref[v] = 1;
        }
        for (int64_t v = 0; v < npoin; ++v)
            if ((bool)refd[v] == (bool)unref[v]) return die("unref set", v);
        if ((int64_t)P.order.size() != NB) return die("order size");
        for (int64_t b = 0; b < NB; ++b)
            if (P.order[b] != (uint32_t)b) return die("order", b);
        std::printf("OK ncolors=0 atomic-only nbatch=%lld\n", (long long)NB);
        return 0;
    }
    // colours: order is a permutation grouped by colour; one colour's
    // batches share no node; exactly one first writer per referenced node, of
    // the lowest colour holding it
    const int nc = (int)P.color_off.size() - 1;
    if (nc < 1 || nc > 32 || P.color_off[0] != 0 || P.color_off[nc] != NB) return die("color_off");
    std::vector<int> colour(NB, -1);
    for (int c = 0; c < nc; ++c)
        for (int64_t i = P.color_off[c]; i < P.color_off[c + 1]; ++i) {
            const uint32_t b = P.order[i];
            if (b >= NB || colour[b] != -1) return die("order", i, b);
            colour[b] = c;
        }
    std::vector<int> seen_in(npoin, -1), first_count(npoin, 0), min_colour(npoin, 99),
        first_colour(npoin, -1);
    for (int c = 0; c < nc; ++c)
        for (int64_t i = P.color_off[c]; i < P.color_off[c + 1]; ++i) {
            const uint32_t b = P.order[i];
            for (uint32_t k = P.uoff[b]; k < P.uoff[b + 1]; ++k) {
                const uint32_t v = P.unode[k] & 0x7fffffffu;
                if (seen_in[v] == c) return die("colour conflict", v, c);
                seen_in[v] = c;
                min_colour[v] = std::min(min_colour[v], c);
                if (P.unode[k] >> 31) {
                    first_count[v]++;
                    first_colour[v] = c;
                }
            }
        }
    std::vector<char> unref(npoin, 0);
    for (uint32_t v : P.unref) {
        if (v >= npoin || unref[v]) return die("unref entry", v);
        unref[v] = 1;
    }
    for (int64_t v = 0; v < npoin; ++v) {
        const bool referenced = seen_in[v] != -1;
        if (referenced == (bool)unref[v]) return die("unref set", v);
        if (referenced && (first_count[v] != 1 || first_colour[v] != min_colour[v]))
            return die("first writer", v, first_count[v]);
    }
    std::printf("OK ncolors=%d nbatch=%lld\n", nc, (long long)NB);
    return 0;
}
