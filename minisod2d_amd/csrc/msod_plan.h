// Assembly plan of a mesh (host code, no HIP): the data the kernels' write
// phase uses (layout described in msod_kernels.h).  Built once per mesh from
// the reference connectivity connec(nelem,nnode) (1-based, Fortran layout):
//   * batches of B consecutive elements;
//   * the sorted distinct nodes of every batch (unode, offsets uoff) and the
//     position of every element node in its batch's list (lidx);
//   * a greedy colouring of the batches such that no two batches of one
//     colour share a node (at most 32 colours), used only by the coloured
//     assembly; a mesh whose batches need more colours (elements in no
//     spatial order) gets no colouring and runs with the atomic assembly
//     alone (colored = false);
//   * bit 31 of unode marks the first writer of a node: the batch of the
//     lowest colour holding it;
//   * the batches grouped by colour (order, color_off) and the nodes no
//     element references (unref).
// Header-only so that a CPU check can compile it on its own.
#pragma once

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <string>
#include <thread>
#include <vector>

namespace msod {

struct HostPlan {
    int B = 1;
    int64_t nbatch = 0;
    std::vector<uint32_t> conn;       // [nelem][nn] 0-based, l lexicographic
    std::vector<uint16_t> lidx;       // [nelem][nn]
    std::vector<uint32_t> uoff;       // [nbatch+1]
    std::vector<uint32_t> unode;      // [uoff[nbatch]]
    std::vector<uint32_t> order;      // [nbatch]
    std::vector<int64_t> color_off;   // [ncolors+1]
    std::vector<uint32_t> unref;
    bool colored = false;             // colouring built (order, color_off, bit 31 valid)
};

template <class F>
void plan_parallel_for(int64_t n, F f)
{
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const unsigned nt = std::min<unsigned>(16, hw);
    if (n < 4096 || nt == 1) {
        f(int64_t(0), n);
        return;
    }
    std::vector<std::thread> th;
    const int64_t chunk = (n + nt - 1) / nt;
    for (unsigned i = 0; i < nt; ++i) {
        const int64_t lo = i * chunk, hi = std::min<int64_t>(n, lo + chunk);
        if (lo < hi) th.emplace_back(f, lo, hi);
    }
    for (auto &t : th) t.join();
}

// raw: connec(nelem,nnode) as stored by the reference; l2g[l]: 0-based Gmsh
// node of lexicographic node l.  Returns false with a message in *err.
inline bool build_host_plan(const int32_t *raw, int64_t nelem, int64_t npoin,
                            const std::vector<int32_t> &l2g, int B, HostPlan &P, std::string *err)
{
    const int nn = (int)l2g.size();
    const int64_t E = nelem * nn;
    if (npoin > 0x7fffffff || E > 0xffffffffLL) {
        *err = "mesh too large for 32-bit node indices";
        return false;
    }
    if (B < 1 || (int64_t)B * nn > 65536) {
        *err = "batch too large for 16-bit node positions";
        return false;
    }
    P.B = B;
    P.nbatch = (nelem + B - 1) / B;
    P.conn.assign(E, 0u);
    std::atomic<int64_t> bad{0};
    plan_parallel_for(nelem, [&](int64_t lo, int64_t hi) {
        int64_t nb = 0;
        for (int64_t e = lo; e < hi; ++e)
            for (int l = 0; l < nn; ++l) {
                const int32_t v = raw[e + nelem * l2g[l]];
                if (v < 1 || v > npoin)
                    ++nb;
                else
                    P.conn[e * nn + l] = (uint32_t)(v - 1);
            }
        bad += nb;
    });
    if (bad.load()) {
        *err = "connec holds " + std::to_string(bad.load()) + " node numbers outside 1..npoin (npoin=" +
               std::to_string(npoin) + ")";
        return false;
    }
    const int64_t NB = P.nbatch;
    const std::vector<uint32_t> &conn = P.conn;

    // distinct nodes per batch: counts, then the lists (each formed in a
    // private buffer, so a batch writes only its own slot of unode)
    std::vector<uint32_t> ucount(NB);
    plan_parallel_for(NB, [&](int64_t lo, int64_t hi) {
        std::vector<uint32_t> tmp;
        for (int64_t b = lo; b < hi; ++b) {
            const int64_t e0 = b * B, e1 = std::min<int64_t>(nelem, e0 + B);
            tmp.assign(conn.begin() + e0 * nn, conn.begin() + e1 * nn);
            std::sort(tmp.begin(), tmp.end());
            ucount[b] = (uint32_t)(std::unique(tmp.begin(), tmp.end()) - tmp.begin());
        }
    });
    P.uoff.assign(NB + 1, 0u);
    for (int64_t b = 0; b < NB; ++b) P.uoff[b + 1] = P.uoff[b] + ucount[b];
    P.unode.assign(P.uoff[NB], 0u);
    P.lidx.assign(E, 0);
    plan_parallel_for(NB, [&](int64_t lo, int64_t hi) {
        std::vector<uint32_t> tmp;
        for (int64_t b = lo; b < hi; ++b) {
            const int64_t e0 = b * B, e1 = std::min<int64_t>(nelem, e0 + B);
            tmp.assign(conn.begin() + e0 * nn, conn.begin() + e1 * nn);
            std::sort(tmp.begin(), tmp.end());
            tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
            const uint32_t U = P.uoff[b + 1] - P.uoff[b];
            if (tmp.size() != U) {
                bad += 1;
                continue;
            }
            uint32_t *u = P.unode.data() + P.uoff[b];
            std::copy(tmp.begin(), tmp.end(), u);
            for (int64_t k = e0 * nn; k < e1 * nn; ++k)
                P.lidx[k] = (uint16_t)(std::lower_bound(u, u + U, conn[k]) - u);
        }
    });
    if (bad.load()) {
        *err = "inconsistent batch node lists";
        return false;
    }

    // greedy colouring: bit c of mask[node] is set once a batch of colour c
    // holds the node.  Past 32 colours the colouring is abandoned: the
    // atomic assembly does not need it.
    std::vector<uint32_t> mask(npoin, 0u);
    std::vector<uint8_t> color(NB);
    int ncolors = 0;
    bool overflow = false;
    for (int64_t b = 0; b < NB && !overflow; ++b) {
        uint32_t forb = 0;
        for (uint32_t k = P.uoff[b]; k < P.uoff[b + 1]; ++k) forb |= mask[P.unode[k]];
        if (forb == 0xffffffffu) {
            overflow = true;
            break;
        }
        const int c = __builtin_ctz(~forb);
        color[b] = (uint8_t)c;
        ncolors = std::max(ncolors, c + 1);
        for (uint32_t k = P.uoff[b]; k < P.uoff[b + 1]; ++k) mask[P.unode[k]] |= 1u << c;
    }
    if (overflow) {
        // atomic assembly only: no colours, no first-writer flags; the
        // unreferenced nodes from a plain reference count
        std::vector<uint8_t> ref(npoin, 0);
        for (uint32_t v : P.unode) ref[v] = 1;
        P.unref.clear();
        for (int64_t i = 0; i < npoin; ++i)
            if (!ref[i]) P.unref.push_back((uint32_t)i);
        P.color_off.assign(2, 0);
        P.color_off[1] = NB;
        P.order.resize(NB);
        for (int64_t b = 0; b < NB; ++b) P.order[b] = (uint32_t)b;
        P.colored = false;
        return true;
    }
    plan_parallel_for(NB, [&](int64_t lo, int64_t hi) {
        for (int64_t b = lo; b < hi; ++b) {
            const uint32_t lower = (1u << color[b]) - 1u;
            for (uint32_t k = P.uoff[b]; k < P.uoff[b + 1]; ++k)
                if ((mask[P.unode[k]] & lower) == 0u) P.unode[k] |= 0x80000000u;
        }
    });
    P.unref.clear();
    for (int64_t i = 0; i < npoin; ++i)
        if (mask[i] == 0u) P.unref.push_back((uint32_t)i);
    P.color_off.assign(ncolors + 1, 0);
    for (int64_t b = 0; b < NB; ++b) P.color_off[color[b] + 1]++;
    for (int c = 0; c < ncolors; ++c) P.color_off[c + 1] += P.color_off[c];
    P.order.assign(NB, 0u);
    std::vector<int64_t> pos(P.color_off.begin(), P.color_off.end() - 1);
    for (int64_t b = 0; b < NB; ++b) P.order[pos[color[b]]++] = (uint32_t)b;
    P.colored = true;
    return true;
}

}  // namespace msod
