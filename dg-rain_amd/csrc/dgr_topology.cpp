// dgr_topology.cpp -- mesh topology of the reference's fort.14 variant,
// built in O(N) with hash lookups instead of the reference's O(edges x
// elements) scans (2DCode/extract_mesh_attributes.c:143-361).  Output is
// identical: the same edge numbering, EdgtoEls pairs and BdryPrescribed codes.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <unordered_map>
#include <vector>

#include "../../include/dgrain.h"

namespace {

inline uint64_t edge_key(int a, int b)
{
    const uint32_t lo = (uint32_t)(a < b ? a : b), hi = (uint32_t)(a < b ? b : a);
    return ((uint64_t)lo << 32) | hi;
}

// position of cell (cx, cy) along the Hilbert curve through a 2^16 x 2^16
// grid (the textbook quadrant recursion: at each level take the quadrant's
// index, then rotate/reflect the cell into that quadrant's frame)
inline uint64_t hilbert_index(uint32_t cx, uint32_t cy)
{
    const uint32_t n = 1u << 16;
    uint64_t d = 0;
    for (uint32_t s = n >> 1; s > 0; s >>= 1) {
        const uint32_t rx = (cx & s) ? 1u : 0u, ry = (cy & s) ? 1u : 0u;
        d += (uint64_t)s * s * ((3u * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) {
                cx = n - 1 - cx;
                cy = n - 1 - cy;
            }
            std::swap(cx, cy);
        }
    }
    return d;
}

}  // namespace

extern "C" dgr_status dgr_locality_order(int num_el, int num_nodes, const int *ev, const double *x, const double *y,
                                         int *el_order, int *node_order)
{
    if (num_el < 0 || num_nodes < 0 || (num_el && (!ev || !el_order)) || (num_nodes && (!x || !y || !node_order)))
        return DGR_ERR_ARG;
    for (long long i = 0; i < 3LL * num_el; ++i)
        if (ev[i] < 0 || ev[i] >= num_nodes) return DGR_ERR_MESH;
    double x0 = INFINITY, x1 = -INFINITY, y0 = INFINITY, y1 = -INFINITY;
    for (int i = 0; i < num_nodes; ++i) {
        if (!std::isfinite(x[i]) || !std::isfinite(y[i])) return DGR_ERR_MESH;
        x0 = std::min(x0, x[i]);
        x1 = std::max(x1, x[i]);
        y0 = std::min(y0, y[i]);
        y1 = std::max(y1, y[i]);
    }
    // one scale for both axes keeps the curve's cells square
    const double span = std::max(x1 - x0, y1 - y0);
    const double scale = span > 0 ? 65535.0 / span : 0.0;
    std::vector<uint64_t> key((size_t)num_el);
    for (int e = 0; e < num_el; ++e) {
        const int *v = ev + 3 * (size_t)e;
        const double cx = (x[v[0]] + x[v[1]] + x[v[2]]) / 3.0, cy = (y[v[0]] + y[v[1]] + y[v[2]]) / 3.0;
        const double gx = std::min(65535.0, std::max(0.0, (cx - x0) * scale));
        const double gy = std::min(65535.0, std::max(0.0, (cy - y0) * scale));
        key[(size_t)e] = hilbert_index((uint32_t)gx, (uint32_t)gy);
    }
    std::iota(el_order, el_order + num_el, 0);
    std::stable_sort(el_order, el_order + num_el, [&](int a, int b) { return key[(size_t)a] < key[(size_t)b]; });
    // nodes in order of first use; nodes no element uses keep their relative
    // order at the end
    std::vector<char> seen((size_t)num_nodes, 0);
    int k = 0;
    for (int i = 0; i < num_el; ++i)
        for (int j = 0; j < 3; ++j) {
            const int v = ev[3 * (size_t)el_order[i] + j];
            if (!seen[(size_t)v]) {
                seen[(size_t)v] = 1;
                node_order[k++] = v;
            }
        }
    for (int v = 0; v < num_nodes; ++v)
        if (!seen[(size_t)v]) node_order[k++] = v;
    return DGR_OK;
}

extern "C" dgr_status dgr_build_topology(int num_el, int num_nodes, const int *ev, const double *x,
                                         const double *y, int n_in1, const int *in1, int n_in2, const int *in2,
                                         int n_out, const int *out, int *e2v, int *e2e, int *bdry,
                                         int *num_edges_out, double *min_edge_length)
{
    if (num_el <= 0 || num_nodes <= 0 || !ev || !x || !y || !e2v || !e2e || !bdry) return DGR_ERR_ARG;
    for (int i = 0; i < 3 * num_el; ++i)
        if (ev[i] < 0 || ev[i] >= num_nodes) return DGR_ERR_MESH;

    // Neighbour nodes of each node in first-seen order: for element el (ascending)
    // and local vertex n, the vertices n+1 and n+2 (extract_mesh_attributes.c:166-214).
    std::vector<int> off(num_nodes + 1, 0);
    for (int i = 0; i < 3 * num_el; ++i) off[ev[i] + 1] += 2;
    for (int i = 0; i < num_nodes; ++i) off[i + 1] += off[i];
    std::vector<int> nbr(off[num_nodes]);
    std::vector<int> cnt(num_nodes, 0);
    for (int el = 0; el < num_el; ++el)
        for (int n = 0; n < 3; ++n) {
            const int node = ev[3 * el + n];
            const int c[2] = {ev[3 * el + (n + 1) % 3], ev[3 * el + (n + 2) % 3]};
            for (int t = 0; t < 2; ++t) {
                bool found = false;
                for (int j = off[node]; j < off[node] + cnt[node]; ++j)
                    if (nbr[j] == c[t]) { found = true; break; }
                if (!found) nbr[off[node] + cnt[node]++] = c[t];
            }
        }

    // Edges (n, m) with m > n, nodes ascending, neighbours in stored order (:225-250).
    const long long ned = (long long)num_nodes + num_el - 1;   // Euler, simply connected (:219)
    long long ec = 0;
    for (int n = 0; n < num_nodes; ++n)
        for (int j = off[n]; j < off[n] + cnt[n]; ++j)
            if (nbr[j] > n) {
                if (ec >= ned) return DGR_ERR_MESH;
                e2v[2 * ec] = n;
                e2v[2 * ec + 1] = nbr[j];
                ++ec;
            }
    if (ec != ned) return DGR_ERR_MESH;   // :252-257

    std::unordered_map<uint64_t, int> idx;
    idx.reserve((size_t)ned * 2);
    for (int e = 0; e < (int)ned; ++e) {
        idx.emplace(edge_key(e2v[2 * e], e2v[2 * e + 1]), e);
        e2e[2 * e] = -1;
        e2e[2 * e + 1] = -1;
    }
    // The first two elements (ascending) that contain each edge; a boundary edge
    // repeats its single element (:259-295).
    for (int el = 0; el < num_el; ++el)
        for (int k = 0; k < 3; ++k) {
            auto it = idx.find(edge_key(ev[3 * el + k], ev[3 * el + (k + 1) % 3]));
            if (it == idx.end()) continue;
            const int e = it->second;
            if (e2e[2 * e] < 0) e2e[2 * e] = el;
            else if (e2e[2 * e + 1] < 0 && e2e[2 * e] != el) e2e[2 * e + 1] = el;
        }
    double mel = 9999999;   // extract_mesh_attributes.c:25
    for (int e = 0; e < (int)ned; ++e) {
        if (e2e[2 * e] >= 0 && e2e[2 * e + 1] < 0) e2e[2 * e + 1] = e2e[2 * e];
        const int n0 = e2v[2 * e], n1 = e2v[2 * e + 1];
        const double len = std::sqrt((x[n1] - x[n0]) * (x[n1] - x[n0]) + (y[n1] - y[n0]) * (y[n1] - y[n0]));
        mel = (mel < len) ? mel : len;
        bdry[e] = 0;
    }
    // BdryPrescribed: second inflow list (3) over first inflow list (1) over outflow (2) (:305-352)
    const int *lists[3] = {out, in1, in2};
    const int nlist[3] = {n_out, n_in1, n_in2};
    const int code[3] = {2, 1, 3};
    for (int l = 0; l < 3; ++l)
        for (int k = 0; k < nlist[l]; ++k) {
            auto it = idx.find(edge_key(lists[l][2 * k], lists[l][2 * k + 1]));
            if (it != idx.end()) bdry[it->second] = code[l];
        }
    if (num_edges_out) *num_edges_out = (int)ned;
    if (min_edge_length) *min_edge_length = mel;
    return DGR_OK;
}
