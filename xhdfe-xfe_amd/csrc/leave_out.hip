// Articulation points of an undirected graph in CSR form on the device: the
// cut-vertex search of the leave-out connected set (articulation_points,
// src/akm_kss.cpp:444-493, used by build_leave_out_sample's pruning loop
// :736-870).  The reference runs Tarjan's depth-first search on the host; a
// depth-first order does not parallelize, so this follows Tarjan and Vishkin's
// biconnectivity method on a breadth-first spanning forest (R. E. Tarjan and
// U. Vishkin, "An efficient parallel biconnectivity algorithm", SIAM J.
// Comput. 14(4), 1985):
//   1. component labels by min-label propagation with pointer jumping; the
//      smallest vertex of each component is its root;
//   2. multi-source BFS depths; parent(v) = the smallest neighbour one level up;
//   3. subtree sizes bottom-up, preorder numbers top-down (children in
//      adjacency order): subtree(v) = [pre(v), pre(v) + size(v));
//   4. low(v) / high(v) = min / max of pre over subtree(v) and the neighbours
//      of its vertices, the edge to parent(v) excepted (bottom-up);
//   5. classes of tree edges (tree edge (parent(w), w) named w): a non-tree
//      edge {u, w} between unrelated vertices joins u and w; a tree edge
//      (v, w) with v not a root joins w and v when low(w) < pre(v) or
//      high(w) >= pre(v) + size(v).  The classes are the biconnected
//      components (restricted to tree edges);
//   6. a vertex is an articulation point iff its tree edges lie in two or more
//      classes.
// Integer arithmetic only: the result is the graph's (unique) set of cut vertices.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <string>
#include <vector>

#include "absorb_kernels.h"
#include "common.h"
#include "xfe_hip.h"

namespace xfe {
namespace {

constexpr int kLoBlock = 256;

int lo_grid(long long work) {
    long long g = (work + kLoBlock - 1) / kLoBlock;
    return static_cast<int>(std::max(1LL, std::min(g, 8192LL)));
}

#define LO_FOR(v) for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
// one wavefront per vertex (the loop variable is uniform across the wave), lane = lane index
#define LO_FOR_WAVE(v)                                                                     \
    for (int v = blockIdx.x * (blockDim.x / kWave) + static_cast<int>(threadIdx.x) / kWave; v < n; \
         v += gridDim.x * (blockDim.x / kWave))

__device__ __forceinline__ int wave_min_int(int m) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) m = min(m, __shfl_xor(m, off, kWave));
    return m;
}

__device__ __forceinline__ int wave_or_int(int m) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) m |= __shfl_xor(m, off, kWave);
    return m;
}

__global__ void lo_iota_kernel(int* label, int n) {
    LO_FOR(v) label[v] = v;
}

// label[v] = min(label[v], label of every neighbour); *changed when lowered
// (a wavefront per vertex: high-degree vertices do not serialize a launch)
__global__ void lo_prop_kernel(const int* __restrict__ ptr, const int* __restrict__ adj, int* label, int n,
                               int* changed) {
    const int lane = threadIdx.x & (kWave - 1);
    LO_FOR_WAVE(v) {
        int m = label[v];
        for (int q = ptr[v] + lane; q < ptr[v + 1]; q += kWave) m = min(m, label[adj[q]]);
        m = wave_min_int(m);
        if (lane == 0 && m < label[v]) {
            atomicMin(&label[v], m);
            *changed = 1;
        }
    }
}

// pointer jumping: label[v] = label[label[v]]
__global__ void lo_jump_kernel(int* label, int n) {
    LO_FOR(v) {
        const int l = label[v];
        const int ll = label[l];
        if (ll < l) atomicMin(&label[v], ll);
    }
}

__global__ void lo_roots_kernel(const int* __restrict__ label, int* depth, int n) {
    LO_FOR(v) depth[v] = label[v] == v ? 0 : -1;
}

// one BFS level: neighbours of depth-d vertices without a depth get d + 1
// (a wavefront per vertex)
__global__ void lo_bfs_kernel(const int* __restrict__ ptr, const int* __restrict__ adj, int* depth, int n, int d,
                              int* changed) {
    const int lane = threadIdx.x & (kWave - 1);
    LO_FOR_WAVE(v) {
        if (depth[v] != d) continue;
        int found = 0;
        for (int q = ptr[v] + lane; q < ptr[v + 1]; q += kWave) {
            const int x = adj[q];
            if (depth[x] == -1 && atomicCAS(&depth[x], -1, d + 1) == -1) found = 1;
        }
        found = wave_or_int(found);
        if (lane == 0 && found) *changed = 1;
    }
}

// parent = smallest neighbour one level up; size = 1 (a wavefront per vertex,
// the lanes over its neighbours and a wave minimum)
__global__ void lo_parent_kernel(const int* __restrict__ ptr, const int* __restrict__ adj,
                                 const int* __restrict__ depth, int* parent, int* size, int n) {
    const int lane = threadIdx.x & (kWave - 1);
    LO_FOR_WAVE(v) {
        const int dv = depth[v];
        int p = INT_MAX;
        if (dv > 0) {
            for (int q = ptr[v] + lane; q < ptr[v + 1]; q += kWave) {
                const int x = adj[q];
                if (depth[x] == dv - 1) p = min(p, x);
            }
            p = wave_min_int(p);
        }
        if (lane == 0) {
            parent[v] = (dv > 0 && p != INT_MAX) ? p : -1;
            size[v] = 1;
        }
    }
}

// bottom-up: size[parent] += size[v] for the vertices of depth d
__global__ void lo_size_kernel(const int* __restrict__ depth, const int* __restrict__ parent, int* size, int n,
                               int d) {
    LO_FOR(v) if (depth[v] == d) atomicAdd(&size[parent[v]], size[v]);
}

// top-down: the children of each depth-d vertex p get consecutive intervals
// after pre[p], in adjacency order (a wavefront per vertex: each chunk of 64
// neighbours takes an exclusive prefix sum of its children's sizes)
__global__ void lo_pre_kernel(const int* __restrict__ ptr, const int* __restrict__ adj,
                              const int* __restrict__ depth, const int* __restrict__ parent,
                              const int* __restrict__ size, int* pre, int n, int d) {
    const int lane = threadIdx.x & (kWave - 1);
    LO_FOR_WAVE(p) {
        if (depth[p] != d) continue;
        if (d == 0 && lane == 0) pre[p] = 0;
        int next = (d == 0 ? 0 : pre[p]) + 1;
        for (int q0 = ptr[p]; q0 < ptr[p + 1]; q0 += kWave) {
            const int q = q0 + lane;
            int x = -1, sz = 0;
            if (q < ptr[p + 1]) {
                x = adj[q];
                if (parent[x] == p) sz = size[x];
            }
            int incl = sz;  // inclusive scan over the lanes
#pragma unroll
            for (int off = 1; off < kWave; off <<= 1) {
                const int t = __shfl_up(incl, off, kWave);
                if (lane >= off) incl += t;
            }
            if (sz > 0) pre[x] = next + incl - sz;
            next += __shfl(incl, kWave - 1, kWave);
        }
    }
}

// own pre and the pre of every neighbour except the parent (a wavefront per
// vertex: wave minimum and maximum over its neighbours)
__global__ void lo_local_kernel(const int* __restrict__ ptr, const int* __restrict__ adj,
                                const int* __restrict__ parent, const int* __restrict__ pre, int* low, int* high,
                                int n) {
    const int lane = threadIdx.x & (kWave - 1);
    LO_FOR_WAVE(v) {
        const int pv = parent[v];
        int lo = pre[v], hi = lo;
        for (int q = ptr[v] + lane; q < ptr[v + 1]; q += kWave) {
            const int x = adj[q];
            if (x == pv) continue;
            const int px = pre[x];
            lo = min(lo, px);
            hi = max(hi, px);
        }
        lo = wave_min_int(lo);
        hi = -wave_min_int(-hi);
        if (lane == 0) {
            low[v] = lo;
            high[v] = hi;
        }
    }
}

// bottom-up subtree min / max
__global__ void lo_lowhigh_kernel(const int* __restrict__ depth, const int* __restrict__ parent, int* low,
                                  int* high, int n, int d) {
    LO_FOR(v) {
        if (depth[v] != d) continue;
        atomicMin(&low[parent[v]], low[v]);
        atomicMax(&high[parent[v]], high[v]);
    }
}

struct Tree {
    const int* ptr;
    const int* adj;
    const int* depth;
    const int* parent;
    const int* pre;
    const int* size;
    const int* low;
    const int* high;
};

// tree edge (parent(w), w), parent(w) = v not a root, joins v's own tree edge
__device__ __forceinline__ bool lo_rule2(const Tree& t, int w) {
    const int v = t.parent[w];
    return v >= 0 && t.depth[v] > 0 && (t.low[w] < t.pre[v] || t.high[w] >= t.pre[v] + t.size[v]);
}

// one propagation sweep over the tree-edge classes (labels named by vertex;
// a wavefront per vertex, the lanes over its neighbours)
__global__ void lo_class_kernel(Tree t, int* cls, int n, int* changed) {
    const int lane = threadIdx.x & (kWave - 1);
    LO_FOR_WAVE(u) {
        if (t.depth[u] <= 0) continue;
        int m = cls[u];
        const int pu = t.parent[u];
        if (lane == 0 && lo_rule2(t, u)) m = min(m, cls[pu]);
        for (int q = t.ptr[u] + lane; q < t.ptr[u + 1]; q += kWave) {
            const int w = t.adj[q];
            if (t.parent[w] == u) {  // child: rule 2 seen from the parent's edge
                if (lo_rule2(t, w)) m = min(m, cls[w]);
                continue;
            }
            if (w == pu || t.depth[w] <= 0) continue;
            const bool related = (t.pre[u] <= t.pre[w] && t.pre[w] < t.pre[u] + t.size[u]) ||
                                 (t.pre[w] <= t.pre[u] && t.pre[u] < t.pre[w] + t.size[w]);
            if (!related) m = min(m, cls[w]);  // rule 1
        }
        m = wave_min_int(m);
        if (lane == 0 && m < cls[u]) {
            atomicMin(&cls[u], m);
            *changed = 1;
        }
    }
}

__global__ void lo_rootmin_kernel(const int* __restrict__ depth, const int* __restrict__ parent,
                                  const int* __restrict__ cls, int* rootmin, int n) {
    LO_FOR(w) {
        const int v = parent[w];
        if (v >= 0 && depth[v] == 0) atomicMin(&rootmin[v], cls[w]);
    }
}

__global__ void lo_test_kernel(const int* __restrict__ depth, const int* __restrict__ parent,
                               const int* __restrict__ cls, const int* __restrict__ rootmin, int* artic, int n) {
    LO_FOR(w) {
        const int v = parent[w];
        if (v < 0) continue;
        const int ref = depth[v] > 0 ? cls[v] : rootmin[v];
        if (cls[w] != ref) artic[v] = 1;
    }
}

#undef LO_FOR
#undef LO_FOR_WAVE

}  // namespace
}  // namespace xfe

using namespace xfe;

extern "C" {

int xfe_articulation_points(int n_vertices, const int* adj_ptr, const int* adj, unsigned char* artic) {
    clear_error();
    if (n_vertices < 0 || (n_vertices > 0 && (!adj_ptr || !artic))) {
        set_error("xfe_articulation_points: invalid arguments");
        return 0;
    }
    if (n_vertices == 0) return 1;
    const int n = n_vertices;
    if (adj_ptr[0] != 0) {
        set_error("xfe_articulation_points: adj_ptr[0] must be 0");
        return 0;
    }
    for (int v = 0; v < n; ++v)
        if (adj_ptr[v + 1] < adj_ptr[v]) {
            set_error("xfe_articulation_points: adj_ptr must be non-decreasing");
            return 0;
        }
    const int E = adj_ptr[n];
    if (E > 0 && !adj) {
        set_error("xfe_articulation_points: invalid arguments");
        return 0;
    }
    for (int q = 0; q < E; ++q)
        if (adj[q] < 0 || adj[q] >= n) {
            set_error("xfe_articulation_points: neighbour outside [0, n_vertices)");
            return 0;
        }
    {
        // a repeated neighbour would give a child two preorder numbers (two lanes write
        // it), so the simple-graph precondition is checked: one stamp per vertex
        std::vector<int> stamp(n, -1);
        for (int v = 0; v < n; ++v)
            for (int q = adj_ptr[v]; q < adj_ptr[v + 1]; ++q) {
                const int u = adj[q];
                if (u == v) {
                    set_error("xfe_articulation_points: self loop at vertex " + std::to_string(v));
                    return 0;
                }
                if (stamp[u] == v) {
                    set_error("xfe_articulation_points: repeated edge " + std::to_string(v) + "-" +
                              std::to_string(u));
                    return 0;
                }
                stamp[u] = v;
            }
    }
    hipStream_t s = nullptr;
    DevBuf<int> d_ptr, d_adj, label, depth, parent, size, pre, low, high, cls, rootmin, art, flag;
    auto release_all = [&]() {
        for (DevBuf<int>* b : {&d_ptr, &d_adj, &label, &depth, &parent, &size, &pre, &low, &high, &cls, &rootmin,
                               &art, &flag})
            b->release();
        if (s) (void)hipStreamDestroy(s);
        s = nullptr;
    };
    try {
        (void)hipGetLastError();
        XFE_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        d_ptr.ensure(n + 1); d_adj.ensure(std::max(E, 1));
        label.ensure(n); depth.ensure(n); parent.ensure(n); size.ensure(n); pre.ensure(n);
        low.ensure(n); high.ensure(n); cls.ensure(n); rootmin.ensure(n); art.ensure(n); flag.ensure(1);
        XFE_CHECK(hipMemcpyAsync(d_ptr.p, adj_ptr, sizeof(int) * (n + 1), hipMemcpyHostToDevice, s));
        if (E > 0) XFE_CHECK(hipMemcpyAsync(d_adj.p, adj, sizeof(int) * E, hipMemcpyHostToDevice, s));
        const int g = lo_grid(n);
        const int gw = lo_grid(static_cast<long long>(n) * kWave);  // one wavefront per vertex
        int changed = 0;
        // every sweep lowers some label or is the last one: at most n sweeps
        auto until_stable = [&](auto&& sweep) {
            for (int it = 0; it <= n; ++it) {
                XFE_CHECK(hipMemsetAsync(flag.p, 0, sizeof(int), s));
                sweep();
                XFE_CHECK(hipGetLastError());
                XFE_CHECK(hipMemcpyAsync(&changed, flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
                XFE_CHECK(hipStreamSynchronize(s));
                if (!changed) return;
            }
            throw Error("xfe_articulation_points: label propagation did not settle");
        };
        // 1. components
        hipLaunchKernelGGL(lo_iota_kernel, dim3(g), dim3(kLoBlock), 0, s, label.p, n);
        until_stable([&] {
            hipLaunchKernelGGL(lo_prop_kernel, dim3(gw), dim3(kLoBlock), 0, s, d_ptr.p, d_adj.p, label.p, n, flag.p);
            hipLaunchKernelGGL(lo_jump_kernel, dim3(g), dim3(kLoBlock), 0, s, label.p, n);
        });
        // 2. BFS forest from the component roots; maxd ends as the deepest level
        hipLaunchKernelGGL(lo_roots_kernel, dim3(g), dim3(kLoBlock), 0, s, label.p, depth.p, n);
        int maxd = 0;
        for (; maxd < n; ++maxd) {
            XFE_CHECK(hipMemsetAsync(flag.p, 0, sizeof(int), s));
            hipLaunchKernelGGL(lo_bfs_kernel, dim3(gw), dim3(kLoBlock), 0, s, d_ptr.p, d_adj.p, depth.p, n, maxd,
                               flag.p);
            XFE_CHECK(hipGetLastError());
            XFE_CHECK(hipMemcpyAsync(&changed, flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
            XFE_CHECK(hipStreamSynchronize(s));
            if (!changed) break;
        }
        hipLaunchKernelGGL(lo_parent_kernel, dim3(gw), dim3(kLoBlock), 0, s, d_ptr.p, d_adj.p, depth.p, parent.p,
                           size.p, n);
        // 3. sizes bottom-up, preorder top-down
        for (int d = maxd; d >= 1; --d)
            hipLaunchKernelGGL(lo_size_kernel, dim3(g), dim3(kLoBlock), 0, s, depth.p, parent.p, size.p, n, d);
        for (int d = 0; d <= maxd; ++d)
            hipLaunchKernelGGL(lo_pre_kernel, dim3(gw), dim3(kLoBlock), 0, s, d_ptr.p, d_adj.p, depth.p, parent.p,
                               size.p, pre.p, n, d);
        // 4. low / high
        hipLaunchKernelGGL(lo_local_kernel, dim3(gw), dim3(kLoBlock), 0, s, d_ptr.p, d_adj.p, parent.p, pre.p,
                           low.p, high.p, n);
        for (int d = maxd; d >= 1; --d)
            hipLaunchKernelGGL(lo_lowhigh_kernel, dim3(g), dim3(kLoBlock), 0, s, depth.p, parent.p, low.p, high.p, n,
                               d);
        XFE_CHECK(hipGetLastError());
        // 5. classes of tree edges
        Tree t{d_ptr.p, d_adj.p, depth.p, parent.p, pre.p, size.p, low.p, high.p};
        hipLaunchKernelGGL(lo_iota_kernel, dim3(g), dim3(kLoBlock), 0, s, cls.p, n);
        until_stable([&] {
            hipLaunchKernelGGL(lo_class_kernel, dim3(gw), dim3(kLoBlock), 0, s, t, cls.p, n, flag.p);
            hipLaunchKernelGGL(lo_jump_kernel, dim3(g), dim3(kLoBlock), 0, s, cls.p, n);
        });
        // 6. two or more classes among a vertex's tree edges
        std::vector<int> h(n, INT_MAX);
        XFE_CHECK(hipMemcpyAsync(rootmin.p, h.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
        XFE_CHECK(hipMemsetAsync(art.p, 0, sizeof(int) * n, s));
        hipLaunchKernelGGL(lo_rootmin_kernel, dim3(g), dim3(kLoBlock), 0, s, depth.p, parent.p, cls.p, rootmin.p, n);
        hipLaunchKernelGGL(lo_test_kernel, dim3(g), dim3(kLoBlock), 0, s, depth.p, parent.p, cls.p, rootmin.p, art.p,
                           n);
        XFE_CHECK(hipGetLastError());
        XFE_CHECK(hipMemcpyAsync(h.data(), art.p, sizeof(int) * n, hipMemcpyDeviceToHost, s));
        XFE_CHECK(hipStreamSynchronize(s));
        for (int v = 0; v < n; ++v) artic[v] = static_cast<unsigned char>(h[v] != 0);
        release_all();
        return 1;
    } catch (const std::exception& e) {
        set_error(e.what());
        release_all();
        return 0;
    }
}

}  // extern "C"

namespace {

struct UnionFindHost {
    std::vector<int> p;
    explicit UnionFindHost(int n) : p(n) {
        for (int i = 0; i < n; ++i) p[i] = i;
    }
    int find(int x) {
        while (p[x] != x) {
            p[x] = p[p[x]];
            x = p[x];
        }
        return x;
    }
    void unite(int a, int b) {
        a = find(a);
        b = find(b);
        if (a != b) p[std::max(a, b)] = std::min(a, b);
    }
};

// dense labels in first-appearance row order (relabel_dense, src/akm_kss.cpp:383):
// a lookup table over [min, max] when the range is at most 4n + 1024, else the
// rows sorted by id (stable), each id's first row found, and the ids ranked by
// that row -- the same labels without a table the size of the id range.
int dense_first_appearance(const int* ids, int n, std::vector<int>& out) {
    out.resize(n);
    if (n == 0) return 0;
    const auto mm = std::minmax_element(ids, ids + n);
    const long long lo = *mm.first, range = static_cast<long long>(*mm.second) - lo + 1;
    if (range <= 4LL * n + 1024) {
        std::vector<int> lut(static_cast<size_t>(range), -1);
        int next = 0;
        for (int i = 0; i < n; ++i) {
            int& slot = lut[static_cast<size_t>(ids[i] - lo)];
            if (slot < 0) slot = next++;
            out[i] = slot;
        }
        return next;
    }
    std::vector<int> rows(n);
    for (int i = 0; i < n; ++i) rows[i] = i;
    std::stable_sort(rows.begin(), rows.end(), [&](int a, int b) { return ids[a] < ids[b]; });
    // first row of each distinct id, in id order; the rows of one id share it
    std::vector<int> first_of_run;   // first row per distinct id
    std::vector<int> run_of_row(n);  // distinct-id index of each row
    for (int k = 0; k < n; ++k) {
        const int r = rows[k];
        if (k == 0 || ids[r] != ids[rows[k - 1]]) first_of_run.push_back(r);
        run_of_row[r] = static_cast<int>(first_of_run.size()) - 1;
    }
    const int m = static_cast<int>(first_of_run.size());
    std::vector<int> runs(m);
    for (int j = 0; j < m; ++j) runs[j] = j;
    std::sort(runs.begin(), runs.end(), [&](int a, int b) { return first_of_run[a] < first_of_run[b]; });
    std::vector<int> label(m);
    for (int j = 0; j < m; ++j) label[runs[j]] = j;
    for (int i = 0; i < n; ++i) out[i] = label[run_of_row[i]];
    return m;
}

}  // namespace

extern "C" {

int xfe_akm_leave_out_set(const int* worker_ids, const int* firm_ids, int n, const long long* fweights,
                          unsigned char* keep, xfe_leave_out_result* out) {
    clear_error();
    if (n < 0 || !out || (n > 0 && (!worker_ids || !firm_ids || !keep))) {
        set_error("xfe_akm_leave_out_set: invalid arguments");
        return 0;
    }
    std::memset(out, 0, sizeof(*out));
    out->n_obs_input = n;
    if (n == 0) return 1;
    for (int i = 0; i < n; ++i)
        if (worker_ids[i] < 0 || firm_ids[i] < 0 || (fweights && fweights[i] < 0)) {
            set_error("xfe_akm_leave_out_set: ids and frequency weights must be non-negative");
            return 0;
        }
    try {
        std::vector<int> w0, f0;
        const int W0 = dense_first_appearance(worker_ids, n, w0);
        const int F0 = dense_first_appearance(firm_ids, n, f0);
        // match table: rows by (worker, firm), stable (the device sort of the match keys)
        std::vector<uint64_t> key(n);
        for (int i = 0; i < n; ++i)
            key[i] = static_cast<uint64_t>(w0[i]) * static_cast<uint64_t>(F0) + static_cast<uint64_t>(f0[i]);
        std::vector<int64_t> order(n);
        if (!xfe_akm_stable_sort_u64(key.data(), order.data(), static_cast<size_t>(n)))
            throw Error(std::string("match-key sort: ") + xfe_last_error());
        out->gpu_used = 1;
        std::vector<int> mw, mf, row_match(n);
        std::vector<long long> mc;
        for (int k = 0; k < n; ++k) {
            const int row = static_cast<int>(order[k]);
            if (mw.empty() || mw.back() != w0[row] || mf.back() != f0[row]) {
                mw.push_back(w0[row]);
                mf.push_back(f0[row]);
                mc.push_back(0);
            }
            mc.back() += fweights ? fweights[row] : 1;
            row_match[row] = static_cast<int>(mw.size()) - 1;
        }
        const int M = static_cast<int>(mw.size());
        std::vector<unsigned char> active(M, 1);
        std::vector<int> firm_orig(F0);
        for (int i = 0; i < n; ++i) firm_orig[f0[i]] = firm_ids[i];
        // largest component: most firms, then most person-year rows, then the
        // smallest original firm id (src/akm_kss.cpp:636-720)
        auto largest_cc = [&]() {
            UnionFindHost uf(W0 + F0);
            for (int m = 0; m < M; ++m)
                if (active[m]) uf.unite(mw[m], W0 + mf[m]);
            std::vector<long long> obs(W0 + F0, 0);
            std::vector<int> firms(W0 + F0, 0), minf(W0 + F0, INT_MAX);
            std::vector<unsigned char> seen(F0, 0);
            for (int m = 0; m < M; ++m) {
                if (!active[m]) continue;
                obs[uf.find(mw[m])] += mc[m];
                if (!seen[mf[m]]) {
                    seen[mf[m]] = 1;
                    const int r = uf.find(W0 + mf[m]);
                    ++firms[r];
                    minf[r] = std::min(minf[r], firm_orig[mf[m]]);
                }
            }
            int best = -1;
            for (int r = 0; r < W0 + F0; ++r) {
                if (firms[r] == 0 && obs[r] == 0) continue;
                if (best < 0 || firms[r] > firms[best] ||
                    (firms[r] == firms[best] && (obs[r] > obs[best] || (obs[r] == obs[best] && minf[r] < minf[best]))))
                    best = r;
            }
            if (best < 0) return;
            for (int m = 0; m < M; ++m)
                if (active[m] && uf.find(mw[m]) != best) active[m] = 0;
        };
        largest_cc();
        for (int m = 0; m < M; ++m)
            if (active[m]) out->n_obs_connected += mc[m];
        // pruning: drop movers that are cut vertices of the mover-firm graph
        // until none is (src/akm_kss.cpp:736-870), the cut-vertex search on the device
        std::vector<int> deg(W0), mover_idx(W0), firm_idx(F0);
        while (true) {
            std::fill(deg.begin(), deg.end(), 0);
            for (int m = 0; m < M; ++m)
                if (active[m]) ++deg[mw[m]];
            std::fill(mover_idx.begin(), mover_idx.end(), -1);
            std::fill(firm_idx.begin(), firm_idx.end(), -1);
            int nmov = 0, nfrm = 0;
            long long nedges = 0;
            for (int m = 0; m < M; ++m) {
                if (!active[m] || deg[mw[m]] < 2) continue;
                ++nedges;
                if (mover_idx[mw[m]] < 0) mover_idx[mw[m]] = nmov++;
                if (firm_idx[mf[m]] < 0) firm_idx[mf[m]] = nfrm++;
            }
            if (nmov == 0) break;
            const int nv = nmov + nfrm;
            std::vector<int> ptr(nv + 1, 0), adj(2 * nedges);
            for (int m = 0; m < M; ++m) {
                if (!active[m] || deg[mw[m]] < 2) continue;
                ++ptr[mover_idx[mw[m]] + 1];
                ++ptr[nmov + firm_idx[mf[m]] + 1];
            }
            for (int v = 0; v < nv; ++v) ptr[v + 1] += ptr[v];
            std::vector<int> fill(ptr.begin(), ptr.end() - 1);
            for (int m = 0; m < M; ++m) {
                if (!active[m] || deg[mw[m]] < 2) continue;
                const int a = mover_idx[mw[m]], b = nmov + firm_idx[mf[m]];
                adj[fill[a]++] = b;
                adj[fill[b]++] = a;
            }
            std::vector<unsigned char> art(nv);
            if (!xfe_articulation_points(nv, ptr.data(), adj.data(), art.data()))
                throw Error(std::string("cut-vertex search: ") + xfe_last_error());
            bool any = false;
            std::vector<unsigned char> bad(W0, 0);
            for (int w = 0; w < W0; ++w)
                if (mover_idx[w] >= 0 && art[mover_idx[w]]) bad[w] = 1, any = true;
            if (!any) break;
            for (int m = 0; m < M; ++m)
                if (active[m] && bad[mw[m]]) active[m] = 0;
            largest_cc();
            ++out->prune_iterations;
        }
        // workers observed once leave (src/akm_kss.cpp:858-869)
        std::vector<long long> wobs(W0, 0);
        for (int m = 0; m < M; ++m)
            if (active[m]) wobs[mw[m]] += mc[m];
        for (int m = 0; m < M; ++m)
            if (active[m] && wobs[mw[m]] == 1) active[m] = 0;
        std::vector<int> wdeg(W0, 0);
        std::vector<unsigned char> fseen(F0, 0);
        for (int m = 0; m < M; ++m) {
            if (!active[m]) continue;
            ++out->n_matches;
            ++wdeg[mw[m]];
            if (!fseen[mf[m]]) fseen[mf[m]] = 1, ++out->n_firms;
        }
        for (int w = 0; w < W0; ++w) {
            if (wdeg[w] > 0) ++out->n_workers;
            if (wdeg[w] >= 2) ++out->n_movers;
            if (wdeg[w] == 1) ++out->n_stayers;
        }
        for (int i = 0; i < n; ++i) {
            keep[i] = active[row_match[i]] ? 1 : 0;
            if (keep[i]) out->n_obs += fweights ? fweights[i] : 1;
        }
        return 1;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 0;
    }
}

}  // extern "C"
