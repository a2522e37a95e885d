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

__global__ void lo_iota_kernel(int* label, int n) {
    LO_FOR(v) label[v] = v;
}

// label[v] = min(label[v], label of every neighbour); *changed when lowered
__global__ void lo_prop_kernel(const int* __restrict__ ptr, const int* __restrict__ adj, int* label, int n,
                               int* changed) {
    LO_FOR(v) {
        int m = label[v];
        for (int q = ptr[v]; q < ptr[v + 1]; ++q) m = min(m, label[adj[q]]);
        if (m < label[v]) {
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
__global__ void lo_bfs_kernel(const int* __restrict__ ptr, const int* __restrict__ adj, int* depth, int n, int d,
                              int* changed) {
    LO_FOR(v) {
        if (depth[v] != d) continue;
        for (int q = ptr[v]; q < ptr[v + 1]; ++q) {
            const int x = adj[q];
            if (depth[x] == -1 && atomicCAS(&depth[x], -1, d + 1) == -1) *changed = 1;
        }
    }
}

// parent = smallest neighbour one level up; size = 1
__global__ void lo_parent_kernel(const int* __restrict__ ptr, const int* __restrict__ adj,
                                 const int* __restrict__ depth, int* parent, int* size, int n) {
    LO_FOR(v) {
        int p = -1;
        if (depth[v] > 0) {
            for (int q = ptr[v]; q < ptr[v + 1]; ++q) {
                const int x = adj[q];
                if (depth[x] == depth[v] - 1 && (p < 0 || x < p)) p = x;
            }
        }
        parent[v] = p;
        size[v] = 1;
    }
}

// bottom-up: size[parent] += size[v] for the vertices of depth d
__global__ void lo_size_kernel(const int* __restrict__ depth, const int* __restrict__ parent, int* size, int n,
                               int d) {
    LO_FOR(v) if (depth[v] == d) atomicAdd(&size[parent[v]], size[v]);
}

// top-down: the children of each depth-d vertex p get consecutive intervals
// after pre[p], in adjacency order
__global__ void lo_pre_kernel(const int* __restrict__ ptr, const int* __restrict__ adj,
                              const int* __restrict__ depth, const int* __restrict__ parent,
                              const int* __restrict__ size, int* pre, int n, int d) {
    LO_FOR(p) {
        if (depth[p] != d) continue;
        if (d == 0) pre[p] = 0;
        int next = pre[p] + 1;
        for (int q = ptr[p]; q < ptr[p + 1]; ++q) {
            const int x = adj[q];
            if (parent[x] == p) {
                pre[x] = next;
                next += size[x];
            }
        }
    }
}

// own pre and the pre of every neighbour except the parent
__global__ void lo_local_kernel(const int* __restrict__ ptr, const int* __restrict__ adj,
                                const int* __restrict__ parent, const int* __restrict__ pre, int* low, int* high,
                                int n) {
    LO_FOR(v) {
        int lo = pre[v], hi = pre[v];
        for (int q = ptr[v]; q < ptr[v + 1]; ++q) {
            const int x = adj[q];
            if (x == parent[v]) continue;
            lo = min(lo, pre[x]);
            hi = max(hi, pre[x]);
        }
        low[v] = lo;
        high[v] = hi;
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

// one propagation sweep over the tree-edge classes (labels named by vertex)
__global__ void lo_class_kernel(Tree t, int* cls, int n, int* changed) {
    LO_FOR(u) {
        if (t.depth[u] <= 0) continue;
        int m = cls[u];
        const int pu = t.parent[u];
        if (lo_rule2(t, u)) m = min(m, cls[pu]);
        for (int q = t.ptr[u]; q < t.ptr[u + 1]; ++q) {
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
        if (m < cls[u]) {
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
            hipLaunchKernelGGL(lo_prop_kernel, dim3(g), dim3(kLoBlock), 0, s, d_ptr.p, d_adj.p, label.p, n, flag.p);
            hipLaunchKernelGGL(lo_jump_kernel, dim3(g), dim3(kLoBlock), 0, s, label.p, n);
        });
        // 2. BFS forest from the component roots; maxd ends as the deepest level
        hipLaunchKernelGGL(lo_roots_kernel, dim3(g), dim3(kLoBlock), 0, s, label.p, depth.p, n);
        int maxd = 0;
        for (; maxd < n; ++maxd) {
            XFE_CHECK(hipMemsetAsync(flag.p, 0, sizeof(int), s));
            hipLaunchKernelGGL(lo_bfs_kernel, dim3(g), dim3(kLoBlock), 0, s, d_ptr.p, d_adj.p, depth.p, n, maxd,
                               flag.p);
            XFE_CHECK(hipGetLastError());
            XFE_CHECK(hipMemcpyAsync(&changed, flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
            XFE_CHECK(hipStreamSynchronize(s));
            if (!changed) break;
        }
        hipLaunchKernelGGL(lo_parent_kernel, dim3(g), dim3(kLoBlock), 0, s, d_ptr.p, d_adj.p, depth.p, parent.p,
                           size.p, n);
        // 3. sizes bottom-up, preorder top-down
        for (int d = maxd; d >= 1; --d)
            hipLaunchKernelGGL(lo_size_kernel, dim3(g), dim3(kLoBlock), 0, s, depth.p, parent.p, size.p, n, d);
        for (int d = 0; d <= maxd; ++d)
            hipLaunchKernelGGL(lo_pre_kernel, dim3(g), dim3(kLoBlock), 0, s, d_ptr.p, d_adj.p, depth.p, parent.p,
                               size.p, pre.p, n, d);
        // 4. low / high
        hipLaunchKernelGGL(lo_local_kernel, dim3(g), dim3(kLoBlock), 0, s, d_ptr.p, d_adj.p, parent.p, pre.p, low.p,
                           high.p, n);
        for (int d = maxd; d >= 1; --d)
            hipLaunchKernelGGL(lo_lowhigh_kernel, dim3(g), dim3(kLoBlock), 0, s, depth.p, parent.p, low.p, high.p, n,
                               d);
        XFE_CHECK(hipGetLastError());
        // 5. classes of tree edges
        Tree t{d_ptr.p, d_adj.p, depth.p, parent.p, pre.p, size.p, low.p, high.p};
        hipLaunchKernelGGL(lo_iota_kernel, dim3(g), dim3(kLoBlock), 0, s, cls.p, n);
        until_stable([&] {
            hipLaunchKernelGGL(lo_class_kernel, dim3(g), dim3(kLoBlock), 0, s, t, cls.p, n, flag.p);
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
