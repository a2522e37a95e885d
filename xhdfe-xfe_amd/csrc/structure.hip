// Integer structure kernels on the device: iterative singleton dropping,
// bipartite mobility-group components with the reference's labelling, the
// reghdfe FE degrees of freedom, plus the FP64 Gram / cluster-meat reductions
// of the OLS step.  All integer results are bit-exact with the reference CPU
// code they restate (cited per function); none of them depends on atomic
// ordering (counts, XOR, min-hooking are order independent).
#include <hip/hip_runtime.h>
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <climits>
#include <cstring>
#include <vector>

#include "common.h"
#include "xfe_hip.h"

namespace xfe {

namespace {

constexpr int kB = 256;

int grid_of(long long work) {
    long long g = (work + kB - 1) / kB;
    if (g < 1) g = 1;
    if (g > 4096) g = 4096;
    return static_cast<int>(g);
}

// Per-call device context for the structure kernels (default-priority stream).
struct Ctx {
    hipStream_t s = nullptr;
    std::vector<void*> allocs;
    Ctx() { XFE_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)); }
    ~Ctx() {
        if (s) (void)hipStreamSynchronize(s);
        for (void* p : allocs) (void)hipFree(p);
        if (s) (void)hipStreamDestroy(s);
    }
    template <typename T>
    T* alloc(size_t count) {
        void* p = nullptr;
        XFE_CHECK(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
        allocs.push_back(p);
        return static_cast<T*>(p);
    }
    template <typename T>
    T* upload(const T* h, size_t count) {
        T* d = alloc<T>(count);
        XFE_CHECK(hipMemcpyAsync(d, h, count * sizeof(T), hipMemcpyHostToDevice, s));
        return d;
    }
    void sync() { XFE_CHECK(hipStreamSynchronize(s)); }
};

// ------------------------------------------------------------------ kernels

__global__ void minmax_kernel(const int* v, int n, int* out /* [min, max] */) {
    int mn = INT_MAX, mx = INT_MIN;
    for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) {
        mn = min(mn, v[i]);
        mx = max(mx, v[i]);
    }
    for (int off = 32; off > 0; off >>= 1) {
        mn = min(mn, __shfl_down(mn, off, 64));
        mx = max(mx, __shfl_down(mx, off, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&out[0], mn);
        atomicMax(&out[1], mx);
    }
}

__global__ void fill_int_kernel(int* v, long long n, int value) {
    for (long long i = blockIdx.x * (long long)kB + threadIdx.x; i < n; i += (long long)gridDim.x * kB)
        v[i] = value;
}

__global__ void offset_kernel(const int* raw, int n, int mn, int* out) {
    for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) out[i] = raw[i] - mn;
}

__global__ void iota_kernel(int* v, int n) {
    for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) v[i] = i;
}

// first row of each dense slot
__global__ void first_row_kernel(const int* slot, int n, int* first) {
    for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) atomicMin(&first[slot[i]], i);
}

// run heads of sorted keys -> flag 1 at heads
__global__ void head_flags_kernel(const int* sorted, int n, int* flags) {
    for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB)
        flags[i] = (i == 0 || sorted[i] != sorted[i - 1]) ? 1 : 0;
}

// compact id (run index) of each sorted position scattered back to its row
__global__ void scatter_rank_kernel(const int* run_index_inclusive, const int* rows, int n, int* out) {
    for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB)
        out[rows[i]] = run_index_inclusive[i] - 1;
}

// For each run (distinct key in sorted order) record its first row.
__global__ void run_first_row_kernel(const int* flags, const int* run_inc, const int* rows, int n,
                                     int* first) {
    for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB)
        if (flags[i]) first[run_inc[i] - 1] = rows[i];
}

__global__ void map_kernel(const int* idx, const int* table, int n, int* out) {
    for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) out[i] = table[idx[i]];
}

__global__ void present_flags_kernel(const int* first, int m, int* flags) {
    for (int j = blockIdx.x * kB + threadIdx.x; j < m; j += gridDim.x * kB)
        flags[j] = first[j] != INT_MAX ? 1 : 0;
}

__global__ void scatter_by_value_kernel(const int* keys_rank_order /* slot sorted by first row */,
                                        int count, int* node_of_slot) {
    for (int r = blockIdx.x * kB + threadIdx.x; r < count; r += gridDim.x * kB)
        node_of_slot[keys_rank_order[r]] = r;
}

// counts and XOR of row ids per level (compute_keep_mask_drop_singletons, :4135-4145)
__global__ void count_xor_kernel(const int* gid, int n, int* cnt, int* axr) {
    for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) {
        atomicAdd(&cnt[gid[i]], 1);
        atomicXor(&axr[gid[i]], i);
    }
}

// One round, phase 1: mark every kept row that is alone in some level
// (counts as of the start of the round).
__global__ void mark_singletons_kernel(const int* const* gids, const int* const* cnts, int D, int n,
                                       unsigned char* keep, unsigned char* dropped_now, int* ndrop) {
    int local = 0;
    for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) {
        unsigned char dn = 0;
        if (keep[i]) {
            for (int d = 0; d < D; ++d) {
                if (cnts[d][gids[d][i]] == 1) { dn = 1; break; }
            }
        }
        dropped_now[i] = dn;
        local += dn;
    }
    for (int off = 32; off > 0; off >>= 1) local += __shfl_down(local, off, 64);
    if ((threadIdx.x & 63) == 0 && local) atomicAdd(ndrop, local);
}

// Phase 2: apply the drops (count -1, XOR the row out).
__global__ void apply_drops_kernel(const int* const* gids, int* const* cnts, int* const* axrs, int D,
                                   int n, unsigned char* keep, const unsigned char* dropped_now) {
    for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) {
        if (!dropped_now[i]) continue;
        keep[i] = 0;
        for (int d = 0; d < D; ++d) {
            const int g = gids[d][i];
            atomicSub(&cnts[d][g], 1);
            atomicXor(&axrs[d][g], i);
        }
    }
}

__global__ void init_keep_kernel(unsigned char* keep, int n) {
    for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) keep[i] = 1;
}

// Union-find by minimum-id hooking over the bipartite edge list.
// Loads bypass the non-coherent per-CU cache: other workgroups relink roots
// concurrently with atomicCAS and a stale root would make the CAS retry loop
// spin on the same cached value.
__device__ __forceinline__ int load_parent(const int* parent, int x) {
    return __hip_atomic_load(parent + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int find_root(const int* parent, int x) {
    int p = load_parent(parent, x);
    while (p != x) {
        x = p;
        p = load_parent(parent, x);
    }
    return x;
}

__global__ void hook_kernel(const int* a, const int* b, int n, int Ga, int* parent, int* changed) {
    // parent[v] <= v holds throughout (hooks link a root to a smaller root,
    // compression links a node to its root), so root walks terminate.
    for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) {
        const int u = a[i], v = Ga + b[i];
        while (true) {
            const int ru = find_root(parent, u);
            const int rv = find_root(parent, v);
            if (ru == rv) break;
            const int hi = max(ru, rv), lo = min(ru, rv);
            if (atomicCAS(&parent[hi], hi, lo) == hi) {
                *changed = 1;
                break;
            }
        }
    }
}

__global__ void compress_kernel(int* parent, int m) {
    for (int v = blockIdx.x * kB + threadIdx.x; v < m; v += gridDim.x * kB) parent[v] = find_root(parent, v);
}

// Flag the root of every row's component (roots are a-nodes: the smallest
// node of a component with an edge is on the a side).
__global__ void edge_root_flags_kernel(const int* a, const int* parent, int n, int* flags) {
    for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) flags[parent[a[i]]] = 1;
}

__global__ void label_rows_kernel(const int* a, const int* parent, const int* rank_inc, int n, int* labels) {
    for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB)
        labels[i] = rank_inc[parent[a[i]]];  // inclusive scan of root flags = 1-based rank
}

// fe_is_function_of_refined_fe: min/max of the coarse id per refined level.
__global__ void coarse_range_kernel(const int* refined_slot, const int* coarse, int n, int* cmin, int* cmax) {
    for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) {
        atomicMin(&cmin[refined_slot[i]], coarse[i]);
        atomicMax(&cmax[refined_slot[i]], coarse[i]);
    }
}

__global__ void any_mismatch_kernel(const int* cmin, const int* cmax, int m, int* flag) {
    for (int j = blockIdx.x * kB + threadIdx.x; j < m; j += gridDim.x * kB)
        if (cmin[j] != INT_MAX && cmin[j] != cmax[j]) atomicOr(flag, 1);
}

// ------------------------------------------------------------ host helpers

void inclusive_scan(Ctx& c, const int* in, int* out, int n) {
    size_t bytes = 0;
    XFE_CHECK(rocprim::inclusive_scan(nullptr, bytes, in, out, static_cast<size_t>(n),
                                      rocprim::plus<int>(), c.s));
    void* tmp = c.alloc<unsigned char>(bytes);
    XFE_CHECK(rocprim::inclusive_scan(tmp, bytes, in, out, static_cast<size_t>(n),
                                      rocprim::plus<int>(), c.s));
}

void sort_pairs(Ctx& c, const int* kin, int* kout, const int* vin, int* vout, int n) {
    size_t bytes = 0;
    XFE_CHECK(rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, static_cast<size_t>(n),
                                        0, 32, c.s));
    void* tmp = c.alloc<unsigned char>(bytes);
    XFE_CHECK(rocprim::radix_sort_pairs(tmp, bytes, kin, kout, vin, vout, static_cast<size_t>(n), 0,
                                        32, c.s));
}

void minmax(Ctx& c, const int* d_v, int n, int& mn, int& mx) {
    int* mm = c.alloc<int>(2);
    const int init[2] = {INT_MAX, INT_MIN};
    XFE_CHECK(hipMemcpyAsync(mm, init, sizeof(init), hipMemcpyHostToDevice, c.s));
    hipLaunchKernelGGL(minmax_kernel, dim3(grid_of(n)), dim3(kB), 0, c.s, d_v, n, mm);
    XFE_CHECK(hipGetLastError());
    int h[2];
    XFE_CHECK(hipMemcpyAsync(h, mm, sizeof(h), hipMemcpyDeviceToHost, c.s));
    c.sync();
    mn = h[0];
    mx = h[1];
}

// Any injective map of raw ids to [0, G): dense offset when the id range is
// compact (<= 2n and <= 50M, like the reference's dense paths), else the rank
// in sorted order.  Returns G.
int compact_ids(Ctx& c, const int* d_raw, int n, int* d_out) {
    int mn, mx;
    minmax(c, d_raw, n, mn, mx);
    const long long range = static_cast<long long>(mx) - mn + 1LL;
    if (range > 0 && range <= 50000000LL && range <= 2LL * n) {
        hipLaunchKernelGGL(offset_kernel, dim3(grid_of(n)), dim3(kB), 0, c.s, d_raw, n, mn, d_out);
        XFE_CHECK(hipGetLastError());
        return static_cast<int>(range);
    }
    int* rows = c.alloc<int>(n);
    int* skeys = c.alloc<int>(n);
    int* srows = c.alloc<int>(n);
    int* flags = c.alloc<int>(n);
    int* inc = c.alloc<int>(n);
    hipLaunchKernelGGL(iota_kernel, dim3(grid_of(n)), dim3(kB), 0, c.s, rows, n);
    sort_pairs(c, d_raw, skeys, rows, srows, n);
    hipLaunchKernelGGL(head_flags_kernel, dim3(grid_of(n)), dim3(kB), 0, c.s, skeys, n, flags);
    inclusive_scan(c, flags, inc, n);
    hipLaunchKernelGGL(scatter_rank_kernel, dim3(grid_of(n)), dim3(kB), 0, c.s, inc, srows, n, d_out);
    XFE_CHECK(hipGetLastError());
    int G = 0;
    XFE_CHECK(hipMemcpyAsync(&G, inc + (n - 1), sizeof(int), hipMemcpyDeviceToHost, c.s));
    c.sync();
    return G;
}

// Node numbering of build_lookup_compact (src/hdfe_regressor_v11.cpp:1354-1417):
// raw - min when min >= 0, the range is dense-able and expected == range;
// otherwise numbering by first appearance over rows.  Returns the node count.
int lookup_nodes(Ctx& c, const int* d_raw, int n, int expected, int* d_node) {
    int mn, mx;
    minmax(c, d_raw, n, mn, mx);
    const long long range = static_cast<long long>(mx) - mn + 1LL;
    const bool denseable = mn >= 0 && range > 0 && range <= 50000000LL && range <= 2LL * n;
    if (denseable && expected > 0 && expected == range) {
        hipLaunchKernelGGL(offset_kernel, dim3(grid_of(n)), dim3(kB), 0, c.s, d_raw, n, mn, d_node);
        XFE_CHECK(hipGetLastError());
        return static_cast<int>(range);
    }
    // distinct values -> first row; rank by first row
    int* slot = c.alloc<int>(n);
    int m = 0;
    int* first = nullptr;
    if (denseable) {
        m = static_cast<int>(range);
        hipLaunchKernelGGL(offset_kernel, dim3(grid_of(n)), dim3(kB), 0, c.s, d_raw, n, mn, slot);
        first = c.alloc<int>(m);
        hipLaunchKernelGGL(fill_int_kernel, dim3(grid_of(m)), dim3(kB), 0, c.s, first, (long long)m, INT_MAX);
        hipLaunchKernelGGL(first_row_kernel, dim3(grid_of(n)), dim3(kB), 0, c.s, slot, n, first);
        XFE_CHECK(hipGetLastError());
    } else {
        int* rows = c.alloc<int>(n);
        int* skeys = c.alloc<int>(n);
        int* srows = c.alloc<int>(n);
        int* flags = c.alloc<int>(n);
        int* inc = c.alloc<int>(n);
        hipLaunchKernelGGL(iota_kernel, dim3(grid_of(n)), dim3(kB), 0, c.s, rows, n);
        sort_pairs(c, d_raw, skeys, rows, srows, n);  // stable: equal keys keep row order
        hipLaunchKernelGGL(head_flags_kernel, dim3(grid_of(n)), dim3(kB), 0, c.s, skeys, n, flags);
        inclusive_scan(c, flags, inc, n);
        hipLaunchKernelGGL(scatter_rank_kernel, dim3(grid_of(n)), dim3(kB), 0, c.s, inc, srows, n, slot);
        XFE_CHECK(hipGetLastError());
        XFE_CHECK(hipMemcpyAsync(&m, inc + (n - 1), sizeof(int), hipMemcpyDeviceToHost, c.s));
        c.sync();
        first = c.alloc<int>(m);
        hipLaunchKernelGGL(run_first_row_kernel, dim3(grid_of(n)), dim3(kB), 0, c.s, flags, inc, srows, n, first);
        XFE_CHECK(hipGetLastError());
    }
    // sort slots by first row; absent dense slots carry INT_MAX and sort last
    int* slots = c.alloc<int>(m);
    int* sfirst = c.alloc<int>(m);
    int* sslots = c.alloc<int>(m);
    hipLaunchKernelGGL(iota_kernel, dim3(grid_of(m)), dim3(kB), 0, c.s, slots, m);
    sort_pairs(c, first, sfirst, slots, sslots, m);
    int* node_of_slot = c.alloc<int>(m);
    hipLaunchKernelGGL(scatter_by_value_kernel, dim3(grid_of(m)), dim3(kB), 0, c.s, sslots, m, node_of_slot);
    hipLaunchKernelGGL(map_kernel, dim3(grid_of(n)), dim3(kB), 0, c.s, slot, node_of_slot, n, d_node);
    XFE_CHECK(hipGetLastError());
    // count present slots
    int* pf = c.alloc<int>(m);
    int* pinc = c.alloc<int>(m);
    hipLaunchKernelGGL(present_flags_kernel, dim3(grid_of(m)), dim3(kB), 0, c.s, first, m, pf);
    inclusive_scan(c, pf, pinc, m);
    int present = 0;
    XFE_CHECK(hipMemcpyAsync(&present, pinc + (m - 1), sizeof(int), hipMemcpyDeviceToHost, c.s));
    c.sync();
    return present;
}

// Connected components of the bipartite graph (a-nodes 0..Ga-1, b-nodes
// Ga..Ga+Gb-1).  Minimum-id hooking makes every root the smallest node of its
// component, which is an a-node.  Returns the component count; when d_labels
// is given, writes 1-based ranks of the root a-node (the reference's order).
int components_device(Ctx& c, const int* d_a, const int* d_b, int n, int Ga, int Gb, int* d_labels) {
    const int m = Ga + Gb;
    int* parent = c.alloc<int>(m);
    int* changed = c.alloc<int>(1);
    hipLaunchKernelGGL(iota_kernel, dim3(grid_of(m)), dim3(kB), 0, c.s, parent, m);
    for (int it = 0; it < 10000; ++it) {
        XFE_CHECK(hipMemsetAsync(changed, 0, sizeof(int), c.s));
        hipLaunchKernelGGL(hook_kernel, dim3(grid_of(n)), dim3(kB), 0, c.s, d_a, d_b, n, Ga, parent, changed);
        hipLaunchKernelGGL(compress_kernel, dim3(grid_of(m)), dim3(kB), 0, c.s, parent, m);
        XFE_CHECK(hipGetLastError());
        int h = 0;
        XFE_CHECK(hipMemcpyAsync(&h, changed, sizeof(int), hipMemcpyDeviceToHost, c.s));
        c.sync();
        if (!h) break;
    }
    int* flags = c.alloc<int>(Ga);
    int* inc = c.alloc<int>(Ga);
    XFE_CHECK(hipMemsetAsync(flags, 0, sizeof(int) * Ga, c.s));
    hipLaunchKernelGGL(edge_root_flags_kernel, dim3(grid_of(n)), dim3(kB), 0, c.s, d_a, parent, n, flags);
    inclusive_scan(c, flags, inc, Ga);
    if (d_labels) {
        hipLaunchKernelGGL(label_rows_kernel, dim3(grid_of(n)), dim3(kB), 0, c.s, d_a, parent, inc, n, d_labels);
    }
    XFE_CHECK(hipGetLastError());
    int h = 0;
    XFE_CHECK(hipMemcpyAsync(&h, inc + (Ga - 1), sizeof(int), hipMemcpyDeviceToHost, c.s));
    c.sync();
    return h;
}

bool is_function_of_device(Ctx& c, const int* d_coarse, const int* d_refined, int n) {
    int* slot = c.alloc<int>(n);
    const int m = compact_ids(c, d_refined, n, slot);
    int* cmin = c.alloc<int>(m);
    int* cmax = c.alloc<int>(m);
    int* flag = c.alloc<int>(1);
    hipLaunchKernelGGL(fill_int_kernel, dim3(grid_of(m)), dim3(kB), 0, c.s, cmin, (long long)m, INT_MAX);
    hipLaunchKernelGGL(fill_int_kernel, dim3(grid_of(m)), dim3(kB), 0, c.s, cmax, (long long)m, INT_MIN);
    XFE_CHECK(hipMemsetAsync(flag, 0, sizeof(int), c.s));
    hipLaunchKernelGGL(coarse_range_kernel, dim3(grid_of(n)), dim3(kB), 0, c.s, slot, d_coarse, n, cmin, cmax);
    hipLaunchKernelGGL(any_mismatch_kernel, dim3(grid_of(m)), dim3(kB), 0, c.s, cmin, cmax, m, flag);
    XFE_CHECK(hipGetLastError());
    int h = 0;
    XFE_CHECK(hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, c.s));
    c.sync();
    return h == 0;
}

int dof_device(Ctx& c, const std::vector<const int*>& d_fes, int n, const std::vector<int>& levels,
               int* redundant, int* num_coefs) {
    const int D = static_cast<int>(d_fes.size());
    for (int d = 0; d < D; ++d) {
        redundant[d] = 0;
        num_coefs[d] = std::max(0, levels[d]);
    }
    if (D < 2) return D == 1 ? num_coefs[0] : 0;
    // coarsening (src/hdfe_regressor_v11.cpp:3789-3860)
    std::vector<unsigned char> coarsened(D, 0);
    bool any = false;
    for (int d = 0; d < D; ++d) {
        for (int e = 0; e < D; ++e) {
            if (d == e || levels[e] <= levels[d]) continue;
            if (is_function_of_device(c, d_fes[d], d_fes[e], n)) {
                coarsened[d] = 1;
                any = true;
                break;
            }
        }
    }
    if (any) {
        std::vector<const int*> eff;
        std::vector<int> eff_levels, eff_map;
        for (int d = 0; d < D; ++d)
            if (!coarsened[d]) { eff.push_back(d_fes[d]); eff_levels.push_back(levels[d]); eff_map.push_back(d); }
        if (!eff.empty() && static_cast<int>(eff.size()) < D) {
            std::vector<int> er(eff.size()), ec(eff.size());
            dof_device(c, eff, n, eff_levels, er.data(), ec.data());
            for (int d = 0; d < D; ++d)
                if (coarsened[d]) { redundant[d] = levels[d]; num_coefs[d] = 0; }
            for (size_t q = 0; q < eff.size(); ++q) { redundant[eff_map[q]] = er[q]; num_coefs[eff_map[q]] = ec[q]; }
            int df = 0;
            for (int d = 0; d < D; ++d) df += num_coefs[d];
            return df;
        }
    }
    // pairwise mobility groups (:3914-3970); counts do not depend on node numbering
    std::vector<int*> nodes(D);
    std::vector<int> G(D);
    for (int d = 0; d < D; ++d) {
        nodes[d] = c.alloc<int>(n);
        G[d] = compact_ids(c, d_fes[d], n, nodes[d]);
    }
    for (int d = 1; d < D; ++d) {
        int mx = 0;
        for (int j = 0; j < d; ++j)
            mx = std::max(mx, components_device(c, nodes[j], nodes[d], n, G[j], G[d], nullptr));
        redundant[d] = mx;
        num_coefs[d] = std::max(0, levels[d] - mx);
    }
    int df = 0;
    for (int d = 0; d < D; ++d) df += num_coefs[d];
    return df;
}

// ------------------------------------------------------------ OLS pieces

// Partial Gram of an n x C column-major block: block (tile) computes the
// products of an 8 x 8 column tile over a slice of rows.
__global__ void gram_kernel(const double* V, int n, int C, int ld, int tiles_per_side, double* partials) {
    const int tile = blockIdx.y;
    const int ti = tile / tiles_per_side, tj = tile % tiles_per_side;
    double acc[8][8];
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int b = 0; b < 8; ++b) acc[a][b] = 0.0;
    for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) {
        double xa[8], xb[8];
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            const int ca = ti * 8 + a, cb = tj * 8 + a;
            xa[a] = ca < C ? V[static_cast<size_t>(ca) * ld + i] : 0.0;
            xb[a] = cb < C ? V[static_cast<size_t>(cb) * ld + i] : 0.0;
        }
#pragma unroll
        for (int a = 0; a < 8; ++a)
#pragma unroll
            for (int b = 0; b < 8; ++b) acc[a][b] += xa[a] * xb[b];
    }
    __shared__ double red[4][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int a = 0; a < 8; ++a)
        for (int b = 0; b < 8; ++b) {
            double v = acc[a][b];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
            if (lane == 0) red[wave][a * 8 + b] = v;
        }
    __syncthreads();
    if (threadIdx.x < 64) {
        const double s = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
        partials[(static_cast<size_t>(blockIdx.x) * gridDim.y + tile) * 64 + threadIdx.x] = s;
    }
}

// Cluster scores: S[g, j] += x_ij e_i with e_i = y_i - x_i' beta.
__global__ void scores_kernel(const double* V, int n, int C, int ld, const double* beta, const int* cl,
                              double* S, int ldS) {
    const int k = C - 1;
    for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) {
        double e = V[i];
        for (int j = 0; j < k; ++j) e -= V[static_cast<size_t>(j + 1) * ld + i] * beta[j];
        for (int j = 0; j < k; ++j)
            atomicAdd(&S[static_cast<size_t>(j) * ldS + cl[i]], V[static_cast<size_t>(j + 1) * ld + i] * e);
    }
}

void gram_device(Ctx& c, const double* dV, int n, int C, int ld, double* gram_host) {
    const int T = (C + 7) / 8;
    const int gx = std::min(grid_of(n), 512);
    double* part = c.alloc<double>(static_cast<size_t>(gx) * T * T * 64);
    hipLaunchKernelGGL(gram_kernel, dim3(gx, T * T), dim3(kB), 0, c.s, dV, n, C, ld, T, part);
    XFE_CHECK(hipGetLastError());
    std::vector<double> h(static_cast<size_t>(gx) * T * T * 64);
    XFE_CHECK(hipMemcpyAsync(h.data(), part, h.size() * sizeof(double), hipMemcpyDeviceToHost, c.s));
    c.sync();
    for (int a = 0; a < C; ++a)
        for (int b = 0; b < C; ++b) {
            const int tile = (a / 8) * T + (b / 8);
            double s = 0.0;
            for (int blk = 0; blk < gx; ++blk) s += h[(static_cast<size_t>(blk) * T * T + tile) * 64 + (a % 8) * 8 + (b % 8)];
            gram_host[static_cast<size_t>(b) * C + a] = s;
        }
}

}  // namespace
}  // namespace xfe

using namespace xfe;

extern "C" {

int xfe_drop_singletons(const int* fes, int n, int num_fes, int max_rounds, unsigned char* keep,
                        int* rounds_used) {
    clear_error();
    if (!fes || !keep || n < 0 || num_fes < 0 || num_fes > XFE_MAX_DIMS) {
        set_error("xfe_drop_singletons: invalid arguments");
        return -1;
    }
    if (rounds_used) *rounds_used = 0;
    if (n == 0) return 0;
    if (num_fes == 0) {
        std::memset(keep, 1, n);
        return 0;
    }
    try {
        Ctx c;
        const int D = num_fes;
        std::vector<int*> gids(D), cnts(D), axrs(D);
        for (int d = 0; d < D; ++d) {
            const int* raw = c.upload(fes + static_cast<size_t>(d) * n, n);
            gids[d] = c.alloc<int>(n);
            const int G = compact_ids(c, raw, n, gids[d]);
            cnts[d] = c.alloc<int>(G);
            axrs[d] = c.alloc<int>(G);
            XFE_CHECK(hipMemsetAsync(cnts[d], 0, sizeof(int) * G, c.s));
            XFE_CHECK(hipMemsetAsync(axrs[d], 0, sizeof(int) * G, c.s));
            hipLaunchKernelGGL(count_xor_kernel, dim3(grid_of(n)), dim3(kB), 0, c.s, gids[d], n, cnts[d], axrs[d]);
            XFE_CHECK(hipGetLastError());
        }
        int** d_gids = c.upload(gids.data(), D);
        int** d_cnts = c.upload(cnts.data(), D);
        int** d_axrs = c.upload(axrs.data(), D);
        unsigned char* d_keep = c.alloc<unsigned char>(n);
        unsigned char* d_now = c.alloc<unsigned char>(n);
        int* d_nd = c.alloc<int>(1);
        hipLaunchKernelGGL(init_keep_kernel, dim3(grid_of(n)), dim3(kB), 0, c.s, d_keep, n);
        const int iters = std::max(1, max_rounds);
        int dropped = 0, rounds = 0;
        for (int it = 0; it < iters; ++it) {
            XFE_CHECK(hipMemsetAsync(d_nd, 0, sizeof(int), c.s));
            hipLaunchKernelGGL(mark_singletons_kernel, dim3(grid_of(n)), dim3(kB), 0, c.s,
                               const_cast<const int* const*>(d_gids), const_cast<const int* const*>(d_cnts), D, n, d_keep, d_now, d_nd);
            XFE_CHECK(hipGetLastError());
            int nd = 0;
            XFE_CHECK(hipMemcpyAsync(&nd, d_nd, sizeof(int), hipMemcpyDeviceToHost, c.s));
            c.sync();
            if (nd == 0) break;
            hipLaunchKernelGGL(apply_drops_kernel, dim3(grid_of(n)), dim3(kB), 0, c.s, const_cast<const int* const*>(d_gids), d_cnts,
                               d_axrs, D, n, d_keep, d_now);
            XFE_CHECK(hipGetLastError());
            dropped += nd;
            rounds = it + 1;
        }
        XFE_CHECK(hipMemcpyAsync(keep, d_keep, n, hipMemcpyDeviceToHost, c.s));
        c.sync();
        if (rounds_used) *rounds_used = rounds;
        return dropped;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

int xfe_bipartite_components(const int* a, const int* b, int n, int expected_a, int expected_b,
                             int* labels) {
    clear_error();
    if (!a || !b || n < 0) {
        set_error("xfe_bipartite_components: invalid arguments");
        return -1;
    }
    if (n == 0) return 0;
    try {
        Ctx c;
        const int* da = c.upload(a, n);
        const int* db = c.upload(b, n);
        int* na = c.alloc<int>(n);
        int* nb = c.alloc<int>(n);
        const int Ga = lookup_nodes(c, da, n, expected_a, na);
        (void)expected_b;  // the b-side numbering does not affect counts or labels
        const int Gb = compact_ids(c, db, n, nb);
        int* dl = labels ? c.alloc<int>(n) : nullptr;
        const int comps = components_device(c, na, nb, n, Ga, Gb, dl);
        if (labels) {
            XFE_CHECK(hipMemcpyAsync(labels, dl, sizeof(int) * n, hipMemcpyDeviceToHost, c.s));
            c.sync();
        }
        return comps;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

int xfe_fe_dof(const int* fes, int n, int num_fes, const int* levels, int* redundant, int* num_coefs) {
    clear_error();
    if (!fes || !levels || !redundant || !num_coefs || n < 0 || num_fes < 0 || num_fes > XFE_MAX_DIMS) {
        set_error("xfe_fe_dof: invalid arguments");
        return -1;
    }
    if (num_fes == 0) return 0;
    try {
        Ctx c;
        std::vector<const int*> d_fes(num_fes);
        std::vector<int> lv(levels, levels + num_fes);
        for (int d = 0; d < num_fes; ++d) d_fes[d] = c.upload(fes + static_cast<size_t>(d) * n, n);
        return dof_device(c, d_fes, n, lv, redundant, num_coefs);
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

int xfe_gram(const double* V, int n, int C, double* gram) {
    clear_error();
    if (!V || !gram || n < 0 || C <= 0 || C > XFE_MAX_COLS) {
        set_error("xfe_gram: invalid arguments");
        return -1;
    }
    try {
        Ctx c;
        const double* dV = c.upload(V, static_cast<size_t>(n) * C);
        gram_device(c, dV, n, C, n, gram);
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

int xfe_cluster_meat(const double* V, int n, int C, const double* beta, const int* cluster,
                     int num_clusters, double* meat) {
    clear_error();
    if (!V || !beta || !cluster || !meat || n < 0 || C < 2 || C > XFE_MAX_COLS || num_clusters <= 0) {
        set_error("xfe_cluster_meat: invalid arguments");
        return -1;
    }
    try {
        Ctx c;
        const int k = C - 1;
        const double* dV = c.upload(V, static_cast<size_t>(n) * C);
        const double* db = c.upload(beta, k);
        const int* dcl = c.upload(cluster, n);
        const int ldS = num_clusters;
        double* S = c.alloc<double>(static_cast<size_t>(ldS) * k);
        XFE_CHECK(hipMemsetAsync(S, 0, sizeof(double) * static_cast<size_t>(ldS) * k, c.s));
        hipLaunchKernelGGL(scores_kernel, dim3(grid_of(n)), dim3(kB), 0, c.s, dV, n, C, n, db, dcl, S, ldS);
        XFE_CHECK(hipGetLastError());
        gram_device(c, S, num_clusters, k, ldS, meat);
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

}  // extern "C"
