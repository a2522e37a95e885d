// FE indexer and level weight sums on the device: the C ABI entry points
// xfe_build_indexer, xfe_build_indexer_device and xfe_weight_sums.
//
// The indexer is the reference's build_indexer (src/fe_absorption.cpp:845-940)
// with finalize_contiguous_groups (:547-590): raw int32 ids become compact ids
// in ascending raw-id order.  When the id range is at most min(2n, 50M) the id
// is raw - min (levels absent from the data keep their slots) unless fewer than
// 90 % of the slots are used, in which case the used slots are ranked; a wider
// range ranks the distinct ids.  Integer work only, so the result is the
// reference's bit for bit.  Device form: one min/max reduction; the dense path
// marks used slots in a byte table and ranks them by an exclusive scan; the
// sparse path radix-sorts (id, row) pairs, flags run heads and scans them.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_reduce.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <climits>
#include <cstring>

#include "common.h"
#include "xfe_hip.h"

namespace xfe {
namespace {

constexpr int kIxBlock = 256;
constexpr int kIxWave = 64;
constexpr long long kDenseRangeCap = 50000000LL;  // src/fe_absorption.cpp:867

int ix_grid(long long m) {
    const long long g = (m + kIxBlock - 1) / kIxBlock;
    return static_cast<int>(std::max(1LL, std::min(g, 4096LL)));
}

#define IX_FOR(i, n) for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < (n); \
                          i += static_cast<long long>(gridDim.x) * blockDim.x)

__global__ void ix_minmax_kernel(const int* __restrict__ raw, int n, int* mm) {
    int lo = INT_MAX, hi = INT_MIN;
    IX_FOR(i, n) {
        const int v = raw[i];
        lo = min(lo, v);
        hi = max(hi, v);
    }
#pragma unroll
    for (int off = kIxWave / 2; off > 0; off >>= 1) {
        lo = min(lo, __shfl_xor(lo, off, kIxWave));
        hi = max(hi, __shfl_xor(hi, off, kIxWave));
    }
    if ((threadIdx.x & (kIxWave - 1)) == 0) {
        atomicMin(&mm[0], lo);
        atomicMax(&mm[1], hi);
    }
}

__global__ void ix_init_minmax_kernel(int* mm) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        mm[0] = INT_MAX;
        mm[1] = INT_MIN;
    }
}

// dense path: id = raw - min; the slot is marked used (every writer stores 1)
__global__ void ix_dense_kernel(const int* __restrict__ raw, int n, int mn, int* __restrict__ gid, int* used) {
    IX_FOR(i, n) {
        const int m = raw[i] - mn;
        gid[i] = m;
        used[m] = 1;
    }
}

__global__ void ix_remap_kernel(int* gid, int n, const int* __restrict__ rank) {
    IX_FOR(i, n) gid[i] = rank[gid[i]];
}

__global__ void ix_iota_kernel(int* v, int n) {
    IX_FOR(i, n) v[i] = static_cast<int>(i);
}

// sparse path: run heads of the sorted ids
__global__ void ix_head_kernel(const int* __restrict__ skey, int n, int* head) {
    IX_FOR(k, n) head[k] = (k == 0 || skey[k] != skey[k - 1]) ? 1 : 0;
}

__global__ void ix_scatter_kernel(const int* __restrict__ pos, const int* __restrict__ inc, int n, int* gid) {
    IX_FOR(k, n) gid[pos[k]] = inc[k] - 1;
}

// finalize_contiguous_groups: ids non-decreasing over the rows
__global__ void ix_descent_kernel(const int* __restrict__ gid, int n, int* flag) {
    IX_FOR(i, n) if (i > 0 && gid[i] < gid[i - 1]) *flag = 1;
}

__global__ void ix_wsum_kernel(const int* __restrict__ gid, const double* __restrict__ w, int n, double* out) {
    IX_FOR(i, n) atomicAdd(&out[gid[i]], w ? w[i] : 1.0);
}

__global__ void ix_range_check_kernel(const int* __restrict__ gid, int n, int G, int* flag) {
    IX_FOR(i, n) if (gid[i] < 0 || gid[i] >= G) *flag = 1;
}

struct Stream {
    hipStream_t s = nullptr;
    Stream() { XFE_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)); }
    ~Stream() {
        if (s) (void)hipStreamDestroy(s);
    }
};

template <typename T>
struct Scoped : DevBuf<T> {
    ~Scoped() { this->release(); }
};

// Device indexer on device buffers; returns the level counts and the
// contiguity flag through host pointers.
void build_indexer_dev(const int* d_raw, int n, int* d_gid, int* num_groups, int* levels_present, int* contiguous,
                       hipStream_t s) {
    *num_groups = 0;
    *levels_present = 0;
    *contiguous = 0;
    if (n <= 0) return;
    const int grid = ix_grid(n);
    Scoped<int> mm, flag;
    mm.ensure(2);
    hipLaunchKernelGGL(ix_init_minmax_kernel, dim3(1), dim3(64), 0, s, mm.p);
    hipLaunchKernelGGL(ix_minmax_kernel, dim3(grid), dim3(kIxBlock), 0, s, d_raw, n, mm.p);
    XFE_CHECK(hipGetLastError());
    int hmm[2];
    XFE_CHECK(hipMemcpyAsync(hmm, mm.p, sizeof(hmm), hipMemcpyDeviceToHost, s));
    XFE_CHECK(hipStreamSynchronize(s));
    const long long range = static_cast<long long>(hmm[1]) - hmm[0] + 1LL;
    Scoped<unsigned char> tmp;
    size_t tmp_bytes = 0;
    if (range > 0 && range <= kDenseRangeCap && range <= 2LL * n) {
        const int groups = static_cast<int>(range);
        Scoped<int> used, rank, total;
        used.ensure(groups);
        rank.ensure(groups);
        total.ensure(1);
        XFE_CHECK(hipMemsetAsync(used.p, 0, sizeof(int) * groups, s));
        hipLaunchKernelGGL(ix_dense_kernel, dim3(grid), dim3(kIxBlock), 0, s, d_raw, n, hmm[0], d_gid, used.p);
        XFE_CHECK(hipGetLastError());
        // rank of every used slot (exclusive scan); unique = last rank + last flag
        XFE_CHECK(rocprim::exclusive_scan(nullptr, tmp_bytes, used.p, rank.p, 0, static_cast<size_t>(groups),
                                          rocprim::plus<int>(), s));
        tmp.ensure(tmp_bytes);
        XFE_CHECK(rocprim::exclusive_scan(tmp.p, tmp_bytes, used.p, rank.p, 0, static_cast<size_t>(groups),
                                          rocprim::plus<int>(), s));
        int last[2] = {0, 0};
        XFE_CHECK(hipMemcpyAsync(&last[0], rank.p + (groups - 1), sizeof(int), hipMemcpyDeviceToHost, s));
        XFE_CHECK(hipMemcpyAsync(&last[1], used.p + (groups - 1), sizeof(int), hipMemcpyDeviceToHost, s));
        XFE_CHECK(hipStreamSynchronize(s));
        const int unique = last[0] + last[1];
        *num_groups = groups;
        *levels_present = unique;
        // fewer than 90 % of the slots used: rank them (src/fe_absorption.cpp:887-902)
        if (static_cast<long long>(unique) * 10LL < range * 9LL) {
            hipLaunchKernelGGL(ix_remap_kernel, dim3(grid), dim3(kIxBlock), 0, s, d_gid, n, rank.p);
            XFE_CHECK(hipGetLastError());
            *num_groups = unique;
            *levels_present = unique;
        }
    } else {
        Scoped<int> iota, skey, pos, head, inc;
        iota.ensure(n); skey.ensure(n); pos.ensure(n); head.ensure(n); inc.ensure(n);
        hipLaunchKernelGGL(ix_iota_kernel, dim3(grid), dim3(kIxBlock), 0, s, iota.p, n);
        XFE_CHECK(rocprim::radix_sort_pairs(nullptr, tmp_bytes, d_raw, skey.p, iota.p, pos.p,
                                            static_cast<size_t>(n), 0, 32, s));
        tmp.ensure(tmp_bytes);
        XFE_CHECK(rocprim::radix_sort_pairs(tmp.p, tmp_bytes, d_raw, skey.p, iota.p, pos.p,
                                            static_cast<size_t>(n), 0, 32, s));
        hipLaunchKernelGGL(ix_head_kernel, dim3(grid), dim3(kIxBlock), 0, s, skey.p, n, head.p);
        size_t scan_bytes = 0;
        XFE_CHECK(rocprim::inclusive_scan(nullptr, scan_bytes, head.p, inc.p, static_cast<size_t>(n),
                                          rocprim::plus<int>(), s));
        tmp.ensure(scan_bytes);
        XFE_CHECK(rocprim::inclusive_scan(tmp.p, scan_bytes, head.p, inc.p, static_cast<size_t>(n),
                                          rocprim::plus<int>(), s));
        hipLaunchKernelGGL(ix_scatter_kernel, dim3(grid), dim3(kIxBlock), 0, s, pos.p, inc.p, n, d_gid);
        XFE_CHECK(hipGetLastError());
        int groups = 0;
        XFE_CHECK(hipMemcpyAsync(&groups, inc.p + (n - 1), sizeof(int), hipMemcpyDeviceToHost, s));
        XFE_CHECK(hipStreamSynchronize(s));
        *num_groups = groups;
        *levels_present = groups;
    }
    flag.ensure(1);
    XFE_CHECK(hipMemsetAsync(flag.p, 0, sizeof(int), s));
    hipLaunchKernelGGL(ix_descent_kernel, dim3(grid), dim3(kIxBlock), 0, s, d_gid, n, flag.p);
    XFE_CHECK(hipGetLastError());
    int descent = 0;
    XFE_CHECK(hipMemcpyAsync(&descent, flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
    XFE_CHECK(hipStreamSynchronize(s));
    *contiguous = (descent == 0 && *num_groups > 0) ? 1 : 0;
}

}  // namespace
}  // namespace xfe

using namespace xfe;

extern "C" {

int xfe_build_indexer(const int* raw, int n, int* group_ids, int* num_groups, int* num_levels_present,
                      int* contiguous) {
    clear_error();
    if (n < 0 || !num_groups || !num_levels_present || !contiguous || (n > 0 && (!raw || !group_ids))) {
        set_error("xfe_build_indexer: invalid arguments");
        return 0;
    }
    try {
        (void)hipGetLastError();
        Stream st;
        Scoped<int> d_raw, d_gid;
        d_raw.ensure(n);
        d_gid.ensure(n);
        if (n > 0) XFE_CHECK(hipMemcpyAsync(d_raw.p, raw, sizeof(int) * n, hipMemcpyHostToDevice, st.s));
        build_indexer_dev(d_raw.p, n, d_gid.p, num_groups, num_levels_present, contiguous, st.s);
        if (n > 0) XFE_CHECK(hipMemcpyAsync(group_ids, d_gid.p, sizeof(int) * n, hipMemcpyDeviceToHost, st.s));
        XFE_CHECK(hipStreamSynchronize(st.s));
        return 1;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 0;
    }
}

int xfe_build_indexer_device(const int* d_raw, int n, int* d_group_ids, int* num_groups, int* num_levels_present,
                             int* contiguous) {
    clear_error();
    if (n < 0 || !num_groups || !num_levels_present || !contiguous || (n > 0 && (!d_raw || !d_group_ids))) {
        set_error("xfe_build_indexer_device: invalid arguments");
        return 0;
    }
    try {
        (void)hipGetLastError();
        XFE_CHECK(hipDeviceSynchronize());  // the caller's pending writes of d_raw
        Stream st;
        build_indexer_dev(d_raw, n, d_group_ids, num_groups, num_levels_present, contiguous, st.s);
        XFE_CHECK(hipStreamSynchronize(st.s));
        return 1;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 0;
    }
}

int xfe_weight_sums(const int* group_ids, int n, int num_groups, const double* weights, double* out) {
    clear_error();
    if (n < 0 || num_groups <= 0 || !out || (n > 0 && !group_ids)) {
        set_error("xfe_weight_sums: invalid arguments");
        return 0;
    }
    try {
        (void)hipGetLastError();
        Stream st;
        Scoped<int> d_gid, flag;
        Scoped<double> d_w, d_out;
        d_gid.ensure(n);
        d_out.ensure(num_groups);
        flag.ensure(1);
        if (weights) d_w.ensure(n);
        if (n > 0) XFE_CHECK(hipMemcpyAsync(d_gid.p, group_ids, sizeof(int) * n, hipMemcpyHostToDevice, st.s));
        if (weights && n > 0)
            XFE_CHECK(hipMemcpyAsync(d_w.p, weights, sizeof(double) * n, hipMemcpyHostToDevice, st.s));
        XFE_CHECK(hipMemsetAsync(d_out.p, 0, sizeof(double) * num_groups, st.s));
        XFE_CHECK(hipMemsetAsync(flag.p, 0, sizeof(int), st.s));
        if (n > 0) {
            const int grid = ix_grid(n);
            hipLaunchKernelGGL(ix_range_check_kernel, dim3(grid), dim3(kIxBlock), 0, st.s, d_gid.p, n, num_groups,
                               flag.p);
            int bad = 0;
            XFE_CHECK(hipMemcpyAsync(&bad, flag.p, sizeof(int), hipMemcpyDeviceToHost, st.s));
            XFE_CHECK(hipStreamSynchronize(st.s));
            if (bad) throw Error("xfe_weight_sums: group id outside [0, num_groups)");
            hipLaunchKernelGGL(ix_wsum_kernel, dim3(grid), dim3(kIxBlock), 0, st.s, d_gid.p,
                               weights ? d_w.p : nullptr, n, d_out.p);
            XFE_CHECK(hipGetLastError());
        }
        XFE_CHECK(hipMemcpyAsync(out, d_out.p, sizeof(double) * num_groups, hipMemcpyDeviceToHost, st.s));
        XFE_CHECK(hipStreamSynchronize(st.s));
        return 1;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 0;
    }
}

}  // extern "C"
