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
// marks used slots in an int table and ranks them by an exclusive scan; the
// sparse path radix-sorts (id, row) pairs, flags run heads and scans them.
//
// xfe_build_indexer_device_noalloc runs the same kernels with their scratch in
// a file-scope device array (scratch, below) instead of device allocations.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_reduce.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <climits>
#include <cstring>
#include <mutex>
#include <string>

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

__global__ void ix_fill_kernel(int* v, long long m, int value) {
    IX_FOR(i, m) v[i] = value;
}

// Scratch of one indexer call: the min/max pair, a flag, five int arrays of n
// (the dense path uses the first four as two tables of up to 2n slots) and the
// rocPRIM temporary storage.
struct IxScratch {
    int* mm = nullptr;
    int* flag = nullptr;
    int* a[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    void* tmp = nullptr;
    size_t tmp_bytes = 0;
};

// rocPRIM temporary bytes for n rows (host-side size queries, no allocation)
size_t ix_tmp_bytes(int n, hipStream_t s) {
    size_t a = 0, b = 0, c = 0;
    const size_t m = static_cast<size_t>(std::max(n, 1));
    XFE_CHECK(rocprim::exclusive_scan(nullptr, a, static_cast<int*>(nullptr), static_cast<int*>(nullptr), 0,
                                      2 * m, rocprim::plus<int>(), s));
    XFE_CHECK(rocprim::radix_sort_pairs(nullptr, b, static_cast<const int*>(nullptr), static_cast<int*>(nullptr),
                                        static_cast<const int*>(nullptr), static_cast<int*>(nullptr), m, 0, 32, s));
    XFE_CHECK(rocprim::inclusive_scan(nullptr, c, static_cast<int*>(nullptr), static_cast<int*>(nullptr), m,
                                      rocprim::plus<int>(), s));
    return std::max(a, std::max(b, c));
}

// Bytes of scratch an indexer call over n rows needs (256-byte aligned pieces).
size_t ix_scratch_bytes(int n, hipStream_t s) {
    const size_t arr = (static_cast<size_t>(std::max(n, 1)) * sizeof(int) + 255) / 256 * 256;
    return 256 + 5 * arr + (ix_tmp_bytes(n, s) + 255) / 256 * 256;
}

IxScratch ix_carve(unsigned char* base, int n, hipStream_t s) {
    IxScratch x;
    const size_t arr = (static_cast<size_t>(std::max(n, 1)) * sizeof(int) + 255) / 256 * 256;
    x.mm = reinterpret_cast<int*>(base);
    x.flag = x.mm + 2;
    unsigned char* p = base + 256;
    for (int k = 0; k < 5; ++k, p += arr) x.a[k] = reinterpret_cast<int*>(p);
    x.tmp = p;
    x.tmp_bytes = ix_tmp_bytes(n, s);
    return x;
}

// Device indexer on device buffers with the given scratch; returns the level
// counts and the contiguity flag through host pointers.  Clears what it needs
// with fill kernels (no memset nodes).
void build_indexer_dev(const int* d_raw, int n, int* d_gid, int* num_groups, int* levels_present, int* contiguous,
                       const IxScratch& x, hipStream_t s) {
    *num_groups = 0;
    *levels_present = 0;
    *contiguous = 0;
    if (n <= 0) return;
    const int grid = ix_grid(n);
    hipLaunchKernelGGL(ix_init_minmax_kernel, dim3(1), dim3(64), 0, s, x.mm);
    hipLaunchKernelGGL(ix_minmax_kernel, dim3(grid), dim3(kIxBlock), 0, s, d_raw, n, x.mm);
    XFE_CHECK(hipGetLastError());
    int hmm[2];
    XFE_CHECK(hipMemcpyAsync(hmm, x.mm, sizeof(hmm), hipMemcpyDeviceToHost, s));
    XFE_CHECK(hipStreamSynchronize(s));
    const long long range = static_cast<long long>(hmm[1]) - hmm[0] + 1LL;
    size_t tmp_bytes = x.tmp_bytes;
    if (range > 0 && range <= kDenseRangeCap && range <= 2LL * n) {
        const int groups = static_cast<int>(range);
        int* used = x.a[0];  // 2n slots over a[0], a[1]
        int* rank = x.a[2];  // 2n slots over a[2], a[3]
        hipLaunchKernelGGL(ix_fill_kernel, dim3(ix_grid(groups)), dim3(kIxBlock), 0, s, used,
                           static_cast<long long>(groups), 0);
        hipLaunchKernelGGL(ix_dense_kernel, dim3(grid), dim3(kIxBlock), 0, s, d_raw, n, hmm[0], d_gid, used);
        XFE_CHECK(hipGetLastError());
        // rank of every used slot (exclusive scan); unique = last rank + last flag
        XFE_CHECK(rocprim::exclusive_scan(x.tmp, tmp_bytes, used, rank, 0, static_cast<size_t>(groups),
                                          rocprim::plus<int>(), s));
        int last[2] = {0, 0};
        XFE_CHECK(hipMemcpyAsync(&last[0], rank + (groups - 1), sizeof(int), hipMemcpyDeviceToHost, s));
        XFE_CHECK(hipMemcpyAsync(&last[1], used + (groups - 1), sizeof(int), hipMemcpyDeviceToHost, s));
        XFE_CHECK(hipStreamSynchronize(s));
        const int unique = last[0] + last[1];
        *num_groups = groups;
        *levels_present = unique;
        // fewer than 90 % of the slots used: rank them (src/fe_absorption.cpp:887-902)
        if (static_cast<long long>(unique) * 10LL < range * 9LL) {
            hipLaunchKernelGGL(ix_remap_kernel, dim3(grid), dim3(kIxBlock), 0, s, d_gid, n, rank);
            XFE_CHECK(hipGetLastError());
            *num_groups = unique;
            *levels_present = unique;
        }
    } else {
        int *iota = x.a[0], *skey = x.a[1], *pos = x.a[2], *head = x.a[3], *inc = x.a[4];
        hipLaunchKernelGGL(ix_iota_kernel, dim3(grid), dim3(kIxBlock), 0, s, iota, n);
        XFE_CHECK(rocprim::radix_sort_pairs(x.tmp, tmp_bytes, d_raw, skey, iota, pos, static_cast<size_t>(n), 0, 32,
                                            s));
        hipLaunchKernelGGL(ix_head_kernel, dim3(grid), dim3(kIxBlock), 0, s, skey, n, head);
        tmp_bytes = x.tmp_bytes;
        XFE_CHECK(rocprim::inclusive_scan(x.tmp, tmp_bytes, head, inc, static_cast<size_t>(n), rocprim::plus<int>(),
                                          s));
        hipLaunchKernelGGL(ix_scatter_kernel, dim3(grid), dim3(kIxBlock), 0, s, pos, inc, n, d_gid);
        XFE_CHECK(hipGetLastError());
        int groups = 0;
        XFE_CHECK(hipMemcpyAsync(&groups, inc + (n - 1), sizeof(int), hipMemcpyDeviceToHost, s));
        XFE_CHECK(hipStreamSynchronize(s));
        *num_groups = groups;
        *levels_present = groups;
    }
    hipLaunchKernelGGL(ix_fill_kernel, dim3(1), dim3(64), 0, s, x.flag, 1LL, 0);
    hipLaunchKernelGGL(ix_descent_kernel, dim3(grid), dim3(kIxBlock), 0, s, d_gid, n, x.flag);
    XFE_CHECK(hipGetLastError());
    int descent = 0;
    XFE_CHECK(hipMemcpyAsync(&descent, x.flag, sizeof(int), hipMemcpyDeviceToHost, s));
    XFE_CHECK(hipStreamSynchronize(s));
    *contiguous = (descent == 0 && *num_groups > 0) ? 1 : 0;
}

// The allocating form: scratch from the device allocator for this call.
void build_indexer_alloc(const int* d_raw, int n, int* d_gid, int* num_groups, int* levels_present, int* contiguous,
                         hipStream_t s) {
    Scoped<unsigned char> mem;
    mem.ensure(ix_scratch_bytes(n, s));
    build_indexer_dev(d_raw, n, d_gid, num_groups, levels_present, contiguous, ix_carve(mem.p, n, s), s);
    XFE_CHECK(hipStreamSynchronize(s));  // before the scratch is freed
}

// ---- the allocation-free variant ----
// Maximum rows of one xfe_build_indexer_device_noalloc call, and the bytes of
// its scratch: 256 + 5 x 4 MiB of int tables plus rocPRIM's temporary storage
// for 1M keys, which the first call checks against kNoAllocScratchBytes.
constexpr int kNoAllocMaxRows = 1 << 20;
constexpr size_t kNoAllocScratchBytes = size_t(40) << 20;

// One copy per device, held until the process ends.  Host code uses the
// address hipGetSymbolAddress returns for the current device (noalloc_scratch()),
// never this name.
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// Address of the current device's scratch, looked up once per device.
unsigned char* noalloc_scratch() {
    int dev = 0;
    XFE_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw Error("xfe_build_indexer_device_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        XFE_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw Error("xfe_build_indexer_device_noalloc: no scratch address");
        g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
    }
    return g_scratch_ptr[dev];
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
        build_indexer_alloc(d_raw.p, n, d_gid.p, num_groups, num_levels_present, contiguous, st.s);
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
        build_indexer_alloc(d_raw, n, d_group_ids, num_groups, num_levels_present, contiguous, st.s);
        XFE_CHECK(hipStreamSynchronize(st.s));
        return 1;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 0;
    }
}

int xfe_build_indexer_device_noalloc(const int* d_raw, int n, int* d_group_ids, int* num_groups,
                                     int* num_levels_present, int* contiguous) {
    clear_error();
    if (n < 0 || !num_groups || !num_levels_present || !contiguous || (n > 0 && (!d_raw || !d_group_ids))) {
        set_error("xfe_build_indexer_device_noalloc: invalid arguments");
        return 0;
    }
    if (n > kNoAllocMaxRows) {
        set_error("xfe_build_indexer_device_noalloc: " + std::to_string(n) + " rows exceed the maximum of " +
                  std::to_string(kNoAllocMaxRows));
        return 0;
    }
    try {
        (void)hipGetLastError();
        unsigned char* base = noalloc_scratch();
        hipStream_t s = nullptr;  // the null stream: no stream object is created
        if (ix_scratch_bytes(n, s) > kNoAllocScratchBytes)
            throw Error("xfe_build_indexer_device_noalloc: scratch too small for " + std::to_string(n) + " rows");
        XFE_CHECK(hipDeviceSynchronize());  // the caller's pending writes of d_raw
        build_indexer_dev(d_raw, n, d_group_ids, num_groups, num_levels_present, contiguous, ix_carve(base, n, s), s);
        XFE_CHECK(hipStreamSynchronize(s));
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
