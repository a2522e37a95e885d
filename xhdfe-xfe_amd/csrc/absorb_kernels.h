// Launch interface of the absorber kernels (absorb_kernels.hip).
#ifndef XFE_ABSORB_KERNELS_H
#define XFE_ABSORB_KERNELS_H

#include <hip/hip_runtime.h>

namespace xfe {

constexpr int kBlock = 256;
constexpr int kWave = 64;

// how a pass accumulates level sums of the dimension it starts (0: no reduce)
enum { kRedNone = 0, kRedLds = 1, kRedSorted = 2, kRedGlobal = 3, kRedSpell = 4 };
constexpr int kTileRows = 512;  // rows per level-aligned tile (tile_kernels.hip)
constexpr int kXcdCopies = 8;  // XCD-private accumulator copies (MI355X: 8 XCDs)

// Arguments of the fused streaming pass (see pass_kernel).
struct PassArgs {
    int n = 0;
    int ld = 0;
    int C = 0;   // total columns (stride of the level-major tables)
    int c0 = 0;  // first column handled by this launch
    const double* src = nullptr;
    double* dst = nullptr;
    const double* P = nullptr;  // optional IT extrapolation v = P - coef (P - Q)
    const double* Q = nullptr;
    double coef = 0.0;
    double* xout = nullptr;
    const int* gid_a = nullptr;  // projection being finished
    const double* means_a = nullptr;
    const int* gid_r = nullptr;  // projection being started
    double* sums_r = nullptr;
    int G_r = 0;
    const double* w = nullptr;
    const double* it_x = nullptr;
    const double* it_gx = nullptr;
    double* partials = nullptr;
    int one = 0;           // values are 1 (weight sums)
    int xcd_copies = 1;    // kRedGlobal: private accumulator copies
    size_t copy_stride = 0;
};

// Arguments of the tiled pass (tile_kernels.hip).
struct TileArgs {
    int n = 0;
    int ld = 0;
    int C = 0;
    int c0 = 0;
    const double* src = nullptr;
    double* dst = nullptr;
    const double* P = nullptr;
    const double* Q = nullptr;
    double coef = 0.0;
    double* xout = nullptr;
    const int* gid_a = nullptr;
    const double* means_a = nullptr;
    const int* tile_start = nullptr;  // num_tiles + 1 row offsets
    int num_tiles = 0;
    const int* gid_p = nullptr;       // the level-sorted dimension
    const double* inv_p = nullptr;
    const int* gid_r = nullptr;
    int G_r = 0;
    double* sums_r = nullptr;         // kRedLds flush target
    double* spell_sums = nullptr;     // kRedSpell output (spells x C)
    const int* tile_spell_base = nullptr;
    const double* w = nullptr;
    const double* it_x = nullptr;
    const double* it_gx = nullptr;
    double* partials = nullptr;
    double* maxp_partials = nullptr;  // per-block max |mean| of p (polish)
};

hipError_t launch_tile_pass(const TileArgs& a, int cc, int rr, bool pp, int nq, int grid,
                            hipStream_t s);
hipError_t launch_spell_pull(const double* S, int C, int c0, int cc, const int* order,
                             const int* skey, int M, double* sums, int grid, hipStream_t s);
hipError_t launch_level_count(const int* gid, int n, int* cnt, int grid, hipStream_t s);
hipError_t launch_max_int(const int* v, int m, int* out, int grid, hipStream_t s);
hipError_t launch_tile_start(const int* gid, int n, int step, int num_tiles, int* tile_start,
                             int grid, hipStream_t s);
hipError_t launch_spell_heads(const int* gp, const int* gr, int n, int* flags, int grid,
                              hipStream_t s);
hipError_t launch_spell_keys(const int* flags, const int* inc, const int* gr, int n, int* skey,
                             int grid, hipStream_t s);
hipError_t launch_tile_spell_base(const int* tile_start, int num_tiles, const int* inc, int n,
                                  int* base, int grid, hipStream_t s);

hipError_t launch_pass(const PassArgs& a, int cc, int rmode, int nq, int grid, hipStream_t s);
hipError_t launch_finalize(double* sums, double* means, const double* inv, int G, int C,
                           int copies, double* partials, int grid, hipStream_t s);
hipError_t launch_reduce_partials(const double* partials, int count, int nq, int mode, double* out,
                                  hipStream_t s);
hipError_t launch_it_stats(const double* x, const double* gx, const double* ggx, int n, int C,
                           int ld, double* partials, int grid, hipStream_t s);
hipError_t launch_col_dot(const double* a, const double* b, int n, int C, int ld, double* partials,
                          int grid, hipStream_t s);
hipError_t launch_col_axpy(double* y, const double* a, const double* b, const double* alpha, int n,
                           int C, int ld, double sign, int grid, hipStream_t s);
hipError_t launch_sub(double* y, const double* a, const double* b, int n, int C, int ld, int grid,
                      hipStream_t s);
hipError_t launch_inverse(const double* ws, double* inv, int G, int grid, hipStream_t s);
hipError_t launch_iota(int* v, int n, int grid, hipStream_t s);
hipError_t launch_gather_ids(const int* src, const int* perm, int* dst, int n, int grid,
                             hipStream_t s);
hipError_t launch_gather_cols(const double* src, int ld_src, const int* perm, double* dst,
                              int ld_dst, int n, int C, int grid, hipStream_t s);
hipError_t launch_scatter_cols(const double* src, int ld_src, const int* perm, double* dst,
                               int ld_dst, int n, int C, int grid, hipStream_t s);
hipError_t launch_gather_vec(const double* src, const int* perm, double* dst, int n, int grid,
                             hipStream_t s);
hipError_t launch_sorted_check(const int* gid, int n, int* flag, int grid, hipStream_t s);

}  // namespace xfe

#endif
