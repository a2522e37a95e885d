// Drop-in for the reference's src/akm_kss_cuda.cu: the hdfe::akm functions of
// include/hdfe/akm_kss_cuda.hpp over the C ABI of include/xfe_hip.h.  The
// callers in src/akm_kss.cpp (:536, :594, :933, :1154, :1323, :1659-1707) are
// unchanged; a nullptr context or a false return sends them to their CPU
// solver, as before.  Links libxfe_hip.so.
#include <cstddef>
#include <cstdint>

#include "hdfe/akm_kss_cuda.hpp"
#include "xfe_hip.h"

namespace hdfe {
namespace akm {

struct AkmCudaContext {
    xfe_akm_context* c;
};

bool akm_cuda_available() { return xfe_akm_available() != 0; }

bool akm_cuda_stable_sort_u64(const std::uint64_t* keys, std::int64_t* order_out, std::size_t n) {
    return xfe_akm_stable_sort_u64(keys, order_out, n) != 0;
}

AkmCudaContext* akm_cuda_create(int n_workers, int n_firms, const int* m_w, const int* m_f, const double* m_c,
                                std::size_t n_matches, const std::int64_t* w_ptr, const std::int64_t* f_ptr,
                                const std::int64_t* f_matches, const double* Dw, const double* Df,
                                const double* inv_diagS) {
    xfe_akm_context* c =
        xfe_akm_create(n_workers, n_firms, m_w, m_f, m_c, n_matches, w_ptr, f_ptr, f_matches, Dw, Df, inv_diagS);
    return c ? new AkmCudaContext{c} : nullptr;
}

void akm_cuda_destroy(AkmCudaContext* ctx) {
    if (!ctx) return;
    xfe_akm_destroy(ctx->c);
    delete ctx;
}

int akm_cuda_solve_S(AkmCudaContext* ctx, const double* rhs, double* z, double tol, int max_iter) {
    return ctx ? xfe_akm_solve_S(ctx->c, rhs, z, tol, max_iter) : -1;
}

int akm_cuda_solve_S_multi(AkmCudaContext* ctx, const double* rhs_pack, double* z_pack, int nb, double tol,
                           int max_iter, int* iters_out) {
    return ctx ? xfe_akm_solve_S_multi(ctx->c, rhs_pack, z_pack, nb, tol, max_iter, iters_out) : -1;
}

int akm_cuda_max_lanes() { return xfe_akm_max_lanes(); }

double* akm_cuda_pack_buffer(AkmCudaContext* ctx, int nb) { return ctx ? xfe_akm_pack_buffer(ctx->c, nb) : nullptr; }

}  // namespace akm
}  // namespace hdfe
