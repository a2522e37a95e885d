// Forwarder for xhdfe::schwarz_demean_raw (include/schwarz_demean.hpp:30) to
// the device's xfe_schwarz_demean.  It replaces the definition in
// src/schwarz_demean.cpp:419-438 (that file is then left out of the build, or
// its definition renamed), so the caller (src/fe_absorption.cpp:5185-5215) is
// unchanged.  Same flag, same stopping rule (checked on the result).
#include <cstdint>
#include <vector>

#include "schwarz_demean.hpp"
#include "xfe_hip.h"

namespace xhdfe {

bool schwarz_demean_raw(double* Y, double* X, int64_t n, int k, const std::vector<const int32_t*>& fe,
                        const std::vector<int>& nlev, int threads, double tol, int max_iter, int* iters_out,
                        int collapse) {
    if (fe.size() != nlev.size()) return false;
    return xfe_schwarz_demean(Y, X, n, k, fe.data(), nlev.data(), static_cast<int>(fe.size()), threads, tol, max_iter,
                              iters_out, collapse) != 0;
}

}  // namespace xhdfe
