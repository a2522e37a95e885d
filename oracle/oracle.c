/* CPU oracle for the fixed-effects absorber hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C restatement of the reference CPU algorithms of xhdfe-xfe.  It is the
 * checker for the HIP product: only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  The product never links it and
 * never falls back to it.
 *
 * What pins each function (the checks live in tests/test_oracle_*.py):
 *   or_build_indexer    -- reference checker (oracle/_ref) level counts + brute force
 *                          (sorted unique raw ids) on random inputs.
 *   or_weight_sums      -- brute force (np.bincount).
 *   or_project          -- closed form: exact group demeaning (one-way FE).
 *   or_absorb           -- (i) the reference checker's within transform on the same
 *                          inputs (C0 panel, 2- and 3-way, GS and SGS), (ii) the sparse
 *                          direct least-squares residual (scipy, grounded normal
 *                          equations) on small panels, (iii) invariants D'v~ = 0 and
 *                          idempotence.
 *   or_drop_singletons  -- reference checker (kept sample) + brute force repeated
 *                          drop on tiny inputs.
 *   or_components       -- reference checker groupvar() labels + BFS with rank-by-min
 *                          node labels on tiny graphs.
 *   or_dof_reghdfe      -- reference checker df_a / fe_redundant + dense rank of the
 *                          dummy matrix on tiny cases.
 * Iteration counts are trajectory details: they are compared with the reference
 * checker where the path is the same (C0) but are not a parity claim.
 */
#include "oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OR_MAX_DIMS 8

/* ------------------------------------------------------------------ indexer */

static int cmp_pair(const void* a, const void* b) {
    const int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
    return (x > y) - (x < y);
}

/* build_indexer, src/fe_absorption.cpp:845-940 (dense path :873-904, sorted
 * rank path :914-934) and finalize_contiguous_groups :576-620. */
int or_build_indexer(int n, const int* raw, int* gid, int* num_groups, int* levels_present,
                     int* contiguous) {
    *num_groups = 0;
    *levels_present = 0;
    *contiguous = 0;
    if (n <= 0) return 0;
    int mn = raw[0], mx = raw[0];
    for (int i = 1; i < n; ++i) {
        if (raw[i] < mn) mn = raw[i];
        else if (raw[i] > mx) mx = raw[i];
    }
    const long long range = (long long)mx - (long long)mn + 1LL;
    if (range > 0 && range <= 50000000LL && range <= 2LL * n) {
        const int groups = (int)range;
        unsigned char* seen = (unsigned char*)calloc((size_t)groups, 1);
        int unique = 0;
        for (int i = 0; i < n; ++i) {
            const int m = raw[i] - mn;
            gid[i] = m;
            if (!seen[m]) { seen[m] = 1; ++unique; }
        }
        *num_groups = groups;
        *levels_present = unique;
        if ((long long)unique * 10LL < range * 9LL) {
            int* remap = (int*)malloc(sizeof(int) * (size_t)groups);
            int next = 0;
            for (int g = 0; g < groups; ++g) remap[g] = seen[g] ? next++ : -1;
            for (int i = 0; i < n; ++i) gid[i] = remap[gid[i]];
            *num_groups = next;
            *levels_present = next;
            free(remap);
        }
        free(seen);
    } else {
        int64_t* keyed = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
        for (int i = 0; i < n; ++i)
            keyed[i] = ((int64_t)raw[i] << 32) | (int64_t)(uint32_t)i;
        qsort(keyed, (size_t)n, sizeof(int64_t), cmp_pair);
        int next = -1;
        int prev = 0;
        for (int k = 0; k < n; ++k) {
            const int key = (int)(keyed[k] >> 32);
            const int pos = (int)(keyed[k] & 0xffffffffLL);
            if (next < 0 || key != prev) { ++next; prev = key; }
            gid[pos] = next;
        }
        *num_groups = next + 1;
        *levels_present = next + 1;
        free(keyed);
    }
    int sorted = 1;
    for (int i = 1; i < n && sorted; ++i) sorted = gid[i] >= gid[i - 1];
    *contiguous = sorted;
    return 0;
}

/* compute_weight_sums, src/fe_absorption.cpp:942-1037 (counts for unit weights). */
void or_weight_sums(int n, const int* gid, int G, const double* w, double* out) {
    memset(out, 0, sizeof(double) * (size_t)G);
    for (int i = 0; i < n; ++i) out[gid[i]] += w ? w[i] : 1.0;
}

/* ------------------------------------------------------------- projections */

#define OR_OBSPART_MAX_GROUPS 2048 /* kObsPartMaxGroups, src/fe_absorption.cpp:6377 */

typedef struct {
    int n, S, D;
    const int* gid[OR_MAX_DIMS];
    int G[OR_MAX_DIMS];
    double* wsum[OR_MAX_DIMS];
    double* wsinv[OR_MAX_DIMS];
    double* means[OR_MAX_DIMS]; /* G x S, level-major */
    int* ptr[OR_MAX_DIMS];      /* CSR: rows of each level in ascending row order */
    int* idx[OR_MAX_DIMS];
    /* per-dimension strategy of the packed path (src/fe_absorption.cpp:6373-6400):
     * 0 contiguous (the rows of each level are consecutive), 1 CSR gather,
     * 2 observation-partitioned (few levels, not contiguous): per-chunk sums */
    int mode[OR_MAX_DIMS];
    double* tls[OR_MAX_DIMS];   /* chunks x G x S partial sums (mode 2) */
    int chunks;
    const double* w;
    int threads;
} or_ws;

static void ws_init(or_ws* ws, int n, int S, int D, const int* gids, const int* G,
                    const double* w, int threads) {
    memset(ws, 0, sizeof(*ws));
    ws->n = n; ws->S = S; ws->D = D; ws->w = w;
    ws->threads = threads > 0 ? threads : 1;
    ws->chunks = 4 * ws->threads;
    for (int d = 0; d < D; ++d) {
        const int* g = gids + (size_t)d * n;
        const int Gd = G[d];
        ws->gid[d] = g;
        ws->G[d] = Gd;
        ws->wsum[d] = (double*)malloc(sizeof(double) * (size_t)(Gd > 0 ? Gd : 1));
        ws->wsinv[d] = (double*)malloc(sizeof(double) * (size_t)(Gd > 0 ? Gd : 1));
        ws->means[d] = (double*)calloc((size_t)(Gd > 0 ? Gd : 1) * (size_t)S, sizeof(double));
        or_weight_sums(n, g, Gd, w, ws->wsum[d]);
        /* FeWorkspace::refresh_weight_sums_inv (src/fe_absorption.cpp:662): 1/sum, 0 if sum <= 0 */
        for (int k = 0; k < Gd; ++k) ws->wsinv[d][k] = ws->wsum[d][k] > 0.0 ? 1.0 / ws->wsum[d][k] : 0.0;
        int* ptr = (int*)calloc((size_t)Gd + 1, sizeof(int));
        int* idx = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
        for (int i = 0; i < n; ++i) ptr[g[i] + 1]++;
        for (int k = 0; k < Gd; ++k) ptr[k + 1] += ptr[k];
        int* fill = (int*)malloc(sizeof(int) * (size_t)(Gd > 0 ? Gd : 1));
        for (int k = 0; k < Gd; ++k) fill[k] = ptr[k];
        for (int i = 0; i < n; ++i) idx[fill[g[i]]++] = i;
        free(fill);
        ws->ptr[d] = ptr;
        ws->idx[d] = idx;
        int contiguous = 1;
        for (int p = 0; p < n && contiguous; ++p) contiguous = idx[p] == p;
        ws->mode[d] = contiguous ? 0 : (Gd <= OR_OBSPART_MAX_GROUPS ? 2 : 1);
        if (ws->mode[d] == 2)
            ws->tls[d] = (double*)malloc(sizeof(double) * (size_t)ws->chunks * (size_t)(Gd > 0 ? Gd : 1) * (size_t)S);
    }
}

static void ws_free(or_ws* ws) {
    for (int d = 0; d < ws->D; ++d) {
        free(ws->wsum[d]); free(ws->wsinv[d]); free(ws->means[d]);
        free(ws->ptr[d]); free(ws->idx[d]); free(ws->tls[d]);
    }
}

/* One projection on the packed row-interleaved buffer cur[i*S + s]: per level
 * the (weighted) sum of its rows in row order, the mean, then the subtraction
 * (packed_demean_grouped, src/fe_absorption.cpp:5719-5800; the
 * observation-partitioned variant for few non-contiguous levels merges
 * per-chunk sums in chunk order).  S is a compile-time constant in the
 * specializations below.  Optionally returns the post-projection sum of
 * squares over all lanes. */
static inline __attribute__((always_inline)) void demean_dim_s(or_ws* ws, double* cur, int d,
                                                              double* out_sumsq, const int S) {
    const int Gd = ws->G[d];
    const int n = ws->n;
    const int* ptr = ws->ptr[d];
    const int* idx = ws->idx[d];
    const int* gid = ws->gid[d];
    const double* w = ws->w;
    const double* inv = ws->wsinv[d];
    double* means = ws->means[d];
    double sumsq = 0.0;
    if (ws->mode[d] == 2) {
        const int chunks = ws->chunks;
        double* tls = ws->tls[d];
#pragma omp parallel for schedule(static) num_threads(ws->threads)
        for (int c = 0; c < chunks; ++c) {
            double* loc = tls + (size_t)c * Gd * S;
            memset(loc, 0, sizeof(double) * (size_t)Gd * S);
            const int i0 = (int)(((long long)n * c) / chunks), i1 = (int)(((long long)n * (c + 1)) / chunks);
            for (int i = i0; i < i1; ++i) {
                double* a = loc + (size_t)gid[i] * S;
                const double* row = cur + (size_t)i * S;
                if (w) { const double wi = w[i]; for (int s = 0; s < S; ++s) a[s] += wi * row[s]; }
                else for (int s = 0; s < S; ++s) a[s] += row[s];
            }
        }
        for (int g = 0; g < Gd; ++g)
            for (int s = 0; s < S; ++s) {
                double t = 0.0;
                for (int c = 0; c < chunks; ++c) t += tls[((size_t)c * Gd + g) * S + s];
                means[(size_t)g * S + s] = t * inv[g];
            }
#pragma omp parallel for schedule(static) num_threads(ws->threads) reduction(+ : sumsq)
        for (int i = 0; i < n; ++i) {
            double* row = cur + (size_t)i * S;
            const double* m = means + (size_t)gid[i] * S;
            for (int s = 0; s < S; ++s) {
                row[s] -= m[s];
                sumsq += row[s] * row[s];
            }
        }
        if (out_sumsq) *out_sumsq = sumsq;
        return;
    }
    const int contiguous = ws->mode[d] == 0;
#pragma omp parallel for schedule(static) num_threads(ws->threads) reduction(+ : sumsq)
    for (int g = 0; g < Gd; ++g) {
        double acc[64];
        for (int s = 0; s < S; ++s) acc[s] = 0.0;
        for (int p = ptr[g]; p < ptr[g + 1]; ++p) {
            const int i = contiguous ? p : idx[p];
            if (!contiguous && p + 16 < n) __builtin_prefetch(cur + (size_t)idx[p + 16] * S, 0, 1);
            const double* row = cur + (size_t)i * S;
            if (w) { const double wi = w[i]; for (int s = 0; s < S; ++s) acc[s] += wi * row[s]; }
            else for (int s = 0; s < S; ++s) acc[s] += row[s];
        }
        double* m = means + (size_t)g * S;
        for (int s = 0; s < S; ++s) m[s] = acc[s] * inv[g];
        for (int p = ptr[g]; p < ptr[g + 1]; ++p) {
            double* row = cur + (size_t)(contiguous ? p : idx[p]) * S;
            for (int s = 0; s < S; ++s) {
                row[s] -= m[s];
                sumsq += row[s] * row[s];
            }
        }
    }
    if (out_sumsq) *out_sumsq = sumsq;
}

static void demean_dim(or_ws* ws, double* cur, int d, double* out_sumsq) {
    switch (ws->S) {
        case 1: demean_dim_s(ws, cur, d, out_sumsq, 1); break;
        case 2: demean_dim_s(ws, cur, d, out_sumsq, 2); break;
        case 3: demean_dim_s(ws, cur, d, out_sumsq, 3); break;
        case 4: demean_dim_s(ws, cur, d, out_sumsq, 4); break;
        case 5: demean_dim_s(ws, cur, d, out_sumsq, 5); break;
        case 6: demean_dim_s(ws, cur, d, out_sumsq, 6); break;
        case 7: demean_dim_s(ws, cur, d, out_sumsq, 7); break;
        case 8: demean_dim_s(ws, cur, d, out_sumsq, 8); break;
        default: demean_dim_s(ws, cur, d, out_sumsq, ws->S); break;
    }
}

/* packed_sweep, src/fe_absorption.cpp:6423-6443: forward over `order`, then
 * backward when symmetric; the sum of squares comes from the last projection. */
static void sweep(or_ws* ws, double* cur, const int* order, int len, int symmetric,
                  double* out_sumsq) {
    double sumsq = 0.0;
    for (int p = 0; p < len; ++p) {
        const int last = out_sumsq && !symmetric && p + 1 == len;
        demean_dim(ws, cur, order[p], last ? &sumsq : NULL);
    }
    if (symmetric) {
        for (int p = len; p-- > 0;) {
            const int last = out_sumsq && p == 0;
            demean_dim(ws, cur, order[p], last ? &sumsq : NULL);
        }
    }
    if (out_sumsq) *out_sumsq = sumsq;
}

static double sumsq_all(const double* v, size_t m) {
    double s = 0.0;
    for (size_t i = 0; i < m; ++i) s += v[i] * v[i];
    return s;
}

/* ------------------------------------------------------------ IT accelerator */

typedef struct { double vprod, ssq, d1sq; } it_stats;

/* irons_tuck_accumulate, src/fe_absorption.cpp:5242-5275, over fixed chunks
 * whose partial sums are added in chunk order (the reference's deterministic
 * chunking), so the result does not depend on the thread count. */
#define OR_VEC_CHUNKS 64
static it_stats it_accumulate_t(const double* x, const double* gx, const double* ggx, size_t m, int threads) {
    double pv[OR_VEC_CHUNKS], ps[OR_VEC_CHUNKS], pd[OR_VEC_CHUNKS];
#pragma omp parallel for schedule(static) num_threads(threads)
    for (int c = 0; c < OR_VEC_CHUNKS; ++c) {
        const size_t i0 = m * (size_t)c / OR_VEC_CHUNKS, i1 = m * (size_t)(c + 1) / OR_VEC_CHUNKS;
        double v = 0.0, q = 0.0, d = 0.0;
        for (size_t i = i0; i < i1; ++i) {
            const double dg = ggx[i] - gx[i];
            const double d2 = dg - gx[i] + x[i];
            v += dg * d2;
            q += d2 * d2;
            d += dg * dg;
        }
        pv[c] = v; ps[c] = q; pd[c] = d;
    }
    it_stats st = {0.0, 0.0, 0.0};
    for (int c = 0; c < OR_VEC_CHUNKS; ++c) { st.vprod += pv[c]; st.ssq += ps[c]; st.d1sq += pd[c]; }
    return st;
}

static void par_copy(double* dst, const double* src, size_t m, int threads) {
#pragma omp parallel for schedule(static) num_threads(threads)
    for (int c = 0; c < OR_VEC_CHUNKS; ++c) {
        const size_t i0 = m * (size_t)c / OR_VEC_CHUNKS, i1 = m * (size_t)(c + 1) / OR_VEC_CHUNKS;
        memcpy(dst + i0, src + i0, sizeof(double) * (i1 - i0));
    }
}

static double sumsq_t(const double* v, size_t m, int threads) {
    double part[OR_VEC_CHUNKS];
#pragma omp parallel for schedule(static) num_threads(threads)
    for (int c = 0; c < OR_VEC_CHUNKS; ++c) {
        const size_t i0 = m * (size_t)c / OR_VEC_CHUNKS, i1 = m * (size_t)(c + 1) / OR_VEC_CHUNKS;
        double t = 0.0;
        for (size_t i = i0; i < i1; ++i) t += v[i] * v[i];
        part[c] = t;
    }
    double s = 0.0;
    for (int c = 0; c < OR_VEC_CHUNKS; ++c) s += part[c];
    return s;
}

typedef struct {
    or_ws* ws;
    int symmetric;
    int honest;
    double honest_tol;
    double conv_tol;
    size_t m;
    double *cur, *gx, *ggx, *acc_a, *acc_b, *acc_c;
    int diverged;
} accel_ctx;

/* packed_accel_phase, src/fe_absorption.cpp:6599-6782 (Irons-Tuck with grand
 * acceleration every 4 iterations, divergence safeguard, CG hand-off). */
static int accel_phase(accel_ctx* c, const int* order, int len, int max_iter, int* iter_used) {
    const size_t m = c->m;
    *iter_used = 0;
    if (max_iter <= 0 || len == 0) return 0;
    const double kDiverge = 4.0, kBail = 1.5;
    const int cg_iter_cap = 1000, grand_every = 4;
    int last_check = -1;
    const int th = c->ws->threads;
    double prev_norm = sqrt(sumsq_t(c->cur, m, th));
    int grand = 0;
    double best = prev_norm;
    int suspended = 0;
    par_copy(c->gx, c->cur, m, th);
    sweep(c->ws, c->gx, order, len, c->symmetric, NULL);
    for (int iter = 0; iter < max_iter; ++iter) {
        par_copy(c->ggx, c->gx, m, th);
        sweep(c->ws, c->ggx, order, len, c->symmetric, NULL);
        const it_stats st = it_accumulate_t(c->cur, c->gx, c->ggx, m, th);
        if (c->honest && sqrt(st.d1sq) <= c->honest_tol * fmax(1.0, prev_norm)) {
            par_copy(c->cur, c->ggx, m, th);
            *iter_used = iter + 1;
            return 1;
        }
        if (st.ssq == 0.0) {
            par_copy(c->cur, c->gx, m, th);
            *iter_used = iter + 1;
            return 1;
        }
        double coef = st.vprod / st.ssq;
        if (suspended) coef = 0.0;
        /* irons_tuck_update_broadcast (:5366-5394) */
#pragma omp parallel for schedule(static) num_threads(th)
        for (size_t i = 0; i < m; ++i) {
            const double nx = c->ggx[i] - coef * (c->ggx[i] - c->gx[i]);
            c->cur[i] = nx;
            c->gx[i] = nx;
        }
        double sumsq = 0.0;
        sweep(c->ws, c->gx, order, len, c->symmetric, &sumsq);
        {
            const double curr = sqrt(sumsq);
            const double denom = fmax(1.0, prev_norm);
            const double rel_change = fabs(curr - prev_norm) / denom;
            const int step = iter - last_check;
            last_check = iter;
            prev_norm = curr;
            if (curr <= best) {
                best = curr;
                suspended = 0;
            } else {
                if (curr > kBail * best) {
                    c->diverged = 1;
                    par_copy(c->cur, c->gx, m, th);
                    *iter_used = iter + 1;
                    return 0;
                }
                if (curr > kDiverge * best) {
                    if (!suspended) grand = 0;
                    suspended = 1;
                }
            }
            if (!c->honest && step > 0 && rel_change / (double)step < c->conv_tol) {
                par_copy(c->cur, c->gx, m, th);
                *iter_used = iter + 1;
                return 1;
            }
        }
        if (c->honest && cg_iter_cap > 0 && !c->diverged && iter + 1 >= cg_iter_cap) {
            c->diverged = 1;
            par_copy(c->cur, c->gx, m, th);
            *iter_used = iter + 1;
            return 0;
        }
        if (!suspended && grand_every > 0 && (iter + 1) % grand_every == 0) {
            ++grand;
            if (grand == 1) {
                par_copy(c->acc_a, c->gx, m, th);
            } else if (grand == 2) {
                par_copy(c->acc_b, c->gx, m, th);
            } else {
                par_copy(c->acc_c, c->gx, m, th);
                const it_stats as = it_accumulate_t(c->acc_a, c->acc_b, c->acc_c, m, th);
                if (as.ssq != 0.0) {
                    const double ac = as.vprod / as.ssq;
                    /* irons_tuck_update (:5277-5298) */
#pragma omp parallel for schedule(static) num_threads(th)
                    for (size_t i = 0; i < m; ++i)
                        c->acc_a[i] = c->acc_c[i] - ac * (c->acc_c[i] - c->acc_b[i]);
                }
                par_copy(c->cur, c->acc_a, m, th);
                par_copy(c->gx, c->cur, m, th);
                sweep(c->ws, c->gx, order, len, c->symmetric, NULL);
                grand = 0;
            }
        }
    }
    par_copy(c->cur, c->gx, m, th);
    *iter_used = max_iter;
    return 0;
}

/* ------------------------------------------------------------------- CG */

static int lane_exact_zero(const double* v, int n, int S, int k) {
    for (int i = 0; i < n; ++i)
        if (v[(size_t)i * S + k] != 0.0) return 0;
    return 1;
}

/* run_packed_cg, src/fe_absorption.cpp:6790-6973: per-lane CG on the
 * symmetric operator (I - G) with G the symmetric sweep; stops when
 * max_k sqrt(alpha_k ssr_k / improvement_k) <= tol.  Returns converged. */
static int run_cg(or_ws* ws, double* cur, const int* order, int len, int max_iter, double tol,
                  int* iters_out) {
    const int n = ws->n, S = ws->S;
    const size_t m = (size_t)n * S;
    double* r = (double*)malloc(sizeof(double) * m);
    double* u = (double*)malloc(sizeof(double) * m);
    double* v = (double*)malloc(sizeof(double) * m);
    double improvement[64], ssr[64], ssr_old[64], alpha[64], beta[64], recent[64], denom[64],
        next_imp[64];
    unsigned char solved[64];
    memset(solved, 0, sizeof(solved));
#define COLSUMSQ(out, a)                                   \
    do {                                                   \
        for (int k_ = 0; k_ < S; ++k_) (out)[k_] = 0.0;    \
        for (size_t i_ = 0; i_ < m; ++i_) (out)[i_ % S] += (a)[i_] * (a)[i_]; \
    } while (0)
#define PROJ(src, dst)                                                    \
    do {                                                                  \
        memcpy((dst), (src), sizeof(double) * m);                         \
        sweep(ws, (dst), order, len, 1, NULL);                            \
        for (size_t i_ = 0; i_ < m; ++i_) (dst)[i_] = (src)[i_] - (dst)[i_]; \
    } while (0)
    COLSUMSQ(improvement, cur);
    PROJ(cur, r);
    COLSUMSQ(ssr, r);
    memcpy(u, r, sizeof(double) * m);
    for (int k = 0; k < S; ++k) recent[k] = 0.0;
    int converged = 0, iters = 0;
    for (int iter = 1; iter <= max_iter; ++iter) {
        PROJ(u, v);
        for (int k = 0; k < S; ++k) denom[k] = 0.0;
        for (size_t i = 0; i < m; ++i) denom[i % S] += u[i] * v[i];
        int breakdown = 0;
        for (int k = 0; k < S; ++k) { alpha[k] = 0.0; next_imp[k] = improvement[k]; }
        for (int k = 0; k < S && !breakdown; ++k) {
            if (!isfinite(improvement[k]) || !isfinite(ssr[k]) || ssr[k] < 0.0) { breakdown = 1; break; }
            if (solved[k]) { if (ssr[k] != 0.0) breakdown = 1; recent[k] = 0.0; continue; }
            if (ssr[k] == 0.0) {
                if (!lane_exact_zero(r, n, S, k)) { breakdown = 1; break; }
                solved[k] = 1; recent[k] = 0.0; continue;
            }
            if (!isfinite(denom[k]) || denom[k] <= 0.0) { breakdown = 1; break; }
            alpha[k] = ssr[k] / denom[k];
            recent[k] = alpha[k] * ssr[k];
            next_imp[k] = improvement[k] - recent[k];
            if (!isfinite(alpha[k]) || alpha[k] <= 0.0 || !isfinite(recent[k]) || recent[k] <= 0.0 ||
                !isfinite(next_imp[k])) { breakdown = 1; break; }
        }
        iters = iter;
        if (breakdown) break;
        for (int k = 0; k < S; ++k) improvement[k] = next_imp[k];
        for (size_t i = 0; i < m; ++i) {
            cur[i] -= alpha[i % S] * u[i];
            r[i] -= alpha[i % S] * v[i];
        }
        memcpy(ssr_old, ssr, sizeof(double) * S);
        COLSUMSQ(ssr, r);
        for (int k = 0; k < S && !breakdown; ++k) {
            beta[k] = 0.0;
            if (!isfinite(ssr[k]) || ssr[k] < 0.0 || !isfinite(ssr_old[k]) || ssr_old[k] < 0.0) { breakdown = 1; break; }
            if (solved[k]) { if (ssr[k] != 0.0) breakdown = 1; continue; }
            if (ssr[k] == 0.0) {
                if (!lane_exact_zero(r, n, S, k)) { breakdown = 1; break; }
                solved[k] = 1; continue;
            }
            if (ssr_old[k] <= 0.0) { breakdown = 1; break; }
            beta[k] = ssr[k] / ssr_old[k];
            if (!isfinite(beta[k]) || beta[k] <= 0.0) { breakdown = 1; break; }
        }
        if (breakdown) break;
        for (size_t i = 0; i < m; ++i) u[i] = r[i] + beta[i % S] * u[i];
        double update_error = 0.0;
        for (int k = 0; k < S; ++k) {
            if (solved[k]) continue;
            const double den = fabs(improvement[k]);
            if (!(recent[k] > 0.0) || !(den > 0.0)) { breakdown = 1; break; }
            const double e = sqrt(recent[k] / den);
            if (!isfinite(e)) { breakdown = 1; break; }
            if (e > update_error) update_error = e;
        }
        if (breakdown) break;
        if (update_error <= tol) { converged = 1; break; }
    }
#undef COLSUMSQ
#undef PROJ
    free(r); free(u); free(v);
    *iters_out = iters > 0 ? iters : max_iter;
    return converged;
}

/* ------------------------------------------------------- sweep order etc. */

typedef struct { int num_groups; double mean, cv, singleton_ratio, max_val; } prof;

/* profile_fe, src/fe_absorption.cpp:4945-4982. */
static prof profile(const double* ws, int G, int unit) {
    prof p = {G, 0.0, 0.0, 0.0, 0.0};
    if (G <= 0) return p;
    double sum = 0.0, sq = 0.0, mx = -INFINITY;
    int single = 0;
    for (int g = 0; g < G; ++g) {
        sum += ws[g]; sq += ws[g] * ws[g];
        if (ws[g] > mx) mx = ws[g];
        if (ws[g] <= 1.0 + 1e-12) ++single;
    }
    p.mean = sum / G;
    const double var = fmax(0.0, sq / G - p.mean * p.mean);
    p.cv = p.mean > 0.0 ? sqrt(var) / p.mean : 0.0;
    p.max_val = isfinite(mx) ? mx : 0.0;
    p.singleton_ratio = unit ? (double)single / G : 0.0;
    return p;
}

/* order_by_profile, src/fe_absorption.cpp:5472-5491 (stable sort). */
static void order_by_profile(const prof* p, int D, int* order) {
    for (int i = 0; i < D; ++i) order[i] = i;
    for (int i = 1; i < D; ++i) {
        const int key = order[i];
        int j = i - 1;
        while (j >= 0) {
            const prof* a = &p[key];
            const prof* b = &p[order[j]];
            int before;
            if (a->num_groups != b->num_groups) before = a->num_groups > b->num_groups;
            else if (a->singleton_ratio != b->singleton_ratio) before = a->singleton_ratio < b->singleton_ratio;
            else if (a->cv != b->cv) before = a->cv < b->cv;
            else before = key < order[j];
            if (!before) break;
            order[j + 1] = order[j];
            --j;
        }
        order[j + 1] = key;
    }
}

static int next_perm(int* a, int n) {
    int i = n - 2;
    while (i >= 0 && a[i] >= a[i + 1]) --i;
    if (i < 0) return 0;
    int j = n - 1;
    while (a[j] <= a[i]) --j;
    int t = a[i]; a[i] = a[j]; a[j] = t;
    for (int l = i + 1, r = n - 1; l < r; ++l, --r) { t = a[l]; a[l] = a[r]; a[r] = t; }
    return 1;
}

/* select_sweep_order, src/fe_absorption.cpp:5396-5470: try every permutation
 * (2 <= D <= 4) for two GS passes over y's first min(n, 200000) rows and keep
 * the one with the smallest remaining norm.  Trajectory detail only. */
static void select_order(const double* y, int n, int D, const int* gids, const int* G,
                         const double* w, int* order) {
    for (int d = 0; d < D; ++d) order[d] = d;
    if (D <= 1 || D > 4) return;
    const int np = n < 200000 ? n : 200000;
    if (np <= 0) return;
    double* tmp = (double*)malloc(sizeof(double) * (size_t)np);
    int perm[4];
    for (int d = 0; d < D; ++d) perm[d] = d;
    double best = INFINITY;
    do {
        memcpy(tmp, y, sizeof(double) * (size_t)np);
        for (int it = 0; it < 2; ++it)
            for (int p = 0; p < D; ++p) or_project(np, 1, tmp, gids + (size_t)perm[p] * n, G[perm[p]], w);
        const double nrm = sqrt(sumsq_all(tmp, (size_t)np));
        if (nrm < best) { best = nrm; memcpy(order, perm, sizeof(int) * (size_t)D); }
    } while (next_perm(perm, D));
    free(tmp);
}

void or_project(int n, int C, double* V, const int* gid, int G, const double* w) {
    double* sums = (double*)calloc((size_t)(G > 0 ? G : 1) * (size_t)C, sizeof(double));
    double* ws = (double*)calloc((size_t)(G > 0 ? G : 1), sizeof(double));
    for (int i = 0; i < n; ++i) {
        const double wi = w ? w[i] : 1.0;
        ws[gid[i]] += wi;
        for (int c = 0; c < C; ++c) sums[(size_t)gid[i] * C + c] += wi * V[(size_t)c * n + i];
    }
    for (int i = 0; i < n; ++i) {
        const int g = gid[i];
        const double inv = ws[g] > 0.0 ? 1.0 / ws[g] : 0.0;
        for (int c = 0; c < C; ++c) V[(size_t)c * n + i] -= sums[(size_t)g * C + c] * inv;
    }
    free(sums);
    free(ws);
}

/* --------------------------------------------------------------- absorber */

/* absorb_fixed_effects, src/fe_absorption.cpp:5993-8100 (packed Irons-Tuck
 * path :6331-7088 and the plain-MAP path :7368-7412), polish :7016-7068. */
int or_absorb(int n, int C, double* V, int D, const int* gids, const int* G, const double* w,
              const or_options* opt, or_result* res) {
    memset(res, 0, sizeof(*res));
    if (D <= 0) { res->converged = 1; return 0; }
    if (D > OR_MAX_DIMS || C > 64 || C <= 0) return -1;
    const int S = C;
    const size_t m = (size_t)n * S;
    const int honest = opt->tolerance_mode == 1;
    /* effective_absorption_tolerance :339-345 */
    const double tol_eff = (opt->tol > 0.0 && opt->tolerance_mode == 1) ? fmin(opt->tol, 1e-9) : opt->tol;
    or_ws ws;
    ws_init(&ws, n, S, D, gids, G, w, opt->threads);
    prof profs[OR_MAX_DIMS];
    for (int d = 0; d < D; ++d) profs[d] = profile(ws.wsum[d], G[d], w == NULL);
    int order[OR_MAX_DIMS];
    if (opt->sweep_order) memcpy(order, opt->sweep_order, sizeof(int) * (size_t)D);
    else if (D > 2) order_by_profile(profs, D, order);
    else select_order(V, n, D, gids, G, w, order);
    memcpy(res->sweep_order, order, sizeof(int) * (size_t)D);
    const int symmetric = opt->symmetric && D > 1;

    /* acceleration gate :6204-6232 */
    double min_occ = -1.0;
    for (int d = 0; d < D; ++d)
        if (profs[d].num_groups > 0 && (min_occ < 0.0 || profs[d].mean < min_occ)) min_occ = profs[d].mean;
    const int ill = D >= 2 && min_occ >= 0.0 && min_occ <= 12.0;
    const int use_accel = D >= 2 && (n >= 200000 || ill);
    const int budget = (use_accel && n < 200000) ? ((opt->max_iter * 2) / 5 > 1 ? (opt->max_iter * 2) / 5 : 1)
                                                 : opt->max_iter;
    /* two-stage warm-up :6293-6331 with classify_problem :5495-5513 */
    int iter_warmup = 0, iter_two = 0, two_order[2] = {-1, -1}, two_len = 0;
    if (D >= 3 && n >= 200000) {
        int first = -1, second = -1, third = -1;
        for (int d = 0; d < D; ++d) {
            const int g = profs[d].num_groups;
            if (g > first) { third = second; second = first; first = g; }
            else if (g > second) { third = second; second = g; }
            else if (g > third) { third = g; }
        }
        if (first > 0 && third > 0 && (double)third <= 0.05 * (double)first) {
            double msing = 0.0, mcv = 0.0;
            for (int d = 0; d < D; ++d) { msing = fmax(msing, profs[d].singleton_ratio); mcv = fmax(mcv, profs[d].cv); }
            int diff; /* 0 easy 1 medium 2 hard 3 extreme */
            if (D >= 4 && (msing >= 0.30 || mcv >= 2.0)) diff = 3;
            else if (D >= 4 || msing >= 0.20 || mcv >= 1.5) diff = 2;
            else if (D >= 3 || msing >= 0.10 || mcv >= 1.0) diff = 1;
            else diff = 0;
            if (diff == 3) { iter_warmup = 5; iter_two = 5; }
            else if (diff >= 2) { iter_warmup = 8; iter_two = 8; }
            else { iter_warmup = 5; iter_two = 8; }
            /* select_two_largest_dims :5515-5545 + filter_order :5547-5566 */
            int best = -1, sec = -1, bg = -1, sg = -1;
            for (int d = 0; d < D; ++d) {
                const int g = profs[d].num_groups;
                if (g > bg) { sec = best; sg = bg; best = d; bg = g; }
                else if (g > sg) { sec = d; sg = g; }
            }
            for (int p = 0; p < D; ++p)
                if (order[p] == best || order[p] == sec) two_order[two_len++] = order[p];
            if (two_len != 2) { iter_warmup = 0; iter_two = 0; two_len = 0; }
        }
    }
    /* CG gates: explicit reghdfe criterion, auto dominant-share (:5214-5240, share 0.37),
     * path-like two-way (:2039-2058). */
    int total_levels = 0;
    for (int d = 0; d < D; ++d) total_levels += G[d];
    int use_cg = 0;
    if (w == NULL) {
        if (opt->force_cg && D >= 2) use_cg = 1;
        if (opt->tolerance_mode == 1 && D >= 3) {
            double share = 0.0;
            for (int d = 0; d < D; ++d) share = fmax(share, profs[d].max_val / (double)n);
            if (share >= 0.37) use_cg = 1;
        }
        if (D == 2 && n > 0 && n <= 20000 && total_levels >= n - 2 && total_levels <= n + 2) use_cg = 1;
    }

    /* pack [y | X] column-major -> row-interleaved */
    double* cur = (double*)malloc(sizeof(double) * m);
    for (int i = 0; i < n; ++i)
        for (int s = 0; s < S; ++s) cur[(size_t)i * S + s] = V[(size_t)s * n + i];

    int converged = 0;
    if (use_accel || use_cg) {
        res->used_accel = 1;
        accel_ctx c;
        memset(&c, 0, sizeof(c));
        c.ws = &ws; c.symmetric = symmetric; c.honest = honest; c.honest_tol = opt->tol;
        c.conv_tol = tol_eff; c.m = m; c.cur = cur;
        c.gx = (double*)malloc(sizeof(double) * m);
        c.ggx = (double*)malloc(sizeof(double) * m);
        c.acc_a = (double*)malloc(sizeof(double) * m);
        c.acc_b = (double*)malloc(sizeof(double) * m);
        c.acc_c = (double*)malloc(sizeof(double) * m);
        if (use_cg) {
            res->used_cg = 1;
            converged = run_cg(&ws, cur, order, D, opt->max_iter, opt->tol, &res->iterations);
        } else {
            int remaining = budget, total = 0, used = 0;
            if (iter_warmup > 0 && remaining > 0) {
                converged = accel_phase(&c, order, D, iter_warmup < remaining ? iter_warmup : remaining, &used);
                total += used; remaining -= used;
            }
            if (!converged && !c.diverged && two_len == 2 && iter_two > 0 && remaining > 0) {
                accel_phase(&c, two_order, 2, iter_two < remaining ? iter_two : remaining, &used);
                total += used; remaining -= used;
            }
            if (!converged && !c.diverged && remaining > 0) {
                converged = accel_phase(&c, order, D, remaining, &used);
                total += used; remaining -= used;
            }
            if (!converged && c.diverged) {
                res->used_cg = 1;
                converged = run_cg(&ws, cur, order, D, opt->max_iter, opt->tol, &res->iterations);
            } else if (converged) {
                res->iterations = total;
            }
        }
        free(c.gx); free(c.ggx); free(c.acc_a); free(c.acc_b); free(c.acc_c);
    } else {
        /* plain MAP :7368-7412: stop when |‖v_k‖-‖v_{k-1}‖| / max(1,‖v_{k-1}‖) < tol_eff */
        double prev = sqrt(sumsq_all(cur, m));
        int last_check = -1;
        for (int iter = 0; iter < opt->max_iter; ++iter) {
            double sq = 0.0;
            sweep(&ws, cur, order, D, symmetric, &sq);
            const double curr = sqrt(sq);
            const double rel = fabs(curr - prev) / fmax(1.0, prev);
            const int step = iter - last_check;
            last_check = iter;
            prev = curr;
            if (step > 0 && rel / (double)step < tol_eff) {
                res->iterations = iter + 1;
                converged = 1;
                break;
            }
        }
    }
    if (!converged) {
        res->iterations = opt->max_iter;
        res->converged = 0;
    } else {
        res->converged = 1;
        if (res->iterations == 0) res->iterations = 1;
        /* tightening polish :7016-7068 */
        const int strict = opt->tolerance_mode == 2;
        const int max_polish = strict ? (opt->max_iter - res->iterations > 0 ? opt->max_iter - res->iterations : 0) : 6;
        const double polish_tol = strict ? opt->tol : (tol_eff > 0.0 ? tol_eff * 0.01 : tol_eff);
        double final_max = 0.0;
        int done = 0;
        for (int p = 0; p < max_polish; ++p) {
            double mx = 0.0;
            for (int q = 0; q < D; ++q) {
                const int d = order[q];
                demean_dim(&ws, cur, d, NULL);
                for (size_t j = 0; j < (size_t)G[d] * S; ++j) mx = fmax(mx, fabs(ws.means[d][j]));
            }
            if (symmetric) {
                for (int q = D; q-- > 0;) {
                    const int d = order[q];
                    demean_dim(&ws, cur, d, NULL);
                    for (size_t j = 0; j < (size_t)G[d] * S; ++j) mx = fmax(mx, fabs(ws.means[d][j]));
                }
            }
            final_max = mx;
            ++done;
            if (mx <= polish_tol) break;
        }
        res->polish_sweeps = done;
        res->polish_final_max = final_max;
    }
    for (int i = 0; i < n; ++i)
        for (int s = 0; s < S; ++s) V[(size_t)s * n + i] = cur[(size_t)i * S + s];
    free(cur);
    ws_free(&ws);
    return 0;
}

/* ------------------------------------------------------------- singletons */

/* compute_keep_mask_drop_singletons, src/hdfe_regressor_v11.cpp:4056-4302:
 * per-level count and XOR of member row ids; each round drops every row that
 * is alone in some level (w.r.t. the counts at the start of the round). */
int or_drop_singletons(int n, int D, const int* fes, int max_rounds, unsigned char* keep,
                       int* rounds_used) {
    *rounds_used = 0;
    if (n <= 0 || D <= 0) { for (int i = 0; i < n; ++i) keep[i] = 1; return 0; }
    int* gid = (int*)malloc(sizeof(int) * (size_t)n * D);
    int* G = (int*)malloc(sizeof(int) * (size_t)D);
    for (int d = 0; d < D; ++d) {
        int lp, cont;
        or_build_indexer(n, fes + (size_t)d * n, gid + (size_t)d * n, &G[d], &lp, &cont);
    }
    int** cnt = (int**)malloc(sizeof(int*) * (size_t)D);
    int** axr = (int**)malloc(sizeof(int*) * (size_t)D);
    for (int d = 0; d < D; ++d) {
        cnt[d] = (int*)calloc((size_t)G[d] + 1, sizeof(int));
        axr[d] = (int*)calloc((size_t)G[d] + 1, sizeof(int));
        for (int i = 0; i < n; ++i) {
            const int g = gid[(size_t)d * n + i];
            cnt[d][g] += 1;
            axr[d][g] ^= i;
        }
    }
    int* cur = (int*)malloc(sizeof(int) * ((size_t)n * D + 1));
    int* nxt = (int*)malloc(sizeof(int) * ((size_t)n * D + 1));
    int* drop = (int*)malloc(sizeof(int) * ((size_t)n + 1));
    int ncur = 0;
    for (int d = 0; d < D; ++d)
        for (int g = 0; g < G[d]; ++g)
            if (cnt[d][g] == 1) cur[ncur++] = axr[d][g];
    for (int i = 0; i < n; ++i) keep[i] = 1;
    int dropped = 0;
    const int iters = max_rounds > 1 ? max_rounds : 1;
    for (int it = 0; it < iters; ++it) {
        if (ncur == 0) break;
        int nd = 0;
        for (int c = 0; c < ncur; ++c) {
            const int i = cur[c];
            if (!keep[i]) continue;
            int single = 0;
            for (int d = 0; d < D && !single; ++d) single = cnt[d][gid[(size_t)d * n + i]] == 1;
            if (single) { keep[i] = 0; drop[nd++] = i; }
        }
        if (nd == 0) break;
        *rounds_used = it + 1;
        dropped += nd;
        for (int c = 0; c < nd; ++c)
            for (int d = 0; d < D; ++d) {
                const int g = gid[(size_t)d * n + drop[c]];
                cnt[d][g] -= 1;
                axr[d][g] ^= drop[c];
            }
        int nn = 0;
        for (int c = 0; c < nd; ++c)
            for (int d = 0; d < D; ++d) {
                const int g = gid[(size_t)d * n + drop[c]];
                if (cnt[d][g] == 1) nxt[nn++] = axr[d][g];
            }
        int* t = cur; cur = nxt; nxt = t;
        ncur = nn;
    }
    for (int d = 0; d < D; ++d) { free(cnt[d]); free(axr[d]); }
    free(cnt); free(axr); free(cur); free(nxt); free(drop); free(gid); free(G);
    return dropped;
}

/* ------------------------------------------------------------- components */

/* build_lookup_compact, src/hdfe_regressor_v11.cpp:1354-1417: dense (raw-min)
 * when the caller's level count equals the id range, otherwise numbering by
 * first appearance over the rows. */
static int lookup_compact(int n, const int* raw, int expected, int* node) {
    if (n == 0) return 0;
    int mn = raw[0], mx = raw[0];
    for (int i = 1; i < n; ++i) { if (raw[i] < mn) mn = raw[i]; else if (raw[i] > mx) mx = raw[i]; }
    const long long range = (long long)mx - (long long)mn + 1LL;
    if (mn >= 0 && range > 0 && range <= 50000000LL && range <= 2LL * n) {
        if (expected > 0 && expected == range) {
            for (int i = 0; i < n; ++i) node[i] = raw[i] - mn;
            return (int)range;
        }
        int* map = (int*)malloc(sizeof(int) * (size_t)range);
        for (long long k = 0; k < range; ++k) map[k] = -1;
        int next = 0;
        for (int i = 0; i < n; ++i) {
            int* slot = &map[raw[i] - mn];
            if (*slot < 0) *slot = next++;
            node[i] = *slot;
        }
        free(map);
        return next;
    }
    /* sparse: first appearance (same numbering as the unordered_map path) */
    int64_t* keyed = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    for (int i = 0; i < n; ++i) keyed[i] = ((int64_t)raw[i] << 32) | (int64_t)(uint32_t)i;
    qsort(keyed, (size_t)n, sizeof(int64_t), cmp_pair);
    int* first_row = (int*)malloc(sizeof(int) * (size_t)n);
    int* key_of = (int*)malloc(sizeof(int) * (size_t)n);
    int nk = 0;
    for (int k = 0; k < n; ++k) {
        const int key = (int)(keyed[k] >> 32);
        const int pos = (int)(keyed[k] & 0xffffffffLL);
        if (k == 0 || key != (int)(keyed[k - 1] >> 32)) { first_row[nk] = pos; ++nk; }
        key_of[pos] = nk - 1;
    }
    /* rank keys by first row */
    int64_t* fr = (int64_t*)malloc(sizeof(int64_t) * (size_t)nk);
    for (int k = 0; k < nk; ++k) fr[k] = ((int64_t)first_row[k] << 32) | k;
    qsort(fr, (size_t)nk, sizeof(int64_t), cmp_pair);
    int* rank = (int*)malloc(sizeof(int) * (size_t)nk);
    for (int r = 0; r < nk; ++r) rank[fr[r] & 0xffffffffLL] = r;
    for (int i = 0; i < n; ++i) node[i] = rank[key_of[i]];
    free(keyed); free(first_row); free(key_of); free(fr); free(rank);
    return nk;
}

static int uf_find(int* parent, int x) {
    while (parent[x] != x) { parent[x] = parent[parent[x]]; x = parent[x]; }
    return x;
}

int or_components(int n, const int* a, const int* b, int expected_a, int expected_b, int* labels) {
    if (n <= 0) return 0;
    int* na = (int*)malloc(sizeof(int) * (size_t)n);
    int* nb = (int*)malloc(sizeof(int) * (size_t)n);
    const int Ga = lookup_compact(n, a, expected_a, na);
    const int Gb = lookup_compact(n, b, expected_b, nb);
    if (Ga <= 0 || Gb <= 0) { free(na); free(nb); return 0; }
    const int total = Ga + Gb;
    int* parent = (int*)malloc(sizeof(int) * (size_t)total);
    unsigned char* rk = (unsigned char*)calloc((size_t)total, 1);
    for (int v = 0; v < total; ++v) parent[v] = v;
    /* UnionFind (:1586-1620): union by rank, path halving */
    for (int i = 0; i < n; ++i) {
        int ra = uf_find(parent, na[i]), rb = uf_find(parent, Ga + nb[i]);
        if (ra == rb) continue;
        if (rk[ra] < rk[rb]) parent[ra] = rb;
        else if (rk[ra] > rk[rb]) parent[rb] = ra;
        else { parent[rb] = ra; rk[ra]++; }
    }
    int* comp = (int*)malloc(sizeof(int) * (size_t)total);
    for (int v = 0; v < total; ++v) comp[v] = -1;
    int ncomp = 0;
    for (int v = 0; v < total; ++v) {
        const int r = uf_find(parent, v);
        if (comp[r] < 0) comp[r] = ncomp++;
    }
    if (labels)
        for (int i = 0; i < n; ++i) labels[i] = comp[uf_find(parent, na[i])] + 1;
    free(na); free(nb); free(parent); free(rk); free(comp);
    return ncomp;
}

/* ------------------------------------------------------------------- DoF */

int or_fe_is_function_of(int n, const int* coarse, const int* refined) {
    if (n == 0) return 1;
    int mn = refined[0], mx = refined[0];
    for (int i = 1; i < n; ++i) { if (refined[i] < mn) mn = refined[i]; if (refined[i] > mx) mx = refined[i]; }
    const long long range = (long long)mx - (long long)mn + 1LL;
    int ok = 1;
    if (mn >= 0 && range > 0 && range <= 50000000LL && range <= 2LL * n) {
        int* map = (int*)malloc(sizeof(int) * (size_t)range);
        unsigned char* seen = (unsigned char*)calloc((size_t)range, 1);
        for (int i = 0; i < n && ok; ++i) {
            const int k = refined[i] - mn;
            if (!seen[k]) { seen[k] = 1; map[k] = coarse[i]; }
            else if (map[k] != coarse[i]) ok = 0;
        }
        free(map); free(seen);
        return ok;
    }
    int64_t* keyed = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    for (int i = 0; i < n; ++i) keyed[i] = ((int64_t)refined[i] << 32) | (int64_t)(uint32_t)i;
    qsort(keyed, (size_t)n, sizeof(int64_t), cmp_pair);
    for (int k = 1; k < n && ok; ++k) {
        const int rk0 = (int)(keyed[k - 1] >> 32), rk1 = (int)(keyed[k] >> 32);
        if (rk0 == rk1 && coarse[keyed[k - 1] & 0xffffffffLL] != coarse[keyed[k] & 0xffffffffLL]) ok = 0;
    }
    free(keyed);
    return ok;
}

/* compute_fe_dof_reghdfe, src/hdfe_regressor_v11.cpp:3743-3986, method
 * "all"/"pairwise" without groupvar: coarsening of FEs that are functions of a
 * finer FE, then redundant[d] = max_{j<d} components(fe_j, fe_d). */
int or_dof_reghdfe(int n, int D, const int* fes, const int* levels, int* redundant, int* num_coefs) {
    if (D <= 0) return 0;
    for (int d = 0; d < D; ++d) { redundant[d] = 0; num_coefs[d] = levels[d] > 0 ? levels[d] : 0; }
    if (D >= 2) {
        unsigned char coarsened[OR_MAX_DIMS] = {0};
        int any = 0;
        for (int d = 0; d < D; ++d)
            for (int e = 0; e < D; ++e) {
                if (d == e || levels[e] <= levels[d]) continue;
                if (or_fe_is_function_of(n, fes + (size_t)d * n, fes + (size_t)e * n)) {
                    coarsened[d] = 1; any = 1; break;
                }
            }
        if (any) {
            int eff_map[OR_MAX_DIMS], ne = 0;
            int eff_levels[OR_MAX_DIMS];
            for (int d = 0; d < D; ++d) if (!coarsened[d]) { eff_map[ne] = d; eff_levels[ne] = levels[d]; ++ne; }
            if (ne > 0 && ne < D) {
                int* sub = (int*)malloc(sizeof(int) * (size_t)n * ne);
                for (int q = 0; q < ne; ++q) memcpy(sub + (size_t)q * n, fes + (size_t)eff_map[q] * n, sizeof(int) * (size_t)n);
                int er[OR_MAX_DIMS], ec[OR_MAX_DIMS];
                or_dof_reghdfe(n, ne, sub, eff_levels, er, ec);
                free(sub);
                for (int d = 0; d < D; ++d) if (coarsened[d]) { redundant[d] = levels[d]; num_coefs[d] = 0; }
                for (int q = 0; q < ne; ++q) { redundant[eff_map[q]] = er[q]; num_coefs[eff_map[q]] = ec[q]; }
                int df = 0;
                for (int d = 0; d < D; ++d) df += num_coefs[d];
                return df;
            }
        }
        for (int d = 1; d < D; ++d) {
            int mx = 0;
            for (int j = 0; j < d; ++j) {
                const int c = or_components(n, fes + (size_t)j * n, fes + (size_t)d * n, levels[j], levels[d], NULL);
                if (c > mx) mx = c;
            }
            redundant[d] = mx;
            num_coefs[d] = levels[d] - mx > 0 ? levels[d] - mx : 0;
        }
    }
    int df = 0;
    for (int d = 0; d < D; ++d) df += num_coefs[d];
    return df;
}
