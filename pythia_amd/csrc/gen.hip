// Device-side synthetic column generation (formulas in include/pgpu_gen.h).
// Inputs are produced directly in HBM so timed regions start with resident
// data; the oracle evaluates the same formulas on the host.
#include "pgpu_internal.h"
#include "../../include/pgpu_gen.h"

namespace {

template <int COL>
__global__ __launch_bounds__(256) void k_gen(pgpu_gen_spec spec, uint64_t n, void* out) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) {
        const uint64_t i = spec.offset + t;
        if constexpr (COL == PGPU_GEN_C1_RKEY) {
            ((int32_t*)out)[t] = pgpu_c1_rkey(spec.seed, i, spec.n_build);
        } else if constexpr (COL == PGPU_GEN_C1_RPAY) {
            ((int32_t*)out)[t] = pgpu_c1_rpay(spec.seed, i);
        } else if constexpr (COL == PGPU_GEN_C1_SFK) {
            ((int32_t*)out)[t] = pgpu_c1_sfk(spec.seed, i, spec.n_build);
        } else if constexpr (COL == PGPU_GEN_C1_SPAY) {
            ((int32_t*)out)[t] = pgpu_c1_spay(spec.seed, i);
        } else if constexpr (COL == PGPU_GEN_C2_RKEY) {
            ((int64_t*)out)[t] = pgpu_c2_rkey(spec.seed, spec.sub, i);
        } else if constexpr (COL == PGPU_GEN_C2_RPAY) {
            ((int64_t*)out)[t] = pgpu_c2_rpay(spec.seed, spec.sub, i);
        } else if constexpr (COL == PGPU_GEN_C2_SFK_UNIFORM) {
            uint64_t row = pgpu_c2_srow_uniform(spec.seed, spec.sub, i, spec.n_build);
            ((int64_t*)out)[t] = pgpu_c2_rkey(spec.seed, spec.sub, row);
        } else if constexpr (COL == PGPU_GEN_C2_SFK_ZIPF) {
            uint64_t row = pgpu_c2_srow_zipf(spec.seed, spec.sub, i, spec.n_build, spec.d_cdf);
            ((int64_t*)out)[t] = pgpu_c2_rkey(spec.seed, spec.sub, row);
        } else if constexpr (COL == PGPU_GEN_C3_D_ORDERKEY) {
            ((int32_t*)out)[t] = pgpu_c3_d_orderkey(spec.seed, i, spec.n_build);
        } else if constexpr (COL == PGPU_GEN_C3_D_CUSTKEY) {
            ((int32_t*)out)[t] = pgpu_c3_d_custkey(spec.seed, i, PGPU_C3_CUSTS);
        } else if constexpr (COL == PGPU_GEN_C3_F_ORDERKEY) {
            ((int32_t*)out)[t] = pgpu_c3_f_orderkey(spec.seed, i, spec.n_build);
        } else if constexpr (COL == PGPU_GEN_C3_F_PARTKEY) {
            ((int32_t*)out)[t] = pgpu_c3_f_partkey(spec.seed, i);
        } else if constexpr (COL == PGPU_GEN_C3_F_QUANTITY) {
            ((int32_t*)out)[t] = pgpu_c3_f_quantity(spec.seed, i);
        } else if constexpr (COL == PGPU_GEN_C3_F_PRICE) {
            ((double*)out)[t] = pgpu_c3_f_price(spec.seed, i);
        } else if constexpr (COL == PGPU_GEN_C4_RKEY) {
            ((int32_t*)out)[t] = pgpu_c4_rkey(spec.seed, spec.sub, i, spec.n_build);
        } else if constexpr (COL == PGPU_GEN_C4_RPAY) {
            ((int32_t*)out)[t] = pgpu_c4_rpay(spec.seed, spec.sub, i);
        } else if constexpr (COL == PGPU_GEN_C4_SFK) {
            ((int32_t*)out)[t] = pgpu_c4_sfk(spec.seed, spec.sub, i, spec.n_build);
        } else if constexpr (COL == PGPU_GEN_C4_SPAY) {
            ((int32_t*)out)[t] = pgpu_c4_spay(spec.seed, spec.sub, i);
        }
    }
}

template <int COL>
int launch(pgpu_ctx* ctx, const pgpu_gen_spec& s, uint64_t n, void* out) {
    const uint64_t blocks = std::min<uint64_t>(pgpu::ceil_div(n, 256), 256 * 64);
    hipLaunchKernelGGL(k_gen<COL>, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, s, n, out);
    PGPU_LAUNCH_CHECK();
    return PGPU_OK;
}

}  // namespace

extern "C" int pgpu_gen(pgpu_ctx* ctx, int column, const pgpu_gen_spec* spec, uint64_t n, void* d_out) {
    using namespace pgpu;
    if (!ctx || !spec || (!d_out && n)) return fail(PGPU_EINVAL, "pgpu_gen: null argument");
    if (n == 0) return PGPU_OK;
    const bool needs_build = column == PGPU_GEN_C1_RKEY || column == PGPU_GEN_C1_SFK ||
                             column == PGPU_GEN_C2_SFK_UNIFORM || column == PGPU_GEN_C2_SFK_ZIPF ||
                             column == PGPU_GEN_C3_D_ORDERKEY || column == PGPU_GEN_C3_F_ORDERKEY ||
                             column == PGPU_GEN_C4_RKEY || column == PGPU_GEN_C4_SFK;
    if (needs_build && spec->n_build == 0) return fail(PGPU_EINVAL, "pgpu_gen: column %d needs n_build > 0", column);
    if (column == PGPU_GEN_C2_SFK_ZIPF && !spec->d_cdf) return fail(PGPU_EINVAL, "pgpu_gen: Zipf column needs d_cdf");
    PGPU_TRY(ctx_activate(ctx));
    switch (column) {
#define CASE(c) \
    case c: return launch<c>(ctx, *spec, n, d_out);
        CASE(PGPU_GEN_C1_RKEY)
        CASE(PGPU_GEN_C1_RPAY)
        CASE(PGPU_GEN_C1_SFK)
        CASE(PGPU_GEN_C1_SPAY)
        CASE(PGPU_GEN_C2_RKEY)
        CASE(PGPU_GEN_C2_RPAY)
        CASE(PGPU_GEN_C2_SFK_UNIFORM)
        CASE(PGPU_GEN_C2_SFK_ZIPF)
        CASE(PGPU_GEN_C3_D_ORDERKEY)
        CASE(PGPU_GEN_C3_D_CUSTKEY)
        CASE(PGPU_GEN_C3_F_ORDERKEY)
        CASE(PGPU_GEN_C3_F_PARTKEY)
        CASE(PGPU_GEN_C3_F_QUANTITY)
        CASE(PGPU_GEN_C3_F_PRICE)
        CASE(PGPU_GEN_C4_RKEY)
        CASE(PGPU_GEN_C4_RPAY)
        CASE(PGPU_GEN_C4_SFK)
        CASE(PGPU_GEN_C4_SPAY)
#undef CASE
        default:
            return fail(PGPU_EINVAL, "pgpu_gen: unknown column %d", column);
    }
}

// Host-only helpers (the formulas they use are not compiled for the device).
#if !defined(__HIP_DEVICE_COMPILE__)
extern "C" int pgpu_zipf_cdf(uint64_t n, double s, double* cdf) {
    using namespace pgpu;
    if (!cdf || n == 0 || !(s > 0)) return fail(PGPU_EINVAL, "pgpu_zipf_cdf: need n > 0, s > 0 and an output array");
    pgpu_zipf_cdf_fill(cdf, n, s);
    return PGPU_OK;
}

extern "C" int pgpu_sizes(uint64_t seed, uint64_t sub, uint64_t first, uint64_t count, double lo, double hi,
                          uint64_t* out) {
    using namespace pgpu;
    if ((!out && count) || !(lo <= hi) || hi > 40) return fail(PGPU_EINVAL, "pgpu_sizes: bad argument");
    for (uint64_t i = 0; i < count; ++i) out[i] = pgpu_loguniform_size(seed, sub, first + i, lo, hi);
    return PGPU_OK;
}

extern "C" int pgpu_c5_trace(uint64_t seed, uint64_t count, uint32_t ntables, double lo, double hi, double rate,
                             uint32_t* tables, uint64_t* ms, double* offsets) {
    using namespace pgpu;
    if ((count && (!tables || !ms || !offsets)) || ntables == 0 || !(rate > 0) || !(lo <= hi) || hi > 40)
        return fail(PGPU_EINVAL, "pgpu_c5_trace: bad argument");
    double t = 0;
    for (uint64_t q = 0; q < count; ++q) {
        tables[q] = pgpu_c5_table(seed, q, ntables);
        ms[q] = pgpu_loguniform_size(seed, 6, q, lo, hi);
        t += pgpu_c5_gap(seed, (uint64_t)llround(rate), q, rate);
        offsets[q] = t;
    }
    return PGPU_OK;
}
#endif
