/*
 * pgpu_gen.h - counter-based synthetic input generators for the hash-join /
 * hash-aggregation workloads.
 *
 * Every value is a pure function of (seed, stream, index), so the host, the
 * oracle and the GPU produce bit-identical inputs independently, in any
 * order and in parallel.  The core is SplitMix64 (Steele, Lea & Flood,
 * "Fast splittable pseudorandom number generators", OOPSLA 2014): output i of
 * a stream is mix64(streamkey + (i+1) * golden).
 *
 * Portable C99: included by gcc (oracle, host code) and by hipcc (kernels).
 * Host-only helpers (the Zipf CDF, log-uniform sizes) use libm and are kept
 * out of device code.
 *
 * Workload shapes (sizes use K = 2^10, M = 2^20 where radix designs want
 * powers of two):
 *   C1  R = 2^20 x (int32 key = permutation of 0..2^20-1, int32 payload),
 *       S = 2^24 x (int32 fk uniform over R keys, int32 payload).
 *   C2  R = 2^24 x (int64 unique key = bijective mix of the row index,
 *       int64 payload U[0,2^31)), S = 2^28 x int64 fk, uniform or Zipf(s)
 *       over R's rows (ranks mapped through a seeded permutation).
 *   C3  D = 15e6 x (int32 orderkey = permutation of 1..15e6, int32 custkey
 *       U[1,1.5e6]); F = 60e6 x (int32 orderkey U[1,15e6], int32 partkey
 *       U[1,2e6], int32 quantity U[1,50], FP64 price = integer cents / 100
 *       with cents U[qty*90000, qty*200000]).
 *   C4  1024 int32 joins, build n_i = round(2^U(10,22)), keys = permutation
 *       of 0..n_i-1; probe 8 n_i rows, 10% drawn from the disjoint range
 *       [n_i, n_i + 2^30).
 *   C5  64 resident int64 tables n_t = round(2^U(16,23)); probe batches
 *       m_q = round(2^U(14,25)) uniform over the chosen table's keys.
 */
#ifndef PGPU_GEN_H
#define PGPU_GEN_H

#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define PGPU_HD __host__ __device__ static inline
#else
#define PGPU_HD static inline
#endif

/* Stream identifiers: base seed 42 with fixed stream ids. */
enum {
    PGPU_STREAM_R = 1,        /* build keys / permutations            */
    PGPU_STREAM_S = 2,        /* probe foreign keys                   */
    PGPU_STREAM_SIZES = 3,    /* C4/C5 relation sizes                 */
    PGPU_STREAM_ARRIVALS = 4, /* C5 Poisson arrivals                  */
    PGPU_STREAM_ZIPF = 5,     /* rank -> row permutation              */
    PGPU_STREAM_TABLEID = 6,  /* C5 table choice                      */
    PGPU_STREAM_RPAY = 7,     /* build payloads / dimension attribute */
    PGPU_STREAM_SPAY = 8,     /* probe payloads / fact attribute      */
    PGPU_STREAM_QTY = 9,      /* C3 quantity                          */
    PGPU_STREAM_PRICE = 10,   /* C3 price                             */
    PGPU_STREAM_MISS = 11     /* C4 non-matching probe keys           */
};

#define PGPU_DEFAULT_SEED 42ULL
#define PGPU_GOLDEN 0x9E3779B97F4A7C15ULL

/* SplitMix64 finalizer: a bijection on 64-bit words. */
PGPU_HD uint64_t pgpu_mix64(uint64_t z)
{
    z ^= z >> 30;
    z *= 0xBF58476D1CE4E5B9ULL;
    z ^= z >> 27;
    z *= 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return z;
}

/* Key of stream `stream` (optionally sub-stream `sub`, e.g. the join index
 * of a batch) under `seed`. */
PGPU_HD uint64_t pgpu_stream_key(uint64_t seed, uint64_t stream, uint64_t sub)
{
    return pgpu_mix64(pgpu_mix64(seed * PGPU_GOLDEN + stream) ^
                      (sub + 1) * 0xD1B54A32D192ED03ULL);
}

/* i-th 64-bit output of the stream whose key is `key`. */
PGPU_HD uint64_t pgpu_rand(uint64_t key, uint64_t i)
{
    return pgpu_mix64(key + (i + 1) * PGPU_GOLDEN);
}

PGPU_HD uint64_t pgpu_mulhi64(uint64_t a, uint64_t b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(a, b);
#else
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

/* Uniform integer in [0, n) from a 64-bit random word (multiply-high). */
PGPU_HD uint64_t pgpu_uniform(uint64_t r, uint64_t n)
{
    return pgpu_mulhi64(r, n);
}

/* Uniform double in [0, 1) with 53 random bits. */
PGPU_HD double pgpu_unit(uint64_t r)
{
    return (double)(r >> 11) * (1.0 / 9007199254740992.0);
}

/* Smallest k with 2^k >= n (n >= 1). */
PGPU_HD uint32_t pgpu_ceil_log2(uint64_t n)
{
    uint32_t k = 0;
    while (k < 64 && (1ULL << k) < n)
        ++k;
    return k;
}

/* One application of a keyed bijection on [0, 2^bits): odd multiply, add,
 * xorshift, odd multiply, xorshift.  Each step is invertible modulo 2^bits. */
PGPU_HD uint64_t pgpu_perm_round(uint64_t x, uint64_t key, uint32_t bits)
{
    const uint64_t mask = (bits >= 64) ? ~0ULL : ((1ULL << bits) - 1);
    const uint32_t h = bits > 1 ? bits / 2 : 1;
    const uint64_t a = (pgpu_mix64(key) | 1ULL);
    const uint64_t b = pgpu_mix64(key ^ 0x5851F42D4C957F2DULL);
    const uint64_t c = (pgpu_mix64(key + PGPU_GOLDEN) | 1ULL);
    if (bits == 0)
        return 0;
    x = (x * a + b) & mask;
    x ^= x >> h;
    x = (x * c) & mask;
    x ^= x >> h;
    x = (x * a + c) & mask;
    return x;
}

/* Seeded permutation of [0, n): cycle-walk the power-of-two bijection until
 * the value lands inside [0, n).  Terminates because the orbit of x returns
 * to x; expected iterations < 2. */
PGPU_HD uint64_t pgpu_perm(uint64_t x, uint64_t n, uint64_t key)
{
    const uint32_t bits = pgpu_ceil_log2(n);
    do {
        x = pgpu_perm_round(x, key, bits);
    } while (x >= n);
    return x;
}

/* ---------------- C1: int32 keys --------------------------------------- */
PGPU_HD int32_t pgpu_c1_rkey(uint64_t seed, uint64_t i, uint64_t nR)
{
    return (int32_t)pgpu_perm(i, nR, pgpu_stream_key(seed, PGPU_STREAM_R, 0));
}
PGPU_HD int32_t pgpu_c1_rpay(uint64_t seed, uint64_t i)
{
    return (int32_t)(uint32_t)(pgpu_rand(pgpu_stream_key(seed, PGPU_STREAM_RPAY, 0), i) >> 32);
}
PGPU_HD int32_t pgpu_c1_sfk(uint64_t seed, uint64_t j, uint64_t nR)
{
    return (int32_t)pgpu_uniform(pgpu_rand(pgpu_stream_key(seed, PGPU_STREAM_S, 0), j), nR);
}
PGPU_HD int32_t pgpu_c1_spay(uint64_t seed, uint64_t j)
{
    return (int32_t)(uint32_t)(pgpu_rand(pgpu_stream_key(seed, PGPU_STREAM_SPAY, 0), j) >> 32);
}

/* ---------------- C2 / C5: int64 keys ---------------------------------- */
/* Unique int64 key of build row i of table `sub`: the bijection mix64 of
 * (i + streamkey), so keys are distinct and not dense. */
PGPU_HD int64_t pgpu_c2_rkey(uint64_t seed, uint64_t sub, uint64_t i)
{
    return (int64_t)pgpu_mix64(i + pgpu_stream_key(seed, PGPU_STREAM_R, sub));
}
PGPU_HD int64_t pgpu_c2_rpay(uint64_t seed, uint64_t sub, uint64_t i)
{
    return (int64_t)(pgpu_rand(pgpu_stream_key(seed, PGPU_STREAM_RPAY, sub), i) >> 33);
}
/* Build row referenced by probe row j (uniform). */
PGPU_HD uint64_t pgpu_c2_srow_uniform(uint64_t seed, uint64_t sub, uint64_t j, uint64_t nR)
{
    return pgpu_uniform(pgpu_rand(pgpu_stream_key(seed, PGPU_STREAM_S, sub), j), nR);
}
/* Zipf: rank = upper_bound(cdf, u); row = perm(rank).  `cdf` holds nR
 * non-decreasing doubles with cdf[nR-1] == 1.0. */
PGPU_HD uint64_t pgpu_zipf_rank(const double* cdf, uint64_t n, double u)
{
    uint64_t lo = 0, hi = n - 1; /* answer in [lo, hi]; cdf[n-1] = 1 > u */
    while (lo < hi) {
        uint64_t mid = lo + ((hi - lo) >> 1);
        if (cdf[mid] > u)
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}
PGPU_HD uint64_t pgpu_c2_srow_zipf(uint64_t seed, uint64_t sub, uint64_t j, uint64_t nR,
                                   const double* cdf)
{
    double u = pgpu_unit(pgpu_rand(pgpu_stream_key(seed, PGPU_STREAM_S, sub), j));
    uint64_t rank = pgpu_zipf_rank(cdf, nR, u);
    return pgpu_perm(rank, nR, pgpu_stream_key(seed, PGPU_STREAM_ZIPF, sub));
}

/* ---------------- C3: join + group-by ---------------------------------- */
#define PGPU_C3_ORDERS 15000000ULL
#define PGPU_C3_CUSTS 1500000ULL
#define PGPU_C3_PARTS 2000000ULL
PGPU_HD int32_t pgpu_c3_d_orderkey(uint64_t seed, uint64_t i, uint64_t nD)
{
    return (int32_t)(1 + pgpu_perm(i, nD, pgpu_stream_key(seed, PGPU_STREAM_R, 3)));
}
PGPU_HD int32_t pgpu_c3_d_custkey(uint64_t seed, uint64_t i, uint64_t ncust)
{
    return (int32_t)(1 + pgpu_uniform(pgpu_rand(pgpu_stream_key(seed, PGPU_STREAM_RPAY, 3), i), ncust));
}
PGPU_HD int32_t pgpu_c3_f_orderkey(uint64_t seed, uint64_t j, uint64_t norders)
{
    return (int32_t)(1 + pgpu_uniform(pgpu_rand(pgpu_stream_key(seed, PGPU_STREAM_S, 3), j), norders));
}
PGPU_HD int32_t pgpu_c3_f_partkey(uint64_t seed, uint64_t j)
{
    return (int32_t)(1 + pgpu_uniform(pgpu_rand(pgpu_stream_key(seed, PGPU_STREAM_SPAY, 3), j), PGPU_C3_PARTS));
}
PGPU_HD int32_t pgpu_c3_f_quantity(uint64_t seed, uint64_t j)
{
    return (int32_t)(1 + pgpu_uniform(pgpu_rand(pgpu_stream_key(seed, PGPU_STREAM_QTY, 3), j), 50));
}
/* Price in integer cents: quantity * U(900.00, 2000.00) rounded to 0.01. */
PGPU_HD int64_t pgpu_c3_f_cents(uint64_t seed, uint64_t j)
{
    uint64_t q = (uint64_t)pgpu_c3_f_quantity(seed, j);
    return (int64_t)(q * 90000ULL +
                     pgpu_uniform(pgpu_rand(pgpu_stream_key(seed, PGPU_STREAM_PRICE, 3), j),
                                  q * 110000ULL + 1));
}
PGPU_HD double pgpu_c3_f_price(uint64_t seed, uint64_t j)
{
    return (double)pgpu_c3_f_cents(seed, j) / 100.0;
}

/* ---------------- C4: batched int32 joins ------------------------------ */
PGPU_HD int32_t pgpu_c4_rkey(uint64_t seed, uint64_t join, uint64_t i, uint64_t n)
{
    return (int32_t)pgpu_perm(i, n, pgpu_stream_key(seed, PGPU_STREAM_R, 16 + join));
}
PGPU_HD int32_t pgpu_c4_rpay(uint64_t seed, uint64_t join, uint64_t i)
{
    return (int32_t)(pgpu_rand(pgpu_stream_key(seed, PGPU_STREAM_RPAY, 16 + join), i) >> 33);
}
PGPU_HD int32_t pgpu_c4_spay(uint64_t seed, uint64_t join, uint64_t j)
{
    return (int32_t)(pgpu_rand(pgpu_stream_key(seed, PGPU_STREAM_SPAY, 16 + join), j) >> 33);
}
/* 10% of probe rows carry a key from [n, n + 2^30), never in the build. */
PGPU_HD int32_t pgpu_c4_sfk(uint64_t seed, uint64_t join, uint64_t j, uint64_t n)
{
    uint64_t r = pgpu_rand(pgpu_stream_key(seed, PGPU_STREAM_S, 16 + join), j);
    if (pgpu_uniform(r, 10) == 0)
        return (int32_t)(n + pgpu_uniform(pgpu_rand(pgpu_stream_key(seed, PGPU_STREAM_MISS, 16 + join), j),
                                          1ULL << 30));
    return (int32_t)pgpu_uniform(pgpu_mix64(r ^ 0x2545F4914F6CDD1DULL), n);
}

/* ---------------- host-only helpers ------------------------------------ */
#if !defined(__HIP_DEVICE_COMPILE__)
#include <math.h>

/* Zipf(s) CDF over ranks 1..n, normalised, last entry forced to 1.0.
 * Summation is sequential in rank order so every caller gets identical
 * doubles. */
static inline void pgpu_zipf_cdf_fill(double* cdf, uint64_t n, double s)
{
    double acc = 0.0;
    uint64_t r;
    for (r = 0; r < n; ++r) {
        acc += pow((double)(r + 1), -s);
        cdf[r] = acc;
    }
    for (r = 0; r < n; ++r)
        cdf[r] /= acc;
    if (n)
        cdf[n - 1] = 1.0;
}

/* round(2^U(lo, hi)) for item `idx` of a sized family (C4 joins, C5 tables
 * and probe batches).  `sub` separates families. */
static inline uint64_t pgpu_loguniform_size(uint64_t seed, uint64_t sub, uint64_t idx,
                                            double lo, double hi)
{
    double u = pgpu_unit(pgpu_rand(pgpu_stream_key(seed, PGPU_STREAM_SIZES, sub), idx));
    return (uint64_t)llround(exp2(lo + (hi - lo) * u));
}

/* C5: the table a query probes (uniform over ntables). */
static inline uint32_t pgpu_c5_table(uint64_t seed, uint64_t q, uint64_t ntables)
{
    return (uint32_t)pgpu_uniform(pgpu_rand(pgpu_stream_key(seed, PGPU_STREAM_TABLEID, 0), q), ntables);
}

/* C5: exponential inter-arrival gap (seconds) before query q at `rate`
 * queries/s; `sub` separates streams of different rates. */
static inline double pgpu_c5_gap(uint64_t seed, uint64_t sub, uint64_t q, double rate)
{
    double u = pgpu_unit(pgpu_rand(pgpu_stream_key(seed, PGPU_STREAM_ARRIVALS, sub), q));
    return -log1p(-u) / rate;
}
#endif

#endif /* PGPU_GEN_H */
