/*
 * oracle.c - TEST INFRASTRUCTURE ONLY: CPU restatement of the reference's
 * hash-join / hash-aggregation path.  See oracle.h for the contract.  Every
 * function names the reference file:line it follows.  Nothing here is
 * linked into libpgpu.so; the product path never reaches this code.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../include/pgpu_gen.h"
#include <stdio.h>

/* ===================================================================== */
/* schema (schema.cpp:82-117): packed offsets, no alignment padding       */
/* ===================================================================== */
void orc_schema_add(orc_schema* s, int type, int size)
{
    int w = 0;
    switch (type) {
    case ORC_INT: w = 4; break;
    case ORC_LONG: w = 8; break;
    case ORC_DEC: w = 8; break;
    case ORC_DATE: w = 8; break;
    case ORC_CHAR: w = size; break;
    default: return;
    }
    s->type[s->ncols] = type;
    s->size[s->ncols] = w;
    s->offset[s->ncols] = s->tuplesize;
    s->tuplesize += w;
    s->ncols++;
}

static int64_t get_num(const orc_schema* s, const void* tup, int col)
{
    const char* p = (const char*)tup + s->offset[col];
    if (s->size[col] == 4) {
        int32_t v;
        memcpy(&v, p, 4);
        return v;
    } else {
        int64_t v;
        memcpy(&v, p, 8);
        return v;
    }
}

static double get_dec(const orc_schema* s, const void* tup, int col)
{
    double v;
    memcpy(&v, (const char*)tup + s->offset[col], 8);
    return v;
}

/* Typed equality of a column pair, as Comparator::Equal / NotEqual with
 * mixed int/long/date compares by value (comparator.cpp:43-413). */
static int cols_equal(const orc_schema* a, const void* ta, int ca, const orc_schema* b, const void* tb, int cb)
{
    int ta_ = a->type[ca], tb_ = b->type[cb];
    if (ta_ == ORC_DEC || tb_ == ORC_DEC) {
        double x = ta_ == ORC_DEC ? get_dec(a, ta, ca) : (double)get_num(a, ta, ca);
        double y = tb_ == ORC_DEC ? get_dec(b, tb, cb) : (double)get_num(b, tb, cb);
        return x == y;
    }
    if (ta_ == ORC_CHAR || tb_ == ORC_CHAR) {
        int n = a->size[ca] < b->size[cb] ? a->size[ca] : b->size[cb];
        return strncmp((const char*)ta + a->offset[ca], (const char*)tb + b->offset[cb], n) == 0;
    }
    return get_num(a, ta, ca) == get_num(b, tb, cb);
}

/* ===================================================================== */
/* hash functions                                                          */
/* ===================================================================== */
/* hash.cpp:47-56, with the same arithmetic shift of a negative int. */
unsigned orc_getlogarithm(unsigned k)
{
    --k;
    int m = (int)(1u << 31);
    for (unsigned i = 0; i < 32; ++i) {
        if ((unsigned)(m >> i) & k)
            return 32 - i;
    }
    return 0;
}

/* ValueHasher::numericalize (hash.h:215-234): 4 or 8 bytes only. */
static int numericalize(const void* p, int size, int64_t* out)
{
    if (size == 4) {
        int32_t v;
        memcpy(&v, p, 4);
        *out = v;
        return 0;
    }
    if (size == 8) {
        memcpy(out, p, 8);
        return 0;
    }
    return -1;
}

int orc_hasher_init(orc_hasher* h, int fn, unsigned buckets, int64_t rmin, int64_t rmax, int skipbits,
                    const orc_schema* s, int fieldmin, int fieldmax)
{
    memset(h, 0, sizeof *h);
    h->fn = fn;
    if (fn == ORC_H_ALWAYSZERO) { /* hash.cpp:108-112 */
        h->buckets = 1;
        return 0;
    }
    if (buckets == 0) return -1; /* MissingParameterException, hash.cpp:61-64 */
    if (fn != ORC_H_BYTES && fn != ORC_H_TPCHQ1MAGIC) {
        if (fieldmin != fieldmax) return -2; /* composite value hash */
        int t = s->type[fieldmin];
        if (t != ORC_INT && t != ORC_LONG && t != ORC_DATE) return -3; /* hash.cpp:158-168 */
    }
    /* range bounds are read as 32-bit int from the plan (hash.cpp:178-186) */
    rmin = (int32_t)rmin;
    rmax = (int32_t)rmax;
    unsigned lg = orc_getlogarithm(buckets);
    switch (fn) {
    case ORC_H_MODULO: /* hash.h:266-293 */
    case ORC_H_WILLIS:
    case ORC_H_TPCHORDERKEY:
        h->k = (int64_t)((1u << lg) - 1);
        h->buckets = (unsigned)h->k + 1;
        break;
    case ORC_H_PARAMMODULO: /* hash.h:301-338 */
    case ORC_H_KNUTH:       /* hash.h:348-369 (offset ignored by hash()) */
        h->k = (int64_t)((1u << lg) - 1);
        h->k <<= skipbits;
        h->skipbits = skipbits;
        h->min = (uint32_t)rmin; /* read as unsigned int offset */
        h->buckets = (unsigned)(h->k >> skipbits) + 1;
        break;
    case ORC_H_RANGE: /* hash.h:240-261 */
        h->k = lg;
        h->min = rmin;
        h->max = rmax;
        h->buckets = 1u << lg;
        break;
    case ORC_H_EXACTRANGE: /* hash.h:444-479 */
        h->min = rmin;
        h->max = rmax;
        h->k = buckets;
        h->bucketrange = (h->max - h->min + h->k) / h->k;
        h->buckets = buckets;
        break;
    case ORC_H_BYTES: /* hash.h:150-193 */
        h->k = lg;
        h->buckets = 1u << lg;
        break;
    case ORC_H_TPCHQ1MAGIC: /* hash.h:129-140: four buckets whatever `buckets` says */
        h->k = 2;
        h->buckets = 4;
        break;
    default:
        return -4;
    }
    /* field span (hash.cpp:223-233) */
    h->offset = s->offset[fieldmin];
    h->size = 0;
    for (int i = fieldmin; i <= fieldmax; ++i) h->size += s->size[i];
    return 0;
}

static unsigned willis(int64_t value, int64_t mask)
{
    /* hash.h:386-401: signed shifts are arithmetic, additions wrap */
    uint64_t l = (uint64_t)value;
    l = (~l) + (l << 21);
    l = l ^ (uint64_t)((int64_t)l >> 24);
    l = (l + (l << 3)) + (l << 8);
    l = l ^ (uint64_t)((int64_t)l >> 14);
    l = (l + (l << 2)) + (l << 4);
    l = l ^ (uint64_t)((int64_t)l >> 28);
    l = l + (l << 31);
    int64_t sl = (int64_t)l;
    sl = (sl > 0) ? sl : (int64_t)(0 - (uint64_t)sl);
    return (unsigned)(sl & mask);
}

unsigned orc_hash_value(const orc_hasher* h, int64_t v)
{
    switch (h->fn) {
    case ORC_H_MODULO: return (unsigned)(v & h->k);
    case ORC_H_PARAMMODULO: return (unsigned)(((v - h->min) & h->k) >> h->skipbits);
    case ORC_H_KNUTH: return (unsigned)((((uint64_t)v * 2654435761ull) & (uint64_t)h->k) >> h->skipbits);
    case ORC_H_RANGE: {
        int64_t val = v - h->min;
        val = (int64_t)((uint64_t)val << h->k);
        return (unsigned)(val / (h->max - h->min + 1));
    }
    case ORC_H_EXACTRANGE: return (unsigned)((v - h->min) / h->bucketrange);
    case ORC_H_ALWAYSZERO: return 0;
    case ORC_H_WILLIS: return willis(v, h->k);
    case ORC_H_TPCHORDERKEY: return (unsigned)((((v >> 2) & ~7LL) | (v & 7)) & h->k);
    default: return 0;
    }
}

/* ByteHasher::hash (hash.h:157-189): FNV-1a 64 with signed-char XOR. */
static unsigned bytes_hash(const orc_hasher* h, const void* start, int size)
{
    uint64_t hash = 14695981039346656037ull;
    const signed char* p = (const signed char*)start;
    for (int i = 0; i < size; ++i) {
        hash ^= (uint64_t)(int64_t)p[i];
        hash += (hash << 1) + (hash << 4) + (hash << 5) + (hash << 7) + (hash << 8) + (hash << 40);
    }
    uint64_t mask = (h->k >= 64) ? ~0ull : ((1ull << h->k) - 1);
    hash = ((uint64_t)((int64_t)hash >> h->k) ^ hash) & mask;
    return (unsigned)hash;
}

unsigned orc_hash_tuple(const orc_hasher* h, const void* tuple)
{
    const char* p = (const char*)tuple + h->offset;
    if (h->fn == ORC_H_ALWAYSZERO) return 0;
    if (h->fn == ORC_H_BYTES) return bytes_hash(h, p, h->size);
    if (h->fn == ORC_H_TPCHQ1MAGIC) { /* hash.h:136-140: bits of the first 4 bytes of the span */
        uint32_t v;
        memcpy(&v, p, 4);
        return (((v >> 4) | (v >> 16)) & 0x1u) | ((v >> 1) & 0x2u);
    }
    int64_t v = 0;
    if (numericalize(p, h->size, &v) != 0) return 0;
    return orc_hash_value(h, v);
}

int64_t orc_exactrange_minimum(const orc_hasher* h, unsigned bucket)
{
    if ((int64_t)bucket == h->k) return h->max + 1;
    return (int64_t)bucket * h->bucketrange + h->min;
}

/* hash.cpp:237-267: bits split MSB-first; the last pass takes the rest. */
int orc_generate(unsigned buckets, int skipbits, unsigned passes, unsigned* out_buckets, int* out_skip)
{
    /* buckets() of the parameterised hasher is the power of two 2^lg, and
     * totalbitsset = getlogarithm(buckets() - 1) */
    unsigned lg = orc_getlogarithm(buckets);
    unsigned total = orc_getlogarithm((1u << lg) - 1);
    unsigned per = total / passes;
    unsigned i;
    for (i = 0; i + 1 < passes; ++i) {
        out_buckets[i] = 1u << per;
        out_skip[i] = skipbits + (int)total - (int)((i + 1) * per);
    }
    per = total - (passes - 1) * per;
    out_buckets[passes - 1] = 1u << per;
    out_skip[passes - 1] = skipbits;
    return (int)passes;
}

/* ===================================================================== */
/* chained-bucket hash table (util/hashtable.h:44-281, .cpp:57-309)        */
/* ===================================================================== */
typedef struct orc_bh {
    volatile int lock;
    unsigned used; /* bytes used, 0..bucksize (uint16 in the reference) */
    struct orc_bh* next;
} orc_bh;

struct orc_htable {
    unsigned nbuckets, bucksize, tuplesize;
    char* base;
    size_t stride;
    volatile uint64_t spills;
    pthread_mutex_t arena_mu;
    char** arenas;
    size_t narenas, arena_cap, arena_used;
};

static orc_bh* header(orc_htable* t, unsigned b)
{
    return (orc_bh*)(t->base + (size_t)b * t->stride);
}

orc_htable* orc_ht_create(unsigned nbuckets, unsigned bucksize, unsigned tuplesize)
{
    if (bucksize > 65535 || tuplesize > bucksize) return NULL; /* `used` is a uint16 */
    orc_htable* t = (orc_htable*)calloc(1, sizeof *t);
    t->nbuckets = nbuckets;
    t->bucksize = bucksize;
    t->tuplesize = tuplesize;
    t->stride = (sizeof(orc_bh) + bucksize + 15) & ~(size_t)15;
    t->base = (char*)calloc(nbuckets, t->stride); /* bucketclear: used = 0, next = 0 */
    pthread_mutex_init(&t->arena_mu, NULL);
    return t;
}

void orc_ht_destroy(orc_htable* t)
{
    if (!t) return;
    for (size_t i = 0; i < t->narenas; ++i) free(t->arenas[i]);
    free(t->arenas);
    free(t->base);
    pthread_mutex_destroy(&t->arena_mu);
    free(t);
}

/* Overflow chunks come from a bump arena, like the 1 GiB lookaside arena of
 * numaallocate_local (util/numaallocate.cpp:264-273). */
static orc_bh* new_chunk(orc_htable* t)
{
    size_t sz = (sizeof(orc_bh) + t->bucksize + 15) & ~(size_t)15;
    pthread_mutex_lock(&t->arena_mu);
    if (t->narenas == 0 || t->arena_used + sz > t->arena_cap) {
        size_t cap = (size_t)64 << 20;
        if (cap < sz) cap = sz;
        t->arenas = (char**)realloc(t->arenas, (t->narenas + 1) * sizeof(char*));
        t->arenas[t->narenas++] = (char*)malloc(cap);
        t->arena_cap = cap;
        t->arena_used = 0;
    }
    orc_bh* c = (orc_bh*)(t->arenas[t->narenas - 1] + t->arena_used);
    t->arena_used += sz;
    pthread_mutex_unlock(&t->arena_mu);
    c->lock = 0;
    c->used = 0;
    c->next = NULL;
    return c;
}

/* HashTable::allocate (util/hashtable.cpp:230-273): first chunk with room
 * in chain order, else a new chunk linked at the tail. */
void* orc_ht_allocate(orc_htable* t, unsigned bucket)
{
    orc_bh* last = NULL;
    for (orc_bh* bh = header(t, bucket); bh; bh = bh->next) {
        if (bh->used + t->tuplesize <= t->bucksize) {
            void* loc = (char*)bh + sizeof(orc_bh) + bh->used;
            bh->used += t->tuplesize;
            return loc;
        }
        last = bh;
    }
    __atomic_add_fetch(&t->spills, 1, __ATOMIC_RELAXED);
    orc_bh* n = new_chunk(t);
    n->used = t->tuplesize;
    last->next = n;
    return (char*)n + sizeof(orc_bh);
}

static void lock_bucket(orc_htable* t, unsigned b)
{
    orc_bh* bh = header(t, b);
    while (__atomic_exchange_n(&bh->lock, 1, __ATOMIC_ACQUIRE))
        while (bh->lock) __builtin_ia32_pause();
}
static void unlock_bucket(orc_htable* t, unsigned b)
{
    __atomic_store_n(&header(t, b)->lock, 0, __ATOMIC_RELEASE);
}

/* HashTable::Iterator::next / placeIterator (util/hashtable.h:116-161). */
unsigned orc_ht_bucket_tuples(orc_htable* t, unsigned bucket, void** out, unsigned cap)
{
    unsigned n = 0;
    for (orc_bh* bh = header(t, bucket); bh; bh = bh->next) {
        if (bh != header(t, bucket) && bh->used == 0) break; /* empty chunk ends the walk */
        char* d = (char*)bh + sizeof(orc_bh);
        for (unsigned off = 0; off < bh->used; off += t->tuplesize) {
            if (n < cap) out[n] = d + off;
            ++n;
        }
    }
    return n;
}

uint64_t orc_ht_spills(const orc_htable* t) { return t->spills; }

/* HashTable::serialize / deserialize (util/hashtable.cpp:99-201) for a table
 * of one partition: the image of its bucket array, each bucket a 16-byte
 * BucketHeader (util/hashtable.h:244-267: the 1-byte lock at offset 0, the
 * uint16 `used` at offset 2, the overflow pointer at offset 8, NULL) followed
 * by `bucksize` bytes of tuples, buckets (16 + bucksize) bytes apart, written
 * into a file (the reference's creatememsegment makes /dev/shm segments or
 * files the same way).  A table that spilled has overflow pointers that
 * cannot be written: refused, as the reference asserts spills == 0.  Bytes
 * the reference leaves unwritten (header padding, bucket space past `used`)
 * are zero here and ignored on reading. */
#define ORC_SEG_BH 16
int orc_ht_serialize(orc_htable* t, const char* path)
{
    if (t->spills) return -1;
    FILE* f = fopen(path, "wb");
    if (!f) return -2;
    const size_t stride = ORC_SEG_BH + t->bucksize;
    char* buf = (char*)calloc(1, stride);
    int rc = 0;
    for (unsigned b = 0; b < t->nbuckets && !rc; ++b) {
        orc_bh* bh = header(t, b);
        memset(buf, 0, stride);
        uint16_t used = (uint16_t)bh->used;
        memcpy(buf + 2, &used, 2);
        memcpy(buf + ORC_SEG_BH, (char*)bh + sizeof(orc_bh), bh->used);
        if (fwrite(buf, 1, stride, f) != stride) rc = -3;
    }
    free(buf);
    if (fclose(f) != 0 && !rc) rc = -3;
    return rc;
}

int orc_ht_deserialize(orc_htable* t, const char* path)
{
    FILE* f = fopen(path, "rb");
    if (!f) return -2;
    const size_t stride = ORC_SEG_BH + t->bucksize;
    char* buf = (char*)malloc(stride);
    int rc = 0;
    for (unsigned b = 0; b < t->nbuckets && !rc; ++b) {
        if (fread(buf, 1, stride, f) != stride) {
            rc = -3;
            break;
        }
        uint16_t used;
        uint64_t next;
        memcpy(&used, buf + 2, 2);
        memcpy(&next, buf + 8, 8);
        if (used > t->bucksize || next != 0) {
            rc = -4; /* not an image of a spill-free table of this geometry */
            break;
        }
        orc_bh* bh = header(t, b);
        bh->lock = 0;
        bh->used = used;
        bh->next = NULL;
        memcpy((char*)bh + sizeof(orc_bh), buf + ORC_SEG_BH, used);
    }
    free(buf);
    fclose(f);
    return rc;
}

/* ===================================================================== */
/* hash join (operators/join.cpp:141-675)                                  */
/* ===================================================================== */
void orc_join_outschema(const orc_join_desc* d, orc_schema* out)
{
    memset(out, 0, sizeof *out);
    for (int i = 0; i < d->nproj; ++i) {
        const orc_schema* s = d->proj_side[i] ? &d->probe : &d->build;
        int c = d->proj_col[i];
        orc_schema_add(out, s->type[c], s->size[c]);
    }
}

/* sbuild = [build key, projected build columns] (join.cpp:278-286). */
static void make_sbuild(const orc_join_desc* d, orc_schema* sb)
{
    memset(sb, 0, sizeof *sb);
    orc_schema_add(sb, d->build.type[d->buildjattr], d->build.size[d->buildjattr]);
    for (int i = 0; i < d->nproj; ++i)
        if (!d->proj_side[i]) orc_schema_add(sb, d->build.type[d->proj_col[i]], d->build.size[d->proj_col[i]]);
}

typedef struct join_shared {
    const orc_join_desc* d;
    orc_schema sbuild, out;
    orc_hasher bh, ph;
    orc_htable* ht;
    const char* build;
    const char* probe;
    uint64_t nbuild, nprobe;
    int threads;
    /* materialising */
    char** tout;
    uint64_t* tcount;
    uint64_t cap_per_thread;
    int overflow;
    /* count + sum */
    int sumcol;
    orc_count_sum* tcs;
    pthread_barrier_t bar;
} join_shared;

typedef struct join_arg {
    join_shared* sh;
    int tid;
} join_arg;

/* buildFromPage (join.cpp:641-667). */
static void build_slice(join_shared* sh, uint64_t lo, uint64_t hi)
{
    const orc_join_desc* d = sh->d;
    for (uint64_t i = lo; i < hi; ++i) {
        const char* tup = sh->build + i * d->build.tuplesize;
        unsigned b = orc_hash_tuple(&sh->bh, tup);
        lock_bucket(sh->ht, b);
        char* target = (char*)orc_ht_allocate(sh->ht, b);
        unlock_bucket(sh->ht, b);
        memcpy(target + sh->sbuild.offset[0], tup + d->build.offset[d->buildjattr], sh->sbuild.size[0]);
        int bc = 1;
        for (int j = 0; j < d->nproj; ++j) {
            if (d->proj_side[j]) continue;
            memcpy(target + sh->sbuild.offset[bc], tup + d->build.offset[d->proj_col[j]], sh->sbuild.size[bc]);
            ++bc;
        }
    }
}

/* constructOutputTuple (join.cpp:475-501). */
static void construct(join_shared* sh, const char* tb, const char* tp, char* out)
{
    const orc_join_desc* d = sh->d;
    int bc = 1;
    for (int j = 0; j < d->nproj; ++j) {
        if (!d->proj_side[j]) {
            memcpy(out + sh->out.offset[j], tb + sh->sbuild.offset[bc], sh->out.size[j]);
            ++bc;
        } else {
            memcpy(out + sh->out.offset[j], tp + d->probe.offset[d->proj_col[j]], sh->out.size[j]);
        }
    }
}

/* getNext probe loop (join.cpp:503-562): probe order x chain order. */
static void probe_slice(join_shared* sh, int tid, uint64_t lo, uint64_t hi)
{
    const orc_join_desc* d = sh->d;
    char outbuf[512];
    orc_count_sum cs = {0, 0, 0.0};
    int32_t i32 = 0;
    uint64_t nout = 0;
    for (uint64_t i = lo; i < hi; ++i) {
        const char* tp = sh->probe + i * d->probe.tuplesize;
        unsigned b = orc_hash_tuple(&sh->ph, tp);
        for (orc_bh* bh = header(sh->ht, b); bh; bh = bh->next) {
            const char* dd = (const char*)bh + sizeof(orc_bh);
            for (unsigned off = 0; off < bh->used; off += sh->ht->tuplesize) {
                const char* tb = dd + off;
                if (!cols_equal(&sh->sbuild, tb, 0, &d->probe, tp, d->probejattr)) continue;
                if (sh->tout) {
                    if (nout >= sh->cap_per_thread) {
                        sh->overflow = 1;
                        continue;
                    }
                    construct(sh, tb, tp, sh->tout[tid] + nout * sh->out.tuplesize);
                    ++nout;
                } else {
                    construct(sh, tb, tp, outbuf);
                    cs.count++;
                    int c = sh->sumcol;
                    if (c >= 0) {
                        if (sh->out.type[c] == ORC_DEC)
                            cs.dsum += get_dec(&sh->out, outbuf, c);
                        else if (sh->out.size[c] == 4)
                            i32 = (int32_t)((uint32_t)i32 + (uint32_t)(int32_t)get_num(&sh->out, outbuf, c));
                        else
                            cs.isum = (int64_t)((uint64_t)cs.isum + (uint64_t)get_num(&sh->out, outbuf, c));
                    }
                }
            }
        }
    }
    if (sh->tout) sh->tcount[tid] = nout;
    else {
        if (sh->sumcol >= 0 && sh->out.type[sh->sumcol] != ORC_DEC && sh->out.size[sh->sumcol] == 4)
            cs.isum = i32;
        sh->tcs[tid] = cs;
    }
}

static void* join_worker(void* p)
{
    join_arg* a = (join_arg*)p;
    join_shared* sh = a->sh;
    int T = sh->threads, t = a->tid;
    build_slice(sh, sh->nbuild * t / T, sh->nbuild * (t + 1) / T);
    pthread_barrier_wait(&sh->bar); /* group barrier after the build (join.cpp:434) */
    probe_slice(sh, t, sh->nprobe * t / T, sh->nprobe * (t + 1) / T);
    return NULL;
}

static int join_setup(join_shared* sh, const orc_join_desc* d, const void* build, uint64_t nbuild,
                      const void* probe, uint64_t nprobe)
{
    memset(sh, 0, sizeof *sh);
    sh->d = d;
    make_sbuild(d, &sh->sbuild);
    orc_join_outschema(d, &sh->out);
    if (orc_hasher_init(&sh->bh, d->hashfn, d->buckets, d->rmin, d->rmax, d->skipbits, &d->build, d->buildjattr,
                        d->buildjattr))
        return -1;
    if (orc_hasher_init(&sh->ph, d->hashfn, d->buckets, d->rmin, d->rmax, d->skipbits, &d->probe, d->probejattr,
                        d->probejattr))
        return -1;
    if (sh->bh.buckets != sh->ph.buckets) return -1;
    sh->ht = orc_ht_create(sh->bh.buckets, d->tuplesperbucket * sh->sbuild.tuplesize, sh->sbuild.tuplesize);
    if (!sh->ht) return -1;
    sh->build = (const char*)build;
    sh->probe = (const char*)probe;
    sh->nbuild = nbuild;
    sh->nprobe = nprobe;
    sh->threads = d->threads > 0 ? d->threads : 1;
    return 0;
}

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + ts.tv_nsec * 1e-9;
}

static void run_join(join_shared* sh)
{
    int T = sh->threads;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * T);
    join_arg* args = (join_arg*)malloc(sizeof(join_arg) * T);
    pthread_barrier_init(&sh->bar, NULL, (unsigned)T);
    for (int t = 0; t < T; ++t) {
        args[t].sh = sh;
        args[t].tid = t;
        pthread_create(&th[t], NULL, join_worker, &args[t]);
    }
    for (int t = 0; t < T; ++t) pthread_join(th[t], NULL);
    pthread_barrier_destroy(&sh->bar);
    free(th);
    free(args);
}

int64_t orc_hashjoin(const orc_join_desc* d, const void* build, uint64_t nbuild, const void* probe,
                     uint64_t nprobe, void* out, uint64_t cap, orc_join_stats* st)
{
    join_shared sh;
    if (join_setup(&sh, d, build, nbuild, probe, nprobe)) return -2;
    int T = sh.threads;
    sh.cap_per_thread = cap;
    sh.tout = (char**)calloc(T, sizeof(char*));
    sh.tcount = (uint64_t*)calloc(T, sizeof(uint64_t));
    for (int t = 0; t < T; ++t) sh.tout[t] = (char*)malloc((cap ? cap : 1) * (size_t)sh.out.tuplesize);
    double t0 = now_s();
    run_join(&sh);
    double t1 = now_s();
    int64_t total = 0;
    if (!sh.overflow) {
        for (int t = 0; t < T; ++t) total += (int64_t)sh.tcount[t];
        if ((uint64_t)total > cap) sh.overflow = 1;
    }
    if (!sh.overflow) {
        char* o = (char*)out;
        for (int t = 0; t < T; ++t) {
            memcpy(o, sh.tout[t], sh.tcount[t] * (size_t)sh.out.tuplesize);
            o += sh.tcount[t] * (size_t)sh.out.tuplesize;
        }
    }
    if (st) {
        st->matches = (uint64_t)total;
        st->spills = orc_ht_spills(sh.ht);
        st->build_s = 0;
        st->probe_s = t1 - t0;
    }
    for (int t = 0; t < T; ++t) free(sh.tout[t]);
    free(sh.tout);
    free(sh.tcount);
    orc_ht_destroy(sh.ht);
    return sh.overflow ? -1 : total;
}

int orc_hashjoin_count_sum(const orc_join_desc* d, const void* build, uint64_t nbuild, const void* probe,
                           uint64_t nprobe, int sumcol, orc_count_sum* out, orc_join_stats* st)
{
    join_shared sh;
    if (join_setup(&sh, d, build, nbuild, probe, nprobe)) return -2;
    int T = sh.threads;
    sh.sumcol = sumcol;
    sh.tcs = (orc_count_sum*)calloc(T, sizeof(orc_count_sum));
    double t0 = now_s();
    run_join(&sh);
    double t1 = now_s();
    /* final aggregate above merge: fold the per-thread partials in thread
     * order, in the column's type */
    orc_count_sum r = {0, 0, 0.0};
    int32_t i32 = 0;
    int is32 = sumcol >= 0 && sh.out.type[sumcol] != ORC_DEC && sh.out.size[sumcol] == 4;
    for (int t = 0; t < T; ++t) {
        r.count += sh.tcs[t].count;
        r.dsum += sh.tcs[t].dsum;
        if (is32) i32 = (int32_t)((uint32_t)i32 + (uint32_t)sh.tcs[t].isum);
        else r.isum = (int64_t)((uint64_t)r.isum + (uint64_t)sh.tcs[t].isum);
    }
    if (is32) r.isum = i32;
    *out = r;
    if (st) {
        st->matches = r.count;
        st->spills = orc_ht_spills(sh.ht);
        st->probe_s = t1 - t0;
    }
    free(sh.tcs);
    orc_ht_destroy(sh.ht);
    return 0;
}

/* ===================================================================== */
/* aggregation (operators/genericaggregate.cpp:51-487, aggregatesum.cpp)   */
/* ===================================================================== */
void orc_agg_outschema(const orc_agg_desc* d, orc_schema* out)
{
    memset(out, 0, sizeof *out);
    for (int i = 0; i < d->nfields; ++i) orc_schema_add(out, d->in.type[d->fields[i]], d->in.size[d->fields[i]]);
    if (d->sumfield >= 0) /* accumulator type = input column type (aggregatesum.cpp:46-49) */
        orc_schema_add(out, d->in.type[d->sumfield], d->in.size[d->sumfield]);
    else /* AggregateCount: CT_LONG (aggregatecount.cpp:37-42) */
        orc_schema_add(out, ORC_LONG, 8);
}

/* foldstart / fold (aggregatesum.cpp:53-107, aggregatecount.cpp:44-52). */
static void fold(const orc_agg_desc* d, const orc_schema* os, char* acc, const char* tup, int start)
{
    if (d->sumfield < 0) {
        int64_t c;
        memcpy(&c, acc, 8);
        c = start ? 1 : c + 1;
        memcpy(acc, &c, 8);
        return;
    }
    int col = d->sumfield;
    const char* src = tup + d->in.offset[col];
    if (d->in.type[col] == ORC_DEC) {
        double v, a;
        memcpy(&v, src, 8);
        memcpy(&a, acc, 8);
        a = start ? v : a + v;
        memcpy(acc, &a, 8);
    } else if (d->in.size[col] == 4) {
        uint32_t v, a;
        memcpy(&v, src, 4);
        memcpy(&a, acc, 4);
        a = start ? v : a + v; /* int accumulator wraps mod 2^32 */
        memcpy(acc, &a, 4);
    } else {
        uint64_t v, a;
        memcpy(&v, src, 8);
        memcpy(&a, acc, 8);
        a = start ? v : a + v;
        memcpy(acc, &a, 8);
    }
    (void)os;
}

/* GenericAggregate::remember (genericaggregate.cpp:240-282). */
static void remember(const orc_agg_desc* d, const orc_schema* os, const orc_hasher* h, orc_htable* ht,
                     const char* tup, int locked)
{
    unsigned b = orc_hash_tuple(h, tup);
    if (locked) lock_bucket(ht, b);
    char* found = NULL;
    orc_bh* head = header(ht, b);
    for (orc_bh* bh = head; bh && !found; bh = bh->next) {
        if (bh != head && bh->used == 0) break;
        char* dd = (char*)bh + sizeof(orc_bh);
        for (unsigned off = 0; off < bh->used; off += ht->tuplesize) {
            char* cand = dd + off;
            int eq = 1;
            for (int i = 0; i < d->nfields && eq; ++i) eq = cols_equal(os, cand, i, &d->in, tup, d->fields[i]);
            if (eq) {
                found = cand;
                break;
            }
        }
    }
    if (found) {
        fold(d, os, found + os->offset[d->nfields], tup, 0);
    } else {
        char* c = (char*)orc_ht_allocate(ht, b);
        for (int i = 0; i < d->nfields; ++i) memcpy(c + os->offset[i], tup + d->in.offset[d->fields[i]], os->size[i]);
        fold(d, os, c + os->offset[d->nfields], tup, 1);
    }
    if (locked) unlock_bucket(ht, b);
}

typedef struct agg_shared {
    const orc_agg_desc* d;
    orc_schema os;
    orc_hasher h;
    orc_htable** tables;
    const char* rows;
    uint64_t nrows;
    int threads;
} agg_shared;
typedef struct agg_arg {
    agg_shared* sh;
    int tid;
} agg_arg;

static void* agg_worker(void* p)
{
    agg_arg* a = (agg_arg*)p;
    agg_shared* sh = a->sh;
    int T = sh->threads, t = a->tid;
    orc_htable* ht = sh->d->global ? sh->tables[0] : sh->tables[t];
    uint64_t lo = sh->nrows * t / T, hi = sh->nrows * (t + 1) / T;
    for (uint64_t i = lo; i < hi; ++i)
        remember(sh->d, &sh->os, &sh->h, ht, sh->rows + i * sh->d->in.tuplesize, sh->d->global);
    return NULL;
}

int64_t orc_aggregate(const orc_agg_desc* d, const void* rows, uint64_t nrows, void* out, uint64_t cap)
{
    agg_shared sh;
    memset(&sh, 0, sizeof sh);
    sh.d = d;
    orc_agg_outschema(d, &sh.os);
    int fn = d->nfields == 0 ? ORC_H_ALWAYSZERO : d->hashfn; /* genericaggregate.cpp:80-87 */
    int fmin = d->nfields ? d->fields[0] : 0, fmax = fmin;
    if (fn == ORC_H_BYTES && d->nfields > 1) fmax = d->fields[d->nfields - 1];
    if (orc_hasher_init(&sh.h, fn, d->buckets, d->rmin, d->rmax, 0, &d->in, fmin, fmax)) return -2;
    sh.threads = d->threads > 0 ? d->threads : 1;
    int ntab = d->global ? 1 : sh.threads;
    sh.tables = (orc_htable**)calloc(ntab, sizeof(orc_htable*));
    for (int i = 0; i < ntab; ++i) /* 4 tuples per bucket (genericaggregate.cpp:143-148, :181-189) */
        sh.tables[i] = orc_ht_create(sh.h.buckets, 4 * sh.os.tuplesize, sh.os.tuplesize);
    sh.rows = (const char*)rows;
    sh.nrows = nrows;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * sh.threads);
    agg_arg* args = (agg_arg*)malloc(sizeof(agg_arg) * sh.threads);
    for (int t = 0; t < sh.threads; ++t) {
        args[t].sh = &sh;
        args[t].tid = t;
        pthread_create(&th[t], NULL, agg_worker, &args[t]);
    }
    for (int t = 0; t < sh.threads; ++t) pthread_join(th[t], NULL);
    /* emit (genericaggregate.cpp:419-487): per table, buckets in order,
     * chain order within a bucket */
    int64_t n = 0;
    char* o = (char*)out;
    for (int i = 0; i < ntab; ++i) {
        orc_htable* ht = sh.tables[i];
        for (unsigned b = 0; b < ht->nbuckets; ++b) {
            orc_bh* head = header(ht, b);
            for (orc_bh* bh = head; bh; bh = bh->next) {
                if (bh != head && bh->used == 0) break;
                char* dd = (char*)bh + sizeof(orc_bh);
                for (unsigned off = 0; off < bh->used; off += ht->tuplesize) {
                    if ((uint64_t)n < cap) memcpy(o + (size_t)n * sh.os.tuplesize, dd + off, sh.os.tuplesize);
                    ++n;
                }
            }
        }
        orc_ht_destroy(ht);
    }
    free(sh.tables);
    free(th);
    free(args);
    return (uint64_t)n > cap ? -1 : n;
}

/* ===================================================================== */
/* partition (operators/partition.cpp:166-306): histogram, column-wise     */
/* prefix over threads, scatter in input order.                            */
/* ===================================================================== */
int orc_partition(const orc_schema* s, int attr, int64_t rmin, int64_t rmax, unsigned buckets, const void* rows,
                  uint64_t nrows, void* out, uint64_t* counts)
{
    orc_hasher h;
    if (orc_hasher_init(&h, ORC_H_EXACTRANGE, buckets, rmin, rmax, 0, s, attr, attr)) return -2;
    /* one producing thread per partition as PartitionOp does (buckets ==
     * threads); thread t owns input slice t */
    unsigned T = buckets;
    uint64_t* hist = (uint64_t*)calloc((size_t)T * buckets, sizeof(uint64_t));
    const char* r = (const char*)rows;
    int ts = s->tuplesize;
    for (unsigned t = 0; t < T; ++t)
        for (uint64_t i = nrows * t / T; i < nrows * (t + 1) / T; ++i) {
            unsigned p = orc_hash_tuple(&h, r + i * ts);
            if (p >= buckets) {
                free(hist);
                return -3;
            }
            hist[(size_t)t * buckets + p]++;
        }
    uint64_t* start = (uint64_t*)calloc((size_t)T * buckets, sizeof(uint64_t));
    uint64_t base = 0;
    for (unsigned p = 0; p < buckets; ++p) {
        counts[p] = 0;
        for (unsigned t = 0; t < T; ++t) {
            start[(size_t)t * buckets + p] = base;
            base += hist[(size_t)t * buckets + p];
            counts[p] += hist[(size_t)t * buckets + p];
        }
    }
    char* o = (char*)out;
    for (unsigned t = 0; t < T; ++t)
        for (uint64_t i = nrows * t / T; i < nrows * (t + 1) / T; ++i) {
            unsigned p = orc_hash_tuple(&h, r + i * ts);
            memcpy(o + start[(size_t)t * buckets + p]++ * ts, r + i * ts, ts);
        }
    free(hist);
    free(start);
    return 0;
}

/* ===================================================================== */
/* generators and closed forms                                             */
/* ===================================================================== */
void orc_zipf_cdf(uint64_t n, double s, double* cdf) { pgpu_zipf_cdf_fill(cdf, n, s); }

void orc_gen_c1(uint64_t seed, uint64_t nR, uint64_t nS, int32_t* rkey, int32_t* rpay, int32_t* sfk,
                int32_t* spay)
{
    for (uint64_t i = 0; i < nR; ++i) {
        if (rkey) rkey[i] = pgpu_c1_rkey(seed, i, nR);
        if (rpay) rpay[i] = pgpu_c1_rpay(seed, i);
    }
    for (uint64_t j = 0; j < nS; ++j) {
        if (sfk) sfk[j] = pgpu_c1_sfk(seed, j, nR);
        if (spay) spay[j] = pgpu_c1_spay(seed, j);
    }
}

void orc_gen_c2(uint64_t seed, uint64_t sub, uint64_t nR, uint64_t nS, double zipf_s, int64_t* rkey,
                int64_t* rpay, int64_t* sfk)
{
    for (uint64_t i = 0; i < nR; ++i) {
        if (rkey) rkey[i] = pgpu_c2_rkey(seed, sub, i);
        if (rpay) rpay[i] = pgpu_c2_rpay(seed, sub, i);
    }
    if (!sfk) return;
    double* cdf = NULL;
    if (zipf_s > 0) {
        cdf = (double*)malloc(sizeof(double) * nR);
        pgpu_zipf_cdf_fill(cdf, nR, zipf_s);
    }
    for (uint64_t j = 0; j < nS; ++j) {
        uint64_t row = cdf ? pgpu_c2_srow_zipf(seed, sub, j, nR, cdf) : pgpu_c2_srow_uniform(seed, sub, j, nR);
        sfk[j] = pgpu_c2_rkey(seed, sub, row);
    }
    free(cdf);
}

void orc_gen_c3(uint64_t seed, uint64_t nD, uint64_t nF, uint64_t norders, int32_t* d_orderkey,
                int32_t* d_custkey, int32_t* f_orderkey, double* f_price)
{
    for (uint64_t i = 0; i < nD; ++i) {
        if (d_orderkey) d_orderkey[i] = pgpu_c3_d_orderkey(seed, i, nD);
        if (d_custkey) d_custkey[i] = pgpu_c3_d_custkey(seed, i, PGPU_C3_CUSTS);
    }
    for (uint64_t j = 0; j < nF; ++j) {
        if (f_orderkey) f_orderkey[j] = pgpu_c3_f_orderkey(seed, j, norders);
        if (f_price) f_price[j] = pgpu_c3_f_price(seed, j);
    }
}

typedef struct closed_arg {
    uint64_t seed, sub, nR, lo, hi, first; /* rows first + [lo, hi) */
    const double* cdf;
    const int32_t* inv; /* C1: key -> row */
    int kind;
    orc_cs64 r;
} closed_arg;

static void* closed_worker(void* p)
{
    closed_arg* a = (closed_arg*)p;
    orc_cs64 r = {0, 0, 0};
    for (uint64_t j = a->first + a->lo; j < a->first + a->hi; ++j) {
        if (a->kind == 1) {
            int32_t fk = pgpu_c1_sfk(a->seed, j, a->nR);
            int32_t row = a->inv[fk];
            r.count++;
            r.sum_build += (uint64_t)(int64_t)pgpu_c1_rpay(a->seed, (uint64_t)row);
            r.sum_probe += (uint64_t)(int64_t)pgpu_c1_spay(a->seed, j);
        } else {
            uint64_t row = a->cdf ? pgpu_c2_srow_zipf(a->seed, a->sub, j, a->nR, a->cdf)
                                  : pgpu_c2_srow_uniform(a->seed, a->sub, j, a->nR);
            r.count++;
            r.sum_build += (uint64_t)pgpu_c2_rpay(a->seed, a->sub, row);
        }
    }
    a->r = r;
    return NULL;
}

static void run_closed(closed_arg* proto, uint64_t nS, int threads, orc_cs64* out)
{
    if (threads < 1) threads = 1;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * threads);
    closed_arg* args = (closed_arg*)malloc(sizeof(closed_arg) * threads);
    for (int t = 0; t < threads; ++t) {
        args[t] = *proto;
        args[t].lo = nS * t / threads;
        args[t].hi = nS * (t + 1) / threads;
        pthread_create(&th[t], NULL, closed_worker, &args[t]);
    }
    orc_cs64 r = {0, 0, 0};
    for (int t = 0; t < threads; ++t) {
        pthread_join(th[t], NULL);
        r.count += args[t].r.count;
        r.sum_build += args[t].r.sum_build;
        r.sum_probe += args[t].r.sum_probe;
    }
    *out = r;
    free(th);
    free(args);
}

void orc_c1_closed(uint64_t seed, uint64_t nR, uint64_t nS, int threads, orc_cs64* out)
{
    int32_t* inv = (int32_t*)malloc(sizeof(int32_t) * nR);
    for (uint64_t i = 0; i < nR; ++i) inv[pgpu_c1_rkey(seed, i, nR)] = (int32_t)i;
    closed_arg a;
    memset(&a, 0, sizeof a);
    a.seed = seed;
    a.nR = nR;
    a.inv = inv;
    a.kind = 1;
    run_closed(&a, nS, threads, out);
    free(inv);
}

void orc_c2_closed(uint64_t seed, uint64_t sub, uint64_t nR, uint64_t nS, double zipf_s, int threads,
                   orc_cs64* out)
{
    double* cdf = NULL;
    if (zipf_s > 0) {
        cdf = (double*)malloc(sizeof(double) * nR);
        pgpu_zipf_cdf_fill(cdf, nR, zipf_s);
    }
    closed_arg a;
    memset(&a, 0, sizeof a);
    a.seed = seed;
    a.sub = sub;
    a.nR = nR;
    a.cdf = cdf;
    a.kind = 2;
    run_closed(&a, nS, threads, out);
    free(cdf);
}

void orc_c5_closed(uint64_t seed, uint64_t table, uint64_t n, uint64_t first, uint64_t m, int threads, orc_cs64* out)
{
    closed_arg a;
    memset(&a, 0, sizeof a);
    a.seed = seed;
    a.sub = table;
    a.nR = n;
    a.first = first;
    a.kind = 2;
    run_closed(&a, m, threads, out);
}

typedef struct c3_arg {
    uint64_t seed, lo, hi, norders, nD;
    const int32_t* cust_of_order;
    int64_t* cents;
    uint64_t joined;
} c3_arg;

static void* c3_worker(void* p)
{
    c3_arg* a = (c3_arg*)p;
    for (uint64_t j = a->lo; j < a->hi; ++j) {
        int32_t ok = pgpu_c3_f_orderkey(a->seed, j, a->norders);
        if (ok < 1 || (uint64_t)ok > a->nD) continue; /* D holds orderkeys 1..nD */
        a->cents[a->cust_of_order[ok]] += pgpu_c3_f_cents(a->seed, j);
        a->joined++;
    }
    return NULL;
}

uint64_t orc_c3_closed(uint64_t seed, uint64_t nD, uint64_t nF, uint64_t norders, uint64_t ncust, int threads,
                       int64_t* cents)
{
    int32_t* cust_of_order = (int32_t*)malloc(sizeof(int32_t) * (nD + 1));
    for (uint64_t i = 0; i < nD; ++i)
        cust_of_order[pgpu_c3_d_orderkey(seed, i, nD)] = pgpu_c3_d_custkey(seed, i, PGPU_C3_CUSTS);
    if (threads < 1) threads = 1;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * threads);
    c3_arg* args = (c3_arg*)calloc(threads, sizeof(c3_arg));
    for (int t = 0; t < threads; ++t) {
        args[t].seed = seed;
        args[t].lo = nF * t / threads;
        args[t].hi = nF * (t + 1) / threads;
        args[t].norders = norders;
        args[t].nD = nD;
        args[t].cust_of_order = cust_of_order;
        args[t].cents = (int64_t*)calloc(PGPU_C3_CUSTS + 1, sizeof(int64_t));
        pthread_create(&th[t], NULL, c3_worker, &args[t]);
    }
    uint64_t joined = 0;
    memset(cents, 0, sizeof(int64_t) * (ncust + 1));
    for (int t = 0; t < threads; ++t) {
        pthread_join(th[t], NULL);
        for (uint64_t c = 0; c <= ncust && c <= PGPU_C3_CUSTS; ++c) cents[c] += args[t].cents[c];
        joined += args[t].joined;
        free(args[t].cents);
    }
    free(th);
    free(args);
    free(cust_of_order);
    return joined;
}

void orc_c4_closed(uint64_t seed, uint64_t join, uint64_t n, uint64_t nprobe, orc_cs64* out)
{
    int32_t* inv = (int32_t*)malloc(sizeof(int32_t) * n);
    for (uint64_t i = 0; i < n; ++i) inv[pgpu_c4_rkey(seed, join, i, n)] = (int32_t)i;
    orc_cs64 r = {0, 0, 0};
    for (uint64_t j = 0; j < nprobe; ++j) {
        int32_t fk = pgpu_c4_sfk(seed, join, j, n);
        if (fk < 0 || (uint64_t)fk >= n) continue;
        r.count++;
        r.sum_build += (uint64_t)(int64_t)pgpu_c4_rpay(seed, join, (uint64_t)inv[fk]);
        r.sum_probe += (uint64_t)(int64_t)pgpu_c4_spay(seed, join, j);
    }
    *out = r;
    free(inv);
}

void orc_gen_c4(uint64_t seed, uint64_t join, uint64_t n, uint64_t m, int32_t* rkey, int32_t* rpay, int32_t* sfk,
                int32_t* spay)
{
    for (uint64_t i = 0; i < n; ++i) {
        rkey[i] = pgpu_c4_rkey(seed, join, i, n);
        rpay[i] = pgpu_c4_rpay(seed, join, i);
    }
    for (uint64_t j = 0; j < m; ++j) {
        sfk[j] = pgpu_c4_sfk(seed, join, j, n);
        spay[j] = pgpu_c4_spay(seed, join, j);
    }
}

typedef struct {
    uint64_t seed, first, count;
    const uint64_t *n, *m;
    orc_cs64* out;
    uint64_t* next;
    pthread_mutex_t* mu;
} c4_many_arg;

static void* c4_many_worker(void* p)
{
    c4_many_arg* a = (c4_many_arg*)p;
    for (;;) {
        pthread_mutex_lock(a->mu);
        uint64_t i = (*a->next)++;
        pthread_mutex_unlock(a->mu);
        if (i >= a->count) break;
        orc_c4_closed(a->seed, a->first + i, a->n[i], a->m[i], &a->out[i]);
    }
    return NULL;
}

void orc_c4_closed_many(uint64_t seed, uint64_t first, uint64_t count, const uint64_t* n, const uint64_t* m,
                        int threads, orc_cs64* out)
{
    if (threads < 1) threads = 1;
    pthread_mutex_t mu;
    pthread_mutex_init(&mu, NULL);
    uint64_t next = 0;
    c4_many_arg arg = {seed, first, count, n, m, out, &next, &mu};
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
    for (int t = 0; t < threads; ++t) pthread_create(&th[t], NULL, c4_many_worker, &arg);
    for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
    free(th);
    pthread_mutex_destroy(&mu);
}

/* ===================================================================== */
/* timed multithreaded baseline (the queryhashjoin plan shape)             */
/* ===================================================================== */
double orc_baseline_join(int kb, const void* rkey, const void* rpay, uint64_t nR, const void* skey, uint64_t nS,
                         unsigned buckets, unsigned tuplesperbucket, int threads, orc_cs64* out)
{
    /* Rows are (key, payload) for the build and (key) for the probe, as
     * the parallel scans hand them over; the plan is hashjoin(modulo,
     * buckets, tuplesperbucket, projection [B$1]) -> partial aggregate
     * count + sum -> merge(threads) -> final aggregate. */
    orc_join_desc d;
    memset(&d, 0, sizeof d);
    int ct = kb == 4 ? ORC_INT : ORC_LONG;
    orc_schema_add(&d.build, ct, kb);
    orc_schema_add(&d.build, ct, kb);
    orc_schema_add(&d.probe, ct, kb);
    d.buildjattr = 0;
    d.probejattr = 0;
    d.nproj = 1;
    d.proj_side[0] = 0;
    d.proj_col[0] = 1;
    d.hashfn = ORC_H_MODULO;
    d.buckets = buckets;
    d.tuplesperbucket = tuplesperbucket;
    d.threads = threads;
    /* pack build rows (the scan's page format) outside the clock */
    char* brows = (char*)malloc((size_t)nR * 2 * kb);
    for (uint64_t i = 0; i < nR; ++i) {
        memcpy(brows + i * 2 * kb, (const char*)rkey + i * kb, kb);
        memcpy(brows + i * 2 * kb + kb, (const char*)rpay + i * kb, kb);
    }
    orc_count_sum cs;
    orc_join_stats st;
    double t0 = now_s();
    orc_hashjoin_count_sum(&d, brows, nR, skey, nS, 0, &cs, &st);
    double t1 = now_s();
    free(brows);
    out->count = cs.count;
    out->sum_build = (uint64_t)cs.isum;
    out->sum_probe = 0;
    return t1 - t0;
}
