/*
 * pp_oracle_tmpl.h -- the width- and precision-specific half of the oracle.
 *
 * TEST INFRASTRUCTURE ONLY (see pp_oracle.h).  Included once per (W, REAL)
 * instantiation by pp_oracle.c with W, REAL and SFX defined; every name is
 * suffixed by SFX.
 *
 * Storage follows the reference's canonical layout (pauli_sum/storage.rs:285-295):
 * buckets of a GF(2)-linear hash in ascending order, each strictly ascending in
 * the lexicographic key (x words, then z words).  Each bucket keeps an OR of its
 * keys (a superset support) so a later layer of the same call can skip buckets a
 * generator cannot touch (engine/bucketed/rotation_layer.rs:23-44, :250-254).
 */

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define T(name) CAT(name, SFX)

typedef struct {
    uint64_t x[W];
    uint64_t z[W];
} T(key);

typedef struct {
    T(key) * k;
    REAL* c;
    uint32_t len;
    uint32_t cap;
    uint64_t sx[W]; /* OR of the x words of every key (a superset) */
    uint64_t sz[W];
} T(bucket);

/* Lexicographic order of PauliString (pauli_string.rs:1-20: derived Ord on (x, z)). */
static inline int T(key_cmp)(const T(key) * a, const T(key) * b) {
    for (int w = 0; w < W; ++w)
        if (a->x[w] != b->x[w]) return a->x[w] < b->x[w] ? -1 : 1;
    for (int w = 0; w < W; ++w)
        if (a->z[w] != b->z[w]) return a->z[w] < b->z[w] ? -1 : 1;
    return 0;
}

/* commutes_with (pauli_string.rs:116-123): parity of popcount(x.gz ^ z.gx). */
static inline int T(anticommutes)(const T(key) * k, const uint64_t* gx, const uint64_t* gz) {
    uint64_t f = 0;
    for (int w = 0; w < W; ++w) f ^= (k->x[w] & gz[w]) ^ (k->z[w] & gx[w]);
    return __builtin_popcountll(f) & 1;
}

/* The phase exponent of k * g (pauli_string.rs:73-93), mod 4. */
static inline unsigned T(product_phase)(const T(key) * k, const uint64_t* gx, const uint64_t* gz) {
    unsigned d = 0;
    for (int w = 0; w < W; ++w) {
        uint64_t a = k->x[w], b = k->z[w], c = gx[w], e = gz[w];
        d += 2u * (unsigned)__builtin_popcountll(b & c);
        d += (unsigned)__builtin_popcountll(a & b);
        d += (unsigned)__builtin_popcountll(c & e);
        d -= (unsigned)__builtin_popcountll((a ^ c) & (b ^ e));
    }
    return d & 3u;
}

static inline unsigned T(weight)(const T(key) * k) {
    unsigned n = 0;
    for (int w = 0; w < W; ++w) n += (unsigned)__builtin_popcountll(k->x[w] | k->z[w]);
    return n;
}

/* keep_term of a built-in policy tree (truncation/builtin.rs:27-32, :50-54, :476-478, :604-606),
 * evaluated as the postfix program the device path also uses. */
static inline int T(keep)(const ppo_policy* p, const T(key) * k, REAL c) {
    if (p == NULL || p->n_nodes == 0) return 1;
    int stack[PPO_MAX_NODES];
    int sp = 0;
    for (int i = 0; i < p->n_nodes; ++i) {
        const ppo_node* nd = &p->nodes[i];
        switch (nd->op) {
            case PPO_NODE_TRUE: stack[sp++] = 1; break;
            case PPO_NODE_COEFF: {
                double v = (double)c;
                stack[sp++] = (nd->eps < 0.0) || (v * v > nd->eps * nd->eps);
                break;
            }
            case PPO_NODE_WEIGHT: stack[sp++] = T(weight)(k) <= nd->k; break;
            case PPO_NODE_AND: sp -= 1; stack[sp - 1] = stack[sp - 1] && stack[sp]; break;
            case PPO_NODE_OR: sp -= 1; stack[sp - 1] = stack[sp - 1] || stack[sp]; break;
            default: return 1;
        }
    }
    return stack[0];
}

static inline uint32_t T(bucket_of)(const ppo_sum* s, const T(key) * k, int bits) {
    uint32_t b = 0;
    for (int r = 0; r < bits; ++r) {
        uint64_t f = 0;
        for (int w = 0; w < W; ++w) f ^= (k->x[w] & s->rows_x[r][w]) ^ (k->z[w] & s->rows_z[r][w]);
        b |= (uint32_t)(__builtin_popcountll(f) & 1) << r;
    }
    return b;
}

static void T(bucket_reserve)(T(bucket) * b, uint32_t cap) {
    if (b->cap >= cap) return;
    uint32_t nc = b->cap ? b->cap : 16;
    while (nc < cap) nc = nc + nc / 2 + 16;
    b->k = (T(key) *)realloc(b->k, (size_t)nc * sizeof(T(key)));
    b->c = (REAL*)realloc(b->c, (size_t)nc * sizeof(REAL));
    b->cap = nc;
}

static void T(bucket_support)(T(bucket) * b) {
    for (int w = 0; w < W; ++w) b->sx[w] = b->sz[w] = 0;
    for (uint32_t i = 0; i < b->len; ++i)
        for (int w = 0; w < W; ++w) {
            b->sx[w] |= b->k[i].x[w];
            b->sz[w] |= b->k[i].z[w];
        }
}

static void T(free_buckets)(ppo_sum* s) {
    T(bucket)* bk = (T(bucket)*)s->buckets;
    if (!bk) return;
    size_t nb = (size_t)1 << s->bits;
    for (size_t i = 0; i < nb; ++i) {
        free(bk[i].k);
        free(bk[i].c);
    }
    free(bk);
    s->buckets = NULL;
}

static void T(clone_buckets)(const ppo_sum* src, ppo_sum* dst) {
    size_t nb = (size_t)1 << src->bits;
    const T(bucket)* a = (const T(bucket)*)src->buckets;
    T(bucket)* bk = (T(bucket)*)calloc(nb, sizeof(T(bucket)));
    for (size_t b = 0; b < nb; ++b) {
        /* half again as much room: a layer's merge swaps the bucket's storage into a worker's scratch, which then
         * holds the next pair's output without growing (a realloc per pair serializes the workers on the allocator) */
        T(bucket_reserve)(&bk[b], a[b].len + a[b].len / 2 + 16);
        memcpy(bk[b].k, a[b].k, (size_t)a[b].len * sizeof(T(key)));
        memcpy(bk[b].c, a[b].c, (size_t)a[b].len * sizeof(REAL));
        bk[b].len = a[b].len;
        for (int w = 0; w < W; ++w) {
            bk[b].sx[w] = a[b].sx[w];
            bk[b].sz[w] = a[b].sz[w];
        }
    }
    dst->buckets = bk;
}

/* ---- construction ---------------------------------------------------------------------- */

typedef struct {
    T(key) k;
    REAL c;
    size_t seq;
} T(entry);

static int T(entry_cmp)(const void* a, const void* b) {
    const T(entry)* p = (const T(entry)*)a;
    const T(entry)* q = (const T(entry)*)b;
    int r = T(key_cmp)(&p->k, &q->k);
    if (r) return r;
    return p->seq < q->seq ? -1 : (p->seq > q->seq ? 1 : 0);
}

/* Distribute key-sorted terms into buckets (PauliSum::from_key_sorted, storage.rs:311-336). */
static void T(from_sorted)(ppo_sum* s, size_t n, const T(key) * keys, const REAL* coeffs) {
    size_t nb = (size_t)1 << s->bits;
    T(bucket)* bk = (T(bucket)*)calloc(nb, sizeof(T(bucket)));
    uint32_t* dest = (uint32_t*)malloc((n ? n : 1) * sizeof(uint32_t));
    uint32_t* counts = (uint32_t*)calloc(nb, sizeof(uint32_t));
    for (size_t i = 0; i < n; ++i) {
        dest[i] = T(bucket_of)(s, &keys[i], s->bits);
        counts[dest[i]]++;
    }
    /* half again as much room as the rows, as in clone_buckets */
    for (size_t b = 0; b < nb; ++b) T(bucket_reserve)(&bk[b], counts[b] + counts[b] / 2 + 16);
    for (size_t i = 0; i < n; ++i) {
        T(bucket)* b = &bk[dest[i]];
        b->k[b->len] = keys[i];
        b->c[b->len] = coeffs[i];
        b->len++;
    }
    for (size_t b = 0; b < nb; ++b) T(bucket_support)(&bk[b]);
    s->buckets = bk;
    s->len = n;
    free(dest);
    free(counts);
}

static void T(build)(ppo_sum* s, size_t n, const uint64_t* x, const uint64_t* z, const REAL* c) {
    T(entry)* e = (T(entry)*)malloc((n ? n : 1) * sizeof(T(entry)));
    for (size_t i = 0; i < n; ++i) {
        for (int w = 0; w < W; ++w) {
            e[i].k.x[w] = x[i * W + w];
            e[i].k.z[w] = z[i * W + w];
        }
        e[i].c = c[i];
        e[i].seq = i;
    }
    qsort(e, n, sizeof(T(entry)), T(entry_cmp));
    T(key)* keys = (T(key)*)malloc((n ? n : 1) * sizeof(T(key)));
    REAL* cs = (REAL*)malloc((n ? n : 1) * sizeof(REAL));
    size_t m = 0;
    for (size_t i = 0; i < n;) {
        size_t j = i;
        REAL acc = e[i].c;
        while (j + 1 < n && T(key_cmp)(&e[j + 1].k, &e[i].k) == 0) {
            ++j;
            acc += e[j].c;
        }
        if (acc != (REAL)0) {
            keys[m] = e[i].k;
            cs[m] = acc;
            ++m;
        }
        i = j + 1;
    }
    s->bits = ppo_desired_bits(m, 1024, 128);
    if (s->bits > s->max_bits) s->bits = s->max_bits;
    T(from_sorted)(s, m, keys, cs);
    free(e);
    free(keys);
    free(cs);
}

/* ---- rebucketing (storage.rs:181-203, :456-487) ------------------------------------------ */

static void T(refine)(ppo_sum* s) {
    int new_bit = s->bits;
    size_t nb = (size_t)1 << s->bits;
    T(bucket)* old = (T(bucket)*)s->buckets;
    T(bucket)* bk = (T(bucket)*)calloc(2 * nb, sizeof(T(bucket)));
    for (size_t b = 0; b < nb; ++b) {
        T(bucket)* src = &old[b];
        T(bucket)* lo = &bk[b];
        T(bucket)* hi = &bk[b + nb];
        uint32_t nhi = 0;
        for (uint32_t i = 0; i < src->len; ++i) {
            uint64_t f = 0;
            for (int w = 0; w < W; ++w)
                f ^= (src->k[i].x[w] & s->rows_x[new_bit][w]) ^ (src->k[i].z[w] & s->rows_z[new_bit][w]);
            nhi += (uint32_t)(__builtin_popcountll(f) & 1);
        }
        T(bucket_reserve)(hi, nhi + nhi / 2 + 16);
        /* the lower half stays in the source buffer, compacted in order */
        *lo = *src;
        uint32_t keep = 0;
        for (uint32_t i = 0; i < src->len; ++i) {
            uint64_t f = 0;
            for (int w = 0; w < W; ++w)
                f ^= (src->k[i].x[w] & s->rows_x[new_bit][w]) ^ (src->k[i].z[w] & s->rows_z[new_bit][w]);
            if (__builtin_popcountll(f) & 1) {
                hi->k[hi->len] = src->k[i];
                hi->c[hi->len] = src->c[i];
                hi->len++;
            } else {
                lo->k[keep] = src->k[i];
                lo->c[keep] = src->c[i];
                ++keep;
            }
        }
        lo->len = keep;
        T(bucket_support)(lo);
        T(bucket_support)(hi);
    }
    free(old);
    s->buckets = bk;
    s->bits += 1;
}

static void T(coarsen)(ppo_sum* s) {
    size_t nb = (size_t)1 << s->bits;
    size_t half = nb / 2;
    T(bucket)* old = (T(bucket)*)s->buckets;
    T(bucket)* bk = (T(bucket)*)calloc(half, sizeof(T(bucket)));
    for (size_t b = 0; b < half; ++b) {
        T(bucket)* a = &old[b];
        T(bucket)* c = &old[b + half];
        T(bucket)* o = &bk[b];
        T(bucket_reserve)(o, a->len + c->len);
        uint32_t i = 0, j = 0;
        while (i < a->len && j < c->len) {
            if (T(key_cmp)(&a->k[i], &c->k[j]) < 0) {
                o->k[o->len] = a->k[i];
                o->c[o->len++] = a->c[i++];
            } else {
                o->k[o->len] = c->k[j];
                o->c[o->len++] = c->c[j++];
            }
        }
        while (i < a->len) {
            o->k[o->len] = a->k[i];
            o->c[o->len++] = a->c[i++];
        }
        while (j < c->len) {
            o->k[o->len] = c->k[j];
            o->c[o->len++] = c->c[j++];
        }
        for (int w = 0; w < W; ++w) {
            o->sx[w] = a->sx[w] | c->sx[w];
            o->sz[w] = a->sz[w] | c->sz[w];
        }
        free(a->k);
        free(a->c);
        free(c->k);
        free(c->c);
    }
    free(old);
    s->buckets = bk;
    s->bits -= 1;
}

static void T(rebucket_to)(ppo_sum* s, int want) {
    while (s->bits < want) T(refine)(s);
    while (s->bits > want) T(coarsen)(s);
}

/* ---- the rotation layer (engine/bucketed/rotation_layer.rs:159-229, :386-463; merge.rs) ---- */

typedef struct {
    uint32_t* hits[2];
    uint8_t* dirty[2];
    uint32_t hits_cap[2];
    T(key) * pk[2]; /* partners bound for member j */
    REAL* pc[2];
    uint32_t np[2];
    uint32_t p_cap[2];
    T(key) * tk; /* merge scratch */
    REAL* tc;
    uint32_t t_cap;
    T(bucket) out; /* output buffer swapped with the bucket's */
} T(scratch);

static void* T(grow)(void* p, uint32_t* cap, uint32_t need, size_t elem) {
    if (*cap >= need) return p;
    uint32_t nc = *cap ? *cap : 64;
    while (nc < need) nc *= 2;
    *cap = nc;
    return realloc(p, (size_t)nc * elem);
}

static void T(scratch_free)(T(scratch) * sc) {
    for (int j = 0; j < 2; ++j) {
        free(sc->hits[j]);
        free(sc->dirty[j]);
        free(sc->pk[j]);
        free(sc->pc[j]);
    }
    free(sc->tk);
    free(sc->tc);
    free(sc->out.k);
    free(sc->out.c);
}

typedef struct {
    const uint64_t* gx;
    const uint64_t* gz;
    REAL a0;      /* cos, rounded to REAL */
    REAL a1;      /* sin, rounded to REAL; sign applied per term */
    int clifford; /* the identity residue is dropped */
    /* g's set bits in lexicographic significance order: (half, word, bit) */
    int m;
    int half[8 * W];
    int word[8 * W];
    int bit[8 * W];
} T(rot);

/* A partner's position among partners of other classes: class = k's bits at g's set bits.
 * Within one class k -> k ^ g preserves the key order (two keys of a class first differ
 * outside g), so the sorted partner run is a merge of the class runs. */
static inline unsigned T(class_of)(const T(rot) * r, const T(key) * k) {
    unsigned cl = 0;
    for (int i = 0; i < r->m; ++i) {
        uint64_t word = r->half[i] ? k->z[r->word[i]] : k->x[r->word[i]];
        cl |= (unsigned)((word >> r->bit[i]) & 1) << i;
    }
    return cl;
}

static int T(pair_cmp)(const void* a, const void* b) {
    return T(key_cmp)((const T(key)*)a, (const T(key)*)b);
}

/* Sort the partner run bound for member j; source order is key order. */
static void T(order_partners)(const T(rot) * r, T(scratch) * sc, int j) {
    uint32_t n = sc->np[j];
    if (n < 2) return;
    sc->tk = (T(key)*)T(grow)(sc->tk, &sc->t_cap, n, sizeof(T(key)));
    sc->tc = (REAL*)realloc(sc->tc, (size_t)sc->t_cap * sizeof(REAL));
    if (r->m > 4) {
        /* A generator wider than two qubits: a comparison sort of (key, coeff) pairs. */
        typedef struct {
            T(key) k;
            REAL c;
        } kc;
        kc* tmp = (kc*)malloc((size_t)n * sizeof(kc));
        for (uint32_t i = 0; i < n; ++i) {
            tmp[i].k = sc->pk[j][i];
            tmp[i].c = sc->pc[j][i];
        }
        qsort(tmp, n, sizeof(kc), T(pair_cmp));
        for (uint32_t i = 0; i < n; ++i) {
            sc->pk[j][i] = tmp[i].k;
            sc->pc[j][i] = tmp[i].c;
        }
        free(tmp);
        return;
    }
    unsigned ncls = 1u << r->m;
    uint32_t cnt[17] = {0};
    /* class of a partner is read off the partner key itself: the class of k ^ g is the
       class of k with every bit flipped, a bijection, so grouping either way is the same. */
    for (uint32_t i = 0; i < n; ++i) cnt[T(class_of)(r, &sc->pk[j][i]) + 1]++;
    for (unsigned c = 1; c <= ncls; ++c) cnt[c] += cnt[c - 1];
    uint32_t pos[16];
    for (unsigned c = 0; c < ncls; ++c) pos[c] = cnt[c];
    for (uint32_t i = 0; i < n; ++i) {
        unsigned c = T(class_of)(r, &sc->pk[j][i]);
        sc->tk[pos[c]] = sc->pk[j][i];
        sc->tc[pos[c]] = sc->pc[j][i];
        pos[c]++;
    }
    /* bottom-up merges of the class runs, ping-ponging between tk/tc and pk/pc */
    uint32_t bounds[17];
    unsigned nruns = ncls;
    for (unsigned c = 0; c <= ncls; ++c) bounds[c] = cnt[c];
    T(key)* srck = sc->tk;
    REAL* srcc = sc->tc;
    T(key)* dstk = sc->pk[j];
    REAL* dstc = sc->pc[j];
    while (nruns > 1) {
        unsigned out_runs = 0;
        uint32_t nb[17];
        nb[0] = 0;
        for (unsigned q = 0; q < nruns; q += 2) {
            uint32_t a = bounds[q], ae = bounds[q + 1];
            uint32_t b = ae, be = (q + 1 < nruns) ? bounds[q + 2] : ae;
            uint32_t o = a;
            while (a < ae && b < be) {
                if (T(key_cmp)(&srck[a], &srck[b]) < 0) {
                    dstk[o] = srck[a];
                    dstc[o++] = srcc[a++];
                } else {
                    dstk[o] = srck[b];
                    dstc[o++] = srcc[b++];
                }
            }
            while (a < ae) {
                dstk[o] = srck[a];
                dstc[o++] = srcc[a++];
            }
            while (b < be) {
                dstk[o] = srck[b];
                dstc[o++] = srcc[b++];
            }
            nb[++out_runs] = o;
        }
        for (unsigned q = 0; q <= out_runs; ++q) bounds[q] = nb[q];
        nruns = out_runs;
        T(key)* tk = srck;
        REAL* tcc = srcc;
        srck = dstk;
        srcc = dstc;
        dstk = tk;
        dstc = tcc;
    }
    if (srck != sc->pk[j]) {
        memcpy(sc->pk[j], srck, (size_t)n * sizeof(T(key)));
        memcpy(sc->pc[j], srcc, (size_t)n * sizeof(REAL));
    }
}

/* Scan one member (rotation_layer.rs:159-229): list the anticommuting rows, append each one's
 * partner (k ^ g, c * a1) to the run bound for member `dst`, then scale the row by a0, or zero
 * it on the Clifford arm. */
static void T(scan)(const T(rot) * r, T(bucket) * b, T(scratch) * sc, int j, int dst) {
    sc->hits[j] = (uint32_t*)T(grow)(sc->hits[j], &sc->hits_cap[j], b->len + 1, sizeof(uint32_t));
    sc->dirty[j] = (uint8_t*)realloc(sc->dirty[j], (size_t)sc->hits_cap[j]);
    uint32_t nh = 0;
    for (uint32_t i = 0; i < b->len; ++i) {
        int a = T(anticommutes)(&b->k[i], r->gx, r->gz);
        sc->dirty[j][i] = (uint8_t)a;
        sc->hits[j][nh] = i;
        nh += (uint32_t)a;
    }
    uint32_t need = sc->np[dst] + nh;
    uint32_t cap = sc->p_cap[dst];
    sc->pk[dst] = (T(key)*)T(grow)(sc->pk[dst], &cap, need, sizeof(T(key)));
    sc->pc[dst] = (REAL*)realloc(sc->pc[dst], (size_t)cap * sizeof(REAL));
    sc->p_cap[dst] = cap;
    for (uint32_t h = 0; h < nh; ++h) {
        uint32_t i = sc->hits[j][h];
        T(key) q = b->k[i];
        for (int w = 0; w < W; ++w) {
            q.x[w] ^= r->gx[w];
            q.z[w] ^= r->gz[w];
        }
        /* i * sin * k * g: the phase i^(1 + delta) is real, so a sign */
        unsigned total = (1u + T(product_phase)(&b->k[i], r->gx, r->gz)) & 3u;
        REAL a1 = (total & 2u) ? -r->a1 : r->a1;
        sc->pk[dst][sc->np[dst]] = q;
        sc->pc[dst][sc->np[dst]] = b->c[i] * a1;
        sc->np[dst]++;
    }
    for (uint32_t h = 0; h < nh; ++h) {
        uint32_t i = sc->hits[j][h];
        b->c[i] = r->clifford ? (REAL)0 : b->c[i] * r->a0;
    }
}

/* Merge the sorted partners into the bucket (merge.rs:53-91 and :209-286): a partner landing on a
 * row adds to it, every changed output is kept iff nonzero and keep_term passes; when `known`,
 * rows that did not change already passed keep_term in an earlier layer of this call. */
static void T(merge)(const ppo_policy* pol, T(bucket) * b, T(scratch) * sc, int j, int has_scan, int known) {
    uint32_t np = sc->np[j];
    uint32_t need = b->len + np;
    T(bucket)* o = &sc->out;
    o->len = 0;
    T(bucket_reserve)(o, need ? need : 1);
    for (int w = 0; w < W; ++w) o->sx[w] = o->sz[w] = 0;
    const T(key)* pk = sc->pk[j];
    const REAL* pc = sc->pc[j];
    uint32_t i = 0, q = 0;
#define EMIT(KEY, VAL)                                        \
    do {                                                      \
        o->k[o->len] = (KEY);                                 \
        o->c[o->len] = (VAL);                                 \
        for (int w_ = 0; w_ < W; ++w_) {                      \
            o->sx[w_] |= o->k[o->len].x[w_];                  \
            o->sz[w_] |= o->k[o->len].z[w_];                  \
        }                                                     \
        o->len++;                                             \
    } while (0)
    while (i < b->len || q < np) {
        int ord;
        if (i == b->len) ord = 1;
        else if (q == np) ord = -1;
        else ord = T(key_cmp)(&b->k[i], &pk[q]);
        if (ord < 0) {
            REAL v = b->c[i];
            int changed = !known || (has_scan && sc->dirty[j][i]);
            if (!changed || (v != (REAL)0 && T(keep)(pol, &b->k[i], v))) EMIT(b->k[i], v);
            ++i;
        } else if (ord > 0) {
            REAL v = pc[q];
            if (v != (REAL)0 && T(keep)(pol, &pk[q], v)) EMIT(pk[q], v);
            ++q;
        } else {
            REAL v = b->c[i] + pc[q];
            if (v != (REAL)0 && T(keep)(pol, &b->k[i], v)) EMIT(b->k[i], v);
            ++i;
            ++q;
        }
    }
#undef EMIT
    /* swap buffers: the bucket takes the merged rows, the scratch keeps the old storage */
    T(bucket) tmp = *b;
    *b = *o;
    *o = tmp;
}

static int T(may_touch)(const T(rot) * r, const T(bucket) * b) {
    uint64_t f = 0;
    for (int w = 0; w < W; ++w) f |= (b->sx[w] & r->gz[w]) | (b->sz[w] & r->gx[w]);
    return f != 0;
}

static void T(rot_coset)(const T(rot) * r, const ppo_policy* pol, T(bucket) * bk, uint32_t b0, uint32_t b1,
                         int size, int known, T(scratch) * sc) {
    uint32_t mem[2] = {b0, b1};
    int scanned[2] = {0, 0};
    int changed[2] = {0, 0};
    sc->np[0] = sc->np[1] = 0;
    for (int j = 0; j < size; ++j) {
        T(bucket)* b = &bk[mem[j]];
        if (known && !T(may_touch)(r, b)) continue;
        int dst = size == 2 ? 1 - j : 0;
        uint32_t before = sc->np[dst];
        T(scan)(r, b, sc, j, dst);
        scanned[j] = 1;
        changed[j] = sc->np[dst] != before;
    }
    for (int j = 0; j < size; ++j) {
        T(bucket)* b = &bk[mem[j]];
        if (known && !changed[j] && sc->np[j] == 0) continue;
        T(order_partners)(r, sc, j);
        T(merge)(pol, b, sc, j, scanned[j], known);
    }
}

static void T(rotation_layer)(ppo_sum* s, const uint64_t* gx, const uint64_t* gz, double cosv, double sinv,
                              int clifford, const ppo_policy* pol, int threads) {
    T(rot) r;
    r.gx = gx;
    r.gz = gz;
    r.a0 = (REAL)cosv;
    r.a1 = (REAL)sinv;
    r.clifford = clifford;
    r.m = 0;
    for (int half = 0; half < 2; ++half)
        for (int w = 0; w < W; ++w) {
            uint64_t live = half ? gz[w] : gx[w];
            while (live) {
                int bit = 63 - __builtin_clzll(live);
                if (r.m < 8 * W) {
                    r.half[r.m] = half;
                    r.word[r.m] = w;
                    r.bit[r.m] = bit;
                }
                r.m++;
                live &= ~(1ull << bit);
            }
        }
    T(key) g;
    for (int w = 0; w < W; ++w) {
        g.x[w] = gx[w];
        g.z[w] = gz[w];
    }
    uint32_t delta = T(bucket_of)(s, &g, s->bits);
    size_t nb = (size_t)1 << s->bits;
    size_t ncos = delta ? nb / 2 : nb;
    int hb = delta ? 31 - __builtin_clz(delta) : 0;
    T(bucket)* bk = (T(bucket)*)s->buckets;
    int known = s->known;
#pragma omp parallel num_threads(threads > 0 ? threads : omp_get_max_threads())
    {
        T(scratch) sc;
        memset(&sc, 0, sizeof(sc));
#pragma omp for schedule(dynamic, 8)
        for (size_t p = 0; p < ncos; ++p) {
            if (delta) {
                uint32_t lo = (uint32_t)p & ((1u << hb) - 1u);
                uint32_t b0 = (((uint32_t)p >> hb) << (hb + 1)) | lo;
                T(rot_coset)(&r, pol, bk, b0, b0 ^ delta, 2, known, &sc);
            } else {
                T(rot_coset)(&r, pol, bk, (uint32_t)p, (uint32_t)p, 1, known, &sc);
            }
        }
        T(scratch_free)(&sc);
    }
    size_t len = 0;
    for (size_t b = 0; b < nb; ++b) len += bk[b].len;
    s->len = len;
    s->known = 1;
}

/* Key-preserving table (a rotation whose sine snapped to zero): rescale in place with the zero
 * drop and keep_term (engine/bucketed.rs:646-662). */
static void T(rescale_layer)(ppo_sum* s, const uint64_t* gx, const uint64_t* gz, double a0_anti,
                             const ppo_policy* pol) {
    size_t nb = (size_t)1 << s->bits;
    T(bucket)* bk = (T(bucket)*)s->buckets;
    REAL a0 = (REAL)a0_anti;
    size_t len = 0;
    for (size_t b = 0; b < nb; ++b) {
        T(bucket)* bb = &bk[b];
        uint32_t w = 0;
        for (uint32_t i = 0; i < bb->len; ++i) {
            REAL v = T(anticommutes)(&bb->k[i], gx, gz) ? bb->c[i] * a0 : bb->c[i] * (REAL)1;
            if (v != (REAL)0 && T(keep)(pol, &bb->k[i], v)) {
                bb->k[w] = bb->k[i];
                bb->c[w] = v;
                ++w;
            }
        }
        bb->len = w;
        len += w;
    }
    s->len = len;
    s->known = 0;
}

/* ---- TopN (truncation/builtin.rs:73-129, :152-239) ---------------------------------------- */

static void T(topn)(ppo_sum* s, uint64_t n) {
    size_t total = s->len;
    if (total <= n) return;
    size_t nb = (size_t)1 << s->bits;
    T(bucket)* bk = (T(bucket)*)s->buckets;
    if (n == 0) {
        for (size_t b = 0; b < nb; ++b) bk[b].len = 0;
        s->len = 0;
        s->known = 0;
        return;
    }
    double* m = (double*)malloc(total * sizeof(double));
    size_t t = 0;
    for (size_t b = 0; b < nb; ++b)
        for (uint32_t i = 0; i < bk[b].len; ++i) {
            double v = (double)bk[b].c[i];
            m[t++] = v * v;
        }
    double* sorted = (double*)malloc(total * sizeof(double));
    memcpy(sorted, m, total * sizeof(double));
    qsort(sorted, total, sizeof(double), ppo_cmp_magnitude_desc);
    double boundary = sorted[n - 1];
    free(sorted);
    int nan_boundary = boundary != boundary;
    size_t above = 0;
    if (!nan_boundary)
        for (size_t i = 0; i < total; ++i) above += m[i] > boundary;
    size_t quota = n - above;
    t = 0;
    size_t len = 0;
    for (size_t b = 0; b < nb; ++b) {
        T(bucket)* bb = &bk[b];
        uint32_t w = 0;
        for (uint32_t i = 0; i < bb->len; ++i) {
            double v = m[t++];
            int side; /* 1 above, 0 at, -1 below */
            if (v != v) side = -1;
            else if (nan_boundary) side = 1;
            else side = v > boundary ? 1 : (v == boundary ? 0 : -1);
            int at = (side == 0) && quota > 0;
            quota -= (size_t)at;
            if (side == 1 || at) {
                bb->k[w] = bb->k[i];
                bb->c[w] = bb->c[i];
                ++w;
            }
        }
        bb->len = w;
        len += w;
    }
    free(m);
    s->len = len;
    s->known = 0;
}

/* ---- ApproxTopN (truncation/builtin.rs:241-468, approx_top_n_solo) -------------------------- */

/* The octave of |c|^2: its exponent bits, norm_sqr().to_bits() >> 52 (builtin.rs:276-279).  A NaN of either
 * sign lands in octave 2047, as a positive NaN does in the reference (a negative one would index past its
 * 2048 bins). */
static unsigned T(octave)(double m) {
    uint64_t u;
    memcpy(&u, &m, 8);
    return (unsigned)((u >> 52) & 0x7FFu);
}

static void T(approx_topn)(ppo_sum* s, uint64_t n) {
    size_t total = s->len;
    if (total <= n) return; /* EdgeDecision::KeepAll: nothing dropped, NaN included */
    size_t nb = (size_t)1 << s->bits;
    T(bucket)* bk = (T(bucket)*)s->buckets;
    if (n == 0) { /* EdgeDecision::Clear */
        for (size_t b = 0; b < nb; ++b) bk[b].len = 0;
        s->len = 0;
        s->known = 0;
        return;
    }
    /* octave_histogram, then octave_edge: walk from the top octave down until one does not fit */
    size_t hist[2048];
    memset(hist, 0, sizeof(hist));
    for (size_t b = 0; b < nb; ++b)
        for (uint32_t i = 0; i < bk[b].len; ++i) {
            double v = (double)bk[b].c[i];
            hist[T(octave)(v * v)]++;
        }
    size_t kept_above = 0;
    int bin = -1;
    for (int o = 2047; o >= 0; --o) {
        if (kept_above + hist[o] > n) {
            bin = o;
            break;
        }
        kept_above += hist[o];
    }
    if (bin < 0) return; /* unreachable: more than n terms overflow n at some octave */
    size_t quota = n - kept_above;
    /* retain_with_boundary_quota (builtin.rs:208-239): above the boundary octave kept, the first `quota` at it
     * in bucket order kept, NaN below every octave */
    size_t len = 0;
    for (size_t b = 0; b < nb; ++b) {
        T(bucket)* bb = &bk[b];
        uint32_t w = 0;
        for (uint32_t i = 0; i < bb->len; ++i) {
            double v = (double)bb->c[i];
            double m = v * v;
            int side;
            if (m != m) side = -1;
            else {
                unsigned o = T(octave)(m);
                side = o > (unsigned)bin ? 1 : (o == (unsigned)bin ? 0 : -1);
            }
            int at = (side == 0) && quota > 0;
            quota -= (size_t)at;
            if (side == 1 || at) {
                bb->k[w] = bb->k[i];
                bb->c[w] = bb->c[i];
                ++w;
            }
        }
        bb->len = w;
        len += w;
    }
    s->len = len;
    s->known = 0;
}

/* ---- read-out and export ------------------------------------------------------------------- */

static double T(expectation)(const ppo_sum* s, const uint64_t* ax, const uint64_t* az, const uint64_t* neg) {
    size_t nb = (size_t)1 << s->bits;
    const T(bucket)* bk = (const T(bucket)*)s->buckets;
    double total = 0.0;
    for (size_t b = 0; b < nb; ++b) {
        REAL partial = (REAL)0;
        for (uint32_t i = 0; i < bk[b].len; ++i) {
            const T(key)* k = &bk[b].k[i];
            uint64_t mismatch = 0;
            unsigned sign = 0;
            for (int w = 0; w < W; ++w) {
                uint64_t sup = k->x[w] | k->z[w];
                mismatch |= (k->x[w] ^ (sup & ax[w])) | (k->z[w] ^ (sup & az[w]));
                sign += (unsigned)__builtin_popcountll(sup & neg[w]);
            }
            if (mismatch) continue;
            partial += (sign & 1u) ? -bk[b].c[i] : bk[b].c[i];
        }
        total += (double)partial;
    }
    return total;
}

static void T(export)(const ppo_sum* s, uint64_t* x, uint64_t* z, REAL* c, uint32_t* bucket) {
    size_t nb = (size_t)1 << s->bits;
    const T(bucket)* bk = (const T(bucket)*)s->buckets;
    size_t t = 0;
    for (size_t b = 0; b < nb; ++b)
        for (uint32_t i = 0; i < bk[b].len; ++i, ++t) {
            for (int w = 0; w < W; ++w) {
                x[t * W + w] = bk[b].k[i].x[w];
                z[t * W + w] = bk[b].k[i].z[w];
            }
            c[t] = bk[b].c[i];
            if (bucket) bucket[t] = (uint32_t)b;
        }
}

/* ---- the naive differential oracle (test_support/oracle.rs:21-54) --------------------------- */

static size_t T(naive)(size_t n, const uint64_t* x, const uint64_t* z, const REAL* c, const uint64_t* gx,
                       const uint64_t* gz, double sinv, double cosv, const ppo_policy* pol, uint64_t* ox,
                       uint64_t* oz, REAL* oc) {
    T(entry)* e = (T(entry)*)malloc((2 * n + 1) * sizeof(T(entry)));
    size_t m = 0;
    for (size_t i = 0; i < n; ++i) {
        T(key) k;
        for (int w = 0; w < W; ++w) {
            k.x[w] = x[i * W + w];
            k.z[w] = z[i * W + w];
        }
        if (!T(anticommutes)(&k, gx, gz)) {
            e[m].k = k;
            e[m].c = c[i];
            e[m].seq = m;
            ++m;
            continue;
        }
        /* cos * Q + i sin * Q P (channel/rotation.rs:43-66) */
        e[m].k = k;
        e[m].c = c[i] * (REAL)cosv;
        e[m].seq = m;
        ++m;
        unsigned total = (1u + T(product_phase)(&k, gx, gz)) & 3u;
        REAL signed_c = (total & 2u) ? -c[i] : c[i];
        for (int w = 0; w < W; ++w) {
            k.x[w] ^= gx[w];
            k.z[w] ^= gz[w];
        }
        e[m].k = k;
        e[m].c = signed_c * (REAL)sinv;
        e[m].seq = m;
        ++m;
    }
    qsort(e, m, sizeof(T(entry)), T(entry_cmp));
    size_t out = 0;
    for (size_t i = 0; i < m;) {
        size_t j = i;
        REAL acc = (REAL)0 + e[i].c;
        while (j + 1 < m && T(key_cmp)(&e[j + 1].k, &e[i].k) == 0) {
            ++j;
            acc += e[j].c;
        }
        if (acc != (REAL)0 && T(keep)(pol, &e[i].k, acc)) {
            for (int w = 0; w < W; ++w) {
                ox[out * W + w] = e[i].k.x[w];
                oz[out * W + w] = e[i].k.z[w];
            }
            oc[out] = acc;
            ++out;
        }
        i = j + 1;
    }
    free(e);
    return out;
}

#undef T
#undef CAT
#undef CAT2
