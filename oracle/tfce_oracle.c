/*
 * CPU ORACLE — TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this library, and only as the checker (or as the timed CPU baseline).  The
 * product path (tfce_mediation_amd) never links or calls it.
 *
 * Plain-C restatement of the TFCE enhancement in the reference's
 * tfce_mediation/lib/fast_tfce.hpp:11-95 (the operator bound by
 * tfce_mediation/tfce.pyx:24-45).  Semantics kept from the reference:
 *   - levels: T_0 = max(image); T_{i+1} = fl32(T_i - dT) while T_i >= minT
 *     (fast_tfce.hpp:39), dT = fl32(max / 100) when 0 is passed (:34-36);
 *   - a vertex is active at level T iff image[v] > T (strict, :41);
 *   - when a vertex activates it is joined with every vertex on its adjacency
 *     list that is already active (:44-66).  For a symmetric adjacency this is
 *     plain connected components of {v : image[v] > T};
 *   - per level, HH = powf(T, H) in float (:70) and every member of a component
 *     of size s receives (float)(pow((double)s, (double)E) * HH) (:77), added
 *     to enhn in float, one add per level in descending-T order (:82);
 *   - enhn is accumulated into, never cleared (the callers zero it).
 * Defined where the reference is not:
 *   - max(image) == 0 (or dT rounding to 0) makes the reference loop forever;
 *     here nothing is added and the return value is TFCE_ORACLE_DEGENERATE;
 *   - NaN elements never activate and do not take part in max();
 *   - ties are activated in ascending vertex order (std::sort leaves them
 *     unspecified; only an asymmetric adjacency can observe the difference).
 * pow_mode 0 computes HH with libm powf exactly as the reference does; glibc's
 * powf(T, 2) is one ulp away from the correctly rounded T*T for ~0.06% of
 * float inputs.  pow_mode 1 uses the correctly rounded power
 * (float)pow((double)T, (double)H) instead, which is what the GPU computes.
 *
 * The data structure is my own: a union-find with union by size and path
 * halving over a value-sorted activation order, not the reference's
 * list-splicing disjoint sets.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define TFCE_ORACLE_DEGENERATE (-2)
#define TFCE_ORACLE_NOMEM (-1)

typedef struct {
    float v;
    int32_t i;
} kv_t;

static int cmp_desc(const void *a, const void *b) {
    const kv_t *x = (const kv_t *)a, *y = (const kv_t *)b;
    int xn = isnan(x->v), yn = isnan(y->v);
    if (xn != yn) return xn - yn; /* NaN last */
    if (!xn) {
        if (x->v > y->v) return -1;
        if (x->v < y->v) return 1;
    }
    return (x->i > y->i) - (x->i < y->i);
}

static int32_t uf_find(int32_t *parent, int32_t x) {
    while (parent[x] != x) {
        parent[x] = parent[parent[x]];
        x = parent[x];
    }
    return x;
}

/* Number of levels the reference loop visits for (maxT, minT, deltaT), and the
 * thresholds themselves when out != NULL (at most cap of them). */
int64_t oracle_tfce_levels(float maxT, float minT, float deltaT, float *out, int64_t cap) {
    if (deltaT == 0.0f) deltaT = maxT / 100.0f;
    if (!(maxT >= minT)) return 0;
    if (!(deltaT > 0.0f)) return TFCE_ORACLE_DEGENERATE;
    int64_t n = 0;
    for (float T = maxT; T >= minT; T -= deltaT) {
        if (out && n < cap) out[n] = T;
        n++;
        if (n > (int64_t)1 << 40) return TFCE_ORACLE_DEGENERATE;
    }
    return n;
}

/*
 * enhn += TFCE(image).  row_ptr has V+1 entries, col the adjacency lists.
 * If labels_level >= 0 and labels != NULL, labels[v] receives the smallest
 * vertex index of v's component after level labels_level has been joined
 * (-1 for inactive vertices).  Returns the number of levels visited, or a
 * negative status.
 */
int64_t oracle_tfce(float H, float E, float minT, float deltaT, int32_t V,
                    const int64_t *row_ptr, const int32_t *col,
                    const float *image, float *enhn,
                    int64_t labels_level, int32_t *labels, int pow_mode) {
    if (V <= 0) return 0;
    kv_t *ord = (kv_t *)malloc(sizeof(kv_t) * (size_t)V);
    int32_t *parent = (int32_t *)malloc(sizeof(int32_t) * (size_t)V);
    int64_t *size = (int64_t *)malloc(sizeof(int64_t) * (size_t)V);
    float *inc = (float *)malloc(sizeof(float) * (size_t)V);
    int32_t *active_list = (int32_t *)malloc(sizeof(int32_t) * (size_t)V);
    if (!ord || !parent || !size || !inc || !active_list) {
        free(ord); free(parent); free(size); free(inc); free(active_list);
        return TFCE_ORACLE_NOMEM;
    }
    for (int32_t i = 0; i < V; ++i) {
        ord[i].v = image[i];
        ord[i].i = i;
        parent[i] = -1; /* inactive */
    }
    qsort(ord, (size_t)V, sizeof(kv_t), cmp_desc);
    float maxT = ord[0].v;
    int64_t status = 0;
    if (deltaT == 0.0f) deltaT = maxT / 100.0f;
    if (isnan(maxT) || !(maxT >= minT)) goto done;
    if (!(deltaT > 0.0f)) { status = TFCE_ORACLE_DEGENERATE; goto done; }

    int32_t j = 0, n_active = 0;
    int64_t level = 0;
    for (float T = maxT; T >= minT; T -= deltaT, ++level) {
        while (j < V && ord[j].v > T) {
            int32_t v = ord[j].i;
            parent[v] = v;
            size[v] = 1;
            active_list[n_active++] = v;
            for (int64_t e = row_ptr[v]; e < row_ptr[v + 1]; ++e) {
                int32_t a = col[e];
                if (a < 0 || a >= V || parent[a] < 0) continue;
                int32_t ra = uf_find(parent, a), rv = uf_find(parent, v);
                if (ra == rv) continue;
                if (size[ra] < size[rv]) { int32_t t = ra; ra = rv; rv = t; }
                parent[rv] = ra;
                size[ra] += size[rv];
            }
            ++j;
        }
        float HH = pow_mode ? (float)pow((double)T, (double)H) : powf(T, H);
        for (int32_t q = 0; q < n_active; ++q) {
            int32_t v = active_list[q];
            if (parent[v] == v) inc[v] = (float)(pow((double)size[v], (double)E) * (double)HH);
        }
        for (int32_t q = 0; q < n_active; ++q) {
            int32_t v = active_list[q];
            enhn[v] += inc[uf_find(parent, v)];
        }
        if (labels && level == labels_level) {
            /* canonical label = smallest member index */
            for (int32_t v = 0; v < V; ++v) labels[v] = -1;
            for (int32_t q = 0; q < n_active; ++q) {
                int32_t v = active_list[q];
                int32_t r = uf_find(parent, v);
                if (labels[r] < 0 || v < labels[r]) labels[r] = v;
            }
            for (int32_t q = 0; q < n_active; ++q) {
                int32_t v = active_list[q];
                int32_t r = uf_find(parent, v);
                if (v != r) labels[v] = labels[r];
            }
        }
        status = level + 1;
    }
done:
    free(ord); free(parent); free(size); free(inc); free(active_list);
    return status;
}
