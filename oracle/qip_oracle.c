/*
 * qip_oracle.c -- CPU restatement of the reference's native state-vector kernels.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker that the HIP engine is compared
 * against, and (built a second time with the reference's compiler flags) the CPU baseline
 * timed beside it. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it. It is never linked into, or called by, the product library.
 *
 * Each function restates one reference function loop for loop and cites it. Bit order is the
 * reference's: qubit q is state-index bit (n-1-q) (qip/ext/kronprod.pyx:168,187-189).
 * Index arithmetic is widened from C int to int64_t so the same code runs up to n = 34; for
 * n <= 30 (the reference's limit) the results are the same.
 *
 * Where the reference relies on undefined behaviour the restatement takes the benign reading
 * and says so at the function:
 *   - prob_magnitude's accumulator is uninitialised (kronprod.pyx:322): here it starts at 0;
 *   - func_apply's rest-register loop runs j < n_reg2 over other_indices (func_apply.pyx:109-111):
 *     `faithful=1` keeps that bound (so rest bits beyond n_reg2 are dropped, as in the reference,
 *     and out-of-range reads contribute nothing), `faithful=0` is the intended j < n_other;
 *   - measure_top_probabilities writes one slot past its heap (kronprod.pyx:298-311): the
 *     faithful variant gives that slot real storage initialised to 0.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

typedef double complex cplx;

/* kronprod.pyx:18-19: MatrixType enum values, also the `_kron_struct` ints of operators.py */
enum { OR_OTHER = 0, OR_NUMPY = 1, OR_CMAT = 2, OR_SWAP = 3 };

/* kronprod.pyx:21-25 MatStruct. `child` replaces the raw pointer to a nested struct: it is the
 * index of the nested matrix in the same array (top-level matrices first, nested ones after,
 * exactly as cdot_loop lays out `cmatrices`, kronprod.pyx:61,118-125). */
typedef struct {
    int32_t type;
    int32_t param;
    int32_t child;
    int32_t pad_;
    const cplx *data;
} or_mat;

/* util.pyx:25-26 */
static inline int64_t set_bit(int64_t num, int bit, int64_t value) {
    return num ^ ((-(int64_t)(value != 0) ^ num) & ((int64_t)1 << bit));
}

/* util.pyx:29-30 */
static inline int64_t get_bit(int64_t num, int bit) { return (num >> bit) & 1; }

/* util.pyx:2-23: interleave v_indices (on the set bits of mask) with v_non_indices (on the
 * clear bits), both MSB-first. */
int64_t or_entwine_bit(int n_indices, int n_non_indices, int64_t v_indices, int64_t v_non_indices,
                       int64_t mask) {
    int n = n_indices + n_non_indices;
    int64_t indices_mask = (int64_t)1 << (n_indices - 1 > 0 ? n_indices - 1 : 0);
    int64_t non_indices_mask = (int64_t)1 << (n_non_indices - 1 > 0 ? n_non_indices - 1 : 0);
    int64_t bit_select = (int64_t)1 << (n - 1);
    int64_t m = 0;
    for (int j = 0; j < n; ++j) {
        m <<= 1;
        if ((mask & bit_select) >> ((n - 1) - j)) {
            n_indices -= 1;
            m += (v_indices & indices_mask) >> n_indices;
            indices_mask >>= 1;
        } else {
            n_non_indices -= 1;
            m += (v_non_indices & non_indices_mask) >> n_non_indices;
            non_indices_mask >>= 1;
        }
        bit_select >>= 1;
    }
    return m;
}

/* kronprod.pyx:205-235 calc_mat_entry */
static cplx calc_mat_entry(const or_mat *mats, int32_t which, int64_t mati, int64_t matj) {
    const or_mat *ms = &mats[which];
    if (ms->type == OR_CMAT) {
        int64_t half = (int64_t)ms->param >> 1;
        if (mati < half && matj < half) return mati == matj ? 1.0 : 0.0;
        if (mati >= half && matj >= half) return calc_mat_entry(mats, ms->child, mati - half, matj - half);
        return 0.0;
    } else if (ms->type == OR_SWAP) {
        int64_t lowmask = ((int64_t)1 << ms->param) - 1; /* == ~((-1) << param) */
        if ((mati & lowmask) == (matj >> ms->param) && (matj & lowmask) == (mati >> ms->param)) return 1.0;
        return 0.0;
    } else if (ms->type == OR_NUMPY) {
        return ms->data[mati * ms->param + matj];
    }
    return 0.0;
}

/*
 * kronprod.pyx:41-200 cdot_loop (its hot loop is 155-197). Out-of-place:
 *   out[r] = sum over columns c that agree with r outside the union of all group indices of
 *            prod_g M_g[r|_g, c|_g] * vec[c]
 * Groups are given flattened (`group_idx`, lengths `group_len`); the reference packs them
 * into an int[:,:] (so equal lengths), which is the special case of this layout.
 * `nthreads` <= 0 leaves the OpenMP default (the reference uses prange(schedule='static')).
 */
int or_cdot(int32_t n_groups, const int32_t *group_len, const int32_t *group_idx,
            const or_mat *mats, const cplx *vec, int64_t input_len, int32_t n, cplx *out,
            int64_t output_len, int64_t input_offset, int64_t output_offset, int nthreads) {
    /* flatindices: sorted unique union of all group indices (kronprod.pyx:140-141) */
    int32_t total = 0;
    for (int g = 0; g < n_groups; ++g) total += group_len[g];
    int32_t *flat = (int32_t *)malloc(sizeof(int32_t) * (total > 0 ? total : 1));
    int32_t *gstart = (int32_t *)malloc(sizeof(int32_t) * (n_groups > 0 ? n_groups : 1));
    if (!flat || !gstart) { free(flat); free(gstart); return -1; }
    int32_t nind = 0, acc = 0;
    for (int g = 0; g < n_groups; ++g) {
        gstart[g] = acc;
        for (int j = 0; j < group_len[g]; ++j) {
            int32_t v = group_idx[acc + j];
            int seen = 0;
            for (int t = 0; t < nind; ++t) seen |= (flat[t] == v);
            if (!seen) flat[nind++] = v;
        }
        acc += group_len[g];
    }
    for (int a = 1; a < nind; ++a) { /* insertion sort */
        int32_t v = flat[a];
        int b = a - 1;
        while (b >= 0 && flat[b] > v) { flat[b + 1] = flat[b]; --b; }
        flat[b + 1] = v;
    }
    const int64_t two_nind = (int64_t)1 << nind;

#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
    for (int64_t orow = 0; orow < output_len; ++orow) {
        int64_t row = orow + output_offset;
        int64_t colbits = row;
        cplx s = 0.0;
        for (int64_t i = 0; i < two_nind; ++i) {
            for (int j = 0; j < nind; ++j)
                colbits = set_bit(colbits, (n - 1) - flat[j], get_bit(i, (nind - 1) - j));
            int64_t vecrow = colbits - input_offset;
            if (vecrow < 0 || vecrow >= input_len) continue;
            cplx p = 1.0;
            for (int g = 0; g < n_groups; ++g) {
                int len = group_len[g];
                int64_t submati = 0, submatj = 0;
                for (int mi = 0; mi < len; ++mi) {
                    int32_t matindex = group_idx[gstart[g] + mi];
                    submati = set_bit(submati, (len - 1) - mi, get_bit(row, (n - 1) - matindex));
                    submatj = set_bit(submatj, (len - 1) - mi, get_bit(colbits, (n - 1) - matindex));
                }
                p = p * calc_mat_entry(mats, g, submati, submatj);
                if (p == 0.0) break;
            }
            s = s + p * vec[vecrow];
        }
        out[orow] = s;
    }
    free(flat);
    free(gstart);
    return 0;
}

/*
 * func_apply.pyx:11-112 func_apply: |x>|q>|r> -> |x>|f(x) xor q>|r>, out-of-place, with the
 * function pre-tabulated (`table[x]` = f(x) as C int, func_apply.pyx:66-69). `other_idx` lists
 * the qubits in neither register in ascending order (func_apply.pyx:58-60).
 * faithful=1 keeps the reference's rest-register bound j < n_reg2 (func_apply.pyx:109);
 * faithful=0 uses the intended j < n_other. They agree whenever n_other <= n_reg2.
 */
int or_func_apply(int32_t n_reg1, const int32_t *reg1, int32_t n_reg2, const int32_t *reg2,
                  int32_t n_other, const int32_t *other_idx, const int32_t *table,
                  const cplx *vec, int32_t n, cplx *out, int faithful, int nthreads) {
    const int64_t N = (int64_t)1 << n;
    const int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nt)
    for (int64_t i = 0; i < N; ++i) out[i] = 0.0;

    const int64_t nx = (int64_t)1 << n_reg1;
    const int64_t nq = (int64_t)1 << n_reg2;
    const int64_t nr = (int64_t)1 << n_other;
    const int rest_bits = faithful ? (n_reg2 < n_other ? n_reg2 : n_other) : n_other;
#pragma omp parallel for num_threads(nt)
    for (int64_t x = 0; x < nx; ++x) {
        int64_t y = (int64_t)table[x];
        int64_t zero_index = 0;
        for (int j = 0; j < n_reg1; ++j)
            zero_index = set_bit(zero_index, (n - 1) - reg1[j], get_bit(x, (n_reg1 - 1) - j));
        for (int64_t q = 0; q < nq; ++q) {
            int64_t reg2_value = y ^ q;
            int64_t f_off = 0, q_off = 0;
            for (int j = 0; j < n_reg2; ++j)
                f_off = set_bit(f_off, (n - 1) - reg2[j], get_bit(reg2_value, (n_reg2 - 1) - j));
            for (int j = 0; j < n_reg2; ++j)
                q_off = set_bit(q_off, (n - 1) - reg2[j], get_bit(q, (n_reg2 - 1) - j));
            for (int64_t r = 0; r < nr; ++r) {
                int64_t r_off = 0;
                for (int j = 0; j < rest_bits; ++j)
                    r_off = set_bit(r_off, (n - 1) - other_idx[j], get_bit(r, (n_other - 1) - j));
                out[zero_index | f_off | r_off] = vec[zero_index | q_off | r_off];
            }
        }
    }
    return 0;
}

/* kronprod.pyx:320-326 prob_magnitude (accumulator initialised to 0; the reference leaves it
 * uninitialised). abs(z)**2 is restated as cabs(z)*cabs(z). */
double or_prob_magnitude(const cplx *vec, int64_t len) {
    double acc = 0.0;
    for (int64_t i = 0; i < len; ++i) {
        double a = cabs(vec[i]);
        acc += a * a;
    }
    return acc;
}

/* kronprod.pyx:240-261 measure_probabilities: qubit indices[j] lands on bit j of the output
 * index (LSB-first); zero amplitudes are skipped. Serial, as in the reference. */
void or_measure_probabilities(int32_t k, const int32_t *indices, int32_t n, const cplx *vec,
                              double *p) {
    const int64_t N = (int64_t)1 << n;
    memset(p, 0, sizeof(double) * ((size_t)1 << k));
    for (int64_t i = 0; i < N; ++i) {
        cplx v = vec[i];
        if (v == 0.0) continue;
        int64_t p_index = 0;
        for (int j = 0; j < k; ++j) p_index = set_bit(p_index, j, get_bit(i, (n - 1) - indices[j]));
        double a = cabs(v);
        p[p_index] = p[p_index] + a * a;
    }
}

static int64_t row_mask_of(int32_t k, const int32_t *indices, int32_t n) {
    int64_t row_mask = 0;
    for (int i = 0; i < k; ++i) row_mask = set_bit(row_mask, (n - 1) - indices[i], 1);
    return row_mask;
}

/* Marginal probability of measured value m (entwine order) over the whole state. */
static double marginal(int32_t k, int32_t n, int64_t row_mask, int64_t m, const cplx *vec,
                       int64_t len, int64_t input_offset) {
    double acc = 0.0;
    const int64_t nun = (int64_t)1 << (n - k);
    for (int64_t i = 0; i < nun; ++i) {
        int64_t vi = or_entwine_bit(k, n - k, m, i, row_mask);
        if (input_offset <= vi && vi < len + input_offset) {
            double a = cabs(vec[vi - input_offset]);
            acc += a * a;
        }
    }
    return acc;
}

/*
 * kronprod.pyx:266-315 measure_top_probabilities. faithful=1 restates the heap loop with the
 * slot it writes past the end (ps[top_k], kronprod.pyx:298-311) given real storage holding 0,
 * then the final stable descending sort of the first top_k entries (kronprod.pyx:313).
 * faithful=0 is the intended top-k: the top_k largest marginals, descending, ties by
 * ascending m. Returns the number of entries written (top_k clipped to 2^k).
 */
int32_t or_measure_top_probabilities(int32_t k, const int32_t *indices, int32_t n, int32_t top_k,
                                     const cplx *vec, int faithful, int64_t *ms_out,
                                     double *ps_out) {
    const int64_t nm = (int64_t)1 << k;
    if (top_k > nm) top_k = (int32_t)nm;
    const int64_t row_mask = row_mask_of(k, indices, n);
    const int64_t len = (int64_t)1 << n;
    int64_t *ms = (int64_t *)calloc((size_t)nm + 1, sizeof(int64_t));
    double *ps = (double *)calloc((size_t)nm + 1, sizeof(double));
    if (!ms || !ps) { free(ms); free(ps); return -1; }
    if (faithful) {
        int heap_end = 0;
        for (int64_t m = 0; m < nm; ++m) {
            double vp = marginal(k, n, row_mask, m, vec, len, 0);
            int heap_pos = heap_end;
            int heap_parent = (heap_pos - 1) >> 1;
            if (vp >= ps[heap_pos]) {
                ms[heap_pos] = m;
                ps[heap_pos] = vp;
                while (heap_pos > 0 && ps[heap_pos] > ps[heap_parent]) {
                    int64_t tm = ms[heap_pos]; ms[heap_pos] = ms[heap_parent]; ms[heap_parent] = tm;
                    double tp = ps[heap_pos]; ps[heap_pos] = ps[heap_parent]; ps[heap_parent] = tp;
                    heap_pos = heap_parent;
                    heap_parent = (heap_pos - 1) >> 1;
                }
            }
            if (heap_end < top_k) heap_end += 1;
        }
        /* stable sort, descending by probability (Python sorted(..., reverse=True) keeps the
         * original order of equal keys) */
        for (int a = 1; a < top_k; ++a) {
            int64_t vm = ms[a];
            double vp = ps[a];
            int b = a - 1;
            while (b >= 0 && ps[b] < vp) { ms[b + 1] = ms[b]; ps[b + 1] = ps[b]; --b; }
            ms[b + 1] = vm;
            ps[b + 1] = vp;
        }
        for (int a = 0; a < top_k; ++a) { ms_out[a] = ms[a]; ps_out[a] = ps[a]; }
    } else {
        for (int64_t m = 0; m < nm; ++m) { ms[m] = m; ps[m] = marginal(k, n, row_mask, m, vec, len, 0); }
        /* selection of the top_k by (-p, m) */
        for (int a = 0; a < top_k; ++a) {
            int64_t best = a;
            for (int64_t b = a + 1; b < nm; ++b)
                if (ps[b] > ps[best] || (ps[b] == ps[best] && ms[b] < ms[best])) best = b;
            int64_t tm = ms[a]; ms[a] = ms[best]; ms[best] = tm;
            double tp = ps[a]; ps[a] = ps[best]; ps[best] = tp;
            ms_out[a] = ms[a];
            ps_out[a] = ps[a];
        }
    }
    free(ms);
    free(ps);
    return top_k;
}

/*
 * kronprod.pyx:331-383 soft_measure, with the uniform draw `u` passed in (the reference draws
 * random.random() itself, kronprod.pyx:362). r = u * prob_magnitude(vec) (kronprod.pyx:364-365),
 * then the sequential walk `r -= acc; if r <= 0: break` over p_index ascending
 * (kronprod.pyx:371-381). measured < 0 means None. Returns the raw (unnormalised) acc.
 */
void or_soft_measure(int32_t k, const int32_t *indices, int32_t n, const cplx *vec, int64_t len,
                     int64_t measured, int64_t input_offset, double u, int64_t *m_out,
                     double *p_out) {
    int64_t lo, hi;
    if (measured < 0) { lo = 0; hi = (int64_t)1 << k; }
    else { lo = measured; hi = measured + 1; }
    double r = u;
    if ((int64_t)k < ((int64_t)1 << n)) r = r * or_prob_magnitude(vec, len);
    const int64_t row_mask = row_mask_of(k, indices, n);
    int64_t p_index = lo;
    double acc = 0.0;
    for (p_index = lo; p_index < hi; ++p_index) {
        acc = marginal(k, n, row_mask, p_index, vec, len, input_offset);
        r -= acc;
        if (r <= 0) break;
    }
    if (p_index == hi) p_index = hi - 1; /* Cython leaves the loop variable at its last value */
    *m_out = p_index;
    *p_out = acc;
}

/* kronprod.pyx:439-444: the collapse loop of measure (offsets 0, in/out both of length 2^n):
 * out = vec * sqrt(1/p) where the measured bits equal m, else 0. */
void or_collapse(int32_t k, const int32_t *indices, int32_t n, const cplx *vec, cplx *out,
                 int64_t m, double mult_p) {
    const int64_t row_mask = row_mask_of(k, indices, n);
    const int64_t measured_entwine = or_entwine_bit(k, n - k, m, 0, row_mask);
    const int64_t N = (int64_t)1 << n;
    for (int64_t i = 0; i < N; ++i)
        out[i] = ((i & row_mask) ^ measured_entwine) ? 0.0 : vec[i] * mult_p;
}

/* kronprod.pyx:478-489: the reduce loop of reduce_measure (offsets 0). Every output index i of
 * the 2^n arena reads the reduced amplitude of its low n-k bits, so the reduced state appears
 * tiled 2^k times (the reference does not shrink its arena, qip/backend.py:139-142). */
void or_reduce(int32_t k, const int32_t *indices, int32_t n, const cplx *vec, cplx *out, int64_t m,
               double mult_p) {
    const int64_t row_mask = row_mask_of(k, indices, n);
    const int64_t N = (int64_t)1 << n;
    for (int64_t i = 0; i < N; ++i) {
        int64_t vec_row = or_entwine_bit(k, n - k, m, i, row_mask);
        out[i] = vec[vec_row] * mult_p;
    }
}
