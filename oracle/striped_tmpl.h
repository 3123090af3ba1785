/* striped_tmpl.h -- one element width of the AVX2 striped baseline
 * (included by striped.c with E_BITS 8 or 16).  TEST INFRASTRUCTURE ONLY:
 * the CPU baseline bench.py times beside the GPU, never a product path.
 *
 * Farrar's striped query profile (Farrar, Bioinformatics 23(2), 2007) as
 * the reference's dispatched AVX2 kernels use it:
 *   SW  src/sw_striped_avx2_256_{8,16}.c:140-422 (scores biased by the
 *       width's minimum so the local floor is the saturating add's own;
 *       the column holding the running maximum kept for the end row),
 *   NW/SG src/sg_striped_avx2_256_{8,16}.c:148-450 (signed scores, the
 *       boundary row and column, last-row and last-column ends, and the
 *       NEG_LIMIT / POS_LIMIT saturation window over H, E and F).
 * The lazy-F pass follows the SWPS3 revision both files use: it corrects
 * H (and the running maximum) but not E.  A row of the striped layout
 * sits in lane k, segment i of a column: row = k * seg + i.
 *
 * Returns 0 with *out filled, 1 when the width saturates (out untouched),
 * -1 on allocation failure. */

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, E_BITS)

#if E_BITS == 8
typedef int8_t FN(elem_t);
#define E_LANES 32
#define E_MIN INT8_MIN
#define E_MAX INT8_MAX
#define V_ADDS _mm256_adds_epi8
#define V_SUBS _mm256_subs_epi8
#define V_MAX _mm256_max_epi8
#define V_MIN _mm256_min_epi8
#define V_GT _mm256_cmpgt_epi8
#define V_SET1(x) _mm256_set1_epi8((char)(x))
#else
typedef int16_t FN(elem_t);
#define E_LANES 16
#define E_MIN INT16_MIN
#define E_MAX INT16_MAX
#define V_ADDS _mm256_adds_epi16
#define V_SUBS _mm256_subs_epi16
#define V_MAX _mm256_max_epi16
#define V_MIN _mm256_min_epi16
#define V_GT _mm256_cmpgt_epi16
#define V_SET1(x) _mm256_set1_epi16((short)(x))
#endif
#define E_BYTES ((int)sizeof(FN(elem_t)))

/* every element moved up one lane (towards the high end), `fill` into
 * lane 0: the low 128-bit half's top element crosses into the high half */
__attribute__((target("avx2"))) static inline __m256i FN(up1)(__m256i v, __m256i lane0,
                                                             __m256i fill)
{
    const __m256i lo_up = _mm256_permute2x128_si256(v, v, 0x08); /* (0, v.lo) */
    const __m256i s = _mm256_alignr_epi8(v, lo_up, 16 - E_BYTES);
    return _mm256_blendv_epi8(s, fill, lane0);
}

static inline int FN(clampe)(long v) { return v < E_MIN ? E_MIN : v > E_MAX ? E_MAX : (int)v; }

__attribute__((target("avx2"))) static int FN(st_align)(int alg, int s1b, int s1e, int s2b,
                                                       int s2e, const int *q, int m,
                                                       const int *t, int n, int open, int gap,
                                                       const orc_matrix_t *mx, orc_result_t *out)
{
    typedef FN(elem_t) elem;
    const int L = E_LANES, seg = (m + L - 1) / L, sz = mx->size;
    const size_t vb = sizeof(__m256i) * (size_t)seg;
    __m256i *prof = (__m256i *)aligned_alloc(32, vb * sz);
    __m256i *buf = (__m256i *)aligned_alloc(32, vb * 4);
    if (!prof || !buf) {
        free(prof);
        free(buf);
        return -1;
    }
    __m256i *hs = buf, *hl = buf + seg, *ev = buf + 2 * seg, *hkeep = buf + 3 * seg;

    /* the query profile: for residue c, segment i, lane k the score of row
     * k*seg + i against c (0 past the query) */
    elem *pe = (elem *)prof;
    for (int c = 0; c < sz; ++c)
        for (int i = 0; i < seg; ++i)
            for (int k = 0; k < L; ++k) {
                const int row = k * seg + i;
                pe[((size_t)c * seg + i) * L + k] =
                    (elem)(row < m ? FN(clampe)(mx->matrix[(size_t)q[row] * sz + c]) : 0);
            }

    const __m256i vO = V_SET1(open), vG = V_SET1(gap);
    __m256i lane0 = _mm256_setzero_si256();
    {
        elem one[E_LANES];
        memset(one, 0, sizeof(one));
        one[0] = (elem)-1;
        lane0 = _mm256_loadu_si256((const __m256i *)one);
    }
    int sat = 0;

    if (alg == ORC_SW) {
        /* H, E and F biased by E_MIN: a score of 0 is E_MIN and the local
         * floor comes from saturating arithmetic */
        const __m256i vB = V_SET1(E_MIN);
        const int maxp = E_MAX - (mx->max + 1);
        for (int i = 0; i < seg; ++i) hs[i] = ev[i] = hkeep[i] = vB;
        __m256i vMax = vB, vMaxSeen = vB;
        int best = E_MIN, end_ref = 0;
        for (int j = 0; j < n; ++j) {
            __m256i vF = vB;
            __m256i vH = FN(up1)(hs[seg - 1], lane0, vB);
            __m256i *sw = hl;
            hl = hs;
            hs = sw;
            const __m256i *vP = prof + (size_t)t[j] * seg;
            for (int i = 0; i < seg; ++i) {
                __m256i vE = ev[i];
                vH = V_MAX(V_MAX(V_ADDS(vH, vP[i]), vE), vF);
                hs[i] = vH;
                vMax = V_MAX(vMax, vH);
                vH = V_SUBS(vH, vO);
                ev[i] = V_MAX(V_SUBS(vE, vG), vH);
                vF = V_MAX(V_SUBS(vF, vG), vH);
                vH = hl[i];
            }
            for (int k = 0; k < L; ++k) {
                vF = FN(up1)(vF, lane0, vB);
                for (int i = 0; i < seg; ++i) {
                    __m256i vH2 = V_MAX(hs[i], vF);
                    hs[i] = vH2;
                    vMax = V_MAX(vMax, vH2);
                    vH2 = V_SUBS(vH2, vO);
                    vF = V_SUBS(vF, vG);
                    if (!_mm256_movemask_epi8(V_GT(vF, vH2))) goto lazy_done;
                }
            }
        lazy_done:
            if (_mm256_movemask_epi8(V_GT(vMax, vMaxSeen))) {
                elem lanes[E_LANES];
                _mm256_storeu_si256((__m256i *)lanes, vMax);
                int s = E_MIN;
                for (int k = 0; k < L; ++k) s = lanes[k] > s ? lanes[k] : s;
                if (s > maxp) {
                    sat = 1;
                    break;
                }
                best = s;
                end_ref = j;
                vMaxSeen = V_SET1(s);
                memcpy(hkeep, hs, vb);
            }
        }
        if (!sat) {
            /* the smallest row holding the maximum in its column */
            const elem *h = (const elem *)hkeep;
            int end_query = m - 1;
            for (int x = 0; x < seg * L; ++x) {
                const int row = (x % L) * seg + x / L;
                if (h[x] == best && row < end_query) end_query = row;
            }
            out->score = best - E_MIN;
            out->end_query = end_query;
            out->end_ref = end_ref;
        }
    } else {
        const int neg = (-open < mx->min ? E_MIN + open : E_MIN - mx->min) + 1;
        const int pos = E_MAX - mx->max - 1;
        const int sg = alg == ORC_SG;
        const int b1 = sg && s1b, b2 = sg && s2b, e1 = sg && s1e, e2 = sg && s2e;
        elem *hs0 = (elem *)hs, *ev0 = (elem *)ev;
        for (int i = 0; i < seg; ++i)
            for (int k = 0; k < L; ++k) {
                const long h = b1 ? 0 : -(long)open - (long)gap * (k * seg + i);
                hs0[i * L + k] = (elem)FN(clampe)(h < E_MIN ? E_MIN : h);
                ev0[i * L + k] = (elem)FN(clampe)(h - open < E_MIN ? E_MIN : h - open);
            }
        const __m256i vNeg = V_SET1(neg);
        __m256i vLo = V_SET1(pos), vHi = vNeg;
        const int last_seg = (m - 1) % seg, last_lane = (m - 1) / seg;
        int row_best = neg, end_ref = n - 1;
        for (int j = 0; j < n; ++j) {
            const long up = j == 0 ? 0 : b2 ? 0 : -(long)open - (long)gap * (j - 1);
            __m256i vF = vNeg;
            __m256i vH = FN(up1)(hs[seg - 1], lane0, V_SET1(FN(clampe)(up)));
            __m256i *sw = hl;
            hl = hs;
            hs = sw;
            const __m256i *vP = prof + (size_t)t[j] * seg;
            for (int i = 0; i < seg; ++i) {
                __m256i vE = ev[i];
                vH = V_MAX(V_MAX(V_ADDS(vH, vP[i]), vE), vF);
                hs[i] = vH;
                vHi = V_MAX(vHi, vH);
                vLo = V_MIN(V_MIN(V_MIN(vLo, vH), vE), vF);
                vH = V_SUBS(vH, vO);
                ev[i] = V_MAX(V_SUBS(vE, vG), vH);
                vF = V_MAX(V_SUBS(vF, vG), vH);
                vH = hl[i];
            }
            /* F entering row 0 from the boundary H of this column */
            const long f0 = b2 ? -(long)open : -(long)open - (long)gap * j - open;
            const __m256i vF0 = V_SET1(FN(clampe)(f0));
            for (int k = 0; k < L; ++k) {
                vF = FN(up1)(vF, lane0, vF0);
                for (int i = 0; i < seg; ++i) {
                    __m256i vH2 = V_MAX(hs[i], vF);
                    hs[i] = vH2;
                    vLo = V_MIN(vLo, vH2);
                    vHi = V_MAX(vHi, vH2);
                    vH2 = V_SUBS(vH2, vO);
                    vF = V_SUBS(vF, vG);
                    if (!_mm256_movemask_epi8(V_GT(vF, vH2))) goto lazy_done2;
                }
            }
        lazy_done2:;
            const int last = ((const elem *)hs)[last_seg * L + last_lane];
            if (last > row_best) {
                row_best = last;
                end_ref = j;
            }
        }
        const __m256i bad = _mm256_or_si256(V_GT(vNeg, vLo), V_GT(vHi, V_SET1(pos)));
        if (_mm256_movemask_epi8(bad)) {
            sat = 1;
        } else {
            const elem *h = (const elem *)hs;
            int score = neg, end_query = m - 1;
            if (e2) score = row_best;
            if (e1) {
                for (int x = 0; x < seg * L; ++x) {
                    const int row = (x % L) * seg + x / L;
                    if (row >= m) continue;
                    if (h[x] > score) {
                        score = h[x];
                        end_query = row;
                        end_ref = n - 1;
                    } else if (h[x] == score && end_ref == n - 1 && row < end_query) {
                        end_query = row;
                    }
                }
            }
            if (!e1 && !e2) {
                score = h[last_seg * L + last_lane];
                end_query = m - 1;
                end_ref = n - 1;
            }
            out->score = score;
            out->end_query = end_query;
            out->end_ref = end_ref;
        }
    }
    free(prof);
    free(buf);
    return sat;
}

#undef E_LANES
#undef E_MIN
#undef E_MAX
#undef V_ADDS
#undef V_SUBS
#undef V_MAX
#undef V_MIN
#undef V_GT
#undef V_SET1
#undef E_BYTES
#undef FN
