/*
 * oracle.c -- CPU restatement of parasail's serial recurrences.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Each function cites the
 * reference file:line whose behaviour it restates.  Written from the
 * reference's algorithm description; the loop structure mirrors the
 * reference's i-outer / j-inner scan because the end-position tie-breaks are
 * defined by that scan order.
 */
#include "oracle.h"

#include <ctype.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAX2(a, b) ((a) > (b) ? (a) : (b))

/* ------------------------------------------------------------------ */
/* matrices                                                             */
/* ------------------------------------------------------------------ */

static void build_mapper(int *mapper, const char *alphabet, int size,
                         int case_sensitive)
{
    /* unknown characters map to the trailing '*' column (memory.c:383-395) */
    for (int c = 0; c < 256; ++c) mapper[c] = size;
    for (int i = 0; i < size; ++i) {
        unsigned char a = (unsigned char)alphabet[i];
        if (case_sensitive) {
            mapper[a] = i;
        } else {
            mapper[toupper(a)] = i;
            mapper[tolower(a)] = i;
        }
    }
}

/* parasail_matrix_create_internal, src/memory.c:336-407 */
int orc_matrix_create(orc_matrix_t *out, const char *alphabet, int match,
                      int mismatch, int case_sensitive)
{
    if (!out || !alphabet) return -1;
    int size = (int)strlen(alphabet);
    if (size > 255) return -1;
    int size1 = size + 1;
    memset(out, 0, sizeof(*out));
    out->matrix = (int *)calloc((size_t)size1 * size1, sizeof(int));
    if (!out->matrix) return -1;
    for (int i = 0; i < size; ++i)
        for (int j = 0; j < size; ++j)
            out->matrix[i * size1 + j] = (i == j) ? match : mismatch;
    /* last row and column stay 0 */
    build_mapper(out->mapper, alphabet, size, case_sensitive);
    memcpy(out->alphabet, alphabet, (size_t)size);
    out->alphabet[size] = '*';
    out->alphabet[size1] = '\0';
    out->size = size1;
    out->length = size1;
    out->type = 0;
    out->max = match > mismatch ? match : mismatch;
    out->min = match > mismatch ? mismatch : match;
    return 0;
}

/* parasail_matrix_pssm_create_internal, src/memory.c:421-498 */
int orc_matrix_pssm_create(orc_matrix_t *out, const char *alphabet,
                           const int *values, int length, int case_sensitive)
{
    if (!out || !alphabet || !values || length <= 0) return -1;
    int size = (int)strlen(alphabet);
    if (size > 255) return -1;
    int size1 = size + 1;
    int mn = INT32_MAX, mx = INT32_MIN;
    for (long k = 0; k < (long)size * length; ++k) {
        if (values[k] < mn) mn = values[k];
        if (values[k] > mx) mx = values[k];
    }
    memset(out, 0, sizeof(*out));
    out->matrix = (int *)calloc((size_t)size1 * length, sizeof(int));
    if (!out->matrix) return -1;
    long v = 0;
    for (int r = 0; r < length; ++r) {
        for (int c = 0; c < size; ++c) out->matrix[(long)r * size1 + c] = values[v++];
        out->matrix[(long)r * size1 + size] = mn; /* extra column holds min */
    }
    build_mapper(out->mapper, alphabet, size, case_sensitive);
    memcpy(out->alphabet, alphabet, (size_t)size);
    out->alphabet[size] = '*';
    out->alphabet[size1] = '\0';
    out->size = size1;
    out->length = length;
    out->type = 1;
    out->max = mx;
    out->min = mn;
    return 0;
}

/* BLOSUM62 in the reference's residue order ARNDCQEGHILKMFPSTWYVBZX*
 * (parasail/matrices/blosum62.h:29-69; the published NCBI BLOSUM62 with a
 * '*' row/column of -4 and *,* = 1).  Data, not code: the table has to be
 * identical for results to be identical. */
static const signed char BLOSUM62[24 * 24] = {
     4,-1,-2,-2, 0,-1,-1, 0,-2,-1,-1,-1,-1,-2,-1, 1, 0,-3,-2, 0,-2,-1, 0,-4,
    -1, 5, 0,-2,-3, 1, 0,-2, 0,-3,-2, 2,-1,-3,-2,-1,-1,-3,-2,-3,-1, 0,-1,-4,
    -2, 0, 6, 1,-3, 0, 0, 0, 1,-3,-3, 0,-2,-3,-2, 1, 0,-4,-2,-3, 3, 0,-1,-4,
    -2,-2, 1, 6,-3, 0, 2,-1,-1,-3,-4,-1,-3,-3,-1, 0,-1,-4,-3,-3, 4, 1,-1,-4,
     0,-3,-3,-3, 9,-3,-4,-3,-3,-1,-1,-3,-1,-2,-3,-1,-1,-2,-2,-1,-3,-3,-2,-4,
    -1, 1, 0, 0,-3, 5, 2,-2, 0,-3,-2, 1, 0,-3,-1, 0,-1,-2,-1,-2, 0, 3,-1,-4,
    -1, 0, 0, 2,-4, 2, 5,-2, 0,-3,-3, 1,-2,-3,-1, 0,-1,-3,-2,-2, 1, 4,-1,-4,
     0,-2, 0,-1,-3,-2,-2, 6,-2,-4,-4,-2,-3,-3,-2, 0,-2,-2,-3,-3,-1,-2,-1,-4,
    -2, 0, 1,-1,-3, 0, 0,-2, 8,-3,-3,-1,-2,-1,-2,-1,-2,-2, 2,-3, 0, 0,-1,-4,
    -1,-3,-3,-3,-1,-3,-3,-4,-3, 4, 2,-3, 1, 0,-3,-2,-1,-3,-1, 3,-3,-3,-1,-4,
    -1,-2,-3,-4,-1,-2,-3,-4,-3, 2, 4,-2, 2, 0,-3,-2,-1,-2,-1, 1,-4,-3,-1,-4,
    -1, 2, 0,-1,-3, 1, 1,-2,-1,-3,-2, 5,-1,-3,-1, 0,-1,-3,-2,-2, 0, 1,-1,-4,
    -1,-1,-2,-3,-1, 0,-2,-3,-2, 1, 2,-1, 5, 0,-2,-1,-1,-1,-1, 1,-3,-1,-1,-4,
    -2,-3,-3,-3,-2,-3,-3,-3,-1, 0, 0,-3, 0, 6,-4,-2,-2, 1, 3,-1,-3,-3,-1,-4,
    -1,-2,-2,-1,-3,-1,-1,-2,-2,-3,-3,-1,-2,-4, 7,-1,-1,-4,-3,-2,-2,-1,-2,-4,
     1,-1, 1, 0,-1, 0, 0, 0,-1,-2,-2, 0,-1,-2,-1, 4, 1,-3,-2,-2, 0, 0, 0,-4,
     0,-1, 0,-1,-1,-1,-1,-2,-2,-1,-1,-1,-1,-2,-1, 1, 5,-2,-2, 0,-1,-1, 0,-4,
    -3,-3,-4,-4,-2,-2,-3,-2,-2,-3,-2,-3,-1, 1,-4,-3,-2,11, 2,-3,-4,-3,-2,-4,
    -2,-2,-2,-3,-2,-1,-2,-3, 2,-1,-1,-2,-1, 3,-3,-2,-2, 2, 7,-1,-3,-2,-1,-4,
     0,-3,-3,-3,-1,-2,-2,-3,-3, 3, 1,-2, 1,-1,-2,-2, 0,-3,-1, 4,-3,-2,-1,-4,
    -2,-1, 3, 4,-3, 0, 1,-1, 0,-3,-4, 0,-3,-3,-2, 0,-1,-4,-3,-3, 4, 1,-1,-4,
    -1, 0, 0, 1,-3, 3, 4,-2, 0,-3,-3, 1,-1,-3,-1, 0,-1,-3,-2,-2, 1, 4,-1,-4,
     0,-1,-1,-1,-2,-1,-1,-1,-1,-1,-1,-1,-1,-1,-2, 0, 0,-2,-1,-1,-1,-1,-1,-4,
    -4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4, 1,
};

/* parasail_blosum62 + parasail_blosum_map (blosum62.h:57-69,
 * blosum_map.h:15-32): letters of the alphabet map case-insensitively to
 * their index, everything else to 23 ('*'). */
int orc_matrix_blosum62(orc_matrix_t *out)
{
    static const char *alpha = "ARNDCQEGHILKMFPSTWYVBZX*";
    memset(out, 0, sizeof(*out));
    out->matrix = (int *)malloc(24 * 24 * sizeof(int));
    if (!out->matrix) return -1;
    for (int k = 0; k < 24 * 24; ++k) out->matrix[k] = BLOSUM62[k];
    for (int c = 0; c < 256; ++c) out->mapper[c] = 23;
    for (int i = 0; i < 23; ++i) {
        out->mapper[toupper((unsigned char)alpha[i])] = i;
        out->mapper[tolower((unsigned char)alpha[i])] = i;
    }
    strcpy(out->alphabet, alpha);
    out->size = 24;
    out->length = 24;
    out->type = 0;
    out->max = 11;
    out->min = -4;
    return 0;
}

/* A square matrix from a value table in the layout of the reference's
 * built-in matrices (parasail/matrices/ headers: row-major size*size, alphabet
 * ending in '*', mapper case-insensitive with unknown -> last index). */
int orc_matrix_square_from_values(orc_matrix_t *out, const char *alphabet,
                                  const int *values, int size, int max, int min)
{
    if (!out || !alphabet || !values || size <= 0 || size > 256 ||
        (int)strlen(alphabet) != size)
        return -1;
    memset(out, 0, sizeof(*out));
    out->matrix = (int *)malloc(sizeof(int) * (size_t)size * size);
    if (!out->matrix) return -1;
    memcpy(out->matrix, values, sizeof(int) * (size_t)size * size);
    for (int c = 0; c < 256; ++c) out->mapper[c] = size - 1;
    for (int i = 0; i < size - 1; ++i) {
        out->mapper[toupper((unsigned char)alphabet[i])] = i;
        out->mapper[tolower((unsigned char)alphabet[i])] = i;
    }
    memcpy(out->alphabet, alphabet, (size_t)size);
    out->size = size;
    out->length = size;
    out->type = 0;
    out->max = max;
    out->min = min;
    return 0;
}

/* parasail_matrix_convert_square_to_pssm, src/memory.c:567-621 */
int orc_matrix_convert_square_to_pssm(orc_matrix_t *out, const orc_matrix_t *sq,
                                      const char *s1, int s1Len)
{
    if (!sq || sq->type != 0 || !s1 || s1Len <= 0) return -1;
    memset(out, 0, sizeof(*out));
    out->matrix = (int *)calloc((size_t)sq->size * s1Len, sizeof(int));
    if (!out->matrix) return -1;
    for (int i = 0; i < s1Len; ++i) {
        int r = sq->mapper[(unsigned char)s1[i]];
        memcpy(&out->matrix[(long)sq->size * i], &sq->matrix[(long)sq->size * r],
               sizeof(int) * (size_t)sq->size);
    }
    memcpy(out->mapper, sq->mapper, sizeof(out->mapper));
    memcpy(out->alphabet, sq->alphabet, sizeof(out->alphabet));
    out->size = sq->size;
    out->length = s1Len;
    out->type = 1;
    out->max = sq->max;
    out->min = sq->min;
    return 0;
}

void orc_matrix_free(orc_matrix_t *m)
{
    if (m && m->matrix) {
        free(m->matrix);
        m->matrix = NULL;
    }
}

void orc_map(const orc_matrix_t *mx, const char *s, int len, int *out)
{
    for (int k = 0; k < len; ++k) out[k] = mx->mapper[(unsigned char)s[k]];
}

/* ------------------------------------------------------------------ */
/* serial DP                                                            */
/* ------------------------------------------------------------------ */

/*
 * One restatement of the serial functions, switched on alg and kind:
 *   SW: src/sw.c:29-185, src/sw_stats.c:29-298, src/sw_trace.c:60-163
 *   NW: src/nw.c:29-167, src/nw_stats.c,         src/nw_trace.c
 *   SG: src/sg.c:31-253 (parasail_sg_flags), src/sg_stats.c, src/sg_trace.c
 * The per-cell recurrence (sw.c:131-149): F from the cell above, E from the
 * cell to the left, H = max(diag + M, E, F[, 0]).  Stats carry
 * (matches, similar, length) with source priority diag > F > E and the
 * strict open>extend restart (sw_stats.c:204-236).  Trace codes follow
 * sg_trace.c:144-150 and sw_trace.c:145-147.
 */
int orc_align(int alg, int kind, int s1_beg, int s1_end, int s2_beg, int s2_end,
              const int *s1, int m, const int *s2, int n, int open, int gap,
              const orc_matrix_t *mx, orc_result_t *res, int8_t *trace)
{
    return orc_align_tables(alg, kind, s1_beg, s1_end, s2_beg, s2_end, s1, m, s2, n, open, gap,
                            mx, res, trace, NULL);
}

/* As orc_align; with tabs non-NULL also the PARASAIL_TABLE outputs
 * (sw.c:159-170, sg_stats.c:258-263): every cell's H into tabs[0] and, for
 * ORC_STATS, its matches/similar/length into tabs[1..3], m*n ints row-major. */
int orc_align_tables(int alg, int kind, int s1_beg, int s1_end, int s2_beg, int s2_end,
                     const int *s1, int m, const int *s2, int n, int open, int gap,
                     const orc_matrix_t *mx, orc_result_t *res, int8_t *trace, int **tabs)
{
    /* PARASAIL_CHECK_* (sw.c:47-56) */
    if (!s2 || n <= 0 || open < 0 || gap < 0 || !mx || !res) return -1;
    const int pssm = mx->type == 1;
    if (!pssm && (!s1 || m <= 0)) return -1;
    if (kind == ORC_STATS && !s1) return -1; /* stats compare s1==s2 (sw_stats.c:62-64) */
    if (pssm) m = mx->length;
    if (kind == ORC_TRACE && !trace) return -1;
    if (alg != ORC_SG) s1_beg = s1_end = s2_beg = s2_end = 0;

    int *H = (int *)malloc(sizeof(int) * (size_t)(n + 1) * 8);
    if (!H) return -1;
    int *F = H + (n + 1);
    int *HM = F + (n + 1), *HS = HM + (n + 1), *HL = HS + (n + 1);
    int *FM = HL + (n + 1), *FS = FM + (n + 1), *FL = FS + (n + 1);

    int score = ORC_NEG_INF, end_query = m, end_ref = n;
    int matches = ORC_NEG_INF, similar = ORC_NEG_INF, length = ORC_NEG_INF;
    if (alg == ORC_SG && kind == ORC_STATS) { end_query = m - 1; end_ref = n - 1; } /* sg_stats.c:79-84 */

    /* first row (sw.c:102-112, nw.c:102-112, sg.c:110-124) */
    H[0] = 0;
    F[0] = ORC_NEG_INF;
    HM[0] = HS[0] = HL[0] = FM[0] = FS[0] = FL[0] = 0;
    for (int j = 1; j <= n; ++j) {
        if (alg == ORC_SW || (alg == ORC_SG && s2_beg)) H[j] = 0;
        else H[j] = -open - (j - 1) * gap;
        F[j] = ORC_NEG_INF;
        HM[j] = HS[j] = HL[j] = FM[j] = FS[j] = FL[j] = 0;
    }

    for (int i = 1; i <= m; ++i) {
        const int *matrow = pssm ? &mx->matrix[(long)mx->size * (i - 1)]
                                 : &mx->matrix[(long)mx->size * s1[i - 1]];
        int NH = H[0], NHM = HM[0], NHS = HS[0], NHL = HL[0];
        int WH;
        if (alg == ORC_SW || (alg == ORC_SG && s1_beg)) WH = 0;
        else WH = -open - (i - 1) * gap;
        int WHM = 0, WHS = 0, WHL = 0;
        int E = ORC_NEG_INF, EM = 0, ES = 0, EL = 0;
        H[0] = WH;
        HM[0] = HS[0] = HL[0] = 0;
        for (int j = 1; j <= n; ++j) {
            int NWH = NH, NWM = NHM, NWS = NHS, NWL = NHL;
            NH = H[j];
            NHM = HM[j]; NHS = HS[j]; NHL = HL[j];
            int F_opn = NH - open;
            int F_ext = F[j] - gap;
            F[j] = MAX2(F_opn, F_ext);
            int E_opn = WH - open;
            int E_ext = E - gap;
            E = MAX2(E_opn, E_ext);
            int sub = matrow[s2[j - 1]];
            int H_dag = NWH + sub;
            if (alg == ORC_SW) {
                WH = MAX2(H_dag, 0);
                WH = MAX2(WH, E);
                WH = MAX2(WH, F[j]);
            } else {
                WH = MAX2(H_dag, E);
                WH = MAX2(WH, F[j]);
            }
            if (kind == ORC_STATS) {
                if (F_opn > F_ext) { FM[j] = NHM; FS[j] = NHS; FL[j] = NHL; }
                FL[j] += 1;
                if (E_opn > E_ext) { EM = WHM; ES = WHS; EL = WHL; }
                EL += 1;
                if (WH == H_dag) {
                    WHM = NWM + (s1[i - 1] == s2[j - 1]);
                    WHS = NWS + (sub > 0);
                    WHL = NWL + 1;
                } else if (WH == F[j]) {
                    WHM = FM[j]; WHS = FS[j]; WHL = FL[j];
                } else {
                    WHM = EM; WHS = ES; WHL = EL;
                }
                if (alg == ORC_SW && WH <= 0) { WH = 0; WHM = WHS = WHL = 0; }
            } else if (kind == ORC_TRACE) {
                int8_t code = (F_opn > F_ext) ? ORC_DIAG_F : ORC_DEL_F;
                code |= (E_opn > E_ext) ? ORC_DIAG_E : ORC_INS_E;
                code |= (WH == H_dag) ? ORC_DIAG : (WH == F[j]) ? ORC_DEL : ORC_INS;
                if (alg == ORC_SW && WH == 0) code &= ORC_ZERO_MASK;
                trace[(long)(i - 1) * n + (j - 1)] = code;
            }
            H[j] = WH;
            HM[j] = WHM; HS[j] = WHS; HL[j] = WHL;
            if (tabs) {
                const long at = (long)(i - 1) * n + (j - 1);
                tabs[0][at] = WH;
                if (kind == ORC_STATS) { tabs[1][at] = WHM; tabs[2][at] = WHS; tabs[3][at] = WHL; }
            }

            if (alg == ORC_SW) {
                /* sw.c:150-158: first strictly better cell, ties -> smaller j */
                if (WH > score || (WH == score && j - 1 < end_ref)) {
                    score = WH;
                    matches = WHM; similar = WHS; length = WHL;
                    end_query = i - 1;
                    end_ref = j - 1;
                }
            } else if (alg == ORC_SG && i == m && kind != ORC_STATS) {
                /* last row inside the DP (sg.c:206-223) */
                if (s1_end && s2_end) {
                    if (WH > score || (WH == score && j - 1 < end_ref)) {
                        score = WH; end_query = m - 1; end_ref = j - 1;
                    }
                } else if (s2_end) {
                    if (WH > score) { score = WH; end_query = m - 1; end_ref = j - 1; }
                }
            }
        }
        if (alg == ORC_SG) {
            if (kind == ORC_STATS) {
                /* sg_stats.c:252-259: every row, including the last */
                if (s1_end && WH > score) {
                    score = WH; matches = WHM; similar = WHS; length = WHL;
                    end_query = i - 1; end_ref = n - 1;
                }
            } else if (i < m) {
                /* sg.c:171-175 */
                if (s1_end && WH > score) { score = WH; end_query = i - 1; end_ref = n - 1; }
            } else {
                /* sg.c:231-235 */
                if ((s1_end && WH > score) || (!s1_end && !s2_end)) {
                    score = WH; end_query = m - 1; end_ref = n - 1;
                }
            }
        }
    }

    if (alg == ORC_NW) {
        score = H[n]; end_query = m - 1; end_ref = n - 1;
        matches = HM[n]; similar = HS[n]; length = HL[n];
    } else if (alg == ORC_SG && kind == ORC_STATS) {
        /* sg_stats.c:266-309: the last row after the DP */
        if (s1_end && s2_end) {
            for (int j = 1; j <= n; ++j) {
                if (H[j] > score || (H[j] == score && j - 1 < end_ref)) {
                    score = H[j]; matches = HM[j]; similar = HS[j]; length = HL[j];
                    end_query = m - 1; end_ref = j - 1;
                }
            }
        } else if (s2_end) {
            for (int j = 1; j <= n; ++j) {
                if (H[j] > score) {
                    score = H[j]; matches = HM[j]; similar = HS[j]; length = HL[j];
                    end_query = m - 1; end_ref = j - 1;
                }
            }
        }
        if (!s1_end && !s2_end) {
            score = H[n]; matches = HM[n]; similar = HS[n]; length = HL[n];
            end_query = m - 1; end_ref = n - 1;
        }
    }

    res->score = score;
    res->end_query = end_query;
    res->end_ref = end_ref;
    res->matches = kind == ORC_STATS ? matches : 0;
    res->similar = kind == ORC_STATS ? similar : 0;
    res->length = kind == ORC_STATS ? length : 0;
    free(H);
    return 0;
}

/*
 * parasail_nw_banded, src/nw_banded.c:29-189.  The band is stored in
 * diagonal coordinates: slot i of column j is row colOff+i, and colOff grows
 * by one per column.  Slots whose row falls outside [0,m) keep the value
 * they last held, and the band edges (slots -1 and colLen) stay at -inf.
 */
int orc_nw_banded(const int *s1, int m, const int *s2, int n, int open, int gap,
                  int k, const orc_matrix_t *mx, orc_result_t *res)
{
    if (!s1 || m <= 0 || !s2 || n <= 0 || open < 0 || gap < 0 || k <= 0 || !mx || !res)
        return -1;
    const int pssm = mx->type == 1;
    int low = (m > n) ? n - m - k : m - n - k;
    int colLen = k - low + 1;
    int colOff = (m > n) ? -k : low;
    int *HB = (int *)calloc((size_t)colLen + 2, sizeof(int));
    int *EB = (int *)calloc((size_t)colLen + 2, sizeof(int));
    if (!HB || !EB) { free(HB); free(EB); return -1; }
    int *Hb = HB + 1, *Eb = EB + 1;
    for (int i = -colOff + 1, j = 0; i < colLen; ++i, ++j) {
        Hb[i] = -open - j * gap;
        Eb[i] = ORC_NEG_INF;
    }
    /* these four stores may land on indices outside [-1, colLen]; the
     * reference writes them unguarded (nw_banded.c:126-133), so guard here
     * only where the reference would write out of its buffer */
    if (-colOff - 1 >= -1 && -colOff - 1 <= colLen) { Hb[-colOff - 1] = -open; Eb[-colOff - 1] = ORC_NEG_INF; }
    if (-colOff >= -1 && -colOff <= colLen) { Hb[-colOff] = 0; Eb[-colOff] = ORC_NEG_INF; }
    Hb[-1] = ORC_NEG_INF; Eb[-1] = ORC_NEG_INF;
    Hb[colLen] = ORC_NEG_INF; Eb[colLen] = ORC_NEG_INF;

    for (int j = 0; j < n; ++j) {
        int F = ORC_NEG_INF;
        if (colOff < 0) Hb[-colOff - 1] = -open - j * gap;
        for (int i = 0; i < colLen; ++i) {
            int pos = colOff + i;
            if (pos < 0 || pos >= m) continue;
            int matval = pssm ? mx->matrix[(long)mx->size * pos + s2[j]]
                              : mx->matrix[(long)mx->size * s1[pos] + s2[j]];
            int H_dag = Hb[i] + matval;
            int E_opn = Hb[i + 1] - open;
            int E_ext = Eb[i + 1] - gap;
            int F_opn = Hb[i - 1] - open;
            int F_ext = F - gap;
            Eb[i] = MAX2(E_opn, E_ext);
            F = MAX2(F_opn, F_ext);
            Hb[i] = MAX2(Eb[i], F);
            Hb[i] = MAX2(Hb[i], H_dag);
        }
        colOff += 1;
    }
    res->score = (m > n) ? Hb[-low] : Hb[k];
    res->end_query = m - 1;
    res->end_ref = n - 1;
    res->matches = res->similar = res->length = 0;
    free(HB);
    free(EB);
    return 0;
}

/* ------------------------------------------------------------------ */
/* CIGAR                                                                */
/* ------------------------------------------------------------------ */

static int encode_op(char c)
{
    /* parasail_cigar_encoded_ops, cigar.c:24-57 (BAM op codes "MIDNSHP=XB") */
    switch (c) {
    case 'M': return 0; case 'I': return 1; case 'D': return 2; case 'N': return 3;
    case 'S': return 4; case 'H': return 5; case 'P': return 6; case '=': return 7;
    case 'X': return 8; default: return 0;
    }
}

static int trace_at(const void *t, int striped, int bits, long loc)
{
    if (!striped) return ((const int8_t *)t)[loc];
    switch (bits) {
    case 8: return ((const int8_t *)t)[loc];
    case 16: return ((const int16_t *)t)[loc];
    case 32: return ((const int32_t *)t)[loc];
    default: return (int)((const int64_t *)t)[loc];
    }
}

typedef struct { uint32_t *seq; int len, cap, overflow; } opbuf_t;

static void put_op(opbuf_t *b, uint32_t cnt, char op)
{
    if (b->len < b->cap) b->seq[b->len] = (cnt << 4) | (uint32_t)encode_op(op);
    else b->overflow = 1;
    b->len++;
}

/* WRITE_ANY, cigar_template.c:38-56: I and D are deliberately swapped */
static void flush(opbuf_t *b, uint32_t *mat, uint32_t *mis, uint32_t *del, uint32_t *ins)
{
    if (*mat) put_op(b, *mat, '=');
    else if (*mis) put_op(b, *mis, 'X');
    else if (*del) put_op(b, *del, 'I');
    else if (*ins) put_op(b, *ins, 'D');
    *mat = *mis = *del = *ins = 0;
}

/* parasail_cigar_{,striped_}{8,16,32,64}, src/cigar_template.c:58-263 */
int orc_cigar(const char *seqA, int lena, const char *seqB, int lenb, int is_sg,
              int end_query, int end_ref, const void *trace, int striped,
              int bits, int lanes, int case_sensitive, const char *aliases_in,
              uint32_t *seq, int cap, int *beg_query, int *beg_ref)
{
    char aliases[256];
    size_t naliases = 0;
    uint32_t c_mat = 0, c_mis = 0, c_del = 0, c_ins = 0;
    long i = end_query, j = end_ref;
    int where = ORC_DIAG;
    long segWidth = lanes, segLen = striped ? (lena + lanes - 1) / lanes : 0;
    opbuf_t b = {seq, 0, cap, 0};
    if (aliases_in) {
        naliases = strlen(aliases_in);
        if (naliases % 2 || naliases >= 256) return -1;
        for (size_t a = 0; a < naliases; ++a)
            aliases[a] = case_sensitive ? aliases_in[a] : (char)toupper((unsigned char)aliases_in[a]);
    }
    if (is_sg) {
        /* cigar_template.c:121-142 */
        if (end_query + 1 == lena) {
            for (long kk = lenb - 1; kk > j; --kk) ++c_ins;
        } else if (end_ref + 1 == lenb) {
            for (long kk = lena - 1; kk > i; --kk) ++c_del;
        } else {
            return -1;
        }
    }
    while (i >= 0 || j >= 0) {
        long loc = striped ? j * segLen * segWidth + (i % (segLen ? segLen : 1)) * segWidth + (segLen ? i / segLen : 0)
                           : i * lenb + j;
        if (i < 0) {
            if (0 == c_ins) flush(&b, &c_mat, &c_mis, &c_del, &c_ins);
            while (j >= 0) { ++c_ins; --j; }
            break;
        }
        if (j < 0) {
            if (0 == c_del) flush(&b, &c_mat, &c_mis, &c_del, &c_ins);
            while (i >= 0) { ++c_del; --i; }
            break;
        }
        int code = trace_at(trace, striped, bits, loc);
        if (where == ORC_DIAG) {
            if (code & ORC_DIAG) {
                char a = case_sensitive ? seqA[i] : (char)toupper((unsigned char)seqA[i]);
                char c = case_sensitive ? seqB[j] : (char)toupper((unsigned char)seqB[j]);
                int match = (a == c);
                if (aliases_in) {
                    for (size_t p = 0; p < naliases; p += 2) {
                        if (aliases[p] == a) match |= aliases[p + 1] == c;
                        else if (aliases[p + 1] == a) match |= aliases[p] == c;
                    }
                }
                if (match) {
                    if (0 == c_mat) flush(&b, &c_mat, &c_mis, &c_del, &c_ins);
                    c_mat += 1;
                } else {
                    if (0 == c_mis) flush(&b, &c_mat, &c_mis, &c_del, &c_ins);
                    c_mis += 1;
                }
                --i;
                --j;
            } else if (code & ORC_INS) {
                where = ORC_INS;
            } else if (code & ORC_DEL) {
                where = ORC_DEL;
            } else {
                break; /* zero: local alignment starts here */
            }
        } else if (where == ORC_INS) {
            if (0 == c_ins) flush(&b, &c_mat, &c_mis, &c_del, &c_ins);
            c_ins += 1;
            --j;
            if (code & ORC_DIAG_E) where = ORC_DIAG;
            else if (code & ORC_INS_E) where = ORC_INS;
            else return -1;
        } else { /* ORC_DEL */
            if (0 == c_del) flush(&b, &c_mat, &c_mis, &c_del, &c_ins);
            c_del += 1;
            --i;
            if (code & ORC_DIAG_F) where = ORC_DIAG;
            else if (code & ORC_DEL_F) where = ORC_DEL;
            else return -1;
        }
    }
    flush(&b, &c_mat, &c_mis, &c_del, &c_ins);
    if (b.overflow) return -2;
    /* reverse (parasail_reverse_uint32_t, memory.c:767-779) */
    for (int a = 0, z = b.len - 1; a < z; ++a, --z) {
        uint32_t tmp = seq[a]; seq[a] = seq[z]; seq[z] = tmp;
    }
    *beg_query = (int)(i + 1);
    *beg_ref = (int)(j + 1);
    return b.len;
}

/* parasail_striped_unwind_T, src/striped_unwind.c:5-26 */
void orc_striped_unwind(const void *striped, int bits, int lanes, int m, int n, int *out)
{
    long segLen = (m + lanes - 1) / lanes;
    for (long j = 0; j < n; ++j)
        for (long i = 0; i < m; ++i) {
            long loc = j * segLen * lanes + (i % segLen) * lanes + i / segLen;
            out[i * n + j] = trace_at(striped, 1, bits, loc);
        }
}

/* ------------------------------------------------------------------ */
/* CPU baseline harness                                                 */
/* ------------------------------------------------------------------ */

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts); /* parasail_time, src/time.c:28-75 */
    return ts.tv_sec + ts.tv_nsec * 1e-9;
}

/* tests/test_openmp.c:383-441: parallel for schedule(guided) over pairs,
 * timed around the parallel region only. */
double orc_batch(int alg, int kind, int s1_beg, int s1_end, int s2_beg,
                 int s2_end, int npairs, const int *q, const int64_t *qoff,
                 const int *qlen, const int *t, const int64_t *toff,
                 const int *tlen, int open, int gap, const orc_matrix_t *mx,
                 orc_result_t *out, int threads)
{
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    double t0 = now_s();
#ifdef _OPENMP
#pragma omp parallel for schedule(guided)
#endif
    for (int p = 0; p < npairs; ++p) {
        int8_t *tr = NULL;
        if (kind == ORC_TRACE) tr = (int8_t *)malloc((size_t)qlen[p] * tlen[p]);
        orc_align(alg, kind, s1_beg, s1_end, s2_beg, s2_end, q + qoff[p], qlen[p],
                  t + toff[p], tlen[p], open, gap, mx, &out[p], tr);
        free(tr);
    }
    return now_s() - t0;
}
