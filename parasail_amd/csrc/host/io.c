/* FASTA/FASTQ ingest: the records of a file into parasail_sequences_t, and
 * their residues packed back to back for upload into an HBM-resident batch.
 *
 * parasail_sequences_from_file follows src/parser.c:48-221, whose parsing is
 * contrib/kseq.h:177-218 (klib kseq, vendored in the reference):
 *  - text before the first '>' or '@' is skipped;
 *  - the name runs from after the header character to the first whitespace
 *    character; when that character is not '\n', the rest of the line is the
 *    comment (a trailing '\r' dropped);
 *  - sequence lines are appended until a line starts with '>', '@' or '+';
 *    empty lines are skipped and a trailing '\r' is dropped;
 *  - after '+' (FASTQ) the rest of that line is skipped and quality lines are
 *    appended until the quality is as long as the sequence; a record whose
 *    quality ends short ends the parse, and the records before it are
 *    returned (the reference's loop stops at kseq's -2 without an error);
 *  - the file name "stdin" (any name starting with it) reads standard input.
 * Length statistics are Welford's running mean and M2, stddev = sqrt(M2/(n-1))
 * (parasail/stats.h:36-63).  The file is read through zlib, as the
 * reference's HAVE_ZLIB build reads it (src/parser.c:50-54,86-114): gzip
 * input is decompressed, other input is read as it is.
 * parasail_sequences_pack follows src/parser.c:235-259.
 */
#include <ctype.h>
#include <math.h>
#include <unistd.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <zlib.h>

#include "internal.h"

static char *read_all(gzFile fp, size_t *n)
{
    size_t cap = 1 << 20, len = 0;
    char *buf = (char *)malloc(cap);
    if (!buf) return NULL;
    for (;;) {
        if (len == cap) {
            char *nb = (char *)realloc(buf, cap * 2);
            if (!nb) { free(buf); return NULL; }
            buf = nb;
            cap *= 2;
        }
        const size_t want = cap - len > (1u << 30) ? (1u << 30) : cap - len;
        const int got = gzread(fp, buf + len, (unsigned)want);
        if (got < 0) { free(buf); return NULL; }
        len += (size_t)got;
        if (got == 0) break;
    }
    *n = len;
    return buf;
}

/* a growable byte string */
typedef struct { char *s; size_t l, m; } gstr_t;

static int gput(gstr_t *g, const char *p, size_t n)
{
    if (g->l + n + 1 > g->m) {
        size_t m = g->m ? g->m : 256;
        while (m < g->l + n + 1) m *= 2;
        char *ns = (char *)realloc(g->s, m);
        if (!ns) return -1;
        g->s = ns;
        g->m = m;
    }
    memcpy(g->s + g->l, p, n);
    g->l += n;
    g->s[g->l] = 0;
    return 0;
}

static int set_string(parasail_string_t *d, const gstr_t *g)
{
    d->l = g->l;
    d->s = NULL;
    if (!g->l) return 0;
    d->s = (char *)malloc(g->l + 1);
    if (!d->s) return -1;
    memcpy(d->s, g->s, g->l + 1);
    return 0;
}

/* end of the line starting at i (index of its '\n' or of the buffer end) */
static size_t line_end(const char *b, size_t n, size_t i)
{
    const char *e = (const char *)memchr(b + i, '\n', n - i);
    return e ? (size_t)(e - b) : n;
}

/* the line [i, e) without one trailing '\r' */
static size_t strip_cr(const char *b, size_t i, size_t e)
{
    return e > i && b[e - 1] == '\r' ? e - 1 : e;
}

PHOST_EXPORT parasail_sequences_t *parasail_sequences_from_file(const char *filename)
{
    if (!filename) {
        fprintf(stderr, "parasail_sequences_from_file: failed\n");
        return NULL;
    }
    const int from_stdin = strncmp("stdin", filename, 5) == 0;
    gzFile fp = from_stdin ? gzdopen(dup(fileno(stdin)), "r") : gzopen(filename, "r");
    if (!fp) {
        perror(from_stdin ? "gzdopen" : "gzopen");
        fprintf(stderr, "parasail_sequences_from_file: failed\n");
        return NULL;
    }
    size_t n = 0;
    char *b = read_all(fp, &n);
    gzclose(fp);
    parasail_sequences_t *ret = (parasail_sequences_t *)calloc(1, sizeof(*ret));
    size_t cap = 1000, count = 0, chars = 0;
    parasail_sequence_t *seqs = (parasail_sequence_t *)malloc(sizeof(*seqs) * cap);
    gstr_t name = {0}, comment = {0}, seq = {0}, qual = {0};
    unsigned long sn = 0;
    double mean = 0, m2 = 0, mn = 0, mx = 0;
    int failed = !b || !ret || !seqs;

    size_t i = 0;
    while (!failed && i < n && b[i] != '>' && b[i] != '@') ++i;
    while (!failed && i < n) {
        ++i; /* the header character */
        if (i >= n) break; /* a header character at the end: no record */
        name.l = comment.l = seq.l = qual.l = 0;
        size_t e = line_end(b, n, i), k = i;
        while (k < e && !isspace((unsigned char)b[k])) ++k;
        if (gput(&name, b + i, k - i)) { failed = 1; break; }
        if (k < e) gput(&comment, b + k + 1, strip_cr(b, k + 1, e) - (k + 1));
        i = e < n ? e + 1 : n;
        /* sequence lines up to the next line starting with '>', '@' or '+' */
        while (i < n && b[i] != '>' && b[i] != '@' && b[i] != '+') {
            e = line_end(b, n, i);
            if (e > i && gput(&seq, b + i, strip_cr(b, i, e) - i)) { failed = 1; break; }
            i = e < n ? e + 1 : n;
        }
        if (failed) break;
        int ok = 1;
        if (i < n && b[i] == '+') {
            e = line_end(b, n, i);
            if (e >= n) ok = 0; /* no quality line */
            i = e < n ? e + 1 : n;
            while (ok && i < n && qual.l < seq.l) {
                e = line_end(b, n, i);
                if (gput(&qual, b + i, strip_cr(b, i, e) - i)) { failed = 1; break; }
                i = e < n ? e + 1 : n;
            }
            if (qual.l != seq.l) ok = 0;
            /* the next record starts at the next '>' or '@' */
            while (ok && i < n && b[i] != '>' && b[i] != '@') ++i;
        }
        if (failed || !ok) break;
        parasail_sequence_t *r = &seqs[count];
        memset(r, 0, sizeof(*r));
        if (set_string(&r->name, &name) || set_string(&r->comment, &comment) ||
            set_string(&r->seq, &seq) || set_string(&r->qual, &qual)) {
            ++count;
            failed = 1;
            break;
        }
        const double x = (double)seq.l;
        mn = sn == 0 || x < mn ? x : mn;
        mx = sn == 0 || x > mx ? x : mx;
        ++sn;
        const double d = x - mean;
        mean += d / (double)sn;
        m2 += d * (x - mean);
        chars += seq.l;
        if (++count >= cap) {
            parasail_sequence_t *ns = (parasail_sequence_t *)realloc(seqs, sizeof(*seqs) * cap * 2);
            if (!ns) { failed = 1; break; }
            seqs = ns;
            cap *= 2;
        }
    }
    free(b);
    free(name.s); free(comment.s); free(seq.s); free(qual.s);
    if (ret) {
        ret->seqs = seqs;
        ret->l = count;
        ret->characters = chars;
        ret->shortest = (size_t)mn;
        ret->longest = (size_t)mx;
        ret->mean = (float)mean;
        ret->stddev = (float)sqrt(m2 / (double)(sn - 1));
    }
    if (failed) {
        fprintf(stderr, "parasail_sequences_from_file: failed\n");
        if (ret) parasail_sequences_free(ret);
        else free(seqs);
        return NULL;
    }
    return ret;
}

PHOST_EXPORT void parasail_sequences_free(parasail_sequences_t *sequences)
{
    if (!sequences) return;
    for (size_t i = 0; sequences->seqs && i < sequences->l; ++i) {
        free(sequences->seqs[i].name.s);
        free(sequences->seqs[i].comment.s);
        free(sequences->seqs[i].seq.s);
        free(sequences->seqs[i].qual.s);
    }
    free(sequences->seqs);
    free(sequences);
}

PHOST_EXPORT char *parasail_sequences_pack(const parasail_sequences_t *sequences, size_t *size)
{
    if (!sequences || !size) return NULL;
    char *packed = (char *)malloc(sequences->characters + sequences->l + 1);
    if (!packed) {
        perror("malloc");
        fprintf(stderr, "parasail_sequences_pack: failed\n");
        return NULL;
    }
    size_t off = 0;
    for (size_t i = 0; i < sequences->l; ++i) {
        if (sequences->seqs[i].seq.l) memcpy(packed + off, sequences->seqs[i].seq.s, sequences->seqs[i].seq.l);
        off += sequences->seqs[i].seq.l;
        packed[off++] = '$';
    }
    packed[off] = 0;
    *size = off;
    return packed;
}
