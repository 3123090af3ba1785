/*
 * files.c -- the file layer in front of the alignment path: a FASTA/FASTQ
 * file mapped into memory (parasail_open/close/read), its statistics and
 * its residues packed back to back with a '$' after each record
 * (parasail_stat*, parasail_pack*, with _buffer forms for text already in
 * memory), and substitution matrices or PSSMs read from text files
 * (parasail_matrix_from_file[_case_sensitive]).
 *
 * The reading rules are those of src/io.c:53-800 and :804-1178:
 *  - FASTA: the first line must start with '>'; every later '>' line starts
 *    a record; letters count, line ends ("\n", "\r", "\r\n", "\n\r") are
 *    skipped, any other printable or non-printing character is an error;
 *    the last record must not be empty;
 *  - FASTQ: four-line records ('@' header, one sequence line, '+' line, one
 *    quality line); the statistics sample each record's length when its
 *    sequence line is read and again when the next record starts, and count
 *    no characters, as src/io.c:401-496 does;
 *  - matrix files: '#' lines are comments; the first other line is the
 *    alphabet (letters and '*'); every later line is an optional row letter
 *    (a row without one is a PSSM position, letter '*') followed by one
 *    integer per alphabet letter; identical row and column alphabets make a
 *    square matrix, anything else a PSSM whose length is its row count;
 *    unknown letters map to the alphabet's last index.
 * Reads never pass the end of the text (the reference's line scans assume a
 * final line end).
 */
#include <ctype.h>
#include <fcntl.h>
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include "internal.h"

/* running length statistics (parasail/stats.h:36-63) */
typedef struct {
    unsigned long n;
    double mean, m2, mn, mx;
} lstats_t;

static void lstats_add(lstats_t *s, double x)
{
    s->mn = s->n == 0 || x < s->mn ? x : s->mn;
    s->mx = s->n == 0 || x > s->mx ? x : s->mx;
    ++s->n;
    const double d = x - s->mean;
    s->mean += d / (double)s->n;
    s->m2 += d * (x - s->mean);
}

static parasail_file_stat_t *lstats_out(const lstats_t *s, unsigned long seqs,
                                        unsigned long chars)
{
    parasail_file_stat_t *p = (parasail_file_stat_t *)malloc(sizeof(*p));
    if (!p) {
        perror("malloc");
        return NULL;
    }
    p->sequences = seqs;
    p->characters = chars;
    p->shortest = (unsigned long)s->mn;
    p->longest = (unsigned long)s->mx;
    p->mean = (float)s->mean;
    p->stddev = (float)sqrt(s->m2 / (double)(s->n - 1));
    return p;
}

/* index of the end of the line holding T[i] (its last line-end character,
 * a "\r\n" or "\n\r" pair counted as one end) */
static off_t line_end_at(const char *T, off_t size, off_t i)
{
    while (i < size && T[i] != '\n' && T[i] != '\r') ++i;
    if (i + 1 < size && (T[i + 1] == '\n' || T[i + 1] == '\r')) ++i;
    return i;
}

PHOST_EXPORT parasail_file_t *parasail_open(const char *fname)
{
    if (!fname) {
        fprintf(stderr, "parasail_open: NULL filename\n");
        return NULL;
    }
    const int fd = open(fname, O_RDONLY);
    if (fd == -1) {
        perror("open");
        fprintf(stderr, "parasail_open: cannot open input file `%s'\n", fname);
        return NULL;
    }
    struct stat fs;
    if (fstat(fd, &fs) == -1) {
        perror("fstat");
        fprintf(stderr, "parasail_open: cannont stat input file `%s'\n", fname);
        close(fd);
        return NULL;
    }
    void *buf = mmap(NULL, (size_t)fs.st_size, PROT_READ, MAP_SHARED, fd, 0);
    if (buf == MAP_FAILED) {
        perror("mmap");
        fprintf(stderr, "parasail_open: cannont mmap input file `%s'\n", fname);
        close(fd);
        return NULL;
    }
    parasail_file_t *pf = (parasail_file_t *)malloc(sizeof(*pf));
    if (!pf) {
        perror("malloc");
        munmap(buf, (size_t)fs.st_size);
        close(fd);
        return NULL;
    }
    pf->fd = fd;
    pf->size = fs.st_size;
    pf->buf = (const char *)buf;
    return pf;
}

PHOST_EXPORT void parasail_close(parasail_file_t *pf)
{
    if (!pf) {
        fprintf(stderr, "parasail_close: NULL file\n");
        return;
    }
    if (munmap((void *)pf->buf, (size_t)pf->size) == -1) {
        perror("munmap");
        fprintf(stderr, "parasail_close: cannot munmap file buffer\n");
    }
    if (close(pf->fd) == -1) {
        perror("close");
        fprintf(stderr, "parasail_close: cannot close file descriptor\n");
    }
    free(pf);
}

PHOST_EXPORT int parasail_is_fasta_buffer(const char *buf, off_t size)
{
    if (!buf) {
        fprintf(stderr, "parasail_is_fasta_buffer: NULL pointer\n");
        return -1;
    }
    return size > 0 && buf[0] == '>';
}

PHOST_EXPORT int parasail_is_fastq_buffer(const char *buf, off_t size)
{
    if (!buf) {
        fprintf(stderr, "parasail_is_fastq_buffer: NULL pointer\n");
        return -1;
    }
    return size > 0 && buf[0] == '@';
}

PHOST_EXPORT int parasail_is_fasta(const parasail_file_t *pf)
{
    if (!pf) {
        fprintf(stderr, "parasail_is_fasta: NULL pointer\n");
        return -1;
    }
    return parasail_is_fasta_buffer(pf->buf, pf->size);
}

PHOST_EXPORT int parasail_is_fastq(const parasail_file_t *pf)
{
    if (!pf) {
        fprintf(stderr, "parasail_is_fastq: NULL pointer\n");
        return -1;
    }
    return parasail_is_fastq_buffer(pf->buf, pf->size);
}

/* One pass over FASTA text: statistics, and the packed residues into P when
 * P is given.  0, or -1 with a message from `who`. */
static int fasta_scan(const char *T, off_t size, const char *who, lstats_t *st,
                      unsigned long *seqs, unsigned long *chars, char *P, long *w)
{
    if (size <= 0 || T[0] != '>') {
        fprintf(stderr, "%s: poorly formatted FASTA file\n", who);
        return -1;
    }
    off_t i = line_end_at(T, size, 0) + 1;
    unsigned long seq = 1, c = 0, ctot = 0;
    long k = 0;
    for (; i < size; ++i) {
        const unsigned char ch = (unsigned char)T[i];
        if (ch == '>') {
            ++seq;
            lstats_add(st, (double)c);
            c = 0;
            if (P) P[k++] = '$';
            i = line_end_at(T, size, i);
        } else if (isalpha(ch)) {
            ++c;
            ++ctot;
            if (P) P[k++] = (char)ch;
        } else if (ch == '\n' || ch == '\r') {
            if (i + 1 < size && (T[i + 1] == '\n' || T[i + 1] == '\r')) ++i;
        } else if (isprint(ch)) {
            fprintf(stderr, "%s: non-alpha character ('%c')\n", who, ch);
            return -1;
        } else {
            fprintf(stderr, "%s: non-printing character ('%d')\n", who, T[i]);
            return -1;
        }
    }
    if (c == 0) {
        fprintf(stderr, "%s: empty sequence at end of input\n", who);
        return -1;
    }
    lstats_add(st, (double)c);
    if (P) P[k++] = '$';
    *seqs = seq;
    *chars = ctot;
    if (w) *w = k;
    return 0;
}

/* One pass over FASTQ text (four-line records), as fasta_scan. */
static int fastq_scan(const char *T, off_t size, const char *who, lstats_t *st,
                      unsigned long *seqs, char *P, long *w)
{
    off_t i = 0;
    unsigned long seq = 0, c = 0, line = 0;
    long k = 0;
    int first = 1;
    while (i < size) {
        if (T[i] != '@') {
            fprintf(stderr, "%s: poorly formatted FASTQ file, line %lu\n", who, line);
            return -1;
        }
        ++seq;
        if (first) first = 0;
        else {
            lstats_add(st, (double)c);
            if (P) P[k++] = '$';
        }
        c = 0;
        i = line_end_at(T, size, i) + 1;
        ++line;
        for (; i < size && T[i] != '\n' && T[i] != '\r'; ++i, ++c)
            if (P) P[k++] = T[i];
        if (i + 1 < size && (T[i + 1] == '\n' || T[i + 1] == '\r')) ++i;
        lstats_add(st, (double)c);
        ++i;
        ++line;
        if (i >= size || T[i] != '+') {
            fprintf(stderr, "%s: poorly formatted FASTQ file, line %lu\n", who, line);
            return -1;
        }
        i = line_end_at(T, size, i) + 1;
        ++line;
        i = line_end_at(T, size, i) + 1;
        ++line;
    }
    if (P) P[k++] = '$';
    *seqs = seq;
    if (w) *w = k;
    return 0;
}

PHOST_EXPORT parasail_file_stat_t *parasail_stat_fasta_buffer(const char *T, off_t size)
{
    if (!T) {
        fprintf(stderr, "parasail_stat_fasta_buffer: NULL pointer\n");
        return NULL;
    }
    lstats_t st = {0, 0, 0, 0, 0};
    unsigned long seqs = 0, chars = 0;
    if (fasta_scan(T, size, "parasail_stat_fasta_buffer", &st, &seqs, &chars, NULL, NULL))
        return NULL;
    return lstats_out(&st, seqs, chars);
}

PHOST_EXPORT parasail_file_stat_t *parasail_stat_fastq_buffer(const char *T, off_t size)
{
    if (!T) {
        fprintf(stderr, "parasail_stat_fastq_buffer: NULL pointer\n");
        return NULL;
    }
    lstats_t st = {0, 0, 0, 0, 0};
    unsigned long seqs = 0;
    if (fastq_scan(T, size, "parasail_stat_fastq_buffer", &st, &seqs, NULL, NULL)) return NULL;
    return lstats_out(&st, seqs, 0);
}

PHOST_EXPORT parasail_file_stat_t *parasail_stat_buffer(const char *buf, off_t size)
{
    if (!buf) {
        fprintf(stderr, "parasail_stat_buffer: NULL pointer\n");
        return NULL;
    }
    parasail_file_stat_t *s = NULL;
    if (parasail_is_fasta_buffer(buf, size) == 1) s = parasail_stat_fasta_buffer(buf, size);
    else if (parasail_is_fastq_buffer(buf, size) == 1) s = parasail_stat_fastq_buffer(buf, size);
    else {
        fprintf(stderr, "parasail_stat: cannot determine file format\n");
        return NULL;
    }
    if (!s) fprintf(stderr, "parasail_stat_buffer: failed\n");
    return s;
}

PHOST_EXPORT parasail_file_stat_t *parasail_stat_fasta(const parasail_file_t *pf)
{
    if (!pf) {
        fprintf(stderr, "parasail_stat_fasta: NULL pointer\n");
        return NULL;
    }
    return parasail_stat_fasta_buffer(pf->buf, pf->size);
}

PHOST_EXPORT parasail_file_stat_t *parasail_stat_fastq(const parasail_file_t *pf)
{
    if (!pf) {
        fprintf(stderr, "parasail_stat_fastq: NULL pointer\n");
        return NULL;
    }
    return parasail_stat_fastq_buffer(pf->buf, pf->size);
}

PHOST_EXPORT parasail_file_stat_t *parasail_stat(const parasail_file_t *pf)
{
    if (!pf) {
        fprintf(stderr, "parasail_stat: NULL pointer\n");
        return NULL;
    }
    parasail_file_stat_t *s = NULL;
    if (parasail_is_fasta(pf) == 1) s = parasail_stat_fasta(pf);
    else if (parasail_is_fastq(pf) == 1) s = parasail_stat_fastq(pf);
    else {
        fprintf(stderr, "parasail_stat: cannot determine file format\n");
        return NULL;
    }
    if (!s) fprintf(stderr, "parasail_stat: failed\n");
    return s;
}

PHOST_EXPORT char *parasail_read(const parasail_file_t *pf, long *size)
{
    if (!pf) {
        fprintf(stderr, "parasail_read: NULL pointer\n");
        return NULL;
    }
    if (!size) {
        fprintf(stderr, "parasail_read: NULL size pointer\n");
        return NULL;
    }
    char *b = (char *)malloc((size_t)pf->size + 1);
    if (!b) {
        perror("malloc");
        return NULL;
    }
    memcpy(b, pf->buf, (size_t)pf->size);
    b[pf->size] = 0;
    *size = (long)pf->size;
    return b;
}

PHOST_EXPORT char *parasail_pack_fasta_buffer(const char *T, off_t size, long *packed_size)
{
    if (!T || !packed_size) {
        fprintf(stderr, "parasail_pack_fasta_buffer: NULL pointer\n");
        return NULL;
    }
    lstats_t st = {0, 0, 0, 0, 0};
    unsigned long seqs = 0, chars = 0;
    if (fasta_scan(T, size, "parasail_pack_fasta_buffer", &st, &seqs, &chars, NULL, NULL))
        return NULL;
    char *P = (char *)malloc(chars + seqs + 1);
    if (!P) {
        perror("malloc");
        return NULL;
    }
    long w = 0;
    fasta_scan(T, size, "parasail_pack_fasta_buffer", &st, &seqs, &chars, P, &w);
    P[w] = 0;
    *packed_size = w;
    return P;
}

PHOST_EXPORT char *parasail_pack_fastq_buffer(const char *T, off_t size, long *packed_size)
{
    if (!T || !packed_size) {
        fprintf(stderr, "parasail_pack_fastq_buffer: NULL pointer\n");
        return NULL;
    }
    lstats_t st = {0, 0, 0, 0, 0};
    unsigned long seqs = 0;
    if (fastq_scan(T, size, "parasail_pack_fastq_buffer", &st, &seqs, NULL, NULL)) return NULL;
    /* every residue and separator fits in the text's own length */
    char *P = (char *)malloc((size_t)size + 2);
    if (!P) {
        perror("malloc");
        return NULL;
    }
    long w = 0;
    fastq_scan(T, size, "parasail_pack_fastq_buffer", &st, &seqs, P, &w);
    P[w] = 0;
    *packed_size = w;
    return P;
}

PHOST_EXPORT char *parasail_pack_buffer(const char *buf, off_t size, long *packed_size)
{
    if (!buf || !packed_size) {
        fprintf(stderr, "parasail_pack_buffer: NULL pointer\n");
        return NULL;
    }
    char *p = NULL;
    if (parasail_is_fasta_buffer(buf, size) == 1) p = parasail_pack_fasta_buffer(buf, size, packed_size);
    else if (parasail_is_fastq_buffer(buf, size) == 1)
        p = parasail_pack_fastq_buffer(buf, size, packed_size);
    else {
        fprintf(stderr, "parasail_pack: cannot determine file format\n");
        return NULL;
    }
    if (!p) fprintf(stderr, "parasail_pack_buffer: failed\n");
    return p;
}

PHOST_EXPORT char *parasail_pack_fasta(const parasail_file_t *pf, long *size)
{
    if (!pf || !size) {
        fprintf(stderr, "parasail_pack_fasta: NULL pointer\n");
        return NULL;
    }
    return parasail_pack_fasta_buffer(pf->buf, pf->size, size);
}

PHOST_EXPORT char *parasail_pack_fastq(const parasail_file_t *pf, long *size)
{
    if (!pf || !size) {
        fprintf(stderr, "parasail_pack_fastq: NULL pointer\n");
        return NULL;
    }
    return parasail_pack_fastq_buffer(pf->buf, pf->size, size);
}

PHOST_EXPORT char *parasail_pack(const parasail_file_t *pf, long *size)
{
    if (!pf || !size) {
        fprintf(stderr, "parasail_pack: NULL pointer\n");
        return NULL;
    }
    char *p = NULL;
    if (parasail_is_fasta(pf) == 1) p = parasail_pack_fasta(pf, size);
    else if (parasail_is_fastq(pf) == 1) p = parasail_pack_fastq(pf, size);
    else {
        fprintf(stderr, "parasail_pack: cannot determine file format\n");
        return NULL;
    }
    if (!p) fprintf(stderr, "parasail_pack: failed\n");
    return p;
}

/* ---------------------------------------------------------------- matrices */

/* the integer at T[*i] after blanks: up to 9 characters, an optional '-'
 * then digits (src/io.c:804-846); 1, or -1 if there is none */
static int read_int(const char *T, off_t size, off_t *i_, int *out)
{
    off_t i = *i_;
    while (i < size && (T[i] == ' ' || T[i] == '\t')) ++i;
    if (i >= size || !(isdigit((unsigned char)T[i]) || T[i] == '-')) return -1;
    char tok[10];
    int p = 0;
    tok[p++] = T[i++];
    while (p < 10 && i < size && isdigit((unsigned char)T[i])) tok[p++] = T[i++];
    if (p == 10) return -1;
    tok[p] = 0;
    if (sscanf(tok, "%d", out) != 1) return -1;
    *i_ = i;
    return 1;
}

/* the alphabet line starting at T[i]: letters and '*', blanks between */
static char *read_alphabet(const char *T, off_t size, off_t i)
{
    size_t n = 0;
    for (off_t k = i; k < size && T[k] != '\n' && T[k] != '\r' && T[k] != '#'; ++k) {
        if (T[k] == ' ' || T[k] == '\t') continue;
        if (!isalpha((unsigned char)T[k]) && T[k] != '*') return NULL;
        ++n;
    }
    if (n == 0) return NULL;
    char *a = (char *)malloc(n + 1);
    if (!a) return NULL;
    n = 0;
    for (off_t k = i; k < size && T[k] != '\n' && T[k] != '\r' && T[k] != '#'; ++k)
        if (T[k] != ' ' && T[k] != '\t') a[n++] = T[k];
    a[n] = 0;
    return a;
}

static parasail_matrix_t *matrix_from_file(const char *filename, int case_sensitive)
{
    const char *who = "parasail_matrix_from_file";
    if (!filename) {
        fprintf(stderr, "%s: NULL pointer\n", who);
        return NULL;
    }
    parasail_file_t *pf = parasail_open(filename);
    if (!pf) {
        fprintf(stderr, "%s: parasail_open(%s) failed\n", who, filename);
        return NULL;
    }
    const char *T = pf->buf;
    const off_t size = pf->size;
    char *alphabet = NULL, *query = NULL;
    int *vals = NULL;
    size_t asize = 0, nvals = 0, cap = 0, qn = 0, qcap = 0, rows = 0;
    int mx = INT_MIN, mn = INT_MAX;
    const char *err = NULL;
    for (off_t i = 0; i < size && !err; ++i) {
        const unsigned char ch = (unsigned char)T[i];
        if (ch == '#') {
            i = line_end_at(T, size, i);
        } else if (isalnum(ch) || ch == '*' || ch == '-') {
            if (!alphabet) {
                alphabet = read_alphabet(T, size, i);
                if (!alphabet) { err = "poorly formed matrix file alphabet"; break; }
                asize = strlen(alphabet);
                cap = asize * asize;
                qcap = asize;
                vals = (int *)malloc(sizeof(int) * cap);
                query = (char *)malloc(qcap + 1);
                if (!vals || !query) { err = "cannont malloc buffer for matrix"; break; }
                i = line_end_at(T, size, i);
                continue;
            }
            if (qn >= qcap) {
                char *nq = (char *)realloc(query, qcap * 2 + 1);
                if (!nq) { err = "couldn't grow query size"; break; }
                query = nq;
                qcap *= 2;
            }
            ++rows;
            /* a row letter, or none (a PSSM position) */
            if (isalpha(ch) || ch == '*') query[qn++] = T[i++];
            else query[qn++] = '*';
            for (size_t k = 0; k < asize; ++k) {
                int v = 0;
                if (read_int(T, size, &i, &v) != 1) { err = "poorly formed matrix file"; break; }
                if (nvals >= cap) {
                    int *nv = (int *)realloc(vals, sizeof(int) * cap * 2);
                    if (!nv) { err = "couldn't grow matrix size"; break; }
                    vals = nv;
                    cap *= 2;
                }
                vals[nvals++] = v;
                mx = v > mx ? v : mx;
                mn = v < mn ? v : mn;
            }
            /* the loop's ++i then passes over the character after the
             * row's last number, as src/io.c:1069 does */
        } else if (ch == '\n' || ch == '\r') {
            if (i + 1 < size && (T[i + 1] == '\n' || T[i + 1] == '\r')) ++i;
        } else if (ch == ' ' || ch == '\t') {
        } else if (isprint(ch)) {
            fprintf(stderr, "%s: non-alpha character in matrix file ('%c')\n", who, ch);
            err = "";
        } else {
            fprintf(stderr, "%s: non-printing character in matrix file ('%d')\n", who, T[i]);
            err = "";
        }
    }
    parasail_close(pf);
    int type = PARASAIL_MATRIX_TYPE_SQUARE;
    if (!err && !alphabet) err = "poorly formed matrix file alphabet";
    if (!err) {
        query[qn] = 0;
        if (asize == qn && strcmp(alphabet, query) == 0) {
            if (nvals != asize * asize) err = "matrix is missing values";
            else if (rows != asize) err = "matrix is missing rows";
        } else {
            type = PARASAIL_MATRIX_TYPE_PSSM;
            if (nvals != asize * rows) err = "matrix is missing values";
        }
    }
    int *mapper = err ? NULL : (int *)malloc(sizeof(int) * 256);
    parasail_matrix_t *m = mapper ? (parasail_matrix_t *)malloc(sizeof(*m)) : NULL;
    if (err || !m) {
        if (err && *err) fprintf(stderr, "%s: %s\n", who, err);
        free(alphabet); free(query); free(vals); free(mapper);
        return NULL;
    }
    for (int c = 0; c < 256; ++c) mapper[c] = (int)asize - 1;
    for (size_t c = 0; c < asize; ++c) {
        const unsigned char a = (unsigned char)alphabet[c];
        if (case_sensitive) {
            mapper[a] = (int)c;
        } else {
            mapper[toupper(a)] = (int)c;
            mapper[tolower(a)] = (int)c;
        }
    }
    m->name = filename;
    m->matrix = vals;
    m->mapper = mapper;
    m->size = (int)asize;
    m->max = mx;
    m->min = mn;
    m->user_matrix = vals;
    m->type = type;
    m->length = (int)rows;
    m->alphabet = alphabet;
    m->query = query;
    return m;
}

PHOST_EXPORT parasail_matrix_t *parasail_matrix_from_file(const char *filename)
{
    return matrix_from_file(filename, 0);
}

PHOST_EXPORT parasail_matrix_t *parasail_matrix_from_file_case_sensitive(const char *filename)
{
    return matrix_from_file(filename, 1);
}
