/* MPS / QPS reader producing an sk_model (C11): the input side of the
 * solver for real instances.
 *
 * Follows the reference reader's accepted language and conventions:
 *   c/src/sankhya_mps.c:150-331  (sections, free format first and the
 *                                 column-addressed fixed format as fallback,
 *                                 RANGES / BOUNDS semantics, OBJSENSE MAX)
 *   c/src/sk_core.c:214-216      (|bound| >= 1e29 or infinite -> 1e30)
 *   c/src/sk_core.c:274-378      (QUADOBJ / QMATRIX / QSECTION second pass,
 *                                 every off-diagonal mirrored, full symmetric Q)
 * Names are looked up through an open-addressing hash table instead of
 * linear scans, which keeps large models (Kennington, Netlib) linear time;
 * the first definition of a name wins where the reference's first match
 * does. */
#define _POSIX_C_SOURCE 200809L
#include "skg_internal.h"

#include <ctype.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

/* ------------------------------------------------------------ name table */

typedef struct names {
    char **s;          /* names in order of first appearance */
    int count, cap;
    int *slot;         /* hash slots: index + 1, 0 empty     */
    int nslot;         /* power of two                       */
} names;

static uint64_t hash_str(const char *s)
{
    uint64_t h = 1469598103934665603ULL;   /* FNV-1a */
    while (*s) { h ^= (unsigned char)*s++; h *= 1099511628211ULL; }
    return h;
}

static void names_free(names *t)
{
    int k;
    for (k = 0; k < t->count; ++k) free(t->s[k]);
    free(t->s); free(t->slot);
    memset(t, 0, sizeof(*t));
}

static int names_find(const names *t, const char *s)
{
    uint64_t h;
    if (!t->nslot) return -1;
    for (h = hash_str(s) & (uint64_t)(t->nslot - 1);; h = (h + 1) & (uint64_t)(t->nslot - 1)) {
        const int v = t->slot[h];
        if (!v) return -1;
        if (strcmp(t->s[v - 1], s) == 0) return v - 1;
    }
}

static int names_rehash(names *t, int nslot)
{
    int *slot = (int *)calloc((size_t)nslot, sizeof(int)), k;
    if (!slot) return -1;
    for (k = 0; k < t->count; ++k) {
        uint64_t h = hash_str(t->s[k]) & (uint64_t)(nslot - 1);
        while (slot[h]) h = (h + 1) & (uint64_t)(nslot - 1);
        slot[h] = k + 1;
    }
    free(t->slot);
    t->slot = slot;
    t->nslot = nslot;
    return 0;
}

/* Index of s, added at the end when new; -1 when out of memory. */
static int names_add(names *t, const char *s, int *added)
{
    int k = names_find(t, s);
    uint64_t h;
    if (added) *added = 0;
    if (k >= 0) return k;
    if (2 * (t->count + 1) > t->nslot && names_rehash(t, t->nslot ? 2 * t->nslot : 64)) return -1;
    if (t->count == t->cap) {
        const int cap = t->cap ? 2 * t->cap : 64;
        char **g = (char **)realloc(t->s, sizeof(char *) * (size_t)cap);
        if (!g) return -1;
        t->s = g; t->cap = cap;
    }
    if (!(t->s[t->count] = strdup(s))) return -1;
    h = hash_str(s) & (uint64_t)(t->nslot - 1);
    while (t->slot[h]) h = (h + 1) & (uint64_t)(t->nslot - 1);
    t->slot[h] = t->count + 1;
    if (added) *added = 1;
    return t->count++;
}

/* ------------------------------------------------------------ raw model */

typedef struct triple { int row, col; double v; } triple;
typedef struct named_value { int row; double v; } named_value;   /* RHS / RANGES by row */
typedef struct bound_rec { char type[4]; int col; double v; } bound_rec;

typedef struct raw {
    names rows, cols, free_rows, rhs_seen, range_seen;   /* rows: first definition */
    char **rowname; char *rowtype; int nrow, rowcap;     /* every definition       */
    double *obj;    unsigned char *integer; int col_cap;
    triple *e;      long long ne, ecap;
    named_value *rhs;   int nrhs, rhs_cap;
    named_value *rng;   int nrng, rng_cap;
    bound_rec *bnd;     int nbnd, bnd_cap;
    char *objective;    /* name of the first N row */
    double objshift_rhs;
    int has_obj_rhs, maximize;
} raw;

static void raw_free(raw *r)
{
    names_free(&r->rows); names_free(&r->cols); names_free(&r->free_rows);
    names_free(&r->rhs_seen); names_free(&r->range_seen);
    if (r->rowname) { int k; for (k = 0; k < r->nrow; ++k) free(r->rowname[k]); }
    free(r->rowname); free(r->rowtype); free(r->obj); free(r->integer); free(r->e);
    free(r->rhs); free(r->rng); free(r->bnd); free(r->objective);
    memset(r, 0, sizeof(*r));
}

static int ensure_col(raw *r, int col)
{
    if (col < r->col_cap) return 0;
    {
        int cap = r->col_cap ? r->col_cap : 64;
        double *o;
        unsigned char *t;
        while (cap <= col) cap *= 2;
        o = (double *)realloc(r->obj, sizeof(double) * (size_t)cap);
        if (!o) return -1;
        r->obj = o;
        t = (unsigned char *)realloc(r->integer, (size_t)cap);
        if (!t) return -1;
        r->integer = t;
        memset(r->obj + r->col_cap, 0, sizeof(double) * (size_t)(cap - r->col_cap));
        memset(r->integer + r->col_cap, 0, (size_t)(cap - r->col_cap));
        r->col_cap = cap;
    }
    return 0;
}

/* ------------------------------------------------------------ tokenizer */

enum { MAXTOK = 8, TOKLEN = 64 };

static void strip_quotes(char *t)
{
    const size_t n = strlen(t);
    if (n >= 2 && ((t[0] == '\'' && t[n - 1] == '\'') || (t[0] == '"' && t[n - 1] == '"'))) {
        memmove(t, t + 1, n - 2);
        t[n - 2] = '\0';
    }
}

/* Whitespace tokens (at most MAXTOK), quotes stripped. */
static int tokens_free(char *line, char **tok)
{
    int n = 0;
    char *save = NULL, *t = strtok_r(line, " \t\r\n", &save);
    while (t && n < MAXTOK) { strip_quotes(t); tok[n++] = t; t = strtok_r(NULL, " \t\r\n", &save); }
    return n;
}

/* Fixed-format data record: the six standard fields (columns 2-3, 5-12,
 * 15-22, 25-36, 40-47, 50-61), each trimmed; names may hold blanks; empty
 * fields are skipped so the token vector matches the free-format one. */
static int tokens_fixed(const char *line, char buf[MAXTOK][TOKLEN], char **tok)
{
    static const int lo[6] = {2, 5, 15, 25, 40, 50}, hi[6] = {3, 12, 22, 36, 47, 61};
    int len = (int)strlen(line), f, n = 0;
    while (len > 0 && (line[len - 1] == '\n' || line[len - 1] == '\r')) --len;
    for (f = 0; f < 6; ++f) {
        int a = lo[f] - 1, b = hi[f] < len ? hi[f] : len, k = 0;
        if (lo[f] > len) continue;
        while (a < b && (line[a] == ' ' || line[a] == '\t')) ++a;
        while (b > a && (line[b - 1] == ' ' || line[b - 1] == '\t')) --b;
        if (a >= b) continue;
        while (a < b && k + 1 < TOKLEN) buf[n][k++] = line[a++];
        buf[n][k] = '\0';
        tok[n] = buf[n];
        ++n;
    }
    return n;
}

static int is_kw(const char *t, const char *kw) { return strcasecmp(t, kw) == 0; }

/* ------------------------------------------------------------ LP pass */

enum { S_NONE, S_ROWS, S_COLUMNS, S_RHS, S_RANGES, S_BOUNDS, S_QUAD };

/* Reads the whole line (any length) into *buf; 0 at end of file. */
static int read_line(FILE *f, char **buf, size_t *cap)
{
    return getline(buf, cap, f) > 0;
}

static int pairs_add(names *seen, named_value **arr, int *count, int *cap, raw *r, const char *row, double v)
{
    int added = 0, k;
    if (names_add(seen, row, &added) < 0) return -1;
    if (!added) return 0;                       /* first definition wins */
    k = names_find(&r->rows, row);
    if (*count == *cap) {
        const int nc = *cap ? 2 * *cap : 64;
        named_value *g = (named_value *)realloc(*arr, sizeof(named_value) * (size_t)nc);
        if (!g) return -1;
        *arr = g; *cap = nc;
    }
    (*arr)[*count].row = k;                     /* -1: no such row (ignored) */
    (*arr)[*count].v = v;
    ++*count;
    if (r->objective && strcmp(row, r->objective) == 0 && arr == &r->rhs) {
        r->objshift_rhs = -v;
        r->has_obj_rhs = 1;
    }
    return 0;
}

static int parse_lp(const char *path, int fixed, raw *r)
{
    FILE *f = fopen(path, "r");
    char *line = NULL, fbuf[MAXTOK][TOKLEN], *tok[MAXTOK];
    size_t cap = 0;
    int section = S_NONE, integer_mode = 0, rc = -1;
    long long ecap = 0;
    memset(r, 0, sizeof(*r));
    if (!f) return -1;
    while (read_line(f, &line, &cap)) {
        int nt, header;
        if (line[0] == '*' || line[0] == '\n' || line[0] == '\r' || line[0] == '\0') continue;
        /* a header starts in column 1, a data record is indented */
        header = line[0] != ' ' && line[0] != '\t';
        nt = (fixed && !header) ? tokens_fixed(line, fbuf, tok) : tokens_free(line, tok);
        if (nt == 0) continue;
        if (header && is_kw(tok[0], "NAME")) continue;
        if (header && is_kw(tok[0], "OBJSENSE")) { section = S_NONE; continue; }
        if (section == S_NONE && is_kw(tok[0], "MAX")) { r->maximize = 1; continue; }
        if (section == S_NONE && is_kw(tok[0], "MIN")) { r->maximize = 0; continue; }
        if (header && is_kw(tok[0], "ROWS")) { section = S_ROWS; continue; }
        if (header && is_kw(tok[0], "COLUMNS")) { section = S_COLUMNS; continue; }
        if (header && is_kw(tok[0], "RHS")) { section = S_RHS; continue; }
        if (header && is_kw(tok[0], "RANGES")) { section = S_RANGES; continue; }
        if (header && is_kw(tok[0], "BOUNDS")) { section = S_BOUNDS; continue; }
        if (is_kw(tok[0], "QUADOBJ") || is_kw(tok[0], "QMATRIX") || is_kw(tok[0], "QSECTION")) {
            section = S_QUAD;
            continue;
        }
        if (header && is_kw(tok[0], "ENDATA")) break;
        if (section == S_ROWS) {
            const char t = tok[0][0];
            int added = 0;
            if (nt < 2 || (t != 'N' && t != 'L' && t != 'G' && t != 'E')) goto done;
            if (t == 'N') {
                /* the first N row is the objective, later ones are free rows */
                if (!r->objective) { if (!(r->objective = strdup(tok[1]))) goto done; }
                else if (names_add(&r->free_rows, tok[1], NULL) < 0) goto done;
                continue;
            }
            /* a name defined twice makes a second row that no record can
               address (lookups find the first definition) */
            if (names_add(&r->rows, tok[1], &added) < 0) goto done;
            if (r->nrow == r->rowcap) {
                const int nc = r->rowcap ? 2 * r->rowcap : 64;
                char *g = (char *)realloc(r->rowtype, (size_t)nc);
                char **h;
                if (!g) goto done;
                r->rowtype = g;
                if (!(h = (char **)realloc(r->rowname, sizeof(char *) * (size_t)nc))) goto done;
                r->rowname = h; r->rowcap = nc;
            }
            if (!(r->rowname[r->nrow] = strdup(tok[1]))) goto done;
            r->rowtype[r->nrow++] = t;
        } else if (section == S_COLUMNS) {
            int col, p;
            if (nt >= 3 && is_kw(tok[1], "MARKER")) {
                if (is_kw(tok[2], "INTORG")) integer_mode = 1;
                if (is_kw(tok[2], "INTEND")) integer_mode = 0;
                continue;
            }
            if (nt < 3) goto done;
            if ((col = names_add(&r->cols, tok[0], NULL)) < 0 || ensure_col(r, col)) goto done;
            if (integer_mode) r->integer[col] = 1;
            for (p = 1; p + 1 < nt; p += 2) {
                const double v = strtod(tok[p + 1], NULL);
                int row;
                if (!isfinite(v)) goto done;
                if (r->objective && strcmp(r->objective, tok[p]) == 0) { r->obj[col] += v; continue; }
                row = names_find(&r->rows, tok[p]);
                if (row < 0) {
                    if (names_find(&r->free_rows, tok[p]) >= 0) continue;   /* free row */
                    goto done;
                }
                if (r->ne == ecap) {
                    const long long nc = ecap ? 2 * ecap : 1024;
                    triple *g = (triple *)realloc(r->e, sizeof(triple) * (size_t)nc);
                    if (!g) goto done;
                    r->e = g; ecap = nc;
                }
                r->e[r->ne].row = row; r->e[r->ne].col = col; r->e[r->ne].v = v;
                ++r->ne;
            }
        } else if (section == S_RHS || section == S_RANGES) {
            /* "[set] row value [row value]": an odd count carries the set name */
            int p;
            for (p = nt % 2; p + 1 < nt; p += 2) {
                const double v = strtod(tok[p + 1], NULL);
                const int bad = section == S_RHS ? pairs_add(&r->rhs_seen, &r->rhs, &r->nrhs, &r->rhs_cap, r, tok[p], v)
                                                 : pairs_add(&r->range_seen, &r->rng, &r->nrng, &r->rng_cap, r, tok[p], v);
                if (bad) goto done;
            }
        } else if (section == S_BOUNDS) {
            /* "TYPE [set] COLUMN [VALUE]"; FR / MI / PL / BV take no value */
            int wants, colpos, col;
            if (nt < 2) goto done;
            wants = !(is_kw(tok[0], "FR") || is_kw(tok[0], "MI") || is_kw(tok[0], "PL") || is_kw(tok[0], "BV"));
            if (wants && nt < 3) goto done;
            colpos = wants ? nt - 2 : nt - 1;
            if ((col = names_add(&r->cols, tok[colpos], NULL)) < 0 || ensure_col(r, col)) goto done;
            if (r->nbnd == r->bnd_cap) {
                const int nc = r->bnd_cap ? 2 * r->bnd_cap : 64;
                bound_rec *g = (bound_rec *)realloc(r->bnd, sizeof(bound_rec) * (size_t)nc);
                if (!g) goto done;
                r->bnd = g; r->bnd_cap = nc;
            }
            strncpy(r->bnd[r->nbnd].type, tok[0], 3);
            r->bnd[r->nbnd].type[3] = '\0';
            if (strlen(tok[0]) > 3) r->bnd[r->nbnd].type[0] = '\0';   /* no such type */
            r->bnd[r->nbnd].col = col;
            r->bnd[r->nbnd].v = wants ? strtod(tok[nt - 1], NULL) : 0.0;
            ++r->nbnd;
        }
    }
    rc = (r->nrow > 0 && r->cols.count > 0 && r->objective) ? 0 : -1;
done:
    free(line);
    fclose(f);
    if (rc) raw_free(r);
    return rc;
}

static int cmp_col_row(const void *a, const void *b)
{
    const triple *x = (const triple *)a, *y = (const triple *)b;
    if (x->col != y->col) return x->col < y->col ? -1 : 1;
    if (x->row != y->row) return x->row < y->row ? -1 : 1;
    return 0;
}

/* sk_core.c:214-216: infinite or |v| >= 1e29 becomes the 1e30 sentinel */
static double sentinel_lo(double v) { return (v <= -1e29 || isinf(v)) ? -SK_INFINITY : v; }
static double sentinel_hi(double v) { return (v >= 1e29 || isinf(v)) ? SK_INFINITY : v; }

/* Sorted (col, row) triples, duplicates summed, zero sums dropped, into a
 * CSC of ncol columns; returns nnz or -1. */
static long long triples_to_csc(triple *e, long long ne, int ncol, int **p, int **i, double **x)
{
    long long k = 0, nz = 0;
    qsort(e, (size_t)ne, sizeof(triple), cmp_col_row);
    *p = (int *)calloc((size_t)ncol + 1, sizeof(int));
    *i = (int *)malloc(sizeof(int) * (size_t)(ne > 0 ? ne : 1));
    *x = (double *)malloc(sizeof(double) * (size_t)(ne > 0 ? ne : 1));
    if (!*p || !*i || !*x) { free(*p); free(*i); free(*x); *p = NULL; *i = NULL; *x = NULL; return -1; }
    while (k < ne) {
        const long long b = k;
        double s = 0.0;
        while (k < ne && e[k].col == e[b].col && e[k].row == e[b].row) s += e[k++].v;
        if (s != 0.0) { (*i)[nz] = e[b].row; (*x)[nz] = s; (*p)[e[b].col + 1]++; ++nz; }
    }
    for (k = 0; k < ncol; ++k) (*p)[k + 1] += (*p)[k];
    return nz;
}

static sk_status build_model(raw *r, sk_model *m)
{
    const int nrow = r->nrow, ncol = r->cols.count;
    int *p = NULL, *ix = NULL, i, j;
    double *x = NULL, *rl, *ru, *cl, *cu;
    long long nz = triples_to_csc(r->e, r->ne, ncol, &p, &ix, &x);
    sk_status st;
    if (nz < 0 || nz > 2147483647LL) { free(p); free(ix); free(x); return SK_ERR_MEMORY; }
    if ((st = sk_gpu_model_alloc(m, nrow, ncol, (int)nz)) != SK_OK) { free(p); free(ix); free(x); return st; }
    memcpy(m->A.p, p, sizeof(int) * ((size_t)ncol + 1));
    memcpy(m->A.i, ix, sizeof(int) * (size_t)nz);
    memcpy(m->A.x, x, sizeof(double) * (size_t)nz);
    free(p); free(ix); free(x);
    rl = (double *)malloc(sizeof(double) * ((size_t)nrow + 1));
    ru = (double *)malloc(sizeof(double) * ((size_t)nrow + 1));
    cl = (double *)calloc((size_t)ncol + 1, sizeof(double));
    cu = (double *)malloc(sizeof(double) * ((size_t)ncol + 1));
    if (!rl || !ru || !cl || !cu) { free(rl); free(ru); free(cl); free(cu); sk_gpu_model_free(m); return SK_ERR_MEMORY; }
    for (i = 0; i < nrow; ++i) {
        const char t = r->rowtype[i];
        rl[i] = t == 'L' ? -INFINITY : 0.0;
        ru[i] = t == 'G' ? INFINITY : 0.0;
    }
    for (j = 0; j < ncol; ++j) cu[j] = INFINITY;
    for (i = 0; i < r->nrhs; ++i) {
        const int row = r->rhs[i].row;
        if (row < 0) continue;
        if (r->rowtype[row] == 'L') ru[row] = r->rhs[i].v;
        else if (r->rowtype[row] == 'G') rl[row] = r->rhs[i].v;
        else rl[row] = ru[row] = r->rhs[i].v;
    }
    for (i = 0; i < r->nrng; ++i) {
        const int row = r->rng[i].row;
        const double w = r->rng[i].v;
        double base;
        if (row < 0) continue;
        base = isfinite(rl[row]) ? rl[row] : ru[row];
        if (r->rowtype[row] == 'L') rl[row] = base - fabs(w);
        else if (r->rowtype[row] == 'G') ru[row] = base + fabs(w);
        else if (w >= 0.0) ru[row] = base + fabs(w);
        else rl[row] = base - fabs(w);
    }
    for (j = 0; j < ncol; ++j) { m->c[j] = r->obj[j]; m->vartype[j] = r->integer[j] ? SK_INTEGER : SK_CONTINUOUS; }
    for (i = 0; i < r->nbnd; ++i) {
        const bound_rec *b = &r->bnd[i];
        const int c = b->col;
        /* applied in file order; type names are matched exactly */
        if (!strcmp(b->type, "LO")) cl[c] = b->v;
        else if (!strcmp(b->type, "UP")) cu[c] = b->v;
        else if (!strcmp(b->type, "FX")) cl[c] = cu[c] = b->v;
        else if (!strcmp(b->type, "FR") || !strcmp(b->type, "MI")) cl[c] = -INFINITY;
        else if (!strcmp(b->type, "BV")) { cl[c] = 0.0; cu[c] = 1.0; m->vartype[c] = SK_INTEGER; }
        else if (!strcmp(b->type, "LI")) { cl[c] = b->v; m->vartype[c] = SK_INTEGER; }
        else if (!strcmp(b->type, "UI")) { cu[c] = b->v; m->vartype[c] = SK_INTEGER; }
    }
    m->objshift = r->has_obj_rhs ? r->objshift_rhs : 0.0;
    if (r->maximize) {
        for (j = 0; j < ncol; ++j) m->c[j] = -m->c[j];
        m->objshift = -m->objshift;
    }
    for (i = 0; i < nrow; ++i) { m->rlow[i] = sentinel_lo(rl[i]); m->rupp[i] = sentinel_hi(ru[i]); }
    for (j = 0; j < ncol; ++j) { m->clow[j] = sentinel_lo(cl[j]); m->cupp[j] = sentinel_hi(cu[j]); }
    free(rl); free(ru); free(cl); free(cu);
    m->objsense = 1.0;      /* normalised to minimisation above */
    m->rowname = (char **)calloc((size_t)nrow + 1, sizeof(char *));
    m->colname = (char **)calloc((size_t)ncol + 1, sizeof(char *));
    if (!m->rowname || !m->colname) { sk_gpu_model_free(m); return SK_ERR_MEMORY; }
    /* hand the name strings over to the model */
    for (i = 0; i < nrow; ++i) { m->rowname[i] = r->rowname[i]; r->rowname[i] = NULL; }
    for (j = 0; j < ncol; ++j) { m->colname[j] = r->cols.s[j]; r->cols.s[j] = NULL; }
    return SK_OK;
}

/* ------------------------------------------------------------ QPS pass */

static sk_status parse_quadratic(const char *path, const names *cols, sk_model *m)
{
    FILE *f = fopen(path, "r");
    char *line = NULL;
    size_t cap = 0;
    triple *e = NULL;
    long long ne = 0, ecap = 0;
    int in_q = 0;
    sk_status st = SK_OK;
    if (!f) return SK_ERR_IO;
    while (read_line(f, &line, &cap)) {
        char *tok[4], *save = NULL, *t, *end = NULL;
        int nt = 0, a, b;
        double v;
        if (line[0] == '*' || line[0] == '\n' || line[0] == '\r' || line[0] == '\0') continue;
        for (t = strtok_r(line, " \t\r\n", &save); t && nt < 4; t = strtok_r(NULL, " \t\r\n", &save)) tok[nt++] = t;
        if (!nt) continue;
        if (is_kw(tok[0], "QUADOBJ") || is_kw(tok[0], "QMATRIX") || is_kw(tok[0], "QSECTION")) { in_q = 1; continue; }
        if (is_kw(tok[0], "ENDATA")) break;
        if (is_kw(tok[0], "ROWS") || is_kw(tok[0], "COLUMNS") || is_kw(tok[0], "RHS") || is_kw(tok[0], "RANGES") ||
            is_kw(tok[0], "BOUNDS")) { in_q = 0; continue; }
        if (!in_q) continue;
        if (nt != 3) { st = SK_ERR_STRUCTURE; break; }
        a = names_find(cols, tok[0]);
        b = names_find(cols, tok[1]);
        v = strtod(tok[2], &end);
        if (a < 0 || b < 0 || !end || *end != '\0' || !isfinite(v)) { st = SK_ERR_STRUCTURE; break; }
        /* QPS holds 0.5 x'Qx; every off-diagonal entry is mirrored */
        if (ne + 2 > ecap) {
            const long long nc = ecap ? 2 * ecap : 256;
            triple *g = (triple *)realloc(e, sizeof(triple) * (size_t)nc);
            if (!g) { st = SK_ERR_MEMORY; break; }
            e = g; ecap = nc;
        }
        e[ne].row = b; e[ne].col = a; e[ne].v = v; ++ne;
        if (a != b) { e[ne].row = a; e[ne].col = b; e[ne].v = v; ++ne; }
    }
    free(line);
    fclose(f);
    if (st == SK_OK && ne > 0) {
        sk_csc *q = (sk_csc *)calloc(1, sizeof(sk_csc));
        long long nz;
        if (!q) st = SK_ERR_MEMORY;
        else if ((nz = triples_to_csc(e, ne, m->ncol, &q->p, &q->i, &q->x)) < 0) { free(q); st = SK_ERR_MEMORY; }
        else {
            q->nrow = q->ncol = m->ncol;
            q->nzmax = (int)nz;
            m->Q = q;
        }
    }
    free(e);
    return st;
}

/* c/src/sk_core.c:318-378 sk_read_mps */
sk_status sk_gpu_read_mps(const char *path, sk_model *m)
{
    raw r;
    sk_status st;
    skg_clear_error();
    if (!path || !m) return SK_ERR_ARG;
    /* free format reads almost every file; strict fixed-format names with
       embedded blanks need the column-addressed fallback */
    if (parse_lp(path, 0, &r) && parse_lp(path, 1, &r)) {
        skg_fail("cannot read MPS file %s", path);
        return SK_ERR_IO;
    }
    st = build_model(&r, m);
    if (st == SK_OK) {
        names cols = r.cols;
        /* the column names now live in the model: look them up there */
        int j;
        for (j = 0; j < m->ncol; ++j) cols.s[j] = m->colname[j];
        st = parse_quadratic(path, &cols, m);
        for (j = 0; j < m->ncol; ++j) cols.s[j] = NULL;
        if (st != SK_OK) { sk_gpu_model_free(m); skg_fail("invalid quadratic section in %s", path); }
    } else {
        skg_fail("out of memory reading %s", path);
    }
    raw_free(&r);
    return st;
}
