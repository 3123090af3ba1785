/*
 * lookup.c -- name -> entry point tables for the hip family, with the
 * reference's lookup semantics (src/function_lookup.c:16-129): exact match
 * first, then the name with a "parasail_" prefix, NULL if absent.  The
 * metadata follows parasail/function_lookup.h:17-15680 (alg, type, isa,
 * bits, width, lanes, is_table, is_rowcol, is_trace, is_stats, is_ref).
 */
#include <stdlib.h>
#include <string.h>

#include "internal.h"

#define ALGNAME_(alg, kind) #alg #kind
#define STATS_ONLY_ 0
#define STATS_ONLY__stats 1
#define STATS_ONLY__trace 0
#define TRACE_ONLY_ 0
#define TRACE_ONLY__stats 0
#define TRACE_ONLY__trace 1
/* the reference labels stats entries "sw_stats" and trace entries "sw" */
#define ALGLABEL_(alg) #alg
#define ALGLABEL__stats(alg) #alg "_stats"
#define ALGLABEL__trace(alg) #alg

#define F_ENTRY(alg, kind, strat, w)                                                       \
    {parasail_##alg##kind##_##strat##_hip_64_##w,                                          \
     "parasail_" #alg #kind "_" #strat "_hip_64_" #w,                                      \
     ALGLABEL_##kind(alg), #strat, "hip", "64", #w, 64, 0, 0, TRACE_ONLY_##kind,            \
     STATS_ONLY_##kind, 0},

#define P_ENTRY(alg, kind, strat, w, creator)                                              \
    {parasail_##alg##kind##_##strat##_profile_hip_64_##w, creator,                          \
     "parasail_" #alg #kind "_" #strat "_profile_hip_64_" #w,                              \
     ALGLABEL_##kind(alg), #strat "_profile", "hip", "64", #w, 64, 0, 0, TRACE_ONLY_##kind, \
     STATS_ONLY_##kind, 0},

#define CREATOR_(w) parasail_profile_create_hip_64_##w
#define CREATOR__stats(w) parasail_profile_create_stats_hip_64_##w
#define CREATOR__trace(w) parasail_profile_create_hip_64_##w

#define F_S(alg, kind, strat) \
    F_ENTRY(alg, kind, strat, 8) F_ENTRY(alg, kind, strat, 16) F_ENTRY(alg, kind, strat, 32) F_ENTRY(alg, kind, strat, sat)
#define P_S(alg, kind, strat)                                                              \
    P_ENTRY(alg, kind, strat, 8, CREATOR_##kind(8))                                        \
    P_ENTRY(alg, kind, strat, 16, CREATOR_##kind(16))                                      \
    P_ENTRY(alg, kind, strat, 32, CREATOR_##kind(32))                                      \
    P_ENTRY(alg, kind, strat, sat, CREATOR_##kind(sat))

#define F_ALG(alg) F_S(alg, , striped) F_S(alg, , scan) F_S(alg, _stats, striped) \
    F_S(alg, _stats, scan) F_S(alg, _trace, striped) F_S(alg, _trace, scan)
#define P_ALG(alg) P_S(alg, , striped) P_S(alg, , scan) P_S(alg, _stats, striped) \
    P_S(alg, _stats, scan) P_S(alg, _trace, striped) P_S(alg, _trace, scan)

/* _table / _rowcol entries: kind "" or _stats, tab table (is_table) or
 * rowcol (is_rowcol) */
#define IS_TABLE_table 1
#define IS_TABLE_rowcol 0
#define IS_ROWCOL_table 0
#define IS_ROWCOL_rowcol 1
#define FT_ENTRY(alg, kind, tab, strat, w)                                                  \
    {parasail_##alg##kind##_##tab##_##strat##_hip_64_##w,                                   \
     "parasail_" #alg #kind "_" #tab "_" #strat "_hip_64_" #w,                              \
     ALGLABEL_##kind(alg), #strat, "hip", "64", #w, 64, IS_TABLE_##tab, IS_ROWCOL_##tab, 0,  \
     STATS_ONLY_##kind, 0},
#define PT_ENTRY(alg, kind, tab, strat, w, creator)                                         \
    {parasail_##alg##kind##_##tab##_##strat##_profile_hip_64_##w, creator,                   \
     "parasail_" #alg #kind "_" #tab "_" #strat "_profile_hip_64_" #w,                      \
     ALGLABEL_##kind(alg), #strat "_profile", "hip", "64", #w, 64, IS_TABLE_##tab,           \
     IS_ROWCOL_##tab, 0, STATS_ONLY_##kind, 0},
#define FT_S(alg, kind, tab, strat)                                                        \
    FT_ENTRY(alg, kind, tab, strat, 8) FT_ENTRY(alg, kind, tab, strat, 16)                  \
    FT_ENTRY(alg, kind, tab, strat, 32) FT_ENTRY(alg, kind, tab, strat, sat)
#define PT_S(alg, kind, tab, strat)                                                        \
    PT_ENTRY(alg, kind, tab, strat, 8, CREATOR_##kind(8))                                  \
    PT_ENTRY(alg, kind, tab, strat, 16, CREATOR_##kind(16))                                \
    PT_ENTRY(alg, kind, tab, strat, 32, CREATOR_##kind(32))                                \
    PT_ENTRY(alg, kind, tab, strat, sat, CREATOR_##kind(sat))
#define FT_ALG(alg) FT_S(alg, , table, striped) FT_S(alg, , table, scan)                    \
    FT_S(alg, , rowcol, striped) FT_S(alg, , rowcol, scan) FT_S(alg, _stats, table, striped) \
    FT_S(alg, _stats, table, scan) FT_S(alg, _stats, rowcol, striped)                      \
    FT_S(alg, _stats, rowcol, scan)
#define PT_ALG(alg) PT_S(alg, , table, striped) PT_S(alg, , table, scan)                    \
    PT_S(alg, , rowcol, striped) PT_S(alg, , rowcol, scan) PT_S(alg, _stats, table, striped) \
    PT_S(alg, _stats, table, scan) PT_S(alg, _stats, rowcol, striped)                      \
    PT_S(alg, _stats, rowcol, scan)

static const parasail_function_info_t functions[] = {
    PARASAIL_HIP_ALGS(F_ALG)
    PARASAIL_HIP_ALGS(FT_ALG)
    {NULL, NULL, NULL, NULL, NULL, NULL, NULL, 0, 0, 0, 0, 0, 0},
};

static const parasail_pfunction_info_t pfunctions[] = {
    PARASAIL_HIP_ALGS(P_ALG)
    PARASAIL_HIP_ALGS(PT_ALG)
    {NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL, 0, 0, 0, 0, 0, 0},
};

static char *with_prefix(const char *name)
{
    static const char prefix[] = "parasail_";
    char *s = (char *)malloc(strlen(prefix) + strlen(name) + 1);
    if (!s) return NULL;
    strcpy(s, prefix);
    strcat(s, name);
    return s;
}

const parasail_function_info_t *parasail_lookup_function_info(const char *funcname)
{
    if (!funcname) return NULL;
    for (int pass = 0; pass < 2; ++pass) {
        char *name = pass ? with_prefix(funcname) : (char *)funcname;
        if (!name) return NULL;
        const parasail_function_info_t *hit = NULL;
        for (const parasail_function_info_t *f = functions; f->pointer; ++f)
            if (strcmp(name, f->name) == 0) { hit = f; break; }
        if (pass) free(name);
        if (hit) return hit;
    }
    return NULL;
}

const parasail_pfunction_info_t *parasail_lookup_pfunction_info(const char *funcname)
{
    if (!funcname) return NULL;
    for (int pass = 0; pass < 2; ++pass) {
        char *name = pass ? with_prefix(funcname) : (char *)funcname;
        if (!name) return NULL;
        const parasail_pfunction_info_t *hit = NULL;
        for (const parasail_pfunction_info_t *f = pfunctions; f->pointer; ++f)
            if (strcmp(name, f->name) == 0) { hit = f; break; }
        if (pass) free(name);
        if (hit) return hit;
    }
    return NULL;
}

parasail_function_t *parasail_lookup_function(const char *funcname)
{
    const parasail_function_info_t *f = parasail_lookup_function_info(funcname);
    return f ? f->pointer : NULL;
}

parasail_pfunction_t *parasail_lookup_pfunction(const char *funcname)
{
    const parasail_pfunction_info_t *f = parasail_lookup_pfunction_info(funcname);
    return f ? f->pointer : NULL;
}

parasail_pcreator_t *parasail_lookup_pcreator(const char *funcname)
{
    const parasail_pfunction_info_t *f = parasail_lookup_pfunction_info(funcname);
    return f ? f->creator : NULL;
}
