/*
 * pvparams.h -- C-ABI of the .params reader of the C++ host (libpvhost).
 *
 * Reads OpenPV's params language (the grammar of src/io/parser/params.y and
 * the tokens of params.l, handled by PVParams, src/io/PVParams.cpp):
 *
 *   name = number;                                  (top-level directive)
 *   Keyword "GroupName" = {
 *       key = number;        key = "string";        key = NULL;
 *       key = [n, n, ...];   #include "OtherGroup";  @key = value;
 *   };
 *
 * Numbers follow strtod; true/false read as 1/0, infinity/-infinity as
 * +/-FLT_MAX.  #include copies the parameters of an earlier group and @key
 * overwrites one of them.  Comments start with //.  ParameterSweep blocks are
 * not supported (an error).
 */
#ifndef PVPARAMS_H_
#define PVPARAMS_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pvparams pvparams;

/* Parses a file; returns 0 or a negative value with the message in
 * pvparams_error().  *out must be released with pvparams_free. */
int pvparams_load(const char *path, pvparams **out);
/* Parses params text held in memory. */
int pvparams_parse(const char *text, pvparams **out);
const char *pvparams_error(void);
void pvparams_free(pvparams *p);

int pvparams_num_groups(const pvparams *p);
const char *pvparams_group_keyword(const pvparams *p, int group);
const char *pvparams_group_name(const pvparams *p, int group);
/* Index of the group with this name, or -1. */
int pvparams_find_group(const pvparams *p, const char *name);

/* Numeric parameter: returns 1 and sets *value if present (an array's first
 * element), 0 if absent. */
int pvparams_number(const pvparams *p, int group, const char *key, double *value);
/* String parameter: returns 1 and sets *value (NULL for NULL) if present. */
int pvparams_string(const pvparams *p, int group, const char *key, const char **value);
/* Array parameter: number of values (0 if absent); copies at most max of them. */
int pvparams_array(const pvparams *p, int group, const char *key, double *values, int max);
/* Top-level directive such as debugParsing; returns 1 if present. */
int pvparams_directive(const pvparams *p, const char *key, double *value);

#ifdef __cplusplus
}
#endif

#endif /* PVPARAMS_H_ */
