/*
 * fe_params.c -- the .in parameter-file grammar and typed queries.
 *
 * Semantics follow SRC/PARAMETERS/parameters.c:
 *   - '#' truncates a line (:119-122); the line splits at its first '=' and
 *     both sides are whitespace-trimmed (:125-133, hashTable.c:174-197);
 *   - an empty value is stored as "no value" (:157-167);
 *   - a repeated key is an error that is counted (:136-140);
 *   - typed queries take [min,max] and a requirement status: an absent
 *     PARAM_OPTIONAL key keeps the caller's default and warns (:287-292), an
 *     absent PARAM_MANDATORY key is an error (:294-301), an out-of-range value
 *     is an INVALID error and leaves the variable untouched (:309-315);
 *   - floats are parsed with sscanf("%f"), integers with sscanf("%d");
 *   - keys never queried are reported as unused (:671).
 * Deviations, each where the reference would dereference a null pointer: a
 * blank line is skipped, a non-blank line without '=' and a query of a key
 * with an empty or non-numeric value are counted as errors.
 */
#include <ctype.h>
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include "fe_host.h"

#define P_LINE_MAX 4096

enum { ST_NOT_USED = 0, ST_VALID, ST_INVALID, ST_DEFAULTED, ST_ABSENT };
static const char *st_name[] = {"NOT USED", "IN VALUE", "INVALID", "DEFAULTED", "ABSENT"};

typedef struct {
  char *name;
  char *value;   /* NULL when the .in had an empty value or the key was absent */
  int state;
  char shown[64];
} pentry;

static pentry *tab = NULL;
static int ntab = 0, captab = 0;
static int numErrors = 0;
static char *outbuf = NULL;
static size_t outlen = 0, outcap = 0;

static void out_append(const char *s) {
  size_t n = strlen(s);
  if (outlen + n + 1 > outcap) {
    size_t nc = outcap ? outcap * 2 : 8192;
    while (nc < outlen + n + 1) nc *= 2;
    char *p = (char *)realloc(outbuf, nc);
    if (!p) return;
    outbuf = p;
    outcap = nc;
  }
  memcpy(outbuf + outlen, s, n + 1);
  outlen += n;
}

static char *trim(char *s) {
  while (*s && isspace((unsigned char)*s)) s++;
  size_t n = strlen(s);
  while (n > 0 && isspace((unsigned char)s[n - 1])) s[--n] = 0;
  return s;
}

static pentry *lookup(const char *name) {
  for (int i = 0; i < ntab; i++)
    if (strcmp(tab[i].name, name) == 0) return &tab[i];
  return NULL;
}

static pentry *add(const char *name, const char *value) {
  if (ntab == captab) {
    int nc = captab ? captab * 2 : 128;
    pentry *p = (pentry *)realloc(tab, sizeof(pentry) * nc);
    if (!p) return NULL;
    tab = p;
    captab = nc;
  }
  pentry *e = &tab[ntab++];
  memset(e, 0, sizeof(*e));
  e->name = strdup(name);
  e->value = value ? strdup(value) : NULL;
  e->state = ST_NOT_USED;
  return e;
}

void parameters_init(void) {
  parameters_clean();
  numErrors = 0;
}

void parameters_clean(void) {
  for (int i = 0; i < ntab; i++) { free(tab[i].name); free(tab[i].value); }
  free(tab);
  tab = NULL; ntab = captab = 0;
  free(outbuf);
  outbuf = NULL; outlen = outcap = 0;
}

int parameters_readFile(const char *paramFile) {
  FILE *fp = fopen(paramFile, "r");
  if (!fp) {
    printf("ERROR: Failed to open parameters file, %s\n", paramFile);
    numErrors++;
    return PARAM_ERROR_FILE_NOT_FOUND;
  }
  int rc = PARAM_SUCCESS;
  char line[P_LINE_MAX];
  int lineno = 0;
  while (fgets(line, sizeof line, fp)) {
    lineno++;
    char *hash = strchr(line, '#');
    if (hash) *hash = 0;
    char *eq = strchr(line, '=');
    if (!eq) {
      if (*trim(line)) {
        printf("ERROR: line %d of %s has no '=': %s\n", lineno, paramFile, trim(line));
        numErrors++;
        rc = PARAM_ERROR_INVALID_FLOAT;
      }
      continue;
    }
    *eq = 0;
    char *name = trim(line);
    char *value = trim(eq + 1);
    if (lookup(name)) {
      printf("ERROR: Duplicate entries for parameter name--%s.\n", name);
      numErrors++;
      rc = PARAM_ERROR_DUPLICATE;
      continue;
    }
    add(name, *value ? value : NULL);
  }
  fclose(fp);
  return rc;
}

/* Shared absent-key handling of the typed queries. */
static int absent(const char *name, int reqStatus, const char *deflt) {
  if (reqStatus == PARAM_NOT_REQUIRED) return PARAM_SUCCESS;
  pentry *e = add(name, NULL);
  if (reqStatus == PARAM_OPTIONAL) {
    e->state = ST_DEFAULTED;
    snprintf(e->shown, sizeof e->shown, "%s", deflt);
    printf("WARNING: optional parameter %s defaulted to %s.\n", name, deflt);
    return PARAM_WARN_DEFAULTED;
  }
  e->state = ST_ABSENT;
  snprintf(e->shown, sizeof e->shown, "??");
  numErrors++;
  printf("ERROR: mandatory parameter %s has not been specified.\n", name);
  return PARAM_ERROR_ABSENT;
}

int queryFloatParameter(const char *name, float *var, float min, float max, int reqStatus) {
  pentry *e = lookup(name);
  if (!e) {
    char d[64];
    snprintf(d, sizeof d, "%g", *var);
    return absent(name, reqStatus, d);
  }
  float val;
  if (!e->value || sscanf(e->value, "%f", &val) != 1) {
    e->state = ST_INVALID;
    numErrors++;
    printf("ERROR: parameter '%s' has no numeric value.\n", name);
    return PARAM_ERROR_INVALID_FLOAT;
  }
  if (val < min || val > max) {
    e->state = ST_INVALID;
    numErrors++;
    printf("ERROR: parameter '%s' value %g is outside limits [%g,%g].\n", name, val, min, max);
    return PARAM_ERROR_INVALID_FLOAT;
  }
  e->state = ST_VALID;
  snprintf(e->shown, sizeof e->shown, "%g", val);
  *var = val;
  return PARAM_SUCCESS;
}

int queryDoubleParameter(const char *name, double *var, double min, double max, int reqStatus) {
  pentry *e = lookup(name);
  if (!e) {
    char d[64];
    snprintf(d, sizeof d, "%g", *var);
    return absent(name, reqStatus, d);
  }
  double val;
  if (!e->value || sscanf(e->value, "%lf", &val) != 1) {
    e->state = ST_INVALID;
    numErrors++;
    printf("ERROR: parameter '%s' has no numeric value.\n", name);
    return PARAM_ERROR_INVALID_DOUBLE;
  }
  if (val < min || val > max) {
    e->state = ST_INVALID;
    numErrors++;
    printf("ERROR: parameter '%s' value %g is outside limits [%g,%g].\n", name, val, min, max);
    return PARAM_ERROR_INVALID_DOUBLE;
  }
  e->state = ST_VALID;
  snprintf(e->shown, sizeof e->shown, "%g", val);
  *var = val;
  return PARAM_SUCCESS;
}

int queryIntegerParameter(const char *name, int *var, int min, int max, int reqStatus) {
  pentry *e = lookup(name);
  if (!e) {
    char d[64];
    snprintf(d, sizeof d, "%d", *var);
    return absent(name, reqStatus, d);
  }
  int val;
  if (!e->value || sscanf(e->value, "%d", &val) != 1) {
    e->state = ST_INVALID;
    numErrors++;
    printf("ERROR: parameter '%s' has no integer value.\n", name);
    return PARAM_ERROR_INVALID_INTEGER;
  }
  if (val < min || val > max) {
    e->state = ST_INVALID;
    numErrors++;
    printf("ERROR: parameter '%s' value %d is outside limits [%d,%d].\n", name, val, min, max);
    return PARAM_ERROR_INVALID_INTEGER;
  }
  e->state = ST_VALID;
  snprintf(e->shown, sizeof e->shown, "%d", val);
  *var = val;
  return PARAM_SUCCESS;
}

/* String queries return NULL for an empty value (parameters.c:460-519). */
int queryStringParameter(const char *name, char **str, int reqStatus) {
  *str = NULL;
  pentry *e = lookup(name);
  if (!e) {
    if (reqStatus == PARAM_NOT_REQUIRED) return PARAM_ERROR_NOT_FOUND;
    return absent(name, reqStatus, "(null)");
  }
  e->state = ST_VALID;
  snprintf(e->shown, sizeof e->shown, "%s", e->value ? e->value : "");
  if (e->value) *str = strdup(e->value);
  return PARAM_SUCCESS;
}

int queryFileParameter(const char *name, char **file, int reqStatus) {
  int rc = queryStringParameter(name, file, reqStatus);
  if (rc == PARAM_ERROR_ABSENT || *file == NULL) return rc;
  struct stat sb;
  if (stat(*file, &sb) == -1 && errno == ENOENT) {
    if (reqStatus == PARAM_MANDATORY) {
      numErrors++;
      printf("ERROR: could not find file %s.\n", *file);
      lookup(name)->state = ST_INVALID;
      return PARAM_ERROR_FILE_NOT_FOUND;
    }
    printf("WARNING: could not find file %s.\n", *file);
  }
  return rc;
}

int queryPathParameter(const char *name, char **path, int reqStatus) {
  int rc = queryStringParameter(name, path, reqStatus);
  if (rc == PARAM_ERROR_ABSENT || *path == NULL) return rc;
  struct stat sb;
  if (stat(*path, &sb) == -1 && errno == ENOENT) {
    if (reqStatus == PARAM_MANDATORY) {
      numErrors++;
      printf("ERROR: could not find path %s.\n", *path);
      lookup(name)->state = ST_INVALID;
      return PARAM_ERROR_PATH_NOT_FOUND;
    }
    printf("WARNING: could not find path %s.\n", *path);
  }
  return rc;
}

int getParameterErrors(void) { return numErrors; }

int countUnusedParameters(void) {
  int n = 0;
  for (int i = 0; i < ntab; i++)
    if (tab[i].state == ST_NOT_USED) n++;
  return n;
}

void printUnusedParameters(void) {
  for (int i = 0; i < ntab; i++)
    if (tab[i].state == ST_NOT_USED)
      printf("%s = %s   # %s\n", tab[i].name, tab[i].value ? tab[i].value : "(null)", st_name[ST_NOT_USED]);
}

void printComment(const char *comment) {
  out_append("## ");
  out_append(comment);
  out_append("\n");
}

void printParameter(const char *name, const char *description) {
  pentry *e = lookup(name);
  if (!e) {
    printf("WARNING: parameter %s was not found but was requested for printing.\n", name);
    return;
  }
  if (e->state == ST_NOT_USED) return;
  char line[512];
  int n = snprintf(line, sizeof line, "%s=%s", name, e->shown);
  while (n < 18 && n < (int)sizeof line - 1) line[n++] = ' ';
  line[n] = 0;
  out_append(line);
  out_append(" # ");
  if (e->state != ST_VALID) {
    out_append("(");
    out_append(st_name[e->state]);
    out_append(") ");
  }
  printComment(description);
}

void outputParameters(FILE *out) { fprintf(out, "\n%s\n", outbuf ? outbuf : ""); }
