/*
 * params.c -- the parameter-file readers main.c uses (main.c:20-35,
 * 176-180): whitespace-separated "key value" lines, and lines whose first
 * token is a specifier followed by key=value pairs ("problem_1 density=1.0
 * ...", "flow.params density=... energy=..."). '#' starts a comment line.
 */
#include <ctype.h>
#include <errno.h>
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../../include/flow_arch.h"

/* Finds the first line whose first token equals `key`; copies the rest of
 * the line (after the token) into rest. Returns 1 if found. A line longer
 * than the reader's buffer is an error (read in pieces, its tail would pass
 * for lines of its own). */
static int find_line(const char* key, const char* filename, char* rest,
                     size_t rest_len) {
  FILE* f = fopen(filename, "r");
  if (!f) return -1;
  char line[MAX_STR_LEN * 4];
  int found = 0;
  int lineno = 0;
  while (fgets(line, sizeof(line), f)) {
    ++lineno;
    if (!strchr(line, '\n') && !feof(f)) {
      fclose(f);
      TERMINATE("Line %d of %s is longer than %d characters.", lineno, filename,
                (int)sizeof(line) - 2);
    }
    char* p = line;
    while (*p && isspace((unsigned char)*p)) ++p;
    if (*p == '\0' || *p == '#') continue;
    char* q = p;
    while (*q && !isspace((unsigned char)*q)) ++q;
    const size_t klen = (size_t)(q - p);
    if (klen != strlen(key) || strncmp(p, key, klen) != 0) continue;
    while (*q && isspace((unsigned char)*q)) ++q;
    strncpy(rest, q, rest_len - 1);
    rest[rest_len - 1] = '\0';
    found = 1;
    break;
  }
  fclose(f);
  return found;
}

static void value_token(const char* param_name, const char* filename,
                        char* out, size_t out_len) {
  char rest[MAX_STR_LEN * 4];
  const int r = find_line(param_name, filename, rest, sizeof(rest));
  if (r < 0) TERMINATE("Could not open the parameter file %s.", filename);
  if (r == 0)
    TERMINATE("Could not find the parameter %s in file %s.", param_name,
              filename);
  char* p = rest;
  size_t n = 0;
  while (*p && !isspace((unsigned char)*p) && n + 1 < out_len) out[n++] = *p++;
  out[n] = '\0';
  if (n == 0)
    TERMINATE("The parameter %s in file %s has no value.", param_name,
              filename);
}

int get_int_parameter(const char* param_name, const char* filename) {
  char tok[MAX_STR_LEN];
  value_token(param_name, filename, tok, sizeof(tok));
  char* end = NULL;
  errno = 0;
  const long v = strtol(tok, &end, 10);
  if (*end != '\0')
    TERMINATE("The parameter %s in file %s is not an integer: %s", param_name,
              filename, tok);
  if (errno == ERANGE || v < INT_MIN || v > INT_MAX)
    TERMINATE("The parameter %s in file %s is out of range: %s", param_name,
              filename, tok);
  return (int)v;
}

double get_double_parameter(const char* param_name, const char* filename) {
  char tok[MAX_STR_LEN];
  value_token(param_name, filename, tok, sizeof(tok));
  char* end = NULL;
  errno = 0;
  const double v = strtod(tok, &end);
  if (*end != '\0')
    TERMINATE("The parameter %s in file %s is not a number: %s", param_name,
              filename, tok);
  if (errno == ERANGE && isinf(v))
    TERMINATE("The parameter %s in file %s is out of range: %s", param_name,
              filename, tok);
  return v;
}

int get_key_value_parameter(const char* specifier, const char* filename,
                            char* keys, double* values, int* nkeys) {
  char rest[MAX_STR_LEN * 4];
  *nkeys = 0;
  if (find_line(specifier, filename, rest, sizeof(rest)) != 1) return 0;
  char* save = NULL;
  for (char* tok = strtok_r(rest, " \t\r\n", &save); tok;
       tok = strtok_r(NULL, " \t\r\n", &save)) {
    char* eq = strchr(tok, '=');
    if (!eq)
      TERMINATE("Malformed key=value entry '%s' for %s in %s.", tok,
                specifier, filename);
    if (*nkeys == MAX_KEYS)
      TERMINATE("Too many keys for %s in %s.", specifier, filename);
    *eq = '\0';
    char* key = keys + (size_t)(*nkeys) * MAX_STR_LEN;
    strncpy(key, tok, MAX_STR_LEN - 1);
    key[MAX_STR_LEN - 1] = '\0';
    char* end = NULL;
    values[*nkeys] = strtod(eq + 1, &end);
    if (end == eq + 1 || *end != '\0')
      TERMINATE("Value of %s for %s in %s is not a number.", tok, specifier,
                filename);
    ++*nkeys;
  }
  return 1;
}
