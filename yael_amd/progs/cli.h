/*
 * cli.h -- shared pieces of the kmeans and knn command-line programs.
 *
 * Options are described by a table (name, kind, destination, help line);
 * one parser walks argv against it and one printer writes the help text
 * from it, so a program's flags, defaults and help cannot drift apart.
 * The loader reads a vector set in any of the input forms the programs
 * take (fvecs, bvecs widened to float, text) and the writer stores a
 * result set as fvecs/ivecs or text.  Header-only: each program is one
 * translation unit.
 */
#ifndef YAEL_CLI_H_INCLUDED
#define YAEL_CLI_H_INCLUDED

#include <errno.h>
#include <limits.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/yael_mi355x.h"

/* forms of a vector file on disk */
typedef enum { CLI_FVECS, CLI_BVECS, CLI_IVECS, CLI_TEXT } cli_form;

static const char *cli_form_name(cli_form f) {
  static const char *names[] = {"fvecs", "bvecs", "ivecs", "text"};
  return names[f];
}

typedef enum {
  CLI_SECTION, /* not an option: a heading in the help text */
  CLI_INT,     /* takes an integer value, stored in *(int *)dst */
  CLI_PATH,    /* takes a file name, stored in *(const char **)dst, and sets
                  *form_dst = form */
  CLI_SET      /* takes no value: *(int *)dst = form (used as a constant) */
} cli_kind;

typedef struct {
  const char *name; /* "-k"; for CLI_SECTION the heading */
  cli_kind kind;
  void *dst;
  cli_form *form_dst;
  int form;
  const char *help;
} cli_opt;

static void cli_help(FILE *out, const char *prog, const char *what, const cli_opt *t, int nt) {
  fprintf(out, "Usage: %s [-option value]...\n%s\n", prog, what);
  for (int i = 0; i < nt; i++) {
    if (t[i].kind == CLI_SECTION) {
      fprintf(out, "\n%s:\n", t[i].name);
      continue;
    }
    char lhs[48];
    snprintf(lhs, sizeof lhs, "%s%s", t[i].name,
             t[i].kind == CLI_INT ? " N" : t[i].kind == CLI_PATH ? " FILE" : "");
    fprintf(out, "  %-14s %s\n", lhs, t[i].help);
  }
}

static int cli_int_value(const char *prog, const char *opt, const char *s, int *out) {
  char *end = NULL;
  errno = 0;
  const long v = strtol(s, &end, 10);
  if (errno || end == s || *end != '\0' || v < INT_MIN || v > INT_MAX) {
    fprintf(stderr, "%s: option %s wants an integer, got '%s'\n", prog, opt, s);
    return -1;
  }
  *out = (int)v;
  return 0;
}

/* Returns 0 when every argument was understood, 1 when help was asked for
 * (printed to stdout), -1 on a bad argument (reported on stderr). */
static int cli_parse(const char *prog, const char *what, const cli_opt *t, int nt, int argc,
                     char **argv) {
  for (int a = 1; a < argc; a++) {
    const char *arg = argv[a];
    if (!strcmp(arg, "-h") || !strcmp(arg, "--help")) {
      cli_help(stdout, prog, what, t, nt);
      return 1;
    }
    const cli_opt *o = NULL;
    for (int i = 0; i < nt && !o; i++)
      if (t[i].kind != CLI_SECTION && !strcmp(t[i].name, arg)) o = &t[i];
    if (!o) {
      fprintf(stderr, "%s: unknown option '%s' (-h lists them)\n", prog, arg);
      return -1;
    }
    if (o->kind == CLI_SET) {
      *(int *)o->dst = o->form;
      continue;
    }
    if (a + 1 >= argc) {
      fprintf(stderr, "%s: option %s needs a value\n", prog, arg);
      return -1;
    }
    const char *val = argv[++a];
    if (o->kind == CLI_INT) {
      if (cli_int_value(prog, arg, val, (int *)o->dst) < 0) return -1;
    } else {
      *(const char **)o->dst = val;
      if (o->form_dst) *o->form_dst = (cli_form)o->form;
    }
  }
  return 0;
}

/* Loads a vector set.  *n and *d: 0 means "take it from the file" (binary
 * forms); a positive *n keeps only the first *n vectors.  Text input needs
 * *d and *n.  NaN components are replaced by nan_value (the programs' rule
 * for unusable input) and their count reported when verbose.  Returns a
 * malloc'd n x d array, NULL after reporting an error. */
static float *cli_load(const char *prog, const char *path, cli_form form, int *n, int *d,
                       float nan_value, int verbose) {
  int fn = 0, fd = 0;
  if (form == CLI_TEXT) {
    if (*n <= 0 || *d <= 0) {
      fprintf(stderr, "%s: %s is text, so its vector count and dimension must be given\n", prog,
              path);
      return NULL;
    }
    fn = *n;
    fd = *d;
  } else {
    const long bytes = form == CLI_BVECS ? bvecs_fsize(path, &fd, &fn) : fvecs_fsize(path, &fd, &fn);
    if (bytes <= 0 || fn <= 0 || fd <= 0) {
      fprintf(stderr, "%s: %s holds no usable vectors\n", prog, path);
      return NULL;
    }
    if (*d > 0 && *d != fd) {
      fprintf(stderr, "%s: %s has dimension %d, %d expected\n", prog, path, fd, *d);
      return NULL;
    }
    if (*n > 0) {
      if (*n > fn) {
        fprintf(stderr, "%s: %s has %d vectors, %d asked for\n", prog, path, fn, *n);
        return NULL;
      }
      fn = *n;
    }
    if (verbose)
      fprintf(stderr, "%s: %s: %d vectors of dimension %d (%s, %ld bytes)\n", prog, path, fn, fd,
              cli_form_name(form), bytes);
  }
  float *v = (float *)malloc(sizeof(float) * (size_t)fn * fd);
  if (!v) {
    fprintf(stderr, "%s: no memory for %d x %d floats\n", prog, fn, fd);
    return NULL;
  }
  const int got = form == CLI_TEXT    ? fvecs_read_txt(path, fd, fn, v)
                  : form == CLI_BVECS ? b2fvecs_read(path, fd, fn, v)
                                      : fvecs_read(path, fd, fn, v);
  if (got != fn) {
    fprintf(stderr, "%s: read %d of the %d vectors of %s\n", prog, got, fn, path);
    free(v);
    return NULL;
  }
  const long bad = fvec_purge_nans(v, (long)fn * fd, nan_value);
  if (bad && verbose)
    fprintf(stderr, "%s: %s: %ld NaN components set to %g\n", prog, path, bad, nan_value);
  *n = fn;
  *d = fd;
  return v;
}

/* Writes n vectors of dimension d, floats or ints (is_int), in a binary
 * form or as text.  0 on success, -1 after reporting an error. */
static int cli_store(const char *prog, const char *path, cli_form form, int d, int n,
                     const void *v, int is_int) {
  int wrote;
  if (form == CLI_TEXT)
    wrote = is_int ? ivecs_write_txt(path, d, n, (const int *)v)
                   : fvecs_write_txt(path, d, n, (const float *)v);
  else
    wrote = is_int ? ivecs_write(path, d, n, (const int *)v)
                   : fvecs_write(path, d, n, (const float *)v);
  if (wrote != n) {
    fprintf(stderr, "%s: writing %s failed (%d of %d vectors)\n", prog, path, wrote, n);
    return -1;
  }
  return 0;
}

#endif
