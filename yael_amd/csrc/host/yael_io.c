/*
 * yael_io.c -- vector files in yael's exchange format (doc/file_format.rst:8-23).
 *
 * A file is a plain concatenation of records, with no file header.  A record
 * is one vector: its length d as a 4-byte little-endian int, then its d
 * components -- 4-byte floats (.fvecs), 4-byte ints (.ivecs) or bytes
 * (.bvecs).  The text form holds whitespace-separated numbers, one vector
 * per line.
 *
 * The function names, arguments and return values are those declared in
 * yael/vector.h:181-270, 325, 460 (counts of vectors read or written, -1 on
 * error), so callers of the reference link against these unchanged and the
 * files interchange with the reference's own programs.  Everything here is
 * host code.
 *
 * Design: every binary read goes through one record cursor (yio_cursor)
 * that knows the component size, checks each record's length against the
 * expected dimension and reports the three ways a read can end -- a clean
 * end of file between records, a record of another length, a record cut
 * short.  Readers that allocate size the result from the file length first,
 * so they make one allocation and never copy.
 */
#include <errno.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../../include/yael_mi355x.h"

/* ------------------------------------------------------------------ */
/* record cursor                                                       */

enum { YIO_END = 0, YIO_RECORD = 1, YIO_BADLEN = -1, YIO_CUT = -2, YIO_SYS = -3 };

typedef struct {
  FILE *f;
  const char *path;
  size_t esize; /* bytes per component: 4 (float, int) or 1 (byte) */
} yio_cursor;

static int yio_open(yio_cursor *c, const char *path, size_t esize, const char *fn) {
  c->path = path;
  c->esize = esize;
  c->f = fopen(path, "rb");
  if (!c->f) {
    fprintf(stderr, "%s: cannot open '%s' for reading (%s)\n", fn, path, strerror(errno));
    return -1;
  }
  return 0;
}

/* Reads the next record's length into *len, then, if dst is not NULL, its
 * payload into dst (which must hold want components).  want < 0 accepts any
 * length.  Returns YIO_RECORD, YIO_END (fewer than 4 bytes left: no further
 * record starts, the readers' end-of-data rule), YIO_BADLEN (length differs
 * from want; payload not read), YIO_CUT (the payload stops early) or
 * YIO_SYS. */
static int yio_next(yio_cursor *c, int want, void *dst, int *len) {
  unsigned char w[4];
  const size_t got = fread(w, 1, 4, c->f);
  if (got < 4) return ferror(c->f) ? YIO_SYS : YIO_END;
  int32_t n;
  memcpy(&n, w, 4);
  *len = n;
  if (n < 0 || (want >= 0 && n != want)) return YIO_BADLEN;
  if (!dst) return YIO_RECORD;
  if (fread(dst, c->esize, (size_t)n, c->f) != (size_t)n) return ferror(c->f) ? YIO_SYS : YIO_CUT;
  return YIO_RECORD;
}

static void yio_complain(const yio_cursor *c, const char *fn, int status, long rec, int len,
                         int want) {
  switch (status) {
  case YIO_BADLEN:
    fprintf(stderr, "%s: '%s' record %ld has length %d, expected %d\n", fn, c->path, rec, len,
            want);
    break;
  case YIO_CUT:
    fprintf(stderr, "%s: '%s' ends inside record %ld\n", fn, c->path, rec);
    break;
  default:
    fprintf(stderr, "%s: reading '%s' failed (%s)\n", fn, c->path, strerror(errno));
  }
}

/* Reads up to nmax records of length d into dst (components of esize
 * bytes, stored back to back).  Returns the number read (fewer when the
 * file ends first), -1 on any malformed record. */
static long yio_read_records(const char *path, size_t esize, int d, long nmax, void *dst,
                             const char *fn) {
  yio_cursor c;
  if (yio_open(&c, path, esize, fn) < 0) return -1;
  unsigned char *out = (unsigned char *)dst;
  long i = 0;
  for (; i < nmax; i++) {
    int len = 0;
    const int st = yio_next(&c, d, out + (size_t)i * d * esize, &len);
    if (st == YIO_END) break;
    if (st != YIO_RECORD) {
      yio_complain(&c, fn, st, i, len, d);
      fclose(c.f);
      return -1;
    }
  }
  fclose(c.f);
  return i;
}

/* ------------------------------------------------------------------ */
/* sizes                                                               */

/* Dimension from the first record and the file length.  Returns 1 with
 * *d, *bytes set; 0 for an empty file; -1 after reporting an error. */
static int yio_probe(const char *path, size_t esize, int *d, long *bytes, const char *fn) {
  yio_cursor c;
  if (yio_open(&c, path, esize, fn) < 0) return -1;
  const int st = yio_next(&c, -1, NULL, d);
  *bytes = -1;
  if (st == YIO_RECORD && fseek(c.f, 0, SEEK_END) == 0) *bytes = ftell(c.f);
  fclose(c.f);
  if (st == YIO_END) return 0;
  if (*bytes < 0) {
    fprintf(stderr, "%s: cannot size '%s'\n", fn, path);
    return -1;
  }
  return 1;
}

/* Count from the file length: every record has esize*d + 4 bytes.  Returns
 * the file length in bytes, -1 on error; an empty file is 0 vectors
 * (return 0, *d_out = -1). */
static long yio_fsize(const char *path, size_t esize, int *d_out, int *n_out, const char *fn) {
  *d_out = -1;
  *n_out = -1;
  int d = 0;
  long bytes = 0;
  const int st = yio_probe(path, esize, &d, &bytes, fn);
  if (st <= 0) {
    if (st == 0) *n_out = 0;
    return st;
  }
  const long rec = (long)esize * d + 4;
  if (bytes % rec != 0) {
    fprintf(stderr, "%s: '%s' is %ld bytes, not a whole number of %ld-byte records (d=%d)\n",
            fn, path, bytes, rec, d);
    return -1;
  }
  *d_out = d;
  *n_out = (int)(bytes / rec);
  return bytes;
}

long fvecs_fsize(const char *fname, int *d_out, int *n_out) {
  return yio_fsize(fname, 4, d_out, n_out, "fvecs_fsize");
}
long ivecs_fsize(const char *fname, int *d_out, int *n_out) {
  return yio_fsize(fname, 4, d_out, n_out, "ivecs_fsize");
}
long bvecs_fsize(const char *fname, int *d_out, int *n_out) {
  return yio_fsize(fname, 1, d_out, n_out, "bvecs_fsize");
}

/* ------------------------------------------------------------------ */
/* readers into caller memory                                          */

int fvecs_read(const char *fname, int d, int n, float *v) {
  return (int)yio_read_records(fname, 4, d, n, v, "fvecs_read");
}

/* byte vectors widened to floats, one record at a time through a d-byte
 * buffer */
int b2fvecs_read(const char *fname, int d, int n, float *v) {
  if (d <= 0 || n < 0) {
    fprintf(stderr, "b2fvecs_read: bad shape d=%d n=%d\n", d, n);
    return -1;
  }
  yio_cursor c;
  if (yio_open(&c, fname, 1, "b2fvecs_read") < 0) return -1;
  unsigned char *rec = (unsigned char *)malloc((size_t)d);
  if (!rec) {
    fclose(c.f);
    return -1;
  }
  long i = 0;
  for (; i < n; i++) {
    int len = 0;
    const int st = yio_next(&c, d, rec, &len);
    if (st == YIO_END) break;
    if (st != YIO_RECORD) {
      yio_complain(&c, "b2fvecs_read", st, i, len, d);
      i = -1;
      break;
    }
    float *dst = v + (size_t)i * d;
    for (int j = 0; j < d; j++) dst[j] = (float)rec[j];
  }
  free(rec);
  fclose(c.f);
  return (int)i;
}

/* n*d numbers; returns the number of whole vectors read before the text
 * ran out (a trailing partial vector is read but not counted) */
int fvecs_read_txt(const char *fname, int d, int n, float *v) {
  FILE *f = fopen(fname, "r");
  if (!f) {
    fprintf(stderr, "fvecs_read_txt: cannot open '%s' (%s)\n", fname, strerror(errno));
    return -1;
  }
  const long want = (long)n * d;
  long got = 0;
  while (got < want) {
    const int r = fscanf(f, "%f", v + got);
    if (r == 1) {
      got++;
      continue;
    }
    if (r == EOF && !ferror(f)) break;
    fprintf(stderr, "fvecs_read_txt: '%s': unreadable number after %ld values\n", fname, got);
    fclose(f);
    return -1;
  }
  fclose(f);
  return d > 0 ? (int)(got / d) : 0;
}

/* ------------------------------------------------------------------ */
/* readers that allocate                                               */

/* Every record of a 4-byte-component file: counted from the file length,
 * one malloc, each record checked to have the first record's length.  Up
 * to 3 trailing bytes are no record (the end-of-data rule of yio_next);
 * more, or a record of another length, is an error. */
static int yio_new_read4(const char *fname, int *d_out, void **out, const char *fn) {
  *d_out = -1;
  *out = NULL;
  int d = 0;
  long bytes = 0;
  const int st = yio_probe(fname, 4, &d, &bytes, fn);
  if (st <= 0) return st;
  const long rec = 4L * d + 4, n = bytes / rec;
  if (bytes % rec >= 4) {
    fprintf(stderr, "%s: '%s' ends inside a record (%ld bytes, %ld-byte records)\n", fn, fname,
            bytes, rec);
    return -1;
  }
  void *buf = malloc((size_t)n * d * 4 + 4); /* +4: d == 0 still gives a pointer */
  if (!buf) {
    fprintf(stderr, "%s: no memory for %ld vectors of dimension %d\n", fn, n, d);
    return -1;
  }
  const long got = yio_read_records(fname, 4, d, n, buf, fn);
  if (got != n) {
    if (got >= 0) fprintf(stderr, "%s: '%s' changed while being read\n", fn, fname);
    free(buf);
    return -1;
  }
  *d_out = d;
  *out = buf;
  return (int)n;
}

int fvecs_new_read(const char *fname, int *d_out, float **vf_out) {
  return yio_new_read4(fname, d_out, (void **)vf_out, "fvecs_new_read");
}

int ivecs_new_read(const char *fname, int *d_out, int **vi_out) {
  return yio_new_read4(fname, d_out, (void **)vi_out, "ivecs_new_read");
}

/* ------------------------------------------------------------------ */
/* writers                                                             */

/* One record: the length word, then d components of 4 bytes.  0 when the
 * record is written; -1 when the length word fails, -2 when the payload
 * fails (the two error codes vector.h's ivec_fwrite callers can tell
 * apart). */
static int yio_put4(FILE *f, const void *v, int d, const char *fn) {
  const int32_t len = d;
  if (fwrite(&len, 4, 1, f) != 1) {
    fprintf(stderr, "%s: writing a record length failed (%s)\n", fn, strerror(errno));
    return -1;
  }
  if (fwrite(v, 4, (size_t)d, f) != (size_t)d) {
    fprintf(stderr, "%s: writing %d components failed (%s)\n", fn, d, strerror(errno));
    return -2;
  }
  return 0;
}

int fvec_fwrite(FILE *fo, const float *v, int d) {
  return yio_put4(fo, v, d, "fvec_fwrite") == 0 ? 0 : -1;
}

int ivec_fwrite(FILE *f, const int *v, int d) {
  return yio_put4(f, v, d, "ivec_fwrite");
}

/* the number of records written before the first failure */
int fvecs_fwrite(FILE *fo, int d, int n, const float *vf) {
  int i = 0;
  while (i < n && fvec_fwrite(fo, vf + (size_t)i * d, d) == 0) i++;
  return i;
}

/* every record is attempted; the count asked for is returned */
int ivecs_fwrite(FILE *f, int d, int n, const int *v) {
  for (int i = 0; i < n; i++) (void)ivec_fwrite(f, v + (size_t)i * d, d);
  return n;
}

static FILE *yio_create(const char *fname, const char *mode, const char *fn) {
  FILE *f = fopen(fname, mode);
  if (!f) fprintf(stderr, "%s: cannot create '%s' (%s)\n", fn, fname, strerror(errno));
  return f;
}

int fvecs_write(const char *fname, int d, int n, const float *vf) {
  FILE *f = yio_create(fname, "wb", "fvecs_write");
  if (!f) return -1;
  const int wrote = fvecs_fwrite(f, d, n, vf);
  fclose(f);
  return wrote;
}

int ivecs_write(const char *fname, int d, int n, const int *v) {
  FILE *f = yio_create(fname, "wb", "ivecs_write");
  if (!f) return -1;
  const int wrote = ivecs_fwrite(f, d, n, v);
  fclose(f);
  return wrote;
}

/* Text form: each component printed with the given conversion and a
 * trailing space, a newline after each vector.  Returns the number of
 * vector lines ended (n when every write succeeds). */
static int yio_write_txt(const char *fname, int d, int n, const void *v, int is_int,
                         const char *fn) {
  FILE *f = yio_create(fname, "w", fn);
  if (!f) return -1;
  int lines = 0;
  for (long i = 0; i < n; i++) {
    for (long j = 0; j < d; j++) {
      const size_t at = (size_t)i * d + j;
      if (is_int)
        fprintf(f, "%d ", ((const int *)v)[at]);
      else
        fprintf(f, "%f ", ((const float *)v)[at]);
    }
    lines += fputc('\n', f) == '\n';
  }
  fclose(f);
  return lines;
}

int fvecs_write_txt(const char *fname, int d, int n, const float *vf) {
  return yio_write_txt(fname, d, n, vf, 0, "fvecs_write_txt");
}

int ivecs_write_txt(const char *fname, int d, int n, const int *v) {
  return yio_write_txt(fname, d, n, v, 1, "ivecs_write_txt");
}

/* ------------------------------------------------------------------ */

long fvec_purge_nans(float *v, long n, float replace_value) {
  long replaced = 0;
  for (long i = 0; i < n; i++) {
    if (isnan(v[i])) {
      v[i] = replace_value;
      replaced++;
    }
  }
  return replaced;
}
