/*
 * yael_io.c -- yael's vector files (doc/file_format.rst, vector.c I/O):
 *   fvecs: per vector, int d then d float32;  ivecs: int d then d int32;
 *   bvecs: int d then d uint8;  text: whitespace-separated numbers.
 * Same names, arguments, return values (count of vectors, -1 on error) and
 * output bytes as vector.c, so files interchange with the reference and
 * its command-line programs.  Host code only: no GPU involved.
 */
#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../../include/yael_mi355x.h"

/* vector.c:593-626: d from the first record, n from the file size */
static long yh_vecs_fsize(long unit, const char *fname, int *d_out, int *n_out) {
  *d_out = -1;
  *n_out = -1;
  FILE *f = fopen(fname, "rb");
  if (!f) {
    fprintf(stderr, "xvecs_fsize %s: %s\n", fname, strerror(errno));
    return -1;
  }
  int d;
  if (fread(&d, sizeof d, 1, f) != 1) { /* empty file */
    fclose(f);
    *n_out = 0;
    return 0;
  }
  fseek(f, 0, SEEK_END);
  const long nbytes = ftell(f);
  fclose(f);
  if (d < 0 || nbytes % (unit * d + 4) != 0) {
    fprintf(stderr, "xvecs_size %s: weird file size %ld for vectors of dimension %d\n", fname,
            nbytes, d);
    return -1;
  }
  *d_out = d;
  *n_out = (int)(nbytes / (unit * d + 4));
  return nbytes;
}

/* vector.c:628-641 */
long fvecs_fsize(const char *fname, int *d_out, int *n_out) {
  return yh_vecs_fsize(sizeof(float), fname, d_out, n_out);
}
long ivecs_fsize(const char *fname, int *d_out, int *n_out) {
  return yh_vecs_fsize(sizeof(int), fname, d_out, n_out);
}
long bvecs_fsize(const char *fname, int *d_out, int *n_out) {
  return yh_vecs_fsize(1, fname, d_out, n_out);
}

/* up to n records of dimension d (unit bytes each value) into v; the
 * number read (a short file ends early), -1 on a dimension mismatch */
static int yh_vecs_read(long unit, const char *fname, int d, int n, void *v, const char *who) {
  FILE *f = fopen(fname, "rb");
  if (!f) {
    fprintf(stderr, "%s: could not open %s\n", who, fname);
    perror("");
    return -1;
  }
  long i;
  for (i = 0; i < n; i++) {
    int dd;
    if (fread(&dd, sizeof dd, 1, f) != 1) {
      if (feof(f)) break;
      perror(who);
      fclose(f);
      return -1;
    }
    if (dd != d) {
      fprintf(stderr, "%s error 2: unexpected vector dimension\n", who);
      fclose(f);
      return -1;
    }
    if (fread((char *)v + unit * d * i, unit, d, f) != (size_t)d) {
      fprintf(stderr, "%s error 3\n", who);
      fclose(f);
      return -1;
    }
  }
  fclose(f);
  return (int)i;
}

/* vector.c:882-920 */
int fvecs_read(const char *fname, int d, int n, float *v) {
  return yh_vecs_read(sizeof(float), fname, d, n, v, "fvecs_read");
}

/* vector.c:923-936, 1253-1275, 1339-1362: bvecs converted to float */
int b2fvecs_read(const char *fname, int d, int n, float *v) {
  int dd, nn;
  if (bvecs_fsize(fname, &dd, &nn) < 0 || dd != d || n > nn) {
    fprintf(stderr, "b2fvecs_read: %s does not hold %d vectors of dimension %d\n", fname, n, d);
    return -1;
  }
  unsigned char *b = (unsigned char *)malloc((size_t)n * d + 1);
  if (!b) return -1;
  const int got = yh_vecs_read(1, fname, d, n, b, "b2fvecs_read");
  for (long i = 0; i < (long)(got > 0 ? got : 0) * d; i++) v[i] = b[i];
  free(b);
  return got;
}

/* vector.c:939-962: n*d numbers; the number of complete vectors read */
int fvecs_read_txt(const char *fname, int d, int n, float *v) {
  FILE *f = fopen(fname, "r");
  if (!f) {
    fprintf(stderr, "fvecs_read_txt: could not open %s\n", fname);
    perror("");
    return -1;
  }
  long i;
  for (i = 0; i < (long)n * d; i++) {
    if (fscanf(f, "%f", v + i) != 1) {
      if (feof(f)) break;
      perror("fvecs_read_txt error 1");
      fclose(f);
      return -1;
    }
  }
  fclose(f);
  return (int)(i / d);
}

/* vector.c:651-666, 671-720: all records of a file, dimension from the
 * first (malloc'd; the caller frees) */
static int yh_vecs_new_read(const char *fname, int *d_out, void **out, const char *who) {
  *d_out = -1;
  *out = NULL;
  FILE *f = fopen(fname, "rb");
  if (!f) {
    fprintf(stderr, "%s: could not open %s\n", who, fname);
    perror("");
    return -1;
  }
  long n = 0, cap = 0;
  int d = -1;
  char *buf = NULL;
  for (;;) {
    int dd;
    if (fread(&dd, sizeof dd, 1, f) != 1) {
      if (feof(f)) break;
      perror(who);
      goto fail;
    }
    if (n == 0)
      d = dd;
    else if (dd != d) {
      fprintf(stderr, "%s: non-uniform vectors sizes: new_d=%d d=%d\n", who, dd, d);
      goto fail;
    }
    if (d < 0) goto fail;
    if ((n + 1) * d > cap) {
      cap = cap ? cap * 2 : 1024;
      while ((n + 1) * d > cap) cap *= 2;
      char *nb = (char *)realloc(buf, (size_t)cap * 4);
      if (!nb) goto fail;
      buf = nb;
    }
    if (fread(buf + (size_t)n * d * 4, 4, d, f) != (size_t)d) {
      fprintf(stderr, "%s: truncated record\n", who);
      goto fail;
    }
    n++;
  }
  fclose(f);
  *d_out = d;
  *out = buf;
  return (int)n;
fail:
  fclose(f);
  free(buf);
  return -1;
}

int fvecs_new_read(const char *fname, int *d_out, float **vf_out) {
  return yh_vecs_new_read(fname, d_out, (void **)vf_out, "fvecs_new_read");
}

/* vector.c:1307-1322 */
int ivecs_new_read(const char *fname, int *d_out, int **vi_out) {
  return yh_vecs_new_read(fname, d_out, (void **)vi_out, "ivecs_new_read");
}

/* vector.c:1369-1383: one record; 0, or -1 on a write error */
int fvec_fwrite(FILE *fo, const float *v, int d) {
  if (fwrite(&d, sizeof d, 1, fo) != 1 || fwrite(v, sizeof(float), d, fo) != (size_t)d) {
    perror("fvec_fwrite: write error");
    return -1;
  }
  return 0;
}

/* vector.c:1446-1456: the number of records written */
int fvecs_fwrite(FILE *fo, int d, int n, const float *vf) {
  for (int i = 0; i < n; i++)
    if (fvec_fwrite(fo, vf + (long)i * d, d) < 0) return i;
  return n;
}

/* vector.c:1459-1472 */
int fvecs_write(const char *fname, int d, int n, const float *vf) {
  FILE *fo = fopen(fname, "wb");
  if (!fo) {
    fprintf(stderr, "fvecs_write: cannot open %s for writing", fname);
    perror("");
    return -1;
  }
  const int ret = fvecs_fwrite(fo, d, n, vf);
  fclose(fo);
  return ret;
}

/* vector.c:1475-1492: "%f " per value, a newline per vector; the return
 * value is the sum of the newline writes' counts (n) */
int fvecs_write_txt(const char *fname, int d, int n, const float *vf) {
  FILE *fo = fopen(fname, "w");
  if (!fo) {
    perror("fvecs_write_txt: cannot open file");
    return -1;
  }
  int ret = 0;
  for (int i = 0; i < n; i++) {
    for (int j = 0; j < d; j++) fprintf(fo, "%f ", vf[(long)i * d + j]);
    ret += fprintf(fo, "\n");
  }
  fclose(fo);
  return ret;
}

/* vector.c:1494-1508 */
int ivec_fwrite(FILE *f, const int *v, int d) {
  if (fwrite(&d, sizeof d, 1, f) != 1) {
    perror("ivec_fwrite: write error 1");
    return -1;
  }
  if (fwrite(v, sizeof(int), d, f) != (size_t)d) {
    perror("ivec_fwrite: write error 2");
    return -2;
  }
  return 0;
}

/* vector.c:1513-1520 */
int ivecs_fwrite(FILE *f, int d, int n, const int *v) {
  for (int i = 0; i < n; i++) ivec_fwrite(f, v + (long)i * d, d);
  return n;
}

/* vector.c:1524-1537 */
int ivecs_write(const char *fname, int d, int n, const int *v) {
  FILE *f = fopen(fname, "wb");
  if (!f) {
    perror("ivecs_write");
    return -1;
  }
  const int ret = ivecs_fwrite(f, d, n, v);
  fclose(f);
  return ret;
}

/* vector.c:1279-1295 */
int ivecs_write_txt(const char *fname, int d, int n, const int *v) {
  FILE *fo = fopen(fname, "w");
  if (!fo) {
    perror("ivecs_write_txt: cannot open file");
    return -1;
  }
  int ret = 0;
  for (int i = 0; i < n; i++) {
    for (int j = 0; j < d; j++) fprintf(fo, "%d ", v[(long)i * d + j]);
    ret += fprintf(fo, "\n");
  }
  fclose(fo);
  return ret;
}

/* vector.c:1955-1964 */
long fvec_purge_nans(float *v, long n, float replace_value) {
  long count = 0;
  for (long i = 0; i < n; i++)
    if (isnan(v[i])) {
      v[i] = replace_value;
      count++;
    }
  return count;
}
