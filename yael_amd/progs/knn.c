/*
 * knn -- command-line exhaustive k-NN search on the GPU library, with the
 * options, defaults and output files of yael's progs/knn.c:
 *   -b / -bb / -bt file   base vectors (fvecs / bvecs / text)
 *   -q / -qb / -qt file   query vectors (fvecs / bvecs / text)
 *   -nb #, -nq #, -d #    counts and dimension (text input needs them)
 *   -k # (default 10), -nt #
 *   -onn / -onnt file     k-NN indices, one vector per query (ivecs / text;
 *                         default nn.out)
 *   -odis / -odist file   their squared distances (fvecs / text; default
 *                         dis.out)
 *   -silence / -verbose
 * L2 search with knn_full_thread, then knn_reorder_shortlist, as
 * progs/knn.c:225-226 does; NaN inputs are replaced by 2.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/yael_mi355x.h"

enum { FMT_FVEC, FMT_IVEC, FMT_BVEC, FMT_TEXT };

static void usage(const char *cmd) {
  printf("Usage: %s [options]\n", cmd);
  printf("  Input\n"
         "    -b filename    database to be searched (fvec file format)\n"
         "    -bb filename   database to be searched (bvec format)\n"
         "    -bt filename   database to be searched (text format)\n"
         "    -q filename    query vectors (file in fvec format)\n"
         "    -qb filename   query vectors (file in bvec format)\n"
         "    -qt filename   query vectors (text format)\n"
         "    -nb #          use the first n database vectors (default: all)\n"
         "    -nq #          use the first n query vectors (default: all)\n"
         "    -d #           dimension of the vector (used with -itxt only)\n"
         "  Search parameters\n"
         "    -k #           number of nearest neighbors to retrieve (default: 10)\n"
         "    -nt #          number of threads used. Default: nt=all\n"
         "  Ouput\n"
         "    -onn idxfile   file (fmt=ivec) of k-NN indexes, 1 vec/query (default: nn.out)\n"
         "    -onnt idxfile  same as onn but in text format\n"
         "    -odis disfile  square distances to k-NN, 1 vec/query (default: dis.out)\n"
         "    -odist disfile same as dis but output in text format\n"
         "  Others:\n"
         "    -verbose       better to look at this than chatting on Internet\n"
         "    -silent        shut up!\n\n"
         "Remark: in text format, all the parameters nb, nq and d must be defined\n");
  exit(0);
}

static int intarg(const char *s, const char *opt) {
  int v;
  if (sscanf(s, "%d", &v) != 1) {
    fprintf(stderr, "knn: bad value '%s' for %s\n", s, opt);
    exit(1);
  }
  return v;
}

static const char *fmt_name(int f) {
  return f == FMT_FVEC ? "fvec" : f == FMT_BVEC ? "bvec" : f == FMT_IVEC ? "ivec" : "txt";
}

/* progs/knn.c:48-100: a vector file in one of the input formats, NaNs
 * replaced by 2; *n and *d are read from the file when either is 0 */
static float *read_set(const char *fname, int fmt, int verbose, int *nptr, int *dptr) {
  int n = *nptr, d = *dptr, got = -1;
  float *v = NULL;
  if (fmt == FMT_TEXT) {
    if (n == 0 || d == 0) {
      fprintf(stderr, "knn: text input %s needs -d and the vector count\n", fname);
      exit(1);
    }
    v = (float *)malloc(sizeof(float) * (size_t)n * d + 1);
    got = fvecs_read_txt(fname, d, n, v);
  } else {
    if (d == 0 || n == 0) {
      const long bytes = fmt == FMT_FVEC ? fvecs_fsize(fname, &d, &n) : bvecs_fsize(fname, &d, &n);
      fprintf(stderr, "File %s contains (%ld bytes) %d vectors of dimension %d\n", fname, bytes,
              n, d);
      if (bytes <= 0) exit(1);
    }
    v = (float *)malloc(sizeof(float) * (size_t)n * d + 1);
    got = fmt == FMT_FVEC ? fvecs_read(fname, d, n, v) : b2fvecs_read(fname, d, n, v);
  }
  if (got < n || n < 1 || d < 1) {
    fprintf(stderr, "knn: read %d of %d vectors from %s\n", got, n, fname);
    exit(1);
  }
  const long nnan = fvec_purge_nans(v, (long)n * d, 2);
  if (nnan > 0 && verbose >= 1)
    fprintf(stderr, "found %ld NaN values in what read from %s\n", nnan, fname);
  *nptr = n;
  *dptr = d;
  return v;
}

int main(int argc, char **argv) {
  int k = 10, d = 0, nb = 0, nq = 0, nt = 1, verbose = 1;
  int fmt_b = FMT_FVEC, fmt_q = FMT_FVEC, fmt_nn = FMT_IVEC, fmt_dis = FMT_FVEC;
  const char *fb = NULL, *fq = NULL, *fnn = "nn.out", *fdis = "dis.out";
  if (argc == 1) usage(argv[0]);
  for (int i = 1; i < argc; i++) {
    const char *a = argv[i];
    const int more = i + 1 < argc;
    if (!strcmp(a, "-h") || !strcmp(a, "--help")) usage(argv[0]);
    else if (!strcmp(a, "-silence")) verbose = 0;
    else if (!strcmp(a, "-verbose")) verbose = 2;
    else if (!strcmp(a, "-k") && more) k = intarg(argv[++i], a);
    else if (!strcmp(a, "-d") && more) d = intarg(argv[++i], a);
    else if (!strcmp(a, "-nt") && more) nt = intarg(argv[++i], a);
    else if (!strcmp(a, "-nb") && more) nb = intarg(argv[++i], a);
    else if (!strcmp(a, "-nq") && more) nq = intarg(argv[++i], a);
    else if (!strcmp(a, "-b") && more) fb = argv[++i], fmt_b = FMT_FVEC;
    else if (!strcmp(a, "-bb") && more) fb = argv[++i], fmt_b = FMT_BVEC;
    else if (!strcmp(a, "-bt") && more) fb = argv[++i], fmt_b = FMT_TEXT;
    else if (!strcmp(a, "-q") && more) fq = argv[++i], fmt_q = FMT_FVEC;
    else if (!strcmp(a, "-qb") && more) fq = argv[++i], fmt_q = FMT_BVEC;
    else if (!strcmp(a, "-qt") && more) fq = argv[++i], fmt_q = FMT_TEXT;
    else if (!strcmp(a, "-onn") && more) fnn = argv[++i], fmt_nn = FMT_IVEC;
    else if (!strcmp(a, "-onnt") && more) fnn = argv[++i], fmt_nn = FMT_TEXT;
    else if (!strcmp(a, "-odis") && more) fdis = argv[++i], fmt_dis = FMT_FVEC;
    else if (!strcmp(a, "-odist") && more) fdis = argv[++i], fmt_dis = FMT_TEXT;
  }
  if (!fb || !fq) {
    fprintf(stderr, "knn: a base (-b/-bb/-bt) and queries (-q/-qb/-qt) are required\n");
    return 1;
  }
  fprintf(stderr, "k = %d\nd = %d\nnt = %d\n", k, d, nt);
  if (verbose) {
    fprintf(stderr, "fb = %s  (fmt = %s)\n", fb, fmt_name(fmt_b));
    fprintf(stderr, "fq = %s  (fmt = %s)\n", fq, fmt_name(fmt_q));
    fprintf(stderr, "fnn = %s  (fmt = %s)\n", fnn, fmt_name(fmt_nn));
    fprintf(stderr, "fdis = %s  (fmt = %s)\n", fdis, fmt_name(fmt_dis));
  }
  float *vb = read_set(fb, fmt_b, verbose, &nb, &d);
  float *vq = read_set(fq, fmt_q, verbose, &nq, &d);

  int *idx = (int *)malloc(sizeof(int) * (size_t)k * nq);
  float *dis = (float *)malloc(sizeof(float) * (size_t)k * nq);
  if (yael_knn_full_checked(2, nq, nb, d, k, vb, vq, NULL, idx, dis) != 0 ||
      yael_knn_reorder_shortlist_checked(nq, nb, d, k, vb, vq, idx, dis) != 0) {
    fprintf(stderr, "knn: %s\n", yk_last_error());
    return 1;
  }
  int wrote = fmt_dis == FMT_FVEC ? fvecs_write(fdis, k, nq, dis)
                                  : fvecs_write_txt(fdis, k, nq, dis);
  if (wrote != nq) {
    fprintf(stderr, "knn: could not write %s\n", fdis);
    return 1;
  }
  wrote = fmt_nn == FMT_IVEC ? ivecs_write(fnn, k, nq, idx) : ivecs_write_txt(fnn, k, nq, idx);
  if (wrote != nq) {
    fprintf(stderr, "knn: could not write %s\n", fnn);
    return 1;
  }
  free(idx);
  free(dis);
  free(vb);
  free(vq);
  return 0;
}
