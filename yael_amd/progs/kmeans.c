/*
 * kmeans -- command-line k-means on the GPU library, with the options,
 * defaults, stderr report and output files of yael's progs/kmeans.c:
 *   -i file / -itxt file   input vectors (fvecs / text; text needs -n, -d)
 *   -n #, -d #             number of vectors and dimension (fvecs: read
 *                          from the file unless both are given)
 *   -k #, -niter #, -nredo #, -seed #, -nt #
 *   -o file / -otxt file   output centroids (fvecs / text)
 * NaN input values are replaced by 2 (progs/kmeans.c:144-147).  The
 * clustering is kmeans() with flags = nt (progs/kmeans.c:152): the thread
 * bits are accepted and ignored, the run is not quiet.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/yael_mi355x.h"

#define DEFAULT_NITER 40
#define DEFAULT_NREDO 1

static void usage(const char *cmd) {
  printf("Usage: %s [options]\n", cmd);
  printf("  Input\n"
         "    -i filename     input set of vectors to cluster (fvec file format)\n"
         "    -itxt filename  the same but for an input text file\n"
         "    -n #            use n points from the file. Default: all the vectors\n"
         "    -d #            dimension of the vector (used with -itxt only)\n"
         "  Clustering parameters\n"
         "    -k #            number of centroids to produce\n"
         "    -niter #        maximum number of iterations. Default: %d\n"
         "    -nredo #        number of runs to perform. Default: %d\n"
         "    -seed #         initialize the random generate with a number\n"
         "    -nt #           number of threads used. Default: nt=all\n"
         "  Ouput\n"
         "    -o clustfile    output cluster file (fvec format)\n"
         "    -otxt clustfile output cluster file (text format)\n",
         DEFAULT_NITER, DEFAULT_NREDO);
  exit(0);
}

static int intarg(const char *s, const char *opt) {
  int v;
  if (sscanf(s, "%d", &v) != 1) {
    fprintf(stderr, "kmeans: bad value '%s' for %s\n", s, opt);
    exit(1);
  }
  return v;
}

int main(int argc, char **argv) {
  int k = 10, d = 0, n = 0, niter = DEFAULT_NITER, nredo = DEFAULT_NREDO, nt = 1, seed = 0;
  int txt_in = 0, txt_out = 0;
  const char *fi = NULL, *fo = NULL;
  if (argc == 1) usage(argv[0]);
  for (int i = 1; i < argc; i++) {
    const char *a = argv[i];
    const int more = i + 1 < argc;
    if (!strcmp(a, "-h") || !strcmp(a, "--help")) usage(argv[0]);
    if (!strcmp(a, "-k") && more) k = intarg(argv[++i], a);
    else if (!strcmp(a, "-d") && more) d = intarg(argv[++i], a);
    else if (!strcmp(a, "-n") && more) n = intarg(argv[++i], a);
    else if (!strcmp(a, "-niter") && more) niter = intarg(argv[++i], a);
    else if (!strcmp(a, "-nt") && more) nt = intarg(argv[++i], a);
    else if (!strcmp(a, "-nredo") && more) nredo = intarg(argv[++i], a);
    else if (!strcmp(a, "-seed") && more) seed = intarg(argv[++i], a);
    else if (!strcmp(a, "-i") && more) fi = argv[++i], txt_in = 0;
    else if (!strcmp(a, "-itxt") && more) fi = argv[++i], txt_in = 1;
    else if (!strcmp(a, "-o") && more) fo = argv[++i], txt_out = 0;
    else if (!strcmp(a, "-otxt") && more) fo = argv[++i], txt_out = 1;
  }
  if (!fi || !fo) {
    fprintf(stderr, "kmeans: an input (-i/-itxt) and an output (-o/-otxt) are required\n");
    return 1;
  }
  fprintf(stderr, "k = %d\nd = %d\nn = %d\nniter = %d\nnredo = %d\n", k, d, n, niter, nredo);
  fprintf(stderr, "nt = %d\nseed = %d\n", nt, seed);
  fprintf(stderr, "fi = %s  (fmt = %s)\n", fi, txt_in ? "txt" : "fvec");
  fprintf(stderr, "fo = %s  (fmt = %s)\n", fo, txt_out ? "txt" : "fvec");

  int got;
  float *v;
  if (!txt_in) {
    if (d == 0 || n == 0) {
      const long bytes = fvecs_fsize(fi, &d, &n);
      fprintf(stderr, "File %s contains (%ld bytes) %d vectors of dimension %d\n", fi, bytes, n, d);
      if (bytes <= 0) return 1;
    }
    v = (float *)malloc(sizeof(float) * (size_t)n * d);
    got = v ? fvecs_read(fi, d, n, v) : -1;
  } else {
    v = (float *)malloc(sizeof(float) * (size_t)n * d + 1);
    got = v ? fvecs_read_txt(fi, d, n, v) : -1;
  }
  if (got < n || n < 1 || d < 1) {
    fprintf(stderr, "kmeans: read %d of %d vectors from %s\n", got, n, fi);
    return 1;
  }
  const long nnan = fvec_purge_nans(v, (long)n * d, 2);
  if (nnan > 0) fprintf(stderr, "found %ld NaN values\n", nnan);

  float *cent = (float *)malloc(sizeof(float) * (size_t)k * d);
  int *nassign = (int *)malloc(sizeof(int) * (size_t)k);
  float err = 0;
  const int rc = yael_kmeans_checked(d, n, k, niter, v, nt, seed, nredo, cent, NULL, NULL,
                                     nassign, &err);
  if (rc != 0) {
    fprintf(stderr, "kmeans: %s\n", yk_last_error());
    return 1;
  }
  const int wrote = txt_out ? fvecs_write_txt(fo, d, k, cent) : fvecs_write(fo, d, k, cent);
  if (wrote != k) {
    fprintf(stderr, "kmeans: could not write %s\n", fo);
    return 1;
  }
  free(cent);
  free(nassign);
  free(v);
  return 0;
}
