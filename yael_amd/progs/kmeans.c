/*
 * kmeans -- k-means clustering of a vector file from the command line, on
 * the GPU library.  Same option names and defaults as the reference's
 * progs/kmeans.c (its flag list is the interface scripts rely on): input
 * -i (fvecs) or -itxt (text, needs -n and -d), -n, -d, -k (10), -niter (40),
 * -nredo (1), -seed (0), -nt (accepted, unused), output -o (fvecs) or -otxt
 * (text).  The clustering is one kmeans() call with the thread count as
 * its flags, as the reference program makes it (progs/kmeans.c:152), so the
 * centroids written are the same.  NaN components of the input count as 2.
 */
#include "cli.h"

static const char *what =
    "Clusters the input vectors into k centroids with k-means and writes the\n"
    "centroids, one vector each.";

int main(int argc, char **argv) {
  const char *prog = "kmeans";
  int k = 10, d = 0, n = 0, niter = 40, nredo = 1, seed = 0, nt = 0, verbose = 1;
  const char *in = NULL, *out = NULL;
  cli_form in_form = CLI_FVECS, out_form = CLI_FVECS;

  const cli_opt opts[] = {
      {"Input", CLI_SECTION, NULL, NULL, 0, NULL},
      {"-i", CLI_PATH, &in, &in_form, CLI_FVECS, "vectors to cluster, fvecs"},
      {"-itxt", CLI_PATH, &in, &in_form, CLI_TEXT, "vectors to cluster, text (needs -n, -d)"},
      {"-n", CLI_INT, &n, NULL, 0, "use the first N vectors (default: all)"},
      {"-d", CLI_INT, &d, NULL, 0, "dimension (needed for text input)"},
      {"Clustering", CLI_SECTION, NULL, NULL, 0, NULL},
      {"-k", CLI_INT, &k, NULL, 0, "number of centroids (default 10)"},
      {"-niter", CLI_INT, &niter, NULL, 0, "most Lloyd iterations per run (default 40)"},
      {"-nredo", CLI_INT, &nredo, NULL, 0, "runs; the best one is kept (default 1)"},
      {"-seed", CLI_INT, &seed, NULL, 0, "random seed (default 0)"},
      {"-nt", CLI_INT, &nt, NULL, 0, "host threads (accepted; the GPU does the work)"},
      {"Output", CLI_SECTION, NULL, NULL, 0, NULL},
      {"-o", CLI_PATH, &out, &out_form, CLI_FVECS, "centroids, fvecs"},
      {"-otxt", CLI_PATH, &out, &out_form, CLI_TEXT, "centroids, text"},
  };
  const int nopts = (int)(sizeof opts / sizeof opts[0]);

  if (argc < 2) {
    cli_help(stdout, prog, what, opts, nopts);
    return 0;
  }
  const int pr = cli_parse(prog, what, opts, nopts, argc, argv);
  if (pr) return pr > 0 ? 0 : 2;
  if (!in || !out) {
    fprintf(stderr, "%s: an input file (-i or -itxt) and an output file (-o or -otxt) are "
                    "required\n", prog);
    return 2;
  }
  if (k < 1 || nredo < 1 || niter < 0 || nt < 0) {
    fprintf(stderr, "%s: need k >= 1, nredo >= 1, niter >= 0, nt >= 0\n", prog);
    return 2;
  }

  float *v = cli_load(prog, in, in_form, &n, &d, 2.0f, verbose);
  if (!v) return 1;
  if (k > n) {
    fprintf(stderr, "%s: k=%d exceeds the %d input vectors\n", prog, k, n);
    return 2;
  }
  fprintf(stderr, "%s: %d vectors, d=%d, k=%d, niter=%d, nredo=%d, seed=%d -> %s (%s)\n", prog,
          n, d, k, niter, nredo, seed, out, cli_form_name(out_form));

  float *cent = (float *)malloc(sizeof(float) * (size_t)k * d);
  int *sizes = (int *)malloc(sizeof(int) * (size_t)k);
  if (!cent || !sizes) {
    fprintf(stderr, "%s: no memory for %d centroids\n", prog, k);
    return 1;
  }
  float mean_err = 0;
  if (yael_kmeans_checked(d, n, k, niter, v, nt & 0xffff, seed, nredo, cent, NULL, NULL, sizes,
                          &mean_err) != 0) {
    fprintf(stderr, "%s: clustering failed: %s\n", prog, yk_last_error());
    return 1;
  }
  fprintf(stderr, "%s: mean squared distance to the nearest centroid %g\n", prog, mean_err);
  const int rc = cli_store(prog, out, out_form, d, k, cent, 0);
  free(cent);
  free(sizes);
  free(v);
  return rc ? 1 : 0;
}
