/*
 * knn -- exhaustive k-nearest-neighbour search from the command line, on the
 * GPU library.  Same option names, defaults and output files as the
 * reference's progs/knn.c (its flag list is the interface scripts rely on):
 * base -b/-bb/-bt, queries -q/-qb/-qt (fvecs / bvecs / text), -nb/-nq/-d,
 * -k (10), -nt (accepted, unused), -onn/-onnt (nn.out), -odis/-odist
 * (dis.out), -silence/-verbose.  The search is squared L2 through
 * knn_full_thread followed by knn_reorder_shortlist, the two calls the
 * reference program makes (progs/knn.c:225-226), so the output holds the
 * same neighbours and distances.  NaN components of the inputs count as 2.
 */
#include "cli.h"

static const char *what =
    "Finds, for every query vector, its k nearest base vectors (squared L2)\n"
    "and writes their indices and squared distances, one vector per query.";

int main(int argc, char **argv) {
  const char *prog = "knn";
  int k = 10, d = 0, nb = 0, nq = 0, nt = 0, verbose = 1;
  const char *base = NULL, *queries = NULL, *out_nn = "nn.out", *out_dis = "dis.out";
  cli_form base_form = CLI_FVECS, query_form = CLI_FVECS, nn_form = CLI_IVECS,
           dis_form = CLI_FVECS;

  const cli_opt opts[] = {
      {"Base vectors (searched)", CLI_SECTION, NULL, NULL, 0, NULL},
      {"-b", CLI_PATH, &base, &base_form, CLI_FVECS, "fvecs file"},
      {"-bb", CLI_PATH, &base, &base_form, CLI_BVECS, "bvecs file (bytes, read as floats)"},
      {"-bt", CLI_PATH, &base, &base_form, CLI_TEXT, "text file (needs -nb and -d)"},
      {"-nb", CLI_INT, &nb, NULL, 0, "use the first N base vectors (default: all)"},
      {"Query vectors", CLI_SECTION, NULL, NULL, 0, NULL},
      {"-q", CLI_PATH, &queries, &query_form, CLI_FVECS, "fvecs file"},
      {"-qb", CLI_PATH, &queries, &query_form, CLI_BVECS, "bvecs file (bytes, read as floats)"},
      {"-qt", CLI_PATH, &queries, &query_form, CLI_TEXT, "text file (needs -nq and -d)"},
      {"-nq", CLI_INT, &nq, NULL, 0, "use the first N queries (default: all)"},
      {"-d", CLI_INT, &d, NULL, 0, "dimension (needed for text files)"},
      {"Search", CLI_SECTION, NULL, NULL, 0, NULL},
      {"-k", CLI_INT, &k, NULL, 0, "neighbours per query (default 10)"},
      {"-nt", CLI_INT, &nt, NULL, 0, "host threads (accepted; the GPU does the work)"},
      {"Results", CLI_SECTION, NULL, NULL, 0, NULL},
      {"-onn", CLI_PATH, &out_nn, &nn_form, CLI_IVECS, "neighbour indices, ivecs (nn.out)"},
      {"-onnt", CLI_PATH, &out_nn, &nn_form, CLI_TEXT, "neighbour indices, text"},
      {"-odis", CLI_PATH, &out_dis, &dis_form, CLI_FVECS, "squared distances, fvecs (dis.out)"},
      {"-odist", CLI_PATH, &out_dis, &dis_form, CLI_TEXT, "squared distances, text"},
      {"Messages", CLI_SECTION, NULL, NULL, 0, NULL},
      {"-silence", CLI_SET, &verbose, NULL, 0, "report errors only"},
      {"-verbose", CLI_SET, &verbose, NULL, 2, "report more"},
  };
  const int nopts = (int)(sizeof opts / sizeof opts[0]);

  if (argc < 2) {
    cli_help(stdout, prog, what, opts, nopts);
    return 0;
  }
  const int pr = cli_parse(prog, what, opts, nopts, argc, argv);
  if (pr) return pr > 0 ? 0 : 2;
  if (!base || !queries) {
    fprintf(stderr, "%s: a base file (-b, -bb or -bt) and a query file (-q, -qb or -qt) are "
                    "required\n", prog);
    return 2;
  }
  if (k < 1) {
    fprintf(stderr, "%s: -k must be at least 1\n", prog);
    return 2;
  }

  float *vb = cli_load(prog, base, base_form, &nb, &d, 2.0f, verbose);
  if (!vb) return 1;
  float *vq = cli_load(prog, queries, query_form, &nq, &d, 2.0f, verbose);
  if (!vq) return 1;
  if (k > nb) {
    fprintf(stderr, "%s: k=%d exceeds the %d base vectors\n", prog, k, nb);
    return 2;
  }
  if (verbose)
    fprintf(stderr, "%s: %d queries x %d base vectors, d=%d, k=%d -> %s (%s), %s (%s)\n", prog,
            nq, nb, d, k, out_nn, cli_form_name(nn_form), out_dis, cli_form_name(dis_form));

  int *nn = (int *)malloc(sizeof(int) * (size_t)k * nq);
  float *dis = (float *)malloc(sizeof(float) * (size_t)k * nq);
  if (!nn || !dis) {
    fprintf(stderr, "%s: no memory for %d x %d results\n", prog, nq, k);
    return 1;
  }
  if (yael_knn_full_checked(2, nq, nb, d, k, vb, vq, NULL, nn, dis) != 0 ||
      yael_knn_reorder_shortlist_checked(nq, nb, d, k, vb, vq, nn, dis) != 0) {
    fprintf(stderr, "%s: search failed: %s\n", prog, yk_last_error());
    return 1;
  }
  const int rc = cli_store(prog, out_nn, nn_form, k, nq, nn, 1) |
                 cli_store(prog, out_dis, dis_form, k, nq, dis, 0);
  free(nn);
  free(dis);
  free(vb);
  free(vq);
  return rc ? 1 : 0;
}
