/*
 * bbscite_quality: the solution-quality harness of the reference
 * (verification/__main__.py:54-163; docs/thesis/chapters/evaluation/
 * quality.tex:18-22) over the C ABI.
 *
 *   bbscite_quality score -i input.csv -t tree.newick [-fd A] [-ad B] [-compact]
 *       Log-likelihood of a tree with every cell at its most likely node,
 *       in double precision (verification/__main__.py:54-87).
 *
 *   bbscite_quality test [-inputs K] [-genes G] [-cells C] [-chains R]
 *                        [-steps S] [-fd A] [-ad B] [-missing P] [-seed X]
 *                        [-confidence Q]
 *       K synthetic inputs (the generator of bbs_generate_dataset, seeds
 *       X .. X+K-1) searched by every form of the chain search: the plain
 *       launch with full and with incremental scoring, one batched run of
 *       all inputs, and the continuous-batching server with all inputs
 *       submitted at once. Each form's best tree is scored as above, and so
 *       is the tree that generated the data. Two one-sided t-tests
 *       (verification/__main__.py:90-120) check every form against the
 *       plain launch with the margin of one bit flip,
 *       max(|log A - log(1-B)|, |log B - log(1-A)|). Exit status 0 when
 *       every form is equivalent at the given confidence.
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "bbscite.h"

namespace {

struct Input {
  uint32_t genes = 0, cells = 0;
  std::vector<uint64_t> mut, kno;
  std::vector<uint32_t> true_parent;
};

bool read_file(const std::string &path, std::string &out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  std::stringstream ss;
  ss << f.rdbuf();
  out = ss.str();
  return true;
}

bool load_matrix(const std::string &path, Input &in) {
  std::string text;
  if (!read_file(path, text)) {
    std::cerr << "Error: could not open " << path << std::endl;
    return false;
  }
  std::vector<int8_t> m(BBS_MAX_GENES * BBS_MAX_CELLS);
  if (bbs_parse_csv(text.data(), text.size(), m.data(), &in.genes, &in.cells) != BBS_OK) {
    std::cerr << "Error: " << bbs_last_error(nullptr) << std::endl;
    return false;
  }
  in.mut.assign(2u * in.cells, 0u);
  in.kno.assign(2u * in.cells, 0u);
  bbs_encode_matrix(m.data(), in.genes, in.cells, in.mut.data(), in.kno.data());
  return true;
}

double score_rows(const Input &in, const uint64_t *dm, double a, double b) {
  double ll = 0.0;
  if (bbs_score_attachments(dm, in.genes + 1u, in.mut.data(), in.kno.data(), in.cells, a, b,
                            nullptr, &ll) != BBS_OK) {
    std::cerr << "Error: " << bbs_last_error(nullptr) << std::endl;
    std::exit(2);
  }
  return ll;
}

int cmd_score(int argc, char **argv) {
  std::string in_path, tree_path;
  double a = 6.04e-5, b = 0.4309;
  bool compact = false;
  for (int i = 2; i < argc; i++) {
    std::string f = argv[i];
    if (f == "-i" && i + 1 < argc) in_path = argv[++i];
    else if (f == "-t" && i + 1 < argc) tree_path = argv[++i];
    else if (f == "-fd" && i + 1 < argc) a = std::atof(argv[++i]);
    else if (f == "-ad" && i + 1 < argc) b = std::atof(argv[++i]);
    else if (f == "-compact") compact = true;
    else {
      std::cerr << "Error: unknown or incomplete argument " << f << std::endl;
      return 2;
    }
  }
  if (in_path.empty() || tree_path.empty()) {
    std::cerr << "Error: score needs -i <matrix.csv> and -t <tree.newick>" << std::endl;
    return 2;
  }
  Input in;
  if (!load_matrix(in_path, in)) return 2;
  std::string newick;
  if (!read_file(tree_path, newick)) {
    std::cerr << "Error: could not open " << tree_path << std::endl;
    return 2;
  }
  std::vector<uint32_t> parent(BBS_MAX_NODES);
  uint32_t n = 0;
  if (bbs_newick_to_parent_vector(newick.c_str(), parent.data(), &n) != BBS_OK) {
    std::cerr << "Error: " << bbs_last_error(nullptr) << std::endl;
    return 2;
  }
  if (n != in.genes + 1u) {
    std::cerr << "Error: the tree has " << n << " nodes, the matrix " << in.genes
              << " genes" << std::endl;
    return 2;
  }
  std::vector<uint64_t> am(2u * n), dm(2u * n);
  bbs_parent_vector_to_matrix(parent.data(), n, am.data(), dm.data());
  std::vector<uint32_t> att(in.cells);
  double ll = 0.0;
  bbs_score_attachments(dm.data(), n, in.mut.data(), in.kno.data(), in.cells, a, b, att.data(),
                        &ll);
  if (compact) {
    std::printf("%.17g\n", ll);
  } else {
    std::printf("Mutation tree log-likelihood: %.17g\n", ll);
    std::printf("Attachments:");
    for (uint32_t c = 0; c < in.cells; c++) std::printf(" %u", att[c]);
    std::printf("\n");
  }
  return 0;
}

void report(const char *name, const std::vector<double> &diff, double delta,
            double confidence, bool &all_ok) {
  double mean = 0.0, pl = 1.0, pu = 1.0;
  if (bbs_tost(diff.data(), (uint32_t)diff.size(), delta, &mean, &pl, &pu) != BBS_OK) {
    std::cerr << "Error: " << bbs_last_error(nullptr) << std::endl;
    std::exit(2);
  }
  const double bound = (1.0 - confidence) / 2.0;
  const bool ok = pl < bound && pu < bound;
  all_ok = all_ok && ok;
  int better = 0, equal = 0, worse = 0;
  double lo = 0.0, hi = 0.0;
  for (double d : diff) {
    better += d > 0.0;
    equal += d == 0.0;
    worse += d < 0.0;
    lo = std::min(lo, d);
    hi = std::max(hi, d);
  }
  std::printf("%-28s mean %+.4f (%+.3f bit flips), p_lower %.3g, p_upper %.3g: %s\n", name,
              mean, mean / delta, pl, pu, ok ? "equivalent" : "NOT shown equivalent");
  std::printf("%-28s better %d, equal %d, worse %d, range [%+.3f, %+.3f] bit flips\n", "",
              better, equal, worse, lo / delta, hi / delta);
}

int cmd_test(int argc, char **argv) {
  uint32_t K = 16, G = 63, C = 64, R = 6, S = 100000;
  double a = 1e-6, b = 0.25, missing = 0.25, confidence = 0.98;
  uint64_t seed = 1;
  for (int i = 2; i < argc; i++) {
    std::string f = argv[i];
    auto next = [&]() -> const char * {
      if (i + 1 >= argc) {
        std::cerr << "Error: missing value for " << f << std::endl;
        std::exit(2);
      }
      return argv[++i];
    };
    if (f == "-inputs") K = (uint32_t)std::atoi(next());
    else if (f == "-genes") G = (uint32_t)std::atoi(next());
    else if (f == "-cells") C = (uint32_t)std::atoi(next());
    else if (f == "-chains") R = (uint32_t)std::atoi(next());
    else if (f == "-steps") S = (uint32_t)std::atoi(next());
    else if (f == "-fd") a = std::atof(next());
    else if (f == "-ad") b = std::atof(next());
    else if (f == "-missing") missing = std::atof(next());
    else if (f == "-seed") seed = std::strtoull(next(), nullptr, 10);
    else if (f == "-confidence") confidence = std::atof(next());
    else {
      std::cerr << "Error: unknown argument " << f << std::endl;
      return 2;
    }
  }
  if (K < 2u) {
    std::cerr << "Error: -inputs must be at least 2" << std::endl;
    return 2;
  }
  const uint32_t N = G + 1u;
  std::vector<Input> inputs(K);
  std::vector<bbs_params> params(K);
  std::vector<bbs_job> jobs(K);
  for (uint32_t k = 0; k < K; k++) {
    Input &in = inputs[k];
    in.genes = G;
    in.cells = C;
    std::vector<int8_t> m((size_t)G * C);
    in.true_parent.assign(N, 0u);
    std::vector<uint32_t> att(C);
    if (bbs_generate_dataset(seed + k, G, C, a, b, missing, m.data(), in.true_parent.data(),
                             att.data()) != BBS_OK) {
      std::cerr << "Error: " << bbs_last_error(nullptr) << std::endl;
      return 2;
    }
    in.mut.assign(2u * C, 0u);
    in.kno.assign(2u * C, 0u);
    bbs_encode_matrix(m.data(), G, C, in.mut.data(), in.kno.data());
    bbs_params &p = params[k];
    bbs_params_default(&p);
    p.alpha_mean = (float)a;
    p.beta_mean = (float)b;
    p.n_chains = R;
    p.n_steps = S;
    p.seed = seed + k;
    jobs[k] = bbs_job{p, C, G, in.mut.data(), in.kno.data(), 0u};
  }
  auto fail = [](const char *what, const char *msg) {
    std::cerr << "Error: " << what << ": " << msg << std::endl;
    std::exit(2);
  };
  std::vector<uint64_t> am(2u * N), dm(2u * N);
  std::vector<double> ll_true(K), ll_plain(K), ll_full(K), ll_batch(K), ll_serve(K);
  for (uint32_t k = 0; k < K; k++) {
    std::vector<uint64_t> tam(2u * N), tdm(2u * N);
    bbs_parent_vector_to_matrix(inputs[k].true_parent.data(), N, tam.data(), tdm.data());
    ll_true[k] = score_rows(inputs[k], tdm.data(), a, b);
    for (int mode = 0; mode < 2; mode++) {
      bbs_ctx *ctx = nullptr;
      if (bbs_create(&params[k], C, G, inputs[k].mut.data(), inputs[k].kno.data(), &ctx) !=
          BBS_OK)
        fail("bbs_create", bbs_last_error(nullptr));
      if (mode == 1 && bbs_set_scoring_mode(ctx, BBS_SCORING_FULL) != BBS_OK)
        fail("bbs_set_scoring_mode", bbs_last_error(ctx));
      float ms = 0.f, beta = 0.f, score = 0.f;
      if (bbs_run(ctx, &ms) != BBS_OK) fail("bbs_run", bbs_last_error(ctx));
      if (bbs_get_best(ctx, am.data(), dm.data(), &beta, &score) != BBS_OK)
        fail("bbs_get_best", bbs_last_error(ctx));
      bbs_destroy(ctx);
      (mode == 0 ? ll_plain : ll_full)[k] = score_rows(inputs[k], dm.data(), a, b);
    }
  }
  {
    bbs_batch *bt = nullptr;
    if (bbs_batch_create(jobs.data(), K, nullptr, &bt) != BBS_OK)
      fail("bbs_batch_create", bbs_batch_last_error(nullptr));
    float ms = 0.f;
    if (bbs_batch_run(bt, &ms) != BBS_OK) fail("bbs_batch_run", bbs_batch_last_error(bt));
    for (uint32_t k = 0; k < K; k++) {
      float beta = 0.f, score = 0.f;
      if (bbs_batch_get_best(bt, k, am.data(), dm.data(), &beta, &score) != BBS_OK)
        fail("bbs_batch_get_best", bbs_batch_last_error(bt));
      ll_batch[k] = score_rows(inputs[k], dm.data(), a, b);
    }
    bbs_batch_destroy(bt);
  }
  {
    bbs_server *srv = nullptr;
    if (bbs_server_create(nullptr, &srv) != BBS_OK)
      fail("bbs_server_create", bbs_server_last_error(nullptr));
    std::vector<uint64_t> ids(K);
    for (uint32_t k = 0; k < K; k++)
      if (bbs_server_submit(srv, &jobs[k], &ids[k]) != BBS_OK)
        fail("bbs_server_submit", bbs_server_last_error(srv));
    uint32_t got = 0;
    std::vector<bbs_job_result> res(K);
    while (got < K) {
      if (bbs_server_step(srv, nullptr, nullptr, nullptr) != BBS_OK)
        fail("bbs_server_step", bbs_server_last_error(srv));
      uint32_t n = 0;
      bbs_server_poll(srv, res.data() + got, K - got, &n);
      got += n;
    }
    for (const bbs_job_result &r : res) {
      const uint32_t k = (uint32_t)(std::find(ids.begin(), ids.end(), r.id) - ids.begin());
      ll_serve[k] = score_rows(inputs[k], &r.best_dm[0][0], a, b);
    }
    bbs_server_destroy(srv);
  }
  const double delta =
      std::max(std::fabs(std::log(a) - std::log(1.0 - b)), std::fabs(std::log(b) - std::log(1.0 - a)));
  std::printf("Quality test: %u inputs of %u genes x %u cells, %u chains x %u steps, "
              "alpha %g, beta %g, missing %g; one bit flip = %.4f\n",
              K, G, C, R, S, a, b, missing, delta);
  bool ok = true;
  std::vector<double> d(K);
  for (uint32_t k = 0; k < K; k++) d[k] = ll_full[k] - ll_plain[k];
  report("full vs incremental scoring", d, delta, confidence, ok);
  for (uint32_t k = 0; k < K; k++) d[k] = ll_batch[k] - ll_plain[k];
  report("batched vs plain", d, delta, confidence, ok);
  for (uint32_t k = 0; k < K; k++) d[k] = ll_serve[k] - ll_plain[k];
  report("served vs plain", d, delta, confidence, ok);
  bool unused = true;
  for (uint32_t k = 0; k < K; k++) d[k] = ll_plain[k] - ll_true[k];
  report("plain vs generating tree", d, delta, confidence, unused);
  std::printf("%s\n", ok ? "Every form is equivalent to the plain launch."
                         : "Some form is NOT shown equivalent to the plain launch.");
  return ok ? 0 : 1;
}

} // namespace

int main(int argc, char **argv) {
  if (argc >= 2 && std::strcmp(argv[1], "score") == 0) return cmd_score(argc, argv);
  if (argc >= 2 && std::strcmp(argv[1], "test") == 0) return cmd_test(argc, argv);
  std::cerr << "usage: bbscite_quality score -i input.csv -t tree.newick [-fd A] [-ad B] "
               "[-compact]\n       bbscite_quality test [-inputs K] [-genes G] [-cells C] "
               "[-chains R] [-steps S] [-fd A] [-ad B] [-missing P] [-seed X] "
               "[-confidence Q]"
            << std::endl;
  return 2;
}
