/*
 * bbscite: SCITE-compatible command line over the MI355X chain search.
 *
 * Drop-in for the reference's bbSCITE executable (src/main.cpp:42-215,
 * flags of src/Parameters.hpp:61-281): same flags, same CSV input, same
 * stdout lines and the same five output files
 * (<base>_ml0.gv, _ml0.newick, _ml0_beta.txt, _ml0_am.txt, _ml0_dm.txt).
 * Differences: -seed is honoured (the reference ignores it), the size limit
 * is 127 genes x 128 cells, and moves are drawn on the device (the
 * "Move generation makespan" line reports 0 s).
 */
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <optional>
#include <sstream>
#include <string>
#include <vector>

#include "bbscite.h"

namespace {

struct Cli {
  std::optional<std::string> input_path, output_path_base;
  std::optional<uint32_t> n_chains, chain_length;
  bbs_params p;
};

/* src/Parameters.hpp:61-281 */
bool load_and_verify_args(int argc, char **argv, Cli &cli) {
  bool error = false;
  bbs_params_default(&cli.p);
  auto missing = [&](const char *flag) {
    std::cerr << "Error: Missing argument to parameter " << flag << "." << std::endl;
    error = true;
  };
  for (int i = 1; i < argc; ++i) {
    if (argv[i][0] != '-') continue;
    std::string a = argv[i];
    if (a == "-i") {
      if (i + 1 < argc) cli.input_path = argv[++i]; else missing("-i");
    } else if (a == "-t") {
      std::cerr << "Warning: The -t parameter is not supported by bbSCITE "
                   "(yet). Ignoring the -t parameter."
                << std::endl;
    } else if (a == "-o") {
      if (i + 1 < argc) cli.output_path_base = argv[++i]; else missing("-o");
    } else if (a == "-n") {
      std::cerr << "Warning: The number of genes is inferred from the input "
                   "file. Ignoring parameter -n."
                << std::endl;
    } else if (a == "-m") {
      std::cerr << "Warning: The number of cells is inferred from the input "
                   "file. Ignoring parameter -m."
                << std::endl;
    } else if (a == "-r") {
      if (i + 1 < argc) cli.n_chains = (uint32_t)atoi(argv[++i]); else missing("-r");
    } else if (a == "-l") {
      if (i + 1 < argc) cli.chain_length = (uint32_t)atoi(argv[++i]); else missing("-l");
    } else if (a == "-g") {
      if (i + 1 < argc) cli.p.gamma = (float)atof(argv[++i]); else missing("-g");
    } else if (a == "-fd") {
      if (i + 1 < argc) cli.p.alpha_mean = (float)atof(argv[++i]); else missing("-fd");
    } else if (a == "-ad") {
      if (i + 1 < argc) cli.p.beta_mean = (float)atof(argv[++i]); else missing("-ad");
      if (i + 1 < argc && argv[i + 1][0] != '-') {
        cli.p.beta_mean = (float)((double)cli.p.beta_mean + atof(argv[++i]));
        std::cerr << "Warning: homo- and heterocygous mutations are not "
                     "supported by bbSCITE (yet). Summing false negative "
                     "probabilites."
                  << std::endl;
      }
    } else if (a == "-cc") {
      std::cerr << "Warning: homo- and heterocygous mutations are not supported "
                   "by bbSCITE (yet). -cc argument ignored."
                << std::endl;
    } else if (a == "-e") {
      if (i + 1 < argc) {
        cli.p.prob_beta_change = (float)atof(argv[++i]);
        if (cli.p.prob_beta_change < 0.0f) {
          std::cerr << "Error: Move probabilities may not be negative." << std::endl;
          error = true;
        }
      } else {
        missing("-e");
      }
    } else if (a == "-x") {
      if (i + 1 < argc) cli.p.beta_jump_scaling_chi = (float)atof(argv[++i]); else missing("-x");
    } else if (a == "-sd") {
      if (i + 1 < argc) cli.p.beta_sd = (float)atof(argv[++i]); else missing("-sd");
    } else if (a == "-a") {
      std::cerr << "Warning: The -a parameter is not supported by bbSCITE "
                   "(yet). Ignoring the -a parameter."
                << std::endl;
    } else if (a == "-p") {
      std::cerr << "Warning: The -p parameter is not supported by bbSCITE "
                   "(yet). Ignoring the -p parameter."
                << std::endl;
    } else if (a == "-names") {
      std::cerr << "The -names parameter is not supported by bbSCITE (yet). "
                   "Ignoring the -names parameter."
                << std::endl;
    } else if (a == "-move_probs") {
      float *dst[3] = {&cli.p.prob_prune_n_reattach, &cli.p.prob_swap_nodes,
                       &cli.p.prob_swap_subtrees};
      for (int k = 0; k < 3; k++) {
        if (i + 1 < argc) {
          *dst[k] = (float)atof(argv[++i]);
          if (*dst[k] < 0.0f) {
            std::cerr << "Error: Move probabilities may not be negative." << std::endl;
            error = true;
          }
        } else {
          missing("-move_probs");
        }
      }
    } else if (a == "-seed") {
      if (i + 1 < argc) cli.p.seed = strtoull(argv[++i], nullptr, 10); else missing("-seed");
    } else if (a == "-max_treelist_size") {
      if (i + 1 < argc) ++i; else missing("-max_treelist_size");
    } else if (a == "-no_tree_list") {
      std::cerr << "Warning: The -no_tree_list parameter is not supported by "
                   "bbSCITE (yet). Ignoring the -no_tree_list parameter."
                << std::endl;
    } else if (a == "-s") {
      std::cerr << "Warning: The -s parameter is not supported by bbSCITE "
                   "(yet). Ignoring the -s parameter."
                << std::endl;
    } else if (a == "-transpose") {
      std::cerr << "Warning: The -transpose parameter is not supported by "
                   "bbSCITE (yet). Ignoring the -transpose parameter."
                << std::endl;
    } else {
      std::cerr << "Error: Unknown parameter " << a << "." << std::endl;
      error = true;
    }
  }
  if (!cli.input_path) {
    std::cerr << "Error: Missing input file path. Did you forget to set the -i "
                 "paramater?"
              << std::endl;
    error = true;
  }
  if (!cli.n_chains) {
    std::cerr << "Error: Missing number of markov chains to simulate. Did you "
                 "forget to set the -r parameter?"
              << std::endl;
    error = true;
  } else if (*cli.n_chains == 0u) {
    /* the reference would index the best of zero chains */
    std::cerr << "Error: The number of markov chains (-r) must be at least 1." << std::endl;
    error = true;
  }
  if (!cli.chain_length) {
    std::cerr << "Error: Missing markov chain length. Did you forget to set the "
                 "-l parameter?"
              << std::endl;
    error = true;
  }
  if (cli.p.prob_beta_change + cli.p.prob_prune_n_reattach + cli.p.prob_swap_nodes +
          cli.p.prob_swap_subtrees ==
      0.0f) {
    std::cerr << "Error: The sum of the move type probabilities is zero." << std::endl;
    error = true;
  }
  if (!error) {
    cli.p.n_chains = *cli.n_chains;
    cli.p.n_steps = *cli.chain_length;
  }
  return !error;
}

/* src/Parameters.hpp:306-313 */
std::string output_base(const Cli &cli) {
  if (cli.output_path_base) return *cli.output_path_base;
  const std::string &in = *cli.input_path;
  size_t dot = in.find_last_of('.');
  return in.substr(0, dot);
}

bool write_file(const std::string &path, const std::string &text) {
  std::ofstream f(path);
  f << text;
  return bool(f);
}

} // namespace

int main(int argc, char **argv) {
  Cli cli;
  if (!load_and_verify_args(argc, argv, cli)) {
    std::cerr << "Quitting due to CLI argument errors." << std::endl;
    return 1;
  }
  std::ifstream in(*cli.input_path, std::ios::binary);
  if (!in) {
    std::cerr << "Error: Could not open " << *cli.input_path << std::endl;
    return 1;
  }
  std::stringstream ss;
  ss << in.rdbuf();
  const std::string text = ss.str();

  std::vector<int8_t> matrix(BBS_MAX_GENES * BBS_MAX_CELLS);
  uint32_t n_genes = 0, n_cells = 0;
  if (bbs_parse_csv(text.data(), text.size(), matrix.data(), &n_genes, &n_cells) != BBS_OK) {
    std::cerr << "Error: " << bbs_last_error(nullptr) << std::endl;
    return 1;
  }
  std::vector<uint64_t> is_mutated(2u * n_cells), is_known(2u * n_cells);
  if (bbs_encode_matrix(matrix.data(), n_genes, n_cells, is_mutated.data(),
                        is_known.data()) != BBS_OK) {
    std::cerr << "Error: " << bbs_last_error(nullptr) << std::endl;
    return 1;
  }

  bbs_ctx *ctx = nullptr;
  if (bbs_create(&cli.p, n_cells, n_genes, is_mutated.data(), is_known.data(), &ctx) !=
      BBS_OK) {
    std::cerr << "Error: " << bbs_last_error(nullptr) << std::endl;
    return 1;
  }
  std::cout << "Move generation makespan: 0 s" << std::endl;
  float ms = 0.f;
  if (bbs_run(ctx, &ms) != BBS_OK) {
    std::cerr << "Error: " << bbs_last_error(ctx) << std::endl;
    bbs_destroy(ctx);
    return 1;
  }
  std::cout << "Time elapsed: " << ms << " ms" << std::endl;

  const uint32_t n = n_genes + 1u;
  std::vector<uint64_t> am(2u * n), dm(2u * n);
  float beta = 0.f, score = 0.f;
  if (bbs_get_best(ctx, am.data(), dm.data(), &beta, &score) != BBS_OK) {
    std::cerr << "Error: " << bbs_last_error(ctx) << std::endl;
    bbs_destroy(ctx);
    return 1;
  }
  bbs_destroy(ctx);

  const std::string base = output_base(cli);
  std::vector<char> buf(64u * n + 256u);
  bbs_tree_to_graphviz(am.data(), dm.data(), n, buf.data(), buf.size());
  write_file(base + "_ml0.gv", buf.data());
  bbs_tree_to_newick(am.data(), dm.data(), n, buf.data(), buf.size());
  write_file(base + "_ml0.newick", buf.data());
  {
    std::ofstream f(base + "_ml0_beta.txt");
    f << beta << std::endl;
  }
  std::string am_txt, dm_txt;
  for (uint32_t i = 0; i < n; i++) {
    for (uint32_t j = 0; j < n; j++) {
      am_txt += ((am[2u * i + (j >> 6)] >> (j & 63u)) & 1u) ? '1' : '0';
      dm_txt += ((dm[2u * i + (j >> 6)] >> (j & 63u)) & 1u) ? '1' : '0';
    }
    am_txt += '\n';
    dm_txt += '\n';
  }
  write_file(base + "_ml0_am.txt", am_txt);
  write_file(base + "_ml0_dm.txt", dm_txt);
  return 0;
}
