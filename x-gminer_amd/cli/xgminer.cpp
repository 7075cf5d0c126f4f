// xgminer: command-line front end over libxgm with the reference's option
// names (xgminer/include/cmd_option.h:30-160) and dispatch
// (xgminer/main.cpp:20-58, xgminer/include/glumin.h:34-80).
//
//   xgminer -g <graph> [--algorithm A] [--patternID N | --pattern-name P |
//           --pattern-size S --pattern-adj-mat M] [--repeat R] [--device D]
//           [--strategy S]
//
// <graph> is detected as in xgminer/src/common.cpp:66-103: a directory
// holding the binary CSR (<dir>/graph.meta.txt, .vertex.bin, .edge.bin), a
// prefix of those files, or an edge-list text file (".txt", or no extension
// with an "n m" first line).  Counts print as the reference prints them
// ("Pattern P<id> count: <n>", xgminer/src/glumin.cu:22,44).  DAG patterns
// are oriented on the device after the undirected upload.
#include <sys/stat.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/xgm.h"

namespace {

const char *kUsage =
    "usage: xgminer -g <graph> [options]\n"
    "  -g, --graph PATH        binary CSR directory or prefix, or an edge-list text file\n"
    "  --algorithm NAME        glumin_g2miner | glumin_g2miner_lut | glumin_cliques |\n"
    "                          glumin_cliques_lut | graphminer (default glumin_g2miner_lut)\n"
    "  --patternID N           GLUMIN pattern id (17-19: vertex-induced 3-star / diamond /\n"
    "                          tailed triangle, unique embeddings; 20: 4-clique, 21: 5-clique)\n"
    "  --pattern-name NAME     triangle | clique4 | clique5 | diamond | rectangle | motif3 | motif4\n"
    "  --pattern-size S --pattern-adj-mat M   pattern by adjacency matrix (row-major 0/1 string)\n"
    "  --repeat R              count R times (default 1)\n"
    "  --strategy NAME         auto | hash | bsearch | lut | list | elut | rank | bucket | merge\n"
    "                          (overrides --algorithm)\n"
    "  --device D              HIP device (default 0)\n"
    "  --run-xgminer 1         X-GMiner's bucket-bitmap P2 (--patternID 18; as --strategy bucket)\n"
    "  -h, --help              this text\n"
    "Accepted and ignored (no effect on counts): --dataname --do-validation --use-graphpi-sched\n"
    "  --run-graphpi --run-our-baseline --tunek --setk --prof-workload --use-vert-para\n"
    "  --prof-edgecheck --vert-induced 0\n";

int fail(const std::string &msg) {
  std::fprintf(stderr, "xgminer: %s\n", msg.c_str());
  return 1;
}

bool is_dir(const std::string &p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

bool is_file(const std::string &p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

bool first_line_two_numbers(const std::string &p) {
  std::ifstream f(p);
  std::string line;
  if (!std::getline(f, line)) return false;
  std::istringstream is(line);
  long long a, b;
  std::string rest;
  return (is >> a >> b) && !(is >> rest);
}

// Loads the graph named by path (undirected); status as xgm_status.
int load_graph(const std::string &path, xgm_hgraph **g, std::string &how) {
  const uint32_t und = XGM_CSR_SORTED | XGM_CSR_SYMMETRIC;
  if (is_dir(path)) {
    how = "binary CSR " + path + "/graph";
    return xgm_load_csr((path + "/graph").c_str(), und, g);
  }
  if (is_file(path + ".meta.txt")) {
    how = "binary CSR " + path;
    return xgm_load_csr(path.c_str(), und, g);
  }
  if (is_file(path)) {
    const size_t dot = path.find_last_of('.'), slash = path.find_last_of('/');
    const bool has_ext = dot != std::string::npos && (slash == std::string::npos || dot > slash);
    if ((has_ext && path.substr(dot) == ".txt") || (!has_ext && first_line_two_numbers(path))) {
      how = "edge list " + path;
      return xgm_load_edgelist(path.c_str(), g);
    }
  }
  return XGM_EIO;
}

struct Pattern {
  int id = 0;          // xgm_pattern, or 0 for a motif census
  int motif_k = 0;     // 3 or 4 for a census
  std::string label;   // as printed
};

// Edge-induced patterns by adjacency matrix (vertex order free).
bool pattern_from_matrix(int size, const std::string &m, Pattern &p) {
  if (size < 3 || size > 5 || (int)m.size() != size * size) return false;
  int edges = 0;
  std::vector<int> deg(size, 0);
  for (int i = 0; i < size; ++i)
    for (int j = 0; j < size; ++j) {
      const char c = m[i * size + j];
      if (c != '0' && c != '1') return false;
      if ((c == '1') != (m[j * size + i] == '1') || (i == j && c == '1')) return false;
      if (c == '1') ++deg[i];
      if (c == '1' && i < j) ++edges;
    }
  const int full = size * (size - 1) / 2;
  if (edges == full) {
    p.id = size == 3 ? XGM_TRIANGLE : size == 4 ? XGM_CLIQUE4 : XGM_CLIQUE5;
  } else if (size == 4 && edges == 5) {
    p.id = XGM_DIAMOND_EDGE;
  } else if (size == 4 && edges == 4 && deg[0] == 2 && deg[1] == 2 && deg[2] == 2 && deg[3] == 2) {
    p.id = XGM_CYCLE4_EDGE;
  } else {
    return false;
  }
  p.label = "adj-mat-" + m;
  return true;
}

}  // namespace

int main(int argc, char **argv) {
  std::map<std::string, std::string> opt;
  const char *ignored[] = {"dataname", "do-validation", "use-graphpi-sched", "run-graphpi", "run-our-baseline",
                           "tunek", "setk", "prof-workload", "use-vert-para", "prof-edgecheck", "vert-induced"};
  const char *known[] = {"graph", "algorithm", "patternID", "pattern-name", "pattern-size", "pattern-adj-mat",
                         "repeat", "strategy", "device", "run-xgminer"};
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    if (a == "-h" || a == "--help") {
      std::fputs(kUsage, stdout);
      return 0;
    }
    std::string key, val;
    if (a == "-g") {
      key = "graph";
    } else if (a.rfind("--", 0) == 0) {
      key = a.substr(2);
      const size_t eq = key.find('=');
      if (eq != std::string::npos) {
        val = key.substr(eq + 1);
        key = key.substr(0, eq);
      }
    } else {
      return fail("unexpected argument '" + a + "'\n" + kUsage);
    }
    if (key == "g") key = "graph";
    bool ok = false;
    for (const char *k : known) ok |= key == k;
    for (const char *k : ignored) ok |= key == k;
    if (!ok) return fail("unknown option '--" + key + "'");
    if (val.empty() && a.find('=') == std::string::npos) {
      if (i + 1 >= argc) return fail("option '" + a + "' needs a value");
      val = argv[++i];
    }
    opt[key] = val;
  }
  if (opt.count("vert-induced") && opt["vert-induced"] != "0")
    return fail("vertex-induced single patterns are not supported (use --pattern-name motif3 / motif4)");
  if (!opt.count("graph") || opt["graph"].empty()) return fail(std::string("no graph given\n") + kUsage);

  // pattern
  Pattern pat;
  if (opt.count("pattern-name")) {
    const std::string n = opt["pattern-name"];
    const std::map<std::string, int> ids = {{"triangle", XGM_TRIANGLE}, {"clique3", XGM_TRIANGLE},
                                            {"clique4", XGM_CLIQUE4},   {"clique5", XGM_CLIQUE5},
                                            {"diamond", XGM_DIAMOND_EDGE}, {"rectangle", XGM_CYCLE4_EDGE}};
    if (n == "motif3" || n == "motif4") {
      pat.motif_k = n[5] - '0';
    } else if (ids.count(n)) {
      pat.id = ids.at(n);
    } else {
      return fail("unsupported pattern name '" + n + "'");
    }
    pat.label = n;
  } else if (opt.count("pattern-adj-mat")) {
    const int size = opt.count("pattern-size") ? std::atoi(opt["pattern-size"].c_str()) : 0;
    if (!pattern_from_matrix(size, opt["pattern-adj-mat"], pat))
      return fail("unsupported pattern matrix (size " + std::to_string(size) + ")");
  } else {
    // GLUMIN numbering: local id = patternID - 16 (xgminer/include/glumin.h:36)
    const int pid = opt.count("patternID") ? std::atoi(opt["patternID"].c_str()) : 0;
    const int local = pid - 16;
    // P1-P3 are the vertex-induced 3-star, diamond and tailed triangle in
    // unique-embedding (GraphMiner) semantics; GLUMIN's P2 kernel walks both
    // directions of every edge and reports twice the induced-diamond count
    // (glumin_GM_solver.cu:267).
    if (local == 4) pat.id = XGM_CLIQUE4;
    else if (local == 5) pat.id = XGM_CLIQUE5;
    else if (local >= 1 && local <= 3)
      pat.id = local == 1 ? XGM_STAR3_V : local == 2 ? XGM_DIAMOND_V : XGM_TAILEDTRI_V;
    else
      return fail("unsupported --patternID " + std::to_string(pid) +
                  " (17-19: vertex-induced 3-star / diamond / tailed triangle, 20: 4-clique, 21: 5-clique)");
    pat.label = "P" + std::to_string(local);
  }

  // strategy
  const std::string algo = opt.count("algorithm") ? opt["algorithm"] : "glumin_g2miner_lut";
  int strategy = XGM_STRAT_AUTO;
  const bool clique = pat.id == XGM_CLIQUE4 || pat.id == XGM_CLIQUE5;
  if (algo == "glumin_g2miner" || algo == "glumin_cliques") {
    if (clique) strategy = XGM_STRAT_LIST;
  } else if (algo == "glumin_g2miner_lut" || algo == "glumin_cliques_lut") {
    if (clique) strategy = XGM_STRAT_LUT;
  } else if (algo == "cpu_baseline") {
    return fail("this build has no CPU path; use a GPU algorithm");
  } else if (algo != "graphminer") {
    return fail("unsupported algorithm '" + algo + "'");
  }
  // the reference's X-GMiner path (main.cpp:48-52): its bucket-bitmap P2
  // kernel for vertex-induced diamonds (xgminer_solver.cu:158)
  if (opt.count("run-xgminer") && std::atoi(opt["run-xgminer"].c_str()) != 0) {
    if (pat.id != XGM_DIAMOND_V) return fail("--run-xgminer counts the vertex-induced diamond (--patternID 18)");
    strategy = XGM_STRAT_BUCKET;
  }
  if (opt.count("strategy")) {
    const std::map<std::string, int> s = {{"auto", XGM_STRAT_AUTO}, {"hash", XGM_STRAT_HASH},
                                          {"bsearch", XGM_STRAT_BSEARCH}, {"lut", XGM_STRAT_LUT},
                                          {"list", XGM_STRAT_LIST}, {"elut", XGM_STRAT_ELUT},
                                          {"rank", XGM_STRAT_LUT_RANK}, {"bucket", XGM_STRAT_BUCKET},
                                          {"merge", XGM_STRAT_MERGE}};
    if (!s.count(opt["strategy"])) return fail("unknown strategy '" + opt["strategy"] + "'");
    strategy = s.at(opt["strategy"]);
  }
  const int repeat = opt.count("repeat") ? std::atoi(opt["repeat"].c_str()) : 1;
  if (repeat < 1) return fail("--repeat must be >= 1");
  const int device = opt.count("device") ? std::atoi(opt["device"].c_str()) : 0;

  // graph
  xgm_hgraph *hg = nullptr;
  std::string how;
  auto t0 = std::chrono::steady_clock::now();
  int rc = load_graph(opt["graph"], &hg, how);
  if (rc) return fail("cannot load graph '" + opt["graph"] + "': " + xgm_strerror(rc));
  xgm_csr csr;
  xgm_hgraph_view(hg, &csr);
  const double load_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::printf("Graph: %s, |V| = %lld, |E| = %lld (undirected), loaded in %.3f s\n", how.c_str(),
              (long long)csr.n, (long long)(csr.nnz / 2), load_s);

  xgm_ctx *ctx = nullptr;
  if ((rc = xgm_ctx_create(device, &ctx))) {
    xgm_hgraph_free(hg);
    return fail(std::string("no usable device: ") + xgm_strerror(rc));
  }
  int status = 0;
  xgm_dgraph *und = nullptr, *dag = nullptr;
  auto report = [&](int r, const char *what) {
    std::fprintf(stderr, "xgminer: %s failed: %s (%s)\n", what, xgm_strerror(r), xgm_ctx_last_error(ctx));
    status = 1;
  };
  t0 = std::chrono::steady_clock::now();
  if ((rc = xgm_graph_upload(ctx, &csr, &und))) report(rc, "upload");
  const bool needs_dag = pat.id == XGM_TRIANGLE || clique;
  if (!status && needs_dag && (rc = xgm_graph_orient(ctx, und, &dag))) report(rc, "orientation");
  if (!status)
    std::printf("Upload%s: %.3f s\n", needs_dag ? " + device orientation" : "",
                std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
  xgm_opts o{};
  o.strategy = strategy;
  for (int r = 0; r < repeat && !status; ++r) {
    t0 = std::chrono::steady_clock::now();
    if (pat.motif_k) {
      uint64_t m[6] = {0, 0, 0, 0, 0, 0};
      if ((rc = xgm_motifs(ctx, und, pat.motif_k, m))) {
        report(rc, "motif census");
        break;
      }
      const char *n3[] = {"wedge", "triangle"};
      const char *n4[] = {"3-star", "4-path", "tailed-triangle", "4-cycle", "diamond", "4-clique"};
      const int cnt = pat.motif_k == 3 ? 2 : 6;
      for (int i = 0; i < cnt; ++i)
        std::printf("Pattern %s count: %llu\n", pat.motif_k == 3 ? n3[i] : n4[i], (unsigned long long)m[i]);
    } else {
      uint64_t c = 0;
      if ((rc = xgm_count(ctx, needs_dag ? dag : und, pat.id, &o, &c))) {
        report(rc, "count");
        break;
      }
      std::printf("Pattern %s count: %llu\n", pat.label.c_str(), (unsigned long long)c);
    }
    xgm_profile prof;
    xgm_last_profile(ctx, &prof);
    std::printf("runtime [xgm] = %.6f sec (kernels %.6f sec)\n",
                std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(),
                prof.total_ms * 1e-3);
  }
  xgm_graph_free(dag);
  xgm_graph_free(und);
  xgm_ctx_destroy(ctx);
  xgm_hgraph_free(hg);
  return status;
}
