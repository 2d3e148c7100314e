// hd_exec: command-line driver over the C ABI, with the reference's flag
// names (src/flags.cc:3-57, parsed into RunConfig at src/main.cpp:47-132).
// Inputs (include/hd_io.h): WKT, OFF and PLY files, the reference's point
// cache (".bin": size_t n, then n packed points; -serialize <dir> caches
// parsed WKT/OFF there, src/loaders/loader.h:176-234), or a named synthetic
// workload (include/hd_gen.h). Preprocessing -normalize and -translate
// follow src/loaders/translate_points.h:42-80.
//
// -vary_params runs the hybrid once per combination of -sample_rate_list,
// -eb_only_threshold_list, -max_hit_list and -n_points_cell_list and writes
// one JSON per combination (src/autotune_hausdorff_distance.cu:181-296).
//
//   hd_exec -input1 a.wkt -input2 b.wkt -input_type wkt -n_dims 2 \
//           -variant eb -execution gpu -repeat 5 -serialize cache -json out.json
//   hd_exec -input_type synthetic -input1 gmm_10m_2d -n_dims 2 -variant hybrid
//   hd_exec ... -vary_params -n_points_cell_list 1,2,4,8,16 -json logs/pair
#include <unistd.h>

#include <algorithm>
#include <cctype>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>
#include <ctime>

#include "hd.h"
#include "hd_gen.h"
#include "hd_io.h"

namespace {

struct Flags {
  std::map<std::string, std::string> kv;
  std::string Get(const std::string& k, const std::string& def) const {
    auto it = kv.find(k);
    return it == kv.end() ? def : it->second;
  }
  bool Bool(const std::string& k, bool def) const {
    auto it = kv.find(k);
    if (it == kv.end()) return def;
    std::string v = it->second;
    std::transform(v.begin(), v.end(), v.begin(), ::tolower);
    return v.empty() || v == "true" || v == "1" || v == "yes";
  }
  long Int(const std::string& k, long def) const {
    auto it = kv.find(k);
    return it == kv.end() ? def : std::strtol(it->second.c_str(), nullptr, 10);
  }
  double Dbl(const std::string& k, double def) const {
    auto it = kv.find(k);
    return it == kv.end() ? def : std::strtod(it->second.c_str(), nullptr);
  }
};

const char* kBoolFlags[] = {"check", "normalize", "symmetric", "learned_order",
                            "auto_tune", "auto_tune_n_points_cell",
                            "auto_tune_max_hit", "overwrite", "shuffle",
                            "vary_params"};

bool IsBool(const std::string& k) {
  for (const char* b : kBoolFlags)
    if (k == b) return true;
  return false;
}

// gflags style: -flag=value, -flag value, --flag, -noflag for booleans.
bool Parse(int argc, char** argv, Flags* f) {
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    if (a.size() < 2 || a[0] != '-') {
      fprintf(stderr, "unexpected argument %s\n", argv[i]);
      return false;
    }
    a = a.substr(a[1] == '-' ? 2 : 1);
    const size_t eq = a.find('=');
    if (eq != std::string::npos) {
      f->kv[a.substr(0, eq)] = a.substr(eq + 1);
    } else if (IsBool(a)) {
      f->kv[a] = "true";
    } else if (a.rfind("no", 0) == 0 && IsBool(a.substr(2))) {
      f->kv[a.substr(2)] = "false";
    } else if (i + 1 < argc) {
      f->kv[a] = argv[++i];
    } else {
      fprintf(stderr, "flag -%s needs a value\n", a.c_str());
      return false;
    }
  }
  return true;
}

// A file through hd_load_points (host memory, AoS).
bool LoadFile(const std::string& path, int input_type, int nd, long limit,
              const std::string& serialize, std::vector<float>* out) {
  float* p = nullptr;
  uint64_t n = 0;
  if (hd_load_points(path.c_str(), input_type, nd, (uint64_t)limit,
                     serialize.c_str(), &p, &n) != HD_OK) {
    fprintf(stderr, "%s\n", hd_last_error());
    return false;
  }
  out->assign(p, p + n * nd);
  hd_free_points(p);
  return true;
}

bool Synthetic(const std::string& name, bool second, int nd, long limit,
               std::vector<float>* out) {
  auto clip = [&](uint64_t n) { return std::min<uint64_t>(n, (uint64_t)limit); };
  if (name == "uniform_100k_2d" || name == "uniform_100k_3d") {
    const uint64_t n = clip(100000);
    out->resize(n * nd);
    return hd_gen_uniform(second ? 2 : 1, n, nd, out->data()) == HD_OK;
  }
  if (name == "gmm_10m_2d" || name == "gmm_50m_256_2d") {
    if (nd != 2) return false;
    const bool big = name == "gmm_50m_256_2d";
    const uint64_t n = clip(big ? 50000000 : 10000000);
    std::vector<float> a(2 * n), b(2 * n);
    const int rc = big ? hd_gen_gmm_pair(n, 256, 0.01, 0.002f, 0.001f, n / 100,
                                         10, 11, 12, 13, a.data(), b.data())
                       : hd_gen_gmm_pair(n, 64, 0.01, 0.002f, 0.001f, n / 100, 3,
                                         4, 5, 6, a.data(), b.data());
    if (rc != HD_OK) return false;
    *out = second ? b : a;
    return true;
  }
  if (name == "sphere_20m_15m_3d") {
    if (nd != 3) return false;
    const uint64_t n = clip(second ? 15000000 : 20000000);
    out->resize(3 * n);
    return hd_gen_sphere(second ? 8 : 7, n, 0.005, second ? 2.0 : 0.0,
                         out->data()) == HD_OK;
  }
  return false;
}

template <typename T>
std::vector<T> SplitList(const std::string& text) {
  std::vector<T> out;
  std::stringstream ss(text);
  std::string item;
  while (std::getline(ss, item, ','))
    if (!item.empty()) out.push_back((T)std::strtod(item.c_str(), nullptr));
  return out;
}

std::string JsonString(const std::string& v) {
  std::string esc;
  for (char ch : v) {
    if (ch == '"' || ch == '\\') esc += '\\';
    esc += ch;
  }
  return "\"" + esc + "\"";
}

// The loading options and file paths the library does not see go into the
// stats JSON's "Input" (run_hausdorff_distance.cu:196-199; the paths and
// the time stamp as autotune_hausdorff_distance.cu:96-99,225).
std::string WithInputOptions(std::string js, const Flags& f, long limit,
                             double translate, const std::string& date) {
  const std::string key = "\"Input\": {";
  const size_t at = js.find(key);
  if (at == std::string::npos) return js;
  char extra[160];
  snprintf(extra, sizeof(extra), "\"Normalize\": %s, \"Translate\": %.17g, ",
           f.Bool("normalize", false) ? "true" : "false", translate);
  js.insert(at + key.size(), std::string(extra) + "\"SerializationPrefix\": " +
                                 JsonString(f.Get("serialize", "")) +
                                 ", \"Limit\": " + std::to_string(limit) + ", ");
  const std::string files = "\"Files\": [{";
  const size_t fa = js.find(files, at);
  if (fa != std::string::npos) {
    js.insert(fa + files.size(), "\"Path\": " + JsonString(f.Get("input1", "")) + ", ");
    const size_t fb = js.find("}, {\"NumPoints\"", fa);
    if (fb != std::string::npos)
      js.insert(fb + 4, "\"Path\": " + JsonString(f.Get("input2", f.Get("input1", ""))) + ", ");
  }
  if (!date.empty()) js.insert(1, "\"DateTime\": " + JsonString(date) + ", ");
  return js;
}

std::string NowString() {
  const std::time_t t = std::time(nullptr);
  std::tm tm;
  localtime_r(&t, &tm);
  char buf[32];
  std::strftime(buf, sizeof(buf), "%Y-%m-%d %H:%M:%S", &tm);
  return buf;
}

// The parameter sweep (autotune_hausdorff_distance.cu:181-296): the hybrid
// for every combination, one JSON per combination named
// <json>_sample_rate_%.6f_eb_only_threshold_%u_max_hit_%u_n_points_cell_%u.json,
// existing files skipped unless -overwrite. Sets B of at least 500k points
// sweep only eb_only_threshold 0 (the reference's sample threshold).
int VaryParams(hd_context* ctx, hd_run_config cfg, const Flags& f,
               const std::vector<float>& a, const std::vector<float>& b,
               long limit, double translate) {
  const uint64_t na = a.size() / cfg.n_dims, nb = b.size() / cfg.n_dims;
  const auto rates = SplitList<double>(f.Get("sample_rate_list", "0.01"));
  auto thresholds = SplitList<long>(f.Get("eb_only_threshold_list", "0"));
  const auto hits = SplitList<long>(f.Get("max_hit_list", "256"));
  const auto cells = SplitList<long>(f.Get("n_points_cell_list", "8"));
  if (nb >= 500000) thresholds.assign(1, 0);
  const std::string json = f.Get("json", "");
  const bool overwrite = f.Bool("overwrite", false);
  cfg.variant = HD_VARIANT_HYBRID;
  cfg.flags &= ~HD_FLAG_CHECK;
  const size_t total = rates.size() * thresholds.size() * hits.size() * cells.size();
  size_t done = 0, skipped = 0;
  double best = -1;
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<char> stats(1 << 20);
  for (double rate : rates)
    for (long thr : thresholds)
      for (long hit : hits)
        for (long npc : cells) {
          if (npc <= 0 || rate > 1 || rate < 0) {
            fprintf(stderr, "bad sweep point: sample_rate %g, n_points_cell %ld\n", rate, npc);
            return 1;
          }
          char path[4096];
          snprintf(path, sizeof(path),
                   "%s_sample_rate_%.6f_eb_only_threshold_%u_max_hit_%u_n_points_cell_%u.json",
                   json.c_str(), rate, (unsigned)thr, (unsigned)hit, (unsigned)npc);
          const bool exists = !json.empty() && access(path, R_OK) == 0;
          if (exists && !overwrite) {
            printf("Skip %s\n", path);
            ++skipped;
            continue;
          }
          cfg.sample_rate = (float)rate;
          cfg.eb_only_threshold = (uint32_t)thr;
          cfg.max_hit = (uint32_t)hit;
          cfg.n_points_cell = (uint32_t)npc;
          const std::string date = NowString();
          hd_result res;
          const int rc = hd_run_hausdorff_distance(ctx, &cfg, a.data(), na, b.data(), nb,
                                                   &res, stats.data(), stats.size());
          if (rc != HD_OK) {
            fprintf(stderr, "error %d: %s\n", rc, hd_last_error());
            return 1;
          }
          best = best < 0 ? res.avg_time_ms : std::min(best, res.avg_time_ms);
          printf("Sample rate %g EB only threshold %ld Max hit %ld Points/cell %ld: "
                 "Avg Running Time %.2f ms\n", rate, thr, hit, npc, res.avg_time_ms);
          if (!json.empty()) {
            std::ofstream ofs(path);
            ofs << WithInputOptions(stats.data(), f, limit, translate, date) << "\n";
          }
          ++done;
          const double el = std::chrono::duration<double>(
                                std::chrono::steady_clock::now() - t0).count();
          printf("Progress %.2f %% Remaining Time %.2f s Best Performance %.2f ms\n",
                 100.0 * (done + skipped) / total, el / done * (total - done - skipped), best);
          fflush(stdout);
        }
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  Flags f;
  if (argc == 1 || !Parse(argc, argv, &f)) {
    fprintf(stderr,
            "Usage: hd_exec -input1 A -input2 B -input_type wkt|off|ply|bin|synthetic "
            "-n_dims 2|3 [-variant eb|grid|rt|lbvh|hybrid|nn|rt-hdist] [-execution gpu|cpu] "
            "[-repeat 5] [-seed 0] [-parallelism -1] [-check|-nocheck] [-symmetric] [-learned_order] "
            "[-n_points_cell 8] [-max_hit 256] [-auto_tune] [-auto_tune_n_points_cell] "
            "[-auto_tune_max_hit] [-b_order auto|random|stratified] [-limit N] "
            "[-stats_n_points_cell 10] [-vary_params -sample_rate_list r,.. "
            "-eb_only_threshold_list n,.. -max_hit_list n,.. -n_points_cell_list n,.. "
            "[-overwrite]] "
            "[-normalize] [-translate r] [-serialize dir] [-json out.json]\n");
    return 1;
  }
  hd_run_config cfg;
  hd_run_config_init(&cfg);
  cfg.n_dims = (int)f.Int("n_dims", 0);
  if (cfg.n_dims != 2 && cfg.n_dims != 3) {
    fprintf(stderr, "Wrong number of dimensions, which can only be 2 or 3\n");
    return 1;
  }
  std::string variant = f.Get("variant", "eb");
  std::transform(variant.begin(), variant.end(), variant.begin(), ::tolower);
  if (variant == "eb") cfg.variant = HD_VARIANT_EB;
  else if (variant == "grid") cfg.variant = HD_VARIANT_GRID;
  else if (variant == "rt" || variant == "lbvh") cfg.variant = HD_VARIANT_LBVH;
  else if (variant == "hybrid") cfg.variant = HD_VARIANT_HYBRID;
  else if (variant == "nn") cfg.variant = HD_VARIANT_NN;
  else if (variant == "rt-hdist") cfg.variant = HD_VARIANT_RT_HDIST;
  else {
    fprintf(stderr, "Unknown variant: %s\n", variant.c_str());
    return 1;
  }
  std::string exec = f.Get("execution", "gpu");
  std::transform(exec.begin(), exec.end(), exec.begin(), ::tolower);
  if (exec == "gpu") cfg.execution = HD_EXEC_GPU;
  else if (exec == "cpu") cfg.execution = HD_EXEC_CPU;
  else {
    fprintf(stderr, "Unknown execution: %s\n", exec.c_str());
    return 1;
  }
  cfg.repeat = (int)f.Int("repeat", 5);
  cfg.seed = (int)f.Int("seed", 0);
  cfg.parallelism = (int)f.Int("parallelism", -1);
  cfg.sample_rate = (float)f.Dbl("sample_rate", 0.01);
  cfg.n_points_cell = (uint32_t)f.Int("n_points_cell", 8);
  cfg.max_hit = (uint32_t)f.Int("max_hit", 256);
  cfg.eb_only_threshold = (uint32_t)f.Int("eb_only_threshold", 0);
  cfg.auto_tune = (f.Bool("auto_tune", false) ? HD_AUTO_TUNE_ALL : 0) |
                  (f.Bool("auto_tune_n_points_cell", false) ? HD_AUTO_TUNE_N_POINTS_CELL : 0) |
                  (f.Bool("auto_tune_max_hit", false) ? HD_AUTO_TUNE_MAX_HIT : 0);
  cfg.stats_n_points_cell = (int32_t)f.Int("stats_n_points_cell", 10);
  // -check defaults to on, as the reference's flag (src/flags.cc:49): every
  // run is compared with the host early break (-nocheck turns it off).
  if (f.Bool("check", true)) cfg.flags |= HD_FLAG_CHECK;
  if (f.Bool("symmetric", false)) cfg.flags |= HD_FLAG_SYMMETRIC;
  if (f.Bool("learned_order", false)) cfg.flags |= HD_FLAG_LEARNED_ORDER;
  if (!f.Bool("shuffle", true)) cfg.flags |= HD_FLAG_NO_SHUFFLE;
  // Early break: the visit order of B (auto: stratified with the learned
  // order, else pseudo-random as the reference shuffles it; same result).
  const std::string b_order = f.Get("b_order", "auto");
  if (b_order == "random")
    cfg.flags |= HD_FLAG_B_ORDER_RANDOM;
  else if (b_order == "stratified")
    cfg.flags |= HD_FLAG_B_ORDER_STRATIFIED;
  else if (b_order != "auto") {
    fprintf(stderr, "Unknown b_order: %s (auto, random or stratified)\n", b_order.c_str());
    return 1;
  }

  const long limit = f.Int("limit", 0x7FFFFFFFL);
  const std::string type = f.Get("input_type", "");
  const std::string in1 = f.Get("input1", ""), in2 = f.Get("input2", "");
  const std::string serialize = f.Get("serialize", "");
  std::vector<float> a, b;
  bool ok = false;
  const int file_type = type == "wkt" ? HD_INPUT_WKT : type == "off" ? HD_INPUT_OFF
                        : type == "ply" ? HD_INPUT_PLY : type == "bin" ? HD_INPUT_BIN
                        : -1;
  if (file_type >= 0) {
    ok = LoadFile(in1, file_type, cfg.n_dims, limit, serialize, &a) &&
         LoadFile(in2, file_type, cfg.n_dims, limit, serialize, &b);
  } else if (type == "synthetic") {
    const std::string n2 = in2.empty() ? in1 : in2;
    ok = Synthetic(in1, false, cfg.n_dims, limit, &a) &&
         Synthetic(n2, true, cfg.n_dims, limit, &b);
  } else {
    fprintf(stderr, "Unsupported input type: %s\n", type.c_str());
    return 1;
  }
  if (!ok) {
    fprintf(stderr, "cannot load inputs %s / %s (%s)\n", in1.c_str(),
            in2.c_str(), type.c_str());
    return 1;
  }
  if (f.Bool("normalize", false)) {
    hd_normalize_points(a.data(), a.size() / cfg.n_dims, cfg.n_dims);
    hd_normalize_points(b.data(), b.size() / cfg.n_dims, cfg.n_dims);
  }
  const double tr = f.Dbl("translate", 0.0);
  if (tr != 0) hd_translate_points(b.data(), b.size() / cfg.n_dims, cfg.n_dims, 0, tr);

  const uint64_t na = a.size() / cfg.n_dims, nb = b.size() / cfg.n_dims;
  printf("Points A: %llu Points B: %llu\n", (unsigned long long)na,
         (unsigned long long)nb);
  hd_context* ctx = nullptr;
  if (cfg.execution == HD_EXEC_GPU && hd_context_create(0, &ctx) != HD_OK) {
    fprintf(stderr, "%s\n", hd_last_error());
    return 1;
  }
  if (f.Bool("vary_params", false)) {
    const int vrc = VaryParams(ctx, cfg, f, a, b, limit, tr);
    hd_context_destroy(ctx);
    return vrc;
  }
  hd_result res;
  std::vector<char> stats(1 << 20);
  int rc = hd_run_hausdorff_distance(ctx, &cfg, a.data(), na, b.data(), nb,
                                     &res, stats.data(), stats.size());
  if (rc != HD_OK && rc != HD_ETRUNC) {
    fprintf(stderr, "error %d: %s\n", rc, hd_last_error());
    hd_context_destroy(ctx);
    return 1;
  }
  printf("HausdorffDistance: distance is %.9g (cmax2 bits 0x%08x, argmax %u, "
         "nn %u, direction %d)\n",
         res.hd, res.cmax2_bits, res.argmax, res.nn, res.direction);
  printf("Avg Running Time %.3f ms\n", res.avg_time_ms);
  if (res.check_pass >= 0)
    printf("HausdorffDistance is %s\n", res.check_pass ? "checked" : "WRONG");
  const std::string json = f.Get("json", "");
  if (!json.empty()) {
    std::ofstream ofs(json);
    ofs << WithInputOptions(stats.data(), f, limit, tr, "") << "\n";
  }
  hd_context_destroy(ctx);
  return res.check_pass == 0 ? 2 : 0;
}
