// cli.cpp -- hbj_cli: the reference's command-line surface on this backend.
//
//   hbj_cli [--backend hip] info | vec-test | dot-test | spmv-test |
//           block-jacobi-test | pcg-update-scalars-test
//   hbj_cli run-pcg-case --case-dir D --out-x X --out-metrics M
//                        [--max-iters 2000] [--rel-tol 1e-8] [--abs-tol 0]
//                        [--algorithm pcg|one-reduction]   (no reference counterpart:
//                        hbj_system_set_algorithm; pcg, the reference's, is the default)
//   hbj_cli compare-x --x-ref R --x X [--rel-tol 1e-3] [--abs-tol 1e-4] [--top-k 10]
//   hbj_cli text-to-bin --input T --out-dir D [--force]
//
// Subcommands, defaults, outputs and exit codes follow
// crates/wgpu_solver_backend_cli/src/main.rs:23-95,721-891 and
// crates/wgpu_solver_backend_convert_cli/src/main.rs:16-100.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <map>
#include <string>
#include <vector>

#include "hbj.h"
#include "hbj_io.h"

namespace {

using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t) {
  return std::chrono::duration<double, std::milli>(Clock::now() - t).count();
}

std::string now_rfc3339() {
  char buf[64];
  std::time_t t = std::time(nullptr);
  std::strftime(buf, sizeof buf, "%Y-%m-%dT%H:%M:%SZ", std::gmtime(&t));
  return buf;
}

std::string json_str(const std::string& s) {
  std::string o = "\"";
  for (char c : s) {
    if (c == '"' || c == '\\') o += '\\';
    if (c == '\n') { o += "\\n"; continue; }
    o += c;
  }
  return o + "\"";
}

std::string num(double v) {
  if (!std::isfinite(v)) return "null";
  char b[64];
  std::snprintf(b, sizeof b, "%.9g", v);
  return b;
}

[[noreturn]] void die(const std::string& msg) {
  std::fprintf(stderr, "%s\n", msg.c_str());
  std::exit(2);
}

hbj_ctx* make_ctx() {
  hbj_ctx* ctx = nullptr;
  if (hbj_ctx_create(0, &ctx) != HBJ_OK) die(std::string("Failed to init GPU context: ") + hbj_last_error());
  return ctx;
}

std::string gpu_json(hbj_ctx* ctx) {
  char buf[1024];
  if (hbj_ctx_describe(ctx, buf, sizeof buf) != HBJ_OK) return "null";
  return buf;
}

std::string build_json() {
  const char* rev = std::getenv("GIT_REV");
  return std::string("{\"crate_version\": ") + json_str(hbj_version()) +
         ", \"git_rev\": " + (rev ? json_str(rev) : "null") + "}";
}

void check(bool ok, const std::string& msg) {
  if (!ok) {
    std::fprintf(stderr, "%s\n", msg.c_str());
    std::exit(101);  // the reference's assert! panics
  }
}

// ---- self-tests (cli main.rs:171-443)

void vec_test(hbj_ctx* ctx) {
  const uint32_t n = 1024;
  const float alpha = 3.0f;
  std::vector<float> x(n, 2.0f), y(n, 1.0f);
  check(hbj_op_axpy(ctx, n, alpha, x.data(), y.data()) == HBJ_OK, hbj_last_error());
  const float expected = 1.0f + alpha * 2.0f;
  for (uint32_t i = 0; i < n; ++i)
    check(std::fabs(y[i] - expected) < 1e-6f, "vec-test failed at i=" + std::to_string(i));
  std::printf("VecTest OK: all y[i] == %g\n", expected);
}

void dot_test(hbj_ctx* ctx) {
  const float a[3] = {1, 2, 3}, b[3] = {4, 5, 6};
  float got = 0;
  check(hbj_op_dot(ctx, 3, a, b, &got) == HBJ_OK, hbj_last_error());
  check(std::fabs(got - 32.0f) < 1e-5f, "dot-test failed: got " + num(got) + ", expected 32");
  std::printf("DotTest OK: got %g\n", got);
}

void spmv_test(hbj_ctx* ctx) {
  const uint32_t rp[4] = {0, 2, 4, 6}, ci[6] = {0, 2, 0, 1, 1, 2};
  const float va[6] = {10, 2, 3, 9, 7, 8}, x[3] = {1, 2, 3}, ex[3] = {16, 21, 38};
  float y[3] = {0, 0, 0};
  check(hbj_op_spmv(ctx, 3, 6, rp, ci, va, x, y) == HBJ_OK, hbj_last_error());
  for (int i = 0; i < 3; ++i)
    check(std::fabs(y[i] - ex[i]) < 1e-6f, "spmv-test failed at i=" + std::to_string(i));
  std::printf("SpmvTest OK: y == [%g, %g, %g]\n", y[0], y[1], y[2]);
}

void block_jacobi_test(hbj_ctx* ctx) {
  const uint32_t n = 12, bs[3] = {0, 6, 12};
  std::vector<float> lu(72, 0.0f), r(n), z(n, 0.0f);
  for (int b = 0; b < 2; ++b)
    for (int i = 0; i < 6; ++i) lu[b * 36 + i * 6 + i] = 1.0f;
  for (uint32_t i = 0; i < n; ++i) r[i] = (float)(i + 1);
  check(hbj_op_block_jacobi(ctx, n, lu.data(), bs, 3, r.data(), z.data()) == HBJ_OK, hbj_last_error());
  for (uint32_t i = 0; i < n; ++i)
    check(std::fabs(z[i] - r[i]) < 1e-6f, "block-jacobi-test failed at i=" + std::to_string(i));
  std::printf("BlockJacobiTest OK: z == r (identity blocks)\n");
}

void pcg_update_scalars_test(hbj_ctx* ctx) {
  enum { PAP = 0, RZ_NEW = 1, RZ_OLD = 2, ALPHA = 3, MINUS_ALPHA = 4, BETA = 5 };
  float s[8] = {0};
  s[PAP] = 2.0f;
  s[RZ_OLD] = 10.0f;
  s[RZ_NEW] = 5.0f;
  check(hbj_op_update_scalars(ctx, s, 8, PAP, RZ_NEW, RZ_OLD, ALPHA, MINUS_ALPHA, BETA) == HBJ_OK,
        hbj_last_error());
  check(std::fabs(s[ALPHA] - 5.0f) < 1e-6f, "alpha got " + num(s[ALPHA]) + ", expected 5");
  check(std::fabs(s[MINUS_ALPHA] + 5.0f) < 1e-6f, "minus_alpha got " + num(s[MINUS_ALPHA]));
  check(std::fabs(s[BETA] - 0.5f) < 1e-6f, "beta got " + num(s[BETA]) + ", expected 0.5");
  std::printf("PcgUpdateScalarsTest OK: alpha=%g, minus_alpha=%g, beta=%g\n", s[ALPHA],
              s[MINUS_ALPHA], s[BETA]);
}

// ---- argument parsing

struct Args {
  std::map<std::string, std::string> kv;
  bool has(const char* k) const { return kv.count(k) > 0; }
  std::string get(const char* k, const char* dflt = nullptr) const {
    auto it = kv.find(k);
    if (it != kv.end()) return it->second;
    if (!dflt) die(std::string("error: the following required arguments were not provided: --") + k);
    return dflt;
  }
};

Args parse(int argc, char** argv, int from, const std::vector<std::string>& flags) {
  Args a;
  for (int i = from; i < argc; ++i) {
    std::string k = argv[i];
    if (k.rfind("--", 0) != 0) die("error: unexpected argument '" + k + "'");
    k = k.substr(2);
    bool is_flag = false;
    for (auto& f : flags) is_flag |= (f == k);
    if (is_flag) {
      a.kv[k] = "true";
      continue;
    }
    if (i + 1 >= argc) die("error: a value is required for '--" + k + "'");
    a.kv[k] = argv[++i];
  }
  return a;
}

// ---- run-pcg-case (cli main.rs:477-558, 795-878)

int run_pcg_case(const Args& a) {
  const std::string case_dir = a.get("case-dir");
  const uint32_t max_iters = (uint32_t)std::stoul(a.get("max-iters", "2000"));
  const float rel_tol = std::stof(a.get("rel-tol", "1e-8"));
  const float abs_tol = std::stof(a.get("abs-tol", "0"));
  const std::string out_x = a.get("out-x"), out_metrics = a.get("out-metrics");
  const std::string algorithm = a.get("algorithm", "pcg");
  if (algorithm != "pcg" && algorithm != "one-reduction")
    die("--algorithm must be pcg or one-reduction, got " + algorithm);
  const auto t0 = Clock::now();
  auto tg = Clock::now();
  hbj_ctx* ctx = make_ctx();
  const double t_gpu = ms_since(tg);
  auto tl = Clock::now();
  hbj_case* cs = nullptr;
  std::string err;
  bool converged = false;
  uint32_t n = 0, nnz = 0, iters = 0;
  double t_load = 0, t_setup = 0, t_solve = 0, t_write = 0;
  hbj_result res{};
  res.rel_res_true_fp64 = NAN;
  std::vector<float> x;
  if (hbj_case_load(case_dir.c_str(), &cs) != HBJ_OK) {
    err = hbj_last_error();
  } else {
    t_load = ms_since(tl);
    uint32_t ns;
    const uint32_t *rp, *ci, *bs;
    const float *va, *b, *x0;
    hbj_case_view(cs, &n, &nnz, &ns, &rp, &ci, &va, &b, &x0, &bs);
    auto ts = Clock::now();
    hbj_system* sys = nullptr;
    if (hbj_system_create(ctx, n, nnz, rp, ci, va, bs, ns, &sys) != HBJ_OK ||
        (algorithm == "one-reduction" &&
         hbj_system_set_algorithm(sys, HBJ_ALG_PCG_ONE_REDUCTION) != HBJ_OK)) {
      err = hbj_last_error();
      if (sys) hbj_system_destroy(sys);
    } else {
      t_setup = ms_since(ts);
      x.assign(x0, x0 + n);
      auto tv = Clock::now();
      hbj_status st = hbj_solve(sys, b, x.data(), max_iters, rel_tol, abs_tol, &res);
      t_solve = ms_since(tv);
      if (st == HBJ_OK) {
        converged = true;
        iters = res.iterations;
      } else {
        err = hbj_last_error();
      }
      hbj_system_destroy(sys);
    }
  }
  if (!err.empty()) std::fprintf(stderr, "Solve failed: %s\n", err.c_str());
  auto tw = Clock::now();
  if (converged && hbj_vec_bin_write(out_x.c_str(), x.data(), n) != HBJ_OK)
    die(std::string("Failed to write x.bin: ") + hbj_last_error());
  t_write = ms_since(tw);
  char tol[128];
  std::snprintf(tol, sizeof tol, "\"rel_tol\": %s,\n  \"abs_tol\": %s", num(rel_tol).c_str(),
                num(abs_tol).c_str());
  std::string j = "{\n  \"run_id\": " + json_str(now_rfc3339()) +
                  ",\n  \"command\": \"run_pcg_case\",\n  \"case_dir\": " + json_str(case_dir) +
                  ",\n  \"n\": " + std::to_string(converged ? n : 0) +
                  ",\n  \"nnz\": " + std::to_string(converged ? nnz : 0) +
                  ",\n  \"max_iters\": " + std::to_string(max_iters) + ",\n  " + tol +
                  ",\n  \"iterations\": " + (converged ? std::to_string(iters) : "null") +
                  ",\n  \"converged\": " + (converged ? "true" : "false") +
                  ",\n  \"error\": " + (err.empty() ? "null" : json_str(err)) +
                  ",\n  \"true_rel_residual_fp64\": " + num(res.rel_res_true_fp64) +
                  ",\n  \"algorithm\": " + json_str(algorithm) +
                  ",\n  \"timings_ms\": {\"total\": " + num(ms_since(t0)) + ", \"load_io\": " +
                  num(t_load) + ", \"gpu_init\": " + num(t_gpu) + ", \"gpu_setup\": " +
                  num(t_setup) + ", \"solve\": " + num(t_solve) + ", \"write_out\": " +
                  num(t_write) + ", \"device_solve\": " + num(res.solve_ms) + "}" +
                  ",\n  \"gpu\": " + gpu_json(ctx) + ",\n  \"build\": " + build_json() + "\n}";
  FILE* f = nullptr;
  {
    std::string p = out_metrics;
    const size_t slash = p.find_last_of('/');
    if (slash != std::string::npos && slash > 0)
      if (std::system(("mkdir -p -- '" + p.substr(0, slash) + "'").c_str()) != 0)
        die("Failed to write metrics: create_dir_all " + p.substr(0, slash));
    f = std::fopen(p.c_str(), "w");
  }
  if (!f) die("Failed to write metrics: " + out_metrics);
  std::fputs(j.c_str(), f);
  std::fclose(f);
  std::printf("%s\n", j.c_str());
  if (cs) hbj_case_free(cs);
  hbj_ctx_destroy(ctx);
  return 0;
}

// ---- compare-x (cli main.rs:664-719)

int compare_x(const Args& a) {
  const std::string pr = a.get("x-ref"), px = a.get("x");
  const float rel_tol = std::stof(a.get("rel-tol", "1e-3"));
  const float abs_tol = std::stof(a.get("abs-tol", "1e-4"));
  const uint32_t top_k = (uint32_t)std::stoul(a.get("top-k", "10"));
  uint32_t nr = 0, nx = 0;
  if (hbj_vec_bin_read(pr.c_str(), nullptr, 0, &nr) != HBJ_OK) die(hbj_last_error());
  if (hbj_vec_bin_read(px.c_str(), nullptr, 0, &nx) != HBJ_OK) die(hbj_last_error());
  std::vector<float> xr(nr), x(nx);
  if (hbj_vec_bin_read(pr.c_str(), xr.data(), nr, &nr) != HBJ_OK) die(hbj_last_error());
  if (hbj_vec_bin_read(px.c_str(), x.data(), nx, &nx) != HBJ_OK) die(hbj_last_error());
  if (nr != nx) die("Length mismatch: x_ref len " + std::to_string(nr) + " vs x len " + std::to_string(nx));
  hbj_compare_stats s{};
  std::vector<uint32_t> worst(top_k ? top_k : 1);
  hbj_compare_x(xr.data(), x.data(), nr, top_k, &s, worst.data());
  const bool pass_rel = s.rel_l2 <= (double)rel_tol, pass_rmse = s.rmse <= (double)abs_tol;
  const bool pass = pass_rel || pass_rmse;
  std::printf("CompareX:\n");
  std::printf("  n                 : %u\n", s.n);
  std::printf("  rtol (rel L2)      : %.3e\n", rel_tol);
  std::printf("  atol (RMSE)        : %.3e\n", abs_tol);
  std::printf("  ||x_ref||_2        : %.9e\n", s.l2_ref);
  std::printf("  ||x - x_ref||_2    : %.9e\n", s.l2_err);
  std::printf("  rel_l2             : %.9e\n", s.rel_l2);
  std::printf("  rmse               : %.9e\n", s.rmse);
  std::printf("  max_abs_err (diag) : %.9e\n", s.max_abs_err);
  std::printf("  pass_rel_l2        : %s\n", pass_rel ? "true" : "false");
  std::printf("  pass_rmse          : %s\n", pass_rmse ? "true" : "false");
  std::printf("  pass               : %s\n", pass ? "true" : "false");
  if (s.n_worst) {
    std::printf("\nTop %u worst entries (by abs error):\n", s.n_worst);
    for (uint32_t k = 0; k < s.n_worst; ++k) {
      const uint32_t i = worst[k];
      std::printf("  i=%-8u x_ref=%.9e  x=%.9e  abs=%.9e\n", i, xr[i], x[i], std::fabs(x[i] - xr[i]));
    }
  }
  if (!pass) die("CompareX failed tolerances (rel_l2 and rmse)");
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  int i = 1;
  if (i + 1 < argc && std::strcmp(argv[i], "--backend") == 0) i += 2;  // only HIP exists here
  if (i >= argc) die("usage: hbj_cli [--backend hip] <info|vec-test|dot-test|spmv-test|"
                     "block-jacobi-test|pcg-update-scalars-test|run-pcg-case|compare-x|text-to-bin> ...");
  const std::string cmd = argv[i];
  if (cmd == "compare-x") return compare_x(parse(argc, argv, i + 1, {}));
  if (cmd == "text-to-bin") {
    Args a = parse(argc, argv, i + 1, {"force"});
    const std::string out = a.get("out-dir");
    if (hbj_text_to_bin(a.get("input").c_str(), out.c_str(), a.has("force")) != HBJ_OK)
      die(std::string("Error: ") + hbj_last_error());
    std::printf("Wrote binaries into %s\n", out.c_str());
    return 0;
  }
  if (cmd == "run-pcg-case") return run_pcg_case(parse(argc, argv, i + 1, {}));
  hbj_ctx* ctx = make_ctx();
  if (cmd == "info") {
    std::printf("%s\n", gpu_json(ctx).c_str());
    std::printf("{\n  \"run_id\": %s,\n  \"command\": \"info\",\n  \"gpu\": %s,\n  \"build\": %s\n}\n",
                json_str(now_rfc3339()).c_str(), gpu_json(ctx).c_str(), build_json().c_str());
  } else if (cmd == "vec-test") {
    vec_test(ctx);
  } else if (cmd == "dot-test") {
    dot_test(ctx);
  } else if (cmd == "spmv-test") {
    spmv_test(ctx);
  } else if (cmd == "block-jacobi-test") {
    block_jacobi_test(ctx);
  } else if (cmd == "pcg-update-scalars-test") {
    pcg_update_scalars_test(ctx);
  } else {
    die("error: unrecognized subcommand '" + cmd + "'");
  }
  hbj_ctx_destroy(ctx);
  return 0;
}
