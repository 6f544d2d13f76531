// warpjq-amd: the `warpjq <query> [files...]` command line over the MI355X
// backend (crates/warpjq-cli/src/main.rs:46-96, run.rs:16-150), running every
// query on the GPU through warpjq_amd_run_paths.
//
//   flags        --csv, -c/--count, --backend auto|gpu|cpu, --strict,
//                --skip-invalid, -j/--threads, --chunk-size, --max-line-bytes,
//                --stats (main.rs:46-96)
//   inputs       files read as one stream; none, or "-", reads stdin
//                (run.rs:67-83)
//   exit codes   0 ok; 1 no rows from a non-aggregate query (like grep), or a
//                runtime error; 2 a query or usage error (main.rs:155-190,
//                run.rs:95-101)
//   stderr       the malformed / type-error summary and --stats line
//                (run.rs:104-150)
//   gen          `warpjq-amd gen --preset nginx|nested --size SIZE -o FILE
//                --seed N`: the seeded generators of this backend's workloads
//
// This build has one engine, the GPU: `--backend cpu` is refused (the
// reference's CPU engine is not part of it) and `auto` means the GPU.
#include <cerrno>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/warpjq_amd.h"
#include "../../include/warpjq_kernels.h"

namespace {

const char *USAGE =
    "usage: warpjq-amd [OPTIONS] <QUERY> [FILES]...\n"
    "       warpjq-amd gen [--preset nginx|nested] [--size SIZE] [-o FILE] [--seed N]\n"
    "\n"
    "Runs a jq-subset query over NDJSON on the GPU. Reads stdin when no file is given.\n"
    "\n"
    "options:\n"
    "  --csv                  emit CSV instead of NDJSON\n"
    "  --input-csv            read one CSV input (header line, then one record per line) instead of NDJSON\n"
    "  -c, --count            print only how many rows the query produced\n"
    "  --backend <auto|gpu|cpu>  engine (this build runs the GPU; cpu is refused)\n"
    "  --strict               stop at the first malformed line\n"
    "  --skip-invalid         skip malformed lines without a warning\n"
    "  -j, --threads <N>      accepted for compatibility (host copies use a fixed pool)\n"
    "  --chunk-size <SIZE>    bytes per pipeline chunk (default 256MB; the device takes at most 64MB)\n"
    "  --max-line-bytes <SIZE>  reject any single line longer than this (default 64MB)\n"
    "  --stats                print a timing and throughput summary to stderr\n"
    "  -h, --help             print this help\n";

// "256MB", "1GB", "4096", "1.5gb" (main.rs:217-248). False on a bad size.
bool parse_size(const std::string &in, uint64_t &out, std::string &err) {
  std::string t;
  for (char ch : in)
    if (ch != ' ' && ch != '\t') t.push_back((char)tolower((unsigned char)ch));
  uint64_t mult = 1;
  auto strip = [&](const char *a, const char *b, uint64_t m) {
    for (const char *suf : {a, b}) {
      const size_t n = strlen(suf);
      if (t.size() >= n && t.compare(t.size() - n, n, suf) == 0) {
        t.resize(t.size() - n);
        mult = m;
        return true;
      }
    }
    return false;
  };
  if (!strip("tb", "t", 1ull << 40) && !strip("gb", "g", 1ull << 30) && !strip("mb", "m", 1ull << 20) &&
      !strip("kb", "k", 1ull << 10))
    strip("b", "b", 1);
  char *end = nullptr;
  errno = 0;
  const double v = t.empty() ? NAN : strtod(t.c_str(), &end);
  if (t.empty() || !end || *end || errno) {
    err = "could not read `" + in + "` as a size (try 256MB or 1GB)";
    return false;
  }
  if (!std::isfinite(v)) {
    err = "`" + in + "` is not a usable size (try 256MB or 1GB)";
    return false;
  }
  if (v < 0) {
    err = "size cannot be negative: `" + in + "`";
    return false;
  }
  out = (uint64_t)(v * (double)mult);
  return true;
}

// main.rs:251-267
std::string human_bytes(uint64_t n) {
  const char *units[] = {"B", "KB", "MB", "GB", "TB"};
  double v = (double)n;
  int i = 0;
  while (v >= 1024.0 && i + 1 < 5) {
    v /= 1024.0;
    i++;
  }
  char b[64];
  if (i == 0) snprintf(b, sizeof b, "%llu B", (unsigned long long)n);
  else if (v >= 100.0) snprintf(b, sizeof b, "%.0f %s", v, units[i]);
  else snprintf(b, sizeof b, "%.1f %s", v, units[i]);
  return b;
}

int usage_error(const std::string &msg) {
  fprintf(stderr, "warpjq-amd: %s\n\nTry 'warpjq-amd --help'.\n", msg.c_str());
  return 2;
}

bool write_all(const uint8_t *p, uint64_t n) {
  while (n) {
    const size_t k = fwrite(p, 1, n, stdout);
    if (k == 0) return false;
    p += k;
    n -= k;
  }
  return fflush(stdout) == 0;
}

int run_gen(int argc, char **argv) {
  std::string preset = "nginx", size = "1GB", outp;
  uint64_t seed = 0;
  for (int i = 2; i < argc; i++) {
    std::string a = argv[i];
    auto val = [&](std::string &dst) {
      if (i + 1 >= argc) return false;
      dst = argv[++i];
      return true;
    };
    std::string s;
    if (a == "--preset") {
      if (!val(preset)) return usage_error("--preset needs a value");
    } else if (a == "--size") {
      if (!val(size)) return usage_error("--size needs a value");
    } else if (a == "-o" || a == "--output") {
      if (!val(outp)) return usage_error("-o needs a value");
    } else if (a == "--seed") {
      if (!val(s)) return usage_error("--seed needs a value");
      seed = strtoull(s.c_str(), nullptr, 10);
    } else if (a == "--list") {
      printf("Presets:\n  nginx        flat access-log records (ts, host, method, path, status, bytes, ua)\n"
             "  nested       nested records with escaped keys, users, float and int bytes\n");
      return 0;
    } else {
      return usage_error("unexpected argument `" + a + "` for gen");
    }
  }
  uint64_t n = 0;
  std::string e;
  if (!parse_size(size, n, e)) return usage_error(e);
  if (n == 0) return usage_error("--size must be greater than zero");
  uint8_t *buf = nullptr;
  uint64_t got = 0, lines = 0, extra[8] = {0};
  int rc;
  if (preset == "nginx") rc = warpjq_amd_gen_nginx(seed, n, 16, 0, &buf, &got, &lines, extra);
  else if (preset == "nested") rc = warpjq_amd_gen_nested(seed, n, 16, 0, &buf, &got, &lines, extra);
  else return usage_error("unknown preset `" + preset + "`; try one of: nginx, nested");
  if (rc) {
    fprintf(stderr, "warpjq-amd: out of memory generating %s\n", human_bytes(n).c_str());
    return 1;
  }
  FILE *f = outp.empty() ? stdout : fopen(outp.c_str(), "wb");
  if (!f) {
    fprintf(stderr, "warpjq-amd: could not create `%s`: %s\n", outp.c_str(), strerror(errno));
    warpjq_amd_free(buf);
    return 1;
  }
  const bool ok = fwrite(buf, 1, got, f) == got;
  warpjq_amd_free(buf);
  if (f != stdout) fclose(f);
  else fflush(stdout);
  if (!ok) {
    fprintf(stderr, "warpjq-amd: writing the output failed\n");
    return 1;
  }
  if (!outp.empty())
    fprintf(stderr, "warpjq-amd: wrote %s (%llu lines) of `%s` data with seed %llu\n", human_bytes(got).c_str(),
            (unsigned long long)lines, preset.c_str(), (unsigned long long)seed);
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc >= 2 && !strcmp(argv[1], "gen")) return run_gen(argc, argv);
  std::string query;
  bool have_query = false, csv = false, count = false, strict = false, skip = false, stats = false, input_csv = false;
  std::string backend = "auto", chunk = "256MB", max_line = "64MB";
  std::vector<const char *> files;
  for (int i = 1; i < argc; i++) {
    std::string a = argv[i];
    auto val = [&](std::string &dst) {
      const size_t eq = a.find('=');
      if (a.rfind("--", 0) == 0 && eq != std::string::npos) {
        dst = a.substr(eq + 1);
        return true;
      }
      if (i + 1 >= argc) return false;
      dst = argv[++i];
      return true;
    };
    const std::string key = a.rfind("--", 0) == 0 ? a.substr(0, a.find('=')) : a;
    std::string ignored;
    if (key == "-h" || key == "--help") {
      fputs(USAGE, stdout);
      return 0;
    } else if (key == "--csv") {
      csv = true;
    } else if (key == "--input-csv") {
      input_csv = true;
    } else if (key == "-c" || key == "--count") {
      count = true;
    } else if (key == "--strict") {
      strict = true;
    } else if (key == "--skip-invalid") {
      skip = true;
    } else if (key == "--stats") {
      stats = true;
    } else if (key == "--backend") {
      if (!val(backend)) return usage_error("--backend needs a value");
    } else if (key == "-j" || key == "--threads") {
      if (!val(ignored)) return usage_error(key + " needs a value");
    } else if (key == "--chunk-size") {
      if (!val(chunk)) return usage_error("--chunk-size needs a value");
    } else if (key == "--max-line-bytes") {
      if (!val(max_line)) return usage_error("--max-line-bytes needs a value");
    } else if (a.size() > 1 && a[0] == '-' && a != "-") {
      return usage_error("unexpected argument `" + a + "`");
    } else if (!have_query) {
      query = a;
      have_query = true;
    } else {
      files.push_back(argv[i]);
    }
  }
  if (!have_query) return usage_error("no query given");
  if (csv && count) return usage_error("the argument '--csv' cannot be used with '--count'");
  if (input_csv && files.size() > 1) return usage_error("--input-csv reads one input");
  if (strict && skip) return usage_error("the argument '--strict' cannot be used with '--skip-invalid'");
  if (backend != "auto" && backend != "gpu" && backend != "cpu")
    return usage_error("invalid value '" + backend + "' for '--backend' (possible values: auto, gpu, cpu)");
  if (backend == "cpu") return usage_error("--backend cpu: this build has only the GPU engine");
  uint64_t chunk_bytes = 0, max_line_bytes = 0;
  std::string e;
  if (!parse_size(chunk, chunk_bytes, e) || !parse_size(max_line, max_line_bytes, e)) {
    fprintf(stderr, "warpjq-amd: %s\n", e.c_str());
    return 1;
  }
  if (chunk_bytes == 0) {
    fprintf(stderr, "warpjq-amd: --chunk-size must be greater than zero\n");
    return 1;
  }

  // a query error says so before any device work (run.rs:17)
  char err[1024] = {0};
  warpjq_amd_program *prog = warpjq_amd_compile(query.c_str(), csv ? 1 : 0, err, sizeof err);
  if (!prog) {
    fprintf(stderr, "warpjq-amd: %s\n", err);
    return 2;
  }
  const bool aggregate = warpjq_amd_program_view(prog)->output_kind == WARPJQ_OUT_AGG;
  warpjq_amd_program_free(prog);
  // --backend gpu with no usable GPU fails before reading (run.rs:45-55)
  if (warpjq_probe(err, sizeof err) != WARPJQ_OK) {
    fprintf(stderr, "warpjq-amd: --backend %s was requested but the GPU is not usable: %s\n", backend.c_str(), err);
    return 1;
  }

  const int format = csv ? 1 : count ? 2 : 0;
  const int on_invalid = strict ? 2 : skip ? 1 : 0;
  uint8_t *out = nullptr;
  uint64_t out_len = 0;
  warpjq_amd_stats st;
  const auto t0 = std::chrono::steady_clock::now();
  int rc;
  if (input_csv) {
    // the whole input, then the device transcodes its records to NDJSON
    std::vector<uint8_t> buf;
    FILE *f = files.empty() || strcmp(files[0], "-") == 0 ? stdin : fopen(files[0], "rb");
    if (!f) {
      fprintf(stderr, "warpjq-amd: could not open %s: %s\n", files[0], strerror(errno));
      return 1;
    }
    uint8_t tmp[1 << 16];
    size_t got;
    while ((got = fread(tmp, 1, sizeof tmp, f)) > 0) buf.insert(buf.end(), tmp, tmp + got);
    if (f != stdin) fclose(f);
    rc = warpjq_amd_run_csv(query.c_str(), buf.data(), buf.size(), format, on_invalid, chunk_bytes, max_line_bytes,
                            &out, &out_len, &st, err, sizeof err);
  } else {
    rc = warpjq_amd_run_paths(query.c_str(), files.data(), (uint32_t)files.size(), format, on_invalid, chunk_bytes,
                              max_line_bytes, &out, &out_len, &st, err, sizeof err);
  }
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  const std::string name = files.empty() ? "<stdin>" : files.size() == 1 ? files[0]
                                                        : std::string(files[0]) + " (+" +
                                                              std::to_string(files.size() - 1) + " more)";
  if (rc) {
    if (rc == 1) {
      fprintf(stderr, "warpjq-amd: %s\n", err);
      return 2;
    }
    if (rc == 2)  // --strict (error.rs: MalformedLine)
      fprintf(stderr,
              "warpjq-amd: malformed JSON on line %llu of %s: %s\n  help: pass --skip-invalid to drop bad lines "
              "instead of stopping\n",
              (unsigned long long)st.err_line, name.c_str(), err);
    else if (rc == 4)
      fprintf(stderr, "warpjq-amd: line %llu of %s: %s\n", (unsigned long long)st.err_line, name.c_str(), err);
    else
      fprintf(stderr, "warpjq-amd: %s\n", err);
    return 1;
  }
  if (out_len && !write_all(out, out_len)) {
    warpjq_amd_free(out);
    return errno == EPIPE ? 0 : 1;  // a reader that went away is not an error (main.rs:170-176)
  }
  warpjq_amd_free(out);

  // run.rs:104-150
  if (st.malformed > 0 && on_invalid != 1)
    fprintf(stderr, "warpjq-amd: skipped %llu malformed line%s (pass --strict to stop on the first one)\n",
            (unsigned long long)st.malformed, st.malformed == 1 ? "" : "s");
  if (st.type_errors > 0 && on_invalid != 1)
    fprintf(stderr,
            "warpjq-amd: skipped %llu line%s where a path hit a value of the wrong type (jq reports these as "
            "\"cannot index\" errors)\n",
            (unsigned long long)st.type_errors, st.type_errors == 1 ? "" : "s");
  if (stats) {
    const double gbps = secs > 0 ? (double)st.bytes_in / secs / 1e9 : 0.0;
    fprintf(stderr, "warpjq-amd: gpu | %s in %.3fs = %.2f GB/s | %llu lines read, %llu rows out\n",
            human_bytes(st.bytes_in).c_str(), secs, gbps, (unsigned long long)st.lines_in,
            (unsigned long long)st.lines_out);
    fprintf(stderr, "warpjq-amd: %llu chunks, H2D %.1f ms, kernels %.1f ms\n", (unsigned long long)st.chunks,
            st.h2d_ms, st.kernel_ms);
    if (st.gpu_redone_chunks)
      fprintf(stderr, "warpjq-amd: %llu chunks were split and resubmitted: past a device limit\n",
              (unsigned long long)st.gpu_redone_chunks);
  }
  return (st.lines_out == 0 && !aggregate) ? 1 : 0;
}
