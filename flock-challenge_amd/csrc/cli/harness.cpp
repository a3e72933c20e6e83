// flock_harness WORKER SCRATCH SCORE SUMMARY LOG2 THREADS WARMUP_RUNS RUNS
//
// The benchmark driver of the challenge (benchmark-tools/harness/src/main.rs):
// for each of WARMUP_RUNS + RUNS trials it
//  - empties SCRATCH and starts a fresh WORKER process
//    (`WORKER LOG2 SCRATCH/run-N.ready SCRATCH/run-N.proof`, stdin a pipe,
//    stdout and stderr discarded, RAYON_NUM_THREADS=THREADS and TMPDIR=SCRATCH
//    set in its environment);
//  - waits (up to 300 s) for the ready file, draws a seed from /dev/urandom,
//    starts the clock and writes the seed line to the worker;
//  - waits (up to 900 s) for the published proof, which it opens without
//    following symlinks and reads only if it is a regular file of 1..500,000
//    bytes that does not change while read; publication, not process exit,
//    stops the clock;
//  - kills the worker, then accepts the captured copy only if its bundle
//    header is "FLOCK" 04 02 and flock_verify_proof (next to this program)
//    accepts it: commitment root recomputed from the seed, full verifier.
// The score is 2^LOG2 / median trial seconds; SCORE receives the JSON metrics
// and SUMMARY a Markdown summary, with the same fields as the reference's
// (main.rs:48-67, 382-436).  Percentiles interpolate linearly between order
// statistics (main.rs:127-148).
//
// This program never opens the GPU itself: workers and the verifier are
// separate processes.  Unlike the reference it keeps the caller's environment
// for the worker (the GPU runtime's settings live there) and has no macOS
// sandbox profile argument.
#include <errno.h>
#include <fcntl.h>
#include <ftw.h>
#include <limits.h>
#include <signal.h>
#include <spawn.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/wait.h>
#include <time.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

extern char **environ;

namespace {

const double kStartupTimeout = 300.0;  // seconds (main.rs:20-21)
const double kRunTimeout = 900.0;
const uint64_t kMaxProofBytes = 500000;  // main.rs:27

using Clock = std::chrono::steady_clock;

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

struct Config {
  std::string worker, scratch, score, summary;
  uint32_t log2 = 0;
  uint64_t threads = 0, warmup_runs = 0, runs = 0;
};

struct Trial {
  double seconds;
  size_t proof_bytes;
};

bool parse_u64(const char *s, uint64_t *out) {
  if (!*s) return false;
  uint64_t v = 0;
  for (; *s; ++s) {
    if (*s < '0' || *s > '9') return false;
    const uint64_t d = (uint64_t)(*s - '0');
    if (v > (UINT64_MAX - d) / 10) return false;
    v = v * 10 + d;
  }
  *out = v;
  return true;
}

// Linear interpolation between the order statistics at rank (n - 1) p.
double percentile_seconds(const std::vector<Trial> &trials, double p) {
  if (trials.empty() || !(p >= 0.0 && p <= 1.0)) throw Error("invalid percentile input");
  std::vector<double> s;
  for (const Trial &t : trials) {
    if (!std::isfinite(t.seconds) || t.seconds <= 0.0) throw Error("trial duration must be finite and positive");
    s.push_back(t.seconds);
  }
  std::sort(s.begin(), s.end());
  const double rank = (double)(s.size() - 1) * p;
  const size_t lo = (size_t)std::floor(rank), hi = (size_t)std::ceil(rank);
  return s[lo] + (s[hi] - s[lo]) * (rank - (double)lo);
}

int rm_entry(const char *path, const struct stat *, int, struct FTW *) { return remove(path); }

void reset_scratch(const std::string &path) {
  if (path.empty() || path == "/") throw Error("refusing unsafe scratch path");
  struct stat st;
  if (lstat(path.c_str(), &st) == 0 && nftw(path.c_str(), rm_entry, 64, FTW_DEPTH | FTW_PHYS) != 0)
    throw Error("cannot empty scratch " + path);
  std::string prefix;
  for (size_t i = 0; i <= path.size(); ++i) {
    if (i == path.size() || (path[i] == '/' && i > 0)) {
      prefix = path.substr(0, i);
      if (mkdir(prefix.c_str(), 0755) != 0 && errno != EEXIST) throw Error("cannot create " + prefix);
    }
  }
}

bool is_file(const std::string &path) {
  struct stat st;
  return stat(path.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

void nap() { std::this_thread::sleep_for(std::chrono::microseconds(100)); }

struct Child {
  pid_t pid = -1;
  int stdin_fd = -1;
  int status = 0;
  bool exited = false;
  // true once the process has ended (its status is then in `status`)
  bool poll() {
    if (exited) return true;
    const pid_t r = waitpid(pid, &status, WNOHANG);
    if (r == pid) exited = true;
    return exited;
  }
  void stop() {
    if (stdin_fd >= 0) close(stdin_fd);
    stdin_fd = -1;
    if (pid > 0 && !exited) {
      kill(pid, SIGKILL);
      while (waitpid(pid, &status, 0) < 0 && errno == EINTR) {
      }
      exited = true;
    }
  }
  std::string describe() const {
    char b[64];
    if (WIFEXITED(status))
      snprintf(b, sizeof(b), "exit status: %d", WEXITSTATUS(status));
    else if (WIFSIGNALED(status))
      snprintf(b, sizeof(b), "signal: %d", WTERMSIG(status));
    else
      snprintf(b, sizeof(b), "status %d", status);
    return b;
  }
};

// Start argv[0] with stdin a pipe (when want_stdin) and stdout/stderr
// discarded (when quiet), in the given environment.
Child spawn(const std::vector<std::string> &args, const std::vector<std::string> &env, bool want_stdin, bool quiet) {
  std::vector<char *> argv, envp;
  for (const std::string &a : args) argv.push_back((char *)a.c_str());
  argv.push_back(nullptr);
  for (const std::string &e : env) envp.push_back((char *)e.c_str());
  envp.push_back(nullptr);
  posix_spawn_file_actions_t fa;
  posix_spawn_file_actions_init(&fa);
  int pipefd[2] = {-1, -1};
  if (want_stdin) {
    if (pipe2(pipefd, O_CLOEXEC) != 0) throw Error("pipe failed");
    posix_spawn_file_actions_adddup2(&fa, pipefd[0], 0);
  } else {
    posix_spawn_file_actions_addopen(&fa, 0, "/dev/null", O_RDONLY, 0);
  }
  if (quiet) {
    posix_spawn_file_actions_addopen(&fa, 1, "/dev/null", O_WRONLY, 0);
    posix_spawn_file_actions_addopen(&fa, 2, "/dev/null", O_WRONLY, 0);
  }
  Child c;
  const int rc = posix_spawn(&c.pid, argv[0], &fa, nullptr, argv.data(), envp.data());
  posix_spawn_file_actions_destroy(&fa);
  if (want_stdin) close(pipefd[0]);
  if (rc != 0) {
    if (want_stdin) close(pipefd[1]);
    throw Error("cannot start " + args[0] + ": " + strerror(rc));
  }
  c.stdin_fd = want_stdin ? pipefd[1] : -1;
  return c;
}

// The caller's environment with RAYON_NUM_THREADS and TMPDIR replaced.
std::vector<std::string> worker_env(const Config &cfg) {
  std::vector<std::string> env;
  for (char **e = environ; *e; ++e) {
    if (!strncmp(*e, "RAYON_NUM_THREADS=", 18) || !strncmp(*e, "TMPDIR=", 7)) continue;
    env.push_back(*e);
  }
  env.push_back("RAYON_NUM_THREADS=" + std::to_string(cfg.threads));
  env.push_back("TMPDIR=" + cfg.scratch);
  return env;
}

// The published proof's bytes, or nothing when it is not there yet.
bool capture_proof_if_ready(const std::string &path, std::vector<uint8_t> *out) {
  const int fd = open(path.c_str(), O_RDONLY | O_NONBLOCK | O_NOFOLLOW | O_CLOEXEC);
  if (fd < 0) {
    if (errno == ENOENT) return false;
    throw Error(std::string("cannot open final proof: ") + strerror(errno));
  }
  struct Closer {
    int fd;
    ~Closer() { close(fd); }
  } closer{fd};
  struct stat st0;
  if (fstat(fd, &st0) != 0 || !S_ISREG(st0.st_mode)) throw Error("final proof path is not a regular file");
  if (st0.st_size == 0 || (uint64_t)st0.st_size > kMaxProofBytes)
    throw Error("proof size " + std::to_string((long long)st0.st_size) + " is outside the allowed range");
  std::vector<uint8_t> bytes(kMaxProofBytes + 1);
  size_t got = 0;
  while (got < bytes.size()) {
    const ssize_t r = read(fd, bytes.data() + got, bytes.size() - got);
    if (r < 0) {
      if (errno == EINTR) continue;
      throw Error(std::string("cannot read final proof: ") + strerror(errno));
    }
    if (r == 0) break;
    got += (size_t)r;
  }
  if (got == 0 || got > kMaxProofBytes)
    throw Error("captured proof size " + std::to_string(got) + " is outside the allowed range");
  struct stat st1;
  if (fstat(fd, &st1) != 0 || st1.st_size != st0.st_size || (size_t)st1.st_size != got)
    throw Error("proof file changed while the harness captured it");
  bytes.resize(got);
  *out = std::move(bytes);
  return true;
}

uint64_t fresh_seed() {
  uint8_t b[8];
  FILE *f = fopen("/dev/urandom", "rb");
  if (!f || fread(b, 1, 8, f) != 8) {
    if (f) fclose(f);
    throw Error("cannot read /dev/urandom");
  }
  fclose(f);
  uint64_t s = 0;
  for (int i = 7; i >= 0; --i) s = (s << 8) | b[i];  // little-endian
  return s;
}

std::string dir_of_self() {
  char buf[PATH_MAX];
  const ssize_t n = readlink("/proc/self/exe", buf, sizeof(buf) - 1);
  if (n <= 0) throw Error("cannot locate this program");
  buf[n] = 0;
  std::string p(buf);
  return p.substr(0, p.rfind('/'));
}

void verify_captured(const Config &cfg, uint64_t seed, const std::vector<uint8_t> &bytes, int run) {
  if (bytes.empty() || bytes.size() > kMaxProofBytes)
    throw Error("proof size " + std::to_string(bytes.size()) + " is outside the allowed range");
  if (bytes.size() < 7 || memcmp(bytes.data(), "FLOCK", 5) != 0 || bytes[5] != 4 || bytes[6] != 2)
    throw Error("invalid FLOCK R1CS proof header");
  // the harness's own immutable copy, handed to the verifier process
  const std::string copy = cfg.scratch + "/run-" + std::to_string(run) + ".captured";
  const int fd = open(copy.c_str(), O_WRONLY | O_CREAT | O_EXCL | O_CLOEXEC | O_NOFOLLOW, 0600);
  if (fd < 0) throw Error("cannot write " + copy);
  size_t off = 0;
  while (off < bytes.size()) {
    const ssize_t w = write(fd, bytes.data() + off, bytes.size() - off);
    if (w < 0 && errno == EINTR) continue;
    if (w <= 0) {
      close(fd);
      throw Error("cannot write " + copy);
    }
    off += (size_t)w;
  }
  close(fd);
  std::vector<std::string> env;
  for (char **e = environ; *e; ++e) env.push_back(*e);
  Child v = spawn({dir_of_self() + "/flock_verify_proof", std::to_string(cfg.log2), std::to_string(seed), copy}, env,
                  false, false);
  while (waitpid(v.pid, &v.status, 0) < 0) {
    if (errno != EINTR) throw Error("lost the verifier process");
  }
  v.exited = true;
  if (!WIFEXITED(v.status) || WEXITSTATUS(v.status) != 0)
    throw Error("trusted verifier rejected proof (" + v.describe() + ")");
}

double since(Clock::time_point t0) { return std::chrono::duration<double>(Clock::now() - t0).count(); }

Trial run_trial(const Config &cfg, int run) {
  reset_scratch(cfg.scratch);
  const std::string ready = cfg.scratch + "/run-" + std::to_string(run) + ".ready";
  const std::string proof = cfg.scratch + "/run-" + std::to_string(run) + ".proof";
  Child child = spawn({cfg.worker, std::to_string(cfg.log2), ready, proof}, worker_env(cfg), true, true);
  struct Reaper {
    Child &c;
    ~Reaper() { c.stop(); }
  } reaper{child};

  const Clock::time_point t_start = Clock::now();
  for (;;) {
    if (is_file(ready)) break;
    if (child.poll()) throw Error("worker exited before readiness with " + child.describe());
    if (since(t_start) >= kStartupTimeout) throw Error("worker readiness timed out");
    nap();
  }
  const uint64_t seed = fresh_seed();

  const Clock::time_point t0 = Clock::now();
  const std::string line = std::to_string(seed) + "\n";
  if (write(child.stdin_fd, line.data(), line.size()) != (ssize_t)line.size())
    throw Error("cannot send the seed to the worker");
  std::vector<uint8_t> bytes;
  for (;;) {
    if (capture_proof_if_ready(proof, &bytes)) break;
    if (child.poll()) {
      // the rename and the exit may become visible in either order
      if (capture_proof_if_ready(proof, &bytes)) break;
      throw Error("worker exited before publishing a proof with " + child.describe());
    }
    if (since(t0) >= kRunTimeout) throw Error("worker proof timed out");
    nap();
  }
  const double seconds = since(t0);
  // publication is the scored boundary: reap the worker before verifying
  child.stop();
  verify_captured(cfg, seed, bytes, run);
  reset_scratch(cfg.scratch);
  return Trial{seconds, bytes.size()};
}

std::string json_list(const std::vector<Trial> &t) {
  std::string s = "[";
  for (size_t i = 0; i < t.size(); ++i) {
    char b[64];
    snprintf(b, sizeof(b), "%s\n      %.17g", i ? "," : "", t[i].seconds);
    s += b;
  }
  s += t.empty() ? "]" : "\n    ]";
  return s;
}

void write_text(const std::string &path, const std::string &text) {
  FILE *f = fopen(path.c_str(), "w");
  if (!f || fwrite(text.data(), 1, text.size(), f) != text.size()) {
    if (f) fclose(f);
    throw Error("cannot write " + path);
  }
  if (fclose(f) != 0) throw Error("cannot write " + path);
}

void write_results(const Config &cfg, const std::vector<Trial> &warm, const std::vector<Trial> &meas,
                   double throughput, uint64_t batch) {
  const double p10 = percentile_seconds(meas, 0.10), med = percentile_seconds(meas, 0.50),
               p90 = percentile_seconds(meas, 0.90);
  double total = 0;
  for (const Trial &t : meas) total += t.seconds;
  const double aggregate = (double)batch * (double)meas.size() / total;
  size_t proof_bytes = 0;
  for (const Trial &t : warm) proof_bytes = std::max(proof_bytes, t.proof_bytes);
  for (const Trial &t : meas) proof_bytes = std::max(proof_bytes, t.proof_bytes);
  char b[4096];
  snprintf(b, sizeof(b),
           "{\n  \"score\": %.17g,\n  \"metrics\": {\n    \"warmup_trial_seconds\": %s,\n"
           "    \"trial_seconds\": %s,\n    \"p10_seconds\": %.17g,\n    \"median_seconds\": %.17g,\n"
           "    \"aggregate_compressions_per_second\": %.17g,\n    \"p90_p10_latency_ratio\": %.17g,\n"
           "    \"batch_size\": %llu,\n    \"warmup_runs\": %zu,\n    \"measured_runs\": %zu,\n"
           "    \"threads\": %llu,\n    \"proof_bytes\": %zu,\n    \"verified\": true\n  }\n}\n",
           throughput, json_list(warm).c_str(), json_list(meas).c_str(), p10, med, aggregate, p90 / p10,
           (unsigned long long)batch, warm.size(), meas.size(), (unsigned long long)cfg.threads, proof_bytes);
  write_text(cfg.score, b);
  snprintf(b, sizeof(b),
           "# Flock BLAKE3 benchmark\n\n- Batch: `%llu`\n- Threads: `%llu`\n"
           "- Warm-up trials: `%zu` (verified, not scored)\n- Measured trials: `%zu`\n"
           "- Score: **%.3f compressions/second**\n- P10 latency: `%.9f s`\n- Median latency: `%.9f s`\n"
           "- Aggregate throughput: `%.3f compressions/second`\n- P90/P10 latency ratio: `%.6f`\n"
           "- Largest proof: `%zu` bytes\n- Trusted verification: `passed`\n",
           (unsigned long long)batch, (unsigned long long)cfg.threads, warm.size(), meas.size(), throughput, p10, med,
           aggregate, p90 / p10, proof_bytes);
  write_text(cfg.summary, b);
}

void log_trial(const char *phase, size_t index, size_t total, const Trial &t, uint64_t batch) {
  printf("%s_trial=%zu/%zu trial_score=%.3f compressions_per_second seconds=%.9f verified=true "
         "included_in_score=%s\n",
         phase, index, total, (double)batch / t.seconds, t.seconds, strcmp(phase, "measured") ? "false" : "true");
  fflush(stdout);
}

}  // namespace

int main(int argc, char **argv) {
  Config cfg;
  uint64_t log2 = 0;
  if (argc != 9 || !parse_u64(argv[5], &log2) || !parse_u64(argv[6], &cfg.threads) ||
      !parse_u64(argv[7], &cfg.warmup_runs) || !parse_u64(argv[8], &cfg.runs) || log2 < 8 || log2 > 20 ||
      cfg.threads == 0 || cfg.runs == 0) {
    fprintf(stderr, "usage: flock_harness WORKER SCRATCH SCORE SUMMARY LOG2 THREADS WARMUP_RUNS RUNS\n");
    return 2;
  }
  cfg.worker = argv[1];
  cfg.scratch = argv[2];
  cfg.score = argv[3];
  cfg.summary = argv[4];
  cfg.log2 = (uint32_t)log2;
  const uint64_t batch = 1ull << log2;
  signal(SIGPIPE, SIG_IGN);  // a worker that died turns the seed write into an error, not a kill
  try {
    std::vector<Trial> warm, meas;
    const uint64_t total = cfg.warmup_runs + cfg.runs;
    for (uint64_t run = 1; run <= total; ++run) {
      const Trial t = run_trial(cfg, (int)run);
      if (run <= cfg.warmup_runs) {
        log_trial("warmup", run, cfg.warmup_runs, t, batch);
        warm.push_back(t);
      } else {
        log_trial("measured", run - cfg.warmup_runs, cfg.runs, t, batch);
        meas.push_back(t);
      }
    }
    const double throughput = (double)batch / percentile_seconds(meas, 0.50);
    write_results(cfg, warm, meas, throughput, batch);
    printf("score=%.3f compressions_per_second\n", throughput);
  } catch (const std::exception &e) {
    fprintf(stderr, "flock_harness: %s\n", e.what());
    return 1;
  }
  return 0;
}
