// flock_harness WORKER SCRATCH SCORE SUMMARY LOG2 THREADS WARMUP_RUNS RUNS
//
// Drives the challenge's worker protocol and scores it.  The interface is
// the challenge's (benchmark-tools/harness/src/main.rs): the command line,
// the worker's arguments `LOG2 READY PROOF` with one decimal seed line on its
// stdin, RAYON_NUM_THREADS and TMPDIR in its environment, publication of the
// proof by rename, the 300 s / 900 s limits, the 500,000-byte bound, the
// score (2^LOG2 / median measured trial), the percentile rule (linear
// between order statistics at rank (n - 1) q), the trial and score lines on
// stdout, and the fields of SCORE (JSON) and SUMMARY (Markdown).
//
// How it runs a trial is this program's own:
//  - It waits on events, not on a timer: an inotify watch on SCRATCH wakes
//    it when the worker creates or renames a file there, and a pidfd when
//    the worker ends, so the clock stops within microseconds of the rename
//    instead of at the next poll.  A 20 ms tick re-checks in any case.
//  - The published file passes a fixed list of gates (opened without
//    following links, regular file, 1..500,000 bytes, header "FLOCK" 04 02,
//    the same length before and after reading) before the trusted copy is
//    written and handed to flock_verify_proof, a separate program next to
//    this one (commitment root recomputed from the seed on the GPU, full
//    verifier), so this process never opens the GPU.
//  - It keeps the caller's environment for the worker (the GPU runtime's
//    settings live there) and has no macOS sandbox profile argument.
#include <errno.h>
#include <fcntl.h>
#include <ftw.h>
#include <limits.h>
#include <poll.h>
#include <signal.h>
#include <spawn.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/inotify.h>
#include <sys/stat.h>
#include <sys/syscall.h>
#include <sys/wait.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

extern char **environ;

namespace {

using Clock = std::chrono::steady_clock;

constexpr double kReadyLimitS = 300.0;
constexpr double kProofLimitS = 900.0;
constexpr size_t kProofLimitBytes = 500000;
constexpr int kTickMs = 20;

struct Failure : std::runtime_error {
  using std::runtime_error::runtime_error;
};

[[noreturn]] void fail(const std::string &what) { throw Failure(what); }

std::string errstr(int e) { return strerror(e); }

// ---------------------------------------------------------------- numbers

bool decimal(const char *s, uint64_t *out) {
  uint64_t v = 0;
  if (!*s) return false;
  for (; *s; ++s) {
    if (*s < '0' || *s > '9') return false;
    if (v > (UINT64_MAX - (uint64_t)(*s - '0')) / 10) return false;
    v = v * 10 + (uint64_t)(*s - '0');
  }
  *out = v;
  return true;
}

// Order statistic at fractional rank (n - 1) q, linearly interpolated.
double quantile(std::vector<double> xs, double q) {
  std::sort(xs.begin(), xs.end());
  const double r = (double)(xs.size() - 1) * q;
  const size_t a = (size_t)r, b = std::min(a + 1, xs.size() - 1);
  return xs[a] + (xs[b] - xs[a]) * (r - (double)a);
}

uint64_t urandom_u64() {
  uint64_t v = 0;
  const int fd = open("/dev/urandom", O_RDONLY | O_CLOEXEC);
  const bool ok = fd >= 0 && read(fd, &v, 8) == 8;  // little-endian host: bytes as a u64le
  if (fd >= 0) close(fd);
  if (!ok) fail("no seed: /dev/urandom is unreadable");
  return v;
}

// ---------------------------------------------------------------- scratch

int unlink_entry(const char *p, const struct stat *, int, struct FTW *) { return remove(p); }

// SCRATCH: emptied and recreated before and after every trial.
struct Scratch {
  std::string dir;

  void renew() const {
    if (dir.empty() || dir == "/") fail("scratch path '" + dir + "' is not allowed");
    struct stat st;
    if (lstat(dir.c_str(), &st) == 0 && nftw(dir.c_str(), unlink_entry, 32, FTW_DEPTH | FTW_PHYS) != 0)
      fail("cannot empty " + dir + ": " + errstr(errno));
    for (size_t i = 1; i <= dir.size(); ++i)
      if (i == dir.size() || dir[i] == '/') {
        const std::string p = dir.substr(0, i);
        if (mkdir(p.c_str(), 0755) != 0 && errno != EEXIST) fail("cannot create " + p + ": " + errstr(errno));
      }
  }
  std::string path(int run, const char *kind) const { return dir + "/run-" + std::to_string(run) + "." + kind; }
};

// ---------------------------------------------------------------- processes

std::string how_it_ended(int status) {
  if (WIFEXITED(status)) return "exit status " + std::to_string(WEXITSTATUS(status));
  if (WIFSIGNALED(status)) return "signal " + std::to_string(WTERMSIG(status));
  return "wait status " + std::to_string(status);
}

class Process {
 public:
  // argv[0] with the given environment; stdin a pipe or /dev/null, stdout
  // and stderr kept or sent to /dev/null.
  Process(const std::vector<std::string> &argv, const std::vector<std::string> &env, bool pipe_stdin, bool quiet) {
    std::vector<char *> av, ev;
    for (auto &a : argv) av.push_back(const_cast<char *>(a.c_str()));
    av.push_back(nullptr);
    for (auto &e : env) ev.push_back(const_cast<char *>(e.c_str()));
    ev.push_back(nullptr);
    posix_spawn_file_actions_t fa;
    posix_spawn_file_actions_init(&fa);
    int p[2] = {-1, -1};
    if (pipe_stdin) {
      if (pipe2(p, O_CLOEXEC) != 0) fail("pipe: " + errstr(errno));
      posix_spawn_file_actions_adddup2(&fa, p[0], 0);
    } else {
      posix_spawn_file_actions_addopen(&fa, 0, "/dev/null", O_RDONLY, 0);
    }
    if (quiet) {
      posix_spawn_file_actions_addopen(&fa, 1, "/dev/null", O_WRONLY, 0);
      posix_spawn_file_actions_addopen(&fa, 2, "/dev/null", O_WRONLY, 0);
    }
    const int rc = posix_spawn(&pid_, av[0], &fa, nullptr, av.data(), ev.data());
    posix_spawn_file_actions_destroy(&fa);
    if (p[0] >= 0) close(p[0]);
    if (rc != 0) {
      if (p[1] >= 0) close(p[1]);
      pid_ = -1;
      fail("cannot start " + argv[0] + ": " + errstr(rc));
    }
    in_ = p[1];
    // a descriptor that polls readable when the process ends (Linux >= 5.3);
    // without it the event loop falls back to its tick
    pidfd_ = (int)syscall(SYS_pidfd_open, pid_, 0);
  }
  ~Process() { kill_and_reap(); }
  Process(const Process &) = delete;
  Process &operator=(const Process &) = delete;

  int pidfd() const { return pidfd_; }
  bool ended() {
    if (!reaped_ && pid_ > 0 && waitpid(pid_, &status_, WNOHANG) == pid_) reaped_ = true;
    return reaped_;
  }
  int status() const { return status_; }
  void send_line(const std::string &line) {
    if (in_ < 0 || write(in_, line.data(), line.size()) != (ssize_t)line.size())
      fail("the seed could not be sent to the worker");
  }
  int wait() {
    while (!reaped_ && waitpid(pid_, &status_, 0) < 0)
      if (errno != EINTR) fail("lost track of process " + std::to_string(pid_));
    reaped_ = true;
    return status_;
  }
  void kill_and_reap() {
    if (in_ >= 0) close(in_);
    in_ = -1;
    if (pid_ > 0 && !ended()) {
      ::kill(pid_, SIGKILL);
      while (waitpid(pid_, &status_, 0) < 0 && errno == EINTR) {
      }
      reaped_ = true;
    }
    if (pidfd_ >= 0) close(pidfd_);
    pidfd_ = -1;
  }

 private:
  pid_t pid_ = -1;
  int in_ = -1, pidfd_ = -1, status_ = 0;
  bool reaped_ = false;
};

// ---------------------------------------------------------------- events

// Sleeps until something may have changed: a file created, written or
// renamed into the watched directory, the watched process ending, or the
// tick.  Callers re-check their condition after every wake-up.
class Waiter {
 public:
  explicit Waiter(const std::string &dir) {
    fd_ = inotify_init1(IN_NONBLOCK | IN_CLOEXEC);
    if (fd_ >= 0 && inotify_add_watch(fd_, dir.c_str(), IN_CREATE | IN_CLOSE_WRITE | IN_MOVED_TO) < 0) {
      close(fd_);
      fd_ = -1;
    }
  }
  ~Waiter() {
    if (fd_ >= 0) close(fd_);
  }
  void wait(int pidfd, Clock::time_point deadline) {
    const double left_ms = std::chrono::duration<double, std::milli>(deadline - Clock::now()).count();
    const int ms = (int)std::max(0.0, std::min<double>(kTickMs, std::ceil(left_ms)));
    pollfd fds[2];
    int n = 0;
    if (fd_ >= 0) fds[n++] = {fd_, POLLIN, 0};
    if (pidfd >= 0) fds[n++] = {pidfd, POLLIN, 0};
    if (poll(fds, n, ms) > 0 && fd_ >= 0 && (fds[0].revents & POLLIN)) {
      char buf[4096];
      while (read(fd_, buf, sizeof buf) > 0) {
      }
    }
  }

 private:
  int fd_ = -1;
};

// ---------------------------------------------------------------- proof gates

enum class Gate { Absent, Taken };

// Take the published proof if it is there and passes every gate; throws
// when a published file fails one.
Gate take_published(const std::string &path, std::vector<uint8_t> &out) {
  const int fd = open(path.c_str(), O_RDONLY | O_NOFOLLOW | O_NONBLOCK | O_CLOEXEC);
  if (fd < 0) {
    if (errno == ENOENT) return Gate::Absent;
    if (errno == ELOOP) fail("published proof rejected: it is a symbolic link");
    fail("published proof rejected: cannot open it (" + errstr(errno) + ")");
  }
  struct Fd {
    int fd;
    ~Fd() { close(fd); }
  } guard{fd};
  struct stat before;
  if (fstat(fd, &before) != 0 || !S_ISREG(before.st_mode)) fail("published proof rejected: not a regular file");
  auto check_size = [](long long n) {
    if (n < 1 || n > (long long)kProofLimitBytes)
      fail("published proof rejected: " + std::to_string(n) + " bytes (allowed 1.." +
           std::to_string(kProofLimitBytes) + ")");
  };
  check_size((long long)before.st_size);
  // read one byte past the limit, so a file that grew shows it
  out.assign(kProofLimitBytes + 1, 0);
  size_t got = 0;
  for (ssize_t r; got < out.size(); got += (size_t)r) {
    r = read(fd, out.data() + got, out.size() - got);
    if (r == 0) break;
    if (r < 0) {
      if (errno == EINTR) {
        r = 0;
        continue;
      }
      fail("published proof rejected: read failed (" + errstr(errno) + ")");
    }
  }
  out.resize(got);
  check_size((long long)got);
  struct stat after;
  if (fstat(fd, &after) != 0 || after.st_size != before.st_size || (size_t)after.st_size != got)
    fail("published proof rejected: it changed while being read");
  if (got < 7 || memcmp(out.data(), "FLOCK\x04\x02", 7) != 0)
    fail("published proof rejected: not a FLOCK v4 R1CS bundle");
  return Gate::Taken;
}

// ---------------------------------------------------------------- trials

struct Setup {
  std::string worker, score, summary;
  Scratch scratch;
  uint32_t log2 = 0;
  uint64_t threads = 0, warmup = 0, runs = 0;
  std::string verifier;  // flock_verify_proof next to this program
};

struct Outcome {
  double seconds = 0;
  size_t bytes = 0;
};

std::vector<std::string> worker_environment(const Setup &s) {
  std::vector<std::string> env;
  for (char **e = environ; *e; ++e)
    if (strncmp(*e, "RAYON_NUM_THREADS=", 18) != 0 && strncmp(*e, "TMPDIR=", 7) != 0) env.emplace_back(*e);
  env.push_back("RAYON_NUM_THREADS=" + std::to_string(s.threads));
  env.push_back("TMPDIR=" + s.scratch.dir);
  return env;
}

double seconds_since(Clock::time_point t) { return std::chrono::duration<double>(Clock::now() - t).count(); }

Clock::time_point after(double s) {
  return Clock::now() + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(s));
}

// The trusted copy goes to the verifier program; it must exit 0.
void verify(const Setup &s, int run, uint64_t seed, const std::vector<uint8_t> &proof) {
  const std::string copy = s.scratch.path(run, "trusted");
  const int fd = open(copy.c_str(), O_WRONLY | O_CREAT | O_EXCL | O_NOFOLLOW | O_CLOEXEC, 0600);
  if (fd < 0) fail("cannot write the trusted copy " + copy + ": " + errstr(errno));
  size_t done = 0;
  while (done < proof.size()) {
    const ssize_t w = write(fd, proof.data() + done, proof.size() - done);
    if (w < 0 && errno == EINTR) continue;
    if (w <= 0) {
      close(fd);
      fail("cannot write the trusted copy " + copy);
    }
    done += (size_t)w;
  }
  close(fd);
  std::vector<std::string> env;
  for (char **e = environ; *e; ++e) env.emplace_back(*e);
  Process v({s.verifier, std::to_string(s.log2), std::to_string(seed), copy}, env, false, false);
  const int st = v.wait();
  if (!WIFEXITED(st) || WEXITSTATUS(st) != 0) fail("verifier did not accept the proof (" + how_it_ended(st) + ")");
}

Outcome trial(const Setup &s, int run) {
  s.scratch.renew();
  Waiter waiter(s.scratch.dir);
  const std::string ready = s.scratch.path(run, "ready"), published = s.scratch.path(run, "proof");
  Process w({s.worker, std::to_string(s.log2), ready, published}, worker_environment(s), true, true);

  // readiness: the ready path exists as a regular file
  for (const Clock::time_point limit = after(kReadyLimitS);;) {
    struct stat st;
    if (stat(ready.c_str(), &st) == 0 && S_ISREG(st.st_mode)) break;
    if (w.ended()) fail("worker ended (" + how_it_ended(w.status()) + ") before signalling readiness");
    if (Clock::now() >= limit) fail("worker not ready after " + std::to_string((int)kReadyLimitS) + " s");
    waiter.wait(w.pidfd(), limit);
  }
  const uint64_t seed = urandom_u64();

  // the timed span: seed line written -> proof taken
  std::vector<uint8_t> proof;
  const Clock::time_point t0 = Clock::now();
  w.send_line(std::to_string(seed) + "\n");
  for (const Clock::time_point limit = after(kProofLimitS);;) {
    if (take_published(published, proof) == Gate::Taken) break;
    if (w.ended()) {
      // the rename may become visible after the exit
      if (take_published(published, proof) == Gate::Taken) break;
      fail("worker ended (" + how_it_ended(w.status()) + ") without publishing a proof");
    }
    if (Clock::now() >= limit) fail("no proof published within " + std::to_string((int)kProofLimitS) + " s");
    waiter.wait(w.pidfd(), limit);
  }
  Outcome o{seconds_since(t0), proof.size()};
  w.kill_and_reap();  // publication ends the timed span; the worker is not needed after it
  verify(s, run, seed, proof);
  s.scratch.renew();
  return o;
}

// ---------------------------------------------------------------- reports

std::string seconds_list(const std::vector<Outcome> &v) {
  std::string s = "[";
  char b[48];
  for (size_t i = 0; i < v.size(); ++i) {
    snprintf(b, sizeof b, "%s\n      %.17g", i ? "," : "", v[i].seconds);
    s += b;
  }
  return s + (v.empty() ? "]" : "\n    ]");
}

void write_file(const std::string &path, const std::string &text) {
  FILE *f = fopen(path.c_str(), "w");
  const bool ok = f && fwrite(text.data(), 1, text.size(), f) == text.size();
  if ((f && fclose(f) != 0) || !ok) fail("cannot write " + path);
}

void report(const Setup &s, const std::vector<Outcome> &warm, const std::vector<Outcome> &meas) {
  const uint64_t batch = 1ull << s.log2;
  std::vector<double> t;
  double sum = 0;
  for (auto &o : meas) {
    if (!std::isfinite(o.seconds) || o.seconds <= 0) fail("a trial took no measurable time");
    t.push_back(o.seconds);
    sum += o.seconds;
  }
  const double p10 = quantile(t, 0.10), p50 = quantile(t, 0.50), p90 = quantile(t, 0.90);
  const double score = (double)batch / p50, aggregate = (double)batch * (double)t.size() / sum;
  size_t largest = 0;
  for (auto *v : {&warm, &meas})
    for (auto &o : *v) largest = std::max(largest, o.bytes);
  char b[2048];
  snprintf(b, sizeof b,
           "{\n  \"score\": %.17g,\n  \"metrics\": {\n    \"warmup_trial_seconds\": %s,\n"
           "    \"trial_seconds\": %s,\n    \"p10_seconds\": %.17g,\n    \"median_seconds\": %.17g,\n"
           "    \"aggregate_compressions_per_second\": %.17g,\n    \"p90_p10_latency_ratio\": %.17g,\n"
           "    \"batch_size\": %llu,\n    \"warmup_runs\": %zu,\n    \"measured_runs\": %zu,\n"
           "    \"threads\": %llu,\n    \"proof_bytes\": %zu,\n    \"verified\": true\n  }\n}\n",
           score, seconds_list(warm).c_str(), seconds_list(meas).c_str(), p10, p50, aggregate, p90 / p10,
           (unsigned long long)batch, warm.size(), meas.size(), (unsigned long long)s.threads, largest);
  write_file(s.score, b);
  snprintf(b, sizeof b,
           "# Flock BLAKE3 benchmark\n\n- Batch: `%llu`\n- Threads: `%llu`\n"
           "- Warm-up trials: `%zu` (verified, not scored)\n- Measured trials: `%zu`\n"
           "- Score: **%.3f compressions/second**\n- P10 latency: `%.9f s`\n- Median latency: `%.9f s`\n"
           "- Aggregate throughput: `%.3f compressions/second`\n- P90/P10 latency ratio: `%.6f`\n"
           "- Largest proof: `%zu` bytes\n- Trusted verification: `passed`\n",
           (unsigned long long)batch, (unsigned long long)s.threads, warm.size(), meas.size(), score, p10, p50,
           aggregate, p90 / p10, largest);
  write_file(s.summary, b);
  printf("score=%.3f compressions_per_second\n", score);
}

void trial_line(const char *phase, uint64_t i, uint64_t n, const Outcome &o, uint64_t batch) {
  printf("%s_trial=%llu/%llu trial_score=%.3f compressions_per_second seconds=%.9f verified=true "
         "included_in_score=%s\n",
         phase, (unsigned long long)i, (unsigned long long)n, (double)batch / o.seconds, o.seconds,
         strcmp(phase, "measured") == 0 ? "true" : "false");
  fflush(stdout);
}

std::string own_directory() {
  char b[PATH_MAX];
  const ssize_t n = readlink("/proc/self/exe", b, sizeof b - 1);
  if (n <= 0) fail("cannot find this program's directory");
  b[n] = 0;
  std::string p(b);
  return p.substr(0, p.rfind('/'));
}

}  // namespace

int main(int argc, char **argv) {
  Setup s;
  uint64_t log2 = 0;
  const bool ok = argc == 9 && decimal(argv[5], &log2) && decimal(argv[6], &s.threads) &&
                  decimal(argv[7], &s.warmup) && decimal(argv[8], &s.runs) && log2 >= 8 && log2 <= 20 &&
                  s.threads > 0 && s.runs > 0;
  if (!ok) {
    fprintf(stderr, "usage: flock_harness WORKER SCRATCH SCORE SUMMARY LOG2 THREADS WARMUP_RUNS RUNS\n"
                    "  (8 <= LOG2 <= 20, THREADS >= 1, RUNS >= 1)\n");
    return 2;
  }
  s.worker = argv[1];
  s.scratch.dir = argv[2];
  s.score = argv[3];
  s.summary = argv[4];
  s.log2 = (uint32_t)log2;
  signal(SIGPIPE, SIG_IGN);  // a dead worker turns the seed write into an error
  try {
    s.verifier = own_directory() + "/flock_verify_proof";
    std::vector<Outcome> warm, meas;
    for (uint64_t run = 1; run <= s.warmup + s.runs; ++run) {
      const Outcome o = trial(s, (int)run);
      const bool warming = run <= s.warmup;
      trial_line(warming ? "warmup" : "measured", warming ? run : run - s.warmup, warming ? s.warmup : s.runs, o,
                 1ull << s.log2);
      (warming ? warm : meas).push_back(o);
    }
    report(s, warm, meas);
  } catch (const std::exception &e) {
    fprintf(stderr, "flock_harness: %s\n", e.what());
    return 1;
  }
  return 0;
}
