// The schedule store of the autotuner: which variant of a pass runs for a given shape.
//
// The launchers ask for a knob's value under a key that names the launch shape (transform
// length, lines, images, worlds, ...).  A value pinned with lx_tune_force wins; else a decision
// already taken (this process, or read from the cache file); else, when tuning is on and the
// caller's stream is not being captured, the launcher times its candidates on the real inputs and
// records the fastest (lx_tune_record); else the launcher's default.  Decisions are appended to the
// cache file (LENIAX_AMD_TUNE_CACHE, or lx_tune_cache) as "knob key value" lines, so that a later
// process starts with them.  Every candidate of a knob is a complete implementation of the same
// pass (each has its own parity test), so the choice changes the time, not the meaning.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <map>
#include <mutex>
#include <string>

#include "lx_internal.h"

namespace {

struct Store {
  std::mutex mu;
  std::map<std::string, int> forced;     // knob -> value
  std::map<std::string, int> decided;    // "knob key" -> value
  std::string path;
  bool loaded = false;
  int mode = LX_TUNE_OFF;
  bool mode_from_env = false;
};

Store& store() {
  static Store s;
  return s;
}

// first use: the cache path and the mode from the environment, then the cache's decisions
void load_locked(Store& s) {
  if (s.loaded) return;
  s.loaded = true;
  if (!s.mode_from_env) {
    s.mode_from_env = true;
    if (const char* m = getenv("LENIAX_AMD_TUNE")) s.mode = atoi(m);
  }
  if (s.path.empty()) {
    if (const char* p = getenv("LENIAX_AMD_TUNE_CACHE")) s.path = p;
  }
  if (s.path.empty()) return;
  FILE* f = fopen(s.path.c_str(), "r");
  if (!f) return;
  char knob[128], key[384];
  int v = 0;
  while (fscanf(f, "%127s %383s %d", knob, key, &v) == 3) s.decided[std::string(knob) + " " + key] = v;
  fclose(f);
}

}  // namespace

int lx_tune_mode() {
  Store& s = store();
  std::lock_guard<std::mutex> lock(s.mu);
  load_locked(s);
  return s.mode;
}

int lx_tune_lookup(const char* knob, const char* key, int* value) {
  Store& s = store();
  std::lock_guard<std::mutex> lock(s.mu);
  load_locked(s);
  auto f = s.forced.find(knob);
  if (f != s.forced.end()) {
    *value = f->second;
    return 1;
  }
  if (s.mode == LX_TUNE_OFF) return 0;
  auto d = s.decided.find(std::string(knob) + " " + key);
  if (d == s.decided.end()) return 0;
  *value = d->second;
  return 1;
}

void lx_tune_record(const char* knob, const char* key, int value) {
  Store& s = store();
  std::lock_guard<std::mutex> lock(s.mu);
  load_locked(s);
  s.decided[std::string(knob) + " " + key] = value;
  if (s.path.empty()) return;
  if (FILE* f = fopen(s.path.c_str(), "a")) {
    fprintf(f, "%s %s %d\n", knob, key, value);
    fclose(f);
  }
}

namespace {
thread_local int quiet = 0;
}

LxTuneQuiet::LxTuneQuiet() { ++quiet; }
LxTuneQuiet::~LxTuneQuiet() { --quiet; }

bool lx_tune_should_time(hipStream_t st) {
  if (quiet > 0 || lx_tune_mode() != LX_TUNE_ON) return false;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return false;
  return true;
}

extern "C" int lx_autotune(int32_t mode) {
  if (mode < LX_TUNE_OFF || mode > LX_TUNE_CACHED) return lx_fail(LX_EINVAL, "autotune mode %d unknown", mode);
  Store& s = store();
  std::lock_guard<std::mutex> lock(s.mu);
  load_locked(s);
  const int prev = s.mode;
  s.mode = mode;
  return prev;
}

extern "C" int lx_tune_cache(const char* path) {
  Store& s = store();
  std::lock_guard<std::mutex> lock(s.mu);
  s.path = path ? path : "";
  s.loaded = false;
  s.decided.clear();
  load_locked(s);
  return LX_OK;
}

extern "C" int lx_tune_force(const char* knob, int32_t value) {
  if (!knob) return lx_fail(LX_EINVAL, "knob is NULL");
  static const char* known[] = {"conv_staged", "line_xcd", "lines96_tpr", "fused_probe", "graph_steps"};
  bool ok = false;
  for (const char* k : known) ok = ok || strcmp(k, knob) == 0;
  if (!ok) return lx_fail(LX_EINVAL, "unknown knob '%s'", knob);
  Store& s = store();
  std::lock_guard<std::mutex> lock(s.mu);
  if (value < 0) s.forced.erase(knob);
  else s.forced[knob] = value;
  return LX_OK;
}

extern "C" int lx_tune_report(char* buf, size_t n) {
  Store& s = store();
  std::lock_guard<std::mutex> lock(s.mu);
  load_locked(s);
  std::string out;
  for (const auto& kv : s.forced) out += "forced " + kv.first + " " + std::to_string(kv.second) + "\n";
  for (const auto& kv : s.decided) out += kv.first + " " + std::to_string(kv.second) + "\n";
  if (buf && n) {
    strncpy(buf, out.c_str(), n - 1);
    buf[n - 1] = 0;
  }
  return int(out.size());
}
