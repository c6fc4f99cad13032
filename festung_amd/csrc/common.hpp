// Shared definitions of the festung_hip library: status codes, the
// thread-local error message, and small helpers.
#pragma once
#include <cstdint>
#include <cstdio>
#include <cstdarg>
#include <string>
#include <stdexcept>
#include <vector>

#include "festung_hip.h"

namespace festung {

// Thrown inside the library; converted to an int status at the C-ABI.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  throw Error(code, buf);
}

void set_last_error(const std::string& s);

// Run f, translating exceptions into status codes (no exception crosses the C-ABI).
template <class F>
int guard(F&& f) {
  try {
    f();
    set_last_error("");
    return FESTUNG_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_last_error("host allocation failed");
    return FESTUNG_ERR_ALLOC;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return FESTUNG_ERR_INTERNAL;
  }
}

const char* last_error();

// Re-raise the status of a nested C-ABI call made inside guard().
inline void check_call(int rc) {
  if (rc != FESTUNG_OK) throw Error(rc, last_error());
}

inline int nbasis(int p) { return (p + 1) * (p + 2) / 2; }

// Autotuner results (capi.cpp): elements per workgroup per tuning key, for
// the life of the process and through festung_tune_save / festung_tune_load.
struct TuneKey {
  std::string device;  // architecture name and compute-unit count
  int p, ordRK, lim;   // lim: limiter type, -1 without limiter
  int lgK;             // ceil(log2 K)
};
int tune_lookup(const TuneKey& k);  // 0 when the key has no entry
void tune_record(const TuneKey& k, int elemsPerBlock);
void tune_load_env();               // FESTUNG_TUNE_FILE, once per process

}  // namespace festung
