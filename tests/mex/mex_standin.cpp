// Test double of the MEX runtime for festung_amd/mex/festung_mex.cpp (see
// tests/mex/mex.h), plus a small C interface through which the Python tests
// build mxArrays, call mexFunction and read the results (ctypes).
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "mex.h"

struct mxArray_tag {
  mxClassID cls;
  size_t m, n;
  std::vector<double> d;       // double
  std::vector<unsigned char> l;  // logical (0 / 1 bytes, read as mxLogical)
  std::vector<uint64_t> u;     // uint64
  std::string s;               // char
};

namespace {
struct MexError : std::runtime_error {
  std::string id;
  MexError(const std::string& i, const std::string& m) : std::runtime_error(m), id(i) {}
};
int g_lock = 0;
void (*g_exit)(void) = nullptr;
}  // namespace

extern "C" {
bool mxIsChar(const mxArray* a) { return a->cls == mxCHAR_CLASS; }
bool mxIsDouble(const mxArray* a) { return a->cls == mxDOUBLE_CLASS; }
bool mxIsLogical(const mxArray* a) { return a->cls == mxLOGICAL_CLASS; }
bool mxIsUint64(const mxArray* a) { return a->cls == mxUINT64_CLASS; }
bool mxIsComplex(const mxArray*) { return false; }
size_t mxGetM(const mxArray* a) { return a->m; }
size_t mxGetN(const mxArray* a) { return a->n; }
size_t mxGetNumberOfElements(const mxArray* a) { return a->m * a->n; }
double* mxGetPr(const mxArray* a) { return const_cast<double*>(a->d.data()); }
void* mxGetData(const mxArray* a) {
  if (a->cls == mxUINT64_CLASS) return const_cast<uint64_t*>(a->u.data());
  if (a->cls == mxLOGICAL_CLASS) return const_cast<unsigned char*>(a->l.data());
  return const_cast<double*>(a->d.data());
}
mxLogical* mxGetLogicals(const mxArray* a) {
  return reinterpret_cast<mxLogical*>(const_cast<unsigned char*>(a->l.data()));
}
double mxGetScalar(const mxArray* a) {
  if (a->cls == mxDOUBLE_CLASS && !a->d.empty()) return a->d[0];
  if (a->cls == mxUINT64_CLASS && !a->u.empty()) return double(a->u[0]);
  if (a->cls == mxLOGICAL_CLASS && !a->l.empty()) return a->l[0] ? 1.0 : 0.0;
  return 0.0;
}
int mxGetString(const mxArray* a, char* buf, mwSize len) {
  if (a->cls != mxCHAR_CLASS || len == 0) return 1;
  std::snprintf(buf, len, "%s", a->s.c_str());
  return a->s.size() + 1 > len ? 1 : 0;
}
mxArray* mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity) {
  auto* a = new mxArray{mxDOUBLE_CLASS, m, n, {}, {}, {}, {}};
  a->d.assign(m * n, 0.0);
  return a;
}
mxArray* mxCreateDoubleScalar(double v) {
  mxArray* a = mxCreateDoubleMatrix(1, 1, mxREAL);
  a->d[0] = v;
  return a;
}
mxArray* mxCreateLogicalMatrix(mwSize m, mwSize n) {
  auto* a = new mxArray{mxLOGICAL_CLASS, m, n, {}, {}, {}, {}};
  a->l.assign(m * n, 0);
  return a;
}
mxArray* mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID cls, mxComplexity) {
  if (cls == mxDOUBLE_CLASS) return mxCreateDoubleMatrix(m, n, mxREAL);
  if (cls != mxUINT64_CLASS) throw MexError("standin:class", "class not supported by the test double");
  auto* a = new mxArray{mxUINT64_CLASS, m, n, {}, {}, {}, {}};
  a->u.assign(m * n, 0);
  return a;
}
mxArray* mxCreateString(const char* s) { return new mxArray{mxCHAR_CLASS, 1, std::strlen(s), {}, {}, {}, s}; }
void mxDestroyArray(mxArray* a) { delete a; }
void mexLock(void) { ++g_lock; }
void mexUnlock(void) { --g_lock; }
int mexAtExit(void (*fn)(void)) {
  g_exit = fn;
  return 0;
}
void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  throw MexError(id, buf);
}

// ---- interface for the tests ----
mxArray* standin_double(size_t m, size_t n, const double* v) {
  mxArray* a = mxCreateDoubleMatrix(m, n, mxREAL);
  if (v) std::memcpy(a->d.data(), v, m * n * sizeof(double));
  return a;
}
mxArray* standin_logical(size_t m, size_t n, const uint8_t* v) {
  mxArray* a = mxCreateLogicalMatrix(m, n);
  for (size_t i = 0; i < m * n; ++i) a->l[i] = v[i] != 0 ? 1 : 0;
  return a;
}
mxArray* standin_string(const char* s) { return mxCreateString(s); }
int standin_class(const mxArray* a) { return int(a->cls); }
size_t standin_rows(const mxArray* a) { return a->m; }
size_t standin_cols(const mxArray* a) { return a->n; }
void standin_read(const mxArray* a, double* out) {
  const size_t n = a->m * a->n;
  for (size_t i = 0; i < n; ++i)
    out[i] = a->cls == mxDOUBLE_CLASS ? a->d[i] : a->cls == mxLOGICAL_CLASS ? (a->l[i] ? 1.0 : 0.0) : double(a->u[i]);
}
void standin_read_string(const mxArray* a, char* out, size_t len) { std::snprintf(out, len, "%s", a->s.c_str()); }
void standin_free(mxArray* a) { delete a; }
int standin_lock_count(void) { return g_lock; }
void standin_at_exit(void) {
  if (g_exit) g_exit();
}
// mexFunction under the test double: 0 on success, else the error id and
// message in idbuf / msgbuf (what the interpreter would raise)
int standin_call(int nlhs, mxArray** plhs, int nrhs, const mxArray** prhs, char* idbuf, char* msgbuf, size_t len) {
  for (int i = 0; i < nlhs; ++i) plhs[i] = nullptr;
  try {
    mexFunction(nlhs, plhs, nrhs, prhs);
  } catch (const MexError& e) {
    std::snprintf(idbuf, len, "%s", e.id.c_str());
    std::snprintf(msgbuf, len, "%s", e.what());
    return 1;
  }
  return 0;
}
}
