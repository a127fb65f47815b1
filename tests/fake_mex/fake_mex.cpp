// fake_mex.cpp -- implementation of the mex.h stand-in plus a C driver API
// (fm_*) through which Python builds mxArrays, calls a gateway and reads the
// outputs.  Test infrastructure only.
#include "mex.h"

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

struct mxArray_tag {
  mxClassID cls = mxDOUBLE_CLASS;
  bool sparse = false;
  bool complex_ = false;
  size_t m = 0, n = 0;
  std::vector<double> pr, pi;
  std::vector<mwIndex> ir, jc;
  std::vector<uint64_t> u64;
};

namespace {
struct MexError : std::runtime_error {
  std::string id;
  MexError(const std::string& i, const std::string& msg) : std::runtime_error(msg), id(i) {}
};
}  // namespace

extern "C" {

size_t mxGetM(const mxArray* a) { return a->m; }
size_t mxGetN(const mxArray* a) { return a->n; }
double* mxGetPr(const mxArray* a) {
  return const_cast<double*>(a->pr.empty() ? nullptr : a->pr.data());
}
double* mxGetPi(const mxArray* a) {
  return a->complex_ ? const_cast<double*>(a->pi.data()) : nullptr;
}
void* mxGetData(const mxArray* a) {
  if (a->cls == mxUINT64_CLASS) return const_cast<uint64_t*>(a->u64.data());
  return mxGetPr(a);
}
mwIndex* mxGetIr(const mxArray* a) { return const_cast<mwIndex*>(a->ir.data()); }
mwIndex* mxGetJc(const mxArray* a) { return const_cast<mwIndex*>(a->jc.data()); }
mwSize mxGetNzmax(const mxArray* a) { return a->pr.size(); }
int mxIsDouble(const mxArray* a) { return a->cls == mxDOUBLE_CLASS; }
int mxIsSparse(const mxArray* a) { return a->sparse; }
int mxIsUint64(const mxArray* a) { return a->cls == mxUINT64_CLASS; }

mxArray* mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c) {
  mxArray* a = new mxArray_tag();
  a->m = m;
  a->n = n;
  a->pr.assign(m * n, 0.0);
  if (c == mxCOMPLEX) {
    a->complex_ = true;
    a->pi.assign(m * n, 0.0);
  }
  return a;
}

mxArray* mxCreateDoubleScalar(double v) {
  mxArray* a = mxCreateDoubleMatrix(1, 1, mxREAL);
  a->pr[0] = v;
  return a;
}

int mxIsEmpty(const mxArray* a) { return a->m == 0 || a->n == 0; }

mxArray* mxCreateNumericArray(mwSize ndim, const mwSize* dims, mxClassID cls, mxComplexity) {
  mxArray* a = new mxArray_tag();
  a->cls = cls;
  a->m = ndim > 0 ? dims[0] : 0;
  a->n = ndim > 1 ? dims[1] : 1;
  if (cls == mxUINT64_CLASS)
    a->u64.assign(a->m * a->n, 0);
  else
    a->pr.assign(a->m * a->n, 0.0);
  return a;
}

void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...) {
  char buf[2048];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  throw MexError(id, buf);
}

// ------------------------------------------------------------ driver API
mxArray* fm_double(size_t m, size_t n, const double* data) {
  mxArray* a = mxCreateDoubleMatrix(m, n, mxREAL);
  if (data) std::memcpy(a->pr.data(), data, sizeof(double) * m * n);
  return a;
}

mxArray* fm_complex(size_t m, size_t n, const double* re, const double* im) {
  mxArray* a = mxCreateDoubleMatrix(m, n, mxCOMPLEX);
  std::memcpy(a->pr.data(), re, sizeof(double) * m * n);
  std::memcpy(a->pi.data(), im, sizeof(double) * m * n);
  return a;
}

mxArray* fm_sparse(size_t m, size_t n, size_t nnz, const double* pr, const int64_t* ir,
                   const int64_t* jc) {
  mxArray* a = new mxArray_tag();
  a->sparse = true;
  a->m = m;
  a->n = n;
  a->pr.assign(pr, pr + nnz);
  a->ir.assign(ir, ir + nnz);
  a->jc.assign(jc, jc + n + 1);
  return a;
}

mxArray* fm_uint64(uint64_t v) {
  mwSize d[2] = {1, 1};
  mxArray* a = mxCreateNumericArray(2, d, mxUINT64_CLASS, mxREAL);
  a->u64[0] = v;
  return a;
}

uint64_t fm_get_uint64(const mxArray* a) { return a->u64.empty() ? 0 : a->u64[0]; }
size_t fm_get_m(const mxArray* a) { return a->m; }
size_t fm_get_n(const mxArray* a) { return a->n; }
void fm_copy_pr(const mxArray* a, double* dst) {
  std::memcpy(dst, a->pr.data(), sizeof(double) * a->pr.size());
}
void fm_destroy(mxArray* a) { delete a; }

typedef void (*gateway_fn)(int, mxArray**, int, const mxArray**);

// Calls a gateway; returns 0 on success, 1 if it raised (id and message copied out).
int fm_call(gateway_fn fn, int nlhs, mxArray** plhs, int nrhs, const mxArray** prhs, char* id,
            size_t idlen, char* msg, size_t msglen) {
  try {
    fn(nlhs, plhs, nrhs, prhs);
    return 0;
  } catch (const MexError& e) {
    std::snprintf(id, idlen, "%s", e.id.c_str());
    std::snprintf(msg, msglen, "%s", e.what());
    return 1;
  }
}

}  // extern "C"
