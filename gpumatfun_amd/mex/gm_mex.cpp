// gm_mex.cpp -- MATLAB MEX gateways over the C ABI (include/gpumatfun.h).
//
// One source, compiled once per gateway with -DGM_GATEWAY=<n>, gives the
// reference's MEX binaries with the reference's names and calling
// conventions (makefile:27-93, lanczosEngine.m:9-55):
//
//   1 lanczos_create          26 inputs -> uint64 handle   (lanczos_create.cpp)
//   2 lanczos_delete          handle                      (lanczos_delete.cpp)
//   3 build_subspace_mex      handle                      (build_subspace.cpp)
//   4 lin_sys_solves_mex      handle                      (lin_sys_solves.cpp)
//   5 apply_V_mex             handle                      (apply_V.cpp)
//   6 update_U_mex            handle                      (update_U.cpp)
//   7 update_U_monodomain_mex handle                      (update_U_monodomain.cpp)
//   8 get_U_mex               handle -> N x 2             (get_U.cpp)
//   9 get_xsol_mex            handle -> N x 2             (get_xsol.cpp)
//  10 get_aabb_mex            handle -> (m+1) x 4         (called at lanczosEngine.m:26)
//  11 set_y_mex               handle, yreal, yimag        (called at lanczosEngine.m:30)
//  12 eigs_smallest_mex       A, k [, opts] -> Q, lambda  (no reference MEX: replaces
//                             [Q, lambda] = eigs(K, M, k, 'sm') and Q = M^(1/2) Q of
//                             process_nonuniform.m:106-110, on A = M^(-1/2) K M^(-1/2))
//  13 eigs_shift_invert_mex  A, k [, sigma [, opts]] -> Q, lambda  (the same replacement
//                             through the host sparse factorisation, for meshes whose
//                             lambda_max / lambda_k is too large for the GPU filter)
//  14 lambda_max_mex         A -> lambda_max  (replaces eigs(K, M, 1, 'lm') of
//                             process_nonuniform.m:40)
//
// Argument errors use the reference's message identifiers; a non-OK status
// of the library becomes mexErrMsgIdAndTxt("gpumatfun:<status>", message).
#include <mex.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "gpumatfun.h"

#ifndef GM_GATEWAY
#error "compile with -DGM_GATEWAY=<1..14>"
#endif

namespace {

const char* status_name(gm_status s) {
  switch (s) {
    case GM_OK: return "GM_OK";
    case GM_EINVAL: return "GM_EINVAL";
    case GM_ENOMEM: return "GM_ENOMEM";
    case GM_EDEVICE: return "GM_EDEVICE";
    case GM_EBREAKDOWN: return "GM_EBREAKDOWN";
    case GM_EBADHANDLE: return "GM_EBADHANDLE";
    case GM_ESTATE: return "GM_ESTATE";
  }
  return "GM_UNKNOWN";
}

void check(gm_status s, const char* (*last_error)() = gm_last_error) {
  if (s != GM_OK) {
    std::string id = std::string("gpumatfun:") + status_name(s);
    mexErrMsgIdAndTxt(id.c_str(), "%s", last_error());
  }
}

gm_solver* handle_arg(const mxArray* a, const char* gw) {
  if (!mxIsUint64(a) || mxGetM(a) * mxGetN(a) != 1) {
    std::string id = std::string("MATLAB:") + gw + ":inputNotHandle";
    mexErrMsgIdAndTxt(id.c_str(), "First input argument must be a handle.");
  }
  uint64_t v;
  std::memcpy(&v, mxGetData(a), sizeof(v));
  return reinterpret_cast<gm_solver*>(static_cast<uintptr_t>(v));
}

void nargs(int nlhs, int want_lhs, int nrhs, int want_rhs, const char* gw) {
  if (nrhs != want_rhs) {
    std::string id = std::string("MATLAB:") + gw + ":invalidNumInputs";
    mexErrMsgIdAndTxt(id.c_str(), want_rhs == 1 ? "Handle input argument required."
                                               : "Wrong number of input arguments.");
  }
  if (nlhs != want_lhs) {
    std::string id = std::string("MATLAB:") + gw + ":maxlhs";
    mexErrMsgIdAndTxt(id.c_str(), want_lhs == 0 ? "No output arguments are required."
                                               : "One output argument is required.");
  }
}

#if GM_GATEWAY == 1
// Real dense double argument; returns its data and element count.
const double* real_arg(const mxArray* a, const char* badid, const char* badmsg, size_t* count) {
  if (!mxIsDouble(a) || mxIsSparse(a)) mexErrMsgIdAndTxt(badid, "%s", badmsg);
  if (mxGetPi(a) != nullptr)
    mexErrMsgIdAndTxt("MATLAB:lanczos_create:inputNotReal", "Input argument must be a real matrix.");
  if (count) *count = mxGetM(a) * mxGetN(a);
  return mxGetPr(a);
}

double scalar_arg(const mxArray* a, const char* name) {
  std::string id = std::string("MATLAB:lanczos_create:bad_") + name;
  std::string msg = std::string("Invalid ") + name + " input argument.";
  size_t c = 0;
  const double* p = real_arg(a, id.c_str(), msg.c_str(), &c);
  if (c != 1) mexErrMsgIdAndTxt(id.c_str(), "%s", msg.c_str());
  return p[0];
}
#endif

}  // namespace

extern "C" void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
#if GM_GATEWAY == 1
  // lanczos_create(U, A, Q, lambda, N, gamma, gamma_s_real, gamma_s_imag, exp_lambda,
  //                phi_lambda, dt, ode, aode, bode, code, vrest, vamp, vth, vpeak, c1ion,
  //                c2ion, mass, wts1, wts2, shifts, valvec)   lanczos_create.cpp:7-32
  if (nrhs != 26)
    mexErrMsgIdAndTxt("MATLAB:lanczos_create:invalidNumInputs",
                      "Twenty-five input arguments are required.");
  if (nlhs != 1)
    mexErrMsgIdAndTxt("MATLAB:lanczos_create:maxlhs", "One output argument is required.");
  gm_create_desc d;
  std::memset(&d, 0, sizeof(d));
  size_t c = 0;
  const mxArray* A = prhs[1];
  d.U = real_arg(prhs[0], "MATLAB:lanczos_create:badU", "Invalid U input argument.", &c);
  const size_t n = mxGetM(prhs[0]);
  if (mxGetN(prhs[0]) != 2)
    mexErrMsgIdAndTxt("MATLAB:lanczos_create:badU", "Invalid U input argument.");
  if (!mxIsSparse(A))
    mexErrMsgIdAndTxt("MATLAB:lanczos_create:inputNotSparse",
                      "Invalid sparse matrix input argument.");
  if (mxGetM(A) != mxGetN(A))
    mexErrMsgIdAndTxt("MATLAB:lanczos_create:inputNotSquare", "Sparse matrix must be square.");
  if (mxGetPi(A) != nullptr)
    mexErrMsgIdAndTxt("MATLAB:lanczos_create:inputNotReal",
                      "First input argument must be a real matrix.");
  if (mxGetM(A) != n)
    mexErrMsgIdAndTxt("MATLAB:lanczos_create:badU", "Invalid U input argument.");
  // mwIndex -> int32 (lanczos_create.cpp:104-115); nnz from the column pointers
  const mwIndex* jc = mxGetJc(A);
  const mwIndex* ir = mxGetIr(A);
  const size_t nnz = jc[n];
  std::vector<int32_t> ptr(n + 1), ind(nnz);
  for (size_t i = 0; i <= n; ++i) ptr[i] = static_cast<int32_t>(jc[i]);
  for (size_t j = 0; j < nnz; ++j) ind[j] = static_cast<int32_t>(ir[j]);
  d.n = static_cast<int32_t>(n);
  d.sparse_format = GM_SPARSE_CSC;
  d.nnz = static_cast<int64_t>(nnz);
  d.A_values = mxGetPr(A);
  d.A_ptr = ptr.data();
  d.A_ind = ind.data();
  d.Q = real_arg(prhs[2], "MATLAB:lanczos_create:badQ", "Invalid Q input argument.", &c);
  d.num_eigs = static_cast<int32_t>(mxGetN(prhs[2]));
  if (c > 0 && mxGetM(prhs[2]) != n)
    mexErrMsgIdAndTxt("MATLAB:lanczos_create:badQ", "Invalid Q input argument.");
  if (c == 0) d.num_eigs = 0;
  d.lambda = real_arg(prhs[3], "MATLAB:lanczos_create:bad_lambda", "Invalid lambda input argument.", &c);
  if (c != static_cast<size_t>(d.num_eigs))
    mexErrMsgIdAndTxt("MATLAB:lanczos_create:bad_lambda", "Invalid lambda input argument.");
  d.subspace_size = static_cast<int32_t>(scalar_arg(prhs[4], "N"));
  d.gamma = real_arg(prhs[5], "MATLAB:lanczos_create:bad_gamma", "Invalid gamma input argument.", &c);
  d.gamma_size = static_cast<int32_t>(c);
  d.gamma_s_real = real_arg(prhs[6], "MATLAB:lanczos_create:bad_gamma_s_real",
                            "Invalid gamma_s_real input argument.", &c);
  if (mxGetM(prhs[6]) != static_cast<size_t>(d.gamma_size))
    mexErrMsgIdAndTxt("MATLAB:lanczos_create:bad_gamma_s_real", "Invalid gamma_s_real input argument.");
  d.contour_size = static_cast<int32_t>(mxGetN(prhs[6]));
  d.gamma_s_imag = real_arg(prhs[7], "MATLAB:lanczos_create:bad_gamma_s_imag",
                            "Invalid gamma_s_imag input argument.", &c);
  if (mxGetM(prhs[7]) != static_cast<size_t>(d.gamma_size) ||
      mxGetN(prhs[7]) != static_cast<size_t>(d.contour_size))
    mexErrMsgIdAndTxt("MATLAB:lanczos_create:bad_gamma_s_imag", "Invalid gamma_s_imag input argument.");
  d.exp_lambda = real_arg(prhs[8], "MATLAB:lanczos_create:bad_exp_lambda",
                          "Invalid exp_lambda input argument.", &c);
  if (c != static_cast<size_t>(d.num_eigs))
    mexErrMsgIdAndTxt("MATLAB:lanczos_create:bad_exp_lambda", "Invalid exp_lambda input argument.");
  d.phi_lambda = real_arg(prhs[9], "MATLAB:lanczos_create:bad_phi_lambda",
                          "Invalid phi_lambda input argument.", &c);
  if (c != static_cast<size_t>(d.num_eigs))
    mexErrMsgIdAndTxt("MATLAB:lanczos_create:bad_phi_lambda", "Invalid phi_lambda input argument.");
  d.dt = scalar_arg(prhs[10], "dt");
  d.ode = real_arg(prhs[11], "MATLAB:lanczos_create:bad_ode", "Invalid ode input argument.", &c);
  if (c != n) mexErrMsgIdAndTxt("MATLAB:lanczos_create:bad_ode", "Invalid ode input argument.");
  d.aode = scalar_arg(prhs[12], "aode");
  d.bode = scalar_arg(prhs[13], "bode");
  d.code = scalar_arg(prhs[14], "code");
  d.vrest = scalar_arg(prhs[15], "vrest");
  d.vamp = scalar_arg(prhs[16], "vamp");
  d.vth = scalar_arg(prhs[17], "vth");
  d.vpeak = scalar_arg(prhs[18], "vpeak");
  d.c1ion = scalar_arg(prhs[19], "c1ion");
  d.c2ion = scalar_arg(prhs[20], "c2ion");
  d.mass = real_arg(prhs[21], "MATLAB:lanczos_create:bad_mass", "Invalid mass input argument.", &c);
  if (c != n) mexErrMsgIdAndTxt("MATLAB:lanczos_create:bad_mass", "Invalid mass input argument.");
  d.wts1 = real_arg(prhs[22], "MATLAB:lanczos_create:bad_wts1", "Invalid wts1 input argument.", &c);
  if (mxGetN(prhs[22]) != 2 || mxGetM(prhs[22]) != static_cast<size_t>(d.contour_size))
    mexErrMsgIdAndTxt("MATLAB:lanczos_create:bad_wts1", "Invalid wts1 input argument.");
  d.wts2 = real_arg(prhs[23], "MATLAB:lanczos_create:bad_wts2", "Invalid wts2 input argument.", &c);
  if (mxGetN(prhs[23]) != 2 || mxGetM(prhs[23]) != static_cast<size_t>(d.contour_size))
    mexErrMsgIdAndTxt("MATLAB:lanczos_create:bad_wts2", "Invalid wts2 input argument.");
  d.shifts = real_arg(prhs[24], "MATLAB:lanczos_create:bad_shifts", "Invalid shifts input argument a.", &c);
  if (mxGetM(prhs[24]) != 1 && mxGetN(prhs[24]) != 1)
    mexErrMsgIdAndTxt("MATLAB:lanczos_create:bad_shifts", "Invalid shifts input argument b.");
  if (c != 4 * static_cast<size_t>(d.subspace_size) * static_cast<size_t>(d.contour_size))
    mexErrMsgIdAndTxt("MATLAB:lanczos_create:bad_shifts", "Invalid shifts input argument c.");
  d.valvec = real_arg(prhs[25], "MATLAB:lanczos_create:bad_valvec", "Invalid valvec input argument.", &c);
  if (c != 3) mexErrMsgIdAndTxt("MATLAB:lanczos_create:bad_valvec", "Invalid valvec input argument.");
  d.flags = 0;   // reference GPU semantics: update_U clamps u >= 0
  d.device = 0;
  gm_solver* s = nullptr;
  check(gm_lanczos_create(&d, &s));
  mwSize dims[2] = {1, 1};
  plhs[0] = mxCreateNumericArray(2, dims, mxUINT64_CLASS, mxREAL);
  const uint64_t v = static_cast<uint64_t>(reinterpret_cast<uintptr_t>(s));
  std::memcpy(mxGetData(plhs[0]), &v, sizeof(v));
#elif GM_GATEWAY == 2
  nargs(nlhs, 0, nrhs, 1, "lanczos_delete");
  check(gm_lanczos_delete(handle_arg(prhs[0], "lanczos_delete")));
#elif GM_GATEWAY == 3
  nargs(nlhs, 0, nrhs, 1, "build_subspace");
  check(gm_build_subspace(handle_arg(prhs[0], "build_subspace")));
#elif GM_GATEWAY == 4
  nargs(nlhs, 0, nrhs, 1, "lin_sys_solves");
  check(gm_lin_sys_solves(handle_arg(prhs[0], "lin_sys_solves")));
#elif GM_GATEWAY == 5
  nargs(nlhs, 0, nrhs, 1, "apply_V");
  check(gm_apply_V(handle_arg(prhs[0], "apply_V")));
#elif GM_GATEWAY == 6
  nargs(nlhs, 0, nrhs, 1, "update_U");
  check(gm_update_U(handle_arg(prhs[0], "update_U")));
#elif GM_GATEWAY == 7
  nargs(nlhs, 0, nrhs, 1, "update_U_monodomain");
  check(gm_update_U_monodomain(handle_arg(prhs[0], "update_U_monodomain")));
#elif GM_GATEWAY == 8 || GM_GATEWAY == 9 || GM_GATEWAY == 10
#if GM_GATEWAY == 8
  const char* gw = "get_U";
#elif GM_GATEWAY == 9
  const char* gw = "get_xsol";
#else
  const char* gw = "get_aabb";
#endif
  nargs(nlhs, 1, nrhs, 1, gw);
  gm_solver* s = handle_arg(prhs[0], gw);
  int64_t n = 0;
  int32_t m = 0;
  check(gm_get_info(s, &n, nullptr, &m, nullptr, nullptr, nullptr, nullptr));
#if GM_GATEWAY == 10
  plhs[0] = mxCreateDoubleMatrix(static_cast<mwSize>(m + 1), 4, mxREAL);
  check(gm_get_aabb(s, mxGetPr(plhs[0])));
#else
  plhs[0] = mxCreateDoubleMatrix(static_cast<mwSize>(n), 2, mxREAL);
#if GM_GATEWAY == 8
  check(gm_get_U(s, mxGetPr(plhs[0])));
#else
  check(gm_get_xsol(s, mxGetPr(plhs[0])));
#endif
#endif
#elif GM_GATEWAY == 11
  nargs(nlhs, 0, nrhs, 3, "set_y");
  gm_solver* s = handle_arg(prhs[0], "set_y");
  int32_t m = 0, ns = 0;
  check(gm_get_info(s, nullptr, nullptr, &m, nullptr, &ns, nullptr, nullptr));
  for (int i = 1; i <= 2; ++i)
    if (!mxIsDouble(prhs[i]) || mxGetM(prhs[i]) != static_cast<size_t>(m) ||
        mxGetN(prhs[i]) != 2 * static_cast<size_t>(ns))
      mexErrMsgIdAndTxt("MATLAB:set_y:badY", "y must be m x 2*contour_size real.");
  check(gm_set_y(s, mxGetPr(prhs[1]), mxGetPr(prhs[2])));
#elif GM_GATEWAY == 12
  // [Q, lambda] = eigs_smallest_mex(A, k [, [block degree max_passes tol]])
  if (nrhs < 2 || nrhs > 3)
    mexErrMsgIdAndTxt("MATLAB:eigs_smallest:invalidNumInputs",
                      "Two or three input arguments are required.");
  if (nlhs < 1 || nlhs > 2)
    mexErrMsgIdAndTxt("MATLAB:eigs_smallest:maxlhs", "One or two output arguments are required.");
  const mxArray* A = prhs[0];
  if (!mxIsSparse(A) || !mxIsDouble(A))
    mexErrMsgIdAndTxt("MATLAB:eigs_smallest:inputNotSparse", "A must be a sparse double matrix.");
  const size_t n = mxGetM(A);
  if (mxGetN(A) != n || n > 0x7fffffff)
    mexErrMsgIdAndTxt("MATLAB:eigs_smallest:notSquare", "A must be square.");
  if (!mxIsDouble(prhs[1]) || mxGetM(prhs[1]) * mxGetN(prhs[1]) != 1)
    mexErrMsgIdAndTxt("MATLAB:eigs_smallest:bad_k", "k must be a scalar.");
  const int k = static_cast<int>(mxGetPr(prhs[1])[0]);
  if (k < 1 || static_cast<double>(k) != mxGetPr(prhs[1])[0])
    mexErrMsgIdAndTxt("MATLAB:eigs_smallest:bad_k", "k must be a positive integer.");
  double opt[4] = {0.0, 40.0, 200.0, 1e-10};   // block (0: smallest of 16/32/64 >= 2k)
  if (nrhs == 3) {
    if (!mxIsDouble(prhs[2]) || mxGetM(prhs[2]) * mxGetN(prhs[2]) != 4)
      mexErrMsgIdAndTxt("MATLAB:eigs_smallest:bad_opts",
                        "opts must be [block degree max_passes tol].");
    std::memcpy(opt, mxGetPr(prhs[2]), sizeof(opt));
  }
  int block = static_cast<int>(opt[0]);
  if (block == 0) block = 2 * k <= 16 ? 16 : 2 * k <= 32 ? 32 : 64;
  // CSC of a symmetric matrix: the same arrays as its CSR
  const mwIndex* jc = mxGetJc(A);
  const mwIndex* ir = mxGetIr(A);
  const size_t nnz = jc[n];
  std::vector<int32_t> ptr(n + 1), ind(nnz);
  for (size_t i = 0; i <= n; ++i) ptr[i] = static_cast<int32_t>(jc[i]);
  for (size_t j = 0; j < nnz; ++j) ind[j] = static_cast<int32_t>(ir[j]);
  gm_eigs_desc d{};
  d.n = static_cast<int32_t>(n);
  d.nnz = static_cast<int64_t>(nnz);
  d.sparse_format = GM_SPARSE_CSC;
  d.A_values = mxGetPr(A);
  d.A_ptr = ptr.data();
  d.A_ind = ind.data();
  d.num_eigs = k;
  d.block_size = block;
  d.degree = static_cast<int32_t>(opt[1]);
  d.max_passes = static_cast<int32_t>(opt[2]);
  d.tol = opt[3];
  d.upper_bound = 0.0;
  d.seed = 0;
  d.device = 0;
  mxArray* Q = mxCreateDoubleMatrix(static_cast<mwSize>(n), static_cast<mwSize>(k), mxREAL);
  mxArray* lam = mxCreateDoubleMatrix(static_cast<mwSize>(k), 1, mxREAL);
  check(gm_eigs_smallest(&d, mxGetPr(lam), mxGetPr(Q), nullptr, nullptr), gm_eigs_last_error);
  plhs[0] = Q;
  if (nlhs == 2) plhs[1] = lam;
#elif GM_GATEWAY == 13
  // [Q, lambda] = eigs_shift_invert_mex(A, k [, sigma [, [block max_blocks tol]]])
  if (nrhs < 2 || nrhs > 4)
    mexErrMsgIdAndTxt("MATLAB:eigs_shift_invert:invalidNumInputs",
                      "Two to four input arguments are required.");
  if (nlhs < 1 || nlhs > 2)
    mexErrMsgIdAndTxt("MATLAB:eigs_shift_invert:maxlhs",
                      "One or two output arguments are required.");
  const mxArray* A = prhs[0];
  if (!mxIsSparse(A) || !mxIsDouble(A))
    mexErrMsgIdAndTxt("MATLAB:eigs_shift_invert:inputNotSparse",
                      "A must be a sparse double matrix.");
  const size_t n = mxGetM(A);
  if (mxGetN(A) != n || n > 0x7fffffff)
    mexErrMsgIdAndTxt("MATLAB:eigs_shift_invert:notSquare", "A must be square.");
  if (!mxIsDouble(prhs[1]) || mxGetM(prhs[1]) * mxGetN(prhs[1]) != 1)
    mexErrMsgIdAndTxt("MATLAB:eigs_shift_invert:bad_k", "k must be a scalar.");
  const int k = static_cast<int>(mxGetPr(prhs[1])[0]);
  if (k < 1 || static_cast<double>(k) != mxGetPr(prhs[1])[0])
    mexErrMsgIdAndTxt("MATLAB:eigs_shift_invert:bad_k", "k must be a positive integer.");
  const mwIndex* jc = mxGetJc(A);
  const mwIndex* ir = mxGetIr(A);
  const size_t nnz = jc[n];
  const double* val = mxGetPr(A);
  // default shift: -1e-10 times the largest absolute column sum (= row sum)
  double sigma = 0.0;
  for (size_t j = 0; j < n; ++j) {
    double acc = 0.0;
    for (mwIndex e = jc[j]; e < jc[j + 1]; ++e) acc += val[e] < 0 ? -val[e] : val[e];
    sigma = acc > sigma ? acc : sigma;
  }
  sigma *= -1e-10;
  if (nrhs >= 3 && !mxIsEmpty(prhs[2])) {
    if (!mxIsDouble(prhs[2]) || mxGetM(prhs[2]) * mxGetN(prhs[2]) != 1)
      mexErrMsgIdAndTxt("MATLAB:eigs_shift_invert:bad_sigma", "sigma must be a scalar.");
    sigma = mxGetPr(prhs[2])[0];
  }
  double opt[3] = {0.0, 12.0, 1e-14};   // block (0: k + 4, at least 8), blocks per cycle, tol
  if (nrhs == 4) {
    if (!mxIsDouble(prhs[3]) || mxGetM(prhs[3]) * mxGetN(prhs[3]) != 3)
      mexErrMsgIdAndTxt("MATLAB:eigs_shift_invert:bad_opts",
                        "opts must be [block max_blocks tol].");
    std::memcpy(opt, mxGetPr(prhs[3]), sizeof(opt));
  }
  int block = static_cast<int>(opt[0]);
  if (block == 0) block = k + 4 < 8 ? 8 : k + 4;
  if (2 * static_cast<size_t>(block) > n) block = static_cast<int>(n / 2);
  // CSC of a symmetric matrix: the same arrays as its CSR
  std::vector<int32_t> ptr(n + 1), ind(nnz);
  for (size_t i = 0; i <= n; ++i) ptr[i] = static_cast<int32_t>(jc[i]);
  for (size_t j = 0; j < nnz; ++j) ind[j] = static_cast<int32_t>(ir[j]);
  gm_sieigs_desc d{};
  d.n = static_cast<int32_t>(n);
  d.nnz = static_cast<int64_t>(nnz);
  d.A_values = val;
  d.A_ptr = ptr.data();
  d.A_ind = ind.data();
  d.num_eigs = k;
  d.block_size = block;
  d.max_blocks = static_cast<int32_t>(opt[1]);
  d.max_restarts = -1;
  d.tol = opt[2];
  d.sigma = sigma;
  d.seed = 0;
  d.threads = 0;
  mxArray* Q = mxCreateDoubleMatrix(static_cast<mwSize>(n), static_cast<mwSize>(k), mxREAL);
  mxArray* lam = mxCreateDoubleMatrix(static_cast<mwSize>(k), 1, mxREAL);
  check(gm_eigs_shift_invert(&d, mxGetPr(lam), mxGetPr(Q), nullptr, nullptr), gm_chol_last_error);
  plhs[0] = Q;
  if (nlhs == 2) plhs[1] = lam;
#elif GM_GATEWAY == 14
  // lmax = lambda_max_mex(A)
  if (nrhs != 1)
    mexErrMsgIdAndTxt("MATLAB:lambda_max:invalidNumInputs", "One input argument is required.");
  if (nlhs > 1)
    mexErrMsgIdAndTxt("MATLAB:lambda_max:maxlhs", "Too many output arguments.");
  const mxArray* A = prhs[0];
  if (!mxIsSparse(A) || !mxIsDouble(A))
    mexErrMsgIdAndTxt("MATLAB:lambda_max:inputNotSparse", "A must be a sparse double matrix.");
  const size_t n = mxGetM(A);
  if (mxGetN(A) != n || n > 0x7fffffff)
    mexErrMsgIdAndTxt("MATLAB:lambda_max:notSquare", "A must be square.");
  const mwIndex* jc = mxGetJc(A);
  const mwIndex* ir = mxGetIr(A);
  const size_t nnz = jc[n];
  std::vector<int32_t> ptr(n + 1), ind(nnz);
  for (size_t i = 0; i <= n; ++i) ptr[i] = static_cast<int32_t>(jc[i]);
  for (size_t j = 0; j < nnz; ++j) ind[j] = static_cast<int32_t>(ir[j]);
  double lmax = 0.0;
  check(gm_lambda_max(static_cast<int32_t>(n), ptr.data(), ind.data(), mxGetPr(A), 0.0, 0, 0, 0,
                      &lmax, nullptr),
        gm_chol_last_error);
  plhs[0] = mxCreateDoubleScalar(lmax);
#else
#error "unknown GM_GATEWAY"
#endif
}
