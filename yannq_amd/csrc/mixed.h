// Mixed-precision factored SR operator: (O_c^T O_c / Ns + shift) p with the
// products O p and O^T y on bf16 MFMA (v_mfma_f32_16x16x32_bf16), and the
// float conjugate-gradient loop that drives it with device-resident scalars.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace yq {

struct MixedWork {
  int ns = 0, n = 0, m = 0, use_bias = 1;
  int Kp = 0;    // n+1 rounded up to 32 (k-steps of the forward product)
  int Kr = 0;    // n+1 rounded up to 16 (rows of the backward product's result,
                 // of its chunk partials and of the CG-step grid)
  int Np = 0;    // m+1 rounded up to 16 (column tiles)
  int nsp = 0;   // ns rounded up to 512
  int ncg = 0;   // column groups of the forward product (Np/16 over ct)
  size_t dim = 0;
  // operands prepared once per SR build
  uint8_t* sigA = nullptr;   // sigma' bits row-major  nsp x 32 bytes (set: -1)
  uint8_t* sigT = nullptr;   // sigma'^T bits          Kp x nsp / 8 bytes
  float* Trm = nullptr;      // T' row-major           nsp x Np
  float* Tt = nullptr;       // T'^T                   Np x nsp
  // per-product scratch
  uint16_t* vsplit = nullptr;  // 3 x Np x Kp bf16 (hi, mid, lo of V')
  float* ypart = nullptr;      // ncg x nsp forward partial sums
  float* part = nullptr;       // chunks x Kp x Np backward partials
  float* om32 = nullptr;       // column mean in float (dim)
  // CG vectors (float) and device scalars (double)
  // (x, r; direction p and s = (S + shift) p; u = M r and Ap = (S + shift) u)
  float *x = nullptr, *r = nullptr, *p = nullptr, *s = nullptr, *u = nullptr, *Ap = nullptr;
  float* dinv = nullptr;       // Jacobi preconditioner 1 / (diag S + shift)
  double* red = nullptr;       // per-block partials of the fused dot products
  size_t cap_rows = 0, cap_kp = 0, cap_np = 0, cap_dim = 0, cap_red = 0;
  // tile shapes of the products: forward column tiles per group (2, 4, 8) and
  // backward samples per chunk (128, 256, 512); chosen per operand shape by
  // the autotuner when enabled
  int ct = 4, chunk = 256;
  bool autotune = false;
  int pin_ct = 0, pin_chunk = 0;  // shapes fixed by the caller (0: not fixed)
  // CG control block (device), its pinned host mirror and poll events
  double* dctl = nullptr;      // shift, stopping threshold, alpha and gamma slots
  int* ictl = nullptr;         // stop flag, completed iterations, budget
  int* host_ctl = nullptr;     // 2 slots of 4 ints
  int* host_ctl_dev = nullptr; // the same pinned memory as a device address
  hipEvent_t ev[2] = {nullptr, nullptr};
  // captured CG chunks (index: Jacobi preconditioner off / on) and their shape
  bool use_graphs = true;
  int chunk_iters = 8;  // CG iterations per captured graph and host poll (even)
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  long gkey[11] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
};

// Build the bf16 / float operands from the batch (sigma: ns x n int8,
// T32: ns x m float = tanh(theta)).
int mixed_prepare(MixedWork& w, const int8_t* sigma, const float* T32, int ns,
                  int n, int m, int use_bias, hipStream_t st, std::string& err);
// out = (S + shift) v for FP64 vectors on the device, through the float /
// split-bf16 operator (omean64: FP64 column mean of O).
int mixed_apply64(MixedWork& w, double shift, const double* omean64,
                  const double* v, double* out, hipStream_t st, std::string& err);
// Inner CG in float with device scalars: solves (S + shift) x = b to the
// relative residual tol (checked every 8 iterations); b and x are FP64.
// diag64 (diag S in FP64) selects Jacobi preconditioning; null for none.
int mixed_cg(MixedWork& w, double shift, const double* omean64, const double* diag64,
             const double* b, double* x, double tol, int max_iter, int* iters,
             hipStream_t st, std::string& err);
// One machine's inner solve in a grouped call: (S + shift) x = b to the
// relative residual tol within max_iter iterations, as mixed_cg does; jobs
// that are not active are left alone.
struct GroupJob {
  MixedWork* w = nullptr;         // prepared by mixed_prepare
  const double* omean64 = nullptr;
  const double* diag64 = nullptr;  // Jacobi preconditioner, or null
  const double* b = nullptr;
  double* x = nullptr;
  double tol = 0.0;
  int max_iter = 0;
  bool active = false;
  int iters = 0;   // out: completed iterations
  int rc = 0;      // out: YQ_OK or YQ_ERR_NOT_CONVERGED
};
// Device tables, poll buffers and the captured chunk of a grouped CG.
struct MixedGroup {
  int cap = 0, cap_blocks = 0;
  void* desc = nullptr;       // per-machine descriptors (device)
  void* host_desc = nullptr;  // pinned staging
  int* btab = nullptr;        // block tables of the three grouped kernels (int2)
  int* host_btab = nullptr;
  int* ctl = nullptr;         // 2 slots x 4 control ints per machine
  int* host_ctl = nullptr;
  double* stat = nullptr;     // per machine: <b, b>, final <r, r>, iterations
  double* host_stat = nullptr;
  int nf = 0, nbw = 0, nc = 0;  // blocks of the forward, backward and step launches
  int cg_threads = 1024;        // threads per step block
  size_t lds_f = 0, lds_b = 0;
  hipEvent_t ev[2] = {nullptr, nullptr};
  bool use_graphs = true;
  int chunk_iters = 8;
  hipGraphExec_t gexec = nullptr;
  long gkey[7] = {0, 0, 0, 0, 0, 0, 0};
};
// The inner CG of every active job in one set of launches per iteration (a
// grouped forward product, backward product and CG step whose grids
// concatenate the machines' own); each machine computes exactly what
// mixed_cg computes for it.  Synchronises st before returning.
int mixed_cg_group(MixedGroup& g, std::vector<GroupJob>& jobs, double shift,
                   hipStream_t st, std::string& err);
void mixed_group_release(MixedGroup& g);
void mixed_release(MixedWork& w);
// drop the captured CG chunks (after a reallocation or a shape change)
void mixed_drop_graphs(MixedWork& w);

}  // namespace yq
