// Checkpoint wire format of the reference: an RBM<T> written by cereal's
// BinaryOutputArchive (Serializers/SerializeRBM.hpp:21-79 with the Eigen
// matrix save of Serializers/SerializeEigen.hpp:10-35), as Runners/RunRBM.hpp:
// 155-165 writes w%04d.dat (through a std::unique_ptr) and
// Runners/AbstractRunner.hpp:83-94 reads it back.
//
// cereal (github.com/USCiLab/cereal, a git submodule of the reference that is
// not vendored; version unpinned) stores arithmetic values as their raw bytes
// in host order with no header, sizes or padding.  A std::unique_ptr<T>
// first stores one byte (1: an object follows, 0: null); a type registered
// with CEREAL_CLASS_VERSION and a versioned save stores its version as a
// uint32 the first time the type appears in the archive.  The RBM's save then
// writes, in order (little-endian, as on the x86-64 hosts the reference runs
// on):
//   [uint8 valid = 1]                       (only through a unique_ptr)
//   uint32 class version (1)
//   bool useBias (1 byte)
//   uint32 N, uint32 M                      (getN / getM)
//   int32 rows = M, int32 cols = N, M*N entries of W, column-major
//   if useBias: int32 N, int32 1, N entries of a; int32 M, int32 1, M entries of b
// with entries of the machine's scalar type (float, double, or complex as
// (re, im) pairs).  The file does not name the scalar type; the reader has
// to know it, as the reference's template parameter does.
//
// Host code only: encoding and decoding touch no device memory.
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "yq_internal.h"

namespace {

size_t elem_bytes(int elem) {
  switch (elem) {
    case YQ_CKPT_F32: return 4;
    case YQ_CKPT_F64: return 8;
    case YQ_CKPT_C64: return 8;
    case YQ_CKPT_C128: return 16;
    default: return 0;
  }
}
bool elem_complex(int elem) { return elem == YQ_CKPT_C64 || elem == YQ_CKPT_C128; }

struct Writer {
  unsigned char* p;
  void bytes(const void* src, size_t n) {
    memcpy(p, src, n);
    p += n;
  }
  void u8(uint8_t v) { bytes(&v, 1); }
  void u32(uint32_t v) {
    const unsigned char b[4] = {(unsigned char)v, (unsigned char)(v >> 8),
                                (unsigned char)(v >> 16), (unsigned char)(v >> 24)};
    bytes(b, 4);
  }
  void i32(int32_t v) { u32((uint32_t)v); }
  void f64(double v) {
    uint64_t u;
    memcpy(&u, &v, 8);
    for (int i = 0; i < 8; ++i) u8((uint8_t)(u >> (8 * i)));
  }
  void f32(float v) {
    uint32_t u;
    memcpy(&u, &v, 4);
    u32(u);
  }
  // `count` scalars of the file's element type from doubles (complex: 2 each)
  void entries(int elem, const double* v, size_t count) {
    const size_t nd = count * (elem_complex(elem) ? 2 : 1);
    for (size_t i = 0; i < nd; ++i) {
      if (elem == YQ_CKPT_F32 || elem == YQ_CKPT_C64)
        f32((float)v[i]);
      else
        f64(v[i]);
    }
  }
};

struct Reader {
  const unsigned char* p;
  const unsigned char* end;
  bool ok = true;
  bool take(void* dst, size_t n) {
    if (!ok || (size_t)(end - p) < n) return ok = false;
    memcpy(dst, p, n);
    p += n;
    return true;
  }
  uint8_t u8() {
    uint8_t v = 0;
    take(&v, 1);
    return v;
  }
  uint32_t u32() {
    unsigned char b[4] = {0, 0, 0, 0};
    take(b, 4);
    return (uint32_t)b[0] | ((uint32_t)b[1] << 8) | ((uint32_t)b[2] << 16) |
           ((uint32_t)b[3] << 24);
  }
  int32_t i32() { return (int32_t)u32(); }
  double f64() {
    uint64_t u = 0;
    for (int i = 0; i < 8; ++i) u |= (uint64_t)u8() << (8 * i);
    double v;
    memcpy(&v, &u, 8);
    return v;
  }
  float f32() {
    const uint32_t u = u32();
    float v;
    memcpy(&v, &u, 4);
    return v;
  }
  void entries(int elem, double* out, size_t count) {
    const size_t nd = count * (elem_complex(elem) ? 2 : 1);
    for (size_t i = 0; i < nd && ok; ++i)
      out[i] = (elem == YQ_CKPT_F32 || elem == YQ_CKPT_C64) ? (double)f32() : f64();
  }
};

size_t body_bytes(int elem, int n, int m, int use_bias) {
  const size_t e = elem_bytes(elem);
  size_t s = 4 + 1 + 4 + 4;  // version, useBias, N, M
  s += 8 + (size_t)n * m * e;
  if (use_bias) s += 8 + (size_t)n * e + 8 + (size_t)m * e;
  return s;
}

}  // namespace

YQ_API size_t yq_ckpt_rbm_bytes(int elem, int n, int m, int use_bias, int pointer_wrapper) {
  if (!elem_bytes(elem) || n <= 0 || m <= 0) return 0;
  return (pointer_wrapper ? 1 : 0) + body_bytes(elem, n, m, use_bias ? 1 : 0);
}

YQ_API int yq_ckpt_rbm_encode(int elem, int n, int m, int use_bias, const double* params,
                              int pointer_wrapper, unsigned char* out, size_t size) {
  const size_t need = yq_ckpt_rbm_bytes(elem, n, m, use_bias, pointer_wrapper);
  if (!need || !params || !out || size < need) return YQ_ERR_ARG;
  Writer w{out};
  if (pointer_wrapper) w.u8(1);
  w.u32(1);  // CEREAL_CLASS_VERSION(yannq::RBM<T>, 1)
  w.u8(use_bias ? 1 : 0);
  w.u32((uint32_t)n);
  w.u32((uint32_t)m);
  const size_t nm = (size_t)n * m, stride = elem_complex(elem) ? 2 : 1;
  w.i32(m);
  w.i32(n);
  w.entries(elem, params, nm);
  if (use_bias) {
    w.i32(n);
    w.i32(1);
    w.entries(elem, params + stride * nm, n);
    w.i32(m);
    w.i32(1);
    w.entries(elem, params + stride * (nm + n), m);
  }
  return YQ_OK;
}

YQ_API int yq_ckpt_rbm_decode(int elem, const unsigned char* buf, size_t size,
                              int pointer_wrapper, int* n, int* m, int* use_bias,
                              double* params, size_t params_len) {
  if (!elem_bytes(elem) || !buf || !n || !m || !use_bias) return YQ_ERR_ARG;
  Reader r{buf, buf + size};
  if (pointer_wrapper && r.u8() != 1) return YQ_ERR_ARG;  // null pointer or not this form
  r.u32();  // class version: the reference's load ignores it
  const uint8_t ub = r.u8();
  const uint32_t N = r.u32(), M = r.u32();
  if (!r.ok || ub > 1 || N == 0 || M == 0 || N > (1u << 20) || M > (1u << 20))
    return YQ_ERR_ARG;
  // the exact length for these sizes: anything else is another type or layout
  if (size != yq_ckpt_rbm_bytes(elem, (int)N, (int)M, ub, pointer_wrapper)) return YQ_ERR_ARG;
  *n = (int)N;
  *m = (int)M;
  *use_bias = ub;
  if (!params) return YQ_OK;  // size query
  const size_t nm = (size_t)N * M, stride = elem_complex(elem) ? 2 : 1;
  const size_t dim = ub ? nm + N + M : nm;
  if (params_len < dim * stride) return YQ_ERR_ARG;
  if (r.i32() != (int32_t)M || r.i32() != (int32_t)N) return YQ_ERR_ARG;
  r.entries(elem, params, nm);
  if (ub) {
    if (r.i32() != (int32_t)N || r.i32() != 1) return YQ_ERR_ARG;
    r.entries(elem, params + stride * nm, N);
    if (r.i32() != (int32_t)M || r.i32() != 1) return YQ_ERR_ARG;
    r.entries(elem, params + stride * (nm + N), M);
  }
  return r.ok ? YQ_OK : YQ_ERR_ARG;
}
