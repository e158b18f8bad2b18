// Kernel-matrix entry generators evaluated inside the sketch kernels.
//
// row_pt / col_pt fetch the point of a row / column index once, eval forms
// the entry from the two points (and the global indices, for the diagonal
// terms).  The FP64 reciprocal is v_rcp_f64 (relative error up to about
// 2^-24) refined by one Newton step: the error after it is about the square,
// 2^-48, i.e. within 16 units in the last place of the correctly rounded
// quotient (at most 11 measured over 4M inputs across 2^-40 .. 2^20).  A
// second step would round correctly but costs 2 of the 7 VALU operations of
// an entry in the FP64 sketch, which is VALU + MFMA bound.
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace hmat {
namespace kern {

__device__ inline double rcp64(double d) {
  double r = __builtin_amdgcn_rcp(d);
  const double e = fma(-d, r, 1.0);
  return fma(r, e, r);
}

struct GDense {
  const double* A;
  int64_t lda;
  __device__ double row_pt(int64_t) const { return 0.0; }
  __device__ double col_pt(int64_t) const { return 0.0; }
  __device__ double eval(double, double, int64_t i, int64_t j) const { return A[i + j * lda]; }
};
struct GGaussian {  // exp(-(x_i-x_j)^2 c) + p1 delta, c = 1/(2 sigma^2)
  const double* x;
  double c, p1;
  __device__ double row_pt(int64_t i) const { return x[i]; }
  __device__ double col_pt(int64_t j) const { return x[j]; }
  __device__ double eval(double xi, double xj, int64_t i, int64_t j) const {
    const double d = xi - xj;
    return exp(-(d * d) * c) + (i == j ? p1 : 0.0);
  }
};
struct GCauchy {  // 1/(x_i - y_j)
  const double* x;
  const double* y;
  __device__ double row_pt(int64_t i) const { return x[i]; }
  __device__ double col_pt(int64_t j) const { return y[j]; }
  __device__ double eval(double xi, double yj, int64_t, int64_t) const { return rcp64(xi - yj); }
};
struct GRefCauchy {  // 1/(x_i + x_j + 1) + p1 delta     (kernel.h:30-48)
  const double* x;
  double p1;
  __device__ double row_pt(int64_t i) const { return x[i]; }
  __device__ double col_pt(int64_t j) const { return x[j]; }
  __device__ double eval(double xi, double xj, int64_t i, int64_t j) const {
    return rcp64(xi + xj + 1.0) + (i == j ? p1 : 0.0);
  }
};
struct GRefGaussian {  // exp(-(x_i-x_j)^2 / h^2) + p1 delta   (kernel.h:52-65), c = 1/h^2
  const double* x;
  double c, p1;
  __device__ double row_pt(int64_t i) const { return x[i]; }
  __device__ double col_pt(int64_t j) const { return x[j]; }
  __device__ double eval(double xi, double xj, int64_t i, int64_t j) const {
    const double d = xi - xj;
    return exp(-(d * d) * c) + (i == j ? p1 : 0.0);
  }
};
struct GRefCoulomb {  // 1/|x_i-x_j|, 0 on the diagonal, + p1    (kernel.h:69-89)
  const double* x;
  double p1;
  __device__ double row_pt(int64_t i) const { return x[i]; }
  __device__ double col_pt(int64_t j) const { return x[j]; }
  __device__ double eval(double xi, double xj, int64_t i, int64_t j) const {
    return i == j ? p1 : rcp64(fabs(xi - xj));
  }
};

struct GGaussDSigma {  // d/dsigma exp(-d^2/(2 sigma^2)) = exp(-d^2 c) d^2 / sigma^3
  const double* x;
  double c, is3;
  __device__ double row_pt(int64_t i) const { return x[i]; }
  __device__ double col_pt(int64_t j) const { return x[j]; }
  __device__ double eval(double xi, double xj, int64_t, int64_t) const {
    const double d2 = (xi - xj) * (xi - xj);
    return exp(-d2 * c) * d2 * is3;
  }
};

template <class F>
inline void with_gen(const GenDesc& g, F&& f) {
  switch (g.kind) {
    case GenKind::Dense: f(GDense{g.A, g.lda}); break;
    case GenKind::Gaussian: f(GGaussian{g.x, 1.0 / (2.0 * g.p0 * g.p0), g.p1}); break;
    case GenKind::Cauchy: f(GCauchy{g.x, g.y}); break;
    case GenKind::RefCauchy: f(GRefCauchy{g.x, g.p1}); break;
    case GenKind::RefGaussian: f(GRefGaussian{g.x, 1.0 / (g.p0 * g.p0), g.p1}); break;
    case GenKind::RefCoulomb: f(GRefCoulomb{g.x, g.p1}); break;
    case GenKind::GaussianDSigma:
      f(GGaussDSigma{g.x, 1.0 / (2.0 * g.p0 * g.p0), 1.0 / (g.p0 * g.p0 * g.p0)});
      break;
  }
}

// FP32 generators of the mixed-precision range finder.  Their entries only
// steer the basis; the projection that defines the operator is FP64.
struct LDense {
  const double* A;
  int64_t lda;
  __device__ float row_pt(int64_t) const { return 0.f; }
  __device__ float col_pt(int64_t) const { return 0.f; }
  __device__ float eval(float, float, int64_t i, int64_t j) const { return (float)A[i + j * lda]; }
};
struct LGaussian {
  const double* x;
  float c, p1;
  __device__ float row_pt(int64_t i) const { return (float)x[i]; }
  __device__ float col_pt(int64_t j) const { return (float)x[j]; }
  __device__ float eval(float xi, float xj, int64_t i, int64_t j) const {
    const float d = xi - xj;
    return __expf(-(d * d) * c) + (i == j ? p1 : 0.f);
  }
};
struct LCauchy {
  const double* x;
  const double* y;
  __device__ float row_pt(int64_t i) const { return (float)x[i]; }
  __device__ float col_pt(int64_t j) const { return (float)y[j]; }
  __device__ float eval(float xi, float yj, int64_t, int64_t) const {
    return __builtin_amdgcn_rcpf(xi - yj);
  }
};
struct LRefCauchy {
  const double* x;
  float p1;
  __device__ float row_pt(int64_t i) const { return (float)x[i]; }
  __device__ float col_pt(int64_t j) const { return (float)x[j]; }
  __device__ float eval(float xi, float xj, int64_t i, int64_t j) const {
    return __builtin_amdgcn_rcpf(xi + xj + 1.f) + (i == j ? p1 : 0.f);
  }
};
struct LRefGaussian {
  const double* x;
  float c, p1;
  __device__ float row_pt(int64_t i) const { return (float)x[i]; }
  __device__ float col_pt(int64_t j) const { return (float)x[j]; }
  __device__ float eval(float xi, float xj, int64_t i, int64_t j) const {
    const float d = xi - xj;
    return __expf(-(d * d) * c) + (i == j ? p1 : 0.f);
  }
};
struct LRefCoulomb {
  const double* x;
  float p1;
  __device__ float row_pt(int64_t i) const { return (float)x[i]; }
  __device__ float col_pt(int64_t j) const { return (float)x[j]; }
  __device__ float eval(float xi, float xj, int64_t i, int64_t j) const {
    return i == j ? p1 : __builtin_amdgcn_rcpf(fabsf(xi - xj));
  }
};

template <class F>
inline void with_gen_lp(const GenDesc& g, F&& f) {
  switch (g.kind) {
    case GenKind::Dense: f(LDense{g.A, g.lda}); break;
    case GenKind::Gaussian: f(LGaussian{g.x, (float)(1.0 / (2.0 * g.p0 * g.p0)), (float)g.p1}); break;
    case GenKind::Cauchy: f(LCauchy{g.x, g.y}); break;
    case GenKind::RefCauchy: f(LRefCauchy{g.x, (float)g.p1}); break;
    case GenKind::RefGaussian: f(LRefGaussian{g.x, (float)(1.0 / (g.p0 * g.p0)), (float)g.p1}); break;
    case GenKind::RefCoulomb: f(LRefCoulomb{g.x, (float)g.p1}); break;
    default: throw Error(HMAT_ERR_ARG, "kernel has no low-precision generator");
  }
}

}  // namespace kern
}  // namespace hmat
