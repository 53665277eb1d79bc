// Optional libtorch layer: torch::autograd::Function wrappers whose forward
// and backward call the library's C ABI (no reference counterpart; the
// reference has no differentiation support).
//
//   hcblas::gemm_batched(A, B, C?, alpha, beta, transA, transB)
//       C_out[b] = alpha * op(A[b]) * op(B[b]) + beta * C[b]   (row-major,
//       batch-major contiguous tensors; A, B float32 / float64 / bfloat16;
//       C_out float32 for bfloat16 inputs, else the input type)
//       backward: hcblas_gemm_strided_batched_backward (dA, dB, dC_in).
//   hcblas::ger(x, y, A, alpha)
//       A_out = alpha * x * y^T + A (row-major, float32 / float64)
//       backward: hcblas_{s,d}ger_backward (dx, dy), dA = dA_out.
//
// Work is enqueued on torch's current HIP stream through a library handle
// kept per thread, device and stream. Built as a separate shared object linked
// against libhcblas.so and libtorch; loaded with torch.ops.load_library.
#include <torch/torch.h>
#include <torch/library.h>
#include <c10/hip/HIPStream.h>

#include <functional>
#include <stdexcept>
#include <string>
#include <unordered_map>

#include "../../include/hcblas.h"

namespace {

using torch::Tensor;
using torch::autograd::AutogradContext;
using torch::autograd::tensor_list;

void check(hcblasStatus_t st, const char *what) {
  TORCH_CHECK(st == HCBLAS_STATUS_SUCCESS, what, " failed with hcblas status ",
              static_cast<int>(st));
}

// One handle per (thread, device, stream): a handle's workspaces (split-K
// partials, the transposed operand copies of a BF16 backward) are used by the
// work it enqueues, so two streams never share one handle, and work on one
// stream cannot overwrite a workspace that another stream is still reading.
struct StreamKey {
  int device;
  hipStream_t stream;
  bool operator==(const StreamKey &o) const {
    return device == o.device && stream == o.stream;
  }
};
struct StreamKeyHash {
  size_t operator()(const StreamKey &k) const {
    return std::hash<const void *>()(k.stream) * 31u + (size_t)k.device;
  }
};

hcblasHandle_t handle_for(const Tensor &t) {
  thread_local std::unordered_map<StreamKey, hcblasHandle_t, StreamKeyHash>
      handles;
  const int dev = t.get_device();
  const StreamKey key{dev, c10::hip::getCurrentHIPStream(dev).stream()};
  auto it = handles.find(key);
  if (it == handles.end()) {
    c10::DeviceGuard guard(t.device());
    hcblasHandle_t h = nullptr;
    check(hcblasCreate(&h), "hcblasCreate");
    check(hcblasSetStream(h, key.stream), "hcblasSetStream");
    it = handles.emplace(key, h).first;
  }
  return it->second;
}

hcblasDatatype_t dtype_of(const Tensor &t) {
  switch (t.scalar_type()) {
    case torch::kFloat32:
      return HCBLAS_R_32F;
    case torch::kFloat64:
      return HCBLAS_R_64F;
    case torch::kBFloat16:
      return HCBLAS_R_16BF;
    default:
      TORCH_CHECK(false, "hcblas: unsupported dtype ", t.scalar_type());
  }
}

// Accumulation / output type of the GEMM for an input type.
torch::ScalarType out_type(const Tensor &t) {
  return t.scalar_type() == torch::kBFloat16 ? torch::kFloat32
                                             : t.scalar_type();
}

struct GemmShape {
  int64_t batch, M, N, K;
};

GemmShape gemm_shape(const Tensor &A, const Tensor &B, bool ta, bool tb) {
  TORCH_CHECK(A.dim() == 3 && B.dim() == 3, "hcblas.gemm_batched: A and B must be [batch, rows, cols]");
  TORCH_CHECK(A.size(0) == B.size(0), "hcblas.gemm_batched: batch mismatch");
  GemmShape s;
  s.batch = A.size(0);
  s.M = ta ? A.size(2) : A.size(1);
  s.K = ta ? A.size(1) : A.size(2);
  const int64_t kb = tb ? B.size(2) : B.size(1);
  s.N = tb ? B.size(1) : B.size(2);
  TORCH_CHECK(kb == s.K, "hcblas.gemm_batched: inner dimensions differ");
  return s;
}

class GemmBatchedFn : public torch::autograd::Function<GemmBatchedFn> {
 public:
  static Tensor forward(AutogradContext *ctx, Tensor A, Tensor B,
                        c10::optional<Tensor> C, double alpha, double beta,
                        bool ta, bool tb) {
    TORCH_CHECK(A.is_cuda() && B.is_cuda(), "hcblas.gemm_batched: GPU tensors required");
    TORCH_CHECK(A.scalar_type() == B.scalar_type(), "hcblas.gemm_batched: A and B types differ");
    A = A.contiguous();
    B = B.contiguous();
    const GemmShape s = gemm_shape(A, B, ta, tb);
    const auto ot = out_type(A);
    Tensor out;
    const bool has_c = C.has_value() && C->defined();
    if (has_c) {
      TORCH_CHECK(C->scalar_type() == ot && C->sizes() == torch::IntArrayRef({s.batch, s.M, s.N}),
                  "hcblas.gemm_batched: C must be [batch, M, N] of the output type");
      out = C->contiguous().clone();
    } else {
      out = torch::empty({s.batch, s.M, s.N}, A.options().dtype(ot));
    }
    const double b = has_c ? beta : 0.0;
    hcblasHandle_t h = handle_for(A);
    const float af = (float)alpha, bf = (float)b;
    const void *pa = ot == torch::kFloat64 ? (const void *)&alpha : (const void *)&af;
    const void *pb = ot == torch::kFloat64 ? (const void *)&b : (const void *)&bf;
    check(hcblas_gemm_strided_batched(
              h, dtype_of(A), HCBLAS_ROW_MAJOR, ta ? HCBLAS_OP_T : HCBLAS_OP_N,
              tb ? HCBLAS_OP_T : HCBLAS_OP_N, (int)s.M, (int)s.N, (int)s.K, pa,
              A.data_ptr(), A.size(2), 0, A.size(1) * A.size(2), B.data_ptr(),
              B.size(2), 0, B.size(1) * B.size(2), pb, out.data_ptr(), s.N, 0,
              s.M * s.N, (int)s.batch),
          "hcblas_gemm_strided_batched");
    ctx->save_for_backward({A, B});
    ctx->saved_data["alpha"] = alpha;
    ctx->saved_data["beta"] = b;
    ctx->saved_data["ta"] = ta;
    ctx->saved_data["tb"] = tb;
    ctx->saved_data["has_c"] = has_c;
    return out;
  }

  static tensor_list backward(AutogradContext *ctx, tensor_list grads) {
    auto saved = ctx->get_saved_variables();
    Tensor A = saved[0], B = saved[1];
    const double alpha = ctx->saved_data["alpha"].toDouble();
    const double beta = ctx->saved_data["beta"].toDouble();
    const bool ta = ctx->saved_data["ta"].toBool();
    const bool tb = ctx->saved_data["tb"].toBool();
    const bool has_c = ctx->saved_data["has_c"].toBool();
    const GemmShape s = gemm_shape(A, B, ta, tb);
    const auto ot = out_type(A);
    Tensor dC = grads[0].to(ot).contiguous();
    const bool need_a = ctx->needs_input_grad(0), need_b = ctx->needs_input_grad(1);
    const bool need_c = has_c && ctx->needs_input_grad(2);
    // Gradients take the input's type (the library rounds BF16 adjoints
    // once from the FP32 accumulation).
    Tensor dA = need_a ? torch::empty(A.sizes(), A.options()) : Tensor();
    Tensor dB = need_b ? torch::empty(B.sizes(), B.options()) : Tensor();
    Tensor dCin = need_c ? torch::empty_like(dC) : Tensor();
    if (need_a || need_b || need_c) {
      hcblasHandle_t h = handle_for(dC);
      const float af = (float)alpha, bf = (float)beta;
      const void *pa = ot == torch::kFloat64 ? (const void *)&alpha : (const void *)&af;
      const void *pb = ot == torch::kFloat64 ? (const void *)&beta : (const void *)&bf;
      check(hcblas_gemm_strided_batched_backward(
                h, dtype_of(A), HCBLAS_ROW_MAJOR, ta ? HCBLAS_OP_T : HCBLAS_OP_N,
                tb ? HCBLAS_OP_T : HCBLAS_OP_N, (int)s.M, (int)s.N, (int)s.K, pa,
                pb, A.data_ptr(), A.size(2), A.size(1) * A.size(2),
                B.data_ptr(), B.size(2), B.size(1) * B.size(2), dC.data_ptr(),
                s.N, s.M * s.N, need_a ? dA.data_ptr() : nullptr, A.size(2),
                A.size(1) * A.size(2), need_b ? dB.data_ptr() : nullptr,
                B.size(2), B.size(1) * B.size(2),
                need_c ? dCin.data_ptr() : nullptr, s.N, s.M * s.N,
                (int)s.batch),
            "hcblas_gemm_strided_batched_backward");
    }
    return {dA, dB, dCin, Tensor(), Tensor(), Tensor(), Tensor()};
  }
};

class GerFn : public torch::autograd::Function<GerFn> {
 public:
  static Tensor forward(AutogradContext *ctx, Tensor x, Tensor y, Tensor A,
                        double alpha) {
    TORCH_CHECK(x.is_cuda() && y.is_cuda() && A.is_cuda(), "hcblas.ger: GPU tensors required");
    TORCH_CHECK(A.dim() == 2 && x.dim() == 1 && y.dim() == 1 && A.size(0) == x.size(0) &&
                    A.size(1) == y.size(0),
                "hcblas.ger: A must be [len(x), len(y)]");
    TORCH_CHECK(A.scalar_type() == x.scalar_type() && A.scalar_type() == y.scalar_type(),
                "hcblas.ger: types differ");
    x = x.contiguous();
    y = y.contiguous();
    Tensor out = A.contiguous().clone();
    const int M = (int)A.size(0), N = (int)A.size(1);
    hcblasHandle_t h = handle_for(A);
    if (A.scalar_type() == torch::kFloat32) {
      const float a = (float)alpha;
      check(hcblas_sger(h, HCBLAS_ROW_MAJOR, M, N, &a, x.data_ptr<float>(), 0, 1,
                        y.data_ptr<float>(), 0, 1, out.data_ptr<float>(), 0, N),
            "hcblas_sger");
    } else {
      TORCH_CHECK(A.scalar_type() == torch::kFloat64, "hcblas.ger: float32 or float64 only");
      check(hcblas_dger(h, HCBLAS_ROW_MAJOR, M, N, &alpha, x.data_ptr<double>(), 0,
                        1, y.data_ptr<double>(), 0, 1, out.data_ptr<double>(), 0,
                        N),
            "hcblas_dger");
    }
    ctx->save_for_backward({x, y});
    ctx->saved_data["alpha"] = alpha;
    return out;
  }

  static tensor_list backward(AutogradContext *ctx, tensor_list grads) {
    auto saved = ctx->get_saved_variables();
    Tensor x = saved[0], y = saved[1];
    const double alpha = ctx->saved_data["alpha"].toDouble();
    Tensor dA = grads[0].to(x.scalar_type()).contiguous();
    const int M = (int)x.size(0), N = (int)y.size(0);
    const bool need_x = ctx->needs_input_grad(0), need_y = ctx->needs_input_grad(1);
    Tensor dx = need_x ? torch::empty_like(x) : Tensor();
    Tensor dy = need_y ? torch::empty_like(y) : Tensor();
    if (need_x || need_y) {
      hcblasHandle_t h = handle_for(dA);
      if (x.scalar_type() == torch::kFloat32) {
        const float a = (float)alpha;
        check(hcblas_sger_backward(h, HCBLAS_ROW_MAJOR, M, N, &a, x.data_ptr<float>(), 1,
                                   y.data_ptr<float>(), 1, dA.data_ptr<float>(), N,
                                   need_x ? dx.data_ptr<float>() : nullptr,
                                   need_y ? dy.data_ptr<float>() : nullptr),
              "hcblas_sger_backward");
      } else {
        check(hcblas_dger_backward(h, HCBLAS_ROW_MAJOR, M, N, &alpha, x.data_ptr<double>(),
                                   1, y.data_ptr<double>(), 1, dA.data_ptr<double>(), N,
                                   need_x ? dx.data_ptr<double>() : nullptr,
                                   need_y ? dy.data_ptr<double>() : nullptr),
              "hcblas_dger_backward");
      }
    }
    return {dx, dy, ctx->needs_input_grad(2) ? dA : Tensor(), Tensor()};
  }
};

Tensor gemm_batched(const Tensor &A, const Tensor &B, const c10::optional<Tensor> &C,
                    double alpha, double beta, bool transA, bool transB) {
  return GemmBatchedFn::apply(A, B, C, alpha, beta, transA, transB);
}

Tensor ger(const Tensor &x, const Tensor &y, const Tensor &A, double alpha) {
  return GerFn::apply(x, y, A, alpha);
}

}  // namespace

TORCH_LIBRARY(hcblas, m) {
  m.def("gemm_batched(Tensor A, Tensor B, Tensor? C, float alpha, float beta, "
        "bool transA, bool transB) -> Tensor",
        &gemm_batched);
  m.def("ger(Tensor x, Tensor y, Tensor A, float alpha) -> Tensor", &ger);
}
