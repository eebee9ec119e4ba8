// C ABI of the learned warm start: each network is created with a hidden width (32 or 128)
// and every call goes to the build of hyq_mlp.hip for that width (namespaces mlp_h32 and
// mlp_h128, one object file each).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/hyqmom.h"

#define HYQ_MLP_DECLARE(NS)                                                                    \
  namespace NS {                                                                               \
  const char* last_error();                                                                    \
  int64_t blob_size();                                                                         \
  int64_t param_count();                                                                       \
  int create(const float* blob_host, void** out);                                              \
  int destroy(void* h);                                                                        \
  int get_blob(void* h, float* blob_host);                                                     \
  int set_blob(void* h, const float* blob_host);                                               \
  int infer_async(void* h, const float* X, int64_t n, float* P, hipStream_t stream);           \
  int infer_masked_async(void* h, const float* X, int64_t n, float* P, hipStream_t stream);    \
  int infer_bf16_masked_async(void* h, const short* Xb, int64_t n, float* P,                   \
                              hipStream_t stream);                                             \
  int infer_bf16_segments_async(void* h, const short* Xb, const uint32_t* idx,                 \
                                const uint32_t* seg, int64_t nseg, int32_t seg_rows, float* P,   \
                                hipStream_t stream);                                             \
  const float* input_norm(void* h);                                                            \
  int set_shape(void* h, int32_t shape);                                                       \
  int infer_layers(void* h, const float* X, int64_t n, float* P);                              \
  int infer(void* h, const float* X, int64_t n, float* P);                                     \
  int grad(void* h, const float* X, const float* T, int64_t n, float* loss_out,                \
           float* grad_host);                                                                  \
  int train_step(void* h, const float* X, const float* T, int64_t n, float lr,                 \
                 float* loss_out);                                                             \
  }
HYQ_MLP_DECLARE(mlp_h32)
HYQ_MLP_DECLARE(mlp_h128)

struct hyq_mlp {
  int32_t hidden;
  void* net;
};

namespace {
// the width whose error text hyq_mlp_last_error reports: that of the last call (0: an error
// of this file's own)
thread_local int32_t g_last_hidden = 32;
}  // namespace

// Calls fn of m's width (of width 32 for a null m, which then reports the null argument).
#define HYQ_MLP_CALL(m, fn, ...)                                                          \
  ((g_last_hidden = (m) ? (m)->hidden : 32) == 128 ? mlp_h128::fn((m)->net, __VA_ARGS__) \
                                                   : mlp_h32::fn((m) ? (m)->net : nullptr, __VA_ARGS__))

extern "C" {

const char* hyq_mlp_last_error(void) {
  if (g_last_hidden == 0) return "hidden width must be 32 or 128";
  return g_last_hidden == 128 ? mlp_h128::last_error() : mlp_h32::last_error();
}
int64_t hyq_mlp_blob_size(void) { return mlp_h32::blob_size(); }
int64_t hyq_mlp_param_count(void) { return mlp_h32::param_count(); }
int64_t hyq_mlp_blob_size_hidden(int32_t hidden) {
  return hidden == 32 ? mlp_h32::blob_size() : (hidden == 128 ? mlp_h128::blob_size() : -1);
}
int64_t hyq_mlp_param_count_hidden(int32_t hidden) {
  return hidden == 32 ? mlp_h32::param_count() : (hidden == 128 ? mlp_h128::param_count() : -1);
}

int hyq_mlp_create_hidden(int32_t hidden, const float* blob_host, hyq_mlp** out) {
  if (hidden != 32 && hidden != 128) {
    g_last_hidden = 0;
    return HYQ_EINVAL;
  }
  if (!out) {
    g_last_hidden = hidden;
    return hidden == 128 ? mlp_h128::create(blob_host, nullptr) : mlp_h32::create(blob_host, nullptr);
  }
  void* net = nullptr;
  g_last_hidden = hidden;
  const int rc = hidden == 128 ? mlp_h128::create(blob_host, &net) : mlp_h32::create(blob_host, &net);
  if (rc != HYQ_OK) return rc;
  *out = new hyq_mlp{hidden, net};
  return HYQ_OK;
}
int hyq_mlp_create(const float* blob_host, hyq_mlp** out) {
  return hyq_mlp_create_hidden(32, blob_host, out);
}
int32_t hyq_mlp_hidden(hyq_mlp* m) { return m ? m->hidden : 0; }

int hyq_mlp_destroy(hyq_mlp* m) {
  if (!m) return HYQ_OK;
  g_last_hidden = m->hidden;
  const int rc = m->hidden == 128 ? mlp_h128::destroy(m->net) : mlp_h32::destroy(m->net);
  delete m;
  return rc;
}
int hyq_mlp_get_blob(hyq_mlp* m, float* blob_host) { return HYQ_MLP_CALL(m, get_blob, blob_host); }
int hyq_mlp_set_blob(hyq_mlp* m, const float* blob_host) { return HYQ_MLP_CALL(m, set_blob, blob_host); }
int hyq_mlp_infer_async(hyq_mlp* m, const float* X, int64_t n, float* P, hipStream_t stream) {
  return HYQ_MLP_CALL(m, infer_async, X, n, P, stream);
}
int hyq_mlp_infer_masked_async(hyq_mlp* m, const float* X, int64_t n, float* P, hipStream_t stream) {
  return HYQ_MLP_CALL(m, infer_masked_async, X, n, P, stream);
}
int hyq_mlp_infer_bf16_masked_async(hyq_mlp* m, const short* Xb, int64_t n, float* P,
                                    hipStream_t stream) {
  return HYQ_MLP_CALL(m, infer_bf16_masked_async, Xb, n, P, stream);
}
int hyq_mlp_infer_bf16_segments_async(hyq_mlp* m, const short* Xb, const uint32_t* idx,
                                      const uint32_t* seg, int64_t nseg, int32_t seg_rows,
                                      float* P, hipStream_t stream) {
  return HYQ_MLP_CALL(m, infer_bf16_segments_async, Xb, idx, seg, nseg, seg_rows, P, stream);
}
const float* hyq_mlp_input_norm(hyq_mlp* m) {
  if (!m) return nullptr;
  return m->hidden == 128 ? mlp_h128::input_norm(m->net) : mlp_h32::input_norm(m->net);
}
int hyq_mlp_set_shape(hyq_mlp* m, int32_t shape) { return HYQ_MLP_CALL(m, set_shape, shape); }
int hyq_mlp_infer_layers(hyq_mlp* m, const float* X, int64_t n, float* P) {
  return HYQ_MLP_CALL(m, infer_layers, X, n, P);
}
int hyq_mlp_infer(hyq_mlp* m, const float* X, int64_t n, float* P) {
  return HYQ_MLP_CALL(m, infer, X, n, P);
}
int hyq_mlp_grad(hyq_mlp* m, const float* X, const float* T, int64_t n, float* loss_out,
                 float* grad_host) {
  return HYQ_MLP_CALL(m, grad, X, T, n, loss_out, grad_host);
}
int hyq_mlp_train_step(hyq_mlp* m, const float* X, const float* T, int64_t n, float lr,
                       float* loss_out) {
  return HYQ_MLP_CALL(m, train_step, X, T, n, lr, loss_out);
}

}  // extern "C"
