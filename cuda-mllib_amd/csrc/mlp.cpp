// mlp.cpp -- host side of the warm-start MLP (cumf_mlp_* in cumf_mlp.h).
//
// Training step on a batch of B rows (B a multiple of 256), all products
// C = A B^T on the BF16 GEMM kernel:
//   forward   Z1 = H0 W1^T + c1, H1 = gelu(Z1)   (H1 also transposed)
//             Z2 = H1 W2^T + c2, H2 = gelu(Z2)   (H2 also transposed)
//             Y  = H2 W3^T + c3
//   loss      dY = 2 (Y - T) / (B f)            (both layouts)
//   backward  dW3 = dY^T H2, dc3 = colsum dY
//             dZ2 = (dY W3) * gelu'(Z2),  dW2 = dZ2^T H1, dc2 = colsum dZ2
//             dZ1 = (dZ2 W2) * gelu'(Z1), dW1 = dZ1^T H0, dc1 = colsum dZ1
//   Adam on the FP32 master weights, then BF16 copies W^T-major and W-major.
// Weight gradients reduce over the batch with split-K slices summed in
// order (deterministic).
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/cumf_mlp.h"
#include "cumf_internal.h"
#include "mlp_internal.h"
#include "mlp_kernels.h"

using cumf::DevBuf;

namespace {

constexpr int kHidden = 256;
constexpr int kSplitK = 8;
constexpr int kLossBlocks = 256;
constexpr int kPredictChunk = 65536;

int round_up(int v, int m) { return (v + m - 1) / m * m; }

uint64_t splitmix_next(uint64_t& s) {
  s += 0x9E3779B97F4A7C15ull;
  uint64_t z = s;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

}  // namespace

struct cumf_mlp {
  int device = 0;
  int f = 0, K0 = 0, H1 = kHidden, H2 = kHidden, N3 = 0;
  int Kin[3] = {0, 0, 0}, Nout[3] = {0, 0, 0};
  int64_t off[6] = {0, 0, 0, 0, 0, 0};
  int64_t nparams = 0;
  hipStream_t stream = nullptr;
  DevBuf P, Mm, V, G;       // FP32 master weights, Adam moments, gradient
  DevBuf Wt[3], W[3];       // BF16 copies (N x K and K x N)
  int64_t step = 0;
  // Batch workspace (rows = cap).
  int cap = 0;
  DevBuf H0, H0t, Z1, H1b, H1t, Z2, H2b, H2t, Y, T, dY, dYt, dZ2, dZ2t, dZ1t, Wpart, loss_part, perm;
  // Harvested training set.
  DevBuf hfeat, htarg;
  int64_t hcount = 0, hcap = 0;

  int ensure_batch(int rows) {
    if (rows <= cap) return CUMF_OK;
    const size_t B = (size_t)rows;
    CUMF_TRY(H0.alloc(2 * B * K0));
    CUMF_TRY(H0t.alloc(2 * B * K0));
    CUMF_TRY(Z1.alloc(4 * B * H1));
    CUMF_TRY(H1b.alloc(2 * B * H1));
    CUMF_TRY(H1t.alloc(2 * B * H1));
    CUMF_TRY(Z2.alloc(4 * B * H2));
    CUMF_TRY(H2b.alloc(2 * B * H2));
    CUMF_TRY(H2t.alloc(2 * B * H2));
    CUMF_TRY(Y.alloc(4 * B * N3));
    CUMF_TRY(T.alloc(4 * B * f));
    CUMF_TRY(dY.alloc(2 * B * N3));
    CUMF_TRY(dYt.alloc(2 * B * N3));
    CUMF_TRY(dZ2.alloc(2 * B * H2));
    CUMF_TRY(dZ2t.alloc(2 * B * H2));
    CUMF_TRY(dZ1t.alloc(2 * B * H1));
    CUMF_TRY(perm.alloc(8 * B));
    cap = rows;
    return CUMF_OK;
  }

  int publish() {
    for (int l = 0; l < 3; ++l)
      CUMF_HIP_TRY(cumf::launch_mlp_publish(P.as<float>() + off[2 * l], Nout[l], Kin[l], Wt[l].as<__bf16>(),
                                            W[l].as<__bf16>(), stream));
    return CUMF_OK;
  }

  cumf::GemmArgs gemm(const __bf16* A, int64_t lda, const __bf16* B, int64_t ldb, int M, int N, int K, int epi) {
    cumf::GemmArgs a{};
    a.A = A;
    a.lda = lda;
    a.B = B;
    a.ldb = ldb;
    a.M = M;
    a.N = N;
    a.K = K;
    a.epi = epi;
    return a;
  }

  // Forward over `rows` rows of H0-layout features; with `train`, keeps
  // what the backward pass needs, else writes predictions to out (stride
  // ldo, f columns).
  int forward(const __bf16* feats, int rows, bool train, float* out, int64_t ldo) {
    const float* Pf = P.as<float>();
    auto a1 = gemm(feats, K0, Wt[0].as<__bf16>(), K0, rows, H1, K0, cumf::EPI_FWD_GELU);
    a1.bias = Pf + off[1];
    a1.C = train ? Z1.as<float>() : nullptr;
    a1.ldc = H1;
    a1.H = H1b.as<__bf16>();
    a1.ldh = H1;
    a1.Ht = train ? H1t.as<__bf16>() : nullptr;
    a1.ldht = rows;
    CUMF_HIP_TRY(cumf::launch_mlp_gemm(a1, 1, stream));
    auto a2 = gemm(H1b.as<__bf16>(), H1, Wt[1].as<__bf16>(), H1, rows, H2, H1, cumf::EPI_FWD_GELU);
    a2.bias = Pf + off[3];
    a2.C = train ? Z2.as<float>() : nullptr;
    a2.ldc = H2;
    a2.H = H2b.as<__bf16>();
    a2.ldh = H2;
    a2.Ht = train ? H2t.as<__bf16>() : nullptr;
    a2.ldht = rows;
    CUMF_HIP_TRY(cumf::launch_mlp_gemm(a2, 1, stream));
    auto a3 = gemm(H2b.as<__bf16>(), H2, Wt[2].as<__bf16>(), H2, rows, train ? N3 : f, H2, cumf::EPI_LINEAR);
    a3.bias = Pf + off[5];
    a3.C = train ? Y.as<float>() : out;
    a3.ldc = train ? N3 : ldo;
    a3.mlimit = rows;
    a3.nlimit = train ? N3 : f;
    CUMF_HIP_TRY(cumf::launch_mlp_gemm(a3, 1, stream));
    return CUMF_OK;
  }

  int wgrad(const __bf16* Dt, int N, const __bf16* Xt, int K, int B, int layer) {
    auto a = gemm(Dt, B, Xt, B, N, K, B, cumf::EPI_STORE);
    a.C = Wpart.as<float>();
    a.ldc = K;
    a.split_stride = (int64_t)N * K;
    CUMF_HIP_TRY(cumf::launch_mlp_gemm(a, kSplitK, stream));
    CUMF_HIP_TRY(cumf::launch_mlp_reduce(Wpart.as<float>(), (int64_t)N * K, kSplitK, (int64_t)N * K,
                                         G.as<float>() + off[2 * layer], stream));
    CUMF_HIP_TRY(cumf::launch_mlp_rowsum(Dt, N, B, G.as<float>() + off[2 * layer + 1], stream));
    return CUMF_OK;
  }

  // Forward + backward on the batch in H0 / H0t / T; gradient into G.
  int backprop(int B) {
    CUMF_TRY(forward(H0.as<__bf16>(), B, true, nullptr, 0));
    CUMF_HIP_TRY(cumf::launch_mlp_loss(Y.as<float>(), N3, T.as<float>(), f, B, f, N3, dY.as<__bf16>(),
                                       dYt.as<__bf16>(), loss_part.as<double>(), kLossBlocks, stream));
    CUMF_TRY(wgrad(dYt.as<__bf16>(), N3, H2t.as<__bf16>(), H2, B, 2));
    auto d2 = gemm(dY.as<__bf16>(), N3, W[2].as<__bf16>(), N3, B, H2, N3, cumf::EPI_BWD_GELU);
    d2.Z = Z2.as<float>();
    d2.ldz = H2;
    d2.H = dZ2.as<__bf16>();
    d2.ldh = H2;
    d2.Ht = dZ2t.as<__bf16>();
    d2.ldht = B;
    CUMF_HIP_TRY(cumf::launch_mlp_gemm(d2, 1, stream));
    CUMF_TRY(wgrad(dZ2t.as<__bf16>(), H2, H1t.as<__bf16>(), H1, B, 1));
    auto d1 = gemm(dZ2.as<__bf16>(), H2, W[1].as<__bf16>(), H2, B, H1, H2, cumf::EPI_BWD_GELU);
    d1.Z = Z1.as<float>();
    d1.ldz = H1;
    d1.Ht = dZ1t.as<__bf16>();
    d1.ldht = B;
    CUMF_HIP_TRY(cumf::launch_mlp_gemm(d1, 1, stream));
    CUMF_TRY(wgrad(dZ1t.as<__bf16>(), H1, H0t.as<__bf16>(), K0, B, 0));
    return CUMF_OK;
  }

  int adam(float lr) {
    ++step;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = (float)(1.0 / (1.0 - std::pow(0.9, (double)step)));
    const float c2 = (float)(1.0 / (1.0 - std::pow(0.999, (double)step)));
    CUMF_HIP_TRY(cumf::launch_mlp_adam(P.as<float>(), G.as<float>(), Mm.as<float>(), V.as<float>(), nparams, lr, b1,
                                       b2, 1e-8f, c1, c2, stream));
    return publish();
  }

  // Mean squared error of the batch just run (blocks).
  int batch_loss(int B, double* out) {
    std::vector<double> part(kLossBlocks);
    CUMF_HIP_TRY(hipMemcpyAsync(part.data(), loss_part.p, sizeof(double) * kLossBlocks, hipMemcpyDeviceToHost,
                                stream));
    CUMF_HIP_TRY(hipStreamSynchronize(stream));
    double s = 0.0;
    for (double v : part) s += v;
    *out = s / ((double)B * f);
    return CUMF_OK;
  }

  // Host batch -> H0, H0t (BF16) and T.
  int upload_batch(const float* feats, const float* targets, int64_t rows) {
    CUMF_TRY(ensure_batch((int)rows));
    DevBuf tmp;
    CUMF_TRY(tmp.alloc(sizeof(float) * (size_t)rows * K0));
    CUMF_HIP_TRY(hipMemcpyAsync(tmp.p, feats, tmp.bytes, hipMemcpyHostToDevice, stream));
    DevBuf fb;
    CUMF_TRY(fb.alloc(2 * (size_t)rows * K0));
    CUMF_HIP_TRY(cumf::launch_mlp_to_bf16(tmp.as<float>(), fb.as<__bf16>(), rows * K0, stream));
    DevBuf tg;
    if (targets) {
      CUMF_TRY(tg.alloc(sizeof(float) * (size_t)rows * f));
      CUMF_HIP_TRY(hipMemcpyAsync(tg.p, targets, tg.bytes, hipMemcpyHostToDevice, stream));
    }
    std::vector<int64_t> id((size_t)rows);
    std::iota(id.begin(), id.end(), 0);
    CUMF_HIP_TRY(hipMemcpyAsync(perm.p, id.data(), 8 * id.size(), hipMemcpyHostToDevice, stream));
    DevBuf zt;
    if (!targets) {
      CUMF_TRY(zt.alloc(sizeof(float) * (size_t)rows * f));
      CUMF_HIP_TRY(hipMemsetAsync(zt.p, 0, zt.bytes, stream));
    }
    CUMF_HIP_TRY(cumf::launch_mlp_gather(fb.as<__bf16>(), K0, targets ? tg.as<float>() : zt.as<float>(), f,
                                         perm.as<int64_t>(), (int)rows, H0.as<__bf16>(), H0t.as<__bf16>(),
                                         T.as<float>(), stream));
    CUMF_HIP_TRY(hipStreamSynchronize(stream));  // temporaries die here
    return CUMF_OK;
  }
};

namespace cumf {

int mlp_f(const cumf_mlp* m) { return m->f; }
int mlp_k0(const cumf_mlp* m) { return m->K0; }
int mlp_device(const cumf_mlp* m) { return m->device; }

int mlp_prepare_predict(cumf_mlp* m) { return m->ensure_batch(kPredictChunk); }

int mlp_predict_device(cumf_mlp* m, const __bf16* feats, int64_t rows, float* x, int64_t ldx, hipStream_t s) {
  if (m->cap < kPredictChunk) return CUMF_ERR_STATE;
  hipStream_t keep = m->stream;
  m->stream = s;
  int rc = CUMF_OK;
  for (int64_t r0 = 0; rc == CUMF_OK && r0 < rows; r0 += kPredictChunk) {
    const int b = (int)std::min<int64_t>(kPredictChunk, rows - r0);
    rc = m->forward(feats + r0 * m->K0, b, false, x + r0 * ldx, ldx);
  }
  m->stream = keep;
  return rc;
}

int mlp_harvest_reserve(cumf_mlp* m, int64_t rows, __bf16** feats, float** targets) {
  if (m->hcount + rows > m->hcap) {
    const int64_t ncap = std::max<int64_t>(m->hcount + rows, 2 * m->hcap);
    DevBuf nf, nt;
    CUMF_TRY(nf.alloc(2 * (size_t)ncap * m->K0));
    CUMF_TRY(nt.alloc(sizeof(float) * (size_t)ncap * m->f));
    if (m->hcount > 0) {
      CUMF_HIP_TRY(hipMemcpy(nf.p, m->hfeat.p, 2 * (size_t)m->hcount * m->K0, hipMemcpyDeviceToDevice));
      CUMF_HIP_TRY(hipMemcpy(nt.p, m->htarg.p, sizeof(float) * (size_t)m->hcount * m->f, hipMemcpyDeviceToDevice));
    }
    std::swap(m->hfeat.p, nf.p);
    std::swap(m->hfeat.bytes, nf.bytes);
    std::swap(m->htarg.p, nt.p);
    std::swap(m->htarg.bytes, nt.bytes);
    m->hcap = ncap;
  }
  *feats = m->hfeat.as<__bf16>() + m->hcount * m->K0;
  *targets = m->htarg.as<float>() + m->hcount * m->f;
  m->hcount += rows;
  return CUMF_OK;
}

}  // namespace cumf

extern "C" {

int cumf_mlp_create(int device, int f, uint64_t seed, cumf_mlp** out) {
  cumf::clear_last_error();
  if (!out) return CUMF_ERR_INVALID;
  *out = nullptr;
  if (f < 1 || f > 140) {
    cumf::set_last_error("f must be in [1, 140]");
    return CUMF_ERR_INVALID;
  }
  if (device < 0 && hipGetDevice(&device) != hipSuccess) {
    (void)hipGetLastError();
    cumf::set_last_error("no usable GPU device");
    return CUMF_ERR_NODEVICE;
  }
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device >= count) {
    (void)hipGetLastError();
    cumf::set_last_error("no usable GPU device");
    return CUMF_ERR_NODEVICE;
  }
  int prev = -1;
  (void)hipGetDevice(&prev);
  (void)hipSetDevice(device);
  auto* m = new cumf_mlp();
  m->device = device;
  m->f = f;
  m->K0 = round_up(2 * f + 1, 32);
  m->N3 = round_up(f, 32);
  const int Kin[3] = {m->K0, m->H1, m->H2}, Nout[3] = {m->H1, m->H2, m->N3};
  const int kreal[3] = {2 * f + 1, m->H1, m->H2}, nreal[3] = {m->H1, m->H2, f};
  int64_t o = 0;
  for (int l = 0; l < 3; ++l) {
    m->Kin[l] = Kin[l];
    m->Nout[l] = Nout[l];
    m->off[2 * l] = o;
    o += (int64_t)Nout[l] * Kin[l];
    m->off[2 * l + 1] = o;
    o += Nout[l];
  }
  m->nparams = o;
  std::vector<float> p((size_t)o, 0.0f);
  uint64_t s = seed;
  for (int l = 0; l < 3; ++l) {
    const float bound = (float)(1.0 / std::sqrt((double)kreal[l]));
    for (int n = 0; n < nreal[l]; ++n)
      for (int k = 0; k < kreal[l]; ++k) {
        const float u = (float)(splitmix_next(s) >> 40) / (float)(1 << 24);
        p[(size_t)(m->off[2 * l] + (int64_t)n * Kin[l] + k)] = (2.0f * u - 1.0f) * bound;
      }
  }
  size_t wpart = 0;
  for (int l = 0; l < 3; ++l) wpart = std::max(wpart, (size_t)Nout[l] * Kin[l]);
  int rc = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) == hipSuccess ? CUMF_OK : CUMF_ERR_HIP;
  if (rc == CUMF_OK) rc = m->P.alloc(sizeof(float) * o);
  if (rc == CUMF_OK) rc = m->Mm.alloc(sizeof(float) * o);
  if (rc == CUMF_OK) rc = m->V.alloc(sizeof(float) * o);
  if (rc == CUMF_OK) rc = m->G.alloc(sizeof(float) * o);
  if (rc == CUMF_OK) rc = m->Wpart.alloc(sizeof(float) * wpart * kSplitK);
  if (rc == CUMF_OK) rc = m->loss_part.alloc(sizeof(double) * kLossBlocks);
  for (int l = 0; l < 3 && rc == CUMF_OK; ++l) {
    rc = m->Wt[l].alloc(2 * (size_t)Nout[l] * Kin[l]);
    if (rc == CUMF_OK) rc = m->W[l].alloc(2 * (size_t)Nout[l] * Kin[l]);
  }
  if (rc == CUMF_OK) rc = cumf_mlp_set_params(m, p.data());
  if (rc != CUMF_OK) {
    if (rc == CUMF_ERR_HIP) cumf::set_last_error("cumf_mlp_create: HIP error");
    cumf_mlp_destroy(m);
    if (prev >= 0) (void)hipSetDevice(prev);
    return rc;
  }
  if (prev >= 0) (void)hipSetDevice(prev);
  *out = m;
  return CUMF_OK;
}

int cumf_mlp_destroy(cumf_mlp* m) {
  if (!m) return CUMF_ERR_INVALID;
  int prev = -1;
  (void)hipGetDevice(&prev);
  (void)hipSetDevice(m->device);
  if (m->stream) {
    (void)hipStreamSynchronize(m->stream);
    (void)hipStreamDestroy(m->stream);
  }
  delete m;
  if (prev >= 0) (void)hipSetDevice(prev);
  return CUMF_OK;
}

int cumf_mlp_dims(const cumf_mlp* m, int32_t dims[4], int64_t* count) {
  if (!m) return CUMF_ERR_INVALID;
  if (dims) {
    dims[0] = m->K0;
    dims[1] = m->H1;
    dims[2] = m->H2;
    dims[3] = m->N3;
  }
  if (count) *count = m->nparams;
  return CUMF_OK;
}

int cumf_mlp_get_params(cumf_mlp* m, float* params) {
  if (!m || !params) return CUMF_ERR_INVALID;
  CUMF_HIP_TRY(hipSetDevice(m->device));
  CUMF_HIP_TRY(hipMemcpyAsync(params, m->P.p, m->P.bytes, hipMemcpyDeviceToHost, m->stream));
  CUMF_HIP_TRY(hipStreamSynchronize(m->stream));
  return CUMF_OK;
}

int cumf_mlp_set_params(cumf_mlp* m, const float* params) {
  if (!m || !params) return CUMF_ERR_INVALID;
  CUMF_HIP_TRY(hipSetDevice(m->device));
  CUMF_HIP_TRY(hipMemcpyAsync(m->P.p, params, m->P.bytes, hipMemcpyHostToDevice, m->stream));
  CUMF_HIP_TRY(hipMemsetAsync(m->Mm.p, 0, m->Mm.bytes, m->stream));
  CUMF_HIP_TRY(hipMemsetAsync(m->V.p, 0, m->V.bytes, m->stream));
  m->step = 0;
  CUMF_TRY(m->publish());
  CUMF_HIP_TRY(hipStreamSynchronize(m->stream));
  return CUMF_OK;
}

int cumf_mlp_predict_host(cumf_mlp* m, const float* feats, int64_t rows, float* out) {
  if (!m || !feats || !out || rows <= 0 || rows > (1 << 24)) return CUMF_ERR_INVALID;
  CUMF_HIP_TRY(hipSetDevice(m->device));
  CUMF_TRY(m->upload_batch(feats, nullptr, rows));
  DevBuf y;
  CUMF_TRY(y.alloc(sizeof(float) * (size_t)rows * m->f));
  CUMF_TRY(m->forward(m->H0.as<__bf16>(), (int)rows, false, y.as<float>(), m->f));
  CUMF_HIP_TRY(hipMemcpyAsync(out, y.p, y.bytes, hipMemcpyDeviceToHost, m->stream));
  CUMF_HIP_TRY(hipStreamSynchronize(m->stream));
  return CUMF_OK;
}

int cumf_mlp_gradients_host(cumf_mlp* m, const float* feats, const float* targets, int64_t rows, float* grads,
                            double* loss) {
  if (!m || !feats || !targets || rows <= 0 || rows % 256 || rows > (1 << 24)) {
    cumf::set_last_error("rows must be a positive multiple of 256");
    return CUMF_ERR_INVALID;
  }
  CUMF_HIP_TRY(hipSetDevice(m->device));
  CUMF_TRY(m->upload_batch(feats, targets, rows));
  CUMF_TRY(m->backprop((int)rows));
  if (loss) CUMF_TRY(m->batch_loss((int)rows, loss));
  if (grads) {
    CUMF_HIP_TRY(hipMemcpyAsync(grads, m->G.p, m->G.bytes, hipMemcpyDeviceToHost, m->stream));
    CUMF_HIP_TRY(hipStreamSynchronize(m->stream));
  }
  return CUMF_OK;
}

int cumf_mlp_step_host(cumf_mlp* m, const float* feats, const float* targets, int64_t rows, float lr, double* loss) {
  CUMF_TRY(cumf_mlp_gradients_host(m, feats, targets, rows, nullptr, loss));
  CUMF_TRY(m->adam(lr));
  CUMF_HIP_TRY(hipStreamSynchronize(m->stream));
  return CUMF_OK;
}

int64_t cumf_mlp_harvested(const cumf_mlp* m) { return m ? m->hcount : -1; }

int cumf_mlp_clear_harvest(cumf_mlp* m) {
  if (!m) return CUMF_ERR_INVALID;
  m->hcount = 0;
  return CUMF_OK;
}

int cumf_mlp_train(cumf_mlp* m, int epochs, int batch, float lr, uint64_t seed, double* loss) {
  if (!m || epochs < 0 || batch <= 0 || batch % 256 || !(lr > 0.f)) {
    cumf::set_last_error("cumf_mlp_train: epochs >= 0, batch a positive multiple of 256, lr > 0");
    return CUMF_ERR_INVALID;
  }
  if (m->hcount < batch) {
    cumf::set_last_error("cumf_mlp_train: fewer harvested rows than one batch");
    return CUMF_ERR_STATE;
  }
  CUMF_HIP_TRY(hipSetDevice(m->device));
  CUMF_TRY(m->ensure_batch(batch));
  DevBuf dperm;
  CUMF_TRY(dperm.alloc(sizeof(int64_t) * (size_t)m->hcount));
  std::vector<int64_t> order((size_t)m->hcount);
  uint64_t s = seed;
  const int64_t steps = m->hcount / batch;
  double last = 0.0;
  for (int ep = 0; ep < epochs; ++ep) {
    std::iota(order.begin(), order.end(), 0);
    for (int64_t i = (int64_t)order.size() - 1; i > 0; --i) {  // Fisher-Yates
      const int64_t j = (int64_t)(splitmix_next(s) % (uint64_t)(i + 1));
      std::swap(order[(size_t)i], order[(size_t)j]);
    }
    CUMF_HIP_TRY(hipMemcpyAsync(dperm.p, order.data(), dperm.bytes, hipMemcpyHostToDevice, m->stream));
    double ep_loss = 0.0;
    for (int64_t st = 0; st < steps; ++st) {
      CUMF_HIP_TRY(cumf::launch_mlp_gather(m->hfeat.as<__bf16>(), m->K0, m->htarg.as<float>(), m->f,
                                           dperm.as<int64_t>() + st * batch, batch, m->H0.as<__bf16>(),
                                           m->H0t.as<__bf16>(), m->T.as<float>(), m->stream));
      CUMF_TRY(m->backprop(batch));
      CUMF_TRY(m->adam(lr));
      if (st + 1 == steps || (st & 63) == 0) {
        double l = 0.0;
        CUMF_TRY(m->batch_loss(batch, &l));
        ep_loss = l;
      }
    }
    last = ep_loss;
  }
  CUMF_HIP_TRY(hipStreamSynchronize(m->stream));
  if (loss) *loss = last;
  return CUMF_OK;
}

}  // extern "C"
