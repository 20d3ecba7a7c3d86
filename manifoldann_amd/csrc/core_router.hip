// Learned probe router (no reference counterpart): an MLP d -> h (ReLU) -> nlist
// scoring the IVF lists of a query, trained on the GPU with softmax
// cross-entropy against the lists that hold the query's exact neighbours.  Its
// negated scores replace the centroid distances in TopKRouting, i.e. it plugs
// into the reference's AbstractTransform seam (Transform.jl:36-83) as a
// transform returning KMeansAssignment(-scores).
#include <cmath>
#include <cstring>
#include <vector>

#include "core.h"
#include "mlp.h"

namespace mann {

static uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static double unif(uint64_t& s) { return (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

mann_router* router_create(int d, int h, int c, uint64_t seed) {
  auto* r = new mann_router();
  try {
    MANN_HIP(hipGetDevice(&r->device));
    r->d = d;
    r->h = h;
    r->c = c;
    r->nparam = (int64_t)h * d + h + (int64_t)c * h + c;
    MANN_HIP(hipMalloc(&r->w, sizeof(float) * r->nparam));
    MANN_HIP(hipMalloc(&r->m1, sizeof(float) * r->nparam));
    MANN_HIP(hipMalloc(&r->m2, sizeof(float) * r->nparam));
    MANN_HIP(hipMemset(r->m1, 0, sizeof(float) * r->nparam));
    MANN_HIP(hipMemset(r->m2, 0, sizeof(float) * r->nparam));
    // He-uniform weights, zero biases
    std::vector<float> hw(r->nparam, 0.0f);
    uint64_t s = seed;
    const double a1 = std::sqrt(6.0 / d), a2 = std::sqrt(6.0 / h);
    for (int64_t i = 0; i < (int64_t)h * d; ++i) hw[i] = (float)((2.0 * unif(s) - 1.0) * a1);
    float* w2 = hw.data() + (int64_t)h * d + h;
    for (int64_t i = 0; i < (int64_t)c * h; ++i) w2[i] = (float)((2.0 * unif(s) - 1.0) * a2);
    MANN_HIP(hipMemcpy(r->w, hw.data(), sizeof(float) * r->nparam, hipMemcpyHostToDevice));
  } catch (...) {
    delete r;
    throw;
  }
  return r;
}

void router_set_prior(mann_router& r, const mann_ivf* iv, float gamma) {
  float** bufs[] = {&r.pc, &r.pcn, &r.pc2, &r.pb};
  for (float** b : bufs) {
    if (*b) (void)hipFree(*b);
    *b = nullptr;
  }
  r.gamma = 0.0f;
  if (gamma == 0.0f || !iv) return;
  MANN_REQUIRE(r.d <= prior_dmax(), ERR_UNSUPPORTED, "router prior supports d <= 1024");
  MANN_HIP(hipMalloc(&r.pc, sizeof(float) * (size_t)r.c * iv->d16));
  MANN_HIP(hipMalloc(&r.pcn, sizeof(float) * (size_t)r.c));
  MANN_HIP(hipMemcpy(r.pc, iv->c32, sizeof(float) * (size_t)r.c * iv->d16,
                     hipMemcpyDeviceToDevice));
  MANN_HIP(hipMemcpy(r.pcn, iv->cn2, sizeof(float) * (size_t)r.c, hipMemcpyDeviceToDevice));
  r.pld = iv->d16;
  r.gamma = gamma;
  // 2 gamma c (one rounding per element) and -gamma |c|^2 for the tiled GEMM
  const size_t nc = (size_t)r.c * iv->d16;
  std::vector<float> hc(nc), hn(r.c);
  MANN_HIP(hipMemcpy(hc.data(), iv->c32, sizeof(float) * nc, hipMemcpyDeviceToHost));
  MANN_HIP(hipMemcpy(hn.data(), iv->cn2, sizeof(float) * r.c, hipMemcpyDeviceToHost));
  const float g2 = 2.0f * gamma;
  for (float& v : hc) v *= g2;
  for (float& v : hn) v *= -gamma;
  MANN_HIP(hipMalloc(&r.pc2, sizeof(float) * nc));
  MANN_HIP(hipMalloc(&r.pb, sizeof(float) * (size_t)r.c));
  MANN_HIP(hipMemcpy(r.pc2, hc.data(), sizeof(float) * nc, hipMemcpyHostToDevice));
  MANN_HIP(hipMemcpy(r.pb, hn.data(), sizeof(float) * r.c, hipMemcpyHostToDevice));
}

// Z = W2 relu(W1 x + b1) + b2 (+ the centroid prior) for B rows of X (row
// stride ldx); H kept for backward.  The prior has no parameters, so the
// backward pass is unchanged.
static void router_forward(const mann_router& r, const float* X, int64_t ldx, int B, float* H,
                           float* Z, hipStream_t st, bool negate = false) {
  GemmArgs g{};
  g.A = X;
  g.sam = ldx;
  g.sak = 1;
  g.B = r.W1();
  g.sbn = r.d;
  g.sbk = 1;
  g.C = H;
  g.ldc = r.h;
  g.M = B;
  g.N = r.h;
  g.K = r.d;
  g.bias = r.b1();
  g.relu = 1;
  launch_gemm(g, st);
  GemmArgs g2{};
  g2.A = H;
  g2.sam = r.h;
  g2.sak = 1;
  g2.B = r.W2();
  g2.sbn = r.h;
  g2.sbk = 1;
  g2.C = Z;
  g2.ldc = r.c;
  g2.M = B;
  g2.N = r.c;
  g2.K = r.h;
  g2.bias = r.b2();
  const bool prior = r.gamma != 0.0f && r.pc;
  // the last GEMM stores -Z when asked (and when the tiled kernel takes it)
  g2.negate = negate && !prior && gemm_tiled_ok(g2);
  launch_gemm(g2, st);
  bool negated = g2.negate;
  if (prior) {
    // the prior on the MFMA with split operands (FP32-grade products), else
    // the FP32 kernel
    GemmArgs gp{};
    gp.A = X;
    gp.sam = ldx;
    gp.sak = 1;
    gp.B = r.pc2;
    gp.sbn = r.pld;
    gp.sbk = 1;
    gp.C = Z;
    gp.ldc = r.c;
    gp.M = B;
    gp.N = r.c;
    gp.K = r.d;
    gp.bias = r.pb;
    gp.accumulate = 1;
    gp.split = 1;
    if (gemm_tiled_ok(gp)) {
      gp.negate = negate;
      launch_gemm(gp, st);
      negated = negate;
    } else {
      launch_prior(X, ldx, B, r.d, r.pc, r.pld, r.pcn, r.c, r.gamma, Z, r.c, st);
    }
  }
  if (negate && !negated) launch_neg(Z, Z, (int64_t)B * r.c, st);
}

// A GEMM whose operands have unit-stride K after transposing the strided ones
// into thread scratch (a K-major copy: row stride K rounded up to 4), so the
// tiled MFMA kernel takes it; the values, their bf16 rounding and the K order
// of the accumulation are those of the strided form.
static void gemm_k_major(GemmArgs g, const char* ta, const char* tb, hipStream_t st) {
  ThreadCtx& c = ctx();
  const int64_t ldk = round_up(g.K, 4);
  if (g.sak != 1 && g.sam == 1) {
    float* At = c.buf(ta).as<float>((size_t)g.M * ldk);
    launch_transpose(g.A, g.sak, g.K, g.M, At, ldk, st);  // A is K x M with row stride sak
    g.A = At;
    g.sam = ldk;
    g.sak = 1;
  }
  if (g.sbk != 1 && g.sbn == 1) {
    float* Bt = c.buf(tb).as<float>((size_t)g.N * ldk);
    launch_transpose(g.B, g.sbk, g.K, g.N, Bt, ldk, st);
    g.B = Bt;
    g.sbn = ldk;
    g.sbk = 1;
  }
  launch_gemm(g, st);
}

// Gradients of the mean cross-entropy into grad (same layout as the parameters).
// Z is overwritten with dL/dZ.  Returns nothing; loss accumulated into *loss.
static void router_backward(const mann_router& r, const float* X, int64_t ldx, int B,
                            const int* labels, int L, float* H, float* Z, float* dH, float* grad,
                            float* loss, hipStream_t st) {
  launch_softmax_xent(Z, r.c, B, labels, L, loss, st);
  float* gW1 = grad;
  float* gb1 = gW1 + (int64_t)r.h * r.d;
  float* gW2 = gb1 + r.h;
  float* gb2 = gW2 + (int64_t)r.c * r.h;
  GemmArgs a{};  // dW2[i][j] = sum_b G[b][i] H[b][j]
  a.A = Z;
  a.sam = 1;
  a.sak = r.c;
  a.B = H;
  a.sbn = 1;
  a.sbk = r.h;
  a.C = gW2;
  a.ldc = r.h;
  a.M = r.c;
  a.N = r.h;
  a.K = B;
  gemm_k_major(a, "rt_tA", "rt_tB", st);
  launch_colsum(Z, B, r.c, gb2, st);
  GemmArgs b{};  // dH[b][j] = sum_i G[b][i] W2[i][j]
  b.A = Z;
  b.sam = r.c;
  b.sak = 1;
  b.B = r.W2();
  b.sbn = 1;
  b.sbk = r.h;
  b.C = dH;
  b.ldc = r.h;
  b.M = B;
  b.N = r.h;
  b.K = r.c;
  gemm_k_major(b, "rt_tA", "rt_tB", st);
  launch_relu_mask(dH, H, (int64_t)B * r.h, st);
  GemmArgs c{};  // dW1[j][t] = sum_b dH[b][j] X[b][t]
  c.A = dH;
  c.sam = 1;
  c.sak = r.h;
  c.B = X;
  c.sbn = 1;
  c.sbk = ldx;
  c.C = gW1;
  c.ldc = r.d;
  c.M = r.h;
  c.N = r.d;
  c.K = B;
  gemm_k_major(c, "rt_tA", "rt_tB", st);
  launch_colsum(dH, B, r.h, gb1, st);
}

void router_scores(const mann_router& r, const float* X, int64_t ldx, int64_t n, float* out,
                   bool negate) {
  ThreadCtx& c = ctx();
  hipStream_t st = c.stream;
  // balanced chunks of at most 65,536 rows (the hidden layer is kept per chunk)
  const int64_t nch = std::max<int64_t>(1, ceil_div(n, (int64_t)65536));
  const int CH = (int)ceil_div(n, nch);
  float* H = c.buf("rt_H").as<float>((size_t)CH * r.h);
  ScopedTime tm("router_forward", st);
  for (int64_t s = 0; s < n; s += CH) {
    const int B = (int)std::min<int64_t>(CH, n - s);
    router_forward(r, X + s * ldx, ldx, B, H, out + s * r.c, st, negate);
  }
}

void router_grad(const mann_router& r, const float* X, int64_t ldx, int B, const int* labels,
                 int L, float* grad, float* loss) {
  ThreadCtx& c = ctx();
  hipStream_t st = c.stream;
  float* H = c.buf("rt_H").as<float>((size_t)B * r.h);
  float* Z = c.buf("rt_Z").as<float>((size_t)B * r.c);
  float* dH = c.buf("rt_dH").as<float>((size_t)B * r.h);
  fill_f32(loss, 1, 0.0f, st);
  router_forward(r, X, ldx, B, H, Z, st);
  router_backward(r, X, ldx, B, labels, L, H, Z, dH, grad, loss, st);
}

// Mini-batch Adam over n rows of X (d16-strided device rows) with labels
// [n][L] (0-based lists, -1 = none).  Returns the mean loss of the last epoch.
double router_train(mann_router& r, const float* X, int64_t ldx, int64_t n, const int* labels,
                    int L, int epochs, int batch, float lr, uint64_t seed) {
  ThreadCtx& c = ctx();
  hipStream_t st = c.stream;
  float* Xb = c.buf("rt_Xb").as<float>((size_t)batch * ldx);
  int* Lb = c.buf("rt_Lb").as<int>((size_t)batch * L);
  float* grad = c.buf("rt_grad").as<float>(r.nparam);
  float* loss = c.buf("rt_loss").as<float>(1);
  float* H = c.buf("rt_H").as<float>((size_t)batch * r.h);
  float* Z = c.buf("rt_Z").as<float>((size_t)batch * r.c);
  float* dH = c.buf("rt_dH").as<float>((size_t)batch * r.h);
  int* idx = c.buf("rt_idx").as<int>(n);
  std::vector<int> perm(n);
  uint64_t s = seed;
  double last = 0.0;
  ScopedTime tm("router_train", st);
  for (int ep = 0; ep < epochs; ++ep) {
    for (int64_t i = 0; i < n; ++i) perm[i] = (int)i;
    for (int64_t i = n - 1; i >= 1; --i) {
      int64_t j = (int64_t)(unif(s) * (double)(i + 1));
      if (j > i) j = i;
      std::swap(perm[i], perm[j]);
    }
    MANN_HIP(hipMemcpyAsync(idx, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
    fill_f32(loss, 1, 0.0f, st);
    int nbatch = 0;
    for (int64_t b0 = 0; b0 + batch <= n; b0 += batch) {
      gather_rows_f32(X, idx + b0, batch, (int)ldx, Xb, st);
      gather_rows_i32(labels, idx + b0, batch, L, Lb, st);
      router_forward(r, Xb, ldx, batch, H, Z, st);
      router_backward(r, Xb, ldx, batch, Lb, L, H, Z, dH, grad, loss, st);
      r.step += 1;
      launch_adam(r.w, grad, r.m1, r.m2, r.nparam, lr, 0.9f, 0.999f, 1e-8f, r.step, st);
      ++nbatch;
    }
    float hl = 0.0f;
    MANN_HIP(hipMemcpyAsync(&hl, loss, sizeof(float), hipMemcpyDeviceToHost, st));
    MANN_HIP(hipStreamSynchronize(st));
    last = nbatch ? hl / nbatch : 0.0;
  }
  return last;
}

}  // namespace mann
