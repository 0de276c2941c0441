// C ABI of the spectral neural operator surrogate (include/gnlse.h,
// gnlse_sno_*): parameter storage, forward, loss + backward, Adam, and a
// general bf16 GEMM entry point. The network and its kernels are described
// in sno.hpp.
#include "ctx.hpp"
#include "sno.hpp"

namespace {

constexpr int kSnoMaxLayers = 8;

struct SnoLayout {
  long long Wl, bl, Rr[kSnoMaxLayers], Ri[kSnoMaxLayers], Wp[kSnoMaxLayers], bp[kSnoMaxLayers], Wq1, bq1, Wq2, bq2,
      total;
};

unsigned blocks_for(long long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

struct gnlse_sno {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
  long long N = 0, C = 0, K = 0, L = 0, CI = 0, CH = 0, maxB = 0;
  SnoLayout lay{};
  long long adam_t = 0;
  DevBuf w, g, m, v;            // FP32 parameters, gradients, Adam moments
  DevBuf blk, blkT, wp16, wpT16, wq116, wq1T16;  // bf16 operand copies of the weights
  DevBuf Ft, G, Gt, Fn;         // bf16 truncated DFT matrices (Gt = G^T, Fn = Ft^T)
  DevBuf H, Hc, Z, Vm, WsT, Zq, Y1, y;
  DevBuf dy, dH, dZ, dZc, dZ32, dWm, dVt, dblk, part, dZq, dZq32, rel;
};

namespace {

float* W(gnlse_sno* s, long long off) { return (float*)s->w.p + off; }
float* Gd(gnlse_sno* s, long long off) { return (float*)s->g.p + off; }

int sno_refresh_bf16(gnlse_sno* s) {
  const long long per = s->C * s->C;
  for (long long l = 0; l < s->L; ++l) {
    hipLaunchKernelGGL(k_sno_build_blk, dim3(blocks_for(s->K * per)), dim3(256), 0, s->stream, W(s, s->lay.Rr[l]),
                       W(s, s->lay.Ri[l]), (uint16_t*)s->blk.p + l * s->K * 4 * per,
                       (uint16_t*)s->blkT.p + l * s->K * 4 * per, (int)s->K, (int)s->C);
    hipLaunchKernelGGL(k_sno_to_bf16, dim3(blocks_for(per)), dim3(256), 0, s->stream, W(s, s->lay.Wp[l]),
                       (uint16_t*)s->wp16.p + l * per, per);
    hipLaunchKernelGGL(k_sno_to_bf16_t, dim3(blocks_for(per)), dim3(256), 0, s->stream, W(s, s->lay.Wp[l]),
                       (uint16_t*)s->wpT16.p + l * per, (int)s->C, (int)s->C);
  }
  hipLaunchKernelGGL(k_sno_to_bf16, dim3(blocks_for(s->C * s->CH)), dim3(256), 0, s->stream, W(s, s->lay.Wq1),
                     (uint16_t*)s->wq116.p, s->C * s->CH);
  hipLaunchKernelGGL(k_sno_to_bf16_t, dim3(blocks_for(s->C * s->CH)), dim3(256), 0, s->stream, W(s, s->lay.Wq1),
                     (uint16_t*)s->wq1T16.p, (int)s->C, (int)s->CH);
  HIPCHK(hipGetLastError());
  return GNLSE_OK;
}

#define GEMMCHK(expr)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) return fail(GNLSE_ERR_HIP, std::string("gemm: ") + hipGetErrorString(e_)); \
  } while (0)

// sum over points of a[p][i] * b[p][j] into out[i*NB + j]
int sno_outer(gnlse_sno* s, const float* a, const uint16_t* a16, const float* b, long long P, int NA, int NB,
              float* out) {
  const long long nblk = P < 1024 ? P : 1024;
  const long long chunk = (P + nblk - 1) / nblk;
  hipLaunchKernelGGL(k_sno_outer_partial, dim3((unsigned)nblk), dim3(256), 0, s->stream, a, a16, b, (float*)s->part.p,
                     P, NA, NB, chunk);
  hipLaunchKernelGGL(k_sno_sum_parts_wide, dim3((unsigned)(NA * NB)), dim3(256), 0, s->stream,
                     (const float*)s->part.p, out, nblk, (long long)NA * NB);
  HIPCHK(hipGetLastError());
  return GNLSE_OK;
}

// out (M x N) = op(A)^T-style weight gradient over P points, split over K
template <int BM, int BN>
int sno_wgrad(gnlse_sno* s, const uint16_t* A, long long lda, const uint16_t* B, long long ldb, long long M, long long N,
              long long P, float* out) {
  int split = 64;
  while (split > 1 && P % ((long long)split * 32) != 0) split >>= 1;
  GemmShape sh{M, N, P, lda, ldb, 0, 0, split};
  GEMMCHK((gemm_bf16<BM, BN, true, false>(s->stream, A, B, sh, 1, EpiF32{(float*)s->part.p, N, M * N, 0})));
  hipLaunchKernelGGL(k_sno_sum_parts, dim3(blocks_for(M * N)), dim3(256), 0, s->stream, (const float*)s->part.p, out,
                     (long long)split, M * N);
  HIPCHK(hipGetLastError());
  return GNLSE_OK;
}

int sno_forward(gnlse_sno* s, long long B, const float* x, float* y, bool train) {
  const long long N = s->N, C = s->C, K = s->K, P = B * N;
  uint16_t* H = (uint16_t*)s->H.p;
  uint16_t* Hc = (uint16_t*)s->Hc.p;
  hipLaunchKernelGGL((k_sno_lift<64>), dim3((unsigned)(P / 128)), dim3(128), 0, s->stream, x, W(s, s->lay.Wl),
                     W(s, s->lay.bl), H, Hc, N, (int)s->CI);
  HIPCHK(hipGetLastError());
  for (long long l = 0; l < s->L; ++l) {
    const uint16_t* Hl = H + (train ? l : (l & 1)) * P * C;
    uint16_t* Hn = H + (train ? l + 1 : ((l + 1) & 1)) * P * C;
    const uint16_t* Hcl = Hc + (l & 1) * P * C;
    uint16_t* Hcn = Hc + ((l + 1) & 1) * P * C;
    uint16_t* Vml = (uint16_t*)s->Vm.p + (train ? l : 0) * K * B * 2 * C;
    const uint16_t* blkT = (const uint16_t*)s->blkT.p + l * K * 4 * C * C;
    // V = DFT_K(H) of all fields as one GEMM: (2K x N) * (N x B*C), the
    // channel-first activations being a k-contiguous (B*C) x N matrix
    GEMMCHK((gemm_bf16<128, 128, false, true>(s->stream, (const uint16_t*)s->Ft.p, Hcl,
                                              GemmShape{2 * K, B * C, N, N, N, 0, 0, 1}, 1,
                                              EpiToModesAll{Vml, B, C})));
    // Ws = V * [[Rr, Ri], [-Ri, Rr]]: per mode, (B x 2C) * (2C x 2C), stored per field as [C][2K]
    GEMMCHK((gemm_bf16<128, 128, false, true>(s->stream, Vml, blkT,
                                              GemmShape{B, 2 * C, 2 * C, 2 * C, 2 * C, B * 2 * C, 4 * C * C, 1}, K,
                                              EpiToFieldsT{(uint16_t*)s->WsT.p, 2 * K, C})));
    // Z = IDFT_K(Ws) + H Wp + bp (one GEMM over k = 2K + C); H' = gelu(Z)
    const bool last = (l == s->L - 1);
    hipLaunchKernelGGL((k_sno_layer<64>), dim3((unsigned)(N / 128), (unsigned)B), dim3(256), 0, s->stream,
                       (const uint16_t*)s->G.p, (const uint16_t*)s->WsT.p, Hl,
                       (const uint16_t*)s->wpT16.p + l * C * C, (const float*)W(s, s->lay.bp[l]),
                       train ? (float*)s->Z.p + l * P * C : nullptr, Hn, Hcn, N, 2 * K, last ? 1 : 0);
    HIPCHK(hipGetLastError());
  }
  const uint16_t* HL = H + (train ? s->L : (s->L & 1)) * P * C;
  if (!train) {  // hidden layer kept in registers
    hipLaunchKernelGGL((k_sno_head<64, 128>), dim3((unsigned)((P + 127) / 128)), dim3(256), 0, s->stream, HL,
                       (const uint16_t*)s->wq1T16.p, W(s, s->lay.bq1), W(s, s->lay.Wq2), W(s, s->lay.bq2), x, y, P,
                       (int)s->CI);
    HIPCHK(hipGetLastError());
    return GNLSE_OK;
  }
  GEMMCHK((gemm_bf16<128, 128, false, true>(s->stream, HL, (const uint16_t*)s->wq1T16.p,
                                            GemmShape{P, s->CH, C, C, C, 0, 0, 1}, 1,
                                            EpiProj{W(s, s->lay.bq1), train ? (float*)s->Zq.p : nullptr,
                                                    (uint16_t*)s->Y1.p, s->CH})));
  hipLaunchKernelGGL(k_sno_out, dim3(blocks_for(P * 16)), dim3(256), 0, s->stream, (const uint16_t*)s->Y1.p,
                     W(s, s->lay.Wq2), W(s, s->lay.bq2), x, y, P, (int)s->CH, (int)s->CI);
  HIPCHK(hipGetLastError());
  return GNLSE_OK;
}

int sno_backward(gnlse_sno* s, long long B, const float* x) {
  const long long N = s->N, C = s->C, K = s->K, P = B * N, CH = s->CH;
  uint16_t* H = (uint16_t*)s->H.p;
  float* dH = (float*)s->dH.p;
  // output layer
  hipLaunchKernelGGL(k_sno_out_bwd, dim3(blocks_for(P * CH)), dim3(256), 0, s->stream, (const float*)s->dy.p,
                     W(s, s->lay.Wq2), (const float*)s->Zq.p, (uint16_t*)s->dZq.p, (float*)s->dZq32.p, P, (int)CH);
  HIPCHK(hipGetLastError());
  RET_IF(sno_outer(s, nullptr, (const uint16_t*)s->Y1.p, (const float*)s->dy.p, P, (int)CH, 2, Gd(s, s->lay.Wq2)));
  RET_IF(sno_outer(s, nullptr, nullptr, (const float*)s->dy.p, P, 1, 2, Gd(s, s->lay.bq2)));
  RET_IF(sno_outer(s, nullptr, nullptr, (const float*)s->dZq32.p, P, 1, (int)CH, Gd(s, s->lay.bq1)));
  const uint16_t* HL = H + s->L * P * C;
  RET_IF((sno_wgrad<64, 128>(s, HL, C, (const uint16_t*)s->dZq.p, CH, C, CH, P, Gd(s, s->lay.Wq1))));
  // dH = dZq Wq1^T
  GEMMCHK((gemm_bf16<128, 64, false, true>(s->stream, (const uint16_t*)s->dZq.p, (const uint16_t*)s->wq116.p,
                                           GemmShape{P, C, CH, CH, CH, 0, 0, 1}, 1, EpiF32{dH, C, 0, 0})));
  for (long long l = s->L - 1; l >= 0; --l) {
    const uint16_t* Hl = H + l * P * C;
    const uint16_t* Vml = (const uint16_t*)s->Vm.p + l * K * B * 2 * C;
    const uint16_t* blk = (const uint16_t*)s->blk.p + l * K * 4 * C * C;
    uint16_t* dZ = (uint16_t*)s->dZ.p;
    uint16_t* dZc = (uint16_t*)s->dZc.p;
    hipLaunchKernelGGL(k_sno_act_bwd, dim3(blocks_for(P * (C / 8))), dim3(256), 0, s->stream, (const float*)dH,
                       (const float*)s->Z.p + l * P * C, dZ, dZc, (float*)s->dZ32.p, P, N, (int)C,
                       l == s->L - 1 ? 1 : 0);
    HIPCHK(hipGetLastError());
    RET_IF(sno_outer(s, nullptr, nullptr, (const float*)s->dZ32.p, P, 1, (int)C, Gd(s, s->lay.bp[l])));
    RET_IF((sno_wgrad<64, 64>(s, Hl, C, dZ, C, C, C, P, Gd(s, s->lay.Wp[l]))));
    // pointwise path: dH = dZ Wp^T
    GEMMCHK((gemm_bf16<128, 64, false, true>(s->stream, dZ, (const uint16_t*)s->wp16.p + l * C * C,
                                             GemmShape{P, C, C, C, C, 0, 0, 1}, 1, EpiF32{dH, C, 0, 0})));
    // inverse DFT: dWs = G^T dZ per field -> mixing layout (both operands k-contiguous)
    GEMMCHK((gemm_bf16<128, 128, false, true>(s->stream, (const uint16_t*)s->Gt.p, dZc,
                                              GemmShape{2 * K, B * C, N, N, N, 0, 0, 1}, 1,
                                              EpiToModesAll{(uint16_t*)s->dWm.p, B, C})));
    // mixing: dV = dWs Blk^T (stored per field as [C][2K]), dBlk = V^T dWs per mode
    GEMMCHK((gemm_bf16<128, 128, false, true>(s->stream, (const uint16_t*)s->dWm.p, blk,
                                              GemmShape{B, 2 * C, 2 * C, 2 * C, 2 * C, B * 2 * C, 4 * C * C, 1}, K,
                                              EpiToFieldsT{(uint16_t*)s->dVt.p, 2 * K, C})));
    GEMMCHK((gemm_bf16<128, 128, true, false>(s->stream, Vml, (const uint16_t*)s->dWm.p,
                                              GemmShape{2 * C, 2 * C, B, 2 * C, 2 * C, B * 2 * C, B * 2 * C, 1}, K,
                                              EpiF32{(float*)s->dblk.p, 2 * C, 4 * C * C, 0})));
    hipLaunchKernelGGL(k_sno_fold_blk, dim3(blocks_for(K * C * C)), dim3(256), 0, s->stream, (const float*)s->dblk.p,
                       Gd(s, s->lay.Rr[l]), Gd(s, s->lay.Ri[l]), (int)K, (int)C);
    HIPCHK(hipGetLastError());
    // DFT: dH += Ft^T dV per field (Fn = Ft^T, dV as [C][2K])
    GEMMCHK((gemm_bf16<128, 64, false, true>(s->stream, (const uint16_t*)s->Fn.p, (const uint16_t*)s->dVt.p,
                                             GemmShape{N, C, 2 * K, 2 * K, 2 * K, 0, C * 2 * K, 1}, B,
                                             EpiF32{dH, C, N * C, 1})));
  }
  // lift
  RET_IF(sno_outer(s, x, nullptr, dH, P, (int)s->CI, (int)C, Gd(s, s->lay.Wl)));
  RET_IF(sno_outer(s, nullptr, nullptr, dH, P, 1, (int)C, Gd(s, s->lay.bl)));
  return GNLSE_OK;
}

int sno_timed(gnlse_sno* s, bool begin) {
  if (begin) {
    HIPCHK(hipEventRecord(s->ev0, s->stream));
    return GNLSE_OK;
  }
  HIPCHK(hipEventRecord(s->ev1, s->stream));
  HIPCHK(hipEventSynchronize(s->ev1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
  s->last_ms = ms;
  return GNLSE_OK;
}

}  // namespace

extern "C" {

int gnlse_sno_destroy(gnlse_sno* s) {
  if (!s) return GNLSE_OK;
  (void)hipSetDevice(s->device);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  DevBuf* bufs[] = {&s->w,    &s->g,   &s->m,   &s->v,   &s->blk,  &s->blkT, &s->wp16,  &s->wpT16, &s->wq116,
                    &s->wq1T16, &s->Ft, &s->G,   &s->Gt,  &s->Fn,   &s->H,   &s->Hc,   &s->Z,    &s->Vm,    &s->WsT,   &s->Zq,
                    &s->Y1,   &s->y,   &s->dy,  &s->dH,  &s->dZ,   &s->dZc, &s->dZ32, &s->dWm,   &s->dVt,   &s->dblk,
                    &s->part, &s->dZq, &s->dZq32, &s->rel};
  for (DevBuf* b : bufs) release(*b);
  if (s->ev0) (void)hipEventDestroy(s->ev0);
  if (s->ev1) (void)hipEventDestroy(s->ev1);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return GNLSE_OK;
}

int gnlse_sno_create(int device, int64_t n, int width, int modes, int layers, int in_channels, int hidden,
                     int64_t max_batch, gnlse_sno** out) {
  if (!out) return fail(GNLSE_ERR_ARG, "null out");
  *out = nullptr;
  if (n < 128 || n % 128 || width != 64 || modes < 16 || modes % 16 || 2 * modes > n || layers < 1 ||
      layers > kSnoMaxLayers || in_channels < 2 || in_channels > 8 || in_channels * width > 256 || hidden < 64 || hidden % 64 ||
      2 * hidden > 256 || hidden != 128 || max_batch < 1)
    return fail(GNLSE_ERR_ARG, "unsupported surrogate shape");
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
    return fail(GNLSE_ERR_HIP, "no HIP device available");
  HIPCHK(hipSetDevice(device));
  gnlse_sno* s = new gnlse_sno();
  s->device = device;
  s->N = n;
  s->C = width;
  s->K = modes;
  s->L = layers;
  s->CI = in_channels;
  s->CH = hidden;
  s->maxB = max_batch;
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&s->ev0) != hipSuccess ||
      hipEventCreate(&s->ev1) != hipSuccess) {
    gnlse_sno_destroy(s);
    return fail(GNLSE_ERR_HIP, "stream/event creation failed");
  }
  const long long C = width, K = modes, N = n, CH = hidden, B = max_batch, P = B * N, L = layers;
  long long o = 0;
  SnoLayout& y = s->lay;
  y.Wl = o, o += in_channels * C;
  y.bl = o, o += C;
  for (long long l = 0; l < L; ++l) {
    y.Rr[l] = o, o += K * C * C;
    y.Ri[l] = o, o += K * C * C;
    y.Wp[l] = o, o += C * C;
    y.bp[l] = o, o += C;
  }
  y.Wq1 = o, o += C * CH;
  y.bq1 = o, o += CH;
  y.Wq2 = o, o += CH * 2;
  y.bq2 = o, o += 2;
  y.total = o;
  int rc = GNLSE_OK;
  auto al = [&](DevBuf& b, size_t bytes) {
    if (rc == GNLSE_OK) rc = ensure(b, bytes);
  };
  const size_t f4 = sizeof(float), h2 = sizeof(uint16_t);
  al(s->w, o * f4);
  al(s->g, o * f4);
  al(s->m, o * f4);
  al(s->v, o * f4);
  al(s->blk, L * K * 4 * C * C * h2);
  al(s->blkT, L * K * 4 * C * C * h2);
  al(s->wp16, L * C * C * h2);
  al(s->wpT16, L * C * C * h2);
  al(s->wq116, C * CH * h2);
  al(s->wq1T16, C * CH * h2);
  al(s->Ft, 2 * K * N * h2);
  al(s->G, N * 2 * K * h2);
  al(s->H, (L + 1) * P * C * h2);
  al(s->Hc, 2 * P * C * h2);
  al(s->Z, L * P * C * f4);
  al(s->Vm, L * K * B * 2 * C * h2);
  al(s->WsT, B * 2 * K * C * h2);
  al(s->Zq, P * CH * f4);
  al(s->Y1, P * CH * h2);
  al(s->y, P * 2 * f4);
  al(s->dy, P * 2 * f4);
  al(s->dH, P * C * f4);
  al(s->dZ, P * C * h2);
  al(s->dZc, P * C * h2);
  al(s->Gt, 2 * K * N * h2);
  al(s->Fn, N * 2 * K * h2);
  al(s->dZ32, P * C * f4);
  al(s->dWm, K * B * 2 * C * h2);
  al(s->dVt, B * 2 * K * C * h2);
  al(s->dblk, K * 4 * C * C * f4);
  al(s->part, (size_t)64 * C * CH * f4 > (size_t)256 * 512 * f4 ? (size_t)64 * C * CH * f4 : (size_t)256 * 512 * f4);
  al(s->dZq, P * CH * h2);
  al(s->dZq32, P * CH * f4);
  al(s->rel, B * sizeof(double));
  if (rc != GNLSE_OK) {
    gnlse_sno_destroy(s);
    return rc;
  }
  // truncated DFT matrices from exact angles (reduced mod N in integers)
  std::vector<uint16_t> ft(2 * K * N), gm(N * 2 * K), gt(2 * K * N), fn(N * 2 * K);
  const long double two_pi = 6.283185307179586476925286766559005768L;
  for (long long k = 0; k < K; ++k)
    for (long long j = 0; j < N; ++j) {
      const long double a = two_pi * (long double)((k * j) % N) / (long double)N;
      const double c = (double)cosl(a), sn = (double)sinl(a);
      ft[(2 * k) * N + j] = fn[j * 2 * K + 2 * k] = f2bf((float)c);
      ft[(2 * k + 1) * N + j] = fn[j * 2 * K + 2 * k + 1] = f2bf((float)-sn);
      const double ck = (k == 0 ? 1.0 : 2.0) / (double)N;
      gm[j * 2 * K + 2 * k] = gt[(2 * k) * N + j] = f2bf((float)(ck * c));
      gm[j * 2 * K + 2 * k + 1] = gt[(2 * k + 1) * N + j] = f2bf((float)(-ck * sn));
    }
  if (upload_raw(device, s->stream, s->Ft, ft.data(), ft.size() * h2) != GNLSE_OK ||
      upload_raw(device, s->stream, s->G, gm.data(), gm.size() * h2) != GNLSE_OK ||
      upload_raw(device, s->stream, s->Gt, gt.data(), gt.size() * h2) != GNLSE_OK ||
      upload_raw(device, s->stream, s->Fn, fn.data(), fn.size() * h2) != GNLSE_OK ||
      hipMemsetAsync(s->w.p, 0, o * f4, s->stream) != hipSuccess ||
      hipMemsetAsync(s->m.p, 0, o * f4, s->stream) != hipSuccess ||
      hipMemsetAsync(s->v.p, 0, o * f4, s->stream) != hipSuccess || hipStreamSynchronize(s->stream) != hipSuccess) {
    gnlse_sno_destroy(s);
    return fail(GNLSE_ERR_HIP, "surrogate initialisation failed");
  }
  *out = s;
  return GNLSE_OK;
}

int gnlse_sno_param_count(gnlse_sno* s, int64_t* count) {
  if (!s || !count) return fail(GNLSE_ERR_ARG, "null argument");
  *count = s->lay.total;
  return GNLSE_OK;
}

int gnlse_sno_set_params(gnlse_sno* s, const float* params) {
  if (!s || !params) return fail(GNLSE_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(s->device));
  HIPCHK(hipMemcpyAsync(s->w.p, params, s->lay.total * sizeof(float), hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipMemsetAsync(s->m.p, 0, s->lay.total * sizeof(float), s->stream));
  HIPCHK(hipMemsetAsync(s->v.p, 0, s->lay.total * sizeof(float), s->stream));
  s->adam_t = 0;
  RET_IF(sno_refresh_bf16(s));
  HIPCHK(hipStreamSynchronize(s->stream));
  return GNLSE_OK;
}

int gnlse_sno_get_params(gnlse_sno* s, float* params) {
  if (!s || !params) return fail(GNLSE_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(s->device));
  HIPCHK(hipMemcpyAsync(params, s->w.p, s->lay.total * sizeof(float), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return GNLSE_OK;
}

int gnlse_sno_forward(gnlse_sno* s, int64_t batch, const float* d_x, float* d_y) {
  if (!s || !d_x || !d_y || batch < 1 || batch > s->maxB) return fail(GNLSE_ERR_ARG, "invalid surrogate forward arguments");
  HIPCHK(hipSetDevice(s->device));
  RET_IF(sno_timed(s, true));
  RET_IF(sno_forward(s, batch, d_x, d_y, false));
  return sno_timed(s, false);
}

int gnlse_sno_loss_grad(gnlse_sno* s, int64_t batch, const float* d_x, const float* d_t, double* loss, double* rel,
                        float* grad) {
  if (!s || !d_x || !d_t || batch < 1 || batch > s->maxB)
    return fail(GNLSE_ERR_ARG, "invalid surrogate training arguments");
  if (batch % 32) return fail(GNLSE_ERR_ARG, "training batch must be a multiple of 32");
  HIPCHK(hipSetDevice(s->device));
  RET_IF(sno_timed(s, true));
  RET_IF(sno_forward(s, batch, d_x, (float*)s->y.p, true));
  hipLaunchKernelGGL(k_sno_loss, dim3((unsigned)batch), dim3(256), 0, s->stream, (const float*)s->y.p, d_t,
                     (float*)s->dy.p, (double*)s->rel.p, s->N * 2, (long long)batch);
  HIPCHK(hipGetLastError());
  RET_IF(sno_backward(s, batch, d_x));
  RET_IF(sno_timed(s, false));
  std::vector<double> r(batch);
  HIPCHK(hipMemcpy(r.data(), s->rel.p, batch * sizeof(double), hipMemcpyDeviceToHost));
  double acc = 0.0;
  for (double e : r) acc += e * e;
  if (loss) *loss = acc / (double)batch;
  if (rel) std::memcpy(rel, r.data(), batch * sizeof(double));
  if (grad) HIPCHK(hipMemcpy(grad, s->g.p, s->lay.total * sizeof(float), hipMemcpyDeviceToHost));
  return GNLSE_OK;
}

int gnlse_sno_adam(gnlse_sno* s, float lr, float beta1, float beta2, float eps) {
  if (!s) return fail(GNLSE_ERR_ARG, "null surrogate");
  HIPCHK(hipSetDevice(s->device));
  s->adam_t += 1;
  const float c1 = 1.0f - std::pow(beta1, (float)s->adam_t), c2 = 1.0f - std::pow(beta2, (float)s->adam_t);
  hipLaunchKernelGGL(k_sno_adam, dim3(blocks_for(s->lay.total)), dim3(256), 0, s->stream, (float*)s->w.p,
                     (const float*)s->g.p, (float*)s->m.p, (float*)s->v.p, s->lay.total, lr, beta1, beta2, eps, c1, c2);
  HIPCHK(hipGetLastError());
  RET_IF(sno_refresh_bf16(s));
  HIPCHK(hipStreamSynchronize(s->stream));
  return GNLSE_OK;
}

int gnlse_sno_last_ms(gnlse_sno* s, double* ms) {
  if (!s || !ms) return fail(GNLSE_ERR_ARG, "null argument");
  *ms = s->last_ms;
  return GNLSE_OK;
}

int gnlse_sno_encode(gnlse_sno* s, gnlse_ctx* c, int64_t batch, const double* d_A, double scale, const double* d_phase,
                     int channels, int time_domain, float* d_out) {
  if (!s || !d_A || !d_out || batch < 1 || !(scale > 0.0) || channels < 2 || channels > 4)
    return fail(GNLSE_ERR_ARG, "invalid encode arguments");
  RET_IF(check_ready(c, false, false));
  if (c->n != s->N) return fail(GNLSE_ERR_ARG, "context grid differs from the surrogate grid");
  if (c->tab->blocks > 1) return fail(GNLSE_ERR_UNSUPPORTED, "surrogate grids have up to 16384 samples");
  HIPCHK(hipSetDevice(s->device));
  RET_IF(ensure(c->io_b, (size_t)batch * s->N * sizeof(double2)));
  HIPCHK(hipStreamSynchronize(s->stream));
  const unsigned nb = blocks_for(batch * s->N);
  double2* tmp = (double2*)c->io_b.p;
  hipError_t e = hipSuccess;
  if (!time_domain) {
    // pyofss fft, then the natural-order encoding (phase in natural order)
    e = c->tab->fft(c->stream, c->ops, +1, batch, (const double2*)d_A, tmp, nullptr);
    if (e == hipSuccess)
      hipLaunchKernelGGL(k_sno_encode, dim3(nb), dim3(256), 0, c->stream, (const double2*)tmp,
                         (const double2*)d_phase, scale, d_out, (long long)batch, s->N, channels);
  } else if (d_phase) {
    // ifft(phase * fft(A)) (phase in FFT order), then the time encoding
    e = c->tab->fft(c->stream, c->ops, +1, batch, (const double2*)d_A, tmp, nullptr);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_sno_cmul, dim3(nb), dim3(256), 0, c->stream, tmp, (const double2*)d_phase, (long long)batch,
                         s->N);
      e = c->tab->fft(c->stream, c->ops, -1, batch, tmp, tmp, nullptr);
    }
    if (e == hipSuccess)
      hipLaunchKernelGGL(k_sno_pack_time, dim3(nb), dim3(256), 0, c->stream, (const double2*)tmp, scale, d_out,
                         (long long)batch, s->N, channels);
  } else {
    hipLaunchKernelGGL(k_sno_pack_time, dim3(nb), dim3(256), 0, c->stream, (const double2*)d_A, scale, d_out,
                       (long long)batch, s->N, channels);
  }
  if (e != hipSuccess) return fail(GNLSE_ERR_HIP, std::string("fft: ") + hipGetErrorString(e));
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return GNLSE_OK;
}

int gnlse_sno_decode(gnlse_sno* s, gnlse_ctx* c, int64_t batch, const float* d_y, double scale, const double* d_phase,
                     int time_domain, double* d_A) {
  if (!s || !d_y || !d_A || batch < 1 || !(scale > 0.0)) return fail(GNLSE_ERR_ARG, "invalid decode arguments");
  RET_IF(check_ready(c, false, false));
  if (c->n != s->N) return fail(GNLSE_ERR_ARG, "context grid differs from the surrogate grid");
  if (c->tab->blocks > 1) return fail(GNLSE_ERR_UNSUPPORTED, "surrogate grids have up to 16384 samples");
  HIPCHK(hipSetDevice(s->device));
  RET_IF(ensure(c->io_b, (size_t)batch * s->N * sizeof(double2)));
  HIPCHK(hipStreamSynchronize(s->stream));
  const unsigned nb = blocks_for(batch * s->N);
  double2* tmp = (double2*)c->io_b.p;
  hipError_t e = hipSuccess;
  if (!time_domain) {
    hipLaunchKernelGGL(k_sno_decode, dim3(nb), dim3(256), 0, c->stream, d_y, (const double2*)d_phase, scale, tmp,
                       (long long)batch, s->N);
    e = c->tab->fft(c->stream, c->ops, -1, batch, tmp, (double2*)d_A, nullptr);
  } else if (d_phase) {
    hipLaunchKernelGGL(k_sno_unpack_time, dim3(nb), dim3(256), 0, c->stream, d_y, scale, tmp, (long long)batch * s->N);
    e = c->tab->fft(c->stream, c->ops, +1, batch, tmp, tmp, nullptr);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_sno_cmul, dim3(nb), dim3(256), 0, c->stream, tmp, (const double2*)d_phase, (long long)batch,
                         s->N);
      e = c->tab->fft(c->stream, c->ops, -1, batch, tmp, (double2*)d_A, nullptr);
    }
  } else {
    hipLaunchKernelGGL(k_sno_unpack_time, dim3(nb), dim3(256), 0, c->stream, d_y, scale, (double2*)d_A,
                       (long long)batch * s->N);
  }
  if (e != hipSuccess) return fail(GNLSE_ERR_HIP, std::string("fft: ") + hipGetErrorString(e));
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return GNLSE_OK;
}

int gnlse_sno_energy_defect(gnlse_sno* s, int64_t batch, const float* d_x, const float* d_y, double* defect) {
  if (!s || !d_x || !d_y || !defect || batch < 1 || batch > s->maxB) return fail(GNLSE_ERR_ARG, "invalid arguments");
  HIPCHK(hipSetDevice(s->device));
  hipLaunchKernelGGL(k_sno_energy_defect, dim3((unsigned)batch), dim3(256), 0, s->stream, d_x, d_y, (double*)s->rel.p,
                     s->N, (int)s->CI);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(defect, s->rel.p, batch * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return GNLSE_OK;
}

int gnlse_sno_combine(gnlse_sno* s, int64_t batch, const float* d_y1, const float* d_y2, float* d_ym, double* dist) {
  if (!s || !d_y1 || !d_y2 || !d_ym || !dist || batch < 1 || batch > s->maxB)
    return fail(GNLSE_ERR_ARG, "invalid arguments");
  HIPCHK(hipSetDevice(s->device));
  hipLaunchKernelGGL(k_sno_combine, dim3((unsigned)batch), dim3(256), 0, s->stream, d_y1, d_y2, d_ym, (double*)s->rel.p,
                     s->N);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(dist, s->rel.p, batch * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return GNLSE_OK;
}

int gnlse_gemm_bf16(int device, int at, int bt, int64_t M, int64_t N, int64_t K, const void* d_A, int64_t lda,
                    const void* d_B, int64_t ldb, float* d_C, int64_t ldc, int64_t batch, int64_t sA, int64_t sB,
                    int64_t sC, int ksplit) {
  if (!d_A || !d_B || !d_C || M < 1 || N < 1 || K < 32 || batch < 1 || ksplit < 1)
    return fail(GNLSE_ERR_ARG, "invalid gemm arguments");
  HIPCHK(hipSetDevice(device));
  const GemmShape sh{M, N, K, lda, ldb, sA, sB, ksplit};
  const uint16_t* A = (const uint16_t*)d_A;
  const uint16_t* B = (const uint16_t*)d_B;
  const EpiF32 epi{d_C, ldc, sC, 0};
  hipError_t e;
  if (!at && !bt) e = gemm_bf16<128, 128, false, false>(nullptr, A, B, sh, batch, epi);
  else if (at && !bt) e = gemm_bf16<128, 128, true, false>(nullptr, A, B, sh, batch, epi);
  else if (!at && bt) e = gemm_bf16<128, 128, false, true>(nullptr, A, B, sh, batch, epi);
  else return fail(GNLSE_ERR_UNSUPPORTED, "gemm with both operands transposed is not compiled");
  if (e != hipSuccess) return fail(GNLSE_ERR_ARG, std::string("gemm: ") + hipGetErrorString(e));
  HIPCHK(hipDeviceSynchronize());
  return GNLSE_OK;
}

}  // extern "C"
