// Superquadric block layer: blocks -> attached surfels, and its adjoint.
//
// Follows BlockGaussianModel.prepare_scaling_rot and its helpers
// (games/block_mesh_splatting/scene/block_gaussian_model.py:103-121,178-256,
// superquadric.py:10-14,93-101, pytorch.py:28-29, utils/general_utils.py:10-87)
// with the block's mesh given as a regular (eta, omega) grid: face (i, j) is
// the triangle (V[i][j], V[i+1][j], V[i][(j+1) % n_omega]) and carries K
// surfels at fixed barycentric weights.
//
// Per surfel the kernel produces exactly what the rasteriser receives:
//   means3D   = sum_k w_k V_k
//   scales    = exp(log(relu(surfel_scale * (s1, s2)) + 1e-8))   (get_scaling)
//   rotations = normalize(rot_to_quat([v1 v2 v0]))                 (get_rotation)
//   opacities = sigmoid(sq_occ[b])                                  (get_opacity)
//
// The backward evaluates the same code on dual numbers, once per block
// parameter (3 log-scales, 4 quaternion, 3 translation, 2 epsilon logits), so
// the 12 directional derivatives of every output are exact; they are
// contracted with the upstream gradients; the per-block sums are formed in
// a fixed order (per-surfel contributions, then one workgroup per block
// and parameter), so the gradients are deterministic.  Surfels whose
// triangle is nearly degenerate (|n| < 1e-3 |e1| |e2|, or |n| or |V1 - c|
// below 1e-4, where the 1e-8 guards of the normalisations matter: small
// faces at the poles) are generated and differentiated in FP64.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/surfel_rast.h"
#include "sr_device.h"
#include "sr_internal.h"

namespace sb {

template <typename T>
struct Dual {
  T v, d;
  Dual() = default;
  SR_DEV Dual(T v_, T d_ = T(0)) : v(v_), d(d_) {}
};

template <typename T> SR_DEV T val(T x) { return x; }
template <typename T> SR_DEV T val(Dual<T> x) { return x.v; }

template <typename T> SR_DEV Dual<T> operator+(Dual<T> a, Dual<T> b) { return {a.v + b.v, a.d + b.d}; }
template <typename T> SR_DEV Dual<T> operator-(Dual<T> a, Dual<T> b) { return {a.v - b.v, a.d - b.d}; }
template <typename T> SR_DEV Dual<T> operator*(Dual<T> a, Dual<T> b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
template <typename T> SR_DEV Dual<T> operator/(Dual<T> a, Dual<T> b) {
  const T q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
template <typename T> SR_DEV Dual<T> operator+(Dual<T> a, T b) { return {a.v + b, a.d}; }
template <typename T> SR_DEV Dual<T> operator+(T a, Dual<T> b) { return {a + b.v, b.d}; }
template <typename T> SR_DEV Dual<T> operator-(Dual<T> a, T b) { return {a.v - b, a.d}; }
template <typename T> SR_DEV Dual<T> operator-(T a, Dual<T> b) { return {a - b.v, -b.d}; }
template <typename T> SR_DEV Dual<T> operator-(Dual<T> a) { return {-a.v, -a.d}; }
template <typename T> SR_DEV Dual<T> operator*(Dual<T> a, T b) { return {a.v * b, a.d * b}; }
template <typename T> SR_DEV Dual<T> operator*(T a, Dual<T> b) { return {a * b.v, a * b.d}; }
template <typename T> SR_DEV Dual<T> operator/(Dual<T> a, T b) { return {a.v / b, a.d / b}; }

SR_DEV float fsqrt(float x) { return sqrtf(x); }
SR_DEV double fsqrt(double x) { return sqrt(x); }
SR_DEV float fexp(float x) { return expf(x); }
SR_DEV double fexp(double x) { return exp(x); }
SR_DEV float flog(float x) { return logf(x); }
SR_DEV double flog(double x) { return log(x); }
template <typename T> SR_DEV Dual<T> fsqrt(Dual<T> a) {
  const T s = fsqrt(a.v);
  return {s, s > T(0) ? a.d / (T(2) * s) : T(0)};
}
template <typename T> SR_DEV Dual<T> fexp(Dual<T> a) {
  const T e = fexp(a.v);
  return {e, e * a.d};
}
template <typename T> SR_DEV Dual<T> flog(Dual<T> a) { return {flog(a.v), a.d / a.v}; }

template <typename S> SR_DEV S sigmoid(S x) { return S(1) / (S(1) + fexp(-x)); }
template <typename T> SR_DEV Dual<T> sigmoid(Dual<T> x) {
  const T s = T(1) / (T(1) + fexp(-x.v));
  return {s, s * (T(1) - s) * x.d};
}

// sign(c) |c|^e for a constant base c (signed_pow, pytorch.py:28-29); the
// derivative in e is sign(c) |c|^e ln|c|, taken as 0 at c == 0 (its limit).
template <typename T> SR_DEV T spow(T c, T e) {
  if (c == T(0)) return T(0);
  const T m = fexp(e * flog(c < T(0) ? -c : c));
  return c < T(0) ? -m : m;
}
template <typename T> SR_DEV Dual<T> spow(T c, Dual<T> e) {
  if (c == T(0)) return {T(0), T(0)};
  const T lc = flog(c < T(0) ? -c : c);
  const T m = fexp(e.v * lc);
  const T s = c < T(0) ? T(-1) : T(1);
  return {s * m, s * m * lc * e.d};
}

// relu(x) + 1e-8 then exp(log(.)) is the identity up to rounding: keep relu + eps.
SR_DEV float relu_eps(float x, float eps) { return x > 0.0f ? x + eps : eps; }
SR_DEV double relu_eps(double x, double eps) { return x > 0.0 ? x + eps : eps; }
template <typename T> SR_DEV Dual<T> relu_eps(Dual<T> x, T eps) {
  return x.v > T(0) ? Dual<T>{x.v + eps, x.d} : Dual<T>{eps, T(0)};
}

template <typename S> struct V3 {
  S x, y, z;
};
template <typename S> SR_DEV V3<S> operator+(V3<S> a, V3<S> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename S> SR_DEV V3<S> operator-(V3<S> a, V3<S> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename S, typename F> SR_DEV V3<S> scl(V3<S> a, F s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename S> SR_DEV S dot(V3<S> a, V3<S> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename S> SR_DEV V3<S> cross(V3<S> a, V3<S> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// Per-block quantities shared by its surfels.
template <typename S> struct Block {
  S e1, e2;       // superquadric exponents
  S Sx, Sy, Sz;   // axis scales exp(sq_s) + scale_min
  S R[9];         // rotation (row-major) of the normalised quaternion
  S tx, ty, tz;   // translation
};

template <typename S, typename T>
SR_DEV Block<S> make_block(const S* p, T scale_min) {
  Block<S> b;
  b.Sx = fexp(p[0]) + scale_min;
  b.Sy = fexp(p[1]) + scale_min;
  b.Sz = fexp(p[2]) + scale_min;
  // torch.nn.functional.normalize: q / max(|q|, 1e-12)
  S n = fsqrt(p[3] * p[3] + p[4] * p[4] + p[5] * p[5] + p[6] * p[6]);
  if (val(n) < T(1e-12)) n = n - n + T(1e-12);
  const S w = p[3] / n, x = p[4] / n, y = p[5] / n, z = p[6] / n;
  // quaternion_to_rotation_matrix, superquadric.py:93-101 (row-major 3 x 3)
  b.R[0] = T(1) - T(2) * y * y - T(2) * z * z;
  b.R[1] = T(2) * x * y - T(2) * z * w;
  b.R[2] = T(2) * x * z + T(2) * y * w;
  b.R[3] = T(2) * x * y + T(2) * z * w;
  b.R[4] = T(1) - T(2) * x * x - T(2) * z * z;
  b.R[5] = T(2) * y * z - T(2) * x * w;
  b.R[6] = T(2) * x * z - T(2) * y * w;
  b.R[7] = T(2) * y * z + T(2) * x * w;
  b.R[8] = T(1) - T(2) * x * x - T(2) * y * y;
  b.tx = p[7];
  b.ty = p[8];
  b.tz = p[9];
  b.e1 = sigmoid(p[10]) * T(1.8) + T(0.1);
  b.e2 = sigmoid(p[11]) * T(1.8) + T(0.1);
  return b;
}

// Vertex at (eta, omega): ((parametric_sq(eta, omega) * ratio) * S) @ R + t.
template <typename S, typename T>
SR_DEV V3<S> vertex(const Block<S>& b, T eta, T om, T ratio) {
  const S ce = spow(cos(eta), b.e1), se = spow(sin(eta), b.e1);
  const S co = spow(cos(om), b.e2), so = spow(sin(om), b.e2);
  const S vx = ce * so * ratio * b.Sx, vy = se * ratio * b.Sy, vz = ce * co * ratio * b.Sz;
  V3<S> v;
  v.x = vx * b.R[0] + vy * b.R[3] + vz * b.R[6] + b.tx;
  v.y = vx * b.R[1] + vy * b.R[4] + vz * b.R[7] + b.ty;
  v.z = vx * b.R[2] + vy * b.R[5] + vz * b.R[8] + b.tz;
  return v;
}

template <typename S> struct SurfelOut {
  S m[3], s[2], q[4];
};

// rot_to_quat_batch (general_utils.py:34-87) of a matrix given by rows.
template <typename S, typename T>
SR_DEV void rot_to_quat(const S m[9], S q[4]) {
  const S m00 = m[0], m01 = m[1], m02 = m[2], m10 = m[3], m11 = m[4], m12 = m[5], m20 = m[6], m21 = m[7], m22 = m[8];
  S qa[4];
  const S a0 = T(1) + m00 + m11 + m22, a1 = T(1) + m00 - m11 - m22;
  const S a2 = T(1) - m00 + m11 - m22, a3 = T(1) - m00 - m11 + m22;
  qa[0] = val(a0) > T(0) ? fsqrt(a0) : a0 - a0;
  qa[1] = val(a1) > T(0) ? fsqrt(a1) : a1 - a1;
  qa[2] = val(a2) > T(0) ? fsqrt(a2) : a2 - a2;
  qa[3] = val(a3) > T(0) ? fsqrt(a3) : a3 - a3;
  int k = 0;  // argmax, first on ties
  for (int i = 1; i < 4; ++i)
    if (val(qa[i]) > val(qa[k])) k = i;
  S c[4];
  if (k == 0) {
    c[0] = qa[0] * qa[0]; c[1] = m21 - m12; c[2] = m02 - m20; c[3] = m10 - m01;
  } else if (k == 1) {
    c[0] = m21 - m12; c[1] = qa[1] * qa[1]; c[2] = m10 + m01; c[3] = m02 + m20;
  } else if (k == 2) {
    c[0] = m02 - m20; c[1] = m10 + m01; c[2] = qa[2] * qa[2]; c[3] = m12 + m21;
  } else {
    c[0] = m10 - m01; c[1] = m20 + m02; c[2] = m21 + m12; c[3] = qa[3] * qa[3];
  }
  S den = qa[k];
  if (val(den) < T(0.1)) den = den - den + T(0.1);
  den = T(2) * den;
  for (int i = 0; i < 4; ++i) q[i] = c[i] / den;
  if (val(q[0]) < T(0))
    for (int i = 0; i < 4; ++i) q[i] = -q[i];
}

// One surfel of face (i, j), barycentric weights w.  Returns the FP32-level
// degeneracy test of the face (|n| < 1e-3 |e1| |e2|).
template <typename S, typename T>
SR_DEV bool surfel(const Block<S>& b, T eta0, T eta1, T om0, T om1, T ratio, T surfel_scale, T w0, T w1, T w2,
                   SurfelOut<S>& o) {
  const T eps = T(1e-8);
  const V3<S> V0 = vertex(b, eta0, om0, ratio), V1 = vertex(b, eta1, om0, ratio), V2 = vertex(b, eta0, om1, ratio);
  o.m[0] = V0.x * w0 + V1.x * w1 + V2.x * w2;
  o.m[1] = V0.y * w0 + V1.y * w1 + V2.y * w2;
  o.m[2] = V0.z * w0 + V1.z * w1 + V2.z * w2;
  const V3<S> e1 = V1 - V0, e2 = V2 - V0;
  const V3<S> n = cross(e1, e2);
  const S nn = fsqrt(dot(n, n));
  const V3<S> v0 = scl(n, S(T(1)) / (nn + eps));
  const V3<S> c = scl(V0 + V1 + V2, T(1) / T(3));
  V3<S> v1 = V1 - c;
  const S v1n = fsqrt(dot(v1, v1)) + eps;
  v1 = scl(v1, S(T(1)) / v1n);
  const V3<S> v2i = V2 - c;
  V3<S> v2 = (v2i - scl(v0, dot(v2i, v0))) - scl(v1, dot(v2i, v1));
  v2 = scl(v2, S(T(1)) / (fsqrt(dot(v2, v2)) + eps));
  const S s1 = v1n / T(2), s2 = dot(v2i, v2) / T(2);
  o.s[0] = relu_eps(s1 * surfel_scale, eps);
  o.s[1] = relu_eps(s2 * surfel_scale, eps);
  // rotation matrix with columns (v1, v2, v0), given by rows
  const S M[9] = {v1.x, v2.x, v0.x, v1.y, v2.y, v0.y, v1.z, v2.z, v0.z};
  S q[4];
  rot_to_quat<S, T>(M, q);
  S qn = fsqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  if (val(qn) < T(1e-12)) qn = qn - qn + T(1e-12);
  for (int k = 0; k < 4; ++k) o.q[k] = q[k] / qn;
  // Degenerate for FP32: nearly collinear edges, or a face so small that the
  // 1e-8 guards of the normalisations become visible (> 1e-4 relative).
  const T le = val(fsqrt(dot(e1, e1))) * val(fsqrt(dot(e2, e2)));
  return val(nn) < T(1e-3) * le || val(nn) < T(1e-4) || val(v1n) < T(1e-4);
}

struct Args {
  int B, n_eta, n_omega, K;
  const float* eta;    // n_eta + 1
  const float* omega;  // n_omega
  const float* bary;   // K x 3 or NULL (centroid, K = 1)
  float ratio, scale_min, surfel_scale;
  const float* prm;    // B x 12: log-scales(3), quaternion(4), translation(3), eps logits(2)
  const float* occ;    // B
};

SR_DEV void face_of(const Args& a, int sidx, int& b, int& i, int& j, int& k) {
  const int F = a.n_eta * a.n_omega;
  k = sidx % a.K;
  const int f = (sidx / a.K) % F;
  b = sidx / (a.K * F);
  i = f / a.n_omega;
  j = f % a.n_omega;
}

SR_DEV void bary_of(const Args& a, int k, float& w0, float& w1, float& w2) {
  if (a.bary) {
    w0 = a.bary[3 * k];
    w1 = a.bary[3 * k + 1];
    w2 = a.bary[3 * k + 2];
  } else {
    w0 = w1 = w2 = 1.0f / 3.0f;
  }
}

template <typename T>
SR_DEV bool gen_plain(const Args& a, int b, int i, int j, int k, SurfelOut<T>& o) {
  T p[12];
  for (int t = 0; t < 12; ++t) p[t] = (T)a.prm[12 * b + t];
  const Block<T> blk = make_block<T, T>(p, (T)a.scale_min);
  float w0, w1, w2;
  bary_of(a, k, w0, w1, w2);
  return surfel<T, T>(blk, (T)a.eta[i], (T)a.eta[i + 1], (T)a.omega[j], (T)a.omega[(j + 1) % a.n_omega],
                      (T)a.ratio, (T)a.surfel_scale, (T)w0, (T)w1, (T)w2, o);
}

__global__ __launch_bounds__(256) void k_block_fwd(Args a, float* means, float* scales, float* rots, float* opac) {
  const int sidx = blockIdx.x * blockDim.x + threadIdx.x;
  const int P = a.B * a.n_eta * a.n_omega * a.K;
  if (sidx >= P) return;
  int b, i, j, k;
  face_of(a, sidx, b, i, j, k);
  SurfelOut<float> o;
  const bool degenerate = gen_plain<float>(a, b, i, j, k, o);
  if (degenerate) {
    SurfelOut<double> od;
    gen_plain<double>(a, b, i, j, k, od);
    for (int t = 0; t < 3; ++t) o.m[t] = (float)od.m[t];
    for (int t = 0; t < 2; ++t) o.s[t] = (float)od.s[t];
    for (int t = 0; t < 4; ++t) o.q[t] = (float)od.q[t];
  }
  for (int t = 0; t < 3; ++t) means[3 * sidx + t] = o.m[t];
  for (int t = 0; t < 2; ++t) scales[2 * sidx + t] = o.s[t];
  for (int t = 0; t < 4; ++t) rots[4 * sidx + t] = o.q[t];
  opac[sidx] = 1.0f / (1.0f + expf(-a.occ[b]));
}

// Directional derivative of the 9 outputs along block parameter pi,
// contracted with the upstream gradient g (9 values).
template <typename T>
SR_DEV T directional(const Args& a, int b, int i, int j, int k, int pi, const float* g) {
  Dual<T> p[12];
  for (int t = 0; t < 12; ++t) p[t] = Dual<T>{(T)a.prm[12 * b + t], t == pi ? T(1) : T(0)};
  const Block<Dual<T>> blk = make_block<Dual<T>, T>(p, (T)a.scale_min);
  float w0, w1, w2;
  bary_of(a, k, w0, w1, w2);
  SurfelOut<Dual<T>> o;
  surfel<Dual<T>, T>(blk, (T)a.eta[i], (T)a.eta[i + 1], (T)a.omega[j], (T)a.omega[(j + 1) % a.n_omega], (T)a.ratio,
                     (T)a.surfel_scale, (T)w0, (T)w1, (T)w2, o);
  T acc = 0;
  for (int t = 0; t < 3; ++t) acc += o.m[t].d * (T)g[t];
  for (int t = 0; t < 2; ++t) acc += o.s[t].d * (T)g[3 + t];
  for (int t = 0; t < 4; ++t) acc += o.q[t].d * (T)g[5 + t];
  return acc;
}

// Pass 1: every surfel's 13 contributions (the 12 directional derivatives
// contracted with its upstream gradients, and the occupancy term) go to
// contrib[pi * P + sidx], parameter-major so each pass stores contiguously.
__global__ __launch_bounds__(256) void k_block_bwd(Args a, const float* dmeans, const float* dscales,
                                                   const float* drots, const float* dopac, float* contrib) {
  const int sidx = blockIdx.x * blockDim.x + threadIdx.x;
  const int P = a.B * a.n_eta * a.n_omega * a.K;
  if (sidx >= P) return;
  int b = 0, i = 0, j = 0, k = 0;
  float g[9];
  face_of(a, sidx, b, i, j, k);
  for (int t = 0; t < 3; ++t) g[t] = dmeans[3 * sidx + t];
  for (int t = 0; t < 2; ++t) g[3 + t] = dscales[2 * sidx + t];
  for (int t = 0; t < 4; ++t) g[5 + t] = drots[4 * sidx + t];
  SurfelOut<float> o;
  const bool degenerate = gen_plain<float>(a, b, i, j, k, o);
  for (int pi = 0; pi < 12; ++pi)
    contrib[(size_t)pi * P + sidx] =
        degenerate ? (float)directional<double>(a, b, i, j, k, pi, g) : directional<float>(a, b, i, j, k, pi, g);
  const float sg = 1.0f / (1.0f + expf(-a.occ[b]));
  contrib[(size_t)12 * P + sidx] = dopac[sidx] * sg * (1.0f - sg);
}

// Pass 2: one workgroup per (block, parameter) sums the block's contiguous
// run of contributions: each lane a strided partial sum in a fixed order,
// then a fixed LDS tree.  No atomics, so the gradients are the same bits on
// every run.
__global__ __launch_bounds__(256) void k_block_reduce(int P, int per_block, const float* contrib, float* dprm,
                                                      float* docc) {
  __shared__ float s_sum[256];
  const int b = blockIdx.x, pi = blockIdx.y;
  const float* src = contrib + (size_t)pi * P + (size_t)b * per_block;
  float acc = 0.0f;
  for (int t = threadIdx.x; t < per_block; t += 256) acc += src[t];
  s_sum[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s_sum[threadIdx.x] += s_sum[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (pi < 12) dprm[12 * b + pi] = s_sum[0];
    else docc[b] = s_sum[0];
  }
}

// The adjoint on a contribution buffer of 13 P floats.
int run_bwd(const Args& a, long long P, const float* dmeans, const float* dscales, const float* drots,
            const float* dopac, float* dprm, float* docc, float* contrib, hipStream_t st) {
  if (P == 0) {
    if (sr::launch_fill_f32(dprm, 0.0f, 12 * (size_t)a.B, st) != hipSuccess) return SR_ERR_HIP;
    if (sr::launch_fill_f32(docc, 0.0f, (size_t)a.B, st) != hipSuccess) return SR_ERR_HIP;
    return SR_OK;
  }
  hipLaunchKernelGGL(k_block_bwd, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, st, a, dmeans, dscales, drots,
                     dopac, contrib);
  if (hipGetLastError() != hipSuccess) return SR_ERR_HIP;
  hipLaunchKernelGGL(k_block_reduce, dim3((unsigned)a.B, 13), dim3(256), 0, st, (int)P,
                     a.n_eta * a.n_omega * a.K, (const float*)contrib, dprm, docc);
  return hipGetLastError() == hipSuccess ? SR_OK : SR_ERR_HIP;
}

// ---------------------------------------------------------------------------
// Superquadric signed distance of points to every block: implicit_sq with
// as_sdf=2 and safe=True (games/block_mesh_splatting/utils/superquadric.py:
// 25-46, safe_pow pytorch.py:32-33) of the points in each block's frame,
// ((p - T) @ R^T) / (S ratio) (block_gaussian_model.py:529-531,613-614):
//   q = clamp(p_block, -5, 5)
//   x = max(q0^2, 1e-6)^(1/e2), y = max(q1^2, 1e-6)^(1/e1), z = max(q2^2, 1e-6)^(1/e2)
//   sdf = max(max(x + z, 1e-6)^(e2/e1) + y, 1e-6)^(e1/2) - 1
// with S, R, T, e1, e2 from the 12 block parameters as for the surfels.  The
// overlap, coverage and occupancy losses (block_gaussian_model.py:522-582)
// reduce this (B x N) field; its adjoint returns d/d(12 block parameters).
// ---------------------------------------------------------------------------
constexpr float kSdfEps = 1e-6f;  // SQRT_EPS, pytorch.py:6
constexpr int kSdfThreads = 256;
constexpr int kSdfSums = 17;  // per block: sum g_k d_j (9), g_k (3), g_k pinv_k (3), dL/de1, dL/de2

struct SdfBlock {
  float R[9], S[3], t[3], e1, e2;
};

SR_DEV SdfBlock sdf_block(const float* prm, int b, float scale_min) {
  float p[12];
  for (int k = 0; k < 12; ++k) p[k] = prm[12 * b + k];
  const Block<float> blk = make_block<float, float>(p, scale_min);
  SdfBlock o;
  for (int k = 0; k < 9; ++k) o.R[k] = blk.R[k];
  o.S[0] = blk.Sx;
  o.S[1] = blk.Sy;
  o.S[2] = blk.Sz;
  o.t[0] = blk.tx;
  o.t[1] = blk.ty;
  o.t[2] = blk.tz;
  o.e1 = blk.e1;
  o.e2 = blk.e2;
  return o;
}

// Point in the block frame: pinv_k = (sum_j d_j R_kj) / (S_k ratio), d = p - t.
SR_DEV void to_block(const SdfBlock& k, float ratio, const float* pt, float* d, float* pinv) {
  for (int j = 0; j < 3; ++j) d[j] = pt[j] - k.t[j];
  for (int r = 0; r < 3; ++r)
    pinv[r] = ((d[0] * k.R[3 * r] + d[1] * k.R[3 * r + 1]) + d[2] * k.R[3 * r + 2]) / (k.S[r] * ratio);
}

// The signed distance, and with grad != nullptr its derivatives: grad[0..2]
// along pinv, grad[3] along e1, grad[4] along e2 (the clamps pass gradient
// where their input is inside the range, as torch.clamp does).
SR_DEV float implicit_sdf(const float* pinv, float e1, float e2, float* grad) {
  float q[3], sq[3], base[3], pw[3];
  const float ie1 = 1.0f / e1, ie2 = 1.0f / e2;
  const float ex[3] = {ie2, ie1, ie2};
  for (int k = 0; k < 3; ++k) {
    q[k] = fminf(fmaxf(pinv[k], -5.0f), 5.0f);
    sq[k] = q[k] * q[k];
    base[k] = fmaxf(sq[k], kSdfEps);
    pw[k] = powf(base[k], ex[k]);
  }
  const float xz = pw[0] + pw[2];
  const float B = fmaxf(xz, kSdfEps);
  const float r21 = e2 / e1;
  const float Bp = powf(B, r21);
  const float res = Bp + pw[1];
  const float A = fmaxf(res, kSdfEps);
  const float h = e1 / 2.0f;
  const float Ap = powf(A, h);
  if (grad) {
    const float d_res = res >= kSdfEps ? h * powf(A, h - 1.0f) : 0.0f;
    const float d_B = xz >= kSdfEps ? d_res * (r21 * powf(B, r21 - 1.0f)) : 0.0f;
    const float d_r21 = d_res * (Bp * logf(B));
    float d_e1 = 0.5f * (Ap * logf(A)) + d_r21 * (-e2 / (e1 * e1));
    float d_e2 = d_r21 * ie1;
    const float d_pw[3] = {d_B, d_res, d_B};
    for (int k = 0; k < 3; ++k) {
      const float d_base = sq[k] >= kSdfEps ? d_pw[k] * (ex[k] * powf(base[k], ex[k] - 1.0f)) : 0.0f;
      const float d_ex = d_pw[k] * (pw[k] * logf(base[k]));
      if (k == 1) d_e1 += d_ex * (-ie1 * ie1);
      else d_e2 += d_ex * (-ie2 * ie2);
      const bool inside = pinv[k] >= -5.0f && pinv[k] <= 5.0f;
      grad[k] = inside ? d_base * (2.0f * q[k]) : 0.0f;
    }
    grad[3] = d_e1;
    grad[4] = d_e2;
  }
  return Ap - 1.0f;
}

__global__ __launch_bounds__(kSdfThreads) void k_sdf_fwd(const float* prm, float ratio, float scale_min,
                                                         const float* pts, int N, float* sdf) {
  const int b = blockIdx.y, n = blockIdx.x * kSdfThreads + threadIdx.x;
  if (n >= N) return;
  const SdfBlock k = sdf_block(prm, b, scale_min);
  float d[3], pinv[3];
  to_block(k, ratio, pts + 3 * (size_t)n, d, pinv);
  sdf[(size_t)b * N + n] = implicit_sdf(pinv, k.e1, k.e2, nullptr);
}

// Per (block, chunk of 256 points): the 17 sums, each over the chunk's
// points in a fixed tree order.
__global__ __launch_bounds__(kSdfThreads) void k_sdf_bwd(const float* prm, float ratio, float scale_min,
                                                         const float* pts, int N, const float* dsdf, float* part) {
  __shared__ float s_w[kSdfSums][kSdfThreads / 64];
  const int b = blockIdx.y, n = blockIdx.x * kSdfThreads + threadIdx.x;
  float c[kSdfSums];
  for (int i = 0; i < kSdfSums; ++i) c[i] = 0.0f;
  if (n < N) {
    const SdfBlock k = sdf_block(prm, b, scale_min);
    float d[3], pinv[3], gr[5];
    to_block(k, ratio, pts + 3 * (size_t)n, d, pinv);
    implicit_sdf(pinv, k.e1, k.e2, gr);
    const float g = dsdf[(size_t)b * N + n];
    for (int r = 0; r < 3; ++r) {
      const float gk = g * gr[r];
      for (int j = 0; j < 3; ++j) c[3 * r + j] = gk * d[j];
      c[9 + r] = gk;
      c[12 + r] = gk * pinv[r];
    }
    c[15] = g * gr[3];
    c[16] = g * gr[4];
  }
  for (int i = 0; i < kSdfSums; ++i) {
    float v = c[i];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) s_w[i][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < kSdfSums) {
    const float v = (s_w[threadIdx.x][0] + s_w[threadIdx.x][1]) + (s_w[threadIdx.x][2] + s_w[threadIdx.x][3]);
    part[((size_t)b * gridDim.x + blockIdx.x) * kSdfSums + threadIdx.x] = v;
  }
}

// Per block: the chunk sums in a fixed order, then the chain to the 12 block
// parameters through S = exp(s) + scale_min, R(normalised quaternion) and
// e = 1.8 sigmoid(.) + 0.1 (dual numbers through make_block).
__global__ __launch_bounds__(64) void k_sdf_reduce(const float* prm, float ratio, float scale_min, int nchunk,
                                                   const float* part, float* dprm) {
  __shared__ float s_tot[kSdfSums];
  const int b = blockIdx.x;
  if (threadIdx.x < kSdfSums) {
    float acc = 0.0f;
    for (int c = 0; c < nchunk; ++c) acc += part[((size_t)b * nchunk + c) * kSdfSums + threadIdx.x];
    s_tot[threadIdx.x] = acc;
  }
  __syncthreads();
  if (threadIdx.x >= 12) return;
  const SdfBlock k = sdf_block(prm, b, scale_min);
  // d/dR_kj, d/dS_k, d/dt_j, d/de1, d/de2 of the loss
  float dR[9], dS[3], dt[3];
  for (int r = 0; r < 3; ++r) {
    const float sr = k.S[r] * ratio;
    for (int j = 0; j < 3; ++j) dR[3 * r + j] = s_tot[3 * r + j] / sr;
    dS[r] = -s_tot[12 + r] / k.S[r];
  }
  for (int j = 0; j < 3; ++j) {
    float acc = 0.0f;
    for (int r = 0; r < 3; ++r) acc += s_tot[9 + r] * k.R[3 * r + j] / (k.S[r] * ratio);
    dt[j] = -acc;
  }
  const int pi = threadIdx.x;
  Dual<float> p[12];
  for (int t = 0; t < 12; ++t) p[t] = Dual<float>{prm[12 * b + t], t == pi ? 1.0f : 0.0f};
  const Block<Dual<float>> blk = make_block<Dual<float>, float>(p, scale_min);
  float acc = 0.0f;
  for (int t = 0; t < 9; ++t) acc += dR[t] * blk.R[t].d;
  acc += dS[0] * blk.Sx.d + dS[1] * blk.Sy.d + dS[2] * blk.Sz.d;
  acc += dt[0] * blk.tx.d + dt[1] * blk.ty.d + dt[2] * blk.tz.d;
  acc += s_tot[15] * blk.e1.d + s_tot[16] * blk.e2.d;
  dprm[12 * b + pi] = acc;
}

// Scratch of sb_generate_bwd_noalloc: the 13 contributions of up to
// kNoAllocSurfels surfels (13.6 MB), one copy per GPU the library runs on,
// held until the process ends.  In host code `scratch` names a host-side
// stand-in, not device memory: kernels get the address hipGetSymbolAddress
// returns for the current device, cached per device.
constexpr long long kNoAllocSurfels = 262144;
__device__ __align__(256) float scratch[13 * kNoAllocSurfels];

std::mutex g_scratch_mu;
float* g_scratch_ptr[64] = {};

// The device address of `scratch` on the current device (looked up once per
// device), or nullptr with the error set.
float* scratch_on_current_device() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
    sr::set_error(SR_ERR_HIP, "sb_generate_bwd_noalloc: no current HIP device");
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      sr::set_error(SR_ERR_HIP, "sb_generate_bwd_noalloc: hipGetSymbolAddress(scratch) failed");
      return nullptr;
    }
    g_scratch_ptr[dev] = (float*)p;
  }
  return g_scratch_ptr[dev];
}

}  // namespace sb

extern "C" {

int sb_generate_fwd(const sb_blocks* blk, float* means3D, float* scales, float* rotations, float* opacities,
                    sr_stream_t stream) {
  if (!blk || !means3D || !scales || !rotations || !opacities) return SR_ERR_INVALID;
  if (blk->B < 0 || blk->n_eta <= 0 || blk->n_omega <= 0 || blk->K <= 0 || !blk->eta || !blk->omega ||
      !blk->params || !blk->occupancy || (!blk->bary && blk->K != 1))
    return SR_ERR_INVALID;
  sb::Args a{blk->B, blk->n_eta, blk->n_omega, blk->K, blk->eta, blk->omega, blk->bary, blk->ratio,
             blk->scale_min, blk->surfel_scale, blk->params, blk->occupancy};
  const long long P = (long long)blk->B * blk->n_eta * blk->n_omega * blk->K;
  if (P == 0) return SR_OK;
  if (P > 0x7fffffffLL) return SR_ERR_INVALID;
  hipLaunchKernelGGL(sb::k_block_fwd, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, (hipStream_t)stream, a,
                     means3D, scales, rotations, opacities);
  return hipGetLastError() == hipSuccess ? SR_OK : SR_ERR_HIP;
}

size_t sb_bwd_workspace_bytes(const sb_blocks* blk) {
  if (!blk || blk->B < 0 || blk->n_eta <= 0 || blk->n_omega <= 0 || blk->K <= 0) return 0;
  const long long P = (long long)blk->B * blk->n_eta * blk->n_omega * blk->K;
  return (size_t)(P > 0 ? P : 1) * 13 * sizeof(float);
}

int sb_generate_bwd(const sb_blocks* blk, const float* dL_dmeans3D, const float* dL_dscales,
                    const float* dL_drotations, const float* dL_dopacities, float* dL_dparams,
                    float* dL_doccupancy, void* workspace, size_t workspace_bytes, sr_stream_t stream) {
  if (!blk || !dL_dmeans3D || !dL_dscales || !dL_drotations || !dL_dopacities || !dL_dparams || !dL_doccupancy)
    return SR_ERR_INVALID;
  if (blk->B < 0 || blk->n_eta <= 0 || blk->n_omega <= 0 || blk->K <= 0 || !blk->eta || !blk->omega ||
      !blk->params || !blk->occupancy || (!blk->bary && blk->K != 1))
    return SR_ERR_INVALID;
  sb::Args a{blk->B, blk->n_eta, blk->n_omega, blk->K, blk->eta, blk->omega, blk->bary, blk->ratio,
             blk->scale_min, blk->surfel_scale, blk->params, blk->occupancy};
  const long long P = (long long)blk->B * blk->n_eta * blk->n_omega * blk->K;
  if (P > 0x7fffffffLL / 13) return SR_ERR_INVALID;
  if (!workspace || workspace_bytes < sb_bwd_workspace_bytes(blk)) return SR_ERR_INVALID;
  return sb::run_bwd(a, P, dL_dmeans3D, dL_dscales, dL_drotations, dL_dopacities, dL_dparams, dL_doccupancy,
                     (float*)workspace, (hipStream_t)stream);
}

int sb_generate_bwd_noalloc(const sb_blocks* blk, const float* dL_dmeans3D, const float* dL_dscales,
                            const float* dL_drotations, const float* dL_dopacities, float* dL_dparams,
                            float* dL_doccupancy, sr_stream_t stream) {
  if (!blk || !dL_dmeans3D || !dL_dscales || !dL_drotations || !dL_dopacities || !dL_dparams || !dL_doccupancy)
    return sr::set_error(SR_ERR_INVALID, "sb_generate_bwd_noalloc: NULL argument");
  if (blk->B < 0 || blk->n_eta <= 0 || blk->n_omega <= 0 || blk->K <= 0 || !blk->eta || !blk->omega ||
      !blk->params || !blk->occupancy || (!blk->bary && blk->K != 1))
    return sr::set_error(SR_ERR_INVALID, "sb_generate_bwd_noalloc: invalid block description");
  const long long P = (long long)blk->B * blk->n_eta * blk->n_omega * blk->K;
  if (P > sb::kNoAllocSurfels)
    return sr::set_error(SR_ERR_INVALID,
                         "sb_generate_bwd_noalloc: more than 262144 surfels (the fixed scratch's limit); "
                         "use sb_generate_bwd with a workspace");
  float* contrib = sb::scratch_on_current_device();
  if (!contrib) return SR_ERR_HIP;
  sb::Args a{blk->B, blk->n_eta, blk->n_omega, blk->K, blk->eta, blk->omega, blk->bary, blk->ratio,
             blk->scale_min, blk->surfel_scale, blk->params, blk->occupancy};
  // every contribution the reduction reads is written first by the same
  // call, so the scratch needs no clearing
  return sb::run_bwd(a, P, dL_dmeans3D, dL_dscales, dL_drotations, dL_dopacities, dL_dparams, dL_doccupancy,
                     contrib, (hipStream_t)stream);
}

int sb_sdf_fwd(const float* params, int B, float ratio, float scale_min, const float* points, int N, float* sdf,
               sr_stream_t stream) {
  if (B < 0 || N < 0) return sr::set_error(SR_ERR_INVALID, "sb_sdf_fwd: B and N must be >= 0");
  if ((long long)B * N == 0) return SR_OK;
  if (!params || !points || !sdf) return sr::set_error(SR_ERR_INVALID, "sb_sdf_fwd: NULL argument");
  if (B > 65535) return sr::set_error(SR_ERR_INVALID, "sb_sdf_fwd: more than 65535 blocks");
  hipLaunchKernelGGL(sb::k_sdf_fwd, dim3((unsigned)((N + sb::kSdfThreads - 1) / sb::kSdfThreads), (unsigned)B),
                     dim3(sb::kSdfThreads), 0, (hipStream_t)stream, params, ratio, scale_min, points, N, sdf);
  return hipGetLastError() == hipSuccess ? SR_OK : sr::set_error(SR_ERR_HIP, "sb_sdf_fwd: launch failed");
}

size_t sb_sdf_bwd_workspace_bytes(int B, int N) {
  if (B < 0 || N < 0) return 0;
  const size_t nchunk = (size_t)(N + sb::kSdfThreads - 1) / sb::kSdfThreads;
  return (size_t)(B > 0 ? B : 1) * (nchunk > 0 ? nchunk : 1) * sb::kSdfSums * sizeof(float);
}

int sb_sdf_bwd(const float* params, int B, float ratio, float scale_min, const float* points, int N,
               const float* dL_dsdf, float* dL_dparams, void* workspace, size_t workspace_bytes, sr_stream_t stream) {
  if (B < 0 || N < 0) return sr::set_error(SR_ERR_INVALID, "sb_sdf_bwd: B and N must be >= 0");
  if (B == 0) return SR_OK;
  if (!params || !dL_dparams || (N > 0 && (!points || !dL_dsdf)))
    return sr::set_error(SR_ERR_INVALID, "sb_sdf_bwd: NULL argument");
  if (B > 65535) return sr::set_error(SR_ERR_INVALID, "sb_sdf_bwd: more than 65535 blocks");
  hipStream_t st = (hipStream_t)stream;
  if (N == 0) {
    return sr::launch_fill_f32(dL_dparams, 0.0f, 12 * (size_t)B, st) == hipSuccess
               ? SR_OK
               : sr::set_error(SR_ERR_HIP, "sb_sdf_bwd: launch failed");
  }
  if (!workspace || workspace_bytes < sb_sdf_bwd_workspace_bytes(B, N))
    return sr::set_error(SR_ERR_INVALID, "sb_sdf_bwd: workspace smaller than sb_sdf_bwd_workspace_bytes");
  const int nchunk = (N + sb::kSdfThreads - 1) / sb::kSdfThreads;
  hipLaunchKernelGGL(sb::k_sdf_bwd, dim3((unsigned)nchunk, (unsigned)B), dim3(sb::kSdfThreads), 0, st, params, ratio,
                     scale_min, points, N, dL_dsdf, (float*)workspace);
  hipLaunchKernelGGL(sb::k_sdf_reduce, dim3((unsigned)B), dim3(64), 0, st, params, ratio, scale_min, nchunk,
                     (const float*)workspace, dL_dparams);
  return hipGetLastError() == hipSuccess ? SR_OK : sr::set_error(SR_ERR_HIP, "sb_sdf_bwd: launch failed");
}

}  // extern "C"
