// Blend kernels of the _part rasteriser variant (submodules/
// diff-surfel-rasterization_part, "dsp/"), the one renderer/
// gaussian_renderer_2d/__init__.py:86-96 calls: pixel centres at +0.5, up to
// 16 semantic channels blended beside the colour, an eighth aux channel with
// the median weight, and the reference's double-precision constants
// (FilterSize, NEAR_PLANE, FAR_PLANE; dsp/cuda_rasterizer/auxiliary.h:19-36).
//
// Preprocess, binning and the per-quadrant sub-lists are the plain pipeline's
// (the preprocess takes the variant's homography when PreprocessArgs.part is
// set).  Each 8 x 8 quadrant is one wave64, one lane per pixel; its sub-list
// is staged in LDS 64 entries at a time and every lane walks the entries in
// order, exactly as the reference's per-pixel loop (dsp forward.cu:334-432,
// backward.cu:258-468).  The backward sums each entry's partial gradients
// over the quadrant's pixels in pixel order and stores them in the
// (instance, quadrant) slots the backward preprocess reduces: no atomics,
// deterministic gradients.
#include "sr_device.h"
#include "sr_internal.h"

namespace sr {

namespace {

constexpr double kFS = 0.7071067811865476;  // FilterSize
constexpr double kNearD = 0.2, kFarD = 100.0;
constexpr int kMaxSem = 16;

// One (pixel, entry) evaluation, dsp forward.cu:352-385 / backward.cu:293-318.
struct PPair {
  float kx, ky, kz, lx, ly, lz, pz, sx, sy, rho3d, rho2d, dx, dy, depth, G, alpha;
};

SR_DEV bool part_pair(float pixx, float pixy, const float4& q0, const float4& q1, const float4& q2, PPair& e) {
  const float Tux = q0.x, Tuy = q0.y, Tuz = q0.z, Tvx = q0.w, Tvy = q1.x, Tvz = q1.y;
  const float Twx = q1.z, Twy = q1.w, Twz = q2.x;
  e.kx = -Tux + pixx * Twx;
  e.ky = -Tuy + pixx * Twy;
  e.kz = -Tuz + pixx * Twz;
  e.lx = -Tvx + pixy * Twx;
  e.ly = -Tvy + pixy * Twy;
  e.lz = -Tvz + pixy * Twz;
  const float px = e.ky * e.lz - e.kz * e.ly;
  const float py = e.kz * e.lx - e.kx * e.lz;
  e.pz = e.kx * e.ly - e.ky * e.lx;
  if (e.pz == 0.0f) return false;
  e.sx = px / e.pz;
  e.sy = py / e.pz;
  e.rho3d = e.sx * e.sx + e.sy * e.sy;
  e.dx = q2.y - pixx;
  e.dy = q2.z - pixy;
  e.rho2d = (float)((1.0 / (kFS * kFS)) * (double)(e.dx * e.dx + e.dy * e.dy));
  const float rho = fmin_(e.rho3d, e.rho2d);
  e.depth = (e.rho3d <= e.rho2d) ? (e.sx * Twx + e.sy * Twy) + Twz : Twz;
  if ((double)e.depth < kNearD) return false;
  const float power = -0.5f * rho;
  if (power > 0.0f) return false;
  e.G = exp_pinned(power);
  e.alpha = fmin_(0.99f, q2.w * e.G);
  return !(e.alpha < 1.0f / 255.0f);
}

// (FAR d - FAR NEAR) / ((FAR - NEAR) d) in double (dsp forward.cu:393).
SR_DEV float mapped_depth(float d) {
  return (float)((kFarD * (double)d - kFarD * kNearD) / ((kFarD - kNearD) * (double)d));
}

struct Quad {
  int tile, q, bx, by;
  uint32_t rbeg, ntile;
  int cnt;
  const uint2* list;
};

SR_DEV Quad quad_of(const PartArgs& a, int qi) {
  const int B = a.vs.B, QX = B / kQuad, Q = QX * QX;
  Quad d;
  d.tile = qi / Q;
  d.q = qi % Q;
  const ViewInfo& vi = a.vs.v[0];
  const int tx = d.tile % vi.gx, ty = d.tile / vi.gx;
  d.bx = tx * B + (d.q % QX) * kQuad;
  d.by = ty * B + (d.q / QX) * kQuad;
  d.rbeg = a.ranges[2 * d.tile];
  d.ntile = a.ranges[2 * d.tile + 1] - d.rbeg;
  d.cnt = (int)a.sub_count[qi];
  d.list = a.sub + (size_t)Q * d.rbeg + (size_t)d.q * d.ntile;
  return d;
}

}  // namespace

// Forward: one wave per quadrant, lane = pixel (lane & 7, lane >> 3).
__global__ __launch_bounds__(64) void k_part_blend_fwd(PartArgs a) {
  __shared__ float4 s_q[5][64];
  __shared__ uint32_t s_pos[64];
  __shared__ float s_sem[kMaxSem][64];
  const int lane = threadIdx.x;
  const Quad d = quad_of(a, blockIdx.x);
  const ViewInfo& vi = a.vs.v[0];
  const uint32_t pxi = (uint32_t)(d.bx + (lane & 7)), pyi = (uint32_t)(d.by + (lane >> 3));
  const bool inside = pxi < (uint32_t)vi.W && pyi < (uint32_t)vi.H;
  const float pixx = (float)((double)(float)pxi + 0.5), pixy = (float)((double)(float)pyi + 0.5);
  const int S = a.S;
  float T = 1.0f;
  uint32_t last_contributor = 0;
  float C[3] = {0.0f, 0.0f, 0.0f}, N[3] = {0.0f, 0.0f, 0.0f}, Sem[kMaxSem];
#pragma unroll
  for (int ch = 0; ch < kMaxSem; ++ch) Sem[ch] = 0.0f;
  float D = 0.0f, dist1 = 0.0f, dist2 = 0.0f, distortion = 0.0f, median_depth = 0.0f, median_weight = 0.0f;
  float median_contributor = -1.0f;
  bool done = !inside;
  for (int c0 = 0; c0 < d.cnt; c0 += 64) {
    if (__ballot(!done) == 0ull) break;
    const int m = min(64, d.cnt - c0);
    __syncthreads();
    if (lane < m) {
      const uint2 e = d.list[c0 + lane];
      const float4* r = a.rec + (size_t)kRecQuads * e.x;
#pragma unroll
      for (int k = 0; k < 5; ++k) s_q[k][lane] = r[k];
      s_pos[lane] = e.y;
      for (int ch = 0; ch < S; ++ch) s_sem[ch][lane] = a.sem[(size_t)e.x * S + ch];
    }
    __syncthreads();
    for (int j = 0; !done && j < m; ++j) {
      PPair e;
      if (!part_pair(pixx, pixy, s_q[0][j], s_q[1][j], s_q[2][j], e)) continue;
      const float alpha = e.alpha, depth = e.depth;
      const float test_T = T * (1.0f - alpha);
      if (test_T < 0.0001f) {
        done = true;
        continue;
      }
      const float A = 1.0f - T;
      const float md = mapped_depth(depth);
      const float error = ((md * md) * A + dist2) - (2.0f * md) * dist1;
      distortion += (error * alpha) * T;
      const uint32_t contributor = s_pos[j] + 1u;  // 1-based position in the tile list
      if (T > 0.5f) {
        median_depth = depth;
        median_weight = alpha * T;
        median_contributor = (float)contributor;
      }
      const float4 q3 = s_q[3][j], q4 = s_q[4][j];
      N[0] += (q3.x * alpha) * T;
      N[1] += (q3.y * alpha) * T;
      N[2] += (q3.z * alpha) * T;
      D += (depth * alpha) * T;
      dist1 += (md * alpha) * T;
      dist2 += ((md * md) * alpha) * T;
      C[0] += (q3.w * alpha) * T;
      C[1] += (q4.x * alpha) * T;
      C[2] += (q4.y * alpha) * T;
      for (int ch = 0; ch < S; ++ch) Sem[ch] += (s_sem[ch][j] * alpha) * T;
      T = test_T;
      last_contributor = contributor;
    }
  }
  if (!inside) return;
  const size_t HW = (size_t)vi.H * vi.W, pid = (size_t)vi.W * pyi + pxi;
  a.final_T[pid] = T;
  a.n_contrib[pid] = last_contributor;
  for (int ch = 0; ch < 3; ++ch) a.out_color[ch * HW + pid] = C[ch] + T * a.bg[ch];
  for (int ch = 0; ch < S; ++ch) a.out_sem[ch * HW + pid] = Sem[ch];
  a.n_contrib[pid + HW] = f2u(median_contributor);
  a.final_T[pid + HW] = dist1;
  a.final_T[pid + 2 * HW] = dist2;
  a.out_others[pid] = D;
  a.out_others[pid + HW] = 1.0f - T;
  a.out_others[pid + 2 * HW] = N[0];
  a.out_others[pid + 3 * HW] = N[1];
  a.out_others[pid + 4 * HW] = N[2];
  a.out_others[pid + 5 * HW] = median_depth;
  a.out_others[pid + 6 * HW] = distortion;
  a.out_others[pid + 7 * HW] = median_weight;
}

hipError_t launch_part_blend_fwd(const PartArgs& a, hipStream_t s) {
  const unsigned n = (unsigned)(a.vs.T * quads_per_tile(a.vs.B));
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_part_blend_fwd, dim3(n), dim3(64), 0, s, a);
  return hipGetLastError();
}

// Backward: one wave per quadrant; the quadrant's sub-list from back to
// front, 64 entries staged at a time, the wave stepping through the entries
// together (each lane walks its own pixel's chain).  Per entry the 18
// geometry partials and the semantic ones are summed over the quadrant's
// pixels in pixel order.
__global__ __launch_bounds__(64) void k_part_blend_bwd(PartArgs a) {
  constexpr int kG = 18 + kMaxSem;
  __shared__ float4 s_q[5][64];
  __shared__ uint32_t s_pos[64];
  __shared__ float s_sem[kMaxSem][64];
  __shared__ float s_tmp[kG][65];
  __shared__ float s_acc[kG][64];
  const int lane = threadIdx.x;
  const Quad d = quad_of(a, blockIdx.x);
  const ViewInfo& vi = a.vs.v[0];
  const uint32_t pxi = (uint32_t)(d.bx + (lane & 7)), pyi = (uint32_t)(d.by + (lane >> 3));
  const bool inside = pxi < (uint32_t)vi.W && pyi < (uint32_t)vi.H;
  const float pixx = (float)((double)(float)pxi + 0.5), pixy = (float)((double)(float)pyi + 0.5);
  const int S = a.S, NG = 18 + S;
  const size_t HW = (size_t)vi.H * vi.W, pid = (size_t)vi.W * pyi + pxi;
  // per-pixel inputs, dsp backward.cu:196-247
  const float T_final = inside ? a.final_T_in[pid] : 0.0f;
  float T = T_final;
  const uint32_t last_contributor = inside ? a.n_contrib_in[pid] : 0u;
  const int median_contributor = inside ? (int)a.n_contrib_in[pid + HW] : 0;
  float dLp[3] = {0.0f, 0.0f, 0.0f}, dLs[kMaxSem], dLn[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int ch = 0; ch < kMaxSem; ++ch) dLs[ch] = 0.0f;
  float dL_ddepth = 0.0f, dL_daccum = 0.0f, dL_dreg = 0.0f, dL_dmedian = 0.0f, dL_dmaxw = 0.0f;
  float final_D = 0.0f, final_D2 = 0.0f;
  if (inside) {
    for (int ch = 0; ch < 3; ++ch) dLp[ch] = a.dL_dpix[ch * HW + pid];
    for (int ch = 0; ch < S; ++ch) dLs[ch] = a.dL_dsem[ch * HW + pid];
    dL_ddepth = a.dL_dothers[pid];
    dL_daccum = a.dL_dothers[HW + pid];
    for (int ch = 0; ch < 3; ++ch) dLn[ch] = a.dL_dothers[(2 + ch) * HW + pid];
    dL_dmedian = a.dL_dothers[5 * HW + pid];
    dL_dreg = a.dL_dothers[6 * HW + pid];
    dL_dmaxw = a.dL_dothers[7 * HW + pid];
    final_D = a.final_T_in[pid + HW];
    final_D2 = a.final_T_in[pid + 2 * HW];
  }
  const float final_A = 1.0f - T_final;
  float bg_dot = 0.0f;
  for (int ch = 0; ch < 3; ++ch) bg_dot += a.bg[ch] * dLp[ch];
  float accum_rec[3] = {0.0f, 0.0f, 0.0f}, last_color[3] = {0.0f, 0.0f, 0.0f};
  float last_depth = 0.0f, last_normal[3] = {0.0f, 0.0f, 0.0f}, accum_depth_rec = 0.0f, accum_alpha_rec = 0.0f;
  float accum_normal_rec[3] = {0.0f, 0.0f, 0.0f}, last_dL_dT = 0.0f, last_alpha = 0.0f;
  const int nch = (d.cnt + 63) / 64;
  const int Q = quads_per_tile(a.vs.B);
  for (int c = nch - 1; c >= 0; --c) {
    const int m = min(64, d.cnt - c * 64);
    __syncthreads();
    if (lane < m) {
      const uint2 e = d.list[c * 64 + lane];
      const float4* r = a.rec + (size_t)kRecQuads * e.x;
#pragma unroll
      for (int k = 0; k < 5; ++k) s_q[k][lane] = r[k];
      s_pos[lane] = e.y;
      for (int ch = 0; ch < S; ++ch) s_sem[ch][lane] = a.sem[(size_t)e.x * S + ch];
    }
    for (int k = 0; k < NG; ++k) s_acc[k][lane] = 0.0f;
    __syncthreads();
    for (int j = m - 1; j >= 0; --j) {
      const uint32_t pos = s_pos[j];  // 0-based position in the tile list
      float g[kG];
#pragma unroll
      for (int k = 0; k < kG; ++k) g[k] = 0.0f;
      bool took = false;
      PPair e;
      if (inside && pos < last_contributor && part_pair(pixx, pixy, s_q[0][j], s_q[1][j], s_q[2][j], e)) {
        took = true;
        const float4 q1 = s_q[1][j], q2 = s_q[2][j], q3 = s_q[3][j], q4 = s_q[4][j];
        const float G = e.G, alpha = e.alpha, c_d = e.depth;
        T = T / (1.0f - alpha);
        const float dchannel_dcolor = alpha * T;
        const float col[3] = {q3.w, q4.x, q4.y};
        float dL_dalpha = 0.0f;
        for (int ch = 0; ch < 3; ++ch) {
          accum_rec[ch] = last_alpha * last_color[ch] + (1.0f - last_alpha) * accum_rec[ch];
          last_color[ch] = col[ch];
          dL_dalpha += (col[ch] - accum_rec[ch]) * dLp[ch];
          g[15 + ch] = dchannel_dcolor * dLp[ch];
        }
        // semantics get their own gradient only (backward.cu:339-348)
        for (int ch = 0; ch < S; ++ch) g[18 + ch] = dchannel_dcolor * dLs[ch];
        float dL_dz = 0.0f, dL_dweight = 0.0f;
        const float m_d = mapped_depth(c_d);
        const float dmd_dd = (float)((kFarD * kNearD) / (((kFarD - kNearD) * (double)c_d) * (double)c_d));
        if ((int)pos == median_contributor - 1) {
          dL_dz += dL_dmedian;
          dL_dweight += dL_dmaxw;
        }
        dL_dweight += ((final_D2 + (m_d * m_d) * final_A) - (2.0f * m_d) * final_D) * dL_dreg;
        dL_dalpha += dL_dweight - last_dL_dT;
        last_dL_dT = dL_dweight * alpha + (1.0f - alpha) * last_dL_dT;
        const float dL_dmd = ((2.0f * (T * alpha)) * (m_d * final_A - final_D)) * dL_dreg;
        dL_dz += dL_dmd * dmd_dd;
        accum_depth_rec = last_alpha * last_depth + (1.0f - last_alpha) * accum_depth_rec;
        last_depth = c_d;
        dL_dalpha += (c_d - accum_depth_rec) * dL_ddepth;
        accum_alpha_rec = (float)((double)last_alpha * 1.0 + (double)((1.0f - last_alpha) * accum_alpha_rec));
        dL_dalpha += (1.0f - accum_alpha_rec) * dL_daccum;
        const float nrm[3] = {q3.x, q3.y, q3.z};
        for (int ch = 0; ch < 3; ++ch) {
          accum_normal_rec[ch] = last_alpha * last_normal[ch] + (1.0f - last_alpha) * accum_normal_rec[ch];
          last_normal[ch] = nrm[ch];
          dL_dalpha += (nrm[ch] - accum_normal_rec[ch]) * dLn[ch];
          g[11 + ch] = (alpha * T) * dLn[ch];
        }
        dL_dalpha *= T;
        last_alpha = alpha;
        dL_dalpha += (-T_final / (1.0f - alpha)) * bg_dot;
        const float dL_dG = q2.w * dL_dalpha;
        dL_dz += (alpha * T) * dL_ddepth;
        if (e.rho3d <= e.rho2d) {
          const float dL_dsx = (dL_dG * -G) * e.sx + dL_dz * q1.z;
          const float dL_dsy = (dL_dG * -G) * e.sy + dL_dz * q1.w;
          const float dsx_pz = dL_dsx / e.pz, dsy_pz = dL_dsy / e.pz;
          const float dpx = dsx_pz, dpy = dsy_pz, dpz = -(dsx_pz * e.sx + dsy_pz * e.sy);
          const float dkx = e.ly * dpz - e.lz * dpy, dky = e.lz * dpx - e.lx * dpz, dkz = e.lx * dpy - e.ly * dpx;
          const float dlx = dpy * e.kz - dpz * e.ky, dly = dpz * e.kx - dpx * e.kz, dlz = dpx * e.ky - dpy * e.kx;
          g[0] = -dkx;
          g[1] = -dky;
          g[2] = -dkz;
          g[3] = -dlx;
          g[4] = -dly;
          g[5] = -dlz;
          g[6] = (pixx * dkx + pixy * dlx) + dL_dz * e.sx;
          g[7] = (pixx * dky + pixy * dly) + dL_dz * e.sy;
          g[8] = (pixx * dkz + pixy * dlz) + dL_dz * 1.0f;
        } else {
          // -G * FilterInvSquare * d: ((-G * 1) / (FS * FS)) * d in double
          const float dG_ddelx = (float)(((double)(-G) / (kFS * kFS)) * (double)e.dx);
          const float dG_ddely = (float)(((double)(-G) / (kFS * kFS)) * (double)e.dy);
          g[9] = dL_dG * dG_ddelx;
          g[10] = dL_dG * dG_ddely;
          g[8] = dL_dz;
        }
        g[14] = G * dL_dalpha;
      }
      const uint64_t tookm = __ballot(took);
      if (tookm == 0ull) continue;
      for (int k = 0; k < NG; ++k) s_tmp[k][lane] = g[k];
      __syncthreads();
      if (lane < NG) {
        float acc = 0.0f;
        for (uint64_t mm = tookm; mm != 0ull; mm &= mm - 1ull) acc += s_tmp[lane][__builtin_ctzll(mm)];
        s_acc[lane][j] = acc;
      }
      __syncthreads();
    }
    __syncthreads();
    if (lane < m) {
      const size_t slot = slot_of(a.pos_slot, a.pos_quads, d.rbeg + s_pos[lane], d.q);
      bool nz = false;
      for (int k = 0; k < NG; ++k) nz |= s_acc[k][lane] != 0.0f;
      a.part_nz[slot] = nz ? 1 : 0;
      if (nz) {
        float* p = a.part + slot * kPartStride;
        for (int k = 0; k < 18; ++k) p[k] = s_acc[k][lane];
        float* ps = a.sem_part + slot * (size_t)S;
        for (int ch = 0; ch < S; ++ch) ps[ch] = s_acc[18 + ch][lane];
      }
    }
  }
}

hipError_t launch_part_blend_bwd(const PartArgs& a, hipStream_t s) {
  const unsigned n = (unsigned)(a.vs.T * quads_per_tile(a.vs.B));
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_part_blend_bwd, dim3(n), dim3(64), 0, s, a);
  return hipGetLastError();
}

}  // namespace sr
