// Backward kernels of the surfel rasteriser for gfx950 (CDNA4, wave64),
// dsr/cuda_rasterizer/backward.cu:143-630.
//
//   k_blend_bwd        pass 1: per pixel, the back-to-front recurrences over
//                      the pairs it blended, storing each pair's weight,
//                      dL/dalpha and dL/ddepth (a quadrant whose pairs missed
//                      the pair store runs the whole backward here instead)
//   k_blend_bwd_grad   pass 2: per (quadrant, 64-entry chunk), each entry's 18
//                      partial gradients summed over the quadrant's pixels
//   k_preprocess_bwd   per surfel: reduces its records' partial sums over
//                      their tile instances, then the AABB-centre adjoint,
//                      dL/dT -> dL/d(mean, scale, rotation), SH backward and
//                      the densification statistic in dL_dmean2D
//
// The partial sums go to one slot per (tile-list position, quadrant) with
// plain stores and are reduced per record by the backward preprocess: no
// gradient is accumulated with global atomics, and the result does not
// depend on scheduling order.
#include "sr_device.h"
#include "sr_internal.h"
#include "sr_pair.h"

namespace sr {

// Partial-gradient slots: 0..8 dL_dtransMat, 9..10 dL_dmean2D.xy,
// 11..13 dL_dnormal, 14 dL_dopacity, 15..17 dL_dcolors.
constexpr int kNumG = 18;
constexpr int kPre = 8;       // pairs a lane fetches ahead in the per-pair pass
#ifndef SR_CHAIN_PREFETCH
#define SR_CHAIN_PREFETCH 4
#endif
constexpr int kChainPre = SR_CHAIN_PREFETCH;  // stored pairs a lane fetches ahead in the chain pass
// waves per quadrant in the chain pass (64 / kSplitBwd = 8 pixels each,
// 8 lanes per pixel)
constexpr int kSplitBwd = 8;
constexpr int kChainSegMin = 16;  // shortest chain segment a lane replays
#ifndef SR_GRAD_GROUP
#define SR_GRAD_GROUP 4
#endif
// pixels per group in the per-pair gradient sums: an entry's partials are
// summed over each run of kGradGroup blended pixels in pixel order, and
// the runs' sums are added in order
constexpr int kGradGroup = SR_GRAD_GROUP;
static_assert(kGradGroup >= 1 && 64 % kGradGroup == 0, "groups tile the 64 pixels");

// Stores the 18 partial sums (g(k) for slot k) of the sub-list entry at
// tile-list position pos of quadrant q as five float4 in its slot.  Every
// entry of every sub-list sets its flag exactly once per backward and
// stores its sums when any is non-zero, so neither buffer needs clearing.
template <class G>
SR_DEV void store_part(const BlendBwdArgs& a, uint32_t pos, int q, G g) {
  const size_t slot = slot_of(a.pos_slot, a.pos_quads, pos, q);
  bool nz = false;
#pragma unroll
  for (int k = 0; k < 18; ++k) nz |= g(k) != 0.0f;
  a.part_nz[slot] = nz ? 1 : 0;  // most entries blend no pixel of the quadrant: only a flag
  if (!nz) return;
  float4* d = reinterpret_cast<float4*>(a.part + slot * kPartStride);
  d[0] = make_float4(g(0), g(1), g(2), g(3));
  d[1] = make_float4(g(4), g(5), g(6), g(7));
  d[2] = make_float4(g(8), g(9), g(10), g(11));
  d[3] = make_float4(g(12), g(13), g(14), g(15));
  d[4] = make_float4(g(16), g(17), 0.0f, 0.0f);
}

// Per-pixel inputs of the backward blend (backward.cu:168-236).
struct PixelBwd {
  float T_final, final_A, final_D, final_D2, bg_dot;
  uint32_t last_contributor;
  int median_contributor;
  float dLp0, dLp1, dLp2, dLn0, dLn1, dLn2;
  float dL_ddepth, dL_daccum, dL_dreg, dL_dmedian;
};

SR_DEV PixelBwd pixel_bwd(const BlendBwdArgs& a, const ViewInfo& vi, bool inside, uint32_t pxi, uint32_t pyi) {
  PixelBwd p{};
  const size_t HW = (size_t)vi.H * vi.W, pid = (size_t)vi.W * pyi + pxi, pb = vi.pix_base;
  if (inside) {
    const float* final_T = a.final_T + 3 * pb;
    const uint32_t* n_contrib = a.n_contrib + 2 * pb;
    const float* dL_dpix = a.dL_dpix + 3 * pb;
    const float* dL_dothers = a.dL_dothers + 7 * pb;
    p.T_final = final_T[pid];
    p.final_D = final_T[pid + HW];
    p.final_D2 = final_T[pid + 2 * HW];
    p.last_contributor = n_contrib[pid];
    p.median_contributor = (int)n_contrib[pid + HW];
    p.dLp0 = dL_dpix[pid];
    p.dLp1 = dL_dpix[HW + pid];
    p.dLp2 = dL_dpix[2 * HW + pid];
    p.dL_ddepth = dL_dothers[pid];
    p.dL_daccum = dL_dothers[HW + pid];
    p.dLn0 = dL_dothers[2 * HW + pid];
    p.dLn1 = dL_dothers[3 * HW + pid];
    p.dLn2 = dL_dothers[4 * HW + pid];
    p.dL_dmedian = dL_dothers[5 * HW + pid];
    p.dL_dreg = dL_dothers[6 * HW + pid];
  }
  p.final_A = 1.0f - p.T_final;
  float bg_dot = 0.0f;
  bg_dot += a.bg[0] * p.dLp0;
  bg_dot += a.bg[1] * p.dLp1;
  bg_dot += a.bg[2] * p.dLp2;
  p.bg_dot = bg_dot;
  return p;
}

// The back-to-front recurrences of one pixel (backward.cu:316-379): fed the
// pairs it blended in reverse order, each step yields the pair's weight
// w = alpha T, dL/dalpha and dL/ddepth, from which every per-pair gradient
// follows without further pixel state.
struct ChainBwd {
  float T;
  float acc_r0 = 0.0f, acc_r1 = 0.0f, acc_r2 = 0.0f;
  float last_c0 = 0.0f, last_c1 = 0.0f, last_c2 = 0.0f;
  float last_depth = 0.0f, accum_depth_rec = 0.0f, accum_alpha_rec = 0.0f;
  float last_n0 = 0.0f, last_n1 = 0.0f, last_n2 = 0.0f;
  float acc_n0 = 0.0f, acc_n1 = 0.0f, acc_n2 = 0.0f;
  float last_dL_dT = 0.0f, last_alpha = 0.0f;
  // nc = (normal.xyz, r); m_d, m2 = m_d * m_d, dmd_dd = distortion weight,
  // its square and its derivative
  SR_DEV float4 step(const PixelBwd& p, float alpha, float c_d, float m_d, float m2, float dmd_dd,
                     uint32_t contributor, const float4& nc, float g, float bl) {
    T = T / (1.0f - alpha);
    const float w = alpha * T;
    float dL_dalpha = 0.0f;
    acc_r0 = last_alpha * last_c0 + (1.0f - last_alpha) * acc_r0;
    last_c0 = nc.w;
    dL_dalpha += (nc.w - acc_r0) * p.dLp0;
    acc_r1 = last_alpha * last_c1 + (1.0f - last_alpha) * acc_r1;
    last_c1 = g;
    dL_dalpha += (g - acc_r1) * p.dLp1;
    acc_r2 = last_alpha * last_c2 + (1.0f - last_alpha) * acc_r2;
    last_c2 = bl;
    dL_dalpha += (bl - acc_r2) * p.dLp2;
    float dL_dz = 0.0f, dL_dweight = 0.0f;
    if ((int)contributor == p.median_contributor - 1) dL_dz += p.dL_dmedian;
    dL_dweight += ((p.final_D2 + m2 * p.final_A) - (2.0f * m_d) * p.final_D) * p.dL_dreg;
    dL_dalpha += dL_dweight - last_dL_dT;
    last_dL_dT = dL_dweight * alpha + (1.0f - alpha) * last_dL_dT;
    const float dL_dmd = ((2.0f * (T * alpha)) * (m_d * p.final_A - p.final_D)) * p.dL_dreg;
    dL_dz += dL_dmd * dmd_dd;
    accum_depth_rec = last_alpha * last_depth + (1.0f - last_alpha) * accum_depth_rec;
    last_depth = c_d;
    dL_dalpha += (c_d - accum_depth_rec) * p.dL_ddepth;
    accum_alpha_rec = last_alpha + (1.0f - last_alpha) * accum_alpha_rec;
    dL_dalpha += (1.0f - accum_alpha_rec) * p.dL_daccum;
    acc_n0 = last_alpha * last_n0 + (1.0f - last_alpha) * acc_n0;
    last_n0 = nc.x;
    dL_dalpha += (nc.x - acc_n0) * p.dLn0;
    acc_n1 = last_alpha * last_n1 + (1.0f - last_alpha) * acc_n1;
    last_n1 = nc.y;
    dL_dalpha += (nc.y - acc_n1) * p.dLn1;
    acc_n2 = last_alpha * last_n2 + (1.0f - last_alpha) * acc_n2;
    last_n2 = nc.z;
    dL_dalpha += (nc.z - acc_n2) * p.dLn2;
    dL_dalpha *= T;
    last_alpha = alpha;
    dL_dalpha += (-p.T_final / (1.0f - alpha)) * p.bg_dot;
    dL_dz += (alpha * T) * p.dL_ddepth;
    return make_float4(w, dL_dalpha, dL_dz, 0.0f);
  }
  // The state updates of step() alone, with the same operations in the same
  // order (so a chain resumed from a saved state continues bit for bit).
  SR_DEV void advance(const PixelBwd& p, float alpha, float c_d, float m_d, float m2, const float4& nc, float g,
                      float bl) {
    T = T / (1.0f - alpha);
    acc_r0 = last_alpha * last_c0 + (1.0f - last_alpha) * acc_r0;
    last_c0 = nc.w;
    acc_r1 = last_alpha * last_c1 + (1.0f - last_alpha) * acc_r1;
    last_c1 = g;
    acc_r2 = last_alpha * last_c2 + (1.0f - last_alpha) * acc_r2;
    last_c2 = bl;
    float dL_dweight = 0.0f;
    dL_dweight += ((p.final_D2 + m2 * p.final_A) - (2.0f * m_d) * p.final_D) * p.dL_dreg;
    last_dL_dT = dL_dweight * alpha + (1.0f - alpha) * last_dL_dT;
    accum_depth_rec = last_alpha * last_depth + (1.0f - last_alpha) * accum_depth_rec;
    last_depth = c_d;
    accum_alpha_rec = last_alpha + (1.0f - last_alpha) * accum_alpha_rec;
    acc_n0 = last_alpha * last_n0 + (1.0f - last_alpha) * acc_n0;
    last_n0 = nc.x;
    acc_n1 = last_alpha * last_n1 + (1.0f - last_alpha) * acc_n1;
    last_n1 = nc.y;
    acc_n2 = last_alpha * last_n2 + (1.0f - last_alpha) * acc_n2;
    last_n2 = nc.z;
    last_alpha = alpha;
  }
};
static_assert(sizeof(ChainBwd) == 18 * sizeof(float), "chain state is 18 floats");

// ChainBwd's recurrences over one segment of a run, as an affine map of the
// state entering the segment.  The seven colour, depth and normal
// accumulators and accum_alpha_rec follow x <- la l + (1 - la) x (la, l the
// alpha and value of the pair behind), last_dL_dT follows
// y <- dw alpha + (1 - alpha) y, and T <- T / (1 - alpha).  Accumulated from
// the identity over the segment's pairs (back to front), the map takes the
// entering state to the leaving one: x_out = A x_in + b, y_out = A2 y_in + b2,
// T_out = T_in / D.  Folding the maps of the later segments onto the chain's
// initial state gives the state where a segment starts without walking the
// pairs behind it; every path folds in the same order with the same
// operations, so a segment's replay starts from the same bits wherever its
// pairs come from.
struct SegMap {
  float A = 1.0f, A2 = 1.0f, b2 = 0.0f, D = 1.0f;
  float b[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};  // r, g, b, depth, n0, n1, n2, alpha
  // one pair (alpha, distortion weight m_d) after the pair behind it (la,
  // values l: r, g, b, depth, n0, n1, n2)
  SR_DEV void add(const PixelBwd& p, float la, const float* l, float alpha, float m_d, float m2) {
    const float om = 1.0f - la;
#pragma unroll
    for (int x = 0; x < 7; ++x) b[x] = la * l[x] + om * b[x];
    b[7] = la + om * b[7];
    A = om * A;
    const float dw = ((p.final_D2 + m2 * p.final_A) - (2.0f * m_d) * p.final_D) * p.dL_dreg;
    const float oa = 1.0f - alpha;
    b2 = dw * alpha + oa * b2;
    A2 = oa * A2;
    D = D * oa;
  }
  // the map applied to the recurrence part of a chain state
  SR_DEV void apply(ChainBwd& s) const {
    s.T = s.T / D;
    s.acc_r0 = A * s.acc_r0 + b[0];
    s.acc_r1 = A * s.acc_r1 + b[1];
    s.acc_r2 = A * s.acc_r2 + b[2];
    s.accum_depth_rec = A * s.accum_depth_rec + b[3];
    s.acc_n0 = A * s.acc_n0 + b[4];
    s.acc_n1 = A * s.acc_n1 + b[5];
    s.acc_n2 = A * s.acc_n2 + b[6];
    s.accum_alpha_rec = A * s.accum_alpha_rec + b[7];
    s.last_dL_dT = A2 * s.last_dL_dT + b2;
  }
  SR_DEV void save(float* d) const {
    d[0] = A;
    d[1] = A2;
    d[2] = b2;
    d[3] = D;
#pragma unroll
    for (int x = 0; x < 8; ++x) d[4 + x] = b[x];
  }
  SR_DEV void load(const float* d) {
    A = d[0];
    A2 = d[1];
    b2 = d[2];
    D = d[3];
#pragma unroll
    for (int x = 0; x < 8; ++x) b[x] = d[4 + x];
  }
};
constexpr int kSegMapFloats = 12;

// The values of a chain state that a pair sets for the next one (the "last"
// members), in SegMap's order.
SR_DEV void chain_last(const ChainBwd& s, float* l) {
  l[0] = s.last_c0;
  l[1] = s.last_c1;
  l[2] = s.last_c2;
  l[3] = s.last_depth;
  l[4] = s.last_n0;
  l[5] = s.last_n1;
  l[6] = s.last_n2;
}

// The 18 partial gradients of one pair (backward.cu:380-437) from its
// weight w, dL/dalpha and dL/ddepth and its pixel's dL/dpixel (dLp) and
// dL/dnormal (dLn), handed to add(slot, value).
template <class Add>
SR_DEV void pair_grads(float dLp0, float dLp1, float dLp2, float dLn0, float dLn1, float dLn2, const Pair& e,
                       const float4& q1, float opa, float w, float dL_dalpha, float dL_dz, float pixx, float pixy,
                       Add add) {
  add(15, w * dLp0);
  add(16, w * dLp1);
  add(17, w * dLp2);
  add(11, w * dLn0);
  add(12, w * dLn1);
  add(13, w * dLn2);
  const float G = e.G;
  const float dL_dG = opa * dL_dalpha;
  if (e.rho3d <= e.rho2d) {
    const float sx = e.sx, sy = e.sy, pz = e.pz;
    const float kx = e.kx, ky = e.ky, kz = e.kz, lx = e.lx, ly = e.ly, lz = e.lz;
    const float dL_dsx = (dL_dG * -G) * sx + dL_dz * q1.z;
    const float dL_dsy = (dL_dG * -G) * sy + dL_dz * q1.w;
    const float dsx_pz = dL_dsx / pz, dsy_pz = dL_dsy / pz;
    const float dpx = dsx_pz, dpy = dsy_pz, dpz = -(dsx_pz * sx + dsy_pz * sy);
    // dL_dk = cross(l, dL_dp), dL_dl = cross(dL_dp, k)
    const float dkx = ly * dpz - lz * dpy, dky = lz * dpx - lx * dpz, dkz = lx * dpy - ly * dpx;
    const float dlx = dpy * kz - dpz * ky, dly = dpz * kx - dpx * kz, dlz = dpx * ky - dpy * kx;
    add(0, -dkx);
    add(1, -dky);
    add(2, -dkz);
    add(3, -dlx);
    add(4, -dly);
    add(5, -dlz);
    add(6, (pixx * dkx + pixy * dlx) + dL_dz * sx);
    add(7, (pixx * dky + pixy * dly) + dL_dz * sy);
    add(8, (pixx * dkz + pixy * dlz) + dL_dz * 1.0f);
  } else {
    const float dG_ddelx = (-G * 2.0f) * e.dx;
    const float dG_ddely = (-G * 2.0f) * e.dy;
    add(9, dL_dG * dG_ddelx);
    add(10, dL_dG * dG_ddely);
    add(8, dL_dz);
  }
  add(14, G * dL_dalpha);
}

// Backward blend, pass 1 of 2.  One wave64 per 8 x 8 quadrant; each lane
// walks, back to front, the pairs its pixel blended (the accepted pairs in
// front of its last contributor, exactly the reference's), running only the
// per-pixel recurrences and writing each pair's (w, dL/dalpha, dL/ddepth)
// to the pair store.  The pairs come from the store as contiguous per-lane
// streams; nothing is staged and there are no barriers, so the serial
// per-pixel chain carries only the recurrences.  A quadrant some of whose
// pairs missed the store (q_flag) instead runs the whole backward here over
// staged records, recomputing those pairs.
__global__ __launch_bounds__(64) void k_blend_bwd(BlendBwdArgs a) {
  __shared__ float4 s_q0[64], s_q1[64], s_q2[64], s_q3[64], s_q4[64];
  __shared__ uint2 s_e[64];
  __shared__ float s_acc[kNumG][64];
  __shared__ float s_tmp[kNumG][65];
  __shared__ float s_map[8][8][kSegMapFloats];  // per pixel of the wave, each segment's SegMap
  const int B = a.vs.B, QX = B / kQuad, Q = QX * QX;
  if (a.counters[1] & kErrCapacity) return;
  const int qi = (int)a.q_order[blockIdx.x / kSplitBwd], part = blockIdx.x % kSplitBwd;
  const int tile = qi / Q, q = qi % Q;
  const ViewInfo& vi = a.vs.v[view_of_tile(a.vs, (uint32_t)tile)];
  const int lt = tile - (int)vi.tile_base;
  const int tx = lt % vi.gx, ty = lt / vi.gx;
  const int lane = threadIdx.x;
  const int cnt = (int)a.sub_count[qi];
  if (cnt == 0) return;  // no surfel reaches the quadrant: no gradient from it
  const bool staged = a.q_flag[qi] != 0u;
  if (staged && part != 0) return;  // the staged path uses all 64 lanes of part 0
  // pixel of this lane: all 64 in the staged path, 64 / kSplitBwd per part otherwise
  const int pix = staged ? lane : part * (64 / kSplitBwd) + lane;
  const bool lane_ok = staged || lane < 64 / kSplitBwd;
  const uint32_t pxi = (uint32_t)(tx * B + (q % QX) * kQuad + (pix & 7));
  const uint32_t pyi = (uint32_t)(ty * B + (q / QX) * kQuad + (pix >> 3));
  const bool inside = lane_ok && pxi < (uint32_t)vi.W && pyi < (uint32_t)vi.H;
  const float pixx = (float)pxi, pixy = (float)pyi;
  ChainBwd ch;

  if (!staged) {
    // Each pixel's blended pairs are one contiguous run in entry order (the
    // forward cut the run at the pixel's last contributor), walked from the
    // back with the next kChainPre pairs in flight.  A wave holds 8 pixels
    // and gives each 8 lanes: a long run is cut into up to 8 segments, one
    // per lane.  Each segment's lane first runs its pairs through SegMap (the
    // state recurrences as an affine map, about a third of a step); the maps
    // of the segments behind a segment, folded onto the chain's initial
    // state, give the state where it starts; then every lane replays its
    // segment with full steps from that state.  The serial chain of a run of
    // n pairs is two passes over n / 8 pairs instead of one over n.
    const int ps = lane & 7, sg = lane >> 3;
    const int pxl = part * 8 + ps;
    const uint32_t qxi = (uint32_t)(tx * B + (q % QX) * kQuad + (pxl & 7));
    const uint32_t qyi = (uint32_t)(ty * B + (q / QX) * kQuad + (pxl >> 3));
    const bool pin = qxi < (uint32_t)vi.W && qyi < (uint32_t)vi.H;
    const uint2 rg = pin ? a.lane_range[(size_t)qi * 64 + pxl] : make_uint2(0u, 0u);
    const int n = (int)rg.y, f0 = (int)rg.x;
    const int seg_len = max(kChainSegMin, (n + 7) / 8);
    const int nseg = (n + seg_len - 1) / seg_len;
    const PixelBwd px = pixel_bwd(a, vi, n > 0, qxi, qyi);
    // walk pairs f0 + hi down to f0 + lo, calling fn(index, p0, p1, p2)
    auto walk = [&](int lo, int hi, auto&& fn) {
      float4 g0[kChainPre], g1[kChainPre], g2[kChainPre];
      auto fetch = [&](int at) {
#pragma unroll
        for (int k = 0; k < kChainPre; ++k) {
          if (at - k >= lo) {
            const float4* pr = a.ps.rec + kPairQuads * (size_t)(f0 + at - k);
            g0[k] = pr[0];
            g1[k] = pr[1];
            g2[k] = pr[2];
          }
        }
      };
      int f = hi;
      if (f >= lo) fetch(f);
      while (f >= lo) {
        float4 c0[kChainPre], c1[kChainPre], c2[kChainPre];
#pragma unroll
        for (int k = 0; k < kChainPre; ++k) {
          c0[k] = g0[k];
          c1[k] = g1[k];
          c2[k] = g2[k];
        }
        const int at = f;
        f -= kChainPre;
        if (f >= lo) fetch(f);  // next batch in flight
#pragma unroll
        for (int k = 0; k < kChainPre; ++k)
          if (at - k >= lo) fn(at - k, c0[k], c1[k], c2[k]);
      }
    };
    const int lo = sg * seg_len, hi = min(n, lo + seg_len) - 1;
    // alpha and values of the pair behind the segment (none behind the last)
    float la = 0.0f, l[7] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if (sg < nseg - 1) {
      const float4* pr = a.ps.rec + kPairQuads * (size_t)(f0 + hi + 1);
      const float4 e0 = pr[0], e1 = pr[1], e2 = pr[2];
      la = e0.w;
      l[0] = e1.w;
      l[1] = e2.x;
      l[2] = e2.y;
      l[3] = e0.x;
      l[4] = e1.x;
      l[5] = e1.y;
      l[6] = e1.z;
    }
    float* maps = &s_map[ps][0][0];
    if (sg >= 1 && sg < nseg) {  // segment 0's map is never needed
      SegMap m;
      float pa = la, pl[7];
#pragma unroll
      for (int x = 0; x < 7; ++x) pl[x] = l[x];
      walk(lo, hi, [&](int, const float4& c0, const float4& c1, const float4& c2) {
        m.add(px, pa, pl, c0.w, c0.y, c0.z);
        pa = c0.w;
        pl[0] = c1.w;
        pl[1] = c2.x;
        pl[2] = c2.y;
        pl[3] = c0.x;
        pl[4] = c1.x;
        pl[5] = c1.y;
        pl[6] = c1.z;
      });
      m.save(maps + sg * kSegMapFloats);
    }
    __syncthreads();
    if (sg < nseg) {
      ch.T = px.T_final;
      for (int t = nseg - 1; t > sg; --t) {
        SegMap m;
        m.load(maps + t * kSegMapFloats);
        m.apply(ch);
      }
      ch.last_alpha = la;
      ch.last_c0 = l[0];
      ch.last_c1 = l[1];
      ch.last_c2 = l[2];
      ch.last_depth = l[3];
      ch.last_n0 = l[4];
      ch.last_n1 = l[5];
      ch.last_n2 = l[6];
      walk(lo, hi, [&](int k, const float4& c0, const float4& c1, const float4& c2) {
        a.ps.fW[f0 + k] = ch.step(px, c0.w, c0.x, c0.y, c0.z, c2.w, __float_as_uint(c2.z), c1, c2.x, c2.y);
      });
    }
    return;
  }

  // Staged path: per chunk, stage the records, run the recurrences and the
  // gradients together, sum per entry in LDS and store the sums.
  const uint32_t rbeg = a.ranges[2 * tile], ntile = a.ranges[2 * tile + 1] - rbeg;
  const uint2* list = a.sub + (size_t)Q * rbeg + (size_t)q * ntile;
  const uint64_t* masks = a.masks + mask_base(Q, rbeg, ntile, tile, q);
  const uint32_t* frag_off = a.frag_off + mask_base(Q, rbeg, ntile, tile, q);
  const PixelBwd px = pixel_bwd(a, vi, inside, pxi, pyi);
  ch.T = px.T_final;
  // The pixel's run is cut into the stored path's segments (the forward
  // recorded how many pairs the pixel blended): while a segment's pairs go
  // through full steps, its SegMap is accumulated alongside, and where the
  // segment ends the state is reset to the maps folded onto the initial
  // state, exactly as the stored path's segment lanes start, so each pair's
  // outputs have the stored path's bits.
  ChainBwd fold = ch;  // the state the current segment started from
  SegMap seg_map;
  const int n_run = inside ? (int)a.lane_range[(size_t)qi * 64 + pix].y : 0;
  const int seg_len = max(kChainSegMin, (n_run + 7) / 8);
  int seg = (n_run + seg_len - 1) / seg_len - 1, seg_start = seg * seg_len, idx = n_run;
  const int nch = (cnt + 63) / 64;
  const bool active = inside && px.last_contributor > 0;
  const float mcoef = kFar / (kFar - kNear);
  const float dmd_num = kFar * kNear, dmd_den0 = kFar - kNear;
  ChunkRegs nx{};
  uint2 v2 = make_uint2(0u, 0u);
  if (nch > 0) {
    nx = chunk_load(masks, frag_off, a.rec, lane, nch - 1, chunk_entry(list, cnt, lane, nch - 1, nch));
    v2 = chunk_entry(list, cnt, lane, nch - 2, nch);
  }
  for (int c = nch - 1; c >= 0; --c) {
    uint64_t mask = active ? nx.mask : 0ull;
    const int m = min(64, cnt - c * 64);
    __syncthreads();
    if (lane < m) {
      s_e[lane] = nx.v;
      s_q0[lane] = nx.q0;
      s_q1[lane] = nx.q1;
      s_q2[lane] = nx.q2;
      s_q3[lane] = nx.q3;
      s_q4[lane] = nx.q4;
    }
#pragma unroll
    for (int k = 0; k < kNumG; ++k) s_acc[k][lane] = 0.0f;
    __syncthreads();
    if (c > 0) {  // in flight while this chunk runs
      const uint2 ve = v2;
      v2 = chunk_entry(list, cnt, lane, c - 2, nch);
      nx = chunk_load(masks, frag_off, a.rec, lane, c - 1, ve);
    }
    if (__ballot(mask != 0ull) != 0ull) {
      // entries from last to first, the wave in step (each lane still walks
      // its own pixel's entries back to front); an entry's 18 partial
      // gradients are then summed over the pixels that blended it in pixel
      // order from 0, in runs of kGradGroup pixels whose sums are added in
      // order, exactly as the per-pair gradient pass sums a stored
      // quadrant's, so the partials have the same bits whichever quadrants
      // missed the pair store
      for (int j = 63; j >= 0; --j) {
        const bool hit = (mask >> j) & 1ull;
        if (__ballot(hit) == 0ull) continue;
        float gk[kNumG];
#pragma unroll
        for (int k = 0; k < kNumG; ++k) gk[k] = 0.0f;
        bool took = false;
        const uint32_t contributor = s_e[j].y;  // 0-based position in the tile list
        if (hit && contributor < px.last_contributor) {
          const float4 q2 = s_q2[j];
          Pair e;  // a flagged quadrant has no stored pairs: recompute
          if (eval_pair<false>(pixx, pixy, s_q0[j], s_q1[j], q2, 0.0f, e)) {
            const float c_d = e.depth;
            const float m_d = mcoef * (1.0f - kNear / c_d);
            const float dmd_dd = dmd_num / ((dmd_den0 * c_d) * c_d);
            const float4 q3 = s_q3[j], q4 = s_q4[j];
            --idx;  // this pair's place in the run
            if (seg >= 1) {
              float l[7];
              chain_last(ch, l);
              seg_map.add(px, ch.last_alpha, l, e.alpha, m_d, m_d * m_d);
            }
            const float4 g = ch.step(px, e.alpha, c_d, m_d, m_d * m_d, dmd_dd, contributor, q3, q4.x, q4.y);
            if (idx == seg_start && seg >= 1) {  // the next pair starts segment seg - 1
              seg_map.apply(fold);
              ch.T = fold.T;
              ch.acc_r0 = fold.acc_r0;
              ch.acc_r1 = fold.acc_r1;
              ch.acc_r2 = fold.acc_r2;
              ch.accum_depth_rec = fold.accum_depth_rec;
              ch.acc_n0 = fold.acc_n0;
              ch.acc_n1 = fold.acc_n1;
              ch.acc_n2 = fold.acc_n2;
              ch.accum_alpha_rec = fold.accum_alpha_rec;
              ch.last_dL_dT = fold.last_dL_dT;
              seg_map = SegMap();
              --seg;
              seg_start -= seg_len;
            }
            pair_grads(px.dLp0, px.dLp1, px.dLp2, px.dLn0, px.dLn1, px.dLn2, e, s_q1[j], q2.w, g.x, g.y, g.z,
                       pixx, pixy, [&](int k, float v) { gk[k] += v; });
            took = true;
          }
        }
        const uint64_t tookm = __ballot(took);
#pragma unroll
        for (int k = 0; k < kNumG; ++k) s_tmp[k][lane] = gk[k];
        __syncthreads();
        if (lane < kNumG) {  // grouped exactly as the per-pair gradient pass sums
          float acc = 0.0f, grp = 0.0f;
          int in_grp = 0;
          for (uint64_t mm = tookm; mm != 0ull; mm &= mm - 1ull) {
            grp += s_tmp[lane][__builtin_ctzll(mm)];
            if (++in_grp == kGradGroup) {
              acc += grp;
              grp = 0.0f;
              in_grp = 0;
            }
          }
          if (in_grp > 0) acc += grp;
          s_acc[lane][j] = acc;
        }
        __syncthreads();
      }
    }
    __syncthreads();
    if (lane < m) store_part(a, rbeg + s_e[lane].y, q, [&](int k) { return s_acc[k][lane]; });
  }
}

hipError_t launch_blend_bwd(const BlendBwdArgs& a, hipStream_t s) {
  if (a.vs.T == 0) return hipSuccess;
  if (a.vs.B != 8 && a.vs.B != 16 && a.vs.B != 32) return hipErrorInvalidValue;
  const unsigned nblk = (unsigned)(a.vs.T * quads_per_tile(a.vs.B));
  hipLaunchKernelGGL(k_blend_bwd, dim3(nblk * kSplitBwd), dim3(64), 0, s, a);
  return hipGetLastError();
}

// 64 x 64 bit-matrix transpose across a wave: lane i passes row i (bit j =
// M[i][j]) and gets column i back (bit j = M[j][i]).  Recursive block swap:
// at each of the six levels the lanes that differ in bit s exchange the
// off-diagonal s x s blocks of their 2s x 2s block.
SR_DEV uint64_t wave_transpose64(uint64_t x, int lane) {
  uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
  {
    const uint32_t plo = (uint32_t)__shfl_xor((int)lo, 32), phi = (uint32_t)__shfl_xor((int)hi, 32);
    if (lane & 32)
      lo = phi;
    else
      hi = plo;
  }
  auto level = [&](uint32_t v, int sh, uint32_t m) -> uint32_t {
    const uint32_t y = (uint32_t)__shfl_xor((int)v, sh);
    return (lane & sh) ? (((y & ~m) >> sh) | (v & ~m)) : ((v & m) | ((y & m) << sh));
  };
  lo = level(lo, 16, 0x0000FFFFu);
  hi = level(hi, 16, 0x0000FFFFu);
  lo = level(lo, 8, 0x00FF00FFu);
  hi = level(hi, 8, 0x00FF00FFu);
  lo = level(lo, 4, 0x0F0F0F0Fu);
  hi = level(hi, 4, 0x0F0F0F0Fu);
  lo = level(lo, 2, 0x33333333u);
  hi = level(hi, 2, 0x33333333u);
  lo = level(lo, 1, 0x55555555u);
  hi = level(hi, 1, 0x55555555u);
  return (uint64_t)lo | ((uint64_t)hi << 32);
}

// Backward blend, pass 2 of 2.  The per-pair gradients, balanced over the
// whole GPU on the acceptance work list: one wave per (quadrant, 64-entry
// chunk).  A pair's gradients need no pixel recurrence any more (pass 1
// stored its w, dL/dalpha and dL/ddepth).  The acceptance masks (one word
// per pixel), trimmed to the pairs in front of each pixel's last
// contributor, are transposed into one word of pixels per entry; each
// entry's pixels are cut into groups of kGradGroup in pixel order, and the
// groups of all entries are dealt over the lanes, 64 per round, so the
// lanes stay busy however unevenly the entries cover the quadrant.  A lane
// recomputes its group's pairs exactly and sums their 18 partial gradients
// in pixel order; the entry's lane then adds its group sums in order.  Each
// entry's sums go to its (position, quadrant) slot of the partial-gradient
// buffer.  The recompute path of pass 1 groups its sums the same way, so
// the partials have the same bits whichever quadrants missed the pair store.
__global__ __launch_bounds__(64) void k_blend_bwd_grad(BlendBwdArgs a) {
  __shared__ uint4 s_pm[64];    // per pixel: trimmed mask (lo, hi), first stored pair of the chunk
  __shared__ float4 s_pg[64];   // per pixel: dL/dpixel (rgb), dL/dnormal.x
  __shared__ float2 s_pn[64];   // per pixel: dL/dnormal.yz
  __shared__ uint32_t s_pos[64];
  __shared__ float4 s_e0[64], s_e1[64], s_e2[64];  // per entry: record quads 0..2
  __shared__ unsigned long long s_col[64];         // per entry: the pixels that blended it
  __shared__ uint32_t s_tbase[64];                 // per entry: its first group
  __shared__ uint8_t s_towner[64 * (64 / kGradGroup)];  // entry of every group
  __shared__ float s_grp[kNumG][64];               // group sums of the round
  const int B = a.vs.B, QX = B / kQuad, Q = QX * QX;
  const int lane = threadIdx.x;
  if (a.counters[1] & kErrCapacity) return;
  const uint32_t nwork = *a.work_count;
  for (uint32_t w = blockIdx.x; w < nwork; w += gridDim.x) {
    const uint2 item = a.work[nwork - 1u - w];  // from the end: see work_capacity
    const int qi = (int)item.x, c = (int)item.y;
    if (a.q_flag[qi] != 0u) continue;  // done by the staged path of pass 1
    const int tile = qi / Q, q = qi % Q;
    const ViewInfo& vi = a.vs.v[view_of_tile(a.vs, (uint32_t)tile)];
    const int lt = tile - (int)vi.tile_base;
    const int tx = lt % vi.gx, ty = lt / vi.gx;
    const int bx = tx * B + (q % QX) * kQuad, by = ty * B + (q / QX) * kQuad;
    // lane as pixel (lane & 7, lane >> 3) of the quadrant
    const uint32_t pxi = (uint32_t)(bx + (lane & 7)), pyi = (uint32_t)(by + (lane >> 3));
    const bool inside = pxi < (uint32_t)vi.W && pyi < (uint32_t)vi.H;
    const uint32_t rbeg = a.ranges[2 * tile], ntile = a.ranges[2 * tile + 1] - rbeg;
    const int cnt = (int)a.sub_count[qi];
    const uint2* list = a.sub + (size_t)Q * rbeg + (size_t)q * ntile;
    const size_t slot = mask_base(Q, rbeg, ntile, tile, q) + (size_t)c * 64 + lane;
    const int e0 = c * 64;
    const int m = min(64, cnt - e0);
    uint64_t mask = inside ? a.masks[slot] : 0ull;
    const uint32_t foff = mask ? a.frag_off[slot] : 0u;
    uint32_t lc = 0u;
    float4 pg = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float2 pn = make_float2(0.0f, 0.0f);
    if (mask) {
      const size_t HW = (size_t)vi.H * vi.W, pid = (size_t)vi.W * pyi + pxi, pb = vi.pix_base;
      lc = a.n_contrib[2 * pb + pid];
      const float* dL_dpix = a.dL_dpix + 3 * pb;
      const float* dL_dothers = a.dL_dothers + 7 * pb;
      pg = make_float4(dL_dpix[pid], dL_dpix[HW + pid], dL_dpix[2 * HW + pid], dL_dothers[2 * HW + pid]);
      pn = make_float2(dL_dothers[3 * HW + pid], dL_dothers[4 * HW + pid]);
    }
    // lane as entry e0 + lane of the chunk
    const bool has = lane < m;
    const uint2 v = has ? list[e0 + lane] : make_uint2(0u, 0xFFFFFFFFu);
    float4 q0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), q1 = q0, q2 = q0;
    if (has) {
      const float4* r = a.rec + (size_t)kRecQuads * v.x;
      q0 = r[0];
      q1 = r[1];
      q2 = r[2];
    }
    __syncthreads();  // the previous item's readers are done
    s_pos[lane] = v.y;
    s_e0[lane] = q0;
    s_e1[lane] = q1;
    s_e2[lane] = q2;
    __syncthreads();
    if (mask) {
      // entries in front of the last contributor: positions rise along the
      // chunk, so they are the first J, found by binary search
      int J = 0;
#pragma unroll
      for (int st = 64; st > 0; st >>= 1)
        if (J + st <= 64 && s_pos[J + st - 1] < lc) J += st;
      if (J < 64) mask &= (1ull << J) - 1ull;
    }
    s_pm[lane] = make_uint4((uint32_t)mask, (uint32_t)(mask >> 32), foff, 0u);
    s_pg[lane] = pg;
    s_pn[lane] = pn;
    const uint64_t col = wave_transpose64(mask, lane);  // bit p: pixel p blended this lane's entry
    const uint32_t ng = ((uint32_t)__popcll(col) + (kGradGroup - 1)) / kGradGroup;
    uint32_t tincl = ng;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t y = (uint32_t)__shfl_up((int)tincl, off, 64);
      if (lane >= off) tincl += y;
    }
    const uint32_t tbase = tincl - ng, ntask = (uint32_t)__shfl((int)tincl, 63, 64);
    s_col[lane] = col;
    s_tbase[lane] = tbase;
    for (uint32_t i = 0; i < ng; ++i) s_towner[tbase + i] = (uint8_t)lane;
    __syncthreads();
    float acc[kNumG];
#pragma unroll
    for (int k = 0; k < kNumG; ++k) acc[k] = 0.0f;
    for (uint32_t base = 0; base < ntask; base += 64) {  // uniform
      const uint32_t t = base + (uint32_t)lane;
      float g[kNumG];
#pragma unroll
      for (int k = 0; k < kNumG; ++k) g[k] = 0.0f;
      if (t < ntask) {
        const int j = s_towner[t];
        const uint32_t gi = t - s_tbase[j];
        const uint64_t cj = s_col[j];
        // the group's pixels: the set bits of cj from the (gi kGradGroup)-th on
        const int first = select_bit(cj, gi * kGradGroup);
        uint64_t mm = cj & ~((1ull << first) - 1ull);
        const uint64_t below = (1ull << j) - 1ull;  // the entry's rank within a pixel's mask
        int pp[kGradGroup];
        float4 fw[kGradGroup];
#pragma unroll
        for (int k = 0; k < kGradGroup; ++k) {
          pp[k] = -1;
          if (mm != 0ull) {
            const int p = __builtin_ctzll(mm);
            mm &= mm - 1ull;
            pp[k] = p;
            const uint4 pm = s_pm[p];
            const uint64_t mp = (uint64_t)pm.x | ((uint64_t)pm.y << 32);
            fw[k] = a.ps.fW[pm.z + (uint32_t)__popcll(mp & below)];  // (w, dL/dalpha, dL/ddepth)
          }
        }
        const float4 r0 = s_e0[j], r1 = s_e1[j], r2 = s_e2[j];
#pragma unroll
        for (int k = 0; k < kGradGroup; ++k) {
          const int p = pp[k];
          if (p < 0) continue;
          const float4 g4 = s_pg[p];
          const float2 n2 = s_pn[p];
          const float pixx = (float)(bx + (p & 7)), pixy = (float)(by + (p >> 3));
          // the pair itself (s, G, planes) recomputed exactly as the
          // acceptance pass evaluated it: cheaper than fetching it
          Pair e;
          (void)eval_pair<false>(pixx, pixy, r0, r1, r2, 0.0f, e);
          pair_grads(g4.x, g4.y, g4.z, g4.w, n2.x, n2.y, e, r1, r2.w, fw[k].x, fw[k].y, fw[k].z, pixx, pixy,
                     [&](int s, float val) { g[s] += val; });
        }
      }
#pragma unroll
      for (int k = 0; k < kNumG; ++k) s_grp[k][lane] = g[k];
      __syncthreads();
      // the entry's lane adds its groups of this round, in order
      const uint32_t lo = max(tbase, base), hi = min(tbase + ng, base + 64u);
      for (uint32_t u = lo; u < hi; ++u) {
#pragma unroll
        for (int k = 0; k < kNumG; ++k) acc[k] += s_grp[k][u - base];
      }
      __syncthreads();
    }
    if (has) store_part(a, rbeg + v.y, q, [&](int k) { return acc[k]; });
  }
}

hipError_t launch_blend_bwd_grad(const BlendBwdArgs& a, size_t max_items, hipStream_t s) {
  if (max_items == 0) return hipSuccess;
  const unsigned grid = (unsigned)(max_items < 65536 ? max_items : 65536);
  hipLaunchKernelGGL(k_blend_bwd_grad, dim3(grid), dim3(64), 0, s, a);
  return hipGetLastError();
}

struct F3 {
  float x, y, z;
};
SR_DEV F3 f3(float x, float y, float z) { return F3{x, y, z}; }
SR_DEV F3 operator+(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
SR_DEV F3 operator*(float s, F3 a) { return f3(s * a.x, s * a.y, s * a.z); }
SR_DEV F3 mul(F3 v, float s) { return f3(v.x * s, v.y * s, v.z * s); }
SR_DEV F3 had(F3 a, F3 b) { return f3(a.x * b.x, a.y * b.y, a.z * b.z); }
SR_DEV float dot3(F3 a, F3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
SR_DEV float comp(F3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

// SH backward, dsr/cuda_rasterizer/backward.cu:20-139 (same association):
// put(k, c) receives the basis value of each active coefficient k (dL/dsh_k
// = c dL/dRGB), S(k) gives coefficient k; returns dL/d(mean) through the
// view direction.
template <class Get, class Put>
SR_DEV F3 sh_backward_core(int deg, F3 dir_orig, F3 dL_dRGB, Get S, Put put) {
  const float len = sqrtf((dir_orig.x * dir_orig.x + dir_orig.y * dir_orig.y) + dir_orig.z * dir_orig.z);
  const F3 dir = f3(dir_orig.x / len, dir_orig.y / len, dir_orig.z / len);
  F3 dx = f3(0.0f, 0.0f, 0.0f), dy = dx, dz = dx;
  const float x = dir.x, y = dir.y, z = dir.z;
  put(0, kSH_C0);
  if (deg > 0) {
    put(1, -kSH_C1 * y);
    put(2, kSH_C1 * z);
    put(3, -kSH_C1 * x);
    dx = -kSH_C1 * S(3);
    dy = -kSH_C1 * S(1);
    dz = kSH_C1 * S(2);
    if (deg > 1) {
      const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
      put(4, kSH_C2_0 * xy);
      put(5, kSH_C2_1 * yz);
      put(6, kSH_C2_2 * ((2.0f * zz - xx) - yy));
      put(7, kSH_C2_3 * xz);
      put(8, kSH_C2_4 * (xx - yy));
      dx = dx + ((((kSH_C2_0 * y) * S(4) + ((kSH_C2_2 * 2.0f) * -x) * S(6)) + (kSH_C2_3 * z) * S(7)) +
                 ((kSH_C2_4 * 2.0f) * x) * S(8));
      dy = dy + ((((kSH_C2_0 * x) * S(4) + (kSH_C2_1 * z) * S(5)) + ((kSH_C2_2 * 2.0f) * -y) * S(6)) +
                 ((kSH_C2_4 * 2.0f) * -y) * S(8));
      dz = dz + (((kSH_C2_1 * y) * S(5) + (((kSH_C2_2 * 2.0f) * 2.0f) * z) * S(6)) + (kSH_C2_3 * x) * S(7));
      if (deg > 2) {
        put(9, (kSH_C3_0 * y) * (3.0f * xx - yy));
        put(10, (kSH_C3_1 * xy) * z);
        put(11, (kSH_C3_2 * y) * ((4.0f * zz - xx) - yy));
        put(12, (kSH_C3_3 * z) * ((2.0f * zz - 3.0f * xx) - 3.0f * yy));
        put(13, (kSH_C3_4 * x) * ((4.0f * zz - xx) - yy));
        put(14, (kSH_C3_5 * z) * (xx - yy));
        put(15, (kSH_C3_6 * x) * (xx - 3.0f * yy));
        F3 ax = mul(mul(mul(kSH_C3_0 * S(9), 3.0f), 2.0f), xy);
        ax = ax + mul(kSH_C3_1 * S(10), yz);
        ax = ax + mul(mul(kSH_C3_2 * S(11), -2.0f), xy);
        ax = ax + mul(mul(mul(kSH_C3_3 * S(12), -3.0f), 2.0f), xz);
        ax = ax + mul(kSH_C3_4 * S(13), ((-3.0f * xx + 4.0f * zz) - yy));
        ax = ax + mul(mul(kSH_C3_5 * S(14), 2.0f), xz);
        ax = ax + mul(mul(kSH_C3_6 * S(15), 3.0f), (xx - yy));
        F3 ay = mul(mul(kSH_C3_0 * S(9), 3.0f), (xx - yy));
        ay = ay + mul(kSH_C3_1 * S(10), xz);
        ay = ay + mul(kSH_C3_2 * S(11), ((-3.0f * yy + 4.0f * zz) - xx));
        ay = ay + mul(mul(mul(kSH_C3_3 * S(12), -3.0f), 2.0f), yz);
        ay = ay + mul(mul(kSH_C3_4 * S(13), -2.0f), xy);
        ay = ay + mul(mul(kSH_C3_5 * S(14), -2.0f), yz);
        ay = ay + mul(mul(mul(kSH_C3_6 * S(15), -3.0f), 2.0f), xy);
        F3 az = mul(kSH_C3_1 * S(10), xy);
        az = az + mul(mul(mul(kSH_C3_2 * S(11), 4.0f), 2.0f), yz);
        az = az + mul(mul(kSH_C3_3 * S(12), 3.0f), ((2.0f * zz - xx) - yy));
        az = az + mul(mul(mul(kSH_C3_4 * S(13), 4.0f), 2.0f), xz);
        az = az + mul(kSH_C3_5 * S(14), (xx - yy));
        dx = dx + ax;
        dy = dy + ay;
        dz = dz + az;
      }
    }
  }
  const F3 dL_ddir = f3(dot3(dx, dL_dRGB), dot3(dy, dL_dRGB), dot3(dz, dL_dRGB));
  // dnormvdv, auxiliary.h:128-138
  const F3 v = dir_orig, dv = dL_ddir;
  const float sum2 = (v.x * v.x + v.y * v.y) + v.z * v.z;
  const float invsum32 = 1.0f / sqrtf((sum2 * sum2) * sum2);
  F3 r;
  r.x = ((((sum2 - v.x * v.x) * dv.x) - (v.y * v.x) * dv.y) - (v.z * v.x) * dv.z) * invsum32;
  r.y = (((-v.x * v.y) * dv.x + (sum2 - v.y * v.y) * dv.y) - (v.z * v.y) * dv.z) * invsum32;
  r.z = (((-v.x * v.z) * dv.x - (v.y * v.z) * dv.y) + (sum2 - v.z * v.z) * dv.z) * invsum32;
  return r;
}

// SH backward against memory: the first view stores c dL_dRGB (and zeros for
// the inactive coefficients), later views add.  With 16 coefficients and
// 16-byte aligned rows the surfel's 48 coefficients are read and its 48
// gradients written as twelve float4 each (same values, same order of
// operations as the scalar form).
// srow (when not null): the surfel's row staged in LDS by its wave, read
// instead of sh; with to_row the gradients are written back into it (the
// caller stores the wave's rows) instead of to dsh.
template <bool kAdd, bool kRow, bool kToRow>
SR_DEV F3 sh_backward(int deg, int M, const float* sh, F3 dir_orig, F3 dL_dRGB, float* dsh, float4* srow) {
  const int nact = (deg + 1) * (deg + 1);
  if (M == 16 && ((reinterpret_cast<uintptr_t>(sh) | reinterpret_cast<uintptr_t>(dsh)) & 15u) == 0u) {
    float shv[48];
    const float4* s4 = reinterpret_cast<const float4*>(sh);
#pragma unroll
    for (int i = 0; i < 12; ++i) {
      const float4 t = kRow ? srow[i] : s4[i];
      shv[4 * i] = t.x;
      shv[4 * i + 1] = t.y;
      shv[4 * i + 2] = t.z;
      shv[4 * i + 3] = t.w;
    }
    float cv[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) cv[k] = 0.0f;
    const F3 r = sh_backward_core(
        deg, dir_orig, dL_dRGB, [&](int k) { return f3(shv[3 * k], shv[3 * k + 1], shv[3 * k + 2]); },
        [&](int k, float c) { cv[k] = c; });
    float4* d4 = reinterpret_cast<float4*>(dsh);
#pragma unroll
    for (int i = 0; i < 12; ++i) {
      float v[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int j = 4 * i + t, k = j / 3;
        v[t] = k < nact ? cv[k] * comp(dL_dRGB, j % 3) : (kAdd ? -0.0f : 0.0f);
      }
      if (kAdd) {
        const float4 o = d4[i];
        d4[i] = make_float4(o.x + v[0], o.y + v[1], o.z + v[2], o.w + v[3]);
      } else if (kToRow) {
        srow[i] = make_float4(v[0], v[1], v[2], v[3]);
      } else {
        d4[i] = make_float4(v[0], v[1], v[2], v[3]);
      }
    }
    return r;
  }
  if (!kAdd)
    for (int k = nact; k < M; ++k) dsh[3 * k] = dsh[3 * k + 1] = dsh[3 * k + 2] = 0.0f;
  return sh_backward_core(
      deg, dir_orig, dL_dRGB, [&](int k) { return f3(sh[3 * k], sh[3 * k + 1], sh[3 * k + 2]); },
      [&](int k, float c) {
        if (kAdd) {
          dsh[3 * k] += c * dL_dRGB.x;
          dsh[3 * k + 1] += c * dL_dRGB.y;
          dsh[3 * k + 2] += c * dL_dRGB.z;
        } else {
          dsh[3 * k] = c * dL_dRGB.x;
          dsh[3 * k + 1] = c * dL_dRGB.y;
          dsh[3 * k + 2] = c * dL_dRGB.z;
        }
      });
}

// SH backward of one record: stores (first view) or adds c * dL_dRGB to the
// surfel's dL_dsh and returns the gradient of the mean through the view
// direction.
template <bool kRow, bool kToRow>
SR_DEV F3 sh_grad(const PreprocessBwdArgs& a, bool first, int idx, size_t r, const float* campos, float m0,
                  float m1, float m2, F3 dRGB, float4* srow) {
  const uchar4 cl = reinterpret_cast<const uchar4*>(a.clamped)[r];
  if (cl.x) dRGB.x = 0.0f;
  if (cl.y) dRGB.y = 0.0f;
  if (cl.z) dRGB.z = 0.0f;
  const F3 dir_orig = f3(m0 - campos[0], m1 - campos[1], m2 - campos[2]);
  float* dsh = a.dL_dsh + (size_t)idx * a.M * 3;
  const float* sh = a.shs + (size_t)idx * a.M * 3;
  return first ? sh_backward<false, kRow, kToRow>(a.D, a.M, sh, dir_orig, dRGB, dsh, srow)
               : sh_backward<true, kRow, false>(a.D, a.M, sh, dir_orig, dRGB, dsh, srow);
}

// Gradients of one surfel summed over the views that see it.
struct SurfelGrad {
  float m[3], sc[2], q[4];
};

// Record r's 18 partial gradients summed over its slots (one per instance
// and quadrant holding it, a contiguous run in instance order, quadrants in
// order): a fixed order, so the sums are the same on every run.
SR_DEV void reduce_record(const PreprocessBwdArgs& a, size_t r, float (&gs)[18]) {
#pragma unroll
  for (int k = 0; k < 18; ++k) gs[k] = 0.0f;
  if (a.tiles[r] == 0 || a.part == nullptr || (a.counters[1] & kErrCapacity)) return;
  const uint2 rs = a.rslot[r];  // the record's slots: its instances in order, each one's quadrants in order
  const uint32_t s0 = rs.x, s1 = rs.x + rs.y;
  // four flags in flight at a time
  for (uint32_t s = s0; s < s1; s += 4) {
    uint8_t fl[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) fl[t] = a.part_nz[min(s + (uint32_t)t, s1 - 1u)];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      if (s + (uint32_t)t >= s1 || fl[t] == 0) continue;
      const float4* p = reinterpret_cast<const float4*>(a.part + (size_t)(s + (uint32_t)t) * kPartStride);
      const float4 p0 = p[0], p1 = p[1], p2 = p[2], p3 = p[3], p4 = p[4];
      gs[0] += p0.x; gs[1] += p0.y; gs[2] += p0.z; gs[3] += p0.w;
      gs[4] += p1.x; gs[5] += p1.y; gs[6] += p1.z; gs[7] += p1.w;
      gs[8] += p2.x; gs[9] += p2.y; gs[10] += p2.z; gs[11] += p2.w;
      gs[12] += p3.x; gs[13] += p3.y; gs[14] += p3.z; gs[15] += p3.w;
      gs[16] += p4.x; gs[17] += p4.y;
    }
  }
}

// Backward preprocess of surfel idx in view v (record r),
// dsr/cuda_rasterizer/backward.cu:443-630, from the record's reduced blend
// gradients gs.  Adds the view's contributions to g (and to dL_dsh, owned by
// this thread) and writes the record's outputs: dL_dtransMat (with the
// AABB-centre term for a precomputed T), dL_dmean2D (the densification
// statistic), dL_dopacity and dL_dcolors.
template <bool kRow, bool kToRow>
SR_DEV void preprocess_bwd_view(const PreprocessBwdArgs& a, int v, int idx, size_t r, bool first, SurfelGrad& g,
                                const float (&gs)[18], float4* srow) {
  const ViewInfo& vi = a.vs.v[v];
  const float tanx = vi.tanx, tany = vi.tany;
  const float focal_x = (float)vi.W / (2.0f * tanx);
  const float focal_y = (float)vi.H / (2.0f * tany);
  const int W = f2i((focal_x * tanx) * 2.0f);
  const int H = f2i((focal_y * tany) * 2.0f);
  const float* pm = a.pm + 16 * v;
  const float* vm = a.vm + 16 * v;
  const float* campos = a.campos + 3 * v;
  float* gT = a.dL_dtransMat + 9 * r;
  const float m0 = a.means3D[3 * idx], m1 = a.means3D[3 * idx + 1], m2 = a.means3D[3 * idx + 2];
  const bool precomp = a.tprec != nullptr;
  F3 T0, T1, T2, Rc0, Rc1;
  float qw = 0.0f, qx = 0.0f, qy = 0.0f, qz = 0.0f, scx = 0.0f, scy = 0.0f;
  float P4[3][4];
  if (precomp) {
    const float* t = a.tprec + 9 * (size_t)idx;
    T0 = f3(t[0], t[1], t[2]);
    T1 = f3(t[3], t[4], t[5]);
    T2 = f3(t[6], t[7], t[8]);
  } else {
    const float* q = a.rots + 4 * idx;
    const float q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
    const float n2 = ((q3 * q3 + q0 * q0) + q1 * q1) + q2 * q2;
    const float s = 1.0f / sqrtf(n2);
    qw = q0 * s;
    qx = q1 * s;
    qy = q2 * s;
    qz = q3 * s;
    const float w = qw, x = qx, y = qy, z = qz;
    Rc0 = f3(1.0f - 2.0f * (y * y + z * z), 2.0f * (x * y + w * z), 2.0f * (x * z - w * y));
    Rc1 = f3(2.0f * (x * y - w * z), 1.0f - 2.0f * (x * x + z * z), 2.0f * (y * z + w * x));
    scx = a.scales[2 * idx];
    scy = a.scales[2 * idx + 1];
    const F3 L0c = (1.0f * scx) * Rc0, L1c = (1.0f * scy) * Rc1;
    const float hw = (float)W / 2.0f, hw1 = (float)(W - 1) / 2.0f;
    const float hh = (float)H / 2.0f, hh1 = (float)(H - 1) / 2.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      P4[0][i] = pm[4 * i] * hw + pm[4 * i + 3] * hw1;
      P4[1][i] = pm[4 * i + 1] * hh + pm[4 * i + 3] * hh1;
      P4[2][i] = pm[4 * i + 3];
    }
    auto dotM = [&](const F3& u, float u4, int k) {
      return ((u.x * P4[k][0] + u.y * P4[k][1]) + u.z * P4[k][2]) + u4 * P4[k][3];
    };
    const F3 pc = f3(m0, m1, m2);
    T0 = f3(dotM(L0c, 0.0f, 0), dotM(L1c, 0.0f, 0), dotM(pc, 1.0f, 0));
    T1 = f3(dotM(L0c, 0.0f, 1), dotM(L1c, 0.0f, 1), dotM(pc, 1.0f, 1));
    T2 = f3(dotM(L0c, 0.0f, 2), dotM(L1c, 0.0f, 2), dotM(pc, 1.0f, 2));
  }
  F3 dT0 = f3(gs[0], gs[1], gs[2]), dT1 = f3(gs[3], gs[4], gs[5]), dT2 = f3(gs[6], gs[7], gs[8]);
  float* gm2 = a.dL_dmean2D + 3 * r;
  const float gmx = gs[9], gmy = gs[10];
  bool updated = false;
  if (gmx != 0.0f || gmy != 0.0f) {
    const F3 tv = f3(9.0f, 9.0f, -1.0f);
    const float d = (tv.x * (T2.x * T2.x) + tv.y * (T2.y * T2.y)) + tv.z * (T2.z * T2.z);
    const float invd = 1.0f / d;
    const F3 f = f3(tv.x * invd, tv.y * invd, tv.z * invd);
    const F3 dL_dT0 = had(gmx * f, T2);
    const F3 dL_dT1 = had(gmy * f, T2);
    F3 dL_dT3 = had(gmx * f, T0) + had(gmy * f, T1);
    const F3 dL_df = had(gmx * T0, T2) + had(gmy * T1, T2);
    const float dotdf = (dL_df.x * f.x + dL_df.y * f.y) + dL_df.z * f.z;
    const float dL_dd = (float)((double)dotdf * (-1.0 / (double)d));
    const F3 dd_dT3 = f3((tv.x * T2.x) * 2.0f, (tv.y * T2.y) * 2.0f, (tv.z * T2.z) * 2.0f);
    dL_dT3 = dL_dT3 + dL_dd * dd_dT3;
    dT0 = dT0 + dL_dT0;
    dT1 = dT1 + dL_dT1;
    dT2 = dT2 + dL_dT3;
    updated = precomp;
  }
  // dL_dtransMat: the blend's sums, with the AABB-centre term only for a
  // precomputed T
  if (updated) {
    gT[0] = dT0.x; gT[1] = dT0.y; gT[2] = dT0.z;
    gT[3] = dT1.x; gT[4] = dT1.y; gT[5] = dT1.z;
    gT[6] = dT2.x; gT[7] = dT2.y; gT[8] = dT2.z;
  } else {
#pragma unroll
    for (int k = 0; k < 9; ++k) gT[k] = gs[k];
  }
  a.dL_dopacity[r] = gs[14];
  a.dL_dcolors[3 * r] = gs[15];
  a.dL_dcolors[3 * r + 1] = gs[16];
  a.dL_dcolors[3 * r + 2] = gs[17];
  if (!precomp) {
    float dM[3][3];
#pragma unroll
    for (int cc = 0; cc < 3; ++cc)
#pragma unroll
      for (int i = 0; i < 3; ++i)
        dM[cc][i] = (P4[0][i] * comp(dT0, cc) + P4[1][i] * comp(dT1, cc)) + P4[2][i] * comp(dT2, cc);
    const float gn[3] = {gs[11], gs[12], gs[13]};
    F3 dL_dtn = f3((vm[0] * gn[0] + vm[1] * gn[1]) + vm[2] * gn[2], (vm[4] * gn[0] + vm[5] * gn[1]) + vm[6] * gn[2],
                   (vm[8] * gn[0] + vm[9] * gn[1]) + vm[10] * gn[2]);
    // the camera-facing sign the forward chose (backward.cu:550-555 re-derives
    // it from p_view . normal, which the forward's FP32 test gives too; for
    // a surfel whose forward ran in FP64 the forward's decision is the one
    // its outputs used)
    const float mult = a.rec[(size_t)kRecQuads * r + 4].w > 0.0f ? 1.0f : -1.0f;
    dL_dtn = mult * dL_dtn;
    const F3 dRS0 = f3(dM[0][0], dM[0][1], dM[0][2]);
    const F3 dRS1 = f3(dM[1][0], dM[1][1], dM[1][2]);
    const F3 dR0 = f3(dRS0.x * scx, dRS0.y * scx, dRS0.z * scx);
    const F3 dR1 = f3(dRS1.x * scy, dRS1.y * scy, dRS1.z * scy);
    const F3 dR2 = dL_dtn;
    const float vR[3][3] = {{dR0.x, dR0.y, dR0.z}, {dR1.x, dR1.y, dR1.z}, {dR2.x, dR2.y, dR2.z}};
    const float w = qw, x = qx, y = qy, z = qz;
    g.q[0] += 2.0f * ((x * (vR[1][2] - vR[2][1]) + y * (vR[2][0] - vR[0][2])) + z * (vR[0][1] - vR[1][0]));
    g.q[1] += 2.0f * ((((-2.0f * x) * (vR[1][1] + vR[2][2]) + y * (vR[0][1] + vR[1][0])) + z * (vR[0][2] + vR[2][0])) +
                      w * (vR[1][2] - vR[2][1]));
    g.q[2] += 2.0f * (((x * (vR[0][1] + vR[1][0]) - (2.0f * y) * (vR[0][0] + vR[2][2])) + z * (vR[1][2] + vR[2][1])) +
                      w * (vR[2][0] - vR[0][2]));
    g.q[3] += 2.0f * (((x * (vR[0][2] + vR[2][0]) + y * (vR[1][2] + vR[2][1])) - (2.0f * z) * (vR[0][0] + vR[1][1])) +
                      w * (vR[0][1] - vR[1][0]));
    g.sc[0] += (dRS0.x * Rc0.x + dRS0.y * Rc0.y) + dRS0.z * Rc0.z;
    g.sc[1] += (dRS1.x * Rc1.x + dRS1.y * Rc1.y) + dRS1.z * Rc1.z;
    F3 dm3 = f3(dM[2][0], dM[2][1], dM[2][2]);
    if (a.shs != nullptr)
      dm3 = dm3 + sh_grad<kRow, kToRow>(a, first, idx, r, campos, m0, m1, m2, f3(gs[15], gs[16], gs[17]), srow);
    g.m[0] += dm3.x;
    g.m[1] += dm3.y;
    g.m[2] += dm3.z;
  } else if (a.shs != nullptr) {
    const F3 dm3 = sh_grad<kRow, kToRow>(a, first, idx, r, campos, m0, m1, m2, f3(gs[15], gs[16], gs[17]), srow);
    g.m[0] += dm3.x;
    g.m[1] += dm3.y;
    g.m[2] += dm3.z;
  }
  const float depth = a.rec[(size_t)kRecQuads * r + 2].x;  // Tw.z = transMat[8]
  gm2[0] = (float)(((double)(gT[2] * depth) * 0.5) * (double)(float)W);
  gm2[1] = (float)(((double)(gT[5] * depth) * 0.5) * (double)(float)H);
  gm2[2] = 0.0f;
}

// One thread per surfel, looping over the views; with one view every sum is
// 0 + x, so the results equal the single-view formulas bit for bit.  Every
// per-surfel and per-record output is written exactly once (zeros for a
// record no view sees), so none of them needs clearing beforehand.
// Degree-3 SH in 16-byte aligned rows move through LDS: each wave loads its
// 64 surfels' rows (12 KB, contiguous) with coalesced loads and, for one
// view, stores the gradient rows the same way, instead of each lane moving
// its 192-byte row in 16-byte pieces 192 bytes apart.
constexpr int kShRow = 12;  // float4 per degree-3 SH row
template <bool kRow, bool kToRow>
SR_DEV void preprocess_bwd_surfel(const PreprocessBwdArgs& a, int idx, float4* srow);
// Three waves per SIMD (a few registers spill to scratch): with the 48 KB of
// staged rows, two per SIMD was slower (0.086 against 0.071 ms at 300k surfels).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 8))) void k_preprocess_bwd(
    PreprocessBwdArgs a) {
  __shared__ float4 s_sh[256 * kShRow];
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const bool staged = a.shs != nullptr && a.M == 16 &&
                      ((reinterpret_cast<uintptr_t>(a.shs) | reinterpret_cast<uintptr_t>(a.dL_dsh)) & 15u) == 0u;
  const bool to_row = staged && a.vs.V == 1;  // uniform
  const int w0 = blockIdx.x * blockDim.x + wave * 64;  // the wave's first surfel
  const int nrow = min(64, a.P - w0) * kShRow;         // float4 of the wave's rows (<= 0: none)
  float4* wrows = s_sh + wave * 64 * kShRow;
  if (staged) {
    const float4* src = reinterpret_cast<const float4*>(a.shs) + (size_t)w0 * kShRow;
#pragma unroll
    for (int t = 0; t < kShRow; ++t)
      if (t * 64 + lane < nrow) wrows[t * 64 + lane] = src[t * 64 + lane];
  }
  __syncthreads();
  if (idx < a.P) {
    if (to_row) preprocess_bwd_surfel<true, true>(a, idx, wrows + lane * kShRow);
    else if (staged) preprocess_bwd_surfel<true, false>(a, idx, wrows + lane * kShRow);
    else preprocess_bwd_surfel<false, false>(a, idx, wrows);
  }
  if (to_row) {
    __syncthreads();
    float4* dst = reinterpret_cast<float4*>(a.dL_dsh) + (size_t)w0 * kShRow;
#pragma unroll
    for (int t = 0; t < kShRow; ++t)
      if (t * 64 + lane < nrow) dst[t * 64 + lane] = wrows[t * 64 + lane];
  }
}

// Gradients of surfel idx over all views; kRow: its SH row is staged in LDS
// at srow, kToRow: its dL_dsh row goes there too.
template <bool kRow, bool kToRow>
SR_DEV void preprocess_bwd_surfel(const PreprocessBwdArgs& a, int idx, float4* srow) {
  SurfelGrad g{};
  bool any = false;
  const bool sum_rec = a.dL_dopacity_out != a.dL_dopacity;
  float op = 0.0f, c0 = 0.0f, c1 = 0.0f, c2 = 0.0f;
  for (int v = 0; v < a.vs.V; ++v) {
    const size_t r = (size_t)v * a.P + idx;
    if (!(a.radii[r] > 0)) {
#pragma unroll
      for (int k = 0; k < 9; ++k) a.dL_dtransMat[9 * r + k] = 0.0f;
      a.dL_dmean2D[3 * r] = a.dL_dmean2D[3 * r + 1] = a.dL_dmean2D[3 * r + 2] = 0.0f;
      a.dL_dopacity[r] = 0.0f;
      a.dL_dcolors[3 * r] = a.dL_dcolors[3 * r + 1] = a.dL_dcolors[3 * r + 2] = 0.0f;
      continue;
    }
    float gs[18];
    reduce_record(a, r, gs);
    preprocess_bwd_view<kRow, kToRow>(a, v, idx, r, !any, g, gs, srow);
    any = true;
    if (sum_rec) {
      op += gs[14];
      c0 += gs[15];
      c1 += gs[16];
      c2 += gs[17];
    }
  }
  if (!any && a.shs != nullptr) {
    if (kToRow) {
      for (int k = 0; k < kShRow; ++k) srow[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    } else {
      float* dsh = a.dL_dsh + (size_t)idx * a.M * 3;
      for (int k = 0; k < 3 * a.M; ++k) dsh[k] = 0.0f;
    }
  }
  a.dL_drots[4 * idx] = g.q[0];
  a.dL_drots[4 * idx + 1] = g.q[1];
  a.dL_drots[4 * idx + 2] = g.q[2];
  a.dL_drots[4 * idx + 3] = g.q[3];
  a.dL_dscales[2 * idx] = g.sc[0];
  a.dL_dscales[2 * idx + 1] = g.sc[1];
  a.dL_dmean3D[3 * idx] = g.m[0];
  a.dL_dmean3D[3 * idx + 1] = g.m[1];
  a.dL_dmean3D[3 * idx + 2] = g.m[2];
  if (sum_rec) {
    a.dL_dopacity_out[idx] = op;
    a.dL_dcolors_out[3 * idx] = c0;
    a.dL_dcolors_out[3 * idx + 1] = c1;
    a.dL_dcolors_out[3 * idx + 2] = c2;
  }
}

hipError_t launch_preprocess_bwd(const PreprocessBwdArgs& a, hipStream_t s) {
  if (a.P == 0 || a.vs.V == 0) return hipSuccess;
  hipLaunchKernelGGL(k_preprocess_bwd, dim3((a.P + 255) / 256), dim3(256), 0, s, a);
  return hipGetLastError();
}

// Backward preprocess of the _part variant (one view; dsp = submodules/
// diff-surfel-rasterization_part/cuda_rasterizer): per visible surfel the
// blend's partial gradients summed over its instances, then the AABB-centre
// adjoint added into dL_dtransMat for every visible surfel and the
// densification statistic (computeAABB, dsp/backward.cu:621-671), then the
// camera-space homography's adjoint (computeTransMat, :473-545) and the SH
// backward.  Every output is written (zeros for a culled surfel).
__global__ __launch_bounds__(256) void k_part_preprocess_bwd(PreprocessBwdArgs a) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a.P) return;
  const int S = a.S;
  float* gT = a.dL_dtransMat + 9 * idx;
  float* gm2 = a.dL_dmean2D + 3 * idx;
  if (!(a.radii[idx] > 0)) {
    for (int k = 0; k < 9; ++k) gT[k] = 0.0f;
    for (int k = 0; k < 3; ++k) gm2[k] = a.dL_dmean3D[3 * idx + k] = a.dL_dcolors[3 * idx + k] = 0.0f;
    a.dL_dopacity[idx] = 0.0f;
    a.dL_dscales[2 * idx] = a.dL_dscales[2 * idx + 1] = 0.0f;
    for (int k = 0; k < 4; ++k) a.dL_drots[4 * idx + k] = 0.0f;
    for (int ch = 0; ch < S; ++ch) a.dL_dsem[(size_t)idx * S + ch] = 0.0f;
    if (a.shs != nullptr)
      for (int k = 0; k < 3 * a.M; ++k) a.dL_dsh[(size_t)idx * a.M * 3 + k] = 0.0f;
    return;
  }
  float gs[18];
  reduce_record(a, (size_t)idx, gs);
  // semantic partials, in the same (instance, quadrant) order
  {
    float sm[16];
    for (int ch = 0; ch < 16; ++ch) sm[ch] = 0.0f;
    if (S > 0 && a.part != nullptr && a.tiles[idx] > 0 && !(a.counters[1] & kErrCapacity)) {
      const uint2 rs = a.rslot[idx];  // the record's slots, in the order of reduce_record
      for (uint32_t sl = rs.x; sl < rs.x + rs.y; ++sl) {
        if (a.part_nz[sl] == 0) continue;
        for (int ch = 0; ch < S; ++ch) sm[ch] += a.sem_part[(size_t)sl * S + ch];
      }
    }
    for (int ch = 0; ch < S; ++ch) a.dL_dsem[(size_t)idx * S + ch] = sm[ch];
  }
  const ViewInfo& vi = a.vs.v[0];
  const float tanx = vi.tanx, tany = vi.tany;
  const float fx = (float)vi.W / (2.0f * tanx), fy = (float)vi.H / (2.0f * tany);
  const float Wh = fx * tanx, Hh = fy * tany;  // "W", "H" of BACKWARD::preprocess (:685-686)
  const float4* rc = a.rec + (size_t)kRecQuads * idx;
  const float4 r0 = rc[0], r1 = rc[1], r2 = rc[2];
  const F3 T0 = f3(r0.x, r0.y, r0.z), T1 = f3(r0.w, r1.x, r1.y), T3 = f3(r1.z, r1.w, r2.x);
  F3 dT0 = f3(gs[0], gs[1], gs[2]), dT1 = f3(gs[3], gs[4], gs[5]), dT2 = f3(gs[6], gs[7], gs[8]);
  {
    const float gmx = gs[9], gmy = gs[10];
    const F3 tv = f3(1.0f, 1.0f, -1.0f);
    const float d = (tv.x * (T3.x * T3.x) + tv.y * (T3.y * T3.y)) + tv.z * (T3.z * T3.z);
    const float invd = 1.0f / d;
    const F3 f = f3(tv.x * invd, tv.y * invd, tv.z * invd);
    const F3 dL_dT0 = had(gmx * f, T3);
    const F3 dL_dT1 = had(gmy * f, T3);
    F3 dL_dT3 = had(gmx * f, T0) + had(gmy * f, T1);
    const F3 dL_df = had(gmx * T0, T3) + had(gmy * T1, T3);
    const float dotdf = (dL_df.x * f.x + dL_df.y * f.y) + dL_df.z * f.z;
    const float dL_dd = (float)((double)dotdf * (-1.0 / (double)d));
    const F3 dd_dT3 = f3((tv.x * T3.x) * 2.0f, (tv.y * T3.y) * 2.0f, (tv.z * T3.z) * 2.0f);
    dL_dT3 = dL_dT3 + dL_dd * dd_dT3;
    dT0 = dT0 + dL_dT0;
    dT1 = dT1 + dL_dT1;
    dT2 = dT2 + dL_dT3;
  }
  gT[0] = dT0.x; gT[1] = dT0.y; gT[2] = dT0.z;
  gT[3] = dT1.x; gT[4] = dT1.y; gT[5] = dT1.z;
  gT[6] = dT2.x; gT[7] = dT2.y; gT[8] = dT2.z;
  const float z = r2.x;
  gm2[0] = (dT0.z * z) * Wh;
  gm2[1] = (dT1.z * z) * Hh;
  gm2[2] = 0.0f;
  a.dL_dopacity[idx] = gs[14];
  a.dL_dcolors[3 * idx] = gs[15];
  a.dL_dcolors[3 * idx + 1] = gs[16];
  a.dL_dcolors[3 * idx + 2] = gs[17];
  const float m0 = a.means3D[3 * idx], m1 = a.means3D[3 * idx + 1], m2 = a.means3D[3 * idx + 2];
  F3 dm3 = f3(0.0f, 0.0f, 0.0f);
  if (a.tprec == nullptr) {
    const float* vm = a.vm;
    const float cx = fx * tanx, cy = fy * tany;  // intrins.zw of the backward (:596)
    const float* q = a.rots + 4 * idx;
    const float q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
    const float n2 = ((q3 * q3 + q0 * q0) + q1 * q1) + q2 * q2;
    const float s = 1.0f / sqrtf(n2);
    const float w = q0 * s, x = q1 * s, y = q2 * s, zq = q3 * s;
    const F3 Rc0 = f3(1.0f - 2.0f * (y * y + zq * zq), 2.0f * (x * y + w * zq), 2.0f * (x * zq - w * y));
    const F3 Rc1 = f3(2.0f * (x * y - w * zq), 1.0f - 2.0f * (x * x + zq * zq), 2.0f * (y * zq + w * x));
    const F3 Rc2 = f3(2.0f * (x * zq + w * y), 2.0f * (y * zq - w * x), 1.0f - 2.0f * (x * x + y * y));
    const float sx = a.scales[2 * idx], sy = a.scales[2 * idx + 1];
    auto Wt = [&](F3 v) {
      return f3((vm[0] * v.x + vm[1] * v.y) + vm[2] * v.z, (vm[4] * v.x + vm[5] * v.y) + vm[6] * v.z,
                (vm[8] * v.x + vm[9] * v.y) + vm[10] * v.z);
    };
    // dL_dM = (P^T dL_dT^T): column j = (fx dTu_j, fy dTv_j, (cx dTu_j + cy dTv_j) + dTw_j)
    const F3 dM0 = f3(fx * dT0.x, fy * dT1.x, (cx * dT0.x + cy * dT1.x) + dT2.x);
    const F3 dM1 = f3(fx * dT0.y, fy * dT1.y, (cx * dT0.y + cy * dT1.y) + dT2.y);
    const F3 dM2 = f3(fx * dT0.z, fy * dT1.z, (cx * dT0.z + cy * dT1.z) + dT2.z);
    const F3 dRS0 = Wt(dM0), dRS1 = Wt(dM1), dpw = Wt(dM2);
    F3 dL_dtn = Wt(f3(gs[11], gs[12], gs[13]));
    // DUAL_VISIABLE (:528-533): the sign from tn = W R[2] against p_view
    const F3 tn = f3((vm[0] * Rc2.x + vm[4] * Rc2.y) + vm[8] * Rc2.z, (vm[1] * Rc2.x + vm[5] * Rc2.y) + vm[9] * Rc2.z,
                     (vm[2] * Rc2.x + vm[6] * Rc2.y) + vm[10] * Rc2.z);
    const float pvx = ((vm[0] * m0 + vm[4] * m1) + vm[8] * m2) + vm[12];
    const float pvy = ((vm[1] * m0 + vm[5] * m1) + vm[9] * m2) + vm[13];
    const float pvz = ((vm[2] * m0 + vm[6] * m1) + vm[10] * m2) + vm[14];
    const float cosv = ((-tn.x * pvx) + (-tn.y * pvy)) + (-tn.z * pvz);
    dL_dtn = (cosv > 0.0f ? 1.0f : -1.0f) * dL_dtn;
    const F3 dR0 = f3(dRS0.x * sx, dRS0.y * sx, dRS0.z * sx);
    const F3 dR1 = f3(dRS1.x * sy, dRS1.y * sy, dRS1.z * sy);
    const float vR[3][3] = {{dR0.x, dR0.y, dR0.z}, {dR1.x, dR1.y, dR1.z}, {dL_dtn.x, dL_dtn.y, dL_dtn.z}};
    float* gq = a.dL_drots + 4 * idx;
    gq[0] = 2.0f * ((x * (vR[1][2] - vR[2][1]) + y * (vR[2][0] - vR[0][2])) + zq * (vR[0][1] - vR[1][0]));
    gq[1] = 2.0f * ((((-2.0f * x) * (vR[1][1] + vR[2][2]) + y * (vR[0][1] + vR[1][0])) + zq * (vR[0][2] + vR[2][0])) +
                    w * (vR[1][2] - vR[2][1]));
    gq[2] = 2.0f * (((x * (vR[0][1] + vR[1][0]) - (2.0f * y) * (vR[0][0] + vR[2][2])) + zq * (vR[1][2] + vR[2][1])) +
                    w * (vR[2][0] - vR[0][2]));
    gq[3] = 2.0f * (((x * (vR[0][2] + vR[2][0]) + y * (vR[1][2] + vR[2][1])) - (2.0f * zq) * (vR[0][0] + vR[1][1])) +
                    w * (vR[0][1] - vR[1][0]));
    a.dL_dscales[2 * idx] = (dRS0.x * Rc0.x + dRS0.y * Rc0.y) + dRS0.z * Rc0.z;
    a.dL_dscales[2 * idx + 1] = (dRS1.x * Rc1.x + dRS1.y * Rc1.y) + dRS1.z * Rc1.z;
    dm3 = dpw;
  } else {
    // the reference reads scales and rotations it was not given here; with
    // a precomputed T these gradients are zero
    a.dL_dscales[2 * idx] = a.dL_dscales[2 * idx + 1] = 0.0f;
    for (int k = 0; k < 4; ++k) a.dL_drots[4 * idx + k] = 0.0f;
  }
  if (a.shs != nullptr) {
    F3 dRGB = f3(gs[15], gs[16], gs[17]);
    const uchar4 cl = reinterpret_cast<const uchar4*>(a.clamped)[idx];
    if (cl.x) dRGB.x = 0.0f;
    if (cl.y) dRGB.y = 0.0f;
    if (cl.z) dRGB.z = 0.0f;
    const F3 dir_orig = f3(m0 - a.campos[0], m1 - a.campos[1], m2 - a.campos[2]);
    dm3 = dm3 + sh_backward<false, false, false>(a.D, a.M, a.shs + (size_t)idx * a.M * 3, dir_orig, dRGB,
                                                  a.dL_dsh + (size_t)idx * a.M * 3, nullptr);
  }
  a.dL_dmean3D[3 * idx] = dm3.x;
  a.dL_dmean3D[3 * idx + 1] = dm3.y;
  a.dL_dmean3D[3 * idx + 2] = dm3.z;
}

hipError_t launch_part_preprocess_bwd(const PreprocessBwdArgs& a, hipStream_t s) {
  if (a.P == 0) return hipSuccess;
  hipLaunchKernelGGL(k_part_preprocess_bwd, dim3((a.P + 255) / 256), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace sr
