// Image-source-method kernels for gfx950 (stages a, b and c of simulateRIR).
//
// Reference path (DavidDiazGuerra/gpuRIR v1.2.0, src/gpuRIR_cuda.cu):
//   calcAmpTau_kernel :245-375   -> image_eval() below, fused into the accumulation
//   generateRIR_kernel :377-390, reduceRIR_kernel :392-413, cuda_rirGenerator :618-674
//                                 -> ism_accumulate_kernel + ism_finalize_kernel
//   generateRIR_kernel_lut :559-572, create_sinc_texture_lut :578-616
//                                 -> MODE == PREC_LUT (LDS table, FP32 lerp)
//   envPred_kernel :415-443, diffRev_kernel :445-461 (+ cuRAND :872-886)
//                                 -> ism_finalize_kernel (Philox4x32-10 noise)
//
// Design (one work item = one RIR x one image chunk x one time tile, one workgroup at a time):
//   * stage (a) per image in FP64 for the delay (image vector is exact for float
//     inputs; sqrt and one multiply by fs/c), split into an integer sample n0 and
//     a float fraction f;  amplitude, directivities and tap constants in FP32;
//   * images whose window misses the tile are culled; live ones are compacted
//     into an LDS queue with wave ballots;
//   * stage (b) scatters each live image's taps into an LDS-resident partial RIR
//     held in 64-bit fixed point (ds_add_u64: exact, order-independent sums; the
//     scale is a power of two from a bound on every tap of the workgroup).  A wave
//     takes one image at a time and its 64 lanes cover consecutive tap offsets j
//     (conflict-free atomics, per-offset constants in registers); the windowed sinc
//     uses the identities
//        sin(pi (j - f))        = (-1)^(j+1) sin(pi f)
//        cos(2 pi (j - f) / Tw) = cos(2 pi j/Tw) cos(2 pi f/Tw) + sin(2 pi j/Tw) sin(2 pi f/Tw)
//     so a tap is one FMA for the offset, one reciprocal, two FMAs for the window,
//     one product and the conversion to fixed point;
//   * the tile is flushed with coalesced stores, either straight into the output
//     (one chunk per RIR) or into a partial row that partial_sum_kernel adds up
//     in a fixed order (deterministic across runs for a given plan);
//   * ism_finalize_kernel computes the Sabine envelope from the last 10 ms of the
//     early part and writes the diffuse tail;
//   * the adjoint kernels at the end of this file give the gradients.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>

#include "internal.h"

namespace gpurir {

namespace {

constexpr int WG = 256;
constexpr int NWAVE = WG / 64;
constexpr int QCAP = 512;
constexpr int QPAD = 16;  // queue entries read ahead past the live part (4 per wave)
// The B/2 sin values of the queue are stored per wave: image q (taken by wave q % NWAVE) at
// qS[(q % NWAVE) * QSW + q / NWAVE], so one 16-byte read gives a wave its next four images.
constexpr int QSW = (QCAP + QPAD) / NWAVE;
static_assert(QSW % 4 == 0 && QSW >= QCAP / NWAVE + 4, "per-wave S rows: float4-aligned, one group of read-ahead");
// Mixed precision: images (so taps per word) between two moves of the 32-bit sums into FP32.
constexpr int FP16_FLUSH = 1023;
static_assert(FP16_FLUSH >= QCAP, "a queue round must fit between two moves");

// Packed FP16 pair (v_pk_* f16 instructions) for the mixed-precision sinc.
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

__device__ inline float h2_bits(half2_t v) { return __builtin_bit_cast(float, v); }
__device__ inline half2_t bits_h2(float v) { return __builtin_bit_cast(half2_t, v); }
__device__ inline half2_t h2(float lo, float hi) { return half2_t{(_Float16)lo, (_Float16)hi}; }

// Reflection coefficients and diffuse-tail levels read at run time: when set, room k's
// six betas are beta[6k .. 6k+5] and its level is level[k] (device arrays that a caller may
// update between replays of a captured plan, e.g. an optimiser step); when null, the values
// planned into RoomDev.
struct RoomParams {
    const float* beta;
    const float* level;
};

struct AccArgs {
    const RoomDev* rooms;
    const Item* items;
    const float* pos_src;
    const float* pos_rcv;
    const float* orv_src;
    const float* orv_rcv;
    const float4* jtab;   // [n_pass*64] per lane of each pass: (sigma_j j, cos(2 pi j/Tw), sin(2 pi j/Tw), sigma_j)
    const float2* lut;    // LUT mode, polyphase: [16][lut_row] (T[16k+ph], T[16k+ph+1]-T[16k+ph])
    float* out;
    float* part;
    int spkr, mic;
    int J0, nJ;           // tap offsets j = t - floor(tau) that can be in the window: [J0, J0 + nJ)
    int n_pass;           // passes of 64 lanes over j = J0 + 64 p + lane
    unsigned mask_bits;   // bit p: pass p has lanes whose window membership depends on f (0: no test)
    float halfTw;         // support |t - tau| < halfTw
    float inv_tw2;        // 2/Tw
    int lut_center, lut_row, lut_k_lo;  // LUT: center index, row length, first k of a row
    int ntab_max;         // max Nx+Ny+Nz over the rooms of the launch
    int acc_len;          // 64-bit words of the LDS accumulator
    // basis method: reversed Chebyshev tap coefficients [NB][bcl], K steps per lane group
    // of the Toeplitz product, and the length of one basis row of the tile
    const float* ctab;
    int bks, bdl, bcl;
    RoomParams prm;
};

struct FinArgs {
    const RoomDev* rooms;
    const int* rir_room;
    const float* pos_src;
    const float* pos_rcv;
    float* out;
    float w_sz;           // float(10e-3f * fs), src/gpuRIR_cuda.cu:421
    double fs;
    unsigned long long seed;
    unsigned call_offset;
    int noise;
    int crossfade;
    RoomParams prm;
};

// Integer power by squaring (e >= 0); 0^0 = 1 as powf(0,0) in the reference.
// Correctly rounded square root of a positive normal double: the v_rsq_f64 estimate and two
// Goldschmidt/Newton refinements with an exact residual, the sequence the compiler emits for
// sqrt() without its range scaling (for inputs below 2^-767) and its zero/infinity/NaN
// selects.  Image distances squared lie far inside that range, so the result equals sqrt().
__device__ inline double sqrt_normal(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    const double r = fma(-h, g, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    double d = fma(-g, g, x);
    g = fma(d, h, g);
    d = fma(-g, g, x);
    return fma(d, h, g);
}

// 1 / x for a positive normal double: the hardware estimate and two Newton steps (error
// below 2 ulp)
__device__ inline double rcp_normal(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// Square root of a positive normal double for the basis kernels' image delays: the v_rsq_f64
// estimate (relative error below 2^-22) and one Goldschmidt step, which squares that error:
// about 44 correct bits, so a delay of up to 2^13 samples is off by less than 1e-9 sample,
// far below the tap expansion's 1e-6 (sqrt_normal takes two more residual corrections).
__device__ inline double sqrt_delay(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    const double g = x * y, h = 0.5 * y;
    const double r = fma(-h, g, 0.5);
    return fma(g, r, g);
}

// sqrt_delay and the reciprocal square root from the same Goldschmidt step (both ~44 bits)
__device__ inline double sqrt_delay_inv(double x, double& inv) {
    const double y = __builtin_amdgcn_rsq(x);
    const double g = x * y, h = 0.5 * y;
    const double r = fma(-h, g, 0.5);
    inv = 2.0 * fma(h, r, h);
    return fma(g, r, g);
}

__device__ inline double ipow(double b, int e) {
    double r = 1.0;
    while (e > 0) {
        if (e & 1) r *= b;
        b *= b;
        e >>= 1;
    }
    return r;
}

// src/gpuRIR_cuda.cu:139-146 (cosTheta already computed)
__device__ inline float pattern_gain(float c, int p) {
    switch (p) {
        case 1: return c > 0.0f ? 1.0f : 0.0f;
        case 2: return 0.5f + 0.5f * c;
        case 3: return 0.25f + 0.75f * c;
        case 4: return 0.75f + 0.25f * c;
        case 5: return c;
        default: return 0.0f;
    }
}

// Taps are accumulated as 64-bit fixed point with integer LDS atomics
// (ds_add_u64): on gfx950 these run at the rate of plain LDS stores, while
// ds_add_f32 is an order of magnitude slower.  A tap already carries the
// workgroup's power-of-two scale and is rounded to an int32; integer sums are
// exact and independent of the order of the atomics.
__device__ inline void lds_add(long long* p, float v) {
    int q;
    asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(q) : "v"(v));  // floor(v + 0.5) in one instruction
    __hip_atomic_fetch_add(reinterpret_cast<unsigned long long*>(p), (unsigned long long)(long long)q,
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Two fixed-point sums in one 64-bit word: hv in the high 64 - S bits, lv in the low S
// bits.  The word is the two's-complement sum of (qh << S) + ql over the adds, so as long
// as every low sum stays inside (-2^(S-1), 2^(S-1)) and every high sum inside its field,
// both are recovered exactly (unpack_pair): the low field as the sign-extended low S bits,
// the high field as the rest shifted down.  One ds_add_u64 carries two basis rows.
template <int S>
__device__ inline void lds_add_pair(unsigned long long* p, float hv, float lv) {
    int qh, ql;
    asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(qh) : "v"(hv));
    asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(ql) : "v"(lv));
    const unsigned long long w = ((unsigned long long)(long long)qh << S) + (unsigned long long)(long long)ql;
    __hip_atomic_fetch_add(p, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <int K>
__device__ constexpr float exp2_neg() {
    static_assert(K >= 0 && K < 64, "power of two");
    return 1.0f / (float)(1ULL << K);
}

// hi is returned as the high field times 2^S (the word minus its sign-extended low field).
template <int S>
__device__ inline void unpack_pair(unsigned long long w, long long& hi, long long& lo) {
    lo = (long long)((int)((unsigned)w << (32 - S)) >> (32 - S));
    hi = (long long)w - lo;
}

// A 64-bit fixed-point value times 2^-E in FP32 (inv = 2^-E): the two 32-bit halves
// converted separately, within 2 ulp of the correctly rounded value.
__device__ inline float fix64_to_f32(long long v, float inv) {
    return fmaf((float)(int)(v >> 32), inv * 4294967296.0f, (float)(unsigned)v * inv);
}

// Mixed-precision mode: 32-bit fixed point (ds_add_u32, half the LDS traffic of the
// 64-bit sums); the workgroup moves these sums into an FP32 accumulator before
// they can overflow.
__device__ inline void lds_add32(int* p, float v) {
    int q;
    asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(q) : "v"(v));
    __hip_atomic_fetch_add(reinterpret_cast<unsigned*>(p), (unsigned)q, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
}

// sin(pi f) for f in [0, 1]: odd Taylor polynomial of sin(pi g), g = min(f, 1 - f) <= 1/2,
// through g^13 (truncation below 2e-9 relative).
__device__ inline float sinpi01(float f) {
    constexpr double P = 3.14159265358979323846;
    constexpr float c1 = (float)P, c3 = (float)(-P * P * P / 6.0), c5 = (float)(P * P * P * P * P / 120.0),
                    c7 = (float)(-P * P * P * P * P * P * P / 5040.0),
                    c9 = (float)(P * P * P * P * P * P * P * P * P / 362880.0),
                    c11 = (float)(-P * P * P * P * P * P * P * P * P * P * P / 39916800.0),
                    c13 = (float)(P * P * P * P * P * P * P * P * P * P * P * P * P / 6227020800.0);
    const float g = fminf(f, 1.0f - f);
    const float u = g * g;
    float q = fmaf(u, c13, c11);
    q = fmaf(u, q, c9);
    q = fmaf(u, q, c7);
    q = fmaf(u, q, c5);
    q = fmaf(u, q, c3);
    q = fmaf(u, q, c1);
    return g * q;
}

// cos(pi h), sin(pi h) for 0 <= h <= 1/16 (Taylor through the 7th power; error < 1e-10).
__device__ inline void sincospi_small(float h, float* sn, float* cs) {
    constexpr double P = 3.14159265358979323846;
    const float a = (float)P * h;
    const float u = a * a;
    *cs = fmaf(u, fmaf(u, fmaf(u, -1.0f / 720.0f, 1.0f / 24.0f), -0.5f), 1.0f);
    *sn = a * fmaf(u, fmaf(u, fmaf(u, -1.0f / 5040.0f, 1.0f / 120.0f), -1.0f / 6.0f), 1.0f);
}

// Room constants in scalar form.  Kernels never keep a RoomDev copy: any runtime
// index into such a copy (e.g. by axis) makes the compiler place it in scratch.
struct RoomK {
    double L0, L1, L2, fs_over_c, inv_nyz, inv_nz;
    float b0, b1, b2, b3, b4, b5, level;
    int Nx, Ny, Nz, n_rcv, src_off, rcv_off, nISM, nSamples, n_chunks, rir_off;
    long long out_off, part_off;
    unsigned room_id;
};

__device__ inline RoomK load_room(const RoomDev* __restrict__ rooms, int k, const RoomParams& prm) {
    const RoomDev* __restrict__ p = rooms + k;
    RoomK r;
    r.L0 = p->L[0];
    r.L1 = p->L[1];
    r.L2 = p->L[2];
    r.fs_over_c = p->fs_over_c;
    r.inv_nyz = p->inv_nyz;
    r.inv_nz = p->inv_nz;
    r.b0 = p->beta[0];
    r.b1 = p->beta[1];
    r.b2 = p->beta[2];
    r.b3 = p->beta[3];
    r.b4 = p->beta[4];
    r.b5 = p->beta[5];
    r.level = p->level;
    r.Nx = p->nb[0];
    r.Ny = p->nb[1];
    r.Nz = p->nb[2];
    r.n_rcv = p->n_rcv;
    r.src_off = p->src_off;
    r.rcv_off = p->rcv_off;
    r.nISM = p->nISM;
    r.nSamples = p->nSamples;
    r.n_chunks = p->n_chunks;
    r.rir_off = p->rir_off;
    r.out_off = p->out_off;
    r.part_off = p->part_off;
    r.room_id = p->room_id;
    if (prm.beta) {
        const float* b = prm.beta + 6 * k;
        r.b0 = b[0];
        r.b1 = b[1];
        r.b2 = b[2];
        r.b3 = b[3];
        r.b4 = b[4];
        r.b5 = b[5];
    }
    if (prm.level) r.level = prm.level[k];
    return r;
}

// Per-axis tables of one RIR: offs[e] = 2 m L + (1-2q) src - rcv (FP64, exact for
// float inputs) and att[e] = (-b_lo)^|m-q| (-b_hi)^|m| (src/gpuRIR_cuda.cu:336-340,353).
// Axis 0 occupies [0, Nx), axis 1 [Nx, Nx+Ny), axis 2 the rest.
__device__ inline void axis_entry(int n, int N, double L, float sd, float rd, float blo, float bhi, double* offs,
                                  float* att) {
    const int k = n - N / 2;
    const int m = (k + 1) >> 1;  // ceil(k / 2)
    const int q = k & 1;         // |k % 2|
    const double sgn = (double)(1 - 2 * q);
    const double av = 2.0 * (double)m * L;
    const double bv = sgn * (double)sd;
    offs[n] = (av + bv) - (double)rd;
    att[n] = (float)(ipow(-(double)blo, abs(m - q)) * ipow(-(double)bhi, abs(m)));
}

// One loop per axis: every field is read with a constant index (a runtime axis
// index into the room constants would place them in scratch memory).
__device__ inline void build_axis_tables(const RoomK& R, const float* ps, const float* pr, double* offs, float* att,
                                         int tid, int nthreads) {
    for (int n = tid; n < R.Nx; n += nthreads) axis_entry(n, R.Nx, R.L0, ps[0], pr[0], R.b0, R.b1, offs, att);
    for (int n = tid; n < R.Ny; n += nthreads)
        axis_entry(n, R.Ny, R.L1, ps[1], pr[1], R.b2, R.b3, offs + R.Nx, att + R.Nx);
    for (int n = tid; n < R.Nz; n += nthreads)
        axis_entry(n, R.Nz, R.L2, ps[2], pr[2], R.b4, R.b5, offs + R.Nx + R.Ny, att + R.Nx + R.Ny);
}

struct ImageVal {
    double tau;
    float amp;
};

// Polar patterns of the receiver and the source on an image's amplitude
// (src/gpuRIR_cuda.cu:355-368): (nx, ny, nz) is the image, (vx, vy, vz) its vector
// from the receiver, os / orr the orientations with their inverse norms in [3].
__device__ inline float apply_patterns(const RoomK& R, float amp, double vx, double vy, double vz, float distf,
                                       int nx, int ny, int nz, int spkr, int mic, const float os[4],
                                       const float orr[4]) {
    const float inv_d = __builtin_amdgcn_rcpf(distf);
    const float fx = (float)vx, fy = (float)vy, fz = (float)vz;
    if (mic != 0) {
        // doa = rcv -> image, receiver orientation as given
        const float c = (fx * orr[0] + fy * orr[1] + fz * orr[2]) * inv_d * orr[3];
        amp *= pattern_gain(c, mic);
    }
    if (spkr != 0) {
        // doa = image -> rcv, source orientation mirrored per reflected axis
        const int qx = (nx - R.Nx / 2) & 1, qy = (ny - R.Ny / 2) & 1, qz = (nz - R.Nz / 2) & 1;
        const float ox = qx ? -os[0] : os[0], oy = qy ? -os[1] : os[1], oz = qz ? -os[2] : os[2];
        const float c = -(fx * ox + fy * oy + fz * oz) * inv_d * os[3];
        amp *= pattern_gain(c, spkr);
    }
    return amp;
}

// Stage (a) for image i of one RIR (src/gpuRIR_cuda.cu:336-369).
__device__ inline ImageVal image_eval(const RoomK& R, int i, const double* offs, const float* att, int spkr,
                                      int mic, const float os[4], const float orr[4]) {
    const int Nx = R.Nx, Ny = R.Ny, Nz = R.Nz;
    const int nyz = Ny * Nz;
    const int nx = (int)(((double)i + 0.5) * R.inv_nyz);
    const int rem = i - nx * nyz;
    const int ny = (int)(((double)rem + 0.5) * R.inv_nz);
    const int nz = rem - ny * Nz;
    const double vx = offs[nx], vy = offs[Nx + ny], vz = offs[Nx + Ny + nz];
    double d2 = vx * vx;
    d2 = d2 + vy * vy;
    d2 = d2 + vz * vz;
    const double dist = sqrt_normal(d2);  // d2 >= (direct-path distance)^2 > 0
    ImageVal r;
    r.tau = dist * R.fs_over_c;
    const float distf = (float)dist;
    float amp = att[nx] * att[Nx + ny] * att[Nx + Ny + nz];
    amp *= __builtin_amdgcn_rcpf(12.566370614359172f * distf);
    if (spkr != 0 || mic != 0) amp = apply_patterns(R, amp, vx, vy, vz, distf, nx, ny, nz, spkr, mic, os, orr);
    r.amp = amp;
    return r;
}

__device__ inline void load_orient(const float* o, float out[4]) {
    out[0] = o[0];
    out[1] = o[1];
    out[2] = o[2];
    out[3] = __builtin_amdgcn_rsqf(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]);
}

// Stage (b) layout: a wave walks the queue one image at a time; lane l of pass p
// owns tap offset j = J0 + 64 p + l, i.e. sample t = floor(tau) + j.  The 64 lanes of
// an LDS atomic hit 64 consecutive words (no bank conflicts), x = j - f is a single
// rounding of an exact integer minus the fraction, and the per-offset window
// constants stay in registers for the whole kernel:
//   w_j(f) B = B/2 + (cos 2pi j/Tw)(B/2 cos 2pi f/Tw) + (sin 2pi j/Tw)(B/2 sin 2pi f/Tw)
//   tap = sigma_j w_j(f) B / (j - f) = w_j(f) B / x',   x' = sigma_j (j - f) (one FMA),
//   B = amp sin(pi f)/pi,   sigma_j = (-1)^(j+1).
// LUT mode: u = 16 x + center = 16 j + center + floor(-16 f) + frac(-16 f), so all
// lanes share the interpolation weight and read consecutive entries of one phase
// row of the table.
// All taps of one queued image; PMAX is the exact number of 64-lane passes (entry c = (acc index bits, f, B/2 | amp, B/2 cos), s = B/2 sin).
// Mixed precision (MODE == PREC_FP16): passes are paired and both taps of a pair are
// computed with packed FP16 instructions (offset FMA, window FMAs, product; the
// reciprocals are FP16 too), the per-image amplitude B (with the fixed-point scale)
// is applied in FP32 when each tap is converted, and the sum is the same 64-bit
// fixed point as the FP32 mode.  Entry c = (acc index bits, B, (f, f), (cos/2, sin/2))
// with the FP16 cos/sin of 2 pi f / Tw (s is not used), so a queue entry is one 16-byte LDS read.
template <int MODE, int PMAX, bool MASK>
__device__ __forceinline__ void taps_image(const AccArgs& a, long long* acc, const float2* lut, int lane, float4 c,
                                           float s, const float* jf, const float* cj, const float* sj,
                                           const float* hj, const half2_t* J2, const half2_t* H2, const half2_t* C2,
                                           const half2_t* S2) {
    long long* ap = acc + __float_as_int(c.x) + lane;
    const float f = c.y;
    if (MODE == PREC_FP16) {
        // the whole entry is one 16-byte read: (acc index bits, B, (f, f), (cos/2, sin/2))
        const half2_t f2 = bits_h2(c.z), cs = bits_h2(c.w);
        const half2_t hc = __builtin_shufflevector(cs, cs, 0, 0), hs = __builtin_shufflevector(cs, cs, 1, 1);
        const float B = c.y;
        const half2_t half = {(_Float16)0.5f, (_Float16)0.5f};
#pragma unroll
        for (int pp = 0; pp < (PMAX + 1) / 2; pp++) {
            const half2_t xs = __builtin_elementwise_fma(-H2[pp], f2, J2[pp]);
            const half2_t r = {(_Float16)__builtin_amdgcn_rcph(xs.x), (_Float16)__builtin_amdgcn_rcph(xs.y)};
            const half2_t w = __builtin_elementwise_fma(C2[pp], hc, __builtin_elementwise_fma(S2[pp], hs, half));
            const half2_t v = w * r;
            // widening multiply by B in one v_fma_mix_f32 per tap
            float v0 = fmaf((float)v.x, B, 0.0f), v1 = fmaf((float)v.y, B, 0.0f);
            if (MASK) {
                v0 = fabsf((float)xs.x) < a.halfTw ? v0 : 0.0f;
                v1 = fabsf((float)xs.y) < a.halfTw ? v1 : 0.0f;
            }
            int* ap32 = reinterpret_cast<int*>(acc) + __float_as_int(c.x) + lane;
            lds_add32(ap32 + 64 * (2 * pp), v0);
            if (2 * pp + 1 < PMAX) lds_add32(ap32 + 64 * (2 * pp + 1), v1);
        }
    } else if (MODE == PREC_LUT) {
        const float amp = c.z;
        const float m16 = -16.0f * f;
        const float k0 = floorf(m16);
        const float al = m16 - k0;
        const int i0 = a.lut_center + (int)k0;  // >= 0
        const int ph = i0 & 15;
        const float2* row = lut + ph * a.lut_row + ((i0 >> 4) - a.lut_k_lo) + a.J0 + lane;
#pragma unroll
        for (int p = 0; p < PMAX; p++) {
            const float2 t = row[64 * p];
            float v = amp * fmaf(al, t.y, t.x);
            if (MASK) v = (fabsf(jf[p] - f) < a.halfTw) ? v : 0.0f;
            lds_add(ap + 64 * p, v);
        }
    } else {
        // sigma_j folded into the offset: x' = sigma_j (j - f), tap = B w_j / x'
        const float H = c.z, BC = c.w, BS = s;
#pragma unroll
        for (int p = 0; p < PMAX; p++) {
            const float xs = fmaf(-hj[p], f, jf[p]);
            const float r = __builtin_amdgcn_rcpf(xs);
            const float w = fmaf(cj[p], BC, fmaf(sj[p], BS, H));
            float v = w * r;
            if (MASK) v = (fabsf(xs) < a.halfTw) ? v : 0.0f;
            lds_add(ap + 64 * p, v);
        }
    }
}

// One work item: stages (a) and (b) for an image chunk of one RIR over one time tile.
// GRP: windows of more than 12 passes (fs above ~96 kHz) in groups of PMAX = 12 passes: the
// per-lane constants of a group are loaded before the queue is walked for that group, and the
// table of the LUT mode (too large for LDS there) is read from global memory.
template <int MODE, int PMAX, bool MASK, bool GRP>
__device__ __forceinline__ void accumulate_item(const AccArgs& a, int item, unsigned char* smem_raw) {
    const Item it = a.items[item];
    const RoomK R = load_room(a.rooms, it.room, a.prm);
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);

    double* offs = reinterpret_cast<double*>(smem_raw);
    long long* acc = reinterpret_cast<long long*>(offs + a.ntab_max);  // [acc_len] fixed point
    float* att = reinterpret_cast<float*>(acc + a.acc_len);
    // queue of live images: q4[e] = (acc index bits, f, B/2 or amp, B/2 cos), qS[e] = B/2 sin
    float4* q4 = reinterpret_cast<float4*>(att + a.ntab_max);
    float* qS = reinterpret_cast<float*>(q4 + QCAP + QPAD);
    // [0] queue length, [1..3] per-axis max |att| bits, [4..6] per-axis min |offset| bits
    int* qcount = reinterpret_cast<int*>(qS + QCAP + QPAD);
    float2* lut = reinterpret_cast<float2*>(qcount + 8);  // LUT mode: [16 * lut_row] polyphase table

    const int s = it.rir / R.n_rcv, r = it.rir - s * R.n_rcv;
    const float* ps = a.pos_src + 3 * (R.src_off + s);
    const float* pr = a.pos_rcv + 3 * (R.rcv_off + r);
    if (tid < 8) qcount[tid] = tid < 4 ? 0 : 0x7f800000;  // maxima from 0, minima from +inf
    __syncthreads();
    build_axis_tables(R, ps, pr, offs, att, tid, WG);
    for (int e = tid; e < a.acc_len; e += WG) acc[e] = 0;
    if (MODE == PREC_LUT && !GRP)
        for (int e = tid; e < 16 * a.lut_row; e += WG) lut[e] = a.lut[e];
    const float2* tab = GRP ? a.lut : lut;
    const int n_groups = GRP ? a.n_pass / PMAX : 1;
    // per-lane constants of every pass (of group g)
    float jf[PMAX], cj[PMAX], sj[PMAX], hj[PMAX];
    // FP16 mode: the same constants, paired by pass (a lone last pass is paired with itself)
    half2_t J2[(PMAX + 1) / 2], H2[(PMAX + 1) / 2], C2[(PMAX + 1) / 2], S2[(PMAX + 1) / 2];
    auto load_group = [&](int g) {
#pragma unroll
        for (int p = 0; p < PMAX; p++) {
            const float4 c = a.jtab[64 * (PMAX * g + p) + lane];  // the table holds n_groups PMAX passes
            jf[p] = c.x;
            cj[p] = c.y;
            sj[p] = c.z;
            hj[p] = c.w;
        }
        if (MODE == PREC_FP16) {
#pragma unroll
            for (int pp = 0; pp < (PMAX + 1) / 2; pp++) {
                const int lo = 2 * pp, hi = 2 * pp + 1 < PMAX ? 2 * pp + 1 : 2 * pp;
                J2[pp] = h2(jf[lo], jf[hi]);
                H2[pp] = h2(hj[lo], hj[hi]);
                C2[pp] = h2(cj[lo], cj[hi]);
                S2[pp] = h2(sj[lo], sj[hi]);
            }
        }
    };
    load_group(0);
    float os[4] = {0, 0, 0, 0}, orr[4] = {0, 0, 0, 0};
    if (a.spkr != 0) load_orient(a.orv_src + 3 * (R.src_off + s), os);
    if (a.mic != 0) load_orient(a.orv_rcv + 3 * (R.rcv_off + r), orr);

    // fixed-point scale: |tap| <= max|att| / (4 pi d_min), d_min^2 = sum over the axes of the
    // smallest squared offset (no image is closer; directivity gains and the windowed sinc are
    // bounded by 1).  With the source and receiver inside the room d_min is the direct path.
    __syncthreads();  // table entries are written by other threads than the ones reading them here
    for (int e = tid; e < R.Nx + R.Ny + R.Nz; e += WG) {
        const int d = (e < R.Nx) ? 0 : (e < R.Nx + R.Ny ? 1 : 2);
        atomicMax(qcount + 1 + d, __float_as_int(fabsf(att[e])));
        atomicMin(qcount + 4 + d, __float_as_int((float)fabs(offs[e])));
    }
    __syncthreads();
    float scale, inv_scale;
    {
        const float dx = __int_as_float(qcount[4]), dy = __int_as_float(qcount[5]), dz = __int_as_float(qcount[6]);
        const float bound = __int_as_float(qcount[1]) * __int_as_float(qcount[2]) * __int_as_float(qcount[3]) *
                            __builtin_amdgcn_rcpf(12.566370614359172f * sqrtf(dx * dx + dy * dy + dz * dz));
        // |tap| * 2^E < 2^30 (64-bit sums), or < 2^21 in the 32-bit mixed-precision sums,
        // which then take up to FP16_FLUSH = 1023 taps per word (< 2^31 in all) between two
        // moves into FP32.  (The quantum, 2^-20 of the largest tap, is a large part of this
        // mode's error: 2^-16 measured 4e-3 relative L2 on the 128-RIR directional room.)
        const int top = MODE == PREC_FP16 ? 20 : 29;
        int E = (bound > 0.0f && bound < 3.0e38f) ? top - ilogbf(bound) : 0;
        E = min(max(E, -120), 120);
        scale = ldexpf(1.0f, E);
        inv_scale = ldexpf(1.0f, -E);
    }

    const int J1 = a.J0 + a.nJ - 1;
    const int lo_n0 = it.t_lo - J1;        // images with n0 in [lo_n0, hi_n0] touch the tile
    const int hi_n0 = it.t_hi - 1 - a.J0;
    const int acc_shift = J1 - it.t_lo;    // acc index of tap j = n0 + acc_shift + (j - J0)

    int since = 0;  // images queued since the last move of the 32-bit sums (mixed precision)
    for (int r0 = it.i0; r0 < it.i1; r0 += QCAP) {
        if (tid == 0) *qcount = 0;
        __syncthreads();
        // ---- phase A: stage (a) for up to QCAP images, compaction of the live ones
        for (int sub = 0; sub < QCAP / WG; sub++) {
            const int i = r0 + sub * WG + tid;
            bool live = false;
            int n0 = 0;
            float f = 0.0f, amp = 0.0f;
            if (i < it.i1) {
                const ImageVal iv = image_eval(R, i, offs, att, a.spkr, a.mic, os, orr);
                const double fl = floor(iv.tau);
                n0 = (int)fl;
                f = (float)(iv.tau - fl);
                amp = iv.amp;
                live = (n0 >= lo_n0) && (n0 <= hi_n0) && (amp != 0.0f);
            }
            const unsigned long long mask = __ballot(live);
            if (mask != 0ull) {
                const int cnt = __popcll(mask);
                const int pre = __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                                          __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
                int base = 0;
                if (lane == 0) base = atomicAdd(qcount, cnt);
                base = __shfl(base, 0);
                if (live) {
                    const int q = base + pre;
                    if (MODE == PREC_LUT) {
                        q4[q] = make_float4(__int_as_float(n0 + acc_shift), f, amp * scale, 0.0f);
                    } else if (MODE == PREC_FP16) {
                        // f in [2^-12, 1 - 2^-11] keeps 1/x' inside the FP16 range at j = 0 and
                        // j = 1 (an f rounded up to 1.0 would make x' = 0); B uses the FP16 f
                        const float fh = (float)(_Float16)fminf(fmaxf(f, 0x1p-12f), 1.0f - 0x1p-11f);
                        const float B = amp * scale * sinpi01(fh) * 0.31830988618379067f;
                        const float ph = fh * a.inv_tw2;
                        float sph, cph;
                        if (a.inv_tw2 <= 0.0625f) {
                            sincospi_small(ph, &sph, &cph);
                        } else {
                            sph = sinpif(ph);
                            cph = cospif(ph);
                        }
                        q4[q] = make_float4(__int_as_float(n0 + acc_shift), B, h2_bits(h2(fh, fh)),
                                            h2_bits(h2(0.5f * cph, 0.5f * sph)));
                    } else {
                        // f == 0 would make the tap at j = 0 a 0 * inf, and so would f == 1 at j = 1
                        // (a fraction within 2^-25 of 1 rounds up to 1.0f); moving tau by less
                        // than 2^-24 samples changes no tap by more than ~1e-7 relative.
                        const float fc = fminf(fmaxf(f, 0x1p-30f), 0x1.fffffep-1f);
                        const float B = amp * scale * sinpi01(fc) * 0.31830988618379067f;
                        const float ph = fc * a.inv_tw2;  // phi / pi = 2 f / Tw
                        float sph, cph;
                        if (a.inv_tw2 <= 0.0625f) {
                            sincospi_small(ph, &sph, &cph);
                        } else {
                            sph = sinpif(ph);
                            cph = cospif(ph);
                        }
                        q4[q] = make_float4(__int_as_float(n0 + acc_shift), fc, 0.5f * B, 0.5f * B * cph);
                        qS[(q & (NWAVE - 1)) * QSW + q / NWAVE] = 0.5f * B * sph;
                    }
                }
            }
        }
        __syncthreads();
        // ---- phase B: every live image scatters its taps into the LDS tile.  A wave
        // takes two images per step (more independent work between LDS waits) and
        // reads the queue entries of its next pair before issuing this pair's atomics,
        // so the in-order LDS pipe never makes those reads wait behind the atomics.
        // The two register sets X and Y alternate (no copies), and the queue is read
        // through per-lane pointers with immediate offsets; reads past the live part
        // stay inside the queue's padding and their values are never used.  The sine
        // terms of four images come in one read from the wave's S row.
        const int nq = __builtin_amdgcn_readfirstlane(*qcount);
        if (MODE == PREC_FP16) {
            // a word takes at most one tap per image: move the 32-bit sums into the FP32
            // accumulator before this round could take them past FP16_FLUSH taps
            if (since + nq > FP16_FLUSH) {
                // four words per thread and step (acc_len is a multiple of 4): 16-byte LDS accesses
                int4* lo = reinterpret_cast<int4*>(acc);
                float4* hi = reinterpret_cast<float4*>(reinterpret_cast<int*>(acc) + a.acc_len);
                for (int e = tid; e < a.acc_len / 4; e += WG) {
                    const int4 l = lo[e];
                    float4 h = hi[e];
                    h.x += (float)l.x * inv_scale;
                    h.y += (float)l.y * inv_scale;
                    h.z += (float)l.z * inv_scale;
                    h.w += (float)l.w * inv_scale;
                    hi[e] = h;
                    lo[e] = make_int4(0, 0, 0, 0);
                }
                __syncthreads();
                since = 0;
            }
            since += nq;
        }
        for (int gi = 0; gi < n_groups; gi++) {
            if (GRP && gi > 0) load_group(gi);
            const int lg = lane + 64 * PMAX * gi;  // the lane's tap offset in pass 0 of group gi
            if (wave < nq) {
                const float4* pq = q4 + (tid >> 6);
                // this wave's S row: group g holds its images 4g .. 4g + 3 (xA, xB, yA, yB)
                const float4* pS4 = reinterpret_cast<const float4*>(qS + wave * QSW);
                float4 xA = pq[0], xB = pq[NWAVE];
                float4 s4 = pS4[0];
                for (int e = wave;;) {
                    const float4 yA = pq[2 * NWAVE], yB = pq[3 * NWAVE];
                    taps_image<MODE, PMAX, MASK>(a, acc, tab, lg, xA, s4.x, jf, cj, sj, hj, J2, H2, C2, S2);
                    if (e + NWAVE < nq) taps_image<MODE, PMAX, MASK>(a, acc, tab, lg, xB, s4.y, jf, cj, sj, hj, J2, H2, C2, S2);
                    e += 2 * NWAVE;
                    if (e >= nq) break;
                    pq += 2 * NWAVE;
                    xA = pq[2 * NWAVE];
                    xB = pq[3 * NWAVE];
                    const float4 s4n = pS4[1];  // group <= 32: inside the row (QSW = 132 floats)
                    taps_image<MODE, PMAX, MASK>(a, acc, tab, lg, yA, s4.z, jf, cj, sj, hj, J2, H2, C2, S2);
                    if (e + NWAVE < nq) taps_image<MODE, PMAX, MASK>(a, acc, tab, lg, yB, s4.w, jf, cj, sj, hj, J2, H2, C2, S2);
                    e += 2 * NWAVE;
                    if (e >= nq) break;
                    pq += 2 * NWAVE;
                    pS4 += 1;
                    s4 = s4n;
                }
            }
        }
        __syncthreads();
    }
    if (it.i0 >= it.i1) __syncthreads();

    // ---- flush the tile [t_lo, t_hi)
    const long long* src = acc + (J1 - a.J0);
    const int nt = it.t_hi - it.t_lo;
    float* dst;
    if (R.n_chunks == 1)
        dst = a.out + R.out_off + (long long)it.rir * R.nSamples + it.t_lo;
    else
        dst = a.part + R.part_off + ((long long)it.rir * R.n_chunks + it.chunk) * R.nISM + it.t_lo;
    if (MODE == PREC_FP16) {
        const int* lo = reinterpret_cast<const int*>(acc) + (J1 - a.J0);
        const float* hi = reinterpret_cast<const float*>(reinterpret_cast<const int*>(acc) + a.acc_len) + (J1 - a.J0);
        for (int t = tid; t < nt; t += WG) dst[t] = hi[t] + (float)lo[t] * inv_scale;
    } else {
        for (int t = tid; t < nt; t += WG) dst[t] = (float)((double)src[t] * (double)inv_scale);
    }
}

// One workgroup per work item.
template <int MODE, int PMAX, bool MASK, bool GRP>
__global__ __launch_bounds__(WG) void ism_accumulate_kernel(AccArgs a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    accumulate_item<MODE, PMAX, MASK, GRP>(a, blockIdx.x, smem_raw);
}

// Persistent form: a grid sized to the resident workgroups takes items from a
// counter in order (longest first).  counters[0] is the next item, counters[1] the
// number of workgroups done; the last one resets both for the next launch (every
// other workgroup has made its final fetch by then), so a captured graph can be
// replayed without any memset node.
template <int MODE, int PMAX, bool MASK, bool GRP>
__global__ __launch_bounds__(WG) void ism_accumulate_persistent_kernel(AccArgs a, int n_items,
                                                                       unsigned* counters) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ int next;
    for (;;) {
        if (threadIdx.x == 0) next = (int)atomicAdd(counters, 1u);
        __syncthreads();
        const int item = next;
        __syncthreads();  // every thread has read next; the item may overwrite all of LDS
        if (item >= n_items) break;
        accumulate_item<MODE, PMAX, MASK, GRP>(a, item, smem_raw);
    }
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(counters + 1, 1u) == gridDim.x - 1) {
            atomicExch(counters, 0u);
            atomicExch(counters + 1, 0u);
        }
    }
}

// ---------------------------------------------------------------------------
// Basis method (FP32 exact sinc, the default for that mode)
// ---------------------------------------------------------------------------
//
// For an image at tau = n0 + f (0 <= f < 1) the tap at offset j = t - n0 is
// h_j(f) = hann(j - f) sinc(j - f), an entire function of f on [0, 1] for every j
// whose window membership does not depend on f (all j for an even Tw; for an odd Tw
// the two end offsets are cut where their taps are below 1e-7).  Its Chebyshev
// interpolant of degree NB - 1 in u = 2 f - 1,
//     h_j(f) ~ sum_m C[m][j] T_m(u)        (within 5e-7 of h for NB = 8),
// turns the scatter of Tw taps per image into NB sums per integer delay,
//     D_m[n] = sum over images with n0 = n of amp T_m(u),
// followed by NB fixed Tw-tap filters, y[t] = sum_m sum_j C[m][j] D_m[t - j], which are
// Toeplitz products: they run on the matrix cores (v_mfma_f32_16x16x4_f32).
//
// A work item is one RIR over a range of BT-sample tiles, one WGB-thread workgroup.
//   * For a tile only images whose n0 can fall in [t_lo - J1, t_hi - 1 - J0] are
//     visited.  Along an axis the image offsets are 2 m L + c_q (q the parity), so the
//     entries inside a distance bound are contiguous runs of m found in closed form:
//     the rows (x, y) with x^2 + y^2 inside the tile's outer radius are one run of y
//     per parity for each x, and a row's z entries in the shell at most two runs per
//     parity.  Rows are taken WGB at a time (one per thread) and their images are
//     placed by a block-wide prefix sum into rounds of QBK queue entries.  A row marks
//     each of its runs in the round's LDS queue (a key at the run's first entry and at
//     every wave segment start it covers); a wave prefix maximum of the keys gives every
//     entry its row and z entry.  A thread takes QBK / WGB consecutive entries: the delays
//     and amplitudes of all of them first (independent FP64 chains), the sums after.  The
//     delays use the same FP64 operations as image_eval (identical n0 and f).
//   * Each live image adds amp scale T_m(u), m < NB, into D_m (64-bit fixed point,
//     ds_add_u64: exact and order-independent), T_m by the three-term recurrence.
//   * D is converted to FP32 in place (rows padded by 4 floats per 16, so that the
//     16-byte operand reads of the 16 columns of a product fall in distinct banks)
//     and the tile is y = sum_m A_m X_m with A_m[r][k] = C[m][J1 - k + r] (Toeplitz)
//     and X_m[k][b] = D_m[16 b + k] (Hankel): 16 samples by 16 blocks per product,
//     BT / 256 products per tile, NB/4 basis rows per wave, and the four
//     partial sums of a sample added in a fixed order.
constexpr int BT = 512;     // samples per tile
constexpr int WGB = 512;    // threads per workgroup
constexpr int NWB = WGB / 64;
constexpr int QB = 2048;    // queue entries per round
constexpr int NOT = BT / 256;  // 256-sample output blocks of a tile (16 x 16 products)
static_assert(NWB == 4 * NOT, "four basis quarters per output block");
constexpr int BASIS_NB = 8; // Chebyshev terms (FP32: within 5e-7 of each tap)
constexpr int BASIS_NB_FP16 = 6;  // mixed precision away from 16 kHz: 6 terms, within 6e-5 (1e-4 relative L2)
constexpr int BCS = 2;      // columns per thread of the in-place conversion: DL <= BCS * WGB
// Packed basis sums: word w < NB/2 of a column holds basis row w in its high 64 - S bits
// and row NB - 1 - w in its low S bits (lds_add_pair).  The low fields take the rows with
// the smallest filters (coefficient L2 norms 0.83, 0.75, 0.23, 5.4e-2, 9.0e-3, 1.4e-3,
// 1.5e-4, 1.8e-5 for rows 0..7), so each field is as wide as its row's filter needs.  Every
// field's sum is bounded per tile by (images in the tile's shell) x (their largest
// amplitude), and the tile's scale keeps that bound plus the rounding of every image
// inside the narrowest field.  RIRs of PACK_MAX_IMAGES images or more use 64-bit rows.
__host__ __device__ constexpr int pack_lo_bits(int nb, int w) {
    return nb == 8 ? (w == 0 ? 24 : w == 1 ? 25 : w == 2 ? 28 : 30) : (w == 0 ? 26 : w == 1 ? 28 : 30);
}
__host__ __device__ constexpr int pack_field_bits(int nb, int m) {
    return m < nb / 2 ? 64 - pack_lo_bits(nb, m) : pack_lo_bits(nb, nb - 1 - m);
}
constexpr long long PACK_MAX_IMAGES = 1LL << 21;
constexpr long long LONG_TILE_MIN_TILES = 2048;  // launches with fewer 2 BT tiles keep BT
// Lookup-table mode on the same machinery: the table's taps are linear in the weight of
// the 1/16-sample grid, so an image at 16 tau = 16 n0 + p + a (phase p < 16, 0 <= a < 1)
// adds amp (1 - a) to phase row p and amp a to row p + 1 (row 0 of column n0 + 1 for
// p = 15) at column n0, and the tile is 16 Toeplitz filters with the table's entries
// T[16 j - p + center] (src/gpuRIR_cuda.cu:236-239, 578-616).  The rows hold 32-bit
// fixed-point sums (top bit 2^24 of the largest image amplitude: 128 images of it per
// word before the sums could wrap), so 16 rows take the LDS of 8 rows of 64-bit sums.
constexpr int LUT_NB = 16;
constexpr int QB_LUT = 1536;  // queue entries per round in the table mode (its rows and
                              // coefficients take more LDS)

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef short short4_t __attribute__((ext_vector_type(4)));

// Inclusive prefix sum over the 64 lanes of a wave with DPP moves (VALU only):
// row_shr 1, 2, 4, 8 within each row of 16 lanes (lanes shifted in from outside the
// row add 0), then row_bcast:15 adds the last lane of rows 0 and 2 to rows 1 and 3,
// and row_bcast:31 the last lane of row 1 to rows 2 and 3.
__device__ inline int wave_scan_incl(int x) {
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);  // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);  // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);  // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);  // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return x;
}

// Largest of non-negative integers (e.g. the bits of non-negative floats) over a wave, in lane
// 63: the prefix maximum by the same DPP moves as wave_scan_incl (lanes shifted in give 0).
__device__ inline int wave_max_nonneg(int x) {
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false));  // row_shr:1
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false));  // row_shr:2
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false));  // row_shr:4
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false));  // row_shr:8
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false));  // row_bcast:15
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false));  // row_bcast:31
    return x;
}

// Largest of non-negative integers over the lanes below this one (0 in lane 0): the prefix
// maximum, moved up one lane by wave_shr:1.
__device__ inline int wave_excl_max_nonneg(int x) {
    x = wave_max_nonneg(x);
    return __builtin_amdgcn_update_dpp(0, x, 0x138, 0xf, 0xf, false);  // wave_shr:1
}

// Exclusive prefix sum of one int per thread over the WGB threads; *total gets the sum.
// Every thread must call it (two barriers); tmp holds NWB ints.
__device__ inline int block_scan_excl(int v, int* tmp, int* total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int x = wave_scan_incl(v);
    if (lane == 63) tmp[w] = x;
    __syncthreads();
    int before = 0, all = 0;
#pragma unroll
    for (int i = 0; i < NWB; i++) {
        const int t = tmp[i];
        before += i < w ? t : 0;
        all += t;
    }
    __syncthreads();
    *total = all;
    return before + x - v;
}

// block_scan_excl with one barrier: the caller alternates between two tmp arrays, so a
// scan's writes never meet the reads of the scan before it (whose successor's barrier lies
// between them).  Every thread must call it; tmp holds NWB ints.
__device__ inline int block_scan_excl1(int v, int* tmp, int* total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int x = wave_scan_incl(v);
    if (lane == 63) tmp[w] = x;
    __syncthreads();
    int before = 0, all = 0;
#pragma unroll
    for (int i = 0; i < NWB; i++) {
        const int t = tmp[i];
        before += i < w ? t : 0;
        all += t;
    }
    *total = all;
    return before + x - v;
}

// sqrt of a non-negative double for the enumeration's distance bounds (widened by 1e-7 of
// 2L, far above sqrt_delay's 2^-44, so values below 1e-200 may read as 0)
__device__ inline double sqrt_bound(double x) { return x > 1e-200 ? sqrt_delay(x) : 0.0; }

// Image entries of one axis in closed form: entry n has k = n - N/2 and offset
// 2 m L + c_q - (rounding) with m = ceil(k / 2), q = k & 1, c_0 = s - r, c_1 = -s - r
// (axis_entry), increasing in m for each parity.  The entries with
// lo <= |offset| <= hi (lo <= 0: no lower bound) are at most two runs of m per parity.
struct AxisRuns {
    double c[2];
    double inv2L;
    int mlo[2], mhi[2];  // valid m per parity
    int half;            // N / 2
};

__device__ inline int floor_div2(int k) { return k >> 1; }  // arithmetic shift: floor(k / 2)

__device__ inline AxisRuns axis_runs(int N, double L, float sd, float rd) {
    AxisRuns A;
    A.half = N / 2;
    const int kmin = -A.half, kmax = N - 1 - A.half;
    A.c[0] = (double)sd - (double)rd;
    A.c[1] = -(double)sd - (double)rd;
    A.inv2L = 1.0 / (2.0 * L);
    A.mlo[0] = -floor_div2(-kmin);  // k = 2 m
    A.mhi[0] = floor_div2(kmax);
    A.mlo[1] = -floor_div2(-(kmin + 1));  // k = 2 m - 1
    A.mhi[1] = floor_div2(kmax + 1);
    return A;
}

// Entry range (first n, count; n steps by 2) of parity q with offset in [u, v] (u <= v),
// widened by 1e-7 of 2L on each side.
__device__ inline void axis_range(const AxisRuns& A, int q, double u, double v, int& n_first, int& count, int cap_hi) {
    int m0 = (int)ceil((u - A.c[q]) * A.inv2L - 1e-7);
    int m1 = (int)floor((v - A.c[q]) * A.inv2L + 1e-7);
    m0 = max(m0, A.mlo[q]);
    m1 = min(min(m1, A.mhi[q]), cap_hi);
    count = max(m1 - m0 + 1, 0);
    n_first = 2 * m0 - q + A.half;
}

// FP32 -> BF16, rounded to nearest even (finite inputs)
__device__ inline unsigned short f32_to_bf16(float x) {
    unsigned u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

// A 16-bit filter input: FP16 (H16 = 1) or BF16 (H16 = 2), as its bits.
template <int H16>
__device__ inline unsigned short half_bits(float x) {
    if (H16 == 2) return f32_to_bf16(x);
    return __builtin_bit_cast(unsigned short, (_Float16)x);
}

// DLC: the basis row length as a compile-time constant (the LDS offsets of an image's
// eight atomics are then immediates), or 0 for the launch's a.bdl.  H16 (mixed precision):
// the filters take 16-bit inputs (D scaled per tile by a power of two, the coefficients
// rounded) with FP32 accumulation: FP16 (1, v_mfma_f32_16x16x16_f16) or BF16 (2,
// v_mfma_f32_16x16x16_bf16).
// PK: the sums of two basis rows share a 64-bit word (pack_lo_bits), with a scale per tile.
// BTK: samples per tile, BT or (FP16 pairs on packed sums at 16 kHz) 2 BT: half the tiles,
// so half the per-tile enumeration, conversion set-up and barriers, and an image's window
// overlaps a second tile half as often.
template <int NB, int DLC, int H16, bool LUT, bool PK, int BTK = BT>
__device__ __forceinline__ void basis_item(const AccArgs& a, int item, unsigned char* smem_raw) {
    static_assert((NB >= 4 && NB <= 8) || (LUT && NB == LUT_NB && (H16 == 0 || (H16 == 3 && DLC == 640))),
                  "basis rows of a wave quarter");
    // SPL: FP32 filters from FP16 pairs (x = hi + lo 2^-11 for the coefficients and the rows,
    // hi lo' + lo hi' accumulated apart): three v_mfma_f32_16x16x16_f16 per 16 K values
    // instead of four v_mfma_f32_16x16x4_f32 per 4 K values
    // H16 = 4 (the FP16 mode at 16 kHz): the same machinery with the rows and coefficients
    // rounded to FP16 and one product per K step (the hi parts only)
    constexpr bool SPL = H16 == 3 || H16 == 4;
    constexpr bool HI1 = H16 == 4;
    static_assert(!SPL || (NB == 8 && !LUT) || (LUT && NB == LUT_NB && DLC == 640),
                  "FP16 pairs carry the 8-term FP32 expansion, or the table mode's 16 phase rows at 16 kHz");
    static_assert(!PK || (!LUT && NB % 2 == 0), "packed sums pair the Chebyshev rows");
    static_assert(BTK == BT || (BTK == 2 * BT && SPL && PK && DLC == 1152), "long tiles: FP16 pairs at 16 kHz");
    constexpr int NOTK = BTK / 256;        // 256-sample output blocks of a tile
    constexpr int UPW = 4 * NOTK / NWB;    // (basis quarter, block) products per wave
    static_assert(UPW * NWB == 4 * NOTK, "whole products per wave");
    // 16-bit coefficient copies: NCP copies, copy rho shifted by rho entries (four: the four
    // operands of a lane start at a multiple of 4 halves, 8-byte reads; two: 4-byte reads)
    // (the table mode's FP16 pairs: one copy, rows of CL + 2 halves, read as three aligned
    // words and a funnel shift: sixteen rows of copies would not leave two workgroups per CU)
    constexpr int NCP = LUT && SPL ? 1 : BTK == BT ? 4 : 2;
    constexpr int CPAD = NCP == 1 ? 2 : NCP;  // halves after each copy's row
    constexpr int QBK = LUT ? QB_LUT : QB;
    constexpr int QEPT = QBK / WGB;   // consecutive queue entries per thread
    constexpr int QSEG = 64 * QEPT;   // entries per wave
    static_assert(QEPT * WGB == QBK, "whole entries per thread");
    const Item it = a.items[item];
    const RoomK R = load_room(a.rooms, it.room, a.prm);
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int ntab = a.ntab_max, DL = DLC ? DLC : a.bdl, CL = a.bcl, KS = a.bks;
    const int DLP = DL + DL / 4;

    double* offs = reinterpret_cast<double*>(smem_raw);  // [ntab] axis offsets
    float* att = reinterpret_cast<float*>(offs + ntab);  // [ntab]
    int* xrun = reinterpret_cast<int*>(att + ntab);      // [3][ntab + 4] per x entry: y runs (n0, count, n1)
    int* rowStart = xrun + 3 * (ntab + 4);               // [ntab + 4] rows before x entry n
    long long* D = reinterpret_cast<long long*>(rowStart + ntab + 4);  // [NB][DL] fixed point
    int* D32 = reinterpret_cast<int*>(D);                              // LUT: [NB][DL] 32-bit sums
    unsigned long long* Dw = reinterpret_cast<unsigned long long*>(D);  // PK: [NB/2][DL] row pairs
    // the rows as FP32 [NB][DLP] after the conversion fit the 64-bit rows; the LUT's 32-bit
    // rows and the packed pairs are given the larger of the two (16-bit rows fit either)
    const int dreg = (LUT || PK) ? max(NB * DL, (H16 == 1 || H16 == 2) ? 0 : NB * DLP) : 2 * NB * DL;  // floats
    const int nwords = PK ? NB / 2 * DL : NB * DL;  // 64-bit (or LUT 32-bit) words of the sums
    float* ctab = reinterpret_cast<float*>(D) + dreg;                 // [NB][CL] (H16: see chs)
    // H16: NCP copies of the coefficients in 16 bits, [NCP][NB][CL + CPAD] (SPL: hi, then lo)
    unsigned short* chs = reinterpret_cast<unsigned short*>(ctab);
    const int crw = H16 ? (SPL ? 2 : 1) * NCP * NB * (CL + CPAD) / 2 : NB * CL;  // floats of the region
    double* rb2 = reinterpret_cast<double*>(ctab + crw);              // [WGB] row x^2 + y^2
    float* ratt = reinterpret_cast<float*>(rb2 + WGB);                // [WGB] row att_x att_y
    int* rxy = reinterpret_cast<int*>(ratt + WGB);                    // [WGB] row nx | ny << 16
    int* queue = rxy + WGB;                                           // [QBK] run marks of a round
    int* misc = queue + QBK;  // [32]: 8..15 att maxima, min offsets, tile max; 16..31 scan scratch
    float* ybuf = reinterpret_cast<float*>(rb2);  // [4][BTK] (aliases the rows and, for 2 BT, the queue)
    static_assert(4 * BTK * 4 <= WGB * (8 + 4 + 4) + QBK * 4, "ybuf fits the row arrays and the queue");
    float* Df = reinterpret_cast<float*>(D);                          // [NB][DLP] (aliases D)
    unsigned short* Dh = reinterpret_cast<unsigned short*>(D);        // H16: [NB][DLP] (aliases D)

    const int s = it.rir / R.n_rcv, r = it.rir - s * R.n_rcv;
    const float* ps = a.pos_src + 3 * (R.src_off + s);
    const float* pr = a.pos_rcv + 3 * (R.rcv_off + r);
    if (tid < 4) misc[8 + tid] = 0;
    if (tid >= 4 && tid < 7) misc[9 + tid] = 0x7f800000;  // misc[13..15]: per-axis min |offset| bits
    build_axis_tables(R, ps, pr, offs, att, tid, WGB);
    for (int e = tid; e < nwords; e += WGB) {
        if (LUT) D32[e] = 0;
        else D[e] = 0;
    }
    for (int e = tid; e < QBK; e += WGB) queue[e] = 0;  // run marks: zero between rounds
    if (H16) {
        // the plan holds the copies' LDS image (make_half_copies): 16-byte copies
        const uint4* src = reinterpret_cast<const uint4*>(a.ctab);
        uint4* dst = reinterpret_cast<uint4*>(chs);
        for (int e = tid; e < crw / 4; e += WGB) dst[e] = src[e];
        if (tid == 0) misc[12] = 0;
    } else {
        for (int e = tid; e < NB * CL; e += WGB) ctab[e] = a.ctab[e];
    }
    float os[4] = {0, 0, 0, 0}, orr[4] = {0, 0, 0, 0};
    if (a.spkr != 0) load_orient(a.orv_src + 3 * (R.src_off + s), os);
    if (a.mic != 0) load_orient(a.orv_rcv + 3 * (R.rcv_off + r), orr);
    __syncthreads();

    const int Nx = R.Nx, Ny = R.Ny, Nz = R.Nz, ntot = Nx + Ny + Nz;
    for (int e = tid; e < ntot; e += WGB) {
        const int ax = e < Nx ? 0 : (e < Nx + Ny ? 1 : 2);
        atomicMax(misc + 9 + ax, __float_as_int(fabsf(att[e])));
        atomicMin(misc + 13 + ax, __float_as_int((float)fabs(offs[e])));
    }
    const AxisRuns AY = axis_runs(Ny, R.L1, ps[1], pr[1]);
    const AxisRuns AZ = axis_runs(Nz, R.L2, ps[2], pr[2]);
    __syncthreads();
    // |amp T_m| <= max|att| / (4 pi d) for an image at distance d >= d_min, d_min from the
    // smallest offset per axis (the direct path when the source and receiver are inside
    // the room)
    const float att_max = __int_as_float(misc[9]) * __int_as_float(misc[10]) * __int_as_float(misc[11]);
    float dmin;
    {
        const float dx = __int_as_float(misc[13]), dy = __int_as_float(misc[14]), dz = __int_as_float(misc[15]);
        dmin = sqrtf(dx * dx + dy * dy + dz * dz);
    }
    float scale, inv_scale;
    {
        // one scale per item: < 2^30 after scaling, 64-bit sums; the LUT's 32-bit sums have
        // room for 128 images of the bound.  Packed sums take a scale per tile (below).
        const float bound = att_max * __builtin_amdgcn_rcpf(12.566370614359172f * dmin);
        const int top = LUT ? 24 : 29;
        int E = (bound > 0.0f && bound < 3.0e38f) ? top - ilogbf(bound) : 0;
        E = min(max(E, -90), 120);  // 2^32 / scale stays finite in the conversion
        scale = ldexpf(1.0f, E);
        inv_scale = ldexpf(1.0f, -E);
    }
    const double fsc = R.fs_over_c;
    const int J0 = a.J0, J1 = a.J0 + a.nJ - 1;
    const bool pattern = a.spkr != 0 || a.mic != 0;
    float* orow = a.out + R.out_off + (long long)it.rir * R.nSamples;
    constexpr int BIG = 1 << 30;
    int scan_par = 0;  // block_scan_excl1 buffer parity (uniform)

    for (int tl = it.t_lo; tl < it.t_hi; tl += BTK) {
        const int th = min(tl + BTK, it.t_hi);
        const int lo_n0 = tl - J1, hi_n0 = th - 1 - J0;
        // distance shell of those delays, one sample wider on each side (the exact
        // n0 test below decides)
        const double rlo = (double)(lo_n0 - 1) / fsc, rhi = (double)(hi_n0 + 2) / fsc;
        const double d2lo = lo_n0 >= 1 ? rlo * rlo : -1.0, d2hi = rhi * rhi;
        // rows: per x entry, the y entries with x^2 + y^2 <= d2hi (one run per parity)
        int carry = 0;
        for (int ix0 = 0; ix0 < Nx; ix0 += WGB) {
            const int nx = ix0 + tid;
            int c = 0;
            if (nx < Nx) {
                const double X = offs[nx];
                const double rem = d2hi - X * X;
                int f0 = 0, c0 = 0, f1 = 0, c1 = 0;
                if (rem >= 0.0) {
                    const double Y = sqrt_bound(rem);
                    axis_range(AY, 0, -Y, Y, f0, c0, BIG);
                    axis_range(AY, 1, -Y, Y, f1, c1, BIG);
                }
                xrun[nx] = f0;
                xrun[(ntab + 4) + nx] = c0;
                xrun[2 * (ntab + 4) + nx] = f1;
                c = c0 + c1;
            }
            int tot;
            const int ex = block_scan_excl1(c, misc + 16 + 8 * (scan_par++ & 1), &tot);
            if (nx < Nx) rowStart[nx + 1] = carry + ex + c;
            carry += tot;
        }
        if (tid == 0) rowStart[0] = 0;
        __syncthreads();
        const int nrows = carry;
        // row `row` of the tile: its x and y entries, x^2 + y^2, and its z runs (first
        // entry, count) in the distance shell: the image count of the row
        auto row_runs = [&](int row, int* zf, int* zc, double& b2, int& nx, int& ny) -> int {
            int lo = 0, hi = Nx;  // rowStart[lo] <= row < rowStart[hi]
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (rowStart[mid] <= row) lo = mid;
                else hi = mid;
            }
            nx = lo;
            const int j = row - rowStart[lo];
            const int c0 = xrun[(ntab + 4) + nx];
            ny = j < c0 ? xrun[nx] + 2 * j : xrun[2 * (ntab + 4) + nx] + 2 * (j - c0);
            const double vx = offs[nx], vy = offs[Nx + ny];
            b2 = vx * vx;
            b2 = b2 + vy * vy;
            // z entries with d2lo - b2 <= z^2 <= d2hi - b2
            const double V2 = d2hi - b2, U2 = d2lo - b2;
            if (V2 >= 0.0) {
                const double V = sqrt_bound(V2);
                if (U2 <= 0.0) {
                    axis_range(AZ, 0, -V, V, zf[0], zc[0], BIG);
                    axis_range(AZ, 1, -V, V, zf[2], zc[2], BIG);
                } else {
                    const double U = sqrt_bound(U2);
                    // positive run first, then the negative run capped below it
                    // (disjoint even where the widened bounds meet near 0)
                    axis_range(AZ, 0, U, V, zf[1], zc[1], BIG);
                    axis_range(AZ, 0, -V, -U, zf[0], zc[0], (zf[1] - AZ.half) / 2 - 1);
                    axis_range(AZ, 1, U, V, zf[3], zc[3], BIG);
                    axis_range(AZ, 1, -V, -U, zf[2], zc[2], (zf[3] - AZ.half + 1) / 2 - 1);
                }
            }
            return zc[0] + zc[1] + zc[2] + zc[3];
        };
        // PK: the image count bounds every field's sum.  The scale is set from the images of
        // the first row round; a later round whose images make the bound fall by k bits first
        // halves the sums so far k times (each field rounded to nearest), so the sums end at
        // the scale of the tile's whole count without a separate counting pass.
        int seen = 0, E_cur = 0;
        float pscale = 1.0f;  // PK: the scale of row 0 (row m: pscale 2^-(F_0 - F_m))
        for (int r0 = 0; r0 < nrows; r0 += WGB) {
            const int row = r0 + tid;
            int cnt = 0;
            int zf[4] = {0, 0, 0, 0}, zc[4] = {0, 0, 0, 0};
            if (row < nrows) {
                double b2;
                int nx, ny;
                cnt = row_runs(row, zf, zc, b2, nx, ny);
                rb2[tid] = b2;
                ratt[tid] = att[nx] * att[Nx + ny];
                rxy[tid] = nx | (ny << 16);
            }
            int tot;
            const int ex = block_scan_excl1(cnt, misc + 16 + 8 * (scan_par++ & 1), &tot);
            if (PK) {
                const int tile_images = seen + tot;
                seen = tile_images;
                // every image of the tile lies at d >= max(d_min, rlo): |amp| <= amax, and a
                // field's sum is at most tile_images (amax scale + 1/2).  The scale of the
                // narrowest field F_min keeps that below 2^(F_min - 1); a wider field's scale
                // is larger by 2^(F - F_min); row 0 (widest) also stays below 2^30 per image
                // for the int32 conversion.  The factor 1 + 2^-10 covers the FP32 rounding of
                // the amplitudes, the recurrence and the bound itself; one spare unit per row
                // round, the rounding of a halving.
                constexpr int F0 = pack_field_bits(NB, 0), FMIN = pack_lo_bits(NB, 0);
                const float rt = fmaxf(dmin, (float)rlo);
                const float amax = att_max * __builtin_amdgcn_rcpf(12.566370614359172f * rt) * (1.0f + 0x1p-10f);
                const float tb = (float)tile_images * amax * (1.0f + 0x1p-10f);
                const float room_f = 0x1p-1f * ldexpf(1.0f, FMIN) - (float)tile_images - (float)(r0 / WGB) - 2.0f;
                int E = 0;
                if (tile_images > 0 && room_f > 0.0f && tb > 0.0f && amax < 3.0e38f) {
                    const float q = room_f / tb;
                    E = q < 3.0e38f ? ilogbf(q) + (F0 - FMIN) : 120;
                    E = min(E, 29 - ilogbf(amax));
                }
                E = min(max(E, -90), 120);
                if (r0 == 0) {
                    E_cur = E;
                } else if (E < E_cur) {
                    // the rounds so far were added at 2^(E_cur - E) times this scale (the
                    // previous queue round ended at a barrier; the next one's marks barrier
                    // orders these writes before its adds)
                    const int k = min(E_cur - E, 62);
                    const long long half = 1LL << (k - 1);
                    for (int e = tid; e < NB / 2 * DL; e += WGB) {
                        const int w = e / DL;
                        const int S = w == 0 ? pack_lo_bits(NB, 0) : w == 1 ? pack_lo_bits(NB, 1)
                                    : w == 2 ? pack_lo_bits(NB, 2) : pack_lo_bits(NB, 3);
                        const unsigned long long v = Dw[e];
                        const long long lo = (long long)__builtin_amdgcn_sbfe((int)(unsigned)v, 0, S);
                        const long long qh = ((long long)v - lo) >> S;
                        Dw[e] = (unsigned long long)((((qh + half) >> k) << S) + ((lo + half) >> k));
                    }
                    E_cur = E;
                }
                pscale = ldexpf(1.0f, E_cur);
                inv_scale = ldexpf(1.0f, -E_cur);
            }
            for (int q0 = 0; q0 < tot; q0 += QBK) {
                // Entry p of the round (run g of row slot, run start s) has nz = base + 2 p with
                // base = zf[g] - 2 (s - q0) in [-4094, 4095].  Each run writes the key
                // (slot, g, base + 4096) at its first entry in the round and at every wave
                // segment start it covers; keys grow with the position, so an entry's key is
                // the largest one at or before it.
                {
                    int s = ex;
#pragma unroll
                    for (int g = 0; g < 4; g++) {
                        const int a0 = max(s, q0) - q0, b0 = min(s + zc[g], q0 + QBK) - q0;
                        if (a0 < b0) {
                            const int key = ((tid * 4 + g) << 13) | (zf[g] - 2 * (s - q0) + 4096);
                            queue[a0] = key;
                            for (int bnd = (a0 / QSEG + 1) * QSEG; bnd < b0; bnd += QSEG) queue[bnd] = key;
                        }
                        s += zc[g];
                    }
                }
                __syncthreads();
                const int nq = min(QBK, tot - q0);
                // a wave whose entries all lie past the round's end has nothing to do (its
                // queue positions hold no marks)
                if (QSEG * wave < nq) {
                int keys[QEPT];
                int kmax = 0;
#pragma unroll
                for (int i = 0; i < QEPT; i++) {
                    keys[i] = queue[QEPT * tid + i];
                    kmax = max(kmax, keys[i]);
                }
#pragma unroll
                for (int i = 0; i < QEPT; i++) queue[QEPT * tid + i] = 0;
                int key = wave_excl_max_nonneg(kmax);
                // delays and amplitudes of the thread's entries first (independent chains), the
                // sums after; an entry outside the round or the tile's delays gets v = 0
                int en0[QEPT], eph[QEPT];
                float ev[QEPT], eu[QEPT];
#pragma unroll
                for (int i = 0; i < QEPT; i++) {
                    key = max(key, keys[i]);
                    const int pos = QEPT * tid + i;
                    const bool inq = pos < nq;
                    const int slot = inq ? key >> 15 : 0;
                    const int nz = inq ? (key & 8191) - 4096 + 2 * pos : 0;
                    const double vz = offs[Nx + Ny + nz];
                    double d2 = rb2[slot];
                    d2 = d2 + vz * vz;
                    const double dist = sqrt_delay(d2);
                    const double tau = dist * fsc;
                    const double fl = floor(tau);
                    const int n0 = (int)fl;
                    const bool live = inq && n0 >= lo_n0 && n0 <= hi_n0;
                    const float f = (float)(tau - fl);
                    const float distf = (float)dist;
                    float amp = ratt[slot] * att[Nx + Ny + nz];
                    amp *= __builtin_amdgcn_rcpf(12.566370614359172f * distf);
                    if (pattern) {
                        const int xy = rxy[slot];
                        const int nx = xy & 0xffff, ny = xy >> 16;
                        amp = apply_patterns(R, amp, offs[nx], offs[Nx + ny], vz, distf, nx, ny, nz, a.spkr, a.mic,
                                             os, orr);
                    }
                    ev[i] = live ? amp * (PK ? pscale : scale) : 0.0f;
                    en0[i] = n0 - lo_n0;
                    if (LUT) {
                        // 16 (tau - n0) = p + w on the table's grid (exact in FP64)
                        const double g16 = (tau - fl) * 16.0;
                        const double gp = floor(g16);
                        eph[i] = (int)gp;
                        eu[i] = (float)(g16 - gp);
                    } else {
                        eph[i] = 0;
                        eu[i] = fmaf(2.0f, f, -1.0f);
                    }
                }
#pragma unroll
                for (int i = 0; i < QEPT; i++) {
                    const float v = ev[i];
                    if (v == 0.0f) continue;
                    if (LUT) {
                        const int ph = eph[i];
                        const float w = eu[i];
                        int* dp = D32 + ph * DL + en0[i];
                        lds_add32(dp, fmaf(-v, w, v));
                        lds_add32(ph < LUT_NB - 1 ? dp + DL : dp - (LUT_NB - 1) * DL + 1, v * w);
                        continue;
                    }
                    const float u = eu[i];
                    const float u2 = u + u;
                    if (PK) {
                        // T_m by the three-term recurrence at row 0's scale; row m's field
                        // takes 2^-(F_0 - F_m) of it (pack_field_bits)
                        float b[NB];
                        b[0] = v;
                        b[1] = v * u;
#pragma unroll
                        for (int m = 2; m < NB; m++) b[m] = fmaf(u2, b[m - 1], -b[m - 2]);
                        unsigned long long* wp = Dw + en0[i];
                        constexpr int F0 = pack_field_bits(NB, 0);
                        lds_add_pair<pack_lo_bits(NB, 0)>(wp, b[0], b[NB - 1] * exp2_neg<F0 - pack_field_bits(NB, NB - 1)>());
                        lds_add_pair<pack_lo_bits(NB, 1)>(wp + DL, b[1] * exp2_neg<F0 - pack_field_bits(NB, 1)>(),
                                                          b[NB - 2] * exp2_neg<F0 - pack_field_bits(NB, NB - 2)>());
                        lds_add_pair<pack_lo_bits(NB, 2)>(wp + 2 * DL, b[2] * exp2_neg<F0 - pack_field_bits(NB, 2)>(),
                                                          b[NB - 3] * exp2_neg<F0 - pack_field_bits(NB, NB - 3)>());
                        if (NB == 8)
                            lds_add_pair<pack_lo_bits(NB, 3)>(wp + 3 * DL,
                                                              b[3] * exp2_neg<F0 - pack_field_bits(NB, 3)>(),
                                                              b[4] * exp2_neg<F0 - pack_field_bits(NB, 4)>());
                        continue;
                    }
                    long long* dp = D + en0[i];
                    float b0 = v, b1 = v * u;
                    lds_add(dp, b0);
                    lds_add(dp + DL, b1);
#pragma unroll
                    for (int m = 2; m < NB; m++) {
                        const float b2 = fmaf(u2, b1, -b0);
                        lds_add(dp + m * DL, b2);
                        b0 = b1;
                        b1 = b2;
                    }
                }
                }
                __syncthreads();
            }
        }
        __syncthreads();

        // fixed point -> FP32 in output units, in place: thread tid converts columns
        // d = tid + s WGB (s < BCS) of every row into registers, and after the barrier
        // writes them to m DLP + d + 4 (d >> 4)
        float out_scale = 1.0f;  // H16: the power of two that undoes the tile's FP16 scale
        if constexpr (PK && (DLC == 640 || DLC == 1152) && NB == 8) {
            // 16 kHz packed sums: word e = tid + 512 k of the [4][DLC] words (DLC = 640 or
            // 1152), so every thread converts DLC / 128 words; e / DLC is the same across a
            // wave (512 and DLC are multiples of 64).  High fields: the word minus its sign-extended low field
            // (row 0's scale 2^-24 of it for every word); low fields: 2^(F_0 - S) of row 0's.
            constexpr int F0 = pack_field_bits(NB, 0), KW = NB / 2 * DLC / WGB;
            static_assert(KW * WGB == NB / 2 * DLC && DLC % 64 == 0, "whole words per thread, rows per wave");
            float fh[KW], fl[KW];
            float mx = 0.0f;
#pragma unroll
            for (int k = 0; k < KW; k++) {
                const int e = tid + WGB * k;
                const int w = __builtin_amdgcn_readfirstlane(e / DLC);
                const int S = w == 0 ? pack_lo_bits(NB, 0) : w == 1 ? pack_lo_bits(NB, 1) : w == 2 ? pack_lo_bits(NB, 2) : pack_lo_bits(NB, 3);
                const unsigned long long v = Dw[e];
                const int lo = __builtin_amdgcn_sbfe((int)(unsigned)v, 0, S);
                const long long h = (long long)v - (long long)lo;
                fh[k] = fix64_to_f32(h, inv_scale * exp2_neg<64 - F0>());
                fl[k] = (float)lo * (inv_scale * (float)(1 << (F0 - S)));
                mx = fmaxf(mx, fmaxf(fabsf(fh[k]), fabsf(fl[k])));
            }
            if (H16) {
                const int wm = wave_max_nonneg(__float_as_int(mx));
                if (lane == 63) atomicMax(misc + 12, wm);
            }
            __syncthreads();
            float sc = 1.0f;
            if (H16) {
                const float big = __int_as_float(misc[12]);
                const int ex = (big > 0.0f && big < 3.0e38f) ? 14 - ilogbf(big) : 0;
                sc = ldexpf(1.0f, min(max(ex, -120), 120));
                out_scale = ldexpf(1.0f, -min(max(ex, -120), 120) - (SPL ? 14 : 0));
            }
#pragma unroll
            for (int k = 0; k < KW; k++) {
                const int e = tid + WGB * k;
                const int w = __builtin_amdgcn_readfirstlane(e / DLC);
                const int d = e - DLC * w;
                const int col = d + 4 * (d >> 4);
#pragma unroll
                for (int half = 0; half < 2; half++) {
                    const int m = half ? NB - 1 - w : w;
                    const float x = half ? fl[k] : fh[k];
                    if (SPL) {
                        const float xs = x * sc;
                        const _Float16 hi = (_Float16)xs;
                        Dh[m * DLP + col] = __builtin_bit_cast(unsigned short, hi);
                        if (!HI1)
                            Dh[(NB + m) * DLP + col] = __builtin_bit_cast(unsigned short, (_Float16)((xs - (float)hi) * 2048.0f));
                    } else if (H16) {
                        Dh[m * DLP + col] = half_bits<H16>(x * sc);
                    } else {
                        Df[m * DLP + col] = x;
                    }
                }
            }
            __syncthreads();
        } else {
            float fv[NB][BCS];
            float mx = 0.0f;
#pragma unroll
            for (int sl = 0; sl < BCS; sl++) {
                const int d = min(tid + sl * WGB, DL - 1);  // columns past the row are not written
                if (PK) {
                    // row m's scale is 2^-(F_0 - F_m) of row 0's; a high field is the word
                    // minus its sign-extended low field, a multiple of 2^S
                    constexpr int F0 = pack_field_bits(NB, 0);
#pragma unroll
                    for (int w = 0; w < NB / 2; w++) {
                        const unsigned long long v = Dw[w * DL + d];
                        const float ih = inv_scale * (float)(1LL << (F0 - pack_field_bits(NB, w))) *
                                         (w == 0   ? exp2_neg<pack_lo_bits(NB, 0)>()
                                          : w == 1 ? exp2_neg<pack_lo_bits(NB, 1)>()
                                          : w == 2 ? exp2_neg<pack_lo_bits(NB, 2)>()
                                                   : exp2_neg<pack_lo_bits(NB, 3)>());
                        const float il = inv_scale * (float)(1LL << (F0 - pack_field_bits(NB, NB - 1 - w)));
                        long long h, l;
                        if (w == 0) unpack_pair<pack_lo_bits(NB, 0)>(v, h, l);
                        else if (w == 1) unpack_pair<pack_lo_bits(NB, 1)>(v, h, l);
                        else if (w == 2) unpack_pair<pack_lo_bits(NB, 2)>(v, h, l);
                        else unpack_pair<pack_lo_bits(NB, 3)>(v, h, l);
                        fv[w][sl] = fix64_to_f32(h, ih);
                        fv[NB - 1 - w][sl] = (float)(int)l * il;
                    }
                } else {
#pragma unroll
                    for (int m = 0; m < NB; m++) {
                        if (LUT) fv[m][sl] = (float)D32[m * DL + d] * inv_scale;
                        else fv[m][sl] = fix64_to_f32(D[m * DL + d], inv_scale);
                    }
                }
#pragma unroll
                for (int m = 0; m < NB; m++) mx = fmaxf(mx, fabsf(fv[m][sl]));
            }
            if (H16) {
                // non-negative floats order as ints; one atomic per wave (atomics of the lanes
                // of a wave on one LDS word run one lane at a time: 512 per tile had cost ~1 ms
                // of the 4096-room batch)
                const int wm = wave_max_nonneg(__float_as_int(mx));
                if (lane == 63) atomicMax(misc + 12, wm);
            }
            __syncthreads();
            float sc = 1.0f;
            if (H16) {
                // the tile's largest |D| maps into [2^14, 2^15): inside the FP16 range
                const float big = __int_as_float(misc[12]);
                const int e = (big > 0.0f && big < 3.0e38f) ? 14 - ilogbf(big) : 0;
                sc = ldexpf(1.0f, min(max(e, -120), 120));
                out_scale = ldexpf(1.0f, -min(max(e, -120), 120) - (SPL ? 14 : 0));
            }
#pragma unroll
            for (int m = 0; m < NB; m++)
#pragma unroll
                for (int sl = 0; sl < BCS; sl++) {
                    const int d = tid + sl * WGB;
                    if (d < DL) {
                        if (SPL) {
                            const float xs = fv[m][sl] * sc;
                            const _Float16 hi = (_Float16)xs;
                            Dh[m * DLP + d + 4 * (d >> 4)] = __builtin_bit_cast(unsigned short, hi);
                            if (!HI1)
                                Dh[(NB + m) * DLP + d + 4 * (d >> 4)] =
                                    __builtin_bit_cast(unsigned short, (_Float16)((xs - (float)hi) * 2048.0f));
                        } else if (H16) {
                            Dh[m * DLP + d + 4 * (d >> 4)] = half_bits<H16>(fv[m][sl] * sc);
                        }
                        else Df[m * DLP + d + 4 * (d >> 4)] = fv[m][sl];
                    }
                }
            __syncthreads();
        }

        // Toeplitz products.  BTK = BT: wave w takes samples [256 (w % NOT), +256) and basis
        // rows (w / NOT) NB/4 .. +NB/4; BTK = 2 BT: basis quarter w % 4 and the 256-sample
        // blocks w / 4 and w / 4 + 2 (UPW = 2 products, sharing the coefficient operands).
        // Lane (c, g) = (lane & 15, lane >> 4) supplies A[c][k] and X[k][c] for k = g KS + i,
        // i < KS (one k per lane group and step)
        {
            const int mq = BTK == BT ? wave / NOT : (wave & 3);
            const int ot = BTK == BT ? wave % NOT : (wave >> 2);  // product u: block ot + 2 u
            const int c = lane & 15, g = lane >> 4;
            floatx4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            floatx4 accx = {0.0f, 0.0f, 0.0f, 0.0f};  // SPL: the cross terms, 2^-11 of acc's units
            floatx4 accu[UPW];
            if constexpr (SPL && (DLC == 640 || DLC == 1152)) {
                // FP16 pairs at 16 kHz (KS = 36, CL = 160): the wave's two rows are nine K
                // steps each; the operand reads of step q + 1 are issued before the products
                // of step q (three per block)
                constexpr int KSC = 36, CLC = 4 * KSC + 16, NR = NB / 4, NSTEP = NR * KSC / 4;
                constexpr int AHL = NCP * NB * (CLC + CPAD);  // halves from the hi to the lo copies
                const int rho = c & (NCP - 1), m0 = mq * NR;
                const unsigned short* Ap0 = chs + (rho * NB + m0) * (CLC + CPAD) + g * KSC + 16 - c + rho;
                const unsigned short* Bp0 = Dh + m0 * DLP;
                int kbu[UPW];
#pragma unroll
                for (int u = 0; u < UPW; u++) kbu[u] = 256 * (ot + 2 * u) + 16 * c + g * KSC;
                half4_t ah[2], al[2], xh[2][UPW], xl[2][UPW];
                floatx4 ahh[UPW], axx[UPW];
#pragma unroll
                for (int u = 0; u < UPW; u++) {
                    ahh[u] = acc;
                    axx[u] = acc;
                }
                // four coefficients of a lane: 8-byte aligned with four copies, 4-byte with two;
                // one copy: the three aligned words that hold them, funnel-shifted by the odd half
                auto ld_coef = [&](const unsigned short* p) -> half4_t {
                    if constexpr (NCP == 4) {
                        return *reinterpret_cast<const half4_t*>(p);
                    } else if constexpr (NCP == 2) {
                        const unsigned* q = reinterpret_cast<const unsigned*>(p);
                        uint2 v;
                        v.x = q[0];
                        v.y = q[1];
                        return __builtin_bit_cast(half4_t, v);
                    } else {
                        const unsigned od = (unsigned)(p - chs) & 1u;
                        const unsigned* q = reinterpret_cast<const unsigned*>(p - od);
                        const unsigned w0 = q[0], w1 = q[1], w2 = q[2];
                        uint2 v;
                        v.x = __builtin_amdgcn_alignbit(w1, w0, 16u * od);
                        v.y = __builtin_amdgcn_alignbit(w2, w1, 16u * od);
                        return __builtin_bit_cast(half4_t, v);
                    }
                };
                auto load_step = [&](int q, int b) {
                    const int r = q / (KSC / 4), t4 = q % (KSC / 4);
                    const unsigned short* ap = Ap0 + r * (CLC + CPAD) + 4 * t4;
                    ah[b] = ld_coef(ap);
                    if (!HI1) al[b] = ld_coef(ap + AHL);
#pragma unroll
                    for (int u = 0; u < UPW; u++) {
                        const int dl = kbu[u] + 4 * t4;
                        const unsigned short* xp = Bp0 + r * DLP + dl + 4 * (dl >> 4);
                        xh[b][u] = *reinterpret_cast<const half4_t*>(xp);
                        if (!HI1) xl[b][u] = *reinterpret_cast<const half4_t*>(xp + NB * DLP);
                    }
                };
                load_step(0, 0);
#pragma unroll
                for (int q = 0; q < NSTEP; q++) {
                    if (q + 1 < NSTEP) load_step(q + 1, (q + 1) & 1);
                    __builtin_amdgcn_sched_barrier(0);  // the next step's reads stay ahead
#pragma unroll
                    for (int u = 0; u < UPW; u++) {
                        ahh[u] = __builtin_amdgcn_mfma_f32_16x16x16f16(ah[q & 1], xh[q & 1][u], ahh[u], 0, 0, 0);
                        if (!HI1) {
                            axx[u] = __builtin_amdgcn_mfma_f32_16x16x16f16(ah[q & 1], xl[q & 1][u], axx[u], 0, 0, 0);
                            axx[u] = __builtin_amdgcn_mfma_f32_16x16x16f16(al[q & 1], xh[q & 1][u], axx[u], 0, 0, 0);
                        }
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
#pragma unroll
                for (int u = 0; u < UPW; u++) accu[u] = (HI1 ? ahh[u] : ahh[u] + axx[u] * (1.0f / 2048.0f)) * out_scale;
            } else {
if (H16) {
                // four k per lane and instruction: K steps of 4 per lane group, KS / 4 products
                const int rho = c & 3;
#pragma unroll
                for (int mm = 0; mm < (NB + 3) / 4; mm++) {
                    // rows 2 mq, 2 mq + 1 (NB = 8) or mq, mq + 4 (NB = 6: waves 2 and 3 take one)
                    const int m = NB % 4 == 0 ? mq * (NB / 4) + mm : mq + 4 * mm;
                    if (NB % 4 != 0 && m >= NB) break;
                    static_assert(NCP == 4, "the general 16-bit products read four copies");
                    const unsigned short* Ap = chs + (rho * NB + m) * (CL + 4) + g * KS + 16 - c + rho;
                    const unsigned short* Bp = Dh + m * DLP;
                    const int kb = 256 * ot + 16 * c + g * KS;
                    for (int t4 = 0; t4 < KS / 4; t4++) {
                        const int dl = kb + 4 * t4;
                        if (SPL) {
                            const half4_t xh = *reinterpret_cast<const half4_t*>(Bp + dl + 4 * (dl >> 4));
                            const half4_t ah = *reinterpret_cast<const half4_t*>(Ap + 4 * t4);
                            acc = __builtin_amdgcn_mfma_f32_16x16x16f16(ah, xh, acc, 0, 0, 0);
                            if (!HI1) {
                                const half4_t xl = *reinterpret_cast<const half4_t*>(Bp + NB * DLP + dl + 4 * (dl >> 4));
                                const half4_t al = *reinterpret_cast<const half4_t*>(Ap + 4 * NB * (CL + 4) + 4 * t4);
                                accx = __builtin_amdgcn_mfma_f32_16x16x16f16(ah, xl, accx, 0, 0, 0);
                                accx = __builtin_amdgcn_mfma_f32_16x16x16f16(al, xh, accx, 0, 0, 0);
                            }
                        } else if (H16 == 2) {
                            const short4_t xv = *reinterpret_cast<const short4_t*>(Bp + dl + 4 * (dl >> 4));
                            const short4_t av = *reinterpret_cast<const short4_t*>(Ap + 4 * t4);
                            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, xv, acc, 0, 0, 0);
                        } else {
                            const half4_t xv = *reinterpret_cast<const half4_t*>(Bp + dl + 4 * (dl >> 4));
                            const half4_t av = *reinterpret_cast<const half4_t*>(Ap + 4 * t4);
                            acc = __builtin_amdgcn_mfma_f32_16x16x16f16(av, xv, acc, 0, 0, 0);
                        }
                    }
                }
                if (SPL && !HI1) acc = acc + accx * (1.0f / 2048.0f);
                acc = acc * out_scale;
            } else if constexpr (DLC == 640 && NB % 4 == 0) {
                // The 16 kHz row length pins KS = 36: the wave's NB/4 rows are nine K steps each,
                // taken in groups of three.  The operands of group q + 1 (three X vectors and 12
                // coefficients per lane) are read before the 12 products of group q, so each
                // read's latency hides behind 384 cycles of products.
                constexpr int KSC = 36, GQ = 3, NGR = KSC / 4 / GQ, NR = NB / 4;
                const int m0 = mq * NR;
                const float* Ap0 = ctab + m0 * CL + g * KSC + 16 - c;
                const float* Bp0 = Df + m0 * DLP;
                const int kb = 256 * ot + 16 * c + g * KSC;
                float4 xb[2][GQ];
                float ab[2][4 * GQ];
                auto load_group = [&](int q, float4* xv, float* av) {
                    const int r = q / NGR, t0 = (q % NGR) * GQ;
#pragma unroll
                    for (int t = 0; t < GQ; t++) {
                        const int dl = kb + 4 * (t0 + t);
                        xv[t] = *reinterpret_cast<const float4*>(Bp0 + r * DLP + dl + 4 * (dl >> 4));
                    }
#pragma unroll
                    for (int i = 0; i < 4 * GQ; i++) av[i] = Ap0[r * CL + 4 * t0 + i];
                };
                load_group(0, xb[0], ab[0]);
#pragma unroll
                for (int q = 0; q < NR * NGR; q++) {
                    if (q + 1 < NR * NGR) load_group(q + 1, xb[(q + 1) & 1], ab[(q + 1) & 1]);
                    __builtin_amdgcn_sched_barrier(0);  // the next group's reads stay ahead
#pragma unroll
                    for (int t = 0; t < GQ; t++) {
                        const float4 x = xb[q & 1][t];
                        const float* av = ab[q & 1] + 4 * t;
                        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av[0], x.x, acc, 0, 0, 0);
                        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av[1], x.y, acc, 0, 0, 0);
                        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av[2], x.z, acc, 0, 0, 0);
                        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av[3], x.w, acc, 0, 0, 0);
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
            } else
#pragma unroll
            for (int mm = 0; mm < (NB + 3) / 4; mm++) {
                const int m = NB % 4 == 0 ? mq * (NB / 4) + mm : mq + 4 * mm;
                if (NB % 4 != 0 && m >= NB) break;
                const float* Ap = ctab + m * CL + g * KS + 16 - c;
                const float* Bp = Df + m * DLP;
                const int kb = 256 * ot + 16 * c + g * KS;
                // operands of step t4 + 1 are read before the products of step t4; the
                // read after the last step (4 floats past this lane's K range) stays
                // inside the LDS of D and the rows and is not used
                float4 xv = *reinterpret_cast<const float4*>(Bp + kb + 4 * (kb >> 4));
                float a0 = Ap[0], a1 = Ap[1], a2 = Ap[2], a3 = Ap[3];
                for (int t4 = 0; t4 < KS / 4; t4++) {
                    const int dn = kb + 4 * (t4 + 1);
                    const float4 xn = *reinterpret_cast<const float4*>(Bp + dn + 4 * (dn >> 4));
                    const float* an = Ap + 4 * (t4 + 1);
                    const float n0 = an[0], n1 = an[1], n2 = an[2], n3 = an[3];
                    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, xv.x, acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, xv.y, acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a2, xv.z, acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a3, xv.w, acc, 0, 0, 0);
                    xv = xn;
                    a0 = n0;
                    a1 = n1;
                    a2 = n2;
                    a3 = n3;
                }
            }
            accu[0] = acc;  // UPW = 1 on these paths
            }
            // result element (row 4 g + i, column c) of product u is sample 16 c + 4 g + i of
            // block ot + 2 u
#pragma unroll
            for (int u = 0; u < UPW; u++)
                *reinterpret_cast<float4*>(ybuf + mq * BTK + 256 * (ot + 2 * u) + 16 * c + 4 * g) =
                    make_float4(accu[u][0], accu[u][1], accu[u][2], accu[u][3]);
        }
        __syncthreads();
        for (int t = tid; t < th - tl; t += WGB)
            orow[tl + t] = ((ybuf[t] + ybuf[BTK + t]) + ybuf[2 * BTK + t]) + ybuf[3 * BTK + t];
        if (BTK != BT) {
            // ybuf covered the queue: its run marks start from zero again (the queue is next
            // written after the next tile's row scan, whose barrier follows these reads)
            __syncthreads();
            for (int e = tid; e < QBK; e += WGB) queue[e] = 0;
        }
        // no barrier: the sums' region was last read by the products (before the barrier
        // above), and ybuf (the row arrays) is next written after the next tile's row scan
        for (int e = tid; e < nwords; e += WGB) {
            if (LUT) D32[e] = 0;
            else D[e] = 0;
        }
        if (H16 && tid == 0) misc[12] = 0;  // read by every thread before the barrier above
        // the row scan of the next tile has barriers before its first atomic
    }
}

// The FP32 and table-mode kernels with the row length compiled in are held to 128 VGPRs
// (two workgroups per CU); the same cap on the 16-bit-input kernels schedules them worse.
template <int NB, int DLC, int H16, bool LUT, bool PK, int BTK = BT>
__global__ __launch_bounds__(WGB) __attribute__((amdgpu_waves_per_eu(DLC && (H16 == 0 || H16 == 3 || H16 == 4) ? 4 : 1)))
void ism_basis_kernel(AccArgs a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    basis_item<NB, DLC, H16, LUT, PK, BTK>(a, blockIdx.x, smem_raw);
}

// Persistent form, items taken from a counter as in ism_accumulate_persistent_kernel.
template <int NB, int DLC, int H16, bool LUT, bool PK, int BTK = BT>
__global__ __launch_bounds__(WGB) __attribute__((amdgpu_waves_per_eu(DLC && (H16 == 0 || H16 == 3 || H16 == 4) ? 4 : 1)))
void ism_basis_persistent_kernel(AccArgs a, int n_items, unsigned* counters) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ int next;
    for (;;) {
        if (threadIdx.x == 0) next = (int)atomicAdd(counters, 1u);
        __syncthreads();
        const int item = next;
        __syncthreads();
        if (item >= n_items) break;
        basis_item<NB, DLC, H16, LUT, PK, BTK>(a, item, smem_raw);
    }
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(counters + 1, 1u) == gridDim.x - 1) {
            atomicExch(counters, 0u);
            atomicExch(counters + 1, 0u);
        }
    }
}

__device__ inline void philox4x32_10(unsigned c0, unsigned c1, unsigned c2, unsigned c3, unsigned k0, unsigned k1,
                                     unsigned out[4]) {
#pragma unroll
    for (int rr = 0; rr < 10; rr++) {
        // one 32x32->64 multiply (v_mad_u64_u32) per product instead of mul_lo + mul_hi
        const unsigned long long p0 = (unsigned long long)0xD2511F53u * c0;
        const unsigned long long p1 = (unsigned long long)0xCD9E8D57u * c2;
        const unsigned lo0 = (unsigned)p0, hi0 = (unsigned)(p0 >> 32);
        const unsigned lo1 = (unsigned)p1, hi1 = (unsigned)(p1 >> 32);
        const unsigned n0 = hi1 ^ c1 ^ k0;
        const unsigned n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// Noise value u (0..3) of one Philox block.  Logistic (the reference's
// 0.551329 log(u / (1 - u + 1e-6)), src/gpuRIR_cuda.cu:454-455) with
// u = (x+1)/2^32 and 1-u = (2^32-1-x)/2^32 as exact integers, or Box-Muller.
// The ratio is formed with the hardware reciprocal and its logarithm with the hardware
// log2 (absolute error of a few 1e-7 on z, against the 1e-5 bar of the tail).
__device__ inline float tail_noise(const unsigned w[4], int u, int noise) {
    if (noise == 0) {
        const float num = (float)w[u] + 1.0f;
        const float den = (float)(0xFFFFFFFFu - w[u]) + 4294.967296f;
        return (0.551329f * 0.69314718055994531f) * __builtin_amdgcn_logf(num * __builtin_amdgcn_rcpf(den));
    }
    const int pr2 = u >> 1;
    const float lnua = log1pf(-(float)(0xFFFFFFFFu - w[2 * pr2]) * 0x1p-32f);
    const float ub = (float)w[2 * pr2 + 1] * 0x1p-32f;
    const float rad = sqrtf(-2.0f * lnua);
    return (u & 1) ? rad * sinpif(2.0f * ub) : rad * cospif(2.0f * ub);
}

// Sum over the workgroup (every thread gets it); red may be reused right after.
__device__ inline double block_sum(double v, double* red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();  // earlier readers of red are done
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < NWAVE; w++) t += red[w];
    return t;
}

// Sum of the partial rows of RIRs whose images were split into several chunks, in a
// fixed order (every run gives the same bits).  A block (RIR, t0 | WIDE) covers either
// 256 samples with one thread per sample adding all chunks (few chunks), or 64 samples
// with four thread groups adding chunks g, g+4, ... and combined in order (many chunks);
// eight loads are in flight per thread.
constexpr int SUM_WIDE = 1 << 30;
__global__ __launch_bounds__(WG) void partial_sum_kernel(const RoomDev* rooms, const int* rir_room,
                                                         const int2* blocks, const float* part, float* out) {
    __shared__ float red[4][64];
    const int2 b = blocks[blockIdx.x];
    const RoomDev* Rp = rooms + rir_room[b.x];
    const int C = Rp->n_chunks, nISM = Rp->nISM, nS = Rp->nSamples;
    const int m = b.x - Rp->rir_off;
    const bool wide = (b.y & SUM_WIDE) != 0;
    const int t0 = b.y & (SUM_WIDE - 1);
    const float* row = part + Rp->part_off + (long long)m * C * nISM;
    float* orow = out + Rp->out_off + (long long)m * nS;
    if (wide) {
        const int t = t0 + threadIdx.x;
        if (t >= nISM) return;
        const float* pp = row + t;
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = u < C ? pp[(long long)u * nISM] : 0.0f;
        float acc = v[0];
#pragma unroll
        for (int u = 1; u < 8; u++)
            if (u < C) acc += v[u];
        orow[t] = acc;
        return;
    }
    const int l = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int t = t0 + l;
    float acc = 0.0f;
    if (t < nISM) {
        const float* pp = row + t;
        int c = g;
        for (; c + 28 < C; c += 32) {
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; u++) v[u] = pp[(long long)(c + 4 * u) * nISM];
#pragma unroll
            for (int u = 0; u < 8; u++) acc += v[u];
        }
        for (; c < C; c += 4) acc += pp[(long long)c * nISM];
    }
    red[g][l] = acc;
    __syncthreads();
    if (g == 0 && t < nISM) orow[t] = ((red[0][l] + red[1][l]) + red[2][l]) + red[3][l];
}

// One workgroup per RIR: the envelope of src/gpuRIR_cuda.cu:415-443 and the tail of :445-461.
__global__ __launch_bounds__(WG) void ism_finalize_kernel(FinArgs a) {
    __shared__ double red[NWAVE];
    const int g = blockIdx.x;
    const RoomK R = load_room(a.rooms, a.rir_room[g], a.prm);
    const int m = g - R.rir_off;
    const int s = m / R.n_rcv, r = m - s * R.n_rcv;
    const int tid = threadIdx.x;
    const int nISM = R.nISM, nS = R.nSamples, nDiff = nS - nISM;
    float* o = a.out + R.out_off + (long long)m * nS;

    // direct-path delay: image with k = 0 on every axis (src/gpuRIR_cuda.cu:341,368-369)
    const float* ps = a.pos_src + 3 * (R.src_off + s);
    const float* pr = a.pos_rcv + 3 * (R.rcv_off + r);
    double v[3];
    for (int d = 0; d < 3; d++) v[d] = (0.0 + (double)ps[d]) - (double)pr[d];
    double d2 = v[0] * v[0];
    d2 = d2 + v[1] * v[1];
    d2 = d2 + v[2] * v[2];
    const double tau_dp = sqrt(d2) * R.fs_over_c;
    const int w_start = (int)ceil(fmax((double)nISM - (double)a.w_sz, tau_dp));

    // the early part is complete here: single-chunk items wrote it directly and
    // partial_sum_kernel added up the partial rows of multi-chunk RIRs
    double pw = 0.0;
    for (int t = w_start + tid; t < nISM; t += WG) {
        const double x = o[t];
        pw += x * x;
    }
    if (nDiff <= 0) return;
    const double P = block_sum(pw, red) / (double)(nISM - w_start);

    // Sabine T60 (src/gpuRIR_cuda.cu:123-130) and the envelope parameters
    const double L[3] = {R.L0, R.L1, R.L2};
    const double b[6] = {R.b0, R.b1, R.b2, R.b3, R.b4, R.b5};
    const double Sa = ((1.0 - b[0] * b[0]) + (1.0 - b[1] * b[1])) * L[1] * L[2] +
                      ((1.0 - b[2] * b[2]) + (1.0 - b[3] * b[3])) * L[0] * L[2] +
                      ((1.0 - b[4] * b[4]) + (1.0 - b[5] * b[5])) * L[0] * L[1];
    const double T60 = 0.161 * (L[0] * L[1] * L[2]) / Sa;
    const double alpha = -13.8155 / (T60 * a.fs);
    const double w_center = (double)w_start + (double)(nISM - w_start) / 2.0;
    const double A = P / exp(alpha * (w_center - tau_dp));
    const float gain = (float)((double)R.level * sqrt(A));
    const float ha = (float)(0.5 * alpha);
    const float base = (float)(0.5 * alpha * ((double)nISM - tau_dp));

    const unsigned k0 = (unsigned)a.seed, k1 = (unsigned)(a.seed >> 32);
    // optional cross-fade on [nISM - X, nISM): (1 - r) early part + r diffuse model, r = (k + 1/2)/X,
    // with noise indices nDiff + k (the envelope above was fitted on the early part alone)
    const int X = min(a.crossfade, nISM);
    for (int k = tid; k < X; k += WG) {
        const int sd = nDiff + k;
        unsigned w[4];
        philox4x32_10((unsigned)(sd >> 2), (unsigned)m, R.room_id, a.call_offset, k0, k1, w);
        const float z = tail_noise(w, sd & 3, a.noise);
        const int t = nISM - X + k;
        const float env = __builtin_amdgcn_exp2f(fmaf((float)(t - nISM), ha, base) * 1.4426950408889634f);
        const float r = ((float)k + 0.5f) / (float)X;
        o[t] = fmaf(r, gain * env * z, (1.0f - r) * o[t]);
    }
    // Tail: noise block B (samples 4B..4B+3) is one Philox call.  The stores go out as aligned
    // 16-byte groups: group q holds samples off + 4q .. off + 4q + 3 (off puts tail + off on a
    // 16-byte boundary), i.e. the last 4 - off values of block q and the first off of block q + 1.
    // Lane l of a wave draws block B = b0 + l and writes group B - 1, taking block B - 1 from
    // lane l - 1; lane 0 draws the block that the previous range's lane 63 drew, so the ranges
    // of successive waves advance by 63 blocks and every group is written once.
    float* tail = o + nISM;
    const int nblk = (nDiff + 3) >> 2;
    const int lane = tid & 63, wave = tid >> 6;
    // the envelope's ratio between neighbouring samples: one exp2 per four samples
    const float q1 = __builtin_amdgcn_exp2f(ha * 1.4426950408889634f), q2 = q1 * q1, q3 = q2 * q1;
    const int off = (4 - (int)((reinterpret_cast<unsigned long long>(tail) >> 2) & 3)) & 3;
    for (int b0 = 63 * wave; b0 <= nblk; b0 += 63 * NWAVE) {
        const int B = b0 + lane;
        float z[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        if (B < nblk) {
            unsigned w[4];
            philox4x32_10((unsigned)B, (unsigned)m, R.room_id, a.call_offset, k0, k1, w);
#pragma unroll
            for (int u = 0; u < 4; u++) z[u] = 4 * B + u < nDiff ? tail_noise(w, u, a.noise) : 0.0f;
        }
        float p[4];
#pragma unroll
        for (int u = 0; u < 4; u++) p[u] = __shfl_up(z[u], 1);
        if (lane == 0 && b0 > 0) continue;
        float v[4];
        switch (off) {  // uniform over the workgroup
            case 0: v[0] = p[0]; v[1] = p[1]; v[2] = p[2]; v[3] = p[3]; break;
            case 1: v[0] = p[1]; v[1] = p[2]; v[2] = p[3]; v[3] = z[0]; break;
            case 2: v[0] = p[2]; v[1] = p[3]; v[2] = z[0]; v[3] = z[1]; break;
            default: v[0] = p[3]; v[1] = z[0]; v[2] = z[1]; v[3] = z[2]; break;
        }
        const int sd0 = off + 4 * (B - 1);
        // exp(x) = 2^(x log2 e) with the hardware exp2 (relative error ~1e-7) at the group's
        // first sample, times the ratio's powers for the other three
        const float ge = gain * __builtin_amdgcn_exp2f(fmaf((float)sd0, ha, base) * 1.4426950408889634f);
        v[0] = ge * v[0];
        v[1] = (ge * q1) * v[1];
        v[2] = (ge * q2) * v[2];
        v[3] = (ge * q3) * v[3];
        if (sd0 >= 0 && sd0 + 3 < nDiff) {
            *reinterpret_cast<float4*>(tail + sd0) = make_float4(v[0], v[1], v[2], v[3]);
        } else {
#pragma unroll
            for (int u = 0; u < 4; u++)
                if (sd0 + u >= 0 && sd0 + u < nDiff) tail[sd0 + u] = v[u];
        }
    }
}

// Stage (a) listing: n0, fraction and amplitude of every (RIR, image) of one room.
__global__ __launch_bounds__(WG) void image_list_kernel(const RoomDev* rooms, const float* pos_src,
                                                        const float* pos_rcv, const float* orv_src,
                                                        const float* orv_rcv, int spkr, int mic, int ntab_max,
                                                        int* n0_out, float* f_out, float* amp_out) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const RoomK R = load_room(rooms, 0, RoomParams{nullptr, nullptr});
    const int m = blockIdx.y;
    const int s = m / R.n_rcv, r = m - s * R.n_rcv;
    double* offs = reinterpret_cast<double*>(smem_raw);
    float* att = reinterpret_cast<float*>(offs + ntab_max);
    const float* ps = pos_src + 3 * s;
    const float* pr = pos_rcv + 3 * r;
    build_axis_tables(R, ps, pr, offs, att, threadIdx.x, WG);
    float os[4] = {0, 0, 0, 0}, orr[4] = {0, 0, 0, 0};
    if (spkr != 0) load_orient(orv_src + 3 * s, os);
    if (mic != 0) load_orient(orv_rcv + 3 * r, orr);
    __syncthreads();
    const long long N = (long long)R.Nx * R.Ny * R.Nz;
    const int i = blockIdx.x * WG + threadIdx.x;
    if (i < N) {
        const ImageVal iv = image_eval(R, i, offs, att, spkr, mic, os, orr);
        const double fl = floor(iv.tau);
        n0_out[(long long)m * N + i] = (int)fl;
        f_out[(long long)m * N + i] = (float)(iv.tau - fl);
        amp_out[(long long)m * N + i] = iv.amp;
    }
}

// Algorithmic work of a plan: the number of (RIR, image, t) with t in the item's
// tile and |t - tau| < halfTw, summed over all items (stage (a) only).
__global__ __launch_bounds__(WG) void count_taps_kernel(AccArgs a, unsigned long long* total) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ unsigned long long red[NWAVE];
    const Item it = a.items[blockIdx.x];
    const RoomK R = load_room(a.rooms, it.room, a.prm);
    double* offs = reinterpret_cast<double*>(smem_raw);
    float* att = reinterpret_cast<float*>(offs + a.ntab_max);
    const int s = it.rir / R.n_rcv, r = it.rir - s * R.n_rcv;
    build_axis_tables(R, a.pos_src + 3 * (R.src_off + s), a.pos_rcv + 3 * (R.rcv_off + r), offs, att, threadIdx.x,
                      WG);
    __syncthreads();
    const float zero4[4] = {0, 0, 0, 0};
    const double h = (double)a.halfTw;
    unsigned long long cnt = 0;
    for (int i = it.i0 + (int)threadIdx.x; i < it.i1; i += WG) {
        const ImageVal iv = image_eval(R, i, offs, att, 0, 0, zero4, zero4);
        long long t0 = (long long)floor(iv.tau - h) + 1;
        long long t1 = (long long)ceil(iv.tau + h) - 1;
        t0 = t0 > it.t_lo ? t0 : it.t_lo;
        t1 = t1 < it.t_hi - 1 ? t1 : it.t_hi - 1;
        if (t1 >= t0) cnt += (unsigned long long)(t1 - t0 + 1);
    }
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < NWAVE; w++) t += red[w];
        atomicAdd(total, t);
    }
}

// ---------------------------------------------------------------------------
// adjoint (gradients of a loss L(out) with respect to beta, positions, level)
// ---------------------------------------------------------------------------
//
// out[t] = sum_i amp_i h(t - tau_i) for t < nISM, and the tail
// out[nISM + s] = level sqrt(A) exp(alpha (nISM + s - tau_dp) / 2) z_s with
// A = P exp(-alpha (w_c - tau_dp)), P = mean(out[w_start:nISM]^2) (src/gpuRIR_cuda.cu:
// 415-461).  Given g = dL/dout:
//   * tail_adjoint_kernel (one workgroup per RIR) regenerates z_s, and from
//     S1 = sum g_s sqrt(A) E_s z_s and S2 = sum g_s tail_s (nISM + s - tau_dp)/2
//     forms dL/dlevel = S1, dL/dA, dL/dalpha, dL/dtau_dp; dL/dP adds c_w out[t]
//     (c_w = dL/dP 2 / n_w) to the gradient of the early part on the envelope window;
//   * ism_adjoint_kernel (one workgroup per work item) needs, per image, only
//     G0 = sum_t g'_t h(t - tau) and G1 = sum_t g'_t h'(t - tau): dL/damp = G0 and
//     dL/dtau = -amp G1.  The windowed sinc and its derivative are recomputed tap by
//     tap (nothing per image is stored by the forward pass) and the chain rule
//     through amp(beta, v), tau(v), v = 2mL + sigma src - rcv gives per-item sums;
//   * grad_rir_kernel and grad_param_kernel add the per-item and per-RIR terms in a
//     fixed order, so the gradients are identical from run to run.
// The integer envelope window start w_start is held fixed (piecewise constant).

struct AdjArgs {
    const RoomDev* rooms;
    const Item* items;
    const float* pos_src;
    const float* pos_rcv;
    const float* orv_src;
    const float* orv_rcv;
    const float4* ktab;   // [nJ]: (k, cos 2pi k/Tw, sin 2pi k/Tw, sigma_k / 2), k = J0 + u, sigma_k = (-1)^(k+1)
    const float* y;       // forward output
    const float* gy;      // dL/dout, same layout
    const float* rir_cw;  // per RIR: c_w
    const int* rir_ws;    // per RIR: w_start
    double* item_part;    // [n_items][12]: dbeta[6], dsrc[3], drcv[3]
    const int* order;     // launch order of the items (longest first)
    int spkr, mic;
    int J0, nJ;
    float halfTw, inv_tw2, pi_tw;  // Tw/2, 2/Tw, pi/Tw
    int ntab_max, gp_len;
    // basis adjoint: LDS image of the coefficients as FP16 pairs (make_adjoint_half_copies),
    // row length of E, logical length of g' (stored with 8 halves of padding per 16), K steps
    // per lane group of the correlation products, coefficient row length
    const unsigned short* achs;
    int abt, adl, agl, aks, acl;  // abt: delays per tile
    RoomParams prm;
};

constexpr int AQ = 2 * WG;  // adjoint image queue (ring of two rounds)

// Adjoint axis tables: offsets as the forward pass, F = (-b_lo)^e0 (-b_hi)^e1 and its
// derivatives dF/db_lo = -e0 (-b_lo)^(e0-1) (-b_hi)^e1, dF/db_hi likewise.
__device__ inline void axis_entry_adj(int n, int N, double L, float sd, float rd, float blo, float bhi, double* offs,
                                      float* F, float* Flo, float* Fhi) {
    const int k = n - N / 2;
    const int m = (k + 1) >> 1;
    const int q = k & 1;
    const double sgn = (double)(1 - 2 * q);
    offs[n] = (2.0 * (double)m * L + sgn * (double)sd) - (double)rd;
    const int e0 = abs(m - q), e1 = abs(m);
    const double a = -(double)blo, b = -(double)bhi;
    const double pa = ipow(a, e0), pb = ipow(b, e1);
    F[n] = (float)(pa * pb);
    Flo[n] = e0 > 0 ? (float)(-(double)e0 * ipow(a, e0 - 1) * pb) : 0.0f;
    Fhi[n] = e1 > 0 ? (float)(-(double)e1 * pa * ipow(b, e1 - 1)) : 0.0f;
}

// Polar pattern gain and its derivative with respect to cos(theta) (src/gpuRIR_cuda.cu:139-146).
__device__ inline void pattern_gain_d(double c, int p, double* D, double* Dp) {
    switch (p) {
        case 0: *D = 1.0; *Dp = 0.0; break;
        case 1: *D = c > 0.0 ? 1.0 : 0.0; *Dp = 0.0; break;
        case 2: *D = 0.5 + 0.5 * c; *Dp = 0.5; break;
        case 3: *D = 0.25 + 0.75 * c; *Dp = 0.75; break;
        case 4: *D = 0.75 + 0.25 * c; *Dp = 0.25; break;
        default: *D = c; *Dp = 1.0; break;
    }
}

// One tap of G0, G1 at offset k (uniform constants kc = ktab[k - J0]):
//   x = k - f, sin(pi x)/pi = sigma_k S, cos(pi x) = -sigma_k C (S = sin(pi f)/pi, C = cos(pi f)),
//   u = S/x, v = (C + u)/x:  sinc = sigma_k u, sinc' = -sigma_k v,
//   w = 1/2 + ck hc + sk hs,  w' = ck ps - sk pc  (ck, sk = cos, sin 2 pi k/Tw),
//   g h = g sigma_k w u,  g h' = g sigma_k (w' u - w v).
// The sign sigma_k goes with g: the kernel stores g~(t) = (-1)^(t+1) g'(t), which is
// sigma_k g' times (-1)^n0, the image's own sign, applied once per image.  With r = 1/x:
//   g h = S (g r w),  g h' = (g r) (A - S w r),  A = S w' - C w = ck al + sk be + ga,
//   al = S ps - C hc, be = -S pc - C hs, ga = -C/2  (per image),
// so a tap takes 11 VALU instructions, each with at most one scalar operand; G0s collects
// g r w and is scaled by S once per image.
template <bool MASK>
__device__ __forceinline__ void adj_tap(float4 kc, float g, float f, float S, float hc, float hs, float al, float be,
                                        float ga, float halfTw, float& G0s, float& G1) {
    const float x = kc.x - f;
    if (MASK) g = fabsf(x) < halfTw ? g : 0.0f;
    const float r = __builtin_amdgcn_rcpf(x);
    const float w = fmaf(kc.y, hc, fmaf(kc.z, hs, 0.5f));
    const float A = fmaf(kc.y, al, fmaf(kc.z, be, ga));
    const float gr = g * r;
    G0s = fmaf(gr, w, G0s);
    G1 = fmaf(gr, fmaf(-S * w, r, A), G1);
}

// Taps k = 0 and k = 1, where |x| can be small: series for |x| < 1/16,
// sinc = 1 - (pi x)^2/6 + (pi x)^4/120, sinc' = pi^2 x (-1/3 + (pi x)^2/30).
template <bool MASK>
__device__ __forceinline__ void adj_tap_near(float4 kc, float g, float f, float S, float C, float hc, float hs,
                                             float pc, float ps, float halfTw, float& G0, float& G1) {
    const float x = kc.x - f;
    if (MASK) g = fabsf(x) < halfTw ? g : 0.0f;
    const float sg = 2.0f * kc.w;  // sigma_k
    float u, v;
    if (fabsf(x) < 0.0625f) {
        const float px2 = 9.8696044010893586f * x * x;
        const float sinc = fmaf(px2, fmaf(px2, 1.0f / 120.0f, -1.0f / 6.0f), 1.0f);
        const float sincp = 9.8696044010893586f * x * fmaf(px2, 1.0f / 30.0f, -1.0f / 3.0f);
        u = sg * sinc;
        v = -sg * sincp;
    } else {
        const float r = 1.0f / x;
        u = S * r;
        v = (C + u) * r;
    }
    const float ws = sg * fmaf(kc.y, hc, fmaf(kc.z, hs, 0.5f));  // sigma_k w
    const float wps = sg * fmaf(kc.y, ps, -kc.z * pc);             // sigma_k w'
    G0 = fmaf(g * ws, u, G0);
    G1 = fmaf(g, fmaf(wps, u, -ws * v), G1);
}

// Chain rule of one image (indices nx, ny, nz) from G0 = sum g' h and G1 = sum g' h'
// through amp(beta, v), the directivities and tau(v), in FP64, added to acc:
// dbeta[6], dsrc[3], drcv[3].
__device__ __forceinline__ void adj_chain(const RoomK& R, int nx, int ny, int nz, const double* offs, const float* F,
                                          const float* Flo, const float* Fhi, double orx, double ory, double orz,
                                          double osx, double osy, double osz, int spkr, int mic, float G0, float G1,
                                          double acc[12], double d) {
    const int Nx = R.Nx, Ny = R.Ny, Nz = R.Nz;
    const int ey = Nx + ny, ez = Nx + Ny + nz;
    const double vx = offs[nx], vy = offs[ey], vz = offs[ez];
    const double id = rcp_normal(d);
    const double ux = vx * id, uy = vy * id, uz = vz * id;
    const double Fx = F[nx], Fy = F[ey], Fz = F[ez];
    const double Fxy = Fx * Fy;
    const double att = Fxy * Fz;
    const double a0 = id * 0.079577471545947668;  // 1 / (4 pi)
    const double sx = ((nx - Nx / 2) & 1) ? -1.0 : 1.0;
    const double sy = ((ny - Ny / 2) & 1) ? -1.0 : 1.0;
    const double sz = ((nz - Nz / 2) & 1) ? -1.0 : 1.0;
    double Dr = 1.0, Drp = 0.0, Ds = 1.0, Dsp = 0.0, cr = 0.0, cs = 0.0;
    if (mic != 0) {
        cr = ux * orx + uy * ory + uz * orz;
        pattern_gain_d(cr, mic, &Dr, &Drp);
    }
    const double qx = sx * osx, qy = sy * osy, qz = sz * osz;  // mirrored source orientation
    if (spkr != 0) {
        cs = -(ux * qx + uy * qy + uz * qz);
        pattern_gain_d(cs, spkr, &Ds, &Dsp);
    }
    const double geo = Ds * Dr * a0;
    const double amp = att * geo;
    const double g0 = (double)G0, g1 = (double)G1;
    const double gbeta = g0 * geo;
    const double gyz = gbeta * (Fy * Fz), gxz = gbeta * (Fx * Fz), gxy = gbeta * Fxy;
    acc[0] += gyz * (double)Flo[nx];
    acc[1] += gyz * (double)Fhi[nx];
    acc[2] += gxz * (double)Flo[ey];
    acc[3] += gxz * (double)Fhi[ey];
    acc[4] += gxy * (double)Flo[ez];
    acc[5] += gxy * (double)Fhi[ez];
    // d amp / d v = att a0 (-Ds Dr u/d + Ds' Dr dcs/dv + Ds Dr' dcr/dv),
    // dcr/dv = (o_r - cr u)/d, dcs/dv = (-o_s' - cs u)/d;  d tau / d v = (fs/c) u
    const double kr = att * a0 * Ds * Drp * id, ks = att * a0 * Dsp * Dr * id;
    const double kd = -att * a0 * Ds * Dr * id;
    const double ku = g0 * (kd - kr * cr - ks * cs) + (-amp * g1) * R.fs_over_c;
    const double gvx = ku * ux + g0 * (kr * orx - ks * qx);
    const double gvy = ku * uy + g0 * (kr * ory - ks * qy);
    const double gvz = ku * uz + g0 * (kr * orz - ks * qz);
    acc[6] += sx * gvx;
    acc[7] += sy * gvy;
    acc[8] += sz * gvz;
    acc[9] -= gvx;
    acc[10] -= gvy;
    acc[11] -= gvz;
}

// adj_chain with omnidirectional source and receiver (no directivity, no orientation terms;
// id = 1 / d):
// with amp = att / (4 pi d), d amp / d v = -amp v / d^2 and d tau / d v = (fs/c) v / d, the
// position terms are g_v = -amp (G0 / d + G1 fs/c) v / d.
__device__ __forceinline__ void adj_chain_omni(const RoomK& R, int nx, int ny, int nz, const double* offs,
                                               const float* F, const float* Flo, const float* Fhi, float G0, float G1,
                                               double acc[12], double id) {
    const int Nx = R.Nx, Ny = R.Ny, Nz = R.Nz;
    const int ey = Nx + ny, ez = Nx + Ny + nz;
    const double Fx = F[nx], Fy = F[ey], Fz = F[ez];
    const double g0 = (double)G0, g1 = (double)G1;
    const double a0 = id * 0.079577471545947668;  // 1 / (4 pi d)
    const double gbeta = g0 * a0;
    const double Fxy = Fx * Fy;
    const double gyz = gbeta * (Fy * Fz), gxz = gbeta * (Fx * Fz), gxy = gbeta * Fxy;
    acc[0] += gyz * (double)Flo[nx];
    acc[1] += gyz * (double)Fhi[nx];
    acc[2] += gxz * (double)Flo[ey];
    acc[3] += gxz * (double)Fhi[ey];
    acc[4] += gxy * (double)Flo[ez];
    acc[5] += gxy * (double)Fhi[ez];
    const double amp = Fxy * Fz * a0;
    const double k = -amp * fma(g1, R.fs_over_c, g0 * id) * id;
    const double gvx = k * offs[nx], gvy = k * offs[ey], gvz = k * offs[ez];
    acc[6] += ((nx - Nx / 2) & 1) ? -gvx : gvx;
    acc[7] += ((ny - Ny / 2) & 1) ? -gvy : gvy;
    acc[8] += ((nz - Nz / 2) & 1) ? -gvz : gvz;
    acc[9] -= gvx;
    acc[10] -= gvy;
    acc[11] -= gvz;
}

struct AdjImg {
    double ox, oy, oz;  // unit source / receiver orientation (receiver: or*, source: os*)
};

template <bool MASK>
__global__ __launch_bounds__(WG) void ism_adjoint_kernel(AdjArgs a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ double red[NWAVE][12];
    __shared__ int wcnt[NWAVE];
    const int bid = a.order[blockIdx.x];
    const Item it = a.items[bid];
    const RoomK R = load_room(a.rooms, it.room, a.prm);
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int ntab = a.ntab_max;
    double* offs = reinterpret_cast<double*>(smem_raw);
    float* F = reinterpret_cast<float*>(offs + ntab);
    float* Flo = F + ntab;
    float* Fhi = Flo + ntab;
    float* gp = Fhi + ntab;                                  // [gp_len] g' around the tile
    int* qi = reinterpret_cast<int*>(gp + a.gp_len);         // [AQ] image index
    int* qn = qi + AQ;                                       // [AQ] g' index of tap k = J0
    float* qf = reinterpret_cast<float*>(qn + AQ);           // [AQ] fraction

    const int s = it.rir / R.n_rcv, r = it.rir - s * R.n_rcv;
    const int g_rir = R.rir_off + it.rir;
    const float* ps = a.pos_src + 3 * (R.src_off + s);
    const float* pr = a.pos_rcv + 3 * (R.rcv_off + r);
    for (int n = tid; n < R.Nx; n += WG)
        axis_entry_adj(n, R.Nx, R.L0, ps[0], pr[0], R.b0, R.b1, offs, F, Flo, Fhi);
    for (int n = tid; n < R.Ny; n += WG)
        axis_entry_adj(n, R.Ny, R.L1, ps[1], pr[1], R.b2, R.b3, offs + R.Nx, F + R.Nx, Flo + R.Nx, Fhi + R.Nx);
    for (int n = tid; n < R.Nz; n += WG) {
        const int o = R.Nx + R.Ny;
        axis_entry_adj(n, R.Nz, R.L2, ps[2], pr[2], R.b4, R.b5, offs + o, F + o, Flo + o, Fhi + o);
    }
    // g'(t) on [tb, tb + gp_len): dL/dout plus the envelope-window term, zero outside the tile
    const int J1 = a.J0 + a.nJ - 1;
    const int tb = it.t_lo - J1 + a.J0;
    {
        const long long row = R.out_off + (long long)it.rir * R.nSamples;
        const float cw = a.rir_cw[g_rir];
        const int ws = a.rir_ws[g_rir];
        const int lo = max(it.t_lo, 0), hi = min(it.t_hi, R.nISM);
        for (int u = tid; u < a.gp_len; u += WG) {
            const int t = tb + u;
            float g = 0.0f;
            if (t >= lo && t < hi) {
                g = a.gy[row + t];
                if (t >= ws) g = fmaf(cw, a.y[row + t], g);
            }
            gp[u] = (t & 1) ? g : -g;  // (-1)^(t+1) g'(t): see adj_tap
        }
    }
    // unit orientations (sigma applied per image for the source)
    double orx = 0, ory = 0, orz = 0, osx = 0, osy = 0, osz = 0;
    if (a.mic != 0) {
        const float* o = a.orv_rcv + 3 * (R.rcv_off + r);
        const double n = sqrt((double)o[0] * o[0] + (double)o[1] * o[1] + (double)o[2] * o[2]);
        orx = o[0] / n;
        ory = o[1] / n;
        orz = o[2] / n;
    }
    if (a.spkr != 0) {
        const float* o = a.orv_src + 3 * (R.src_off + s);
        const double n = sqrt((double)o[0] * o[0] + (double)o[1] * o[1] + (double)o[2] * o[2]);
        osx = o[0] / n;
        osy = o[1] / n;
        osz = o[2] / n;
    }
    __syncthreads();

    const int lo_n0 = it.t_lo - J1, hi_n0 = it.t_hi - 1 - a.J0;
    double acc[12];
#pragma unroll
    for (int k = 0; k < 12; k++) acc[k] = 0.0;
    const int u0 = -a.J0;  // ktab index of k = 0 (J0 <= 0 < 1 <= J1)

    auto process = [&](int e) {
        const int i = qi[e], nrel = qn[e];
        const float f = qf[e];
        const float S = sinpif(f) * 0.31830988618379067f, C = cospif(f);
        const float ph = f * a.inv_tw2;
        const float cph = cospif(ph), sph = sinpif(ph);
        const float hc = 0.5f * cph, hs = 0.5f * sph, pc = a.pi_tw * cph, pss = a.pi_tw * sph;
        const float al = S * pss - C * hc, be = -S * pc - C * hs, ga = -0.5f * C;
        const float* gk = gp + nrel;
        // (-1)^n0 (gk[u] = (-1)^n0 sigma_k g'); sigma_0 = -1, sigma_1 = +1 for the near taps
        const float pn = ((nrel - a.J0 + tb) & 1) ? -1.0f : 1.0f;
        float G0 = 0.0f, G0s = 0.0f, G1 = 0.0f, G1f = 0.0f;
#pragma unroll 4
        for (int u = 0; u < u0; u++)
            adj_tap<MASK>(a.ktab[u], gk[u], f, S, hc, hs, al, be, ga, a.halfTw, G0s, G1f);
        adj_tap_near<MASK>(a.ktab[u0], -pn * gk[u0], f, S, C, hc, hs, pc, pss, a.halfTw, G0, G1);
        adj_tap_near<MASK>(a.ktab[u0 + 1], pn * gk[u0 + 1], f, S, C, hc, hs, pc, pss, a.halfTw, G0, G1);
#pragma unroll 4
        for (int u = u0 + 2; u < a.nJ; u++)
            adj_tap<MASK>(a.ktab[u], gk[u], f, S, hc, hs, al, be, ga, a.halfTw, G0s, G1f);
        G0 = fmaf(S * pn, G0s, G0);
        G1 = fmaf(pn, G1f, G1);

        // chain rule through amp(beta, v) and tau(v) in FP64
        const int Nx = R.Nx, Ny = R.Ny, Nz = R.Nz;
        const int nx = (int)(((double)i + 0.5) * R.inv_nyz);
        const int rem = i - nx * (Ny * Nz);
        const int ny = (int)(((double)rem + 0.5) * R.inv_nz);
        const int nz = rem - ny * Nz;
        const double dx = offs[nx], dy = offs[Nx + ny], dz = offs[Nx + Ny + nz];
        adj_chain(R, nx, ny, nz, offs, F, Flo, Fhi, orx, ory, orz, osx, osy, osz, a.spkr, a.mic, G0, G1, acc,
                  sqrt_normal(dx * dx + dy * dy + dz * dz));
    };

    // image rounds of WG: phase A lists the images whose taps reach the tile into a
    // ring queue in a fixed order (wave by wave), phase B processes WG entries at a time
    int head = 0, n = 0;  // ring head and live entries (uniform)
    for (int r0 = it.i0; r0 < it.i1; r0 += WG) {
        const int i = r0 + tid;
        bool live = false;
        int n0 = 0;
        float f = 0.0f;
        if (i < it.i1) {
            const int Nyz = R.Ny * R.Nz;
            const int nx = (int)(((double)i + 0.5) * R.inv_nyz);
            const int rem = i - nx * Nyz;
            const int ny = (int)(((double)rem + 0.5) * R.inv_nz);
            const int nz = rem - ny * R.Nz;
            const double vx = offs[nx], vy = offs[R.Nx + ny], vz = offs[R.Nx + R.Ny + nz];
            double d2 = vx * vx;
            d2 = d2 + vy * vy;
            d2 = d2 + vz * vz;
            const double tau = sqrt(d2) * R.fs_over_c;
            const double fl = floor(tau);
            n0 = (int)fl;
            f = (float)(tau - fl);
            live = n0 >= lo_n0 && n0 <= hi_n0;
        }
        const unsigned long long mask = __ballot(live);
        if (lane == 0) wcnt[wave] = __popcll(mask);
        __syncthreads();
        int base = head + n, total = 0;
        for (int w = 0; w < NWAVE; w++) {
            const int c = wcnt[w];
            if (w < wave) base += c;
            total += c;
        }
        if (live) {
            const int pre = __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                                      __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
            const int q = (base + pre) & (AQ - 1);
            qi[q] = i;
            qn[q] = n0 + a.J0 - tb;
            qf[q] = f;
        }
        n += total;
        __syncthreads();
        const bool last = r0 + WG >= it.i1;
        while (n >= WG || (last && n > 0)) {
            const int take = min(n, WG);
            if (tid < take) process((head + tid) & (AQ - 1));
            head = (head + take) & (AQ - 1);
            n -= take;
            __syncthreads();  // slots of this pass are rewritten by the next round
        }
    }

    // fixed-order reduction: lanes, then waves
#pragma unroll
    for (int k = 0; k < 12; k++) {
        double v = acc[k];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
        if (lane == 0) red[wave][k] = v;
    }
    __syncthreads();
    if (tid < 12) {
        double t = 0.0;
        for (int w = 0; w < NWAVE; w++) t += red[w][tid];
        a.item_part[(long long)bid * 12 + tid] = t;
    }
}


// ---------------------------------------------------------------------------
// adjoint, basis method
// ---------------------------------------------------------------------------
// With h_j(f) ~ sum_m C[m][j] T_m(u), u = 2 f - 1 (ADJ_NB = 10 terms: the f-derivative of
// the 8-term expansion of the forward kernel is only within 6e-5 of h', of 10 terms within
// 5e-7), an image at tau = n0 + f has
//   G0 = sum_t g'(t) h(t - tau)  = sum_m T_m(u) E_m[n0],   E_m[n] = sum_j C[m][j] g'(n + j),
//   G1 = sum_t g'(t) h'(t - tau) = -2 sum_m m U_{m-1}(u) E_m[n0]
// (h'(x) = -d h_j / df at x = j - f, du/df = 2, T_m' = m U_{m-1}).  Per tile of 768 delays
// (512 where two workgroups would not fit a CU's LDS) the workgroup builds g' around the
// tile (the images of the tile see their whole window),
// correlates it with the coefficient rows on the matrix cores (E; both operands as
// FP16 pairs x = hi + lo 2^-11, three products hi hi, hi lo, lo hi per K step, 22 bits of
// each operand), enumerates the images of the tile's delay range as the forward basis
// kernel does, and sends each live image's G0, G1 through adj_chain into per-lane FP64 sums,
// reduced per item in a fixed order.
constexpr int ADJ_NB = 10;

// PAT: a source or receiver directivity other than omni (without it, the orientations and
// the pattern terms of the chain rule are compiled out).
template <bool PAT>
__global__ __launch_bounds__(WGB) __attribute__((amdgpu_waves_per_eu(4))) void ism_adjoint_basis_kernel(AdjArgs a) {
    constexpr int NB = ADJ_NB;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ double red[NWB][12];
    const int bid = a.order[blockIdx.x];
    const Item it = a.items[bid];
    const RoomK R = load_room(a.rooms, it.room, a.prm);
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int ntab = a.ntab_max, DL = a.adl, GL = a.agl, nJ = a.nJ, J0 = a.J0, J1 = a.J0 + a.nJ - 1;
    const int KS = a.aks, CL = a.acl;

    double* offs = reinterpret_cast<double*>(smem_raw);  // [ntab]
    float* F = reinterpret_cast<float*>(offs + ntab);    // [ntab] attenuation and its beta derivatives
    float* Flo = F + ntab;
    float* Fhi = Flo + ntab;
    int* xrun = reinterpret_cast<int*>(Fhi + ntab);      // [3][ntab + 4] per x entry: y runs (n0, count, n1)
    int* rowStart = xrun + 3 * (ntab + 4);               // [ntab + 4]
    const int GP = GL + GL / 2;
    // [2][GP] g'(lo_n0 + J0 + u) 2^sg as FP16 pairs (hi, then lo) at u + 8 (u >> 4)
    unsigned short* gph = reinterpret_cast<unsigned short*>(rowStart + ntab + 4);
    constexpr int NCP = 2;                               // shifted coefficient copies
    const int CLP = CL + NCP;
    unsigned short* chs = gph + 2 * GP;                  // [2][NCP][NB][CLP] coefficient FP16 pairs
    float* E = reinterpret_cast<float*>(chs + 2 * NCP * NB * CLP);  // [DL][NB] E_m[lo_n0 + d] at d NB + m
    double* rb2 = reinterpret_cast<double*>(E + NB * DL);  // [WGB] row x^2 + y^2
    int* rxy = reinterpret_cast<int*>(rb2 + WGB);        // [WGB] row nx | ny << 16
    int* queue = rxy + WGB;                              // [QB] run marks
    int* misc = queue + QB;                              // [32] scan scratch (two buffers at 16, 24)
    constexpr int QEPT = QB / WGB;                       // most consecutive queue entries per thread
    for (int e = tid; e < NCP * NB * CLP / 2; e += WGB)  // 2 NCP NB CLP halves in 8-byte words
        reinterpret_cast<uint2*>(chs)[e] = reinterpret_cast<const uint2*>(a.achs)[e];
    for (int e = tid; e < QB; e += WGB) queue[e] = 0;    // run marks: zero between rounds
    const int s = it.rir / R.n_rcv, r = it.rir - s * R.n_rcv;
    const int g_rir = R.rir_off + it.rir;
    const float* ps = a.pos_src + 3 * (R.src_off + s);
    const float* pr = a.pos_rcv + 3 * (R.rcv_off + r);
    for (int n = tid; n < R.Nx; n += WGB)
        axis_entry_adj(n, R.Nx, R.L0, ps[0], pr[0], R.b0, R.b1, offs, F, Flo, Fhi);
    for (int n = tid; n < R.Ny; n += WGB)
        axis_entry_adj(n, R.Ny, R.L1, ps[1], pr[1], R.b2, R.b3, offs + R.Nx, F + R.Nx, Flo + R.Nx, Fhi + R.Nx);
    for (int n = tid; n < R.Nz; n += WGB) {
        const int o = R.Nx + R.Ny;
        axis_entry_adj(n, R.Nz, R.L2, ps[2], pr[2], R.b4, R.b5, offs + o, F + o, Flo + o, Fhi + o);
    }
    double orx = 0, ory = 0, orz = 0, osx = 0, osy = 0, osz = 0;
    const int spkr = PAT ? a.spkr : 0, mic = PAT ? a.mic : 0;
    if (mic != 0) {
        const float* o = a.orv_rcv + 3 * (R.rcv_off + r);
        const double n = sqrt((double)o[0] * o[0] + (double)o[1] * o[1] + (double)o[2] * o[2]);
        orx = o[0] / n;
        ory = o[1] / n;
        orz = o[2] / n;
    }
    if (spkr != 0) {
        const float* o = a.orv_src + 3 * (R.src_off + s);
        const double n = sqrt((double)o[0] * o[0] + (double)o[1] * o[1] + (double)o[2] * o[2]);
        osx = o[0] / n;
        osy = o[1] / n;
        osz = o[2] / n;
    }
    const AxisRuns AY = axis_runs(R.Ny, R.L1, ps[1], pr[1]);
    const AxisRuns AZ = axis_runs(R.Nz, R.L2, ps[2], pr[2]);
    const int Nx = R.Nx, Ny = R.Ny, Nz = R.Nz;
    const double fsc = R.fs_over_c;
    const long long row_off = R.out_off + (long long)it.rir * R.nSamples;
    const float cw = a.rir_cw[g_rir];
    const int ws = a.rir_ws[g_rir];
    constexpr int BIG = 1 << 30;
    int scan_par = 0;  // block_scan_excl1 buffer parity (uniform)
    double acc[12];
#pragma unroll
    for (int k = 0; k < 12; k++) acc[k] = 0.0;
    __syncthreads();

    // The item's output samples [t_lo, t_hi) name its delays: n0 in [t_lo - J1, t_hi - J1), the
    // RIR's last item through nISM - 1 - J0 (the last delay with a tap in the ISM part), so
    // the items of a RIR partition its images and each image is evaluated once, in the tile
    // holding its n0, against g' over its whole window.
    const int n0_end = it.t_hi >= R.nISM ? R.nISM - J0 : it.t_hi - J1;  // exclusive
    const int ABT = a.abt;
    for (int tl = it.t_lo; tl < it.t_hi; tl += ABT) {
        const int lo_n0 = tl - J1, hi_n0 = tl + ABT >= it.t_hi ? n0_end - 1 : tl + ABT - J1 - 1;
        const int nE = hi_n0 - lo_n0 + 1;  // E positions of the tile (<= ABT + nJ - 1 <= DL)
        // g' around the tile: dL/dout plus the envelope-window term on the ISM part, zero
        // outside it (GL <= 4 WGB)
        float gv[4];
        float gm = 0.0f;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int u = tid + i * WGB;
            const int t = lo_n0 + J0 + u;
            float g = 0.0f;
            if (u < GL && t >= 0 && t < R.nISM) {
                g = a.gy[row_off + t];
                if (t >= ws) g = fmaf(cw, a.y[row_off + t], g);
            }
            gv[i] = g;
            gm = fmaxf(gm, fabsf(g));
        }
        // the tile's largest |g'| (non-negative floats order as their bit patterns) scales
        // g' into [2^14, 2^15) for the FP16 pairs
        {
            const int wm = wave_max_nonneg(__float_as_int(gm));
            if (lane == 63) misc[wave] = wm;
        }
        __syncthreads();
        int gmb = 0;
#pragma unroll
        for (int w = 0; w < NWB; w++) gmb = max(gmb, misc[w]);
        const float gmax = __int_as_float(gmb);
        const int sg = gmax > 0.0f ? min(max(14 - ilogbf(gmax), -100), 100) : 0;
        {
            const float gs = ldexpf(1.0f, sg);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int u = tid + i * WGB;
                if (u < GL) {
                    const float x = gv[i] * gs;
                    const _Float16 hi = (_Float16)x;
                    gph[u + 8 * (u >> 4)] = __builtin_bit_cast(unsigned short, hi);
                    gph[GP + u + 8 * (u >> 4)] = __builtin_bit_cast(unsigned short, (_Float16)((x - (float)hi) * 2048.0f));
                }
            }
        }
        __syncthreads();
        // E_m[lo_n0 + d] = sum_k C[m][k] g'(lo_n0 + d + J0 + k): E[16 b + r] = sum_x A[r][x] X[x][b]
        // with A[r][x] = C[x - r] (Toeplitz) and X[x][b] = g'[16 b + x] (Hankel), the forward
        // basis kernel's products with the coefficients unreversed, on v_mfma_f32_16x16x32_f16;
        // one task per (rows m and m + NB/2, 256 positions of E), the waves taking tasks in
        // turn.  Lane (c, g) holds K indices g KS + 8 s + e (e < 8) of step s in both operands:
        // coefficients C[g KS + 8 s + e - c] from copy c & 1 (4-byte aligned reads), g' at
        // 16 c + g KS + 8 s + e (one 16-byte read).
        {
            const int c = lane & 15, g = lane >> 4, rho = c & (NCP - 1);
            const int nota = (nE + 255) / 256, nst = KS / 8;
            constexpr int MH = NB / 2;
            static_assert(NB % 2 == 0, "row pairs");
            const int LO = NCP * NB * CLP;                   // halves from the hi to the lo copies
            const float oscale = ldexpf(1.0f, -(14 + sg));  // coefficients carry 2^14, g' 2^sg
            auto ld_coef = [&](const unsigned short* p) -> half8_t {
                const unsigned* q = reinterpret_cast<const unsigned*>(p);
                uint4 v;
                v.x = q[0];
                v.y = q[1];
                v.z = q[2];
                v.w = q[3];
                return __builtin_bit_cast(half8_t, v);
            };
            // lanes c >= 8 of each row of 16 get v from lane c - 8, lanes c < 8 keep f
            auto shr8_or = [](half8_t f, half8_t v) -> half8_t {
                uint4 fu = __builtin_bit_cast(uint4, f);
                const uint4 vu = __builtin_bit_cast(uint4, v);
                fu.x = __builtin_amdgcn_update_dpp(fu.x, vu.x, 0x118, 0xf, 0xf, false);  // row_shr:8
                fu.y = __builtin_amdgcn_update_dpp(fu.y, vu.y, 0x118, 0xf, 0xf, false);
                fu.z = __builtin_amdgcn_update_dpp(fu.z, vu.z, 0x118, 0xf, 0xf, false);
                fu.w = __builtin_amdgcn_update_dpp(fu.w, vu.w, 0x118, 0xf, 0xf, false);
                return __builtin_bit_cast(half8_t, fu);
            };
            for (int task = wave; task < MH * nota; task += NWB) {
                const int mp = task / nota, ot = task - mp * nota;
                const unsigned short* A0 = chs + (rho * NB + mp) * CLP + g * KS + 16 - c + rho;
                const unsigned short* A1 = A0 + MH * CLP;
                const int kb = 256 * ot + 16 * c + g * KS;
                floatx4 h0 = {0.0f, 0.0f, 0.0f, 0.0f}, x0 = h0, h1 = h0, x1 = h0;
                half8_t a0h = ld_coef(A0), a0l = ld_coef(A0 + LO), a1h = ld_coef(A1), a1l = ld_coef(A1 + LO);
                half8_t xh = *reinterpret_cast<const half8_t*>(gph + kb + 8 * (kb >> 4));
                half8_t xl = *reinterpret_cast<const half8_t*>(gph + GP + kb + 8 * (kb >> 4));
                for (int st = 0; st < nst; st++) {
                    // operands of step st + 1 are read before the products of step st
                    half8_t n0h = a0h, n0l = a0l, n1h = a1h, n1l = a1l, nxh = xh, nxl = xl;
                    if (st + 1 < nst) {
                        const int o = 8 * (st + 1), dn = kb + o;
                        // the Toeplitz operand of lane c at step st + 1 is lane c - 8's at step
                        // st: lanes c >= 8 take it by a row shift, only lanes c < 8 read LDS
                        half8_t f0h = a0h, f0l = a0l, f1h = a1h, f1l = a1l;
                        if (c < 8) {
                            f0h = ld_coef(A0 + o);
                            f0l = ld_coef(A0 + LO + o);
                            f1h = ld_coef(A1 + o);
                            f1l = ld_coef(A1 + LO + o);
                        }
                        n0h = shr8_or(f0h, a0h);
                        n0l = shr8_or(f0l, a0l);
                        n1h = shr8_or(f1h, a1h);
                        n1l = shr8_or(f1l, a1l);
                        nxh = *reinterpret_cast<const half8_t*>(gph + dn + 8 * (dn >> 4));
                        nxl = *reinterpret_cast<const half8_t*>(gph + GP + dn + 8 * (dn >> 4));
                    }
                    h0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(a0h, xh, h0, 0, 0, 0);
                    x0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(a0h, xl, x0, 0, 0, 0);
                    h1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(a1h, xh, h1, 0, 0, 0);
                    x1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(a1h, xl, x1, 0, 0, 0);
                    x0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(a0l, xh, x0, 0, 0, 0);
                    x1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(a1l, xh, x1, 0, 0, 0);
                    a0h = n0h;
                    a0l = n0l;
                    a1h = n1h;
                    a1l = n1l;
                    xh = nxh;
                    xl = nxl;
                }
                // element (row 4 g + i, column c) is position 256 ot + 16 c + 4 g + i
                const int d0 = 256 * ot + 16 * c + 4 * g;
#pragma unroll
                for (int i = 0; i < 4; i++)
                    if (d0 + i < nE) {
                        E[(d0 + i) * NB + mp] = (h0[i] + x0[i] * (1.0f / 2048.0f)) * oscale;
                        E[(d0 + i) * NB + mp + MH] = (h1[i] + x1[i] * (1.0f / 2048.0f)) * oscale;
                    }
            }
        }
        // rows: per x entry, the y entries with x^2 + y^2 <= d2hi (one run per parity)
        const double rlo = (double)(lo_n0 - 1) / fsc, rhi = (double)(hi_n0 + 2) / fsc;
        const double d2lo = lo_n0 >= 1 ? rlo * rlo : -1.0, d2hi = rhi * rhi;
        int carry = 0;
        for (int ix0 = 0; ix0 < Nx; ix0 += WGB) {
            const int nx = ix0 + tid;
            int c = 0;
            if (nx < Nx) {
                const double X = offs[nx];
                const double rem = d2hi - X * X;
                int f0 = 0, c0 = 0, f1 = 0, c1 = 0;
                if (rem >= 0.0) {
                    const double Y = sqrt_bound(rem);
                    axis_range(AY, 0, -Y, Y, f0, c0, BIG);
                    axis_range(AY, 1, -Y, Y, f1, c1, BIG);
                }
                xrun[nx] = f0;
                xrun[(ntab + 4) + nx] = c0;
                xrun[2 * (ntab + 4) + nx] = f1;
                c = c0 + c1;
            }
            int tot;
            const int ex = block_scan_excl1(c, misc + 16 + 8 * (scan_par++ & 1), &tot);
            if (nx < Nx) rowStart[nx + 1] = carry + ex + c;
            carry += tot;
        }
        if (tid == 0) rowStart[0] = 0;
        __syncthreads();
        const int nrows = carry;
        for (int r0 = 0; r0 < nrows; r0 += WGB) {
            const int row = r0 + tid;
            int cnt = 0;
            int zf[4] = {0, 0, 0, 0}, zc[4] = {0, 0, 0, 0};
            if (row < nrows) {
                int lo = 0, hi = Nx;
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (rowStart[mid] <= row) lo = mid;
                    else hi = mid;
                }
                const int nx = lo, j = row - rowStart[lo];
                const int c0 = xrun[(ntab + 4) + nx];
                const int ny = j < c0 ? xrun[nx] + 2 * j : xrun[2 * (ntab + 4) + nx] + 2 * (j - c0);
                const double vx = offs[nx], vy = offs[Nx + ny];
                double b2 = vx * vx;
                b2 = b2 + vy * vy;
                const double V2 = d2hi - b2, U2 = d2lo - b2;
                if (V2 >= 0.0) {
                    const double V = sqrt_bound(V2);
                    if (U2 <= 0.0) {
                        axis_range(AZ, 0, -V, V, zf[0], zc[0], BIG);
                        axis_range(AZ, 1, -V, V, zf[2], zc[2], BIG);
                    } else {
                        const double U = sqrt_bound(U2);
                        axis_range(AZ, 0, U, V, zf[1], zc[1], BIG);
                        axis_range(AZ, 0, -V, -U, zf[0], zc[0], (zf[1] - AZ.half) / 2 - 1);
                        axis_range(AZ, 1, U, V, zf[3], zc[3], BIG);
                        axis_range(AZ, 1, -V, -U, zf[2], zc[2], (zf[3] - AZ.half + 1) / 2 - 1);
                    }
                }
                cnt = zc[0] + zc[1] + zc[2] + zc[3];
                rb2[tid] = b2;
                rxy[tid] = nx | (ny << 16);
            }
            int tot;
            const int ex = block_scan_excl1(cnt, misc + 16 + 8 * (scan_par++ & 1), &tot);
            for (int q0 = 0; q0 < tot; q0 += QB) {
                // run marks as in the forward basis kernel: entry p of the round (run g of row
                // slot, run start s) has nz = base + 2 p, base = zf[g] - 2 (s - q0); each run
                // writes its key (slot, g, base + 4096) at its first entry in the round and at
                // every wave segment start it covers, and an entry's key is the largest one at
                // or before it.  A round gives each thread qe consecutive entries, as few as
                // spread it over all the threads.
                const int nq = min(QB, tot - q0);
                const int qe = (nq + WGB - 1) / WGB, qseg = 64 * qe;
                {
                    int s = ex;
#pragma unroll
                    for (int g = 0; g < 4; g++) {
                        const int a0 = max(s, q0) - q0, b0 = min(s + zc[g], q0 + QB) - q0;
                        if (a0 < b0) {
                            const int key = ((tid * 4 + g) << 13) | (zf[g] - 2 * (s - q0) + 4096);
                            queue[a0] = key;
                            for (int bnd = (a0 / qseg + 1) * qseg; bnd < b0; bnd += qseg) queue[bnd] = key;
                        }
                        s += zc[g];
                    }
                }
                __syncthreads();
                if (qseg * wave < nq) {
                    static_assert(QEPT == 4, "entry keys are selected from four registers");
                    int keys[QEPT];
                    int kmax = 0;
#pragma unroll
                    for (int i = 0; i < QEPT; i++) {
                        keys[i] = i < qe ? queue[qe * tid + i] : 0;
                        kmax = max(kmax, keys[i]);
                    }
#pragma unroll
                    for (int i = 0; i < QEPT; i++)
                        if (i < qe) queue[qe * tid + i] = 0;
                    int key = wave_excl_max_nonneg(kmax);
#pragma unroll
                    for (int i = 0; i < QEPT; i++) {
                        key = max(key, keys[i]);
                        keys[i] = key;
                    }
                    for (int i = 0; i < qe; i++) {
                        const int pos = qe * tid + i;
                        if (pos >= nq) break;
                        const int ki = i == 0 ? keys[0] : i == 1 ? keys[1] : i == 2 ? keys[2] : keys[3];
                        const int slot = ki >> 15, nz = (ki & 8191) - 4096 + 2 * pos;
                        const double vz = offs[Nx + Ny + nz];
                        double d2 = rb2[slot];
                        d2 = d2 + vz * vz;
                        double idist;
                        const double dist = sqrt_delay_inv(d2, idist);
                        const double tau = dist * fsc;
                        const double fl = floor(tau);
                        const int n0 = (int)fl;
                        if (n0 < lo_n0 || n0 > hi_n0) continue;
                        const float u = fmaf(2.0f, (float)(tau - fl), -1.0f);
                        const float u2 = u + u;
                        // E_0 .. E_9 of the delay: five 8-byte reads
                        const float2* ep = reinterpret_cast<const float2*>(E + (n0 - lo_n0) * NB);
                        float e[NB];
#pragma unroll
                        for (int q = 0; q < NB / 2; q++) {
                            const float2 v = ep[q];
                            e[2 * q] = v.x;
                            e[2 * q + 1] = v.y;
                        }
                        // T_m (first kind) and m U_{m-1} (derivative) by their recurrences
                        float t0 = 1.0f, t1 = u, w0 = 1.0f, w1 = u2;  // T_0, T_1, U_0, U_1
                        float G0 = fmaf(t1, e[1], e[0]);
                        float D1 = e[1];  // 1 * U_0 * E_1
#pragma unroll
                        for (int m = 2; m < NB; m++) {
                            const float t2 = fmaf(u2, t1, -t0);
                            const float em = e[m];
                            G0 = fmaf(t2, em, G0);
                            D1 = fmaf((float)m * w1, em, D1);
                            const float w2 = fmaf(u2, w1, -w0);
                            t0 = t1;
                            t1 = t2;
                            w0 = w1;
                            w1 = w2;
                        }
                        const int xy = rxy[slot];
                        if constexpr (PAT)
                            adj_chain(R, xy & 0xffff, xy >> 16, nz, offs, F, Flo, Fhi, orx, ory, orz, osx, osy, osz,
                                      spkr, mic, G0, -2.0f * D1, acc, dist);
                        else
                            adj_chain_omni(R, xy & 0xffff, xy >> 16, nz, offs, F, Flo, Fhi, G0, -2.0f * D1, acc, idist);
                    }
                }
                __syncthreads();
            }
        }
        __syncthreads();
    }

    // fixed-order reduction: lanes, then waves
#pragma unroll
    for (int k = 0; k < 12; k++) {
        double v = acc[k];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
        if (lane == 0) red[wave][k] = v;
    }
    __syncthreads();
    if (tid < 12) {
        double t = 0.0;
        for (int w = 0; w < NWB; w++) t += red[w][tid];
        a.item_part[(long long)bid * 12 + tid] = t;
    }
}

struct TailAdjArgs {
    const RoomDev* rooms;
    const int* rir_room;
    const float* pos_src;
    const float* pos_rcv;
    const float* y;
    const float* gy;
    float* rir_cw;
    int* rir_ws;
    double* rir_tail;  // [n_rirs][13]: dbeta[6], dsrc[3], drcv[3], dlevel
    float w_sz;
    double fs;
    unsigned long long seed;
    unsigned call_offset;
    int noise;
    RoomParams prm;
};

__global__ __launch_bounds__(WG) void tail_adjoint_kernel(TailAdjArgs a) {
    __shared__ double red[NWAVE];
    const int g = blockIdx.x;
    const RoomK R = load_room(a.rooms, a.rir_room[g], a.prm);
    const int m = g - R.rir_off;
    const int s = m / R.n_rcv, r = m - s * R.n_rcv;
    const int tid = threadIdx.x;
    const int nISM = R.nISM, nS = R.nSamples, nDiff = nS - nISM;
    const long long row = R.out_off + (long long)m * nS;
    const float* y = a.y + row;
    const float* gy = a.gy + row;
    double* outp = a.rir_tail + 13LL * g;

    const float* ps = a.pos_src + 3 * (R.src_off + s);
    const float* pr = a.pos_rcv + 3 * (R.rcv_off + r);
    double v[3];
    for (int d = 0; d < 3; d++) v[d] = (0.0 + (double)ps[d]) - (double)pr[d];
    double d2 = v[0] * v[0];
    d2 = d2 + v[1] * v[1];
    d2 = d2 + v[2] * v[2];
    const double d0 = sqrt(d2);
    const double tau_dp = d0 * R.fs_over_c;
    const int w_start = (int)ceil(fmax((double)nISM - (double)a.w_sz, tau_dp));
    if (nDiff <= 0) {
        if (tid < 13) outp[tid] = 0.0;
        if (tid == 0) {
            a.rir_cw[g] = 0.0f;
            a.rir_ws[g] = nISM;
        }
        return;
    }
    double pw = 0.0;
    for (int t = w_start + tid; t < nISM; t += WG) {
        const double x = y[t];
        pw += x * x;
    }
    const double nw = (double)(nISM - w_start);
    const double P = block_sum(pw, red) / nw;
    const double L0 = R.L0, L1 = R.L1, L2 = R.L2;
    const double b[6] = {R.b0, R.b1, R.b2, R.b3, R.b4, R.b5};
    const double Sw[6] = {L1 * L2, L1 * L2, L0 * L2, L0 * L2, L0 * L1, L0 * L1};
    const double Sa = ((1.0 - b[0] * b[0]) + (1.0 - b[1] * b[1])) * L1 * L2 +
                      ((1.0 - b[2] * b[2]) + (1.0 - b[3] * b[3])) * L0 * L2 +
                      ((1.0 - b[4] * b[4]) + (1.0 - b[5] * b[5])) * L0 * L1;
    const double V = L0 * L1 * L2;
    const double T60 = 0.161 * V / Sa;
    const double alpha = -13.8155 / (T60 * a.fs);
    const double w_center = (double)w_start + (double)(nISM - w_start) / 2.0;
    const double A = P / exp(alpha * (w_center - tau_dp));
    const double sA = sqrt(A);
    const double level = (double)R.level;

    const unsigned k0 = (unsigned)a.seed, k1 = (unsigned)(a.seed >> 32);
    double s1 = 0.0, s2 = 0.0;
    const int nblk = (nDiff + 3) >> 2;
    // sA exp(alpha ts / 2) at the thread's sample, advanced by its ratio per sample and per
    // loop step (4 WG samples) instead of an FP64 exp per sample
    const double r1 = exp(0.5 * alpha), rstep = exp(0.5 * alpha * (4.0 * WG));
    double env_b = sA * exp(0.5 * alpha * ((double)(nISM + 4 * tid) - tau_dp));
    for (int blk = tid; blk < nblk; blk += WG) {
        unsigned w[4];
        philox4x32_10((unsigned)blk, (unsigned)m, R.room_id, a.call_offset, k0, k1, w);
        double env = env_b;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int sd = 4 * blk + u;
            if (sd < nDiff) {
                const double z = (double)tail_noise(w, u, a.noise);
                const double ts = (double)(nISM + sd) - tau_dp;
                const double q = env * z;  // tail / level
                const double gq = (double)gy[nISM + sd] * q;
                s1 += gq;
                s2 += gq * ts;
            }
            env *= r1;
        }
        env_b *= rstep;
    }
    s1 = block_sum(s1, red);
    s2 = 0.5 * level * block_sum(s2, red);
    const double gA = A > 0.0 ? level * s1 / (2.0 * A) : 0.0;
    const double galpha = s2 - gA * A * (w_center - tau_dp);
    const double gtau = -0.5 * alpha * level * s1 + gA * A * alpha;
    if (tid == 0) {
        a.rir_cw[g] = P > 0.0 ? (float)(gA * (A / P) * 2.0 / nw) : 0.0f;
        a.rir_ws[g] = w_start;
        // alpha = -13.8155 Sa / (0.161 V fs): d alpha / d b_k = 13.8155 * 2 b_k S_k / (0.161 V fs)
        const double ka = 13.8155 * 2.0 / (0.161 * V * a.fs);
        for (int k = 0; k < 6; k++) outp[k] = galpha * ka * b[k] * Sw[k];
        // tau_dp = |src - rcv| fs/c
        for (int d = 0; d < 3; d++) {
            const double gs = gtau * R.fs_over_c * v[d] / d0;
            outp[6 + d] = gs;
            outp[9 + d] = -gs;
        }
        outp[12] = s1;
    }
}

// Per RIR: its items' partial sums (items of one RIR are contiguous) plus the tail terms.
__global__ __launch_bounds__(WG) void grad_rir_kernel(const double* item_part, const int* rir_item_begin,
                                                      const double* rir_tail, double* rir_grad, int n_rirs) {
    const int g = blockIdx.x * WG + threadIdx.x;
    if (g >= n_rirs) return;
    double t[13];
    for (int k = 0; k < 13; k++) t[k] = 0.0;
    for (int e = rir_item_begin[g]; e < rir_item_begin[g + 1]; e++)
        for (int k = 0; k < 12; k++) t[k] += item_part[(long long)e * 12 + k];
    for (int k = 0; k < 13; k++) rir_grad[(long long)g * 13 + k] = t[k] + rir_tail[(long long)g * 13 + k];
}

// Per room: d beta[6], d level, and the rows of its sources and receivers.
__global__ __launch_bounds__(WG) void grad_param_kernel(const RoomDev* rooms, const double* rir_grad, float* g_beta,
                                                        float* g_src, float* g_rcv, float* g_level) {
    const RoomDev* R = rooms + blockIdx.x;
    const int ns = R->n_src, nr = R->n_rcv, ro = R->rir_off;
    const int n_out = 7 + 3 * ns + 3 * nr;
    for (int o = threadIdx.x; o < n_out; o += WG) {
        double t = 0.0;
        if (o < 7) {
            const int k = o < 6 ? o : 12;
            for (int m = 0; m < ns * nr; m++) t += rir_grad[(long long)(ro + m) * 13 + k];
            if (o < 6) {
                if (g_beta) g_beta[(long long)blockIdx.x * 6 + o] = (float)t;
            } else if (g_level) {
                g_level[blockIdx.x] = (float)t;
            }
        } else if (o < 7 + 3 * ns) {
            const int sidx = (o - 7) / 3, d = (o - 7) % 3;
            for (int rr = 0; rr < nr; rr++) t += rir_grad[(long long)(ro + sidx * nr + rr) * 13 + 6 + d];
            if (g_src) g_src[(long long)(R->src_off + sidx) * 3 + d] = (float)t;
        } else {
            const int ridx = (o - 7 - 3 * ns) / 3, d = (o - 7 - 3 * ns) % 3;
            for (int ss = 0; ss < ns; ss++) t += rir_grad[(long long)(ro + ss * nr + ridx) * 13 + 9 + d];
            if (g_rcv) g_rcv[(long long)(R->rcv_off + ridx) * 3 + d] = (float)t;
        }
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

struct TapTables {
    int J0 = 0, nJ = 0;        // offsets j that can be in the window
    int n_pass = 0;            // ceil(nJ / 64)
    unsigned mask_bits = 0;    // passes that need the |x| < halfTw test
    float halfTw = 0.0f;
    std::vector<float4> jtab;  // [n_pass * 64]: (sigma j, cos, sin, sigma)
    int lut_center = 0, lut_row = 0, lut_k_lo = 0;
    std::vector<float2> lut;   // polyphase [16][lut_row]
    std::vector<double> lut_tab;  // the table itself (FP64), lut_len entries
};

// Tap offsets j = t - floor(tau) that can lie inside the window, and their
// per-offset constants.  Exact mode: |j - f| < Tw/2 for some f in [0,1).  LUT
// mode: |j - f| < Tw/2 (integer division), src/gpuRIR_cuda.cu:637, with the table
// of src/gpuRIR_cuda.cu:578-590 (1/16-sample grid, odd length, zero outside).
TapTables make_tap_tables(int precision, int Tw) {
    TapTables T;
    const bool lut = precision == PREC_LUT;
    if (lut) {
        const int half = Tw / 2;
        T.J0 = -half + 1;
        T.nJ = 2 * half;
        T.halfTw = (float)half;
    } else {
        const double h = Tw / 2.0;
        T.halfTw = (float)h;
        T.J0 = (int)std::floor(-h) + 1;
        const int J1 = (int)std::ceil(h + 1.0) - 1;
        T.nJ = J1 - T.J0 + 1;
    }
    const int J1 = T.J0 + T.nJ - 1;
    // passes are instantiated for 1, 2, 3, 4, 6 and 12, and beyond 12 in groups of 12; other
    // counts are padded with masked lanes up to the next one
    const int need = (T.nJ + 63) / 64;
    T.n_pass = need <= 4 ? need : (need <= 6 ? 6 : 12 * ((need + 11) / 12));
    // a pass needs the window test if it holds padding lanes (j > J1) or, for an odd
    // window (half-integer Tw/2, or integer Tw/2 in LUT mode with odd Tw), the end
    // offsets J0, J1 whose membership depends on f.  For an even window the end
    // offsets are inside for every f, except j = J1 at f = 0 whose tap is exactly 0.
    const bool f_dependent_ends = (Tw % 2) != 0;
    for (int p = 0; p < T.n_pass; p++) {
        const int jlo = T.J0 + 64 * p, jhi = jlo + 63;
        const bool pad = jhi > J1;
        const bool ends = f_dependent_ends && (jlo <= T.J0 || jhi >= J1);
        if (pad || ends) T.mask_bits |= 1u << p;
    }
    T.jtab.resize((size_t)T.n_pass * 64);
    for (int u = 0; u < T.n_pass * 64; u++) {
        const int j = T.J0 + u;
        const double sg = (j & 1) ? 1.0 : -1.0;  // (-1)^(j+1)
        const double th = 2.0 * M_PI * (double)j / (double)Tw;
        T.jtab[u] = make_float4((float)(sg * j), (float)std::cos(th), (float)std::sin(th), (float)sg);
    }
    if (lut) {
        int len = Tw * 16;
        len += (len % 2) ? 0 : 1;
        const int center = len / 2;
        std::vector<double> tab(len + 1, 0.0);
        for (int i = 0; i <= center; i++) {
            const double x = (double)i / 16.0;
            const double sn = (i == 0) ? 1.0 : std::sin(M_PI * x) / (M_PI * x);
            const double h = 0.5 * (1.0 + std::cos(2.0 * M_PI * x / Tw));
            tab[center + i] = tab[center - i] = h * sn;
        }
        // table index of lane j: i = 16 j + center + k0, k0 = floor(-16 f) in [-16, 0];
        // row k = floor(i / 16) spans [k_lo, k_hi] over all lanes of all passes
        auto fdiv = [](int a, int b) { return (a >= 0) ? a / b : -((-a + b - 1) / b); };
        const int i_min = 16 * T.J0 + center - 16;
        const int i_max = 16 * (T.J0 + 64 * T.n_pass - 1) + center;
        T.lut_center = center;
        T.lut_k_lo = fdiv(i_min, 16);
        T.lut_row = fdiv(i_max, 16) - T.lut_k_lo + 2;
        T.lut.assign((size_t)16 * T.lut_row, make_float2(0.0f, 0.0f));
        T.lut_tab.assign(tab.begin(), tab.begin() + len);
        auto val = [&](int i) { return (i >= 0 && i < len) ? tab[i] : 0.0; };  // zero border
        for (int ph = 0; ph < 16; ph++)
            for (int k = 0; k < T.lut_row; k++) {
                const int i = 16 * (T.lut_k_lo + k) + ph;
                T.lut[(size_t)ph * T.lut_row + k] = make_float2((float)val(i), (float)(val(i + 1) - val(i)));
            }
    }
    return T;
}

struct Plan {
    std::vector<RoomDev> rooms;
    std::vector<Item> items;
    std::vector<int> rir_room;
    std::vector<int2> sum_blocks;  // (global RIR, first sample) of 64-sample partial-sum blocks
    long long part_elems = 0;
    int ntab_max = 0;
    int tt_max = 0;
};

constexpr int MIN_CHUNK = 512;
constexpr int MAX_CHUNK = 32768;      // images per work item by default at most (splits the largest rooms)
constexpr int TARGET_ITEMS = 2048;
constexpr int ADJ_TARGET_ITEMS = 512;  // basis adjoint
constexpr long long AUTO_BASIS_MIN_TILES = 256;  // METHOD_AUTO: fewer basis tiles -> tap scatter
constexpr size_t LDS_PER_WG = 53 * 1024;  // three workgroups per CU (160 KiB of LDS)
constexpr int MIN_TILE = 512;

// LDS bytes of one accumulation workgroup for a tile of tt samples
size_t acc_lds_bytes(int tt, int ntab, const TapTables& T) {
    const size_t nt = (size_t)(ntab + 3) / 4 * 4;
    const size_t acc = ((size_t)tt + (size_t)(T.nJ - 1) + (size_t)T.n_pass * 64 + 3) / 4 * 4;
    const size_t lut = T.n_pass <= 12 ? sizeof(float2) * T.lut.size() : 0;  // grouped: read from global memory
    return 12 * nt + 8 * acc + (size_t)(QCAP + QPAD) * 20 + 32 + lut;
}

Plan make_plan(const std::vector<RoomHost>& rh, const SimOptions& opt, const TapTables& Tt, bool sort_items) {
    Plan P;
    int ntab = 0, nism_max = 0;
    for (const RoomHost& h : rh) {
        ntab = std::max(ntab, h.nb[0] + h.nb[1] + h.nb[2]);
        nism_max = std::max(nism_max, n_samples_rule(h.Tdiff, opt.fs));
    }
    // default tile: the whole early part if a workgroup then still fits three per CU
    int tile_cap = opt.tile_samples;
    if (tile_cap <= 0) {
        tile_cap = std::max(nism_max, 1);
        while (tile_cap > MIN_TILE && acc_lds_bytes(tile_cap, ntab, Tt) > LDS_PER_WG) {
            const int ntiles = (nism_max + tile_cap - 1) / tile_cap + 1;
            tile_cap = std::max(MIN_TILE, (nism_max + ntiles - 1) / ntiles);
        }
    }
    double total = 0.0;
    long long total_rirs = 0;
    for (const RoomHost& h : rh) {
        const long long N = (long long)h.nb[0] * h.nb[1] * h.nb[2];
        const int nISM = n_samples_rule(h.Tdiff, opt.fs);
        const int tiles = nISM > 0 ? (nISM + tile_cap - 1) / tile_cap : 0;
        total += (double)h.n_src * h.n_rcv * (double)N * tiles;
        total_rirs += (long long)h.n_src * h.n_rcv;
    }
    long long per_item = (long long)std::ceil(total / (double)TARGET_ITEMS);
    per_item = std::min<long long>(std::max<long long>(per_item, MIN_CHUNK), MAX_CHUNK);
    if (opt.chunk_images > 0) per_item = std::max(opt.chunk_images, 1);
    long long rir_off = 0;
    for (size_t k = 0; k < rh.size(); k++) {
        const RoomHost& h = rh[k];
        RoomDev R;
        std::memset(&R, 0, sizeof(R));
        for (int d = 0; d < 3; d++) {
            R.L[d] = h.L[d];
            R.nb[d] = h.nb[d];
        }
        for (int d = 0; d < 6; d++) R.beta[d] = h.beta[d];
        R.fs_over_c = (double)opt.fs / (double)opt.c;
        R.inv_nyz = 1.0 / ((double)h.nb[1] * (double)h.nb[2]);
        R.inv_nz = 1.0 / (double)h.nb[2];
        R.n_src = h.n_src;
        R.n_rcv = h.n_rcv;
        R.src_off = h.src_off;
        R.rcv_off = h.rcv_off;
        R.nISM = n_samples_rule(h.Tdiff, opt.fs);
        R.nSamples = n_samples_rule(h.Tmax, opt.fs);
        R.rir_off = (int)rir_off;
        R.out_off = h.out_off;
        R.room_id = h.room_id;
        R.level = h.level;
        const int N = h.nb[0] * h.nb[1] * h.nb[2];
        const int M = h.n_src * h.n_rcv;
        const int TT = std::min(R.nISM, tile_cap);
        const int tiles = R.nISM > 0 ? (R.nISM + TT - 1) / TT : 0;
        int C = (int)((N + per_item - 1) / per_item);
        C = std::max(C, 1);
        R.n_chunks = C;
        if (C > 1) {
            R.part_off = P.part_elems;
            P.part_elems += (long long)M * C * R.nISM;
        }
        P.ntab_max = std::max(P.ntab_max, h.nb[0] + h.nb[1] + h.nb[2]);
        P.tt_max = std::max(P.tt_max, TT);
        const int csz = (N + C - 1) / C;
        for (int m = 0; m < M; m++) {
            if (C > 1 && C <= 8)
                for (int t0 = 0; t0 < R.nISM; t0 += 256)
                    P.sum_blocks.push_back(make_int2((int)(rir_off + m), t0 | SUM_WIDE));
            else if (C > 8)
                for (int t0 = 0; t0 < R.nISM; t0 += 64) P.sum_blocks.push_back(make_int2((int)(rir_off + m), t0));
            P.rir_room.push_back((int)k);
            for (int c = 0; c < C; c++)
                for (int tl = 0; tl < tiles; tl++) {
                    Item it;
                    it.room = (int)k;
                    it.rir = m;
                    it.i0 = std::min(N, c * csz);
                    it.i1 = std::min(N, (c + 1) * csz);
                    it.t_lo = tl * TT;
                    it.t_hi = std::min(R.nISM, (tl + 1) * TT);
                    it.chunk = c;
                    it.pad = 0;
                    P.items.push_back(it);
                }
        }
        rir_off += M;
        P.rooms.push_back(R);
    }
    // longest items first, so that the tail of the launch is short
    if (sort_items)
        std::stable_sort(P.items.begin(), P.items.end(),
                         [](const Item& x, const Item& y) { return (x.i1 - x.i0) > (y.i1 - y.i0); });
    (void)total_rirs;
    return P;
}

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Basis method tables: Chebyshev interpolation coefficients of every tap offset
// (FP64 on the host, NB Chebyshev nodes in f), stored reversed in j with 16 zeros in
// front, ctab[m][16 + x] = C[m][J1 - x], for the Toeplitz operand of the products.
struct BasisTables {
    int ks = 0;   // K steps per lane group: 4 ks >= 16 + nJ - 1, ks a multiple of 4
    int cl = 0;   // row length of ctab: 4 ks + 16
    int dl = 0;   // basis row of a tile: >= BT + nJ - 1 and >= BT - 16 + 4 ks, a multiple of 16
    std::vector<float> ctab;
};

double windowed_sinc(double x, int Tw) {
    if (!(std::fabs(x) < Tw / 2.0)) return 0.0;  // src/gpuRIR_cuda.cu:110-121
    const double w = 0.5 * (1.0 + std::cos(2.0 * M_PI * x / Tw));
    return x == 0.0 ? w : w * std::sin(M_PI * x) / (M_PI * x);
}

BasisTables make_basis_tables(const TapTables& T, int Tw, int NB, int bt = BT) {
    BasisTables B;
    B.ks = 4 * ((T.nJ + 15 + 15) / 16);
    B.cl = 4 * B.ks + 16;
    B.dl = (int)align_up((size_t)std::max(bt + T.nJ - 1, bt - 16 + 4 * B.ks), 16);
    B.ctab.assign((size_t)NB * B.cl, 0.0f);
    const int J1 = T.J0 + T.nJ - 1;
    std::vector<double> fn(NB), val(NB);
    for (int i = 0; i < NB; i++) fn[i] = 0.5 + 0.5 * std::cos(M_PI * (i + 0.5) / NB);  // u_i = 2 f_i - 1
    for (int x = 0; x < T.nJ; x++) {
        const int j = J1 - x;
        for (int i = 0; i < NB; i++) val[i] = windowed_sinc((double)j - fn[i], Tw);
        for (int m = 0; m < NB; m++) {
            double c = 0.0;
            for (int i = 0; i < NB; i++) c += val[i] * std::cos(m * M_PI * (i + 0.5) / NB);
            c *= (m == 0 ? 1.0 : 2.0) / NB;
            B.ctab[(size_t)m * B.cl + 16 + x] = (float)c;
        }
    }
    return B;
}

// Lookup-table mode on the basis machinery: phase row p holds the table's entries at the
// taps of an image whose 1/16-sample delay grid point has phase p, reversed in j with 16
// zeros in front like the Chebyshev rows: ctab[p][16 + x] = T[16 (J1 - x) - p + center]
// (zero outside the table), for the offsets j = J1 - x of the table mode's window.  The
// rows are one column longer than the tile needs: the weight of phase p + 1 of an image
// at phase 15 goes to phase 0 of the next column.
BasisTables make_lut_basis_tables(const TapTables& T) {
    BasisTables B;
    B.ks = 4 * ((T.nJ + 15 + 15) / 16);
    B.cl = 4 * B.ks + 16;
    B.dl = (int)align_up((size_t)std::max(BT + T.nJ, BT - 16 + 4 * B.ks), 16);
    B.ctab.assign((size_t)LUT_NB * B.cl, 0.0f);
    const int J1 = T.J0 + T.nJ - 1, len = (int)T.lut_tab.size();
    for (int p = 0; p < LUT_NB; p++)
        for (int x = 0; x < T.nJ; x++) {
            const int i = 16 * (J1 - x) - p + T.lut_center;
            B.ctab[(size_t)p * B.cl + 16 + x] = (i >= 0 && i < len) ? (float)T.lut_tab[i] : 0.0f;
        }
    return B;
}

// LDS image of the 16-bit coefficient copies of the basis kernels with 16-bit filter inputs:
// copy rho < ncp of row m holds C[m][z - rho] at z < cl + ncp (zero outside [0, cl)), so that
// a lane's four operands start at a multiple of ncp halves (ncp = 4: 8-byte reads; 2: 4-byte
// reads).  h16 = 1: FP16; 2: BF16 (rounded to nearest even); 3: FP16 pairs of 2^14 C, the hi
// copies followed by the lo copies (the rounding error of hi times 2^11).  The kernel copies
// the image into LDS unchanged.
std::vector<unsigned short> make_half_copies(const BasisTables& B, int nb, int h16, int ncp = 4) {
    const int rs = B.cl + (ncp == 1 ? 2 : ncp);  // row stride: a single copy keeps rows word-aligned
    const size_t per = (size_t)ncp * nb * rs;
    const bool pairs = h16 == 3 || h16 == 4;  // hi and lo halves (4 reads the hi ones only)
    std::vector<unsigned short> img((pairs ? 2 : 1) * per, 0);
    for (size_t e = 0; e < per; e++) {
        const int rho = (int)(e / ((size_t)nb * rs));
        const int rm = (int)(e - (size_t)rho * nb * rs);
        const int m = rm / rs, z = rm - m * rs, src = z - rho;
        const float cv = (src >= 0 && src < B.cl) ? B.ctab[(size_t)m * B.cl + src] : 0.0f;
        if (pairs) {
            const float vs = cv * 16384.0f;
            const _Float16 hi = (_Float16)vs;
            img[e] = __builtin_bit_cast(unsigned short, hi);
            img[e + per] = __builtin_bit_cast(unsigned short, (_Float16)((vs - (float)hi) * 2048.0f));
        } else if (h16 == 2) {
            unsigned u;
            std::memcpy(&u, &cv, 4);
            u += 0x7fffu + ((u >> 16) & 1u);
            img[e] = (unsigned short)(u >> 16);
        } else {
            img[e] = __builtin_bit_cast(unsigned short, (_Float16)cv);
        }
    }
    return img;
}

// Chebyshev interpolation coefficients of every tap offset for the basis adjoint: row m
// holds 16 zeros, then C[m][k] for k = j - J0 at 16 + k, then zeros up to cl (FP64 on the
// host), the Toeplitz operand of its correlation products.
std::vector<float> make_adjoint_coefficients(const TapTables& T, int Tw, int NB, int cl) {
    std::vector<float> c((size_t)NB * cl, 0.0f);
    std::vector<double> fn(NB), val(NB);
    for (int i = 0; i < NB; i++) fn[i] = 0.5 + 0.5 * std::cos(M_PI * (i + 0.5) / NB);
    for (int k = 0; k < T.nJ; k++) {
        const int j = T.J0 + k;
        for (int i = 0; i < NB; i++) val[i] = windowed_sinc((double)j - fn[i], Tw);
        for (int m = 0; m < NB; m++) {
            double v = 0.0;
            for (int i = 0; i < NB; i++) v += val[i] * std::cos(m * M_PI * (i + 0.5) / NB);
            c[(size_t)m * cl + 16 + k] = (float)(v * (m == 0 ? 1.0 : 2.0) / NB);
        }
    }
    return c;
}

// LDS image of the basis adjoint's coefficients c ([nb][cl], make_adjoint_coefficients) as
// FP16 pairs: x = 2^14 c, hi = x rounded to FP16, lo = (x - hi) 2^11 rounded to FP16; in
// ncp copies, copy rho shifted right by rho halves, rows of cl + ncp: [hi, lo][ncp][nb][cl + ncp].
std::vector<unsigned short> make_adjoint_half_copies(const std::vector<float>& c, int nb, int cl, int ncp) {
    const size_t per = (size_t)ncp * nb * (cl + ncp);
    std::vector<unsigned short> img(2 * per, 0);
    for (size_t e = 0; e < per; e++) {
        const int rho = (int)(e / ((size_t)nb * (cl + ncp)));
        const int rm = (int)(e - (size_t)rho * nb * (cl + ncp));
        const int m = rm / (cl + ncp), z = rm - m * (cl + ncp), src = z - rho;
        const float x = (src >= 0 && src < cl) ? c[(size_t)m * cl + src] * 16384.0f : 0.0f;
        const _Float16 hi = (_Float16)x;
        img[e] = __builtin_bit_cast(unsigned short, hi);
        img[e + per] = __builtin_bit_cast(unsigned short, (_Float16)((x - (float)hi) * 2048.0f));
    }
    return img;
}

size_t basis_lds_bytes(int ntab, const BasisTables& B, int h16, int nb, bool lut, bool pack, int ncp = 4) {
    // 16-bit coefficients in ncp shifted copies (FP16 pairs: hi and lo), or FP32
    const size_t coef =
        h16 ? (h16 == 3 || h16 == 4 ? 4 : 2) * (size_t)ncp * nb * (B.cl + (ncp == 1 ? 2 : ncp)) : 4 * (size_t)nb * B.cl;
    // the region holds the sums and, after the conversion, the FP32 (or 16-bit) rows
    const size_t f32rows = 4 * (size_t)nb * (B.dl + B.dl / 4);
    const size_t rows = lut    ? f32rows
                        : pack ? std::max(4 * (size_t)nb * B.dl, (h16 == 1 || h16 == 2) ? 0 : f32rows)
                               : 8 * (size_t)nb * B.dl;
    return 8 * (size_t)ntab + 4 * (size_t)ntab + 16 * ((size_t)ntab + 4) + rows + coef + (8 + 4 + 4) * (size_t)WGB +
           4 * (size_t)(lut ? QB_LUT : QB) + 4 * 32;
}

// Basis plan: one item per RIR and range of BT-sample tiles (tile_samples rounded up
// to BT; by default as many tiles as keep about TARGET_ITEMS items, so that a small
// launch still spreads its tiles over the device), longest expected work first: the
// images in a shell grow with the cube of its outer radius and with 1 / volume.
Plan make_plan_basis(const std::vector<RoomHost>& rh, const SimOptions& opt, const TapTables& Tt,
                     bool sort_items = true, int bt = BT) {
    Plan P;
    int ntab = 0;
    long long tiles_all = 0;
    for (const RoomHost& h : rh) {
        ntab = std::max(ntab, h.nb[0] + h.nb[1] + h.nb[2]);
        tiles_all += (long long)h.n_src * h.n_rcv * ((n_samples_rule(h.Tdiff, opt.fs) + bt - 1) / bt);
    }
    P.ntab_max = ntab;
    const long long per_default = (long long)bt * std::max<long long>(1, tiles_all / TARGET_ITEMS);
    std::vector<std::pair<double, Item>> work;
    long long rir_off = 0;
    for (size_t k = 0; k < rh.size(); k++) {
        const RoomHost& h = rh[k];
        RoomDev R;
        std::memset(&R, 0, sizeof(R));
        for (int d = 0; d < 3; d++) {
            R.L[d] = h.L[d];
            R.nb[d] = h.nb[d];
        }
        for (int d = 0; d < 6; d++) R.beta[d] = h.beta[d];
        R.fs_over_c = (double)opt.fs / (double)opt.c;
        R.inv_nyz = 1.0 / ((double)h.nb[1] * (double)h.nb[2]);
        R.inv_nz = 1.0 / (double)h.nb[2];
        R.n_src = h.n_src;
        R.n_rcv = h.n_rcv;
        R.src_off = h.src_off;
        R.rcv_off = h.rcv_off;
        R.nISM = n_samples_rule(h.Tdiff, opt.fs);
        R.nSamples = n_samples_rule(h.Tmax, opt.fs);
        R.rir_off = (int)rir_off;
        R.out_off = h.out_off;
        R.room_id = h.room_id;
        R.level = h.level;
        R.n_chunks = 1;
        const int N = h.nb[0] * h.nb[1] * h.nb[2];
        const int M = h.n_src * h.n_rcv;
        int per = opt.tile_samples > 0 ? (int)align_up((size_t)opt.tile_samples, bt)
                                       : (int)std::min<long long>(per_default, 1 << 30);
        per = std::max(per, bt);
        const double vol = (double)h.L[0] * h.L[1] * h.L[2];
        for (int m = 0; m < M; m++) {
            P.rir_room.push_back((int)k);
            for (int t0 = 0; t0 < R.nISM; t0 += per) {
                Item it;
                it.room = (int)k;
                it.rir = m;
                it.i0 = 0;
                it.i1 = N;
                it.t_lo = t0;
                it.t_hi = std::min(R.nISM, t0 + per);
                it.chunk = 0;
                it.pad = 0;
                const double r1 = (double)(it.t_hi + Tt.nJ), r0 = (double)std::max(it.t_lo - Tt.nJ, 0);
                const double tiles = (double)((it.t_hi - it.t_lo + bt - 1) / bt);
                work.emplace_back((r1 * r1 * r1 - r0 * r0 * r0) / vol + 64.0 * tiles, it);
            }
        }
        rir_off += M;
        P.rooms.push_back(R);
    }
    if (sort_items)
        std::stable_sort(work.begin(), work.end(), [](const std::pair<double, Item>& x, const std::pair<double, Item>& y) {
            return x.first > y.first;
        });
    P.items.reserve(work.size());
    for (const auto& w : work) P.items.push_back(w.second);
    P.tt_max = bt;
    return P;
}

}  // namespace

// Host description of one simulation launch (two kernels).  The plan blob
// [rooms | items | rir_room | jtab | lut] is uploaded to device memory once;
// sim_enqueue() only launches kernels, so it can be captured in a hipGraph.
struct SimLaunch {
    Plan P;
    TapTables T;
    BasisTables Bt;
    bool basis = false;  // accumulation by the basis method (FP32, or FP16 filter inputs)
    int h16 = 0;         // basis filter inputs: 16-bit (1 FP16, 2 BF16) or FP16 pairs for FP32 (3)
    bool lut = false;    // table mode on the basis machinery (16 phase rows)
    bool pack = false;   // two basis rows per 64-bit sum (RIRs of fewer than PACK_MAX_IMAGES images)
    std::vector<unsigned short> chs;  // h16: LDS image of the 16-bit coefficient copies (make_half_copies)
    int bt = BT;         // samples per tile of the basis kernel (2 BT: FP16 pairs at 16 kHz, large launches)
    int nb = 8;          // Chebyshev terms of the basis method
    int prec = PREC_FP32;
    int Tw = 0;
    size_t off_rooms = 0, off_items = 0, off_rr = 0, off_jt = 0, off_lut = 0, off_sb = 0, off_ct = 0, plan_bytes = 0;
    std::vector<unsigned char> blob;
    size_t lds = 0;
    int acc_len = 0, ntab = 0;
    SimOptions opt;
    unsigned* counters = nullptr;  // persistent launch: device item counter pair (zero between launches)
    RoomParams prm{nullptr, nullptr};  // betas / levels read at run time (null: the planned values)
};

int sim_prepare(const std::vector<RoomHost>& rooms, const SimOptions& opt, SimLaunch& S) {
    GPURIR_CHECK(opt.precision == PREC_FP32 || opt.precision == PREC_LUT || opt.precision == PREC_FP16 ||
                     opt.precision == PREC_BF16,
                 "unknown precision mode");
    GPURIR_CHECK(opt.fs > 0.0f && opt.c > 0.0f, "fs and c must be positive");
    GPURIR_CHECK(!rooms.empty(), "no rooms");
    for (const RoomHost& h : rooms) {
        GPURIR_CHECK(h.nb[0] > 0 && h.nb[1] > 0 && h.nb[2] > 0, "nb_img must be positive in every dimension");
        GPURIR_CHECK((long long)h.nb[0] * h.nb[1] * h.nb[2] < (1LL << 31), "too many images per RIR");
        GPURIR_CHECK(h.nb[0] + h.nb[1] + h.nb[2] <= 4096, "too many images along the room axes");
        GPURIR_CHECK(h.n_src > 0 && h.n_rcv > 0, "each room needs at least one source and one receiver");
        GPURIR_CHECK(h.Tdiff <= h.Tmax, "Tmax must be equal or greater than Tdiff");
        GPURIR_CHECK(h.Tdiff >= 0.0f, "Tdiff must be non-negative");
    }
    S.opt = opt;
    S.Tw = tw_rule(opt.fs);
    GPURIR_CHECK(S.Tw >= 2, "sampling frequency too low for the 8 ms window");
    S.prec = opt.precision;
    {
        // per (precision, Tw): the LUT mode's table takes ~4e3 transcendental calls
        static std::mutex mu;
        static std::map<std::pair<int, int>, TapTables> cache;
        std::lock_guard<std::mutex> lk(mu);
        const auto key = std::make_pair(S.prec, S.Tw);
        auto f = cache.find(key);
        if (f == cache.end()) f = cache.emplace(key, make_tap_tables(S.prec, S.Tw)).first;
        S.T = f->second;
    }
    GPURIR_CHECK(opt.method >= METHOD_AUTO && opt.method <= METHOD_BASIS, "unknown accumulation method");
    GPURIR_CHECK(S.prec != PREC_BF16 || opt.method != METHOD_TAPS,
                 "the BF16 mode runs on the matrix cores only (method basis or auto)");
    S.basis = opt.method != METHOD_TAPS;
    // basis filters: FP16 (1) or BF16 (2) inputs in the mixed-precision modes, FP16 pairs (3)
    // for FP32
    S.h16 = S.prec == PREC_FP16 ? 1 : (S.prec == PREC_BF16 ? 2 : (S.prec == PREC_FP32 ? 3 : 0));
    S.lut = S.prec == PREC_LUT;
    S.nb = S.lut ? LUT_NB : ((S.h16 == 1 || S.h16 == 2) ? BASIS_NB_FP16 : BASIS_NB);
    if (S.basis && opt.method == METHOD_AUTO && S.prec != PREC_BF16) {
        // a launch of fewer tiles than the device has CUs (e.g. one room with two
        // receivers) is faster as a tap scatter, whose image chunks fill the device
        long long tiles = 0;
        for (const RoomHost& h : rooms)
            tiles += (long long)h.n_src * h.n_rcv * ((n_samples_rule(h.Tdiff, opt.fs) + BT - 1) / BT);
        if (tiles < AUTO_BASIS_MIN_TILES) S.basis = false;
    }
    S.bt = BT;
    if (S.basis) {
        // the tables depend on Tw (and the tile length) only (about 10^4 transcendental
        // calls): built once
        static std::mutex mu;
        static std::map<std::tuple<int, int, int>, BasisTables> cache;
        auto tables = [&](int bt) {
            std::lock_guard<std::mutex> lk(mu);
            const auto key = std::make_tuple(S.Tw, S.nb, bt);
            auto f = cache.find(key);
            if (f == cache.end())
                f = cache.emplace(key, S.lut ? make_lut_basis_tables(S.T) : make_basis_tables(S.T, S.Tw, S.nb, bt)).first;
            return f->second;
        };
        S.Bt = tables(BT);
        // the table mode's 16 phase rows: FP16-pair filters at 16 kHz (one coefficient copy),
        // FP32 ones otherwise
        if (S.lut) S.h16 = S.Bt.dl == 640 ? 3 : 0;
        // the FP16 mode at 16 kHz: the FP32 path's machinery (8 terms, packed sums, long tiles)
        // with the rows and coefficients rounded to FP16 and one product per K step (4)
        if (S.prec == PREC_FP16) {
            const int nb_fp16 = S.nb;
            S.nb = BASIS_NB;
            const BasisTables B8 = tables(BT);
            if (B8.dl == 640) {
                S.h16 = 4;
                S.Bt = B8;
            } else {
                S.nb = nb_fp16;
            }
        }
        int ntab = 0;
        for (const RoomHost& h : rooms) ntab = std::max(ntab, h.nb[0] + h.nb[1] + h.nb[2]);
        long long most_images = 0;
        for (const RoomHost& h : rooms) most_images = std::max(most_images, (long long)h.nb[0] * h.nb[1] * h.nb[2]);
        S.pack = !S.lut && (S.h16 == 0 || S.h16 == 3 || S.h16 == 4) && most_images < PACK_MAX_IMAGES;
        if (S.pack && (S.h16 == 3 || S.h16 == 4) && S.Bt.dl == 640) {
            // FP16 pairs at 16 kHz: tiles of 2 BT samples when the launch has enough of them
            // to fill the device and two workgroups still fit a CU's LDS
            long long tiles2 = 0;
            for (const RoomHost& h : rooms)
                tiles2 += (long long)h.n_src * h.n_rcv * ((n_samples_rule(h.Tdiff, opt.fs) + 2 * BT - 1) / (2 * BT));
            const BasisTables B2 = tables(2 * BT);
            if (B2.dl == 1152 && tiles2 >= LONG_TILE_MIN_TILES &&
                2 * basis_lds_bytes((ntab + 3) / 4 * 4, B2, 3, S.nb, false, true, 2) <= 160 * 1024) {
                S.bt = 2 * BT;
                S.Bt = B2;
            }
        }
        // (2 BT tiles convert with their own DLC / 128 words per thread)
        if (basis_lds_bytes((ntab + 3) / 4 * 4, S.Bt, S.h16, S.nb, S.lut, S.pack, S.lut ? 1 : S.bt == BT ? 4 : 2) >
                160 * 1024 ||
            (S.bt == BT && S.Bt.dl > BCS * WGB)) {
            // the tile's basis rows do not fit the conversion or the LDS (windows of more
            // than 513 taps, i.e. fs above 64 kHz, or very long axis tables): the tap
            // scatter takes the launch
            GPURIR_CHECK(opt.method == METHOD_AUTO && S.prec != PREC_BF16,
                         "basis method: LDS budget exceeded (window longer than 513 taps, fs above 64 kHz)");
            S.basis = false;
        }
    }
    if (S.basis) {
        S.P = make_plan_basis(rooms, opt, S.T, true, S.bt);
        if (S.h16) {
            static std::mutex mu;
            static std::map<std::tuple<int, int, int, int>, std::vector<unsigned short>> cache;
            std::lock_guard<std::mutex> lk(mu);
            const int ncp = S.lut ? 1 : S.bt == BT ? 4 : 2;
            const auto key = std::make_tuple(S.Tw, S.nb, S.h16, ncp);
            auto f = cache.find(key);
            if (f == cache.end()) f = cache.emplace(key, make_half_copies(S.Bt, S.nb, S.h16, ncp)).first;
            S.chs = f->second;
        }
    } else {
        S.Bt = BasisTables();
        S.chs.clear();
        S.bt = BT;
        S.P = make_plan(rooms, opt, S.T, true);
    }
    const Plan& P = S.P;
    const TapTables& T = S.T;
    S.off_rooms = 0;
    S.off_items = align_up(S.off_rooms + sizeof(RoomDev) * P.rooms.size(), 256);
    S.off_rr = align_up(S.off_items + sizeof(Item) * P.items.size(), 256);
    S.off_jt = align_up(S.off_rr + sizeof(int) * P.rir_room.size(), 256);
    S.off_lut = align_up(S.off_jt + sizeof(float4) * T.jtab.size(), 256);
    S.off_sb = align_up(S.off_lut + sizeof(float2) * (T.lut.size() + 1), 256);
    S.off_ct = align_up(S.off_sb + sizeof(int2) * (P.sum_blocks.size() + 1), 256);
    // the coefficients: FP32 rows, or the LDS image of their 16-bit copies
    const size_t ct_bytes = S.chs.empty() ? sizeof(float) * S.Bt.ctab.size() : sizeof(unsigned short) * S.chs.size();
    S.plan_bytes = align_up(S.off_ct + ct_bytes + sizeof(float), 256);
    S.blob.assign(S.plan_bytes, 0);
    std::memcpy(S.blob.data() + S.off_rooms, P.rooms.data(), sizeof(RoomDev) * P.rooms.size());
    if (!P.items.empty()) std::memcpy(S.blob.data() + S.off_items, P.items.data(), sizeof(Item) * P.items.size());
    std::memcpy(S.blob.data() + S.off_rr, P.rir_room.data(), sizeof(int) * P.rir_room.size());
    std::memcpy(S.blob.data() + S.off_jt, T.jtab.data(), sizeof(float4) * T.jtab.size());
    if (!T.lut.empty()) std::memcpy(S.blob.data() + S.off_lut, T.lut.data(), sizeof(float2) * T.lut.size());
    if (!P.sum_blocks.empty())
        std::memcpy(S.blob.data() + S.off_sb, P.sum_blocks.data(), sizeof(int2) * P.sum_blocks.size());
    if (!S.chs.empty())
        std::memcpy(S.blob.data() + S.off_ct, S.chs.data(), ct_bytes);
    else if (!S.Bt.ctab.empty())
        std::memcpy(S.blob.data() + S.off_ct, S.Bt.ctab.data(), sizeof(float) * S.Bt.ctab.size());
    S.ntab = (P.ntab_max + 3) / 4 * 4;  // keeps every LDS array after the tables 16-byte aligned
    // acc index = n0 - t_lo + J1 + (j - J0), n0 <= t_hi - 1 - J0, j - J0 < 64 n_pass
    S.acc_len = (int)align_up((size_t)P.tt_max + (size_t)(T.nJ - 1) + (size_t)T.n_pass * 64, 4);
    S.lds = sizeof(double) * S.ntab + sizeof(float) * S.ntab + sizeof(long long) * S.acc_len +
            (size_t)(QCAP + QPAD) * 5 * 4 + 32 + (T.n_pass <= 12 ? sizeof(float2) * T.lut.size() : 0);
    if (S.basis) S.lds = basis_lds_bytes(S.ntab, S.Bt, S.h16, S.nb, S.lut, S.pack, S.lut ? 1 : S.bt == BT ? 4 : 2);
    GPURIR_CHECK(T.n_pass <= 48, "sampling frequency too high: at most 3072 window taps (fs 384 kHz) are supported");
    GPURIR_CHECK(S.lds <= 160 * 1024, "LDS budget exceeded: reduce the tile length or the image counts");
    return 0;
}

size_t sim_plan_bytes(const SimLaunch& S) { return S.plan_bytes; }
size_t sim_partial_bytes(const SimLaunch& S) { return sizeof(float) * (size_t)S.P.part_elems; }
long long sim_n_rirs(const SimLaunch& S) { return (long long)S.P.rir_room.size(); }

static AccArgs acc_args(const SimLaunch& S, const void* d_plan, float* d_part, const float* d_pos_src,
                        const float* d_pos_rcv, const float* d_orv_src, const float* d_orv_rcv, float* d_out) {
    const unsigned char* dp = static_cast<const unsigned char*>(d_plan);
    AccArgs a;
    a.rooms = reinterpret_cast<const RoomDev*>(dp + S.off_rooms);
    a.items = reinterpret_cast<const Item*>(dp + S.off_items);
    a.pos_src = d_pos_src;
    a.pos_rcv = d_pos_rcv;
    a.orv_src = d_orv_src;
    a.orv_rcv = d_orv_rcv;
    a.jtab = reinterpret_cast<const float4*>(dp + S.off_jt);
    a.lut = reinterpret_cast<const float2*>(dp + S.off_lut);
    a.out = d_out;
    a.part = d_part;
    a.spkr = S.opt.spkr;
    a.mic = S.opt.mic;
    a.J0 = S.T.J0;
    a.nJ = S.T.nJ;
    a.n_pass = S.T.n_pass;
    a.mask_bits = S.T.mask_bits;
    a.lut_center = S.T.lut_center;
    a.lut_row = S.T.lut_row;
    a.lut_k_lo = S.T.lut_k_lo;
    a.halfTw = S.T.halfTw;
    a.inv_tw2 = 2.0f / (float)S.Tw;
    a.ntab_max = S.ntab;
    a.acc_len = S.acc_len;
    a.ctab = reinterpret_cast<const float*>(dp + S.off_ct);
    a.bks = S.Bt.ks;
    a.bdl = S.Bt.dl;
    a.bcl = S.Bt.cl;
    a.prm = S.prm;
    return a;
}

int sim_count_taps(const SimLaunch& S, const void* d_plan, const float* d_pos_src, const float* d_pos_rcv,
                   unsigned long long* d_counter, hipStream_t stream) {
    if (S.P.items.empty()) return 0;
    AccArgs a = acc_args(S, d_plan, nullptr, d_pos_src, d_pos_rcv, nullptr, nullptr, nullptr);
    const size_t lds = (sizeof(double) + sizeof(float)) * S.ntab;
    hipLaunchKernelGGL(count_taps_kernel, dim3((unsigned)S.P.items.size()), dim3(WG), lds, stream, a, d_counter);
    GPURIR_HIP(hipGetLastError());
    return 0;
}

// Raises a kernel's dynamic LDS limit above the 64 KiB default when needed (once per
// device, kernel and size).
static int allow_lds(const void* fn, size_t lds) {
    if (lds <= 64 * 1024) return 0;
    static std::mutex mu;
    static std::map<std::pair<int, const void*>, size_t> done;
    int dev = 0;
    GPURIR_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    size_t& have = done[std::make_pair(dev, fn)];
    if (have < lds) {
        GPURIR_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        have = lds;
    }
    return 0;
}

// Workgroups of a kernel that fit on the whole device at once (occupancy x CUs), cached
// per (device, kernel, LDS bytes): the queries cost more than a small launch.
static int resident_workgroups(const void* fn, size_t lds, int* out, int block = WG) {
    static std::mutex mu;
    static std::map<std::tuple<int, const void*, size_t>, int> cache;
    int dev = 0;
    GPURIR_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    const auto key = std::make_tuple(dev, fn, lds);
    auto f = cache.find(key);
    if (f == cache.end()) {
        int per_cu = 0, n_cu = 0;
        GPURIR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, block, lds));
        GPURIR_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
        f = cache.emplace(key, std::max(per_cu, 1) * std::max(n_cu, 1)).first;
    }
    *out = f->second;
    return 0;
}

int sim_enqueue_stage(const SimLaunch& S, int stage, const void* d_plan, float* d_part, const float* d_pos_src,
                      const float* d_pos_rcv, const float* d_orv_src, const float* d_orv_rcv, float* d_out,
                      hipStream_t stream) {
    const unsigned char* dp = static_cast<const unsigned char*>(d_plan);
    const RoomDev* d_rooms = reinterpret_cast<const RoomDev*>(dp + S.off_rooms);
    if ((stage & 1) && !S.P.items.empty() && S.basis) {
        AccArgs a = acc_args(S, d_plan, d_part, d_pos_src, d_pos_rcv, d_orv_src, d_orv_rcv, d_out);
        // the 16 kHz window (128 taps) has its row length compiled in
        const bool dl640 = S.Bt.dl == 640;
#define GPURIR_BASIS_FN1(K, N, H, U, P) \
    (dl640 ? reinterpret_cast<const void*>(&K<N, 640, H, U, P>) : reinterpret_cast<const void*>(&K<N, 0, H, U, P>))
#define GPURIR_BASIS_FN(K, N, H) \
    (S.pack ? GPURIR_BASIS_FN1(K, N, H, false, true) : GPURIR_BASIS_FN1(K, N, H, false, false))
        const void* fn_once = S.bt != BT   ? (S.h16 == 4 ? reinterpret_cast<const void*>(&ism_basis_kernel<BASIS_NB, 1152, 4, false, true, 2 * BT>)
                                                        : reinterpret_cast<const void*>(&ism_basis_kernel<BASIS_NB, 1152, 3, false, true, 2 * BT>))
                              : S.lut      ? (dl640 ? reinterpret_cast<const void*>(&ism_basis_kernel<LUT_NB, 640, 3, true, false>)
                                                    : reinterpret_cast<const void*>(&ism_basis_kernel<LUT_NB, 0, 0, true, false>))
                              : S.h16 == 2 ? GPURIR_BASIS_FN(ism_basis_kernel, BASIS_NB_FP16, 2)
                              : S.h16 == 1 ? GPURIR_BASIS_FN(ism_basis_kernel, BASIS_NB_FP16, 1)
                              : S.h16 == 3 ? GPURIR_BASIS_FN(ism_basis_kernel, BASIS_NB, 3)
                              : S.h16 == 4 ? GPURIR_BASIS_FN(ism_basis_kernel, BASIS_NB, 4)
                                           : GPURIR_BASIS_FN(ism_basis_kernel, BASIS_NB, 0);
        const void* fn_pers = S.bt != BT   ? (S.h16 == 4 ? reinterpret_cast<const void*>(&ism_basis_persistent_kernel<BASIS_NB, 1152, 4, false, true, 2 * BT>)
                                                        : reinterpret_cast<const void*>(&ism_basis_persistent_kernel<BASIS_NB, 1152, 3, false, true, 2 * BT>))
                              : S.lut      ? (dl640 ? reinterpret_cast<const void*>(&ism_basis_persistent_kernel<LUT_NB, 640, 3, true, false>)
                                                    : reinterpret_cast<const void*>(&ism_basis_persistent_kernel<LUT_NB, 0, 0, true, false>))
                              : S.h16 == 2 ? GPURIR_BASIS_FN(ism_basis_persistent_kernel, BASIS_NB_FP16, 2)
                              : S.h16 == 1 ? GPURIR_BASIS_FN(ism_basis_persistent_kernel, BASIS_NB_FP16, 1)
                              : S.h16 == 3 ? GPURIR_BASIS_FN(ism_basis_persistent_kernel, BASIS_NB, 3)
                              : S.h16 == 4 ? GPURIR_BASIS_FN(ism_basis_persistent_kernel, BASIS_NB, 4)
                                           : GPURIR_BASIS_FN(ism_basis_persistent_kernel, BASIS_NB, 0);
#undef GPURIR_BASIS_FN
#undef GPURIR_BASIS_FN1
        int n_items = (int)S.P.items.size();
        int resident = 0;
        if (S.counters) {
            if (int rc = allow_lds(fn_pers, S.lds)) return rc;
            if (int rc = resident_workgroups(fn_pers, S.lds, &resident, WGB)) return rc;
        }
        if (S.counters && n_items > resident) {
            unsigned* cnt = S.counters;
            void* args[] = {&a, &n_items, &cnt};
            GPURIR_HIP(hipLaunchKernel(fn_pers, dim3((unsigned)resident), dim3(WGB), args, S.lds, stream));
        } else {
            if (int rc = allow_lds(fn_once, S.lds)) return rc;
            void* args[] = {&a};
            GPURIR_HIP(hipLaunchKernel(fn_once, dim3((unsigned)n_items), dim3(WGB), args, S.lds, stream));
        }
    } else if ((stage & 1) && !S.P.items.empty()) {
        AccArgs a = acc_args(S, d_plan, d_part, d_pos_src, d_pos_rcv, d_orv_src, d_orv_rcv, d_out);
        const dim3 grid((unsigned)S.P.items.size());
        const void* fn = nullptr;
        const bool mk = S.T.mask_bits != 0;
#define GPURIR_ACC_FN(M, P) \
    (mk ? reinterpret_cast<const void*>(&ism_accumulate_kernel<M, P, true, false>) \
        : reinterpret_cast<const void*>(&ism_accumulate_kernel<M, P, false, false>))
#define GPURIR_ACC_GRP(M) reinterpret_cast<const void*>(&ism_accumulate_kernel<M, 12, true, true>)
#define GPURIR_ACC_PICK(M)                                         \
    switch (S.T.n_pass) {                                          \
        case 1: fn = GPURIR_ACC_FN(M, 1); break;                   \
        case 2: fn = GPURIR_ACC_FN(M, 2); break;                   \
        case 3: fn = GPURIR_ACC_FN(M, 3); break;                   \
        case 4: fn = GPURIR_ACC_FN(M, 4); break;                   \
        case 6: fn = GPURIR_ACC_FN(M, 6); break;                   \
        case 12: fn = GPURIR_ACC_FN(M, 12); break;                 \
        default: fn = GPURIR_ACC_GRP(M); break;                    \
    }
        if (S.prec == PREC_LUT) {
            GPURIR_ACC_PICK(PREC_LUT)
        } else if (S.prec == PREC_FP16) {
            GPURIR_ACC_PICK(PREC_FP16)
        } else {
            GPURIR_ACC_PICK(PREC_FP32)
        }
#undef GPURIR_ACC_PICK
#undef GPURIR_ACC_FN
#undef GPURIR_ACC_GRP
        const void* fn_once = fn;  // one workgroup per item
        if (S.counters) {
            // the same instantiation in its persistent form
#define GPURIR_ACC_FN(M, P) \
    (mk ? reinterpret_cast<const void*>(&ism_accumulate_persistent_kernel<M, P, true, false>) \
        : reinterpret_cast<const void*>(&ism_accumulate_persistent_kernel<M, P, false, false>))
#define GPURIR_ACC_GRP(M) reinterpret_cast<const void*>(&ism_accumulate_persistent_kernel<M, 12, true, true>)
#define GPURIR_ACC_PICK(M)                                         \
    switch (S.T.n_pass) {                                          \
        case 1: fn = GPURIR_ACC_FN(M, 1); break;                   \
        case 2: fn = GPURIR_ACC_FN(M, 2); break;                   \
        case 3: fn = GPURIR_ACC_FN(M, 3); break;                   \
        case 4: fn = GPURIR_ACC_FN(M, 4); break;                   \
        case 6: fn = GPURIR_ACC_FN(M, 6); break;                   \
        case 12: fn = GPURIR_ACC_FN(M, 12); break;                 \
        default: fn = GPURIR_ACC_GRP(M); break;                    \
    }
            if (S.prec == PREC_LUT) {
                GPURIR_ACC_PICK(PREC_LUT)
            } else if (S.prec == PREC_FP16) {
                GPURIR_ACC_PICK(PREC_FP16)
            } else {
                GPURIR_ACC_PICK(PREC_FP32)
            }
#undef GPURIR_ACC_PICK
#undef GPURIR_ACC_FN
#undef GPURIR_ACC_GRP
        }
        if (int rc = allow_lds(fn, S.lds)) return rc;
        int resident = 0;
        if (S.counters) {
            if (int rc = resident_workgroups(fn, S.lds, &resident)) return rc;
        }
        int n_items = (int)S.P.items.size();
        if (S.counters && n_items > resident) {
            // more items than resident workgroups: each workgroup takes items off the counter
            unsigned* cnt = S.counters;
            void* args[] = {&a, &n_items, &cnt};
            GPURIR_HIP(hipLaunchKernel(fn, dim3((unsigned)resident), dim3(WG), args, S.lds, stream));
        } else if (S.counters) {
            // every item has a resident workgroup of its own: the counter would only add latency
            if (int rc = allow_lds(fn_once, S.lds)) return rc;
            void* args[] = {&a};
            GPURIR_HIP(hipLaunchKernel(fn_once, grid, dim3(WG), args, S.lds, stream));
        } else {
            void* args[] = {&a};
            GPURIR_HIP(hipLaunchKernel(fn, grid, dim3(WG), args, S.lds, stream));
        }
    }
    if (!(stage & 2)) return 0;
    if (!S.P.sum_blocks.empty()) {
        hipLaunchKernelGGL(partial_sum_kernel, dim3((unsigned)S.P.sum_blocks.size()), dim3(WG), 0, stream, d_rooms,
                           reinterpret_cast<const int*>(dp + S.off_rr), reinterpret_cast<const int2*>(dp + S.off_sb),
                           d_part, d_out);
        GPURIR_HIP(hipGetLastError());
    }
    FinArgs f;
    f.rooms = d_rooms;
    f.rir_room = reinterpret_cast<const int*>(dp + S.off_rr);
    f.pos_src = d_pos_src;
    f.pos_rcv = d_pos_rcv;
    f.out = d_out;
    f.w_sz = 10e-3f * S.opt.fs;
    f.fs = (double)S.opt.fs;
    f.seed = S.opt.seed;
    f.call_offset = S.opt.call_offset;
    f.noise = S.opt.noise;
    f.crossfade = S.opt.crossfade;
    f.prm = S.prm;
    hipLaunchKernelGGL(ism_finalize_kernel, dim3((unsigned)S.P.rir_room.size()), dim3(WG), 0, stream, f);
    GPURIR_HIP(hipGetLastError());
    return 0;
}

int sim_enqueue(const SimLaunch& S, const void* d_plan, float* d_part, const float* d_pos_src, const float* d_pos_rcv,
                const float* d_orv_src, const float* d_orv_rcv, float* d_out, hipStream_t stream) {
    return sim_enqueue_stage(S, 3, d_plan, d_part, d_pos_src, d_pos_rcv, d_orv_src, d_orv_rcv, d_out, stream);
}

SimLaunch* sim_new() { return new SimLaunch(); }
void sim_set_counters(SimLaunch& S, unsigned* d_counters) { S.counters = d_counters; }
void sim_set_params(SimLaunch& S, const float* d_beta, const float* d_level) { S.prm = RoomParams{d_beta, d_level}; }
const SimOptions& sim_options_of(const SimLaunch& S) { return S.opt; }
bool sim_is_basis(const SimLaunch& S) { return S.basis; }
void sim_delete(SimLaunch* s) { delete s; }
const void* sim_blob(const SimLaunch& S) { return S.blob.data(); }
void sim_set_stream_opts(SimLaunch& S, unsigned long long seed, unsigned call_offset) {
    S.opt.seed = seed;
    S.opt.call_offset = call_offset;
}

// ---- plans of recent eager launches
// An eager launch plans its rooms on the host and uploads the plan.  Repeated calls for the
// same rooms and options (positions and orientations are read from device memory at run
// time, and the noise stream's seed and call counter are launch arguments) reuse the last
// plans: per device the PLAN_CACHE most recent, each with its own device copy.  A plan's
// device copy is only read after its upload, so launches on any stream may share it; the
// host entry point's positions section is rewritten by each of its (synchronous) calls.
namespace {
struct CachedPlan {
    std::string key;
    SimLaunch S;
    void* d = nullptr;
    size_t bytes = 0;  // plan, then the host entry point's positions section
    hipEvent_t uploaded = nullptr;  // after the plan's upload; launches on any stream wait for it
};
constexpr int PLAN_CACHE = 8;
constexpr int PLAN_CACHE_DEV = 64;
std::mutex g_plan_mu;
std::vector<CachedPlan*> g_plans[PLAN_CACHE_DEV];  // most recently used last

std::string plan_key(const std::vector<RoomHost>& rooms, const SimOptions& o, int n_pos_src, int n_pos_rcv) {
    std::string k;
    auto put = [&k](const void* p, size_t n) { k.append(static_cast<const char*>(p), n); };
    for (const RoomHost& h : rooms) {
        put(h.L, sizeof(h.L));
        put(h.beta, sizeof(h.beta));
        put(h.nb, sizeof(h.nb));
        const int ints[4] = {h.n_src, h.n_rcv, h.src_off, h.rcv_off};
        put(ints, sizeof(ints));
        const float fl[3] = {h.Tdiff, h.Tmax, h.level};
        put(fl, sizeof(fl));
        put(&h.out_off, sizeof(h.out_off));
        put(&h.room_id, sizeof(h.room_id));
    }
    const int oi[10] = {o.precision, o.noise, o.spkr, o.mic, o.chunk_images, o.tile_samples, o.crossfade, o.method,
                        n_pos_src, n_pos_rcv};
    put(oi, sizeof(oi));
    const float of[2] = {o.fs, o.c};
    put(of, sizeof(of));
    return k;
}

// The cached plan of these rooms and options (planned and uploaded on a miss, with
// extra_bytes of device memory after the plan).  The caller holds g_plan_mu while it uses
// the entry.
int cached_plan(const std::vector<RoomHost>& rooms, const SimOptions& opt, int n_pos_src, int n_pos_rcv,
                size_t extra_bytes, hipStream_t stream, CachedPlan** out) {
    int dev = 0;
    GPURIR_HIP(hipGetDevice(&dev));
    GPURIR_CHECK(dev >= 0 && dev < PLAN_CACHE_DEV, "device index out of range");
    const std::string key = plan_key(rooms, opt, n_pos_src, n_pos_rcv);
    std::vector<CachedPlan*>& L = g_plans[dev];
    for (size_t i = 0; i < L.size(); i++) {
        if (L[i]->key == key) {
            CachedPlan* c = L[i];
            L.erase(L.begin() + i);
            L.push_back(c);
            GPURIR_HIP(hipStreamWaitEvent(stream, c->uploaded, 0));
            *out = c;
            return 0;
        }
    }
    CachedPlan* c = new CachedPlan();
    c->key = key;
    if (int rc = sim_prepare(rooms, opt, c->S)) {
        delete c;
        return rc;
    }
    c->bytes = align_up(c->S.plan_bytes, 256) + extra_bytes;
    if (hipMalloc(&c->d, c->bytes) != hipSuccess) {
        (void)hipGetLastError();
        delete c;
        set_error("out of device memory for an eager plan");
        return -3;
    }
    int rc = upload_async(c->d, c->S.blob.data(), c->S.plan_bytes, stream);
    if (!rc && (hipEventCreateWithFlags(&c->uploaded, hipEventDisableTiming) != hipSuccess ||
                hipEventRecord(c->uploaded, stream) != hipSuccess)) {
        (void)hipGetLastError();
        set_error("HIP error while recording a plan upload");
        rc = -2;
    }
    if (rc) {
        if (c->uploaded) (void)hipEventDestroy(c->uploaded);
        (void)hipFree(c->d);
        delete c;
        return rc;
    }
    if (L.size() >= (size_t)PLAN_CACHE) {
        CachedPlan* old = L.front();
        L.erase(L.begin());
        (void)hipFree(old->d);  // waits for the launches that still read it
        (void)hipEventDestroy(old->uploaded);
        delete old;
    }
    L.push_back(c);
    *out = c;
    return 0;
}
}  // namespace

int launch_simulation(const std::vector<RoomHost>& rooms, const float* d_pos_src, const float* d_pos_rcv,
                      const float* d_orv_src, const float* d_orv_rcv, const SimOptions& opt, float* d_out,
                      hipStream_t stream) {
    std::lock_guard<std::mutex> lk(g_plan_mu);
    CachedPlan* c = nullptr;
    if (int rc = cached_plan(rooms, opt, 0, 0, 0, stream, &c)) return rc;
    SimLaunch& S = c->S;
    sim_set_stream_opts(S, opt.seed, opt.call_offset);
    float* dpart = nullptr;
    if (S.P.part_elems > 0) {
        void* pv = nullptr;
        if (int rc = workspace(WS_PARTIAL, sim_partial_bytes(S), &pv)) return rc;
        dpart = static_cast<float*>(pv);
    }
    unsigned* counters = nullptr;  // persistent launch: items taken longest first from a counter
    if (int rc = device_counters(&counters)) return rc;
    // basis launches take one workgroup per item (their persistent form holds more
    // registers and fits fewer workgroups per CU; plans choose it only when timed faster)
    sim_set_counters(S, S.basis ? nullptr : counters);
    return sim_enqueue(S, c->d, dpart, d_pos_src, d_pos_rcv, d_orv_src, d_orv_rcv, d_out, stream);
}

int launch_simulation_host(const std::vector<RoomHost>& rooms, const float* pos_src, int n_src, const float* pos_rcv,
                           int n_rcv, const float* orv_src, const float* orv_rcv, const SimOptions& opt, float* d_out,
                           hipStream_t stream) {
    // after the plan: pos_src | pos_rcv | orv_src | orv_rcv, 256-byte aligned sections
    const size_t bs = align_up(sizeof(float) * 3 * (size_t)n_src, 256), br = align_up(sizeof(float) * 3 * (size_t)n_rcv, 256);
    const size_t o_pr = bs, o_os = o_pr + br, o_or = o_os + bs, total = o_or + br;
    std::vector<unsigned char> h(total, 0);
    std::memcpy(h.data(), pos_src, sizeof(float) * 3 * n_src);
    std::memcpy(h.data() + o_pr, pos_rcv, sizeof(float) * 3 * n_rcv);
    if (orv_src) std::memcpy(h.data() + o_os, orv_src, sizeof(float) * 3 * n_src);
    if (orv_rcv) std::memcpy(h.data() + o_or, orv_rcv, sizeof(float) * 3 * n_rcv);
    std::lock_guard<std::mutex> lk(g_plan_mu);
    CachedPlan* c = nullptr;
    if (int rc = cached_plan(rooms, opt, n_src, n_rcv, total, stream, &c)) return rc;
    SimLaunch& S = c->S;
    sim_set_stream_opts(S, opt.seed, opt.call_offset);
    unsigned char* d = static_cast<unsigned char*>(c->d) + align_up(S.plan_bytes, 256);
    if (int rc = upload_async(d, h.data(), total, stream)) return rc;
    float* dpart = nullptr;
    if (S.P.part_elems > 0) {
        void* pv = nullptr;
        if (int rc = workspace(WS_PARTIAL, sim_partial_bytes(S), &pv)) return rc;
        dpart = static_cast<float*>(pv);
    }
    unsigned* counters = nullptr;
    if (int rc = device_counters(&counters)) return rc;
    sim_set_counters(S, S.basis ? nullptr : counters);
    return sim_enqueue(S, c->d, dpart, reinterpret_cast<const float*>(d), reinterpret_cast<const float*>(d + o_pr),
                       reinterpret_cast<const float*>(d + o_os), reinterpret_cast<const float*>(d + o_or), d_out, stream);
}

// ---- the no-allocation launch (gpurir_simulate_rir_device_noalloc)
// Its plan, partial rows and work counters live in `scratch`, one array per GPU, held for
// the life of the process.  On the host the array's name is not a device address: each
// device's address comes from hipGetSymbolAddress (noalloc_scratch) and only that address
// is handed to copies and kernels.  Calls on one GPU share the array; calls that overlap
// on one GPU race.
namespace {
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEV = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[NOALLOC_MAX_DEV] = {};

// The current device's copy of `scratch`, resolved once per device with that device current.
int noalloc_scratch(void** ptr) {
    int dev = 0;
    GPURIR_HIP(hipGetDevice(&dev));
    GPURIR_CHECK(dev >= 0 && dev < NOALLOC_MAX_DEV, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            (void)hipGetLastError();
            set_error(std::string("hipGetSymbolAddress of the no-allocation scratch failed: ") +
                      hipGetErrorString(e));
            return -2;
        }
        g_scratch_ptr[dev] = p;
    }
    *ptr = g_scratch_ptr[dev];
    return 0;
}

// Fills n words with zero (a kernel rather than a memset node, so that captured graphs
// replay it correctly).
__global__ void zero_words_kernel(unsigned* p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}
}  // namespace

int launch_simulation_noalloc(const std::vector<RoomHost>& rooms, const float* d_pos_src, const float* d_pos_rcv,
                              const float* d_orv_src, const float* d_orv_rcv, const SimOptions& opt, float* d_out,
                              hipStream_t stream) {
    SimLaunch S;
    if (int rc = sim_prepare(rooms, opt, S)) return rc;
    // scratch: counters [0, 256) | plan | partial rows, 256-byte aligned sections
    const size_t o_plan = 256, o_part = align_up(o_plan + S.plan_bytes, 256);
    const size_t need = o_part + align_up(sim_partial_bytes(S), 256);
    if (need > NOALLOC_SCRATCH_BYTES) {
        set_error("this launch needs " + std::to_string(need) + " bytes of device scratch (plan " +
                  std::to_string(S.plan_bytes) + ", partial rows " + std::to_string(sim_partial_bytes(S)) +
                  "); the no-allocation entry point has " + std::to_string(NOALLOC_SCRATCH_BYTES) +
                  " (use gpurir_simulate_rir_device, or fewer receivers per call)");
        return -1;
    }
    void* base = nullptr;
    if (int rc = noalloc_scratch(&base)) return rc;
    unsigned char* b = static_cast<unsigned char*>(base);
    unsigned* counters = reinterpret_cast<unsigned*>(b);
    hipLaunchKernelGGL(zero_words_kernel, dim3(1), dim3(64), 0, stream, counters, 2);
    GPURIR_HIP(hipGetLastError());
    // the plan goes from pageable host memory: the runtime stages it, no pinned buffer is allocated
    GPURIR_HIP(hipMemcpyAsync(b + o_plan, S.blob.data(), S.plan_bytes, hipMemcpyHostToDevice, stream));
    float* dpart = S.P.part_elems > 0 ? reinterpret_cast<float*>(b + o_part) : nullptr;
    sim_set_counters(S, S.basis ? nullptr : counters);
    return sim_enqueue(S, b + o_plan, dpart, d_pos_src, d_pos_rcv, d_orv_src, d_orv_rcv, d_out, stream);
}

int launch_image_list(const RoomHost& room, const float* d_pos_src, const float* d_pos_rcv, const float* d_orv_src,
                      const float* d_orv_rcv, const SimOptions& opt, int* d_n0, float* d_frac, float* d_amp,
                      hipStream_t stream) {
    GPURIR_CHECK(room.nb[0] > 0 && room.nb[1] > 0 && room.nb[2] > 0, "nb_img must be positive in every dimension");
    GPURIR_CHECK(room.nb[0] + room.nb[1] + room.nb[2] <= 4096, "too many images along the room axes");
    std::vector<RoomHost> rh(1, room);
    Plan P = make_plan(rh, opt, make_tap_tables(PREC_FP32, std::max(tw_rule(opt.fs), 2)), false);
    void* dplan = nullptr;
    if (int rc = workspace(WS_PLAN, sizeof(RoomDev), &dplan)) return rc;
    if (int rc = upload_async(dplan, P.rooms.data(), sizeof(RoomDev), stream)) return rc;
    const long long N = (long long)room.nb[0] * room.nb[1] * room.nb[2];
    const int M = room.n_src * room.n_rcv;
    GPURIR_CHECK(M <= 65535, "image listing supports at most 65535 (source, receiver) pairs per call");
    const int ntab = (P.ntab_max + 3) / 4 * 4;
    const size_t lds = (sizeof(double) + sizeof(float)) * ntab;
    hipLaunchKernelGGL(image_list_kernel, dim3((unsigned)((N + WG - 1) / WG), (unsigned)M), dim3(WG), lds, stream,
                       static_cast<const RoomDev*>(dplan), d_pos_src, d_pos_rcv, d_orv_src, d_orv_rcv, opt.spkr,
                       opt.mic, ntab, d_n0, d_frac, d_amp);
    GPURIR_HIP(hipGetLastError());
    return 0;
}


// ---------------------------------------------------------------------------
// adjoint launch
// ---------------------------------------------------------------------------

namespace {
constexpr int ADJ_TILE = 8192;  // ISM samples of g' per adjoint workgroup at most
}

// Host description of one adjoint launch.  Items are those of the forward plan
// for the same rooms, left in (room, RIR, chunk, tile) order so that the items of
// one RIR are contiguous; the blob is [rooms | items | rir_room | ktab | rir_item_begin].
struct AdjLaunch {
    Plan P;
    TapTables T;
    std::vector<float4> ktab;
    std::vector<int> rir_item_begin;
    std::vector<int> order;
    bool basis = false;      // basis adjoint (ism_adjoint_basis_kernel)
    std::vector<unsigned short> achs;  // basis: coefficient FP16 pairs (make_adjoint_half_copies)
    int abt = BT, adl = 0, agl = 0, aks = 0, acl = 0;
    size_t off_rooms = 0, off_items = 0, off_rr = 0, off_kt = 0, off_rib = 0, off_ord = 0, off_act = 0, plan_bytes = 0;
    std::vector<unsigned char> blob;
    size_t lds = 0;
    int ntab = 0, gp_len = 0, Tw = 0;
    bool mask = false;
    SimOptions opt;
    RoomParams prm{nullptr, nullptr};  // betas / levels read at run time (null: the planned values)
};

AdjLaunch* adj_new() { return new AdjLaunch(); }
void adj_delete(AdjLaunch* a) { delete a; }
void adj_set_params(AdjLaunch& A, const float* d_beta, const float* d_level) { A.prm = RoomParams{d_beta, d_level}; }
size_t adj_plan_bytes(const AdjLaunch& A) { return A.plan_bytes; }
const void* adj_blob(const AdjLaunch& A) { return A.blob.data(); }

// workspace: item partials [n_items][12] | rir_tail [n_rirs][13] | rir_grad [n_rirs][13] | cw | ws
size_t adj_ws_bytes(const AdjLaunch& A) {
    const size_t ni = A.P.items.size(), nr = A.P.rir_room.size();
    return align_up(sizeof(double) * 12 * ni, 256) + 2 * align_up(sizeof(double) * 13 * nr, 256) +
           align_up(sizeof(float) * nr, 256) + align_up(sizeof(int) * nr, 256);
}

int adj_prepare(const std::vector<RoomHost>& rooms, const SimOptions& opt_in, AdjLaunch& A) {
    GPURIR_CHECK(opt_in.fs > 0.0f && opt_in.c > 0.0f, "fs and c must be positive");
    GPURIR_CHECK(opt_in.crossfade == 0, "the adjoint does not support a cross-fade");
    GPURIR_CHECK(!rooms.empty(), "no rooms");
    int nism_max = 0;
    for (const RoomHost& h : rooms) {
        GPURIR_CHECK(h.nb[0] > 0 && h.nb[1] > 0 && h.nb[2] > 0, "nb_img must be positive in every dimension");
        GPURIR_CHECK((long long)h.nb[0] * h.nb[1] * h.nb[2] < (1LL << 31), "too many images per RIR");
        GPURIR_CHECK(h.nb[0] + h.nb[1] + h.nb[2] <= 4096, "too many images along the room axes");
        GPURIR_CHECK(h.n_src > 0 && h.n_rcv > 0, "each room needs at least one source and one receiver");
        GPURIR_CHECK(h.Tdiff <= h.Tmax, "Tmax must be equal or greater than Tdiff");
        GPURIR_CHECK(h.Tdiff >= 0.0f, "Tdiff must be non-negative");
        nism_max = std::max(nism_max, n_samples_rule(h.Tdiff, opt_in.fs));
    }
    SimOptions opt = opt_in;
    A.Tw = tw_rule(opt.fs);
    GPURIR_CHECK(A.Tw >= 2, "sampling frequency too low for the 8 ms window");
    A.T = make_tap_tables(PREC_FP32, A.Tw);
    // basis adjoint unless asked for taps, or (automatic) fewer tiles than CUs, or its LDS
    // does not fit
    A.basis = opt.method != METHOD_TAPS;
    if (A.basis) {
        long long tiles = 0;
        int ntab = 0;
        for (const RoomHost& h : rooms) {
            tiles += (long long)h.n_src * h.n_rcv * ((n_samples_rule(h.Tdiff, opt.fs) + BT - 1) / BT);
            ntab = std::max(ntab, h.nb[0] + h.nb[1] + h.nb[2]);
        }
        A.aks = 8 * ((A.T.nJ + 15 + 31) / 32);  // K per lane group: 4 aks >= nJ + 15, steps of 8
        A.acl = 4 * A.aks + 16;
        const size_t nt = align_up((size_t)ntab, 4);
        // tiles of 768 delays where two workgroups still fit a CU's LDS, else 512
        for (int abt : {3 * BT / 2, BT}) {
            A.abt = abt;
            A.adl = (int)align_up((size_t)(abt + A.T.nJ - 1), 16);
            const int nota = (A.adl + 255) / 256;
            A.agl = (int)align_up((size_t)std::max(A.adl + A.T.nJ - 1, 256 * nota - 16 + 4 * A.aks + 4), 16);
            // axis tables, x runs and row starts, g' pairs, coefficient pairs (two copies), E,
            // row arrays, queue, scan scratch
            A.lds = 20 * nt + 16 * (nt + 4) + 4 * ((size_t)A.agl + A.agl / 2) + 8 * (size_t)ADJ_NB * (A.acl + 2) +
                    4 * (size_t)ADJ_NB * A.adl + 12 * (size_t)WGB + 4 * (size_t)QB + 128;
            if (2 * A.lds <= 160 * 1024) break;
        }
        if ((opt.method == METHOD_AUTO && tiles < AUTO_BASIS_MIN_TILES) || A.lds > 150 * 1024 || A.agl > 4 * WGB) {
            GPURIR_CHECK(opt.method == METHOD_AUTO, "basis adjoint: LDS budget exceeded or window too long");
            A.basis = false;
        }
    }
    if (A.basis) {
        SimOptions ob = opt;
        // items of about 1/ADJ_TARGET_ITEMS of the launch's tiles (at least one tile): fewer,
        // longer items than the forward planner's default, since an item's set-up (axis
        // tables, coefficient image, the final reduction) weighs more here than the launch's
        // tail (256 rooms: 0.321 -> 0.296 ms with one item per RIR)
        long long tiles_abt = 0;
        for (const RoomHost& h : rooms)
            tiles_abt += (long long)h.n_src * h.n_rcv * ((n_samples_rule(h.Tdiff, opt.fs) + A.abt - 1) / A.abt);
        ob.tile_samples = A.abt * (int)std::max<long long>(1, tiles_abt / ADJ_TARGET_ITEMS);
        A.opt = opt;
        A.mask = (A.Tw % 2) != 0;
        A.P = make_plan_basis(rooms, ob, A.T, false, A.abt);
        A.achs = make_adjoint_half_copies(make_adjoint_coefficients(A.T, A.Tw, ADJ_NB, A.acl), ADJ_NB, A.acl, 2);
    } else {
        if (opt.tile_samples <= 0) opt.tile_samples = std::max(1, std::min(nism_max, ADJ_TILE));
        if (opt.chunk_images <= 0) {
            // about 8192 items (a few per resident workgroup) so that the launch ends evenly
            double work = 0.0;
            for (const RoomHost& h : rooms) {
                const int nISM = n_samples_rule(h.Tdiff, opt.fs);
                const int tiles = nISM > 0 ? (nISM + opt.tile_samples - 1) / opt.tile_samples : 0;
                work += (double)h.n_src * h.n_rcv * h.nb[0] * h.nb[1] * (double)h.nb[2] * tiles;
            }
            opt.chunk_images = (int)std::min(65536.0, std::max(1024.0, std::ceil(work / 8192.0)));
        }
        A.opt = opt;
        A.mask = (A.Tw % 2) != 0;
        A.P = make_plan(rooms, opt, A.T, false);
    }
    const Plan& P = A.P;
    const int J1 = A.T.J0 + A.T.nJ - 1;
    GPURIR_CHECK(A.T.J0 <= 0 && J1 >= 1, "window too short for the adjoint");
    A.ktab.resize(A.T.nJ);
    for (int u = 0; u < A.T.nJ; u++) {
        const int k = A.T.J0 + u;
        const double sg = (k & 1) ? 1.0 : -1.0;  // (-1)^(k+1)
        const double th = 2.0 * M_PI * (double)k / (double)A.Tw;
        A.ktab[u] = make_float4((float)k, (float)std::cos(th), (float)std::sin(th), (float)(0.5 * sg));
    }
    const int n_rirs = (int)P.rir_room.size();
    A.rir_item_begin.assign(n_rirs + 1, 0);
    {
        // items are in (room, rir, chunk, tile) order; global RIR = rir_off + rir
        std::vector<int> cnt(n_rirs, 0);
        for (const Item& it : P.items) cnt[P.rooms[it.room].rir_off + it.rir]++;
        for (int g = 0; g < n_rirs; g++) A.rir_item_begin[g + 1] = A.rir_item_begin[g] + cnt[g];
    }
    // launch order: longest items first (the partial sums keep the item's own index); basis
    // items by their shell volume over the room volume
    A.order.resize(P.items.size());
    for (size_t e = 0; e < P.items.size(); e++) A.order[e] = (int)e;
    if (A.basis) {
        std::vector<double> wk(P.items.size());
        for (size_t e = 0; e < P.items.size(); e++) {
            const Item& it = P.items[e];
            const RoomDev& R = P.rooms[it.room];
            const double r1 = (double)(it.t_hi + A.T.nJ), r0 = (double)std::max(it.t_lo - A.T.nJ, 0);
            wk[e] = (r1 * r1 * r1 - r0 * r0 * r0) / (R.L[0] * R.L[1] * R.L[2]);
        }
        std::stable_sort(A.order.begin(), A.order.end(), [&](int x, int y) { return wk[x] > wk[y]; });
    } else {
        std::stable_sort(A.order.begin(), A.order.end(), [&](int x, int y) {
            return (P.items[x].i1 - P.items[x].i0) > (P.items[y].i1 - P.items[y].i0);
        });
    }
    A.off_rooms = 0;
    A.off_items = align_up(sizeof(RoomDev) * P.rooms.size(), 256);
    A.off_rr = align_up(A.off_items + sizeof(Item) * P.items.size(), 256);
    A.off_kt = align_up(A.off_rr + sizeof(int) * P.rir_room.size(), 256);
    A.off_rib = align_up(A.off_kt + sizeof(float4) * A.ktab.size(), 256);
    A.off_ord = align_up(A.off_rib + sizeof(int) * A.rir_item_begin.size(), 256);
    A.off_act = align_up(A.off_ord + sizeof(int) * (A.order.size() + 1), 256);
    A.plan_bytes = align_up(A.off_act + sizeof(unsigned short) * (A.achs.size() + 4), 256);
    A.blob.assign(A.plan_bytes, 0);
    std::memcpy(A.blob.data() + A.off_rooms, P.rooms.data(), sizeof(RoomDev) * P.rooms.size());
    if (!P.items.empty()) std::memcpy(A.blob.data() + A.off_items, P.items.data(), sizeof(Item) * P.items.size());
    std::memcpy(A.blob.data() + A.off_rr, P.rir_room.data(), sizeof(int) * P.rir_room.size());
    std::memcpy(A.blob.data() + A.off_kt, A.ktab.data(), sizeof(float4) * A.ktab.size());
    std::memcpy(A.blob.data() + A.off_rib, A.rir_item_begin.data(), sizeof(int) * A.rir_item_begin.size());
    if (!A.order.empty()) std::memcpy(A.blob.data() + A.off_ord, A.order.data(), sizeof(int) * A.order.size());
    if (!A.achs.empty())
        std::memcpy(A.blob.data() + A.off_act, A.achs.data(), sizeof(unsigned short) * A.achs.size());
    A.ntab = (P.ntab_max + 3) / 4 * 4;
    A.gp_len = (int)align_up((size_t)P.tt_max + 2 * (size_t)(A.T.nJ - 1) + 4, 4);
    if (!A.basis)
        A.lds = (sizeof(double) + 3 * sizeof(float)) * A.ntab + sizeof(float) * A.gp_len + (size_t)AQ * 12;
    GPURIR_CHECK(A.lds <= 150 * 1024, "LDS budget exceeded: reduce the tile length or the image counts");
    return 0;
}

int adj_enqueue(const AdjLaunch& A, const void* d_plan, void* d_ws, const float* d_pos_src, const float* d_pos_rcv,
                const float* d_orv_src, const float* d_orv_rcv, const float* d_y, const float* d_gy, float* g_beta,
                float* g_src, float* g_rcv, float* g_level, hipStream_t stream) {
    const unsigned char* dp = static_cast<const unsigned char*>(d_plan);
    const Plan& P = A.P;
    const int n_rirs = (int)P.rir_room.size();
    unsigned char* w = static_cast<unsigned char*>(d_ws);
    double* item_part = reinterpret_cast<double*>(w);
    w += align_up(sizeof(double) * 12 * P.items.size(), 256);
    double* rir_tail = reinterpret_cast<double*>(w);
    w += align_up(sizeof(double) * 13 * n_rirs, 256);
    double* rir_grad = reinterpret_cast<double*>(w);
    w += align_up(sizeof(double) * 13 * n_rirs, 256);
    float* rir_cw = reinterpret_cast<float*>(w);
    w += align_up(sizeof(float) * n_rirs, 256);
    int* rir_ws = reinterpret_cast<int*>(w);
    const RoomDev* d_rooms = reinterpret_cast<const RoomDev*>(dp + A.off_rooms);
    const int* d_rr = reinterpret_cast<const int*>(dp + A.off_rr);

    TailAdjArgs t;
    t.rooms = d_rooms;
    t.rir_room = d_rr;
    t.pos_src = d_pos_src;
    t.pos_rcv = d_pos_rcv;
    t.y = d_y;
    t.gy = d_gy;
    t.rir_cw = rir_cw;
    t.rir_ws = rir_ws;
    t.rir_tail = rir_tail;
    t.w_sz = 10e-3f * A.opt.fs;
    t.fs = (double)A.opt.fs;
    t.seed = A.opt.seed;
    t.call_offset = A.opt.call_offset;
    t.noise = A.opt.noise;
    t.prm = A.prm;
    hipLaunchKernelGGL(tail_adjoint_kernel, dim3((unsigned)n_rirs), dim3(WG), 0, stream, t);
    GPURIR_HIP(hipGetLastError());

    if (!P.items.empty()) {
        AdjArgs a;
        a.rooms = d_rooms;
        a.items = reinterpret_cast<const Item*>(dp + A.off_items);
        a.pos_src = d_pos_src;
        a.pos_rcv = d_pos_rcv;
        a.orv_src = d_orv_src;
        a.orv_rcv = d_orv_rcv;
        a.ktab = reinterpret_cast<const float4*>(dp + A.off_kt);
        a.y = d_y;
        a.gy = d_gy;
        a.rir_cw = rir_cw;
        a.rir_ws = rir_ws;
        a.item_part = item_part;
        a.order = reinterpret_cast<const int*>(dp + A.off_ord);
        a.spkr = A.opt.spkr;
        a.mic = A.opt.mic;
        a.J0 = A.T.J0;
        a.nJ = A.T.nJ;
        a.halfTw = A.T.halfTw;
        a.inv_tw2 = 2.0f / (float)A.Tw;
        a.pi_tw = (float)(M_PI / (double)A.Tw);
        a.ntab_max = A.ntab;
        a.gp_len = A.gp_len;
        a.achs = reinterpret_cast<const unsigned short*>(dp + A.off_act);
        a.abt = A.abt;
        a.adl = A.adl;
        a.agl = A.agl;
        a.aks = A.aks;
        a.acl = A.acl;
        a.prm = A.prm;
        const bool pat = a.spkr != 0 || a.mic != 0;
        const void* fn = A.basis  ? (pat ? reinterpret_cast<const void*>(&ism_adjoint_basis_kernel<true>)
                                         : reinterpret_cast<const void*>(&ism_adjoint_basis_kernel<false>))
                         : A.mask ? reinterpret_cast<const void*>(&ism_adjoint_kernel<true>)
                                  : reinterpret_cast<const void*>(&ism_adjoint_kernel<false>);
        if (int rc = allow_lds(fn, A.lds)) return rc;
        void* args[] = {&a};
        GPURIR_HIP(hipLaunchKernel(fn, dim3((unsigned)P.items.size()), dim3(A.basis ? WGB : WG), args, A.lds,
                                   stream));
    }
    hipLaunchKernelGGL(grad_rir_kernel, dim3((unsigned)((n_rirs + WG - 1) / WG)), dim3(WG), 0, stream, item_part,
                       reinterpret_cast<const int*>(dp + A.off_rib), rir_tail, rir_grad, n_rirs);
    GPURIR_HIP(hipGetLastError());
    hipLaunchKernelGGL(grad_param_kernel, dim3((unsigned)P.rooms.size()), dim3(WG), 0, stream, d_rooms, rir_grad,
                       g_beta, g_src, g_rcv, g_level);
    GPURIR_HIP(hipGetLastError());
    return 0;
}

int launch_adjoint(const std::vector<RoomHost>& rooms, const float* d_pos_src, const float* d_pos_rcv,
                   const float* d_orv_src, const float* d_orv_rcv, const SimOptions& opt, const float* d_y,
                   const float* d_gy, float* g_beta, float* g_src, float* g_rcv, float* g_level, hipStream_t stream) {
    AdjLaunch A;
    if (int rc = adj_prepare(rooms, opt, A)) return rc;
    void* dplan = nullptr;
    if (int rc = workspace(WS_ADJ_PLAN, A.plan_bytes, &dplan)) return rc;
    if (int rc = upload_async(dplan, A.blob.data(), A.plan_bytes, stream)) return rc;
    void* dws = nullptr;
    if (int rc = workspace(WS_GRAD, adj_ws_bytes(A), &dws)) return rc;
    return adj_enqueue(A, dplan, dws, d_pos_src, d_pos_rcv, d_orv_src, d_orv_rcv, d_y, d_gy, g_beta, g_src, g_rcv,
                       g_level, stream);
}
}  // namespace gpurir
