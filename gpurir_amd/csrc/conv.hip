// Stage (d): batched FFT convolution and trajectory overlap-add for gfx950.
//
// Reference path (DavidDiazGuerra/gpuRIR v1.2.0):
//   gpuRIR/__init__.py:177-221 simulateTrajectory (segmentation, float64 OLA on the host)
//   src/gpuRIR_cuda.cu:933-1023 cuda_convolutions (cuFFT R2C/C2R, N_fft = pow2roundup(seg+L-1);
//   here N_fft = 2^a or 3 2^a, whichever is smaller)
//   src/gpuRIR_cuda.cu:463-482  complexPointwiseMulAndScale
//
// Design: a real length-N transform is a complex length-H = N/2 transform of
// z[n] = x[2n] + i x[2n+1] plus an O(H) split.  One workgroup holds one
// length-H complex sequence in LDS (H <= 16384: 136 KiB with one float2 of padding
// per 16 elements) and transforms it in place with a Stockham (self-sorting) FFT:
// passes of radix 16, then one of 8, 4 or 2, then one of 3 when H = 3 2^a; each
// thread reads its butterflies' inputs at stride H/R, applies the pass twiddles, does
// the R-point DFT in registers and writes the outputs at stride Ns after a barrier --
// natural order in and out, about log16(H) + 1 LDS round trips instead of log2(H).
// The inverse is the forward transform of the conjugate (conjugated again on the way
// out).  Segment spectra are computed once per trajectory point and shared by that
// point's receivers through L2.  The overlap-add is a gather: each output sample adds
// the segments that cover it in ascending segment order, in FP64 (deterministic).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/gpurir.h"
#include "internal.h"

namespace gpurir {
namespace {

// The transforms in this file may contract a * b + c into FMAs (the library is built with
// -ffp-contract=off for the image delays, which must round like the CPU oracle's).
#pragma clang fp contract(fast)

constexpr int CWG = 512;

struct ConvArgs {
    const float* src;       // segment source (signal or segment matrix)
    const int* seg_start;   // [P] element offset of segment p in src
    const int* seg_len;     // [P] samples of segment p
    const float* rirs;      // [P][R][L]
    const float2* spec;     // [P][H+1] segment spectra (natural order)
    float2* spec_out;
    const float2* tw;       // [H] e^{-2 pi i m / H}
    const float2* twN;      // [H+1] e^{-2 pi i k / N}
    float* conv;            // [P][R][conv_stride]
    int P, R, L, H;
    int l2, t3;             // H = 2^l2 * 3^t3, t3 in {0, 1}
    int conv_stride;
    // Partitioned convolution when seg_len + L - 1 exceeds one transform: segment p is cut
    // into na blocks of A samples and each RIR into nb blocks of B samples (A + B - 1 <= N);
    // block products go to blk[p][r][a][b][blk_stride] and are added up by conv_assemble_kernel.
    // na = nb = 1: one transform per (p, r), written straight to conv.
    int na, nb, A, B;
    float* blk;
    int blk_stride;
};

__device__ inline float2 cmul(float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline float2 cmulc(float2 a, float2 b) {  // a * conj(b)
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ inline float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }

// LDS position of element i of a transform: one float2 of padding after every 16, so
// the 16 lanes of one LDS cycle hit 16 different bank pairs also when each lane writes
// a run of 16 consecutive elements (the first Stockham pass).
__device__ __forceinline__ int pz(int i) { return i + (i >> 4); }

// W16^m = exp(-2 pi i m / 16), m = 0..7
__device__ constexpr float W16C[8] = {1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508978f,
                                      0.0f, -0.38268343236508978f, -0.70710678118654752f, -0.92387953251128674f};
__device__ constexpr float W16S[8] = {0.0f, -0.38268343236508978f, -0.70710678118654752f, -0.92387953251128674f,
                                      -1.0f, -0.92387953251128674f, -0.70710678118654752f, -0.38268343236508978f};

template <int M>
__device__ __forceinline__ float2 w16mul(float2 a) {
    if (M == 0) return a;
    if (M == 4) return make_float2(a.y, -a.x);  // * (-i)
    return make_float2(a.x * W16C[M] - a.y * W16S[M], a.x * W16S[M] + a.y * W16C[M]);
}

// One radix-2 stage of an R-point decimation-in-frequency DFT held in registers
// (pairs i, i + h inside groups of 2h; twiddle W_{2h}^(i mod h) = W16^((i mod h) 16 / 2h)).
template <int R, int H2, int I>
__device__ __forceinline__ void dif_stage_pairs(float2* v) {
    if constexpr (I < R) {
        if constexpr ((I & H2) == 0) {
            const float2 a = v[I], b = v[I + H2];
            v[I] = cadd(a, b);
            v[I + H2] = w16mul<(I & (H2 - 1)) * (16 / (2 * H2))>(csub(a, b));
        }
        dif_stage_pairs<R, H2, I + 1>(v);
    }
}

template <int R, int H2>
__device__ __forceinline__ void dif_stages(float2* v) {
    if constexpr (H2 >= 1) {
        dif_stage_pairs<R, H2, 0>(v);
        dif_stages<R, H2 / 2>(v);
    }
}

template <int R>
__device__ __forceinline__ constexpr int brev_c(int k) {
    int r = 0;
    for (int b = 1, c = R >> 1; b < R; b <<= 1, c >>= 1)
        if (k & b) r |= c;
    return r;
}

// y_k = sum_r v_r W_R^(r k), natural order in and out (R = 2, 4, 8, 16)
template <int R>
__device__ __forceinline__ void dft_reg(float2* v) {
    dif_stages<R, R / 2>(v);
    float2 t[R];
#pragma unroll
    for (int k = 0; k < R; k++) t[k] = v[brev_c<R>(k)];
#pragma unroll
    for (int k = 0; k < R; k++) v[k] = t[k];
}

// y_k = sum_r v_r W_3^(r k), W_3 = exp(-2 pi i / 3)
__device__ __forceinline__ void dft3(float2* v) {
    constexpr float H3 = 0.86602540378443865f;  // sqrt(3) / 2
    const float2 a = v[0], t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
    const float2 m = make_float2(a.x - 0.5f * t1.x, a.y - 0.5f * t1.y);
    const float2 d = make_float2(H3 * t2.x, H3 * t2.y);
    v[0] = cadd(a, t1);
    v[1] = make_float2(m.x + d.y, m.y - d.x);  // m - i d
    v[2] = make_float2(m.x - d.y, m.y + d.x);  // m + i d
}

template <int R>
__device__ __forceinline__ void dft_any(float2* v) {
    if constexpr (R == 3)
        dft3(v);
    else
        dft_reg<R>(v);
}

// One Stockham pass of radix R over x[0:H] (in place, padded positions pz): butterfly j
// (of H/R) reads x[j + r H/R], multiplies input r by W_H^(r (j mod Ns) H/(Ns R)), does the
// R-point DFT and writes x[(j / Ns) Ns R + (j mod Ns) + r Ns].  twH[m] = W_H^m, m < H.
// Ns is a power of two (the radix-3 pass of an H = 3 2^l2 transform comes last); FIRST
// (Ns == 1) has no twiddles.  Inlined into the kernels, so that x is known to be LDS and
// twH global memory: the twiddle loads of a butterfly are then issued together and do
// not wait behind the LDS traffic.
template <int R, int VPT, bool FIRST>
__device__ __forceinline__ void stockham_pass(float2* x, const float2* __restrict__ twH, int H, int lNs, int step) {
    constexpr int MAXB = VPT / R > 0 ? VPT / R : 1;  // butterflies per thread
    const int nbf = H / R;
    const int Ns = 1 << lNs;
    float2 v[MAXB][R];
#pragma unroll
    for (int q = 0; q < MAXB; q++) {
        const int j = (int)threadIdx.x + q * CWG;
        if (j < nbf) {
            const int k0 = j & (Ns - 1);
            float2 w[R];
            if (!FIRST) {
#pragma unroll
                for (int r = 1; r < R; r++) w[r] = twH[__mul24(k0 * r, step)];
            }
#pragma unroll
            for (int r = 0; r < R; r++) {
                const float2 a = x[pz(j + r * nbf)];
                v[q][r] = (FIRST || r == 0) ? a : cmul(a, w[r]);
            }
            dft_any<R>(v[q]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXB; q++) {
        const int j = (int)threadIdx.x + q * CWG;
        if (j < nbf) {
            const int k0 = j & (Ns - 1);
            const int base = (j - k0) * R + k0;
#pragma unroll
            for (int r = 0; r < R; r++) x[pz(base + r * Ns)] = v[q][r];
        }
    }
    __syncthreads();
}

// Forward DFT of x[0:H] in place (natural order in and out), H = 2^l2 3^t3 with at most
// VPT values per thread in every pass.  VPT sets the register footprint: 16 keeps two
// 512-thread workgroups per CU for H <= 8192 (three for H <= 6144); the largest radix is
// min(16, VPT).  Radix-16 passes first, then one of 8, 4 or 2, then the radix-3 pass.
template <int VPT>
__device__ __forceinline__ void fft_stockham(float2* x, const float2* __restrict__ twH, int H, int l2, int t3) {
    int lNs = 0, bits = l2;
    // step = H / (Ns R) = 2^(l2 - lNs - log2 R) 3^t3
    const int m3 = t3 ? 3 : 1;
    if (bits >= 4) {
        stockham_pass<16, VPT, true>(x, twH, H, 0, 0);
        bits -= 4;
        lNs = 4;
    }
    for (; bits >= 4; bits -= 4, lNs += 4) stockham_pass<16, VPT, false>(x, twH, H, lNs, m3 << (l2 - lNs - 4));
    if (bits == 3) stockham_pass<8, VPT, false>(x, twH, H, lNs, m3 << (l2 - lNs - 3));
    if (bits == 2) stockham_pass<4, VPT, false>(x, twH, H, lNs, m3 << (l2 - lNs - 2));
    if (bits == 1) stockham_pass<2, VPT, false>(x, twH, H, lNs, m3 << (l2 - lNs - 1));
    if (t3) {
        if (l2 == 0)
            stockham_pass<3, VPT, true>(x, twH, H, 0, 0);
        else
            stockham_pass<3, VPT, false>(x, twH, H, l2, 1);
    }
}

// z[n] = x[2n] + i x[2n+1] for x = src[0:len] zero-padded to 2H.
__device__ void load_real(float2* z, const float* __restrict__ src, int len, int H) {
    for (int n = threadIdx.x; n < H; n += blockDim.x) {
        const int e = 2 * n;
        z[pz(n)] = make_float2(e < len ? src[e] : 0.0f, e + 1 < len ? src[e + 1] : 0.0f);
    }
    __syncthreads();
}

// X[k] and X[H-k] of the real transform from the half-length spectrum (natural order).
__device__ inline void split_pair(const float2* z, const float2* __restrict__ twN, int k, int H, float2& Xk,
                                  float2& Xj) {
    const int j = H - k;  // k in [0, H/2], so only j = H wraps (to 0)
    const float2 A = z[pz(k)];
    const float2 Bc = conjf2(z[pz(j == H ? 0 : j)]);
    const float2 E = make_float2(0.5f * (A.x + Bc.x), 0.5f * (A.y + Bc.y));
    const float2 D = csub(A, Bc);
    const float2 O = make_float2(0.5f * D.y, -0.5f * D.x);  // -i D / 2
    const float2 WO = cmul(twN[k], O);
    Xk = cadd(E, WO);
    Xj = conjf2(csub(E, WO));
}

// Samples of block q (of length Q) of a sequence of n samples: 0 for a block past the end.
__device__ inline int block_len(int n, int q, int Q) { return max(0, min(Q, n - q * Q)); }

// One workgroup per (trajectory point p, segment block a): spec[p * na + a].
template <int VPT>
__global__ __launch_bounds__(CWG) void segment_spectrum_kernel(ConvArgs a) {
    extern __shared__ __align__(16) float2 z[];
    const int p = blockIdx.x / a.na, q = blockIdx.x - p * a.na;
    load_real(z, a.src + a.seg_start[p] + (long long)q * a.A, block_len(a.seg_len[p], q, a.A), a.H);
    fft_stockham<VPT>(z, a.tw, a.H, a.l2, a.t3);
    float2* S = a.spec_out + (long long)blockIdx.x * (a.H + 1);
    for (int k = threadIdx.x; k <= a.H / 2; k += blockDim.x) {
        float2 Xk, Xj;
        split_pair(z, a.twN, k, a.H, Xk, Xj);
        S[k] = Xk;
        S[a.H - k] = Xj;
    }
}

// One workgroup per (trajectory point p, receiver r, segment block qa, RIR block qb), job
// ((p R + r) na + qa) nb + qb: conv[p][r][0 : seg_len[p] + L - 1] when na = nb = 1, else the
// product of the two blocks into blk.
template <int VPT>
__global__ __launch_bounds__(CWG) void conv_kernel(ConvArgs a) {
    extern __shared__ __align__(16) float2 z[];
    const int job = blockIdx.x;
    const int qb = job % a.nb, pa = job / a.nb;
    const int qa = pa % a.na, pr = pa / a.na;
    const int p = pr / a.R;
    const int H = a.H;
    const int lb = block_len(a.L, qb, a.B);
    load_real(z, a.rirs + (long long)pr * a.L + (long long)qb * a.B, lb, H);
    fft_stockham<VPT>(z, a.tw, H, a.l2, a.t3);
    const float2* S = a.spec + (long long)(p * a.na + qa) * (H + 1);
    const float scale = 1.0f / (float)H;
    // product and inverse split, pairs (k, H-k) handled by one thread (in place)
    for (int k = threadIdx.x; k <= H / 2; k += blockDim.x) {
        const int j = H - k;
        float2 Xk, Xj;
        split_pair(z, a.twN, k, H, Xk, Xj);
        float2 Yk = cmul(Xk, S[k]), Yj = cmul(Xj, S[j]);
        Yk = make_float2(Yk.x * scale, Yk.y * scale);
        Yj = make_float2(Yj.x * scale, Yj.y * scale);
        const float2 Yjc = conjf2(Yj);
        const float2 E = make_float2(0.5f * (Yk.x + Yjc.x), 0.5f * (Yk.y + Yjc.y));
        const float2 Dm = csub(Yk, Yjc);
        const float2 O = cmulc(make_float2(0.5f * Dm.x, 0.5f * Dm.y), a.twN[k]);  // (Yk - conj Yj) W^-k / 2
        const float2 Zk = make_float2(E.x - O.y, E.y + O.x);                      // E + i O
        const float2 Zj = make_float2(E.x + O.y, -E.y + O.x);                     // conj(E) + i conj(O)
        // both positions were read above by this thread only (pairs are disjoint); the
        // inverse transform is the forward one of the conjugate
        if (k == 0 || 2 * k == H) {  // k == H - k (mod H): one position
            z[pz(k)] = conjf2(Zk);
        } else {
            z[pz(k)] = conjf2(Zk);
            z[pz(j)] = conjf2(Zj);
        }
    }
    __syncthreads();
    fft_stockham<VPT>(z, a.tw, H, a.l2, a.t3);
    float* out;
    int len;
    if (a.na * a.nb == 1) {
        out = a.conv + (long long)pr * a.conv_stride;
        len = a.seg_len[p] + a.L - 1;
    } else {
        out = a.blk + (long long)job * a.blk_stride;
        const int la = block_len(a.seg_len[p], qa, a.A);
        len = la > 0 && lb > 0 ? la + lb - 1 : 0;
    }
    for (int t = threadIdx.x; t < len; t += blockDim.x) {
        const float2 v = z[pz(t >> 1)];
        out[t] = (t & 1) ? -v.y : v.x;
    }
}

// conv[p][r][t] = sum over blocks (qa, qb) of blk[p][r][qa][qb][t - qa A - qb B], in
// ascending (qa, qb) order in FP64: one thread per output sample, rows of the grid's y.
__global__ __launch_bounds__(256) void conv_assemble_kernel(ConvArgs a) {
    const int pr = blockIdx.y, p = pr / a.R;
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int sl = a.seg_len[p];
    if (t >= sl + a.L - 1) return;
    double acc = 0.0;
    for (int qa = 0; qa < a.na; qa++) {
        const int la = block_len(sl, qa, a.A);
        if (la == 0) break;
        for (int qb = 0; qb < a.nb; qb++) {
            const int lb = block_len(a.L, qb, a.B);
            const int u = t - qa * a.A - qb * a.B;
            if (u >= 0 && u < la + lb - 1)
                acc += (double)a.blk[((long long)(pr * a.na + qa) * a.nb + qb) * a.blk_stride + u];
        }
    }
    a.conv[(long long)pr * a.conv_stride + t] = (float)acc;
}

// out[t][r] = sum over segments p covering t of conv[p][r][t - w_ini[p]], FP64, ascending p.
// A workgroup owns 64 consecutive samples (one per lane) and up to 16 receivers at a time
// (wave w takes receivers w, w + 4, w + 8, w + 12): the reads of a wave are 64 consecutive
// floats of one conv row, each thread keeps up to sixteen loads in flight, and the sums go
// through LDS so that the float64 stores of the [t][r] output are contiguous.  The host
// gives each workgroup the segment range of its first sample and the last segment that
// can reach it; the workgroup copies that part of w_ini into LDS (when it has at most
// OLA_MAXSEG entries) and each thread steps from the range to its own sample.
constexpr int OLA_T = 64, OLA_RPT = 4, OLA_RC = 4 * OLA_RPT, OLA_MAXSEG = 512;

// rg = (first segment ending after the workgroup's first sample t0, last segment starting at
// or before t0, last segment starting at or before its last sample); W reads w_ini[p].
template <typename WI>
__device__ __forceinline__ void ola_tile(const float* __restrict__ conv, WI wi, int4 rg, int P, int R, int L,
                                         int conv_stride, long long n_out, double* __restrict__ out, double* tile) {
    const long long t0 = (long long)blockIdx.x * OLA_T;
    const int i = threadIdx.x & (OLA_T - 1), w = threadIdx.x >> 6;
    const long long t = t0 + i;
    const int nt = (int)(n_out - t0 < OLA_T ? n_out - t0 : OLA_T);
    // both bounds are non-decreasing in t; t >= the workgroup's first sample
    int pl = rg.x, ph = rg.y;
    if (t < n_out) {
        while (pl < P && (long long)wi(pl + 1) + L - 1 <= t) pl++;
        while (ph + 1 < P && (long long)wi(ph + 1) <= t) ph++;
    } else {
        ph = pl - 1;
    }
    for (int r0 = 0; r0 < R; r0 += OLA_RC) {
        const int rc = R - r0 < OLA_RC ? R - r0 : OLA_RC;
        double acc[OLA_RPT];
#pragma unroll
        for (int k = 0; k < OLA_RPT; k++) acc[k] = 0.0;
        for (int p = pl; p <= ph; p++) {
            const float* row = conv + ((long long)p * R + r0 + w) * conv_stride + (t - wi(p));
            float v[OLA_RPT];
#pragma unroll
            for (int k = 0; k < OLA_RPT; k++) v[k] = w + 4 * k < rc ? row[(long long)(4 * k) * conv_stride] : 0.0f;
#pragma unroll
            for (int k = 0; k < OLA_RPT; k++) acc[k] += (double)v[k];
        }
#pragma unroll
        for (int k = 0; k < OLA_RPT; k++) tile[i * (OLA_RC + 1) + w + 4 * k] = acc[k];
        __syncthreads();
        for (int e = threadIdx.x; e < nt * rc; e += 256) {
            const int ii = e / rc, u = e - ii * rc;
            out[(t0 + ii) * R + r0 + u] = tile[ii * (OLA_RC + 1) + u];
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void ola_kernel(const float* __restrict__ conv, const int* __restrict__ w_ini,
                                                  const int4* __restrict__ ranges, int P, int R, int L,
                                                  int conv_stride, long long n_out, double* __restrict__ out) {
    __shared__ int wsh[OLA_MAXSEG];
    __shared__ double tile[OLA_T * (OLA_RC + 1)];  // [sample][receiver], padded rows
    const int4 rg = ranges[blockIdx.x];
    const int nst = rg.z + 2 - rg.x;  // w_ini[rg.x .. rg.z + 1] is all a thread reads
    if (nst <= OLA_MAXSEG) {
        for (int e = threadIdx.x; e < nst; e += 256) wsh[e] = w_ini[rg.x + e];
        __syncthreads();
        const int base = rg.x;
        ola_tile(conv, [&](int p) { return wsh[p - base]; }, rg, P, R, L, conv_stride, n_out, out, tile);
    } else {
        ola_tile(conv, [&](int p) { return w_ini[p]; }, rg, P, R, L, conv_stride, n_out, out, tile);
    }
}

// ---- Segment-aligned partitioned convolution (the trajectory's usual case) ----------------
//
// When the segments have S or S + 1 samples (simulateTrajectory's own segmentation) the
// output is cut at the segment starts o_m = w_ini[m] (continued past the signal in steps of
// S_max) and each RIR of point p into blocks at the same boundaries: block j of RIR p covers
// RIR samples [o_{p+j} - o_p, o_{p+j+1} - o_p).  Every product x_p * block_j(h_p) then starts
// at the same window start o_{p+j}, so window m's result is
//   y_m = IFFT( sum_p X_p . FFT(block_{m-p}(h_p)) )
// with one inverse transform per window, and it is shorter than o_{m+2} - o_m (blocks and
// segments differ by at most one sample).  Output block [o_b, o_{b+1}) is
// y_b[0 : S_b] + y_{b-1}[S_{b-1} :], a two-term FP64 sum.  Windows are taken in pairs: the
// blocks of windows 2J and 2J + 1 go into the real and imaginary parts of one complex
// transform of length N = 128 V >= 2 S_max; the products with the (Hermitian) segment
// spectra keep the two apart, so the inverse gives y_2J + i y_2J+1.
//
// A transform unit is 128 threads holding V values each: element tau + 128 i in register i.
// Forward: DFT-V over i in registers, twiddles W_N^(tau k2), then a DFT-128 over tau as
// 16 x 8 (DFT-8 over tau >> 4 after an LDS exchange, twiddles W_128^(a c), DFT-16 over
// tau & 15 after a second exchange, the DFT-16 split over two lanes as two DFT-8s and one
// radix-2 step through a DPP swap).  The spectrum stays in that unit's own (permuted)
// register layout: segment spectra are stored in it and the inverse is the exact adjoint of
// the forward steps in reverse order, so no reordering pass is needed.  The XOR-swizzled
// LDS positions keep every 32-lane group of each exchange on distinct 8-byte bank pairs.
constexpr int UT = 128;

struct WinArgs {
    const float* signal;
    const int* w_ini;      // [P + 1]
    const float* rirs;     // [P][R][L]
    const float2* tw;      // [V + 8][UT] unit twiddles: W_N^(tau k), then W_128^((tau & 15) c)
    float2* spec;          // [P][V][UT] segment spectra / N, unit register layout
    const int* p_lo;       // [n_jobs] first point of window pair J
    const int* job_order;  // [n_conv_wg] job J * R + r of each workgroup, -1: none
    float* yhead;          // [n_jobs][R][ystride] y_2J[0 : S_2J]
    float* ytail;          // [n_jobs][R][ystride] y_2J+1[S_2J+1 : S_2J+1 + S_2J+2]
    double* out;           // [n_out][R]
    int P, R, L, n_sig, s_max, n_jobs, n_conv_wg, ystride;
    long long n_out;
};

__device__ __forceinline__ long long win_start(const WinArgs& a, int m) {
    return m <= a.P ? (long long)a.w_ini[m] : (long long)a.n_sig + (long long)(m - a.P) * a.s_max;
}

__device__ __forceinline__ int e1_pos(int k2, int n1) { return k2 * UT + (n1 ^ ((k2 & 1) << 4)); }
__device__ __forceinline__ int e2_pos(int k2, int c, int a) {
    return k2 * UT + ((c * 16 + a) ^ (((k2 & 1) << 4) ^ ((c >> 1) << 1) ^ ((c & 1) << 3)));
}

__device__ __forceinline__ float swap_xor1(float v) {  // value of lane ^ 1 (quad_perm 1,0,3,2)
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, true));
}

__device__ __forceinline__ float2 swap_ri(float2 v) { return make_float2(v.y, v.x); }

// Raw buffer view of n floats: loads past the end return 0 without a branch (the range
// check of the buffer resource), which is how zero-padded blocks are read.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t float_range(const float* p, int n) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(p), 0, n * (int)sizeof(float), 0x00020000);
}
__device__ __forceinline__ float load_or_zero(__amdgpu_buffer_rsrc_t r, int lane_i, int const_i) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, lane_i * (int)sizeof(float),
                                                                          const_i * (int)sizeof(float), 0));
}

// v, returned through an empty asm that the compiler cannot see into: addresses derived from
// the result are recomputed where it is used, and cannot be hoisted out of a loop or above
// the point where dep is ready.  Keeps the thread's loop-invariant LDS and twiddle addresses
// and its spectrum loads inside the loop, where their registers are briefly live; hoisted
// they hold ~100 VGPRs across the whole loop and cut the occupancy to one or two waves per SIMD.
__device__ __forceinline__ int pinned_v(int v) {
    asm volatile("" : "+v"(v));
    return v;
}
__device__ __forceinline__ float2 pinned_f2(float2 v) {
    asm volatile("" : "+v"(v.x), "+v"(v.y));
    return v;
}
__device__ __forceinline__ int pinned_after(int v, float dep) {
    asm volatile("" : "+v"(v) : "v"(dep));
    return v;
}

// Good-Thomas DFT-24 = 3 x 8 (coprime, no twiddles): input x[(8 n1 + 3 n2) mod 24],
// output X[(16 k1 + 9 k2) mod 24].
__device__ __forceinline__ void dft24(float2* v) {
    float2 b[3][8];
#pragma unroll
    for (int n2 = 0; n2 < 8; n2++) {
        float2 t[3] = {v[(3 * n2) % 24], v[(8 + 3 * n2) % 24], v[(16 + 3 * n2) % 24]};
        dft3(t);
#pragma unroll
        for (int k1 = 0; k1 < 3; k1++) b[k1][n2] = t[k1];
    }
#pragma unroll
    for (int k1 = 0; k1 < 3; k1++) dft_reg<8>(b[k1]);
#pragma unroll
    for (int k1 = 0; k1 < 3; k1++)
#pragma unroll
        for (int k2 = 0; k2 < 8; k2++) v[(16 * k1 + 9 * k2) % 24] = b[k1][k2];
}

// The same transforms for input whose upper half is known to be zero (v[R/2 ..] unread): the
// zero-padded blocks fill at most half of a transform (N >= 2 S_max).
template <bool Z0, bool Z1, bool Z2>
__device__ __forceinline__ void dft3z(float2* v) {  // Zj: v[j] is zero
    constexpr float H3 = 0.86602540378443865f;
    const float2 zero = make_float2(0.0f, 0.0f);
    if constexpr (Z1 && Z2) {
        const float2 a = Z0 ? zero : v[0];
        v[0] = a;
        v[1] = a;
        v[2] = a;
    } else {
        float2 t1, t2;
        if constexpr (Z1) {
            t1 = v[2];
            t2 = make_float2(-v[2].x, -v[2].y);
        } else if constexpr (Z2) {
            t1 = v[1];
            t2 = v[1];
        } else {
            t1 = cadd(v[1], v[2]);
            t2 = csub(v[1], v[2]);
        }
        float2 m;
        if constexpr (Z0) {
            m = make_float2(-0.5f * t1.x, -0.5f * t1.y);
            v[0] = t1;
        } else {
            const float2 a = v[0];
            m = make_float2(a.x - 0.5f * t1.x, a.y - 0.5f * t1.y);
            v[0] = cadd(a, t1);
        }
        const float2 d = make_float2(H3 * t2.x, H3 * t2.y);
        v[1] = make_float2(m.x + d.y, m.y - d.x);
        v[2] = make_float2(m.x - d.y, m.y + d.x);
    }
}

template <int N2 = 0>
__device__ __forceinline__ void pfa24_low_cols(const float2* v, float2 (&b)[3][8]) {
    if constexpr (N2 < 8) {
        constexpr int i0 = (3 * N2) % 24, i1 = (8 + 3 * N2) % 24, i2 = (16 + 3 * N2) % 24;
        float2 t[3] = {i0 < 12 ? v[i0] : make_float2(0.0f, 0.0f), i1 < 12 ? v[i1] : make_float2(0.0f, 0.0f),
                       i2 < 12 ? v[i2] : make_float2(0.0f, 0.0f)};
        dft3z<(i0 >= 12), (i1 >= 12), (i2 >= 12)>(t);
#pragma unroll
        for (int k1 = 0; k1 < 3; k1++) b[k1][N2] = t[k1];
        pfa24_low_cols<N2 + 1>(v, b);
    }
}

__device__ __forceinline__ void dft24_low(float2* v) {
    float2 b[3][8];
    pfa24_low_cols(v, b);
#pragma unroll
    for (int k1 = 0; k1 < 3; k1++) dft_reg<8>(b[k1]);
#pragma unroll
    for (int k1 = 0; k1 < 3; k1++)
#pragma unroll
        for (int k2 = 0; k2 < 8; k2++) v[(16 * k1 + 9 * k2) % 24] = b[k1][k2];
}

template <int R, int I = 0>
__device__ __forceinline__ void dif_first_low(float2* v) {  // first DIF stage, v[R/2 ..] zero
    if constexpr (I < R / 2) {
        v[I + R / 2] = w16mul<I * (16 / R)>(v[I]);
        dif_first_low<R, I + 1>(v);
    }
}

template <int R>
__device__ __forceinline__ void dft_reg_low(float2* v) {
    dif_first_low<R>(v);
    dif_stages<R, R / 4>(v);
    float2 t[R];
#pragma unroll
    for (int k = 0; k < R; k++) t[k] = v[brev_c<R>(k)];
#pragma unroll
    for (int k = 0; k < R; k++) v[k] = t[k];
}

template <int V>
__device__ __forceinline__ void dft_v_low(float2* v) {
    if constexpr (V == 24)
        dft24_low(v);
    else
        dft_reg_low<V>(v);
}

template <int V>
__device__ __forceinline__ void dft_v(float2* v) {
    if constexpr (V == 24)
        dft24(v);
    else
        dft_reg<V>(v);
}

// Unnormalized inverse (conjugate) DFT: swap(DFT(swap(v))), swap(x + iy) = y + ix.
template <int R, typename F>
__device__ __forceinline__ void conj_dft(float2* v, F dft) {
#pragma unroll
    for (int k = 0; k < R; k++) v[k] = swap_ri(v[k]);
    dft(v);
#pragma unroll
    for (int k = 0; k < R; k++) v[k] = swap_ri(v[k]);
}

// The thread's twiddles, made in registers from V/8 + 1 values loaded once per kernel instead
// of 30 table loads per transform (whose latency the transform cannot hide): W_N^(tau k) for
// k < V applied as W_N^(tau j) then W_N^(8 tau h), k = 8 h + j, and W_128^(a c),
// a = tau mod 16, c < 8, with the powers of W_N^tau and W_128^a built by squaring (at most
// four roundings more than a table entry).
template <int V>
struct UnitSeeds {
    float2 w1, wa, wh[V / 8];  // W_N^tau, W_128^a, wh[h] = W_N^(8 h tau) for h >= 1
};

template <int V>
__device__ __forceinline__ UnitSeeds<V> load_seeds(const float2* __restrict__ tw, int tau) {
    UnitSeeds<V> sd;
    const float2* ut = tw + tau;
    sd.w1 = ut[UT];
    sd.wa = ut[(V + 1) * UT];
#pragma unroll
    for (int h = 1; h < V / 8; h++) sd.wh[h] = ut[8 * h * UT];
    return sd;
}

// p[j] = w^j for j = 1..7 (p[0] unused), made where it is used (not hoisted out of a loop,
// where the powers would hold 14 VGPRs throughout)
__device__ __forceinline__ void powers7(float2 w, float2 (&p)[8]) {
    w = pinned_f2(w);
    p[1] = w;
    p[2] = cmul(w, w);
    p[3] = cmul(p[2], w);
    p[4] = cmul(p[2], p[2]);
    p[5] = cmul(p[4], w);
    p[6] = cmul(p[4], p[2]);
    p[7] = cmul(p[4], p[3]);
}

// z[k] *= W_N^(tau k) (CONJ: by its conjugate), k = 8 h + j, as z[k] * W_N^(tau j) * W_N^(8 tau h)
template <int V, bool CONJ>
__device__ __forceinline__ void tau_twiddles(float2 (&z)[V], const UnitSeeds<V>& sd) {
    float2 lo[8];
    powers7(sd.w1, lo);
#pragma unroll
    for (int k = 1; k < V; k++) {
        const int j = k & 7, h = k >> 3;
        if (j != 0) z[k] = CONJ ? cmulc(z[k], lo[j]) : cmul(z[k], lo[j]);
        if (h != 0) z[k] = CONJ ? cmulc(z[k], pinned_f2(sd.wh[h])) : cmul(z[k], pinned_f2(sd.wh[h]));
    }
}

// W16^e for the radix-2 step of the lane-pair DFT-16 (odd lanes), 1 on even lanes.
__device__ __forceinline__ float2 lane_w16(int e, bool odd) {
    return odd ? make_float2(W16C[e], W16S[e]) : make_float2(1.0f, 0.0f);
}

// Forward transform of one unit, in two parts.  unit_forward_front: z[i] = element tau + 128 i
// on entry, the elements past N/2 zero (z[V/2 ..] is not read); runs the DFT-V, the twiddles, the DFT-8 over tau >> 4 and both exchanges, and
// leaves the last stage's input in buf (z is free on return).  unit_forward_back: the lane-pair
// DFT-16 over tau & 15, one chunk of eight outputs at a time, each handed to sink(t, out) --
// out[e] is register 8 t + e of the spectrum in the unit's layout.  buf: V * 128 float2 of LDS
// owned by the unit.  All threads of the workgroup call the front together (it synchronizes
// the workgroup); the back only reads buf.
template <int V>
__device__ __forceinline__ void unit_forward_front(float2 (&z)[V], float2* buf, const UnitSeeds<V>& sd,
                                                   const float2* __restrict__ tw, int tau) {
    tau = pinned_v(tau);
    dft_v_low<V>(z);
    tau_twiddles<V, false>(z, sd);
    const int a = tau & 15, g = tau >> 4;
    float2 w[8];
    const float2* ut = tw + tau;
#pragma unroll
    for (int c = 1; c < 8; c++) w[c] = ut[(V + c) * UT];  // in flight across the barriers
    __syncthreads();  // the previous transform's last reads of buf are done
#pragma unroll
    for (int k2 = 0; k2 < V; k2++) buf[e1_pos(k2, tau)] = z[k2];
    __syncthreads();
#pragma unroll
    for (int t = 0; t < V / 8; t++) {
        // chunk t's reads after chunk t - 1's results: one chunk of temporaries at a time
        const int gt = t == 0 ? g : pinned_after(g, z[8 * t - 1].x);
        float2 u[8];
#pragma unroll
        for (int b = 0; b < 8; b++) u[b] = buf[e1_pos(gt + 8 * t, a + 16 * b)];
        dft_reg<8>(u);
        z[8 * t] = u[0];
#pragma unroll
        for (int c = 1; c < 8; c++) z[8 * t + c] = cmul(u[c], w[c]);
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < V / 8; t++)
#pragma unroll
        for (int c = 0; c < 8; c++) buf[e2_pos(g + 8 * t, c, a)] = z[8 * t + c];
    __syncthreads();
}

template <int V, typename F>
__device__ __forceinline__ void unit_forward_back(const float2* buf, int tau, F&& sink) {
    const bool odd = tau & 1;
    const float sgn = odd ? -1.0f : 1.0f;
    const int cc = (tau >> 1) & 7, gg = tau >> 4;
#pragma unroll
    for (int t = 0; t < V / 8; t++) {
        float2 x[8];
#pragma unroll
        for (int ap = 0; ap < 8; ap++) x[ap] = buf[e2_pos(gg + 8 * t, cc, 2 * ap + (odd ? 1 : 0))];
        dft_reg<8>(x);
        float2 out[8];
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const float2 m = e == 0 ? x[e] : cmul(x[e], lane_w16(e, odd));
            const float2 r = make_float2(swap_xor1(m.x), swap_xor1(m.y));
            out[e] = make_float2(fmaf(sgn, m.x, r.x), fmaf(sgn, m.y, r.y));  // even: P0 + T, odd: P0 - T
        }
        sink(t, out);
    }
}

// The whole forward transform, in place: the spectrum in the unit's register layout on exit.
template <int V>
__device__ __forceinline__ void unit_forward(float2 (&z)[V], float2* buf, const UnitSeeds<V>& sd,
                                             const float2* __restrict__ tw, int tau) {
    unit_forward_front<V>(z, buf, sd, tw, tau);
    unit_forward_back<V>(buf, tau, [&](int t, const float2 (&out)[8]) {
#pragma unroll
        for (int e = 0; e < 8; e++) z[8 * t + e] = out[e];
    });
}

// Adjoint of unit_forward (N times the inverse DFT), in place: the spectrum in the unit's
// register layout on entry, element tau + 128 i in z[i] on exit.
template <int V>
__device__ __forceinline__ void unit_inverse(float2 (&z)[V], float2* buf, const UnitSeeds<V>& sd, int tau) {
    tau = pinned_v(tau);
    const bool odd = tau & 1;
    const float sgn = odd ? -1.0f : 1.0f;
    const int cc = (tau >> 1) & 7, gg = tau >> 4;
    __syncthreads();  // the previous transform's last reads of buf are done
#pragma unroll
    for (int t = 0; t < V / 8; t++) {
        float2 x[8];
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const float2 m = z[8 * t + e];
            const float2 r = make_float2(swap_xor1(m.x), swap_xor1(m.y));
            const float2 q = make_float2(fmaf(sgn, m.x, r.x), fmaf(sgn, m.y, r.y));  // even: X0 + X1, odd: X0 - X1
            x[e] = e == 0 ? q : cmulc(q, lane_w16(e, odd));
        }
        conj_dft<8>(x, [](float2* v) { dft_reg<8>(v); });
#pragma unroll
        for (int ap = 0; ap < 8; ap++) buf[e2_pos(gg + 8 * t, cc, 2 * ap + (odd ? 1 : 0))] = x[ap];
    }
    __syncthreads();
    const int a = tau & 15, g = tau >> 4;
    float2 w[8];
    powers7(sd.wa, w);
#pragma unroll
    for (int t = 0; t < V / 8; t++) {
        float2 u[8];
        u[0] = buf[e2_pos(g + 8 * t, 0, a)];
#pragma unroll
        for (int c = 1; c < 8; c++) u[c] = cmulc(buf[e2_pos(g + 8 * t, c, a)], w[c]);
        conj_dft<8>(u, [](float2* v) { dft_reg<8>(v); });
#pragma unroll
        for (int b = 0; b < 8; b++) z[8 * t + b] = u[b];
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < V / 8; t++)
#pragma unroll
        for (int b = 0; b < 8; b++) buf[e1_pos(g + 8 * t, a + 16 * b)] = z[8 * t + b];
    __syncthreads();
#pragma unroll
    for (int k2 = 0; k2 < V; k2++) z[k2] = buf[e1_pos(k2, tau)];
    tau_twiddles<V, true>(z, sd);
    conj_dft<V>(z, [](float2* v) { dft_v<V>(v); });
}

// One 128-thread workgroup per segment p: spec[p] = FFT_N(x_p zero-padded) / N.
template <int V>
__global__ __launch_bounds__(UT) void window_spectrum_kernel(WinArgs a) {
    __shared__ float2 buf[V * UT];
    const int p = blockIdx.x, tau = threadIdx.x;
    const int w0 = a.w_ini[p], len = a.w_ini[p + 1] - w0;
    const __amdgpu_buffer_rsrc_t x = float_range(a.signal + w0, len);  // len <= N / 2
    float2 z[V];
#pragma unroll
    for (int i = 0; i < V / 2; i++) z[i] = make_float2(load_or_zero(x, tau, UT * i), 0.0f);
    unit_forward<V>(z, buf, load_seeds<V>(a.tw, tau), a.tw, tau);
    const float sc = 1.0f / (float)(V * UT);
    float2* S = a.spec + (size_t)p * V * UT;
#pragma unroll
    for (int i = 0; i < V; i++) S[i * UT + tau] = make_float2(z[i].x * sc, z[i].y * sc);
}

// One 256-thread workgroup per (window pair J, receiver r): two units take alternate points
// of the pair and each accumulates sum_p X_p . FFT(block pair); unit 1 then hands its sum to
// unit 0 through LDS and ends, and unit 0 transforms the total back (s_barrier waits only for
// the waves that have not ended): window 2J is the real part, window 2J + 1 the imaginary.
// amdgpu_waves_per_eu(3): at most 168 VGPRs, three waves per SIMD (three workgroups per CU,
// as the LDS allows); unconstrained, the compiler takes ~173 and halves the waves per SIMD.
template <int V>
__global__ __launch_bounds__(2 * UT) __attribute__((amdgpu_waves_per_eu(3))) void window_conv_kernel(WinArgs a) {
    __shared__ float2 bufs[2][V * UT];
    const int q = a.job_order[blockIdx.x];
    if (q < 0) return;  // whole workgroup
    const int J = q / a.R, r = q - J * a.R;
    const int u = __builtin_amdgcn_readfirstlane(threadIdx.x >> 7);  // unit: the same for a whole wave
    const int tau = threadIdx.x & (UT - 1);
    float2* buf = bufs[u];
    const int m0 = 2 * J;
    const int p_lo = a.p_lo[J], p_hi = min(a.P - 1, m0 + 1);
    const int n_it = p_hi >= p_lo ? (p_hi - p_lo + 2) / 2 : 0;
    const long long o0 = win_start(a, m0), o1 = win_start(a, m0 + 1), o2 = win_start(a, m0 + 2);
    // RIR blocks of point p for windows m0 (real part) and m0 + 1 (imaginary part), zero
    // past their ends and for points outside [p_lo, p_hi]; at most S_max <= N / 2 samples
    auto load_blocks = [&](int p, float2 (&z)[V]) {
        int len0 = 0, len1 = 0;
        long long s0 = 0, s1 = 0;
        if (p <= p_hi) {
            const long long op = a.w_ini[p];
            s0 = o0 - op;
            s1 = o1 - op;
            len0 = (p <= m0 && s0 < a.L) ? (int)(min(o1 - op, (long long)a.L) - s0) : 0;
            len1 = s1 < a.L ? (int)(min(o2 - op, (long long)a.L) - s1) : 0;
        }
        const float* h = a.rirs + ((size_t)min(p, p_hi) * a.R + r) * a.L;
        const __amdgpu_buffer_rsrc_t h0 = float_range(h + (len0 > 0 ? s0 : 0), len0);
        const __amdgpu_buffer_rsrc_t h1 = float_range(h + (len1 > 0 ? s1 : 0), len1);
#pragma unroll
        for (int i = 0; i < V / 2; i++) z[i] = make_float2(load_or_zero(h0, tau, UT * i), load_or_zero(h1, tau, UT * i));
    };
    const UnitSeeds<V> sd = load_seeds<V>(a.tw, tau);
    float2 acc[V];
#pragma unroll
    for (int i = 0; i < V; i++) acc[i] = make_float2(0.0f, 0.0f);
    float2 z[V];
    if (n_it > 0) load_blocks(p_lo + u, z);
    // software pipeline: the next point's blocks load while the last stage of this point's
    // transform runs and its products accumulate
    for (int it = 0; it < n_it; it++) {
        const int p = p_lo + 2 * it + u;
        const bool act = p <= p_hi;  // the same for the whole unit
        unit_forward_front<V>(z, buf, sd, a.tw, tau);
        // the spectrum of point p (two chunks of it), then the next point's blocks: loads return
        // in issue order, so the first chunks do not wait for the HBM reads of the blocks
        // the spectrum of point p chunk by chunk: chunk 0 before the next point's blocks (loads
        // return in issue order, so it does not wait for their HBM reads), chunk t + 1 at the end
        // of chunk t
        const float2* X = a.spec + (size_t)p * V * UT;
        float2 xs[8];
        if (act) {
            const float2* X0 = X + pinned_v(tau);
#pragma unroll
            for (int e = 0; e < 8; e++) xs[e] = X0[e * UT];
        }
        if (it + 1 < n_it) load_blocks(p + 2, z);
        if (act) {
            unit_forward_back<V>(buf, tau, [&](int t, const float2 (&out)[8]) {
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    float2& c = acc[8 * t + e];
                    c.x = fmaf(out[e].x, xs[e].x, fmaf(-out[e].y, xs[e].y, c.x));
                    c.y = fmaf(out[e].x, xs[e].y, fmaf(out[e].y, xs[e].x, c.y));
                }
                if (t + 1 < V / 8) {
                    const float2* Xn = X + 8 * (t + 1) * UT + pinned_after(tau, out[7].x);
#pragma unroll
                    for (int e = 0; e < 8; e++) xs[e] = Xn[e * UT];
                }
            });
        }
    }
    __syncthreads();  // the last transform's reads of both buffers are done
    if (u == 1) {
#pragma unroll
        for (int i = 0; i < V; i++) bufs[1][i * UT + tau] = acc[i];
    }
    __syncthreads();
    if (u == 1) return;
#pragma unroll
    for (int i = 0; i < V; i++) {
        const float2 o = bufs[1][i * UT + tau];
        acc[i] = make_float2(acc[i].x + o.x, acc[i].y + o.y);
    }
    unit_inverse<V>(acc, bufs[0], sd, tau);
    // sample n = tau + 128 i of windows 2J (real part) and 2J + 1 (imaginary part).  Block
    // 2J + 1 of the output lies under these two windows only: it is written here, as the FP64
    // sum of the tail of window 2J and the head of window 2J + 1.  The head of window 2J and
    // the tail of window 2J + 1 go to yhead / ytail for the seam blocks 2J and 2J + 2, which the
    // neighbouring jobs share (window_seam_kernel).
    const int S0 = (int)(o1 - o0), S1 = (int)(o2 - o1), S2 = (int)(win_start(a, m0 + 3) - o2);
    float* head = a.yhead + ((size_t)J * a.R + r) * a.ystride;
    float* tail = a.ytail + ((size_t)J * a.R + r) * a.ystride;
    float* re = reinterpret_cast<float*>(bufs[1]);  // unit 1's buffer: read for the last time before the inverse
#pragma unroll
    for (int i = 0; i < V; i++) {
        const int n = tau + UT * i;
        re[n] = acc[i].x;
        if (n < S0) head[n] = acc[i].x;
        if (n >= S1 && n < S1 + S2) tail[n - S1] = acc[i].y;
    }
    __syncthreads();  // unit 0's two waves
#pragma unroll
    for (int i = 0; i < V; i++) {
        const int k = tau + UT * i;
        if (k < S1 && o1 + k < a.n_out) a.out[(o1 + k) * a.R + r] = (double)re[S0 + k] + (double)acc[i].y;
    }
}

// Seam block 2J (J = 0 .. n_jobs): out[o_2J + k][r] = head_J[k] + tail_{J-1}[k] in FP64, the
// head of window 2J and the tail of window 2J - 1.  A workgroup owns WSEAM_T samples of one
// block and all receivers, eight at a time: the reads run along the rows, the sums go through
// LDS so that the [t][r] stores are contiguous.
constexpr int WSEAM_T = 128;

__global__ __launch_bounds__(256) void window_seam_kernel(WinArgs a) {
    __shared__ double tile[WSEAM_T * 9];
    const int chunks = (a.s_max + WSEAM_T - 1) / WSEAM_T;
    const int J = blockIdx.x / chunks, c = blockIdx.x - J * chunks;
    const long long ob = win_start(a, 2 * J);
    const int k0 = c * WSEAM_T;
    const long long len = min(win_start(a, 2 * J + 1) - ob, a.n_out - ob);  // samples of the block
    const int nt = (int)max(0LL, min((long long)WSEAM_T, len - k0));
    if (nt == 0) return;  // whole workgroup
    const int i = threadIdx.x & (WSEAM_T - 1), j0 = threadIdx.x >> 7;
    const bool has_head = J < a.n_jobs && i < nt, has_tail = J >= 1 && i < nt;
    for (int r0 = 0; r0 < a.R; r0 += 8) {
        const int rc = min(8, a.R - r0);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int j = j0 + 2 * k;
            if (j < rc) {
                const int r = r0 + j;
                double v = 0.0;
                if (has_head) v = (double)a.yhead[((size_t)J * a.R + r) * a.ystride + k0 + i];
                if (has_tail) v += (double)a.ytail[((size_t)(J - 1) * a.R + r) * a.ystride + k0 + i];
                tile[i * 9 + j] = v;
            }
        }
        __syncthreads();
        for (int e = threadIdx.x; e < nt * rc; e += 256) {
            const int ii = e / rc, jj = e - ii * rc;
            a.out[(ob + k0 + ii) * a.R + r0 + jj] = tile[ii * 9 + jj];
        }
        __syncthreads();
    }
}

struct FftTables {
    std::vector<float2> tw, twN;
};

FftTables make_fft_tables(int H) {
    FftTables T;
    T.tw.resize(std::max(H, 1));
    for (int m = 0; m < H; m++) {
        const double a = -2.0 * M_PI * (double)m / (double)H;
        T.tw[m] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    T.twN.resize(H + 1);
    for (int k = 0; k <= H; k++) {
        const double a = -2.0 * M_PI * (double)k / (2.0 * (double)H);
        T.twN[k] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    return T;
}

// Twiddle tables of length-H transforms, built once per (device, H) and kept for the
// life of the process (a few hundred KiB at most).
std::mutex g_fft_mu;
std::map<std::pair<int, int>, float2*> g_fft_tables;

int fft_tables_device(int H, const float2** tw, const float2** twN) {
    int dev = 0;
    GPURIR_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_fft_mu);
    auto key = std::make_pair(dev, H);
    auto f = g_fft_tables.find(key);
    if (f == g_fft_tables.end()) {
        FftTables T = make_fft_tables(H);
        float2* d = nullptr;
        const size_t n = T.tw.size() + T.twN.size();
        if (hipMalloc(reinterpret_cast<void**>(&d), sizeof(float2) * n) != hipSuccess) {
            (void)hipGetLastError();
            set_error("out of device memory for the FFT tables");
            return -3;
        }
        GPURIR_HIP(hipMemcpy(d, T.tw.data(), sizeof(float2) * T.tw.size(), hipMemcpyHostToDevice));
        GPURIR_HIP(hipMemcpy(d + T.tw.size(), T.twN.data(), sizeof(float2) * T.twN.size(), hipMemcpyHostToDevice));
        f = g_fft_tables.emplace(key, d).first;
    }
    *tw = f->second;
    *twN = f->second + std::max(H, 1);
    return 0;
}

// Twiddles of a transform unit of N = 128 V points in the order its threads load them:
// row k < V holds W_N^(tau k) for tau = 0..127, row V + c holds W_128^((tau mod 16) c), so that
// every twiddle load of a wave reads 512 contiguous bytes.  Built once per (device, V).
std::map<std::pair<int, int>, float2*> g_unit_tables;

int unit_tables_device(int V, const float2** tw) {
    int dev = 0;
    GPURIR_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_fft_mu);
    auto key = std::make_pair(dev, V);
    auto f = g_unit_tables.find(key);
    if (f == g_unit_tables.end()) {
        const int N = V * UT;
        std::vector<float2> T((size_t)(V + 8) * UT);
        for (int k = 0; k < V + 8; k++)
            for (int t = 0; t < UT; t++) {
                const long long m = k < V ? (long long)t * k : (long long)V * (t & 15) * (k - V);
                const double ang = -2.0 * M_PI * (double)(m % N) / (double)N;
                T[(size_t)k * UT + t] = make_float2((float)std::cos(ang), (float)std::sin(ang));
            }
        float2* d = nullptr;
        if (hipMalloc(reinterpret_cast<void**>(&d), sizeof(float2) * T.size()) != hipSuccess) {
            (void)hipGetLastError();
            set_error("out of device memory for the FFT tables");
            return -3;
        }
        GPURIR_HIP(hipMemcpy(d, T.data(), sizeof(float2) * T.size(), hipMemcpyHostToDevice));
        f = g_unit_tables.emplace(key, d).first;
    }
    *tw = f->second;
    return 0;
}

// Transform length for a linear convolution of n outputs: the smallest N = 2^a or 3 2^a
// (N even) with N >= n.  The reference rounds up to a power of two (pow2roundup,
// src/gpuRIR_cuda.cu:95-104); any N >= n gives the same linear convolution, and 3 2^a
// saves up to a third of the work (9249 outputs: 12288 instead of 16384).
unsigned fft_length(unsigned n) {
    unsigned p = 2;
    while (p < n) p <<= 1;
    unsigned q = 6;
    while (q < n) q <<= 1;
    return std::min(p, q);
}

size_t al(size_t x) { return (x + 255) / 256 * 256; }

// Shared driver: segments of src given by (seg_start, seg_len) per point.
// Writes conv [P][R][conv_stride] into *conv_out (workspace) and returns the stride.
int run_conv(const float* d_src, const std::vector<int>& seg_start, const std::vector<int>& seg_len, const float* d_rirs,
             int P, int R, int L, hipStream_t stream, float** conv_out, int* conv_stride, void** extra,
             size_t extra_bytes) {
    int seg_max = 0;
    for (int v : seg_len) seg_max = std::max(seg_max, v);
    seg_max = std::max(seg_max, 1);
    GPURIR_CHECK((long long)seg_max + L - 1 < (1LL << 30), "convolution too long");
    // one transform of N <= 32768 points per (p, r) if the linear convolution fits, else
    // blocks of the segment (A samples) and of the RIR (B samples) with A + B - 1 = 32768
    constexpr int NMAX = 32768;
    int A = seg_max, B = L, na = 1, nb = 1;
    if (seg_max + L - 1 > NMAX) {
        if (L <= NMAX / 2) {
            A = NMAX - L + 1;
        } else if (seg_max <= NMAX / 2) {
            B = NMAX - seg_max + 1;
        } else {
            A = NMAX / 2;
            B = NMAX / 2 + 1;
        }
        na = (seg_max + A - 1) / A;
        nb = (L + B - 1) / B;
    }
    const unsigned n_lin = (unsigned)std::max(A + B - 1, 2);
    const unsigned N = fft_length(n_lin);  // <= 32768
    const int H = (int)(N / 2);
    const int t3 = H % 3 == 0 ? 1 : 0;
    int l2 = 0;
    while (((t3 ? 3 : 1) << l2) < H) l2++;
    const float2 *d_tw = nullptr, *d_twN = nullptr;
    if (int rc = fft_tables_device(H, &d_tw, &d_twN)) return rc;
    const int stride = (int)((seg_max + L - 1 + 63) / 64 * 64);
    const int blk_stride = (int)((A + B - 1 + 63) / 64 * 64);
    const long long n_jobs = (long long)P * R * na * nb;
    GPURIR_CHECK(n_jobs < (1LL << 31) && (long long)P * na < (1LL << 31), "too many convolution blocks");
    const size_t b_ss = al(sizeof(int) * P), b_sl = al(sizeof(int) * P);
    const size_t b_spec = al(sizeof(float2) * (size_t)P * na * (H + 1));
    const size_t b_conv = al(sizeof(float) * (size_t)P * R * stride);
    const size_t b_blk = n_jobs > (long long)P * R ? al(sizeof(float) * (size_t)n_jobs * blk_stride) : 0;
    const size_t b_small = b_ss + b_sl;
    void* ws = nullptr;
    if (int rc = workspace(WS_CONV, b_small + b_spec + b_conv + b_blk + al(extra_bytes), &ws)) return rc;
    unsigned char* w = static_cast<unsigned char*>(ws);
    std::vector<unsigned char> host(b_small, 0);
    std::memcpy(host.data(), seg_start.data(), sizeof(int) * P);
    std::memcpy(host.data() + b_ss, seg_len.data(), sizeof(int) * P);
    if (int rc = upload_async(w, host.data(), b_small, stream)) return rc;
    ConvArgs a;
    a.src = d_src;
    a.seg_start = reinterpret_cast<const int*>(w);
    a.seg_len = reinterpret_cast<const int*>(w + b_ss);
    a.tw = d_tw;
    a.twN = d_twN;
    a.spec_out = reinterpret_cast<float2*>(w + b_small);
    a.spec = a.spec_out;
    a.conv = reinterpret_cast<float*>(w + b_small + b_spec);
    a.rirs = d_rirs;
    a.P = P;
    a.R = R;
    a.L = L;
    a.H = H;
    a.l2 = l2;
    a.t3 = t3;
    a.conv_stride = stride;
    a.na = na;
    a.nb = nb;
    a.A = A;
    a.B = B;
    a.blk = b_blk ? reinterpret_cast<float*>(w + b_small + b_spec + b_conv) : nullptr;
    a.blk_stride = blk_stride;
    const size_t lds = sizeof(float2) * (size_t)(H + H / 16);  // padded positions
    const bool big = H > 16 * CWG;  // 32 values per thread
    const void* fs = big ? reinterpret_cast<const void*>(&segment_spectrum_kernel<32>)
                         : reinterpret_cast<const void*>(&segment_spectrum_kernel<16>);
    const void* fc = big ? reinterpret_cast<const void*>(&conv_kernel<32>)
                         : reinterpret_cast<const void*>(&conv_kernel<16>);
    if (lds > 64 * 1024) {
        GPURIR_HIP(hipFuncSetAttribute(fs, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        GPURIR_HIP(hipFuncSetAttribute(fc, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    }
    void* args[] = {&a};
    GPURIR_HIP(hipLaunchKernel(fs, dim3((unsigned)(P * na)), dim3(CWG), args, lds, stream));
    GPURIR_HIP(hipLaunchKernel(fc, dim3((unsigned)n_jobs), dim3(CWG), args, lds, stream));
    if (b_blk) {
        GPURIR_CHECK(P * R <= 65535, "too many (point, receiver) rows for a partitioned convolution");
        hipLaunchKernelGGL(conv_assemble_kernel, dim3((unsigned)((seg_max + L - 1 + 255) / 256), (unsigned)(P * R)),
                           dim3(256), 0, stream, a);
        GPURIR_HIP(hipGetLastError());
    }
    *conv_out = a.conv;
    *conv_stride = stride;
    if (extra) *extra = (extra_bytes > 0) ? (w + b_small + b_spec + b_conv + b_blk) : nullptr;
    return 0;
}

// The segment-aligned partitioned convolution applies when the segments start at 0 and have
// S or S + 1 samples with 256 <= S_max <= 1536 (transforms of 1024, 2048 or 3072 points).
// Shorter segments would pad each block to at least 1024 points, longer ones need larger
// transforms: those go through run_conv and the gather overlap-add.
bool window_path_applies(const int* w_ini, int P) {
    if (w_ini[0] != 0) return false;
    int s_min = w_ini[1] - w_ini[0], s_max = s_min;
    for (int p = 1; p < P; p++) {
        const int s = w_ini[p + 1] - w_ini[p];
        s_min = std::min(s_min, s);
        s_max = std::max(s_max, s);
    }
    return s_min >= 1 && s_max - s_min <= 1 && s_max >= 256 && 2 * s_max <= 24 * UT;
}

template <int V>
int launch_window(WinArgs& a, int n_seam_wg, hipStream_t st) {
    hipLaunchKernelGGL(window_spectrum_kernel<V>, dim3((unsigned)a.P), dim3(UT), 0, st, a);
    GPURIR_HIP(hipGetLastError());
    hipLaunchKernelGGL(window_conv_kernel<V>, dim3((unsigned)a.n_conv_wg), dim3(2 * UT), 0, st, a);
    GPURIR_HIP(hipGetLastError());
    hipLaunchKernelGGL(window_seam_kernel, dim3((unsigned)n_seam_wg), dim3(256), 0, st, a);
    GPURIR_HIP(hipGetLastError());
    return 0;
}

// The device tables of one segmentation (P, R, L, n_signal, w_ini): w_ini | p_lo | job_order, kept per device while calls repeat the segmentation, so that such a call plans
// nothing on the host and uploads nothing.  A different segmentation replaces them (the old
// buffer is freed, which waits for the work that still reads it).
struct WindowPlan {
    std::vector<int> key;
    void* d = nullptr;
    size_t bytes = 0;
    hipEvent_t uploaded = nullptr;  // recorded on up_stream after the upload
    hipStream_t up_stream = nullptr;
    int V = 0, s_max = 0, n_jobs = 0, n_conv_wg = 0, n_seam_wg = 0, ystride = 0;
    size_t off_j = 0, off_q = 0;
};
std::mutex g_win_mu;
std::map<int, WindowPlan> g_win_plans;

int make_window_plan(WindowPlan& W, int n_signal, int P, int R, int L, const int* w_ini, hipStream_t st) {
    int s_max = 0;
    for (int p = 0; p < P; p++) s_max = std::max(s_max, w_ini[p + 1] - w_ini[p]);
    const int V = 2 * s_max <= 8 * UT ? 8 : 2 * s_max <= 16 * UT ? 16 : 24;
    auto o = [&](long long m) -> long long {
        return m <= P ? (long long)w_ini[m] : (long long)n_signal + (m - P) * (long long)s_max;
    };
    const long long n_out = (long long)n_signal + L - 1;
    long long M = P;  // windows 0 .. M-1 receive products
    while (o(M) - o(P - 1) < L) M++;
    const long long n_jobs = (M + 1) / 2;
    GPURIR_CHECK(n_jobs * R < (1LL << 30) && (n_jobs + 1) * ((s_max + WSEAM_T - 1) / WSEAM_T) < (1LL << 30),
                 "trajectory too long");
    std::vector<int> p_lo((size_t)n_jobs);
    for (long long J = 0, p = 0; J < n_jobs; J++) {
        while (p < P && o(2 * J) - o(p) >= L) p++;
        p_lo[(size_t)J] = (int)p;
    }
    // Jobs in dispatch order.  Workgroup b runs on XCD b mod 8; each XCD takes a contiguous
    // range of the full-cost window pairs (the segment spectra of a range stay in its L2),
    // then its share of the cheaper pairs at the ends of the signal, cheapest last: when the
    // jobs do not divide evenly over the CUs, the workgroups dispatched last are the ones that
    // share a CU with two others, and they are then the shortest.
    std::vector<int> n_it((size_t)n_jobs);
    int it_max = 0;
    for (long long J = 0; J < n_jobs; J++) {
        const int hi = (int)std::min<long long>(P - 1, 2 * J + 1);
        n_it[(size_t)J] = hi >= p_lo[(size_t)J] ? (hi - p_lo[(size_t)J] + 2) / 2 : 0;
        it_max = std::max(it_max, n_it[(size_t)J]);
    }
    std::vector<int> full, edge;
    for (long long J = 0; J < n_jobs; J++) (n_it[(size_t)J] == it_max ? full : edge).push_back((int)J);
    std::stable_sort(edge.begin(), edge.end(), [&](int x, int y) { return n_it[(size_t)x] > n_it[(size_t)y]; });
    std::vector<std::vector<int>> per_xcd(8);
    const size_t chunk = (full.size() * R + 7) / 8;
    for (size_t k = 0; k < full.size() * R; k++) per_xcd[k / std::max<size_t>(chunk, 1)].push_back(full[k / R] * R + (int)(k % R));
    for (size_t k = 0; k < edge.size() * R; k++) {
        size_t x = 0;  // the XCD with the fewest jobs so far
        for (size_t y = 1; y < 8; y++)
            if (per_xcd[y].size() < per_xcd[x].size()) x = y;
        per_xcd[x].push_back(edge[k / R] * R + (int)(k % R));
    }
    size_t slots = 0;
    for (auto& v : per_xcd) slots = std::max(slots, v.size());
    std::vector<int> order(8 * slots, -1);
    for (size_t x = 0; x < 8; x++)
        for (size_t k = 0; k < per_xcd[x].size(); k++) order[8 * k + x] = per_xcd[x][k];
    const int n_conv_wg = (int)order.size();
    const size_t b_w = al(sizeof(int) * (P + 1)), b_j = al(sizeof(int) * n_jobs), b_q = al(sizeof(int) * n_conv_wg);
    const size_t bytes = b_w + b_j + b_q;
    if (W.bytes < bytes) {
        if (W.d) GPURIR_HIP(hipFree(W.d));  // waits for the work that still reads the old tables
        W.d = nullptr;
        W.bytes = 0;
        if (hipMalloc(&W.d, bytes) != hipSuccess) {
            (void)hipGetLastError();
            W.d = nullptr;
            W.key.clear();
            set_error("out of device memory for the trajectory tables");
            return -3;
        }
        W.bytes = bytes;
    }
    std::vector<unsigned char> host(bytes, 0);
    std::memcpy(host.data(), w_ini, sizeof(int) * (P + 1));
    std::memcpy(host.data() + b_w, p_lo.data(), sizeof(int) * n_jobs);
    std::memcpy(host.data() + b_w + b_j, order.data(), sizeof(int) * n_conv_wg);
    if (int rc = upload_async(W.d, host.data(), bytes, st)) {
        W.key.clear();
        return rc;
    }
    if (!W.uploaded) GPURIR_HIP(hipEventCreateWithFlags(&W.uploaded, hipEventDisableTiming));
    GPURIR_HIP(hipEventRecord(W.uploaded, st));
    W.up_stream = st;
    W.V = V;
    W.s_max = s_max;
    W.n_jobs = (int)n_jobs;
    W.n_conv_wg = n_conv_wg;
    W.n_seam_wg = (int)((n_jobs + 1) * ((s_max + WSEAM_T - 1) / WSEAM_T));
    W.ystride = (s_max + 63) / 64 * 64;
    W.off_j = b_w;
    W.off_q = b_w + b_j;
    return 0;
}

size_t window_workspace_bytes(const WindowPlan& W, int P, int R);
int window_args(const WindowPlan& W, const float* d_signal, int n_signal, const float* d_rirs, int P, int R, int L,
                double* d_out, void* ws, WinArgs& a);
int launch_window(WinArgs& a, const WindowPlan& W, hipStream_t st);

int window_trajectory(const float* d_signal, int n_signal, const float* d_rirs, int P, int R, int L,
                      const int* w_ini, double* d_out, hipStream_t st) {
    int dev = 0;
    GPURIR_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_win_mu);
    WindowPlan& W = g_win_plans[dev];
    const size_t nk = 4 + (size_t)P + 1;
    const bool hit = W.key.size() == nk && W.key[0] == P && W.key[1] == R && W.key[2] == L && W.key[3] == n_signal &&
                     std::memcmp(W.key.data() + 4, w_ini, sizeof(int) * (P + 1)) == 0;
    if (!hit) {
        W.key.clear();
        if (int rc = make_window_plan(W, n_signal, P, R, L, w_ini, st)) return rc;
        W.key.resize(nk);
        W.key[0] = P;
        W.key[1] = R;
        W.key[2] = L;
        W.key[3] = n_signal;
        std::memcpy(W.key.data() + 4, w_ini, sizeof(int) * (P + 1));
    } else if (W.up_stream != st) {
        GPURIR_HIP(hipStreamWaitEvent(st, W.uploaded, 0));
    }
    void* ws = nullptr;
    if (int rc = workspace(WS_CONV, window_workspace_bytes(W, P, R), &ws)) return rc;
    WinArgs a;
    if (int rc = window_args(W, d_signal, n_signal, d_rirs, P, R, L, d_out, ws, a)) return rc;
    return launch_window(a, W, st);
}

size_t window_workspace_bytes(const WindowPlan& W, int P, int R) {
    return al(sizeof(float2) * (size_t)P * W.V * UT) + 2 * al(sizeof(float) * (size_t)W.n_jobs * R * W.ystride);
}

int window_args(const WindowPlan& W, const float* d_signal, int n_signal, const float* d_rirs, int P, int R, int L,
                double* d_out, void* ws, WinArgs& a) {
    const int V = W.V;
    const float2* d_tw = nullptr;
    if (int rc = unit_tables_device(V, &d_tw)) return rc;
    const size_t b_spec = al(sizeof(float2) * (size_t)P * V * UT);
    const size_t b_y = al(sizeof(float) * (size_t)W.n_jobs * R * W.ystride);
    unsigned char* w = static_cast<unsigned char*>(ws);
    const unsigned char* t = static_cast<const unsigned char*>(W.d);
    a.signal = d_signal;
    a.w_ini = reinterpret_cast<const int*>(t);
    a.rirs = d_rirs;
    a.tw = d_tw;
    a.p_lo = reinterpret_cast<const int*>(t + W.off_j);
    a.job_order = reinterpret_cast<const int*>(t + W.off_q);
    a.spec = reinterpret_cast<float2*>(w);
    a.yhead = reinterpret_cast<float*>(w + b_spec);
    a.ytail = reinterpret_cast<float*>(w + b_spec + b_y);
    a.out = d_out;
    a.P = P;
    a.R = R;
    a.L = L;
    a.n_sig = n_signal;
    a.s_max = W.s_max;
    a.n_jobs = W.n_jobs;
    a.n_conv_wg = W.n_conv_wg;
    a.ystride = W.ystride;
    a.n_out = (long long)n_signal + L - 1;
    return 0;
}

int launch_window(WinArgs& a, const WindowPlan& W, hipStream_t st) {
    if (W.V == 8) return launch_window<8>(a, W.n_seam_wg, st);
    if (W.V == 16) return launch_window<16>(a, W.n_seam_wg, st);
    return launch_window<24>(a, W.n_seam_wg, st);
}

}  // namespace
}  // namespace gpurir

struct gpurir_traj_plan {
    const float* signal;
    const float* rirs;
    double* out;
    int n_signal, P, R, L;
    std::vector<int> w_ini;
    bool window = false;
    gpurir::WindowPlan tables;  // own tables and workspace: every run's kernels read them
    void* ws = nullptr;
    gpurir::WinArgs args;
};

using namespace gpurir;

extern "C" {

int gpurir_conv(const float* segments, int M_src, int seg_len, const float* rirs, int M_rcv, int rir_len, float* out) {
    GPURIR_CHECK(segments && rirs && out, "null pointer argument");
    GPURIR_CHECK(M_src > 0 && M_rcv > 0 && seg_len > 0 && rir_len > 0, "sizes must be positive");
    const hipStream_t st = nullptr;
    const size_t nseg = (size_t)M_src * seg_len, nrir = (size_t)M_src * M_rcv * rir_len;
    void* sv = nullptr;
    if (int rc = workspace(WS_STAGE, sizeof(float) * (nseg + nrir + 128), &sv)) return rc;
    float* d_seg = static_cast<float*>(sv);
    float* d_rir = d_seg + (nseg + 63) / 64 * 64;
    GPURIR_HIP(hipMemcpyAsync(d_seg, segments, sizeof(float) * nseg, hipMemcpyHostToDevice, st));
    GPURIR_HIP(hipMemcpyAsync(d_rir, rirs, sizeof(float) * nrir, hipMemcpyHostToDevice, st));
    std::vector<int> ss(M_src), sl(M_src, seg_len);
    for (int p = 0; p < M_src; p++) ss[p] = p * seg_len;
    float* conv = nullptr;
    int stride = 0;
    if (int rc = run_conv(d_seg, ss, sl, d_rir, M_src, M_rcv, rir_len, st, &conv, &stride, nullptr, 0)) return rc;
    const int len = seg_len + rir_len - 1;
    GPURIR_HIP(hipMemcpy2DAsync(out, sizeof(float) * len, conv, sizeof(float) * stride, sizeof(float) * len,
                                (size_t)M_src * M_rcv, hipMemcpyDeviceToHost, st));
    GPURIR_HIP(hipStreamSynchronize(st));
    return 0;
}

namespace {
int check_trajectory_args(const float* d_signal, int n_signal, const float* d_rirs, int n_pts, int n_rcv, int rir_len,
                          const int* w_ini, double* d_out) {
    GPURIR_CHECK(d_signal && d_rirs && w_ini && d_out, "null pointer argument");
    GPURIR_CHECK(n_signal > 0 && n_pts > 0 && n_rcv > 0 && rir_len > 0, "sizes must be positive");
    GPURIR_CHECK(w_ini[0] >= 0 && w_ini[n_pts] == n_signal, "w_ini must start at >= 0 and end at n_signal");
    for (int p = 0; p < n_pts; p++) GPURIR_CHECK(w_ini[p + 1] >= w_ini[p], "segment starts must be non-decreasing");
    return 0;
}

void free_traj_plan(gpurir_traj_plan* p) {
    if (p->tables.uploaded) (void)hipEventDestroy(p->tables.uploaded);
    if (p->tables.d) (void)hipFree(p->tables.d);
    if (p->ws) (void)hipFree(p->ws);
    delete p;
}
}  // namespace

int gpurir_trajectory_plan_create(const float* d_signal, int n_signal, const float* d_rirs, int n_pts, int n_rcv,
                                  int rir_len, const int* w_ini, double* d_out, gpurir_traj_plan_t* plan) {
    GPURIR_CHECK(plan != nullptr, "null plan pointer");
    *plan = nullptr;
    if (int rc = check_trajectory_args(d_signal, n_signal, d_rirs, n_pts, n_rcv, rir_len, w_ini, d_out)) return rc;
    gpurir_traj_plan* p = new gpurir_traj_plan();
    p->signal = d_signal;
    p->rirs = d_rirs;
    p->out = d_out;
    p->n_signal = n_signal;
    p->P = n_pts;
    p->R = n_rcv;
    p->L = rir_len;
    p->w_ini.assign(w_ini, w_ini + n_pts + 1);
    p->window = window_path_applies(w_ini, n_pts);
    if (p->window) {
        if (int rc = make_window_plan(p->tables, n_signal, n_pts, n_rcv, rir_len, w_ini, nullptr)) {
            free_traj_plan(p);
            return rc;
        }
        if (hipMalloc(&p->ws, window_workspace_bytes(p->tables, n_pts, n_rcv)) != hipSuccess) {
            (void)hipGetLastError();
            p->ws = nullptr;
            free_traj_plan(p);
            set_error("out of device memory for the trajectory plan");
            return -3;
        }
        if (int rc = window_args(p->tables, d_signal, n_signal, d_rirs, n_pts, n_rcv, rir_len, d_out, p->ws, p->args)) {
            free_traj_plan(p);
            return rc;
        }
        if (hipStreamSynchronize(nullptr) != hipSuccess) {  // the tables are on the device
            free_traj_plan(p);
            GPURIR_HIP(hipGetLastError());
        }
    }
    *plan = p;
    return 0;
}

int gpurir_trajectory_plan_run(gpurir_traj_plan_t p, void* stream) {
    GPURIR_CHECK(p != nullptr, "null plan");
    const hipStream_t st = static_cast<hipStream_t>(stream);
    if (!p->window)
        return gpurir_trajectory_device(p->signal, p->n_signal, p->rirs, p->P, p->R, p->L, p->w_ini.data(), p->out,
                                        stream);
    // three direct launches: a graph launch of the same three kernels cost 0.05-0.2 ms of host
    // time per replay, more than the kernels themselves (48 us on 1024 RIRs of 8000 samples)
    return launch_window(p->args, p->tables, st);
}

int gpurir_trajectory_plan_destroy(gpurir_traj_plan_t p) {
    if (!p) return 0;
    (void)hipDeviceSynchronize();  // no replay may still read the plan's buffers
    free_traj_plan(p);
    return 0;
}

int gpurir_trajectory_device(const float* d_signal, int n_signal, const float* d_rirs, int n_pts, int n_rcv,
                             int rir_len, const int* w_ini, double* d_out, void* stream) {
    if (int rc = check_trajectory_args(d_signal, n_signal, d_rirs, n_pts, n_rcv, rir_len, w_ini, d_out)) return rc;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    if (window_path_applies(w_ini, n_pts))
        return window_trajectory(d_signal, n_signal, d_rirs, n_pts, n_rcv, rir_len, w_ini, d_out, st);
    std::vector<int> ss(n_pts), sl(n_pts);
    for (int p = 0; p < n_pts; p++) {
        ss[p] = w_ini[p];
        sl[p] = w_ini[p + 1] - w_ini[p];
    }
    float* conv = nullptr;
    int stride = 0;
    void* extra = nullptr;
    const long long n_out = (long long)n_signal + rir_len - 1;
    GPURIR_CHECK(n_out / OLA_T < (1LL << 28), "trajectory output too long");
    const int n_wg = (int)((n_out + OLA_T - 1) / OLA_T);
    // per workgroup: (first segment ending after t0, last starting at or before t0, last
    // starting at or before t1), by three monotone sweeps
    const size_t b_w = al(sizeof(int) * (n_pts + 1));
    std::vector<unsigned char> host(b_w + sizeof(int4) * (size_t)n_wg, 0);
    std::memcpy(host.data(), w_ini, sizeof(int) * (n_pts + 1));
    int4* rg = reinterpret_cast<int4*>(host.data() + b_w);
    {
        int pa = 0, pb = -1, pc = -1;
        for (int b = 0; b < n_wg; b++) {
            const long long t0 = (long long)b * OLA_T, t1 = std::min(t0 + OLA_T - 1, n_out - 1);
            while (pa < n_pts && (long long)w_ini[pa + 1] + rir_len - 1 <= t0) pa++;
            while (pb + 1 < n_pts && (long long)w_ini[pb + 1] <= t0) pb++;
            while (pc + 1 < n_pts && (long long)w_ini[pc + 1] <= t1) pc++;
            rg[b] = make_int4(pa, pb, pc, 0);
        }
    }
    if (int rc = run_conv(d_signal, ss, sl, d_rirs, n_pts, n_rcv, rir_len, st, &conv, &stride, &extra, host.size()))
        return rc;
    if (int rc = upload_async(extra, host.data(), host.size(), st)) return rc;
    unsigned char* ex = static_cast<unsigned char*>(extra);
    hipLaunchKernelGGL(ola_kernel, dim3((unsigned)n_wg), dim3(256), 0, st, conv, reinterpret_cast<const int*>(ex),
                       reinterpret_cast<const int4*>(ex + b_w), n_pts, n_rcv, rir_len, stride, n_out, d_out);
    GPURIR_HIP(hipGetLastError());
    return 0;
}

}  // extern "C"
