/*
 * rir_oracle.c -- FP64 CPU restatement of gpuRIR's simulation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP kernels of
 * gpurir_amd are compared against, and the CPU baseline that bench.py times.
 * Nothing in the gpurir_amd package may link, load or call it.
 *
 * Every function cites the reference code (DavidDiazGuerra/gpuRIR v1.2.0) it
 * restates.  All arithmetic is IEEE double unless a comment says otherwise;
 * the file is compiled with -ffp-contract=off so that the delay of every image
 * is computed with exactly the operation sequence the GPU kernel uses (the
 * image vector is exact in FP64 for float inputs, then three squares, two
 * adds, a correctly rounded sqrt and one multiply by fs/c).
 *
 * Modes of the early (image-source) part:
 *   ORC_EXACT_GATHER  literal transcription of generateRIR_kernel: every
 *                     sample t is a sum over ALL images (O(M*T*N), tiny inputs)
 *   ORC_EXACT_SCATTER windowed: every image touches only |t-tau| < Tw/2
 *   ORC_LUT           the lookup-table variant (1/16-sample table + lerp)
 *   ORC_FP32_REF      the reference's float32 arithmetic (gather order), used
 *                     to state the reference's own error against FP64
 *
 * The diffuse tail uses Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) as
 * its noise source instead of the reference's cuRAND default generator (a
 * closed library whose stream cannot be restated); the envelope follows the
 * reference exactly.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_EXACT_GATHER 0
#define ORC_EXACT_SCATTER 1
#define ORC_LUT 2
#define ORC_FP32_REF 3

/* Polar pattern codes: gpuRIR_cuda.h:3-9, gpuRIR/__init__.py:13-20 */
#define DIR_OMNI 0
#define DIR_HOMNI 1
#define DIR_CARD 2
#define DIR_HYPCARD 3
#define DIR_SUBCARD 4
#define DIR_BIDIR 5

static const double ORC_PI = 3.14159265358979323846;

/* ------------------------------------------------------------------------ */
/* Sample-count rules: gpuRIR_cuda.cu:831-835 and python_bind.cpp:82-83.  */
/* Tmax and Fs arrive as C floats; their product is a float.                 */
/* ------------------------------------------------------------------------ */
int orc_n_samples(float T, float fs) {
    float p = T * fs;               /* float product, as in the reference */
    int n = (int)ceilf(p);
    n += n % 2;                     /* forced even */
    return n;
}

/* Window length: gpuRIR_cuda.cu:629  Tw = (int) round(8e-3f * Fs) */
int orc_tw(float fs) {
    float p = 8e-3f * fs;
    return (int)roundf(p);
}

/* ------------------------------------------------------------------------ */
/* Directivity: gpuRIR_cuda.cu:132-147                                       */
/* ------------------------------------------------------------------------ */
static double orc_directivity(const double doa[3], const double orv[3], int pattern) {
    if (pattern == DIR_OMNI) return 1.0;
    double c = doa[0] * orv[0] + doa[1] * orv[1] + doa[2] * orv[2];
    c /= sqrt(doa[0] * doa[0] + doa[1] * doa[1] + doa[2] * doa[2]);
    c /= sqrt(orv[0] * orv[0] + orv[1] * orv[1] + orv[2] * orv[2]);
    switch (pattern) {
        case DIR_HOMNI: return (c > 0.0) ? 1.0 : 0.0;
        case DIR_CARD: return 0.5 + 0.5 * c;
        case DIR_HYPCARD: return 0.25 + 0.75 * c;
        case DIR_SUBCARD: return 0.75 + 0.25 * c;
        case DIR_BIDIR: return c;
        default: return 0.0; /* reference prints "Invalid polar pattern." and returns 0 */
    }
}

/* ------------------------------------------------------------------------ */
/* Image index map for one axis: gpuRIR_cuda.cu:336-341                     */
/*   k = n - N/2 (integer division), m = ceil(k/2), q = |k % 2|             */
/* ------------------------------------------------------------------------ */
static void orc_axis_index(int n, int N, int* m, int* q) {
    int k = n - N / 2;
    *m = (int)ceil(k / 2.0);
    *q = abs(k % 2);
}

typedef struct {
    double amp;   /* signed amplitude (attenuation, 1/(4 pi d), directivities) */
    double tau;   /* delay in samples */
} orc_image_t;

/*
 * Amplitude and delay of one image for one (source, receiver) pair.
 * gpuRIR_cuda.cu:271-277 (negated betas), :336-342 (index map, attenuation),
 * :351-366 (image vector, distance, directivities, amp, tau).
 * The delay is dist * (fs / c); the reference writes dist / c * Fs in float.
 */
static orc_image_t orc_image(const double L[3], const double beta[6], const double src[3],
                             const double rcv[3], const double osrc[3], const double orcv[3],
                             int spkr, int mic, const int nb[3], int nx, int ny, int nz,
                             double fs_over_c) {
    int n[3] = {nx, ny, nz};
    double att = 1.0, v[3], osrc_img[3], im2rcv[3];
    for (int d = 0; d < 3; d++) {
        int m, q;
        orc_axis_index(n[d], nb[d], &m, &q);
        att *= pow(-beta[2 * d], (double)abs(m - q)) * pow(-beta[2 * d + 1], (double)abs(m));
        double sgn = (double)(1 - 2 * q);
        double a = 2.0 * (double)m * L[d];
        double b = sgn * src[d];
        v[d] = (a + b) - rcv[d];
        osrc_img[d] = sgn * osrc[d];
        im2rcv[d] = -v[d];
    }
    double d2 = v[0] * v[0];
    d2 = d2 + v[1] * v[1];
    d2 = d2 + v[2] * v[2];
    double dist = sqrt(d2);
    orc_image_t r;
    r.amp = att / (4.0 * ORC_PI * dist);
    if (spkr != DIR_OMNI) r.amp *= orc_directivity(im2rcv, osrc_img, spkr);
    if (mic != DIR_OMNI) r.amp *= orc_directivity(v, orcv, mic);
    r.tau = dist * fs_over_c;
    return r;
}

/* Windowed sinc tap, gpuRIR_cuda.cu:110-121: Hann(x) * sinc(pi x) for |x| < Tw/2 */
static double orc_tap(double x, double Tw) {
    if (!(fabs(x) < Tw / 2.0)) return 0.0;
    double w = 0.5 * (1.0 + cos(2.0 * ORC_PI * x / Tw));
    double s = (x == 0.0) ? 1.0 : sin(ORC_PI * x) / (ORC_PI * x);
    return w * s;
}

/* ------------------------------------------------------------------------ */
/* Sinc lookup table: gpuRIR_cuda.cu:578-590 (table), :26 (oversampling 16),*/
/* :632-637 (odd length, center, support |t-tau| < Tw/2 integer),           */
/* :236-239 (lookup at (t-tau)*16 + center with linear interpolation).      */
/* ------------------------------------------------------------------------ */
typedef struct {
    int len, center, half_tw_int;
    double* tab;
} orc_lut_t;

static orc_lut_t orc_lut_make(int Tw) {
    orc_lut_t L;
    L.len = Tw * 16;
    L.len += (L.len % 2) ? 0 : 1;
    L.center = L.len / 2;
    L.half_tw_int = Tw / 2;
    L.tab = (double*)calloc((size_t)L.len + 1, sizeof(double));
    for (int i = 0; i <= L.center; i++) {
        double x = (double)i / 16.0;
        double s = (x == 0.0) ? 1.0 : sin(ORC_PI * x) / (ORC_PI * x);
        double h = 0.5 * (1.0 + cos(2.0 * ORC_PI * x / Tw));
        L.tab[L.center + i] = h * s;
        L.tab[L.center - i] = h * s;
    }
    return L;
}

static double orc_lut_tap(const orc_lut_t* L, double x) {
    if (!(fabs(x) < (double)L->half_tw_int)) return 0.0;
    double u = x * 16.0 + (double)L->center;
    double fl = floor(u);
    int i = (int)fl;
    double a = u - fl;
    double t0 = (i >= 0 && i < L->len) ? L->tab[i] : 0.0;
    double t1 = (i + 1 >= 0 && i + 1 < L->len) ? L->tab[i + 1] : 0.0;
    return t0 + a * (t1 - t0);
}

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon et al., SC'11).                                    */
/* ------------------------------------------------------------------------ */
void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Noise value of tail sample s of one RIR.
 * counter = {s/4, rir index in room, room index, call offset}, key = seed.
 * kind 0: unit-variance logistic from u in (0,1]  (gpuRIR_cuda.cu:454-455)
 * kind 1: standard Gaussian by Box-Muller on two words of the block. */
double orc_noise(uint64_t seed, uint32_t rir, uint32_t room, uint32_t offset, uint32_t s, int kind) {
    uint32_t ctr[4] = {s >> 2, rir, room, offset};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t w[4];
    orc_philox4x32_10(ctr, key, w);
    uint32_t x = w[s & 3u];
    if (kind == 0) {
        double u = ((double)x + 1.0) / 4294967296.0;          /* (0, 1]  */
        double one_m_u = (double)(0xFFFFFFFFu - x) / 4294967296.0; /* 1 - u exactly */
        return 0.551329 * log(u / (one_m_u + 1e-6));
    } else {
        uint32_t pr = (s & 3u) >> 1;
        double ua = ((double)w[2 * pr] + 1.0) / 4294967296.0;   /* (0, 1] */
        double ub = (double)w[2 * pr + 1] / 4294967296.0;       /* [0, 1) */
        double rad = sqrt(-2.0 * log(ua));
        return (s & 1u) ? rad * sin(2.0 * ORC_PI * ub) : rad * cos(2.0 * ORC_PI * ub);
    }
}

/* ------------------------------------------------------------------------ */
/* Envelope of the diffuse tail: gpuRIR_cuda.cu:415-443 (envPred_kernel)     */
/* and SabineT60 :123-130.  Literal constants are the reference's.           */
/* ------------------------------------------------------------------------ */
typedef struct {
    double A, alpha, tau_dp, w_center;
    int w_start;
} orc_env_t;

static double orc_sabine_t60(const double L[3], const double b[6]) {
    double Sa = ((1.0 - b[0] * b[0]) + (1.0 - b[1] * b[1])) * L[1] * L[2] +
                ((1.0 - b[2] * b[2]) + (1.0 - b[3] * b[3])) * L[0] * L[2] +
                ((1.0 - b[4] * b[4]) + (1.0 - b[5] * b[5])) * L[0] * L[1];
    double V = L[0] * L[1] * L[2];
    return 0.161 * V / Sa;
}

static orc_env_t orc_envelope(const double* rir_ism, int nISM, double tau_dp, float fs_f,
                              const double L[3], const double beta[6]) {
    orc_env_t e;
    double w_sz = (double)(10e-3f * fs_f);      /* float product, :421 */
    double ws = fmax((double)nISM - w_sz, tau_dp);
    e.w_start = (int)ceil(ws);
    e.w_center = (double)e.w_start + (double)(nISM - e.w_start) / 2.0;
    double P = 0.0;
    for (int t = e.w_start; t < nISM; t++) P += rir_ism[t] * rir_ism[t];
    P /= (double)(nISM - e.w_start);
    double T60 = orc_sabine_t60(L, beta);
    e.alpha = -13.8155 / (T60 * (double)fs_f);
    e.A = P / exp(e.alpha * (e.w_center - tau_dp));
    e.tau_dp = tau_dp;
    return e;
}

/* ------------------------------------------------------------------------ */
/* Float32 restatement of the reference arithmetic (gather order), used only */
/* to state the reference's own float error.  gpuRIR_cuda.cu:245-390.       */
/* ------------------------------------------------------------------------ */
static void orc_fp32_ref_rir(const float Lf[3], const float bf[6], const float srcf[3],
                             const float rcvf[3], const float osf[3], const float orf[3],
                             int spkr, int mic, const int nb[3], float fs, float c,
                             int nISM, double* out, float* tau_dp_out) {
    const float PIf = 3.141592654f;
    int N = nb[0] * nb[1] * nb[2];
    float* amp = (float*)malloc(sizeof(float) * (size_t)N);
    float* tau = (float*)malloc(sizeof(float) * (size_t)N);
    float beta[6];
    for (int i = 0; i < 6; i++) beta[i] = -bf[i];
    for (int nx = 0; nx < nb[0]; nx++)
        for (int ny = 0; ny < nb[1]; ny++)
            for (int nz = 0; nz < nb[2]; nz++) {
                int n[3] = {nx, ny, nz};
                float att = 1.0f, cp[3];
                int ci[3], ri[3], dp = 1;
                for (int d = 0; d < 3; d++) {
                    ci[d] = (int)ceilf((n[d] - nb[d] / 2) / 2.0f);
                    cp[d] = ci[d] * 2 * Lf[d];
                    ri[d] = abs((n[d] - nb[d] / 2) % 2);
                    att *= powf(beta[d * 2], (float)abs(ci[d] - ri[d])) * powf(beta[d * 2 + 1], (float)abs(ci[d]));
                    dp = dp && (ci[d] == 0) && (ri[d] == 0);
                }
                float v[3], im[3], os[3], dist = 0.0f;
                for (int d = 0; d < 3; d++) {
                    v[d] = cp[d] + (1 - 2 * ri[d]) * srcf[d] - rcvf[d];
                    dist += v[d] * v[d];
                    os[d] = (1 - 2 * ri[d]) * osf[d];
                    im[d] = -v[d];
                }
                dist = sqrtf(dist);
                float a = att / (4 * PIf * dist);
                if (spkr != DIR_OMNI) {
                    double dd[3] = {im[0], im[1], im[2]}, oo[3] = {os[0], os[1], os[2]};
                    a *= (float)orc_directivity(dd, oo, spkr);
                }
                if (mic != DIR_OMNI) {
                    double dd[3] = {v[0], v[1], v[2]}, oo[3] = {orf[0], orf[1], orf[2]};
                    a *= (float)orc_directivity(dd, oo, mic);
                }
                int idx = nx * nb[1] * nb[2] + ny * nb[2] + nz;
                amp[idx] = a;
                tau[idx] = dist / c * fs;
                if (dp) *tau_dp_out = tau[idx];
            }
    float Tw = (float)orc_tw(fs);
    for (int t = 0; t < nISM; t++) {
        float acc = 0.0f;
        for (int i = 0; i < N; i++) {
            float x = (float)t - tau[i];
            if (fabsf(x) < Tw / 2) {
                float w = 0.5f * (1.0f + cosf(2.0f * PIf * x / Tw));
                float xs = x * PIf;
                float s = (xs == 0.0f) ? 1.0f : sinf(xs) / xs;
                acc += w * amp[i] * s;
            }
        }
        out[t] = acc;
    }
    free(amp);
    free(tau);
}

/* ------------------------------------------------------------------------ */
/* Per-image listing (stage a), for "identical image indices and offsets".   */
/* Writes for RIR m = s*M_rcv + r and image i = (nx*Ny + ny)*Nz + nz:        */
/*   n0[m*N + i] = floor(tau), frac[m*N + i] = tau - n0, amp[m*N + i].      */
/* ------------------------------------------------------------------------ */
int orc_images(const float room[3], const float beta_f[6], const float* pos_src, int M_src,
               const float* pos_rcv, int M_rcv, const float* orv_src, const float* orv_rcv,
               int spkr, int mic, const int nb[3], float fs, float c,
               int64_t* n0, double* frac, double* amp, double* tau_dp) {
    double L[3] = {room[0], room[1], room[2]};
    double beta[6];
    for (int i = 0; i < 6; i++) beta[i] = beta_f[i];
    double fs_over_c = (double)fs / (double)c;
    long N = (long)nb[0] * nb[1] * nb[2];
#pragma omp parallel for collapse(2) schedule(dynamic)
    for (int s = 0; s < M_src; s++)
        for (int r = 0; r < M_rcv; r++) {
            double src[3], rcv[3], os[3], orr[3];
            for (int d = 0; d < 3; d++) {
                src[d] = pos_src[s * 3 + d];
                rcv[d] = pos_rcv[r * 3 + d];
                os[d] = orv_src[s * 3 + d];
                orr[d] = orv_rcv[r * 3 + d];
            }
            long m = (long)s * M_rcv + r;
            for (int nx = 0; nx < nb[0]; nx++)
                for (int ny = 0; ny < nb[1]; ny++)
                    for (int nz = 0; nz < nb[2]; nz++) {
                        orc_image_t im = orc_image(L, beta, src, rcv, os, orr, spkr, mic, nb, nx, ny, nz, fs_over_c);
                        long i = ((long)nx * nb[1] + ny) * nb[2] + nz;
                        double fl = floor(im.tau);
                        n0[m * N + i] = (int64_t)fl;
                        frac[m * N + i] = im.tau - fl;
                        amp[m * N + i] = im.amp;
                    }
            if (tau_dp) {
                orc_image_t dp = orc_image(L, beta, src, rcv, os, orr, spkr, mic, nb, nb[0] / 2, nb[1] / 2, nb[2] / 2, fs_over_c);
                tau_dp[m] = dp.tau;
            }
        }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Whole simulateRIR: gpuRIR_cuda.cu:766-931.                               */
/* out is [M_src][M_rcv][nSamples] doubles: ISM part then the tail.          */
/* rir_base: index of the first RIR of this room in the noise counter        */
/* (0 for a single-room call).  level multiplies the tail (default 1).       */
/* ------------------------------------------------------------------------ */
/* Core with double-precision geometry (used directly by the gradient checks, which
 * perturb positions and betas by less than a float ulp).  room/beta/positions/orientations
 * are doubles; fs, c, Tdiff and Tmax keep the reference's float semantics. */
int orc_simulate_rir_d(const double room[3], const double beta_d[6], const double* pos_src, int M_src,
                       const double* pos_rcv, int M_rcv, const double* orv_src, const double* orv_rcv,
                       int spkr, int mic, const int nb[3], float Tdiff, float Tmax, float fs, float c,
                       int mode, int noise_kind, uint64_t seed, uint32_t room_index, uint32_t call_offset,
                       double level, int crossfade, double* out, double* env_out,
                       const float* f32_room, const float* f32_beta, const float* f32_src, const float* f32_rcv,
                       const float* f32_osrc, const float* f32_orcv) {
    int nISM = orc_n_samples(Tdiff, fs);
    int nSamples = orc_n_samples(Tmax, fs);
    int nDiff = nSamples - nISM;
    int Tw = orc_tw(fs);
    double L[3] = {room[0], room[1], room[2]};
    double beta[6];
    for (int i = 0; i < 6; i++) beta[i] = beta_d[i];
    double fs_over_c = (double)fs / (double)c;
    long N = (long)nb[0] * nb[1] * nb[2];
    int M = M_src * M_rcv;
    orc_lut_t lut = {0, 0, 0, NULL};
    if (mode == ORC_LUT) lut = orc_lut_make(Tw);

    /* image chunks are fixed (independent of the thread count) so that the
       partial sums are always added in the same order */
    const long CH = 4096;
    long nch = (N + CH - 1) / CH;
    double* part = (double*)calloc((size_t)M * (size_t)(nch > 0 ? nch : 1) * (size_t)(nISM > 0 ? nISM : 1), sizeof(double));

    if (mode == ORC_FP32_REF) {
#pragma omp parallel for schedule(dynamic)
        for (int m = 0; m < M; m++) {
            int s = m / M_rcv, r = m % M_rcv;
            float tdp = 0.0f;
            orc_fp32_ref_rir(f32_room, f32_beta, f32_src + 3 * s, f32_rcv + 3 * r, f32_osrc + 3 * s,
                             f32_orcv + 3 * r, spkr, mic, nb, fs, c, nISM, part + (size_t)m * nch * nISM, &tdp);
            (void)tdp;
            for (long k = 1; k < nch; k++) memset(part + ((size_t)m * nch + k) * nISM, 0, sizeof(double) * nISM);
        }
    } else {
#pragma omp parallel for collapse(2) schedule(dynamic)
        for (int m = 0; m < M; m++)
            for (long k = 0; k < nch; k++) {
                int s = m / M_rcv, r = m % M_rcv;
                double src[3], rcv[3], os[3], orr[3];
                for (int d = 0; d < 3; d++) {
                    src[d] = pos_src[s * 3 + d];
                    rcv[d] = pos_rcv[r * 3 + d];
                    os[d] = orv_src[s * 3 + d];
                    orr[d] = orv_rcv[r * 3 + d];
                }
                double* acc = part + ((size_t)m * nch + k) * nISM;
                long i1 = (k + 1) * CH < N ? (k + 1) * CH : N;
                if (mode == ORC_EXACT_GATHER) {
                    /* literal: every t sums every image (gpuRIR_cuda.cu:377-390) */
                    orc_image_t* ims = (orc_image_t*)malloc(sizeof(orc_image_t) * CH);
                    for (long i = k * CH; i < i1; i++) {
                        int nz = (int)(i % nb[2]), ny = (int)((i / nb[2]) % nb[1]), nx = (int)(i / ((long)nb[1] * nb[2]));
                        ims[i - k * CH] = orc_image(L, beta, src, rcv, os, orr, spkr, mic, nb, nx, ny, nz, fs_over_c);
                    }
                    for (int t = 0; t < nISM; t++) {
                        double a = 0.0;
                        for (long i = k * CH; i < i1; i++) a += ims[i - k * CH].amp * orc_tap((double)t - ims[i - k * CH].tau, (double)Tw);
                        acc[t] = a;
                    }
                    free(ims);
                } else {
                    double half = (mode == ORC_LUT) ? (double)lut.half_tw_int : (double)Tw / 2.0;
                    for (long i = k * CH; i < i1; i++) {
                        int nz = (int)(i % nb[2]), ny = (int)((i / nb[2]) % nb[1]), nx = (int)(i / ((long)nb[1] * nb[2]));
                        orc_image_t im = orc_image(L, beta, src, rcv, os, orr, spkr, mic, nb, nx, ny, nz, fs_over_c);
                        long t0 = (long)ceil(im.tau - half);
                        long t1 = (long)floor(im.tau + half);
                        if (t0 < 0) t0 = 0;
                        if (t1 > nISM - 1) t1 = nISM - 1;
                        for (long t = t0; t <= t1; t++) {
                            double x = (double)t - im.tau;
                            double h = (mode == ORC_LUT) ? orc_lut_tap(&lut, x) : orc_tap(x, (double)Tw);
                            acc[t] += im.amp * h;
                        }
                    }
                }
            }
    }

#pragma omp parallel for schedule(static)
    for (int m = 0; m < M; m++) {
        int s = m / M_rcv, r = m % M_rcv;
        double* o = out + (size_t)m * nSamples;
        for (int t = 0; t < nISM; t++) {
            double a = 0.0;
            for (long k = 0; k < nch; k++) a += part[((size_t)m * nch + k) * nISM + t];
            o[t] = a;
        }
        double src[3], rcv[3], os[3] = {0, 0, 0}, orr[3] = {0, 0, 0};
        for (int d = 0; d < 3; d++) {
            src[d] = pos_src[s * 3 + d];
            rcv[d] = pos_rcv[r * 3 + d];
        }
        /* direct path: image with k = 0 on every axis (gpuRIR_cuda.cu:341,368-369) */
        orc_image_t dp = orc_image(L, beta, src, rcv, os, orr, DIR_OMNI, DIR_OMNI, nb, nb[0] / 2, nb[1] / 2, nb[2] / 2, fs_over_c);
        orc_env_t e = orc_envelope(o, nISM, dp.tau, fs, L, beta);
        if (env_out) {
            env_out[3 * m + 0] = e.A;
            env_out[3 * m + 1] = e.alpha;
            env_out[3 * m + 2] = e.tau_dp;
        }
        /* tail: gpuRIR_cuda.cu:445-461 */
        for (int sd = 0; sd < nDiff; sd++) {
            double env = e.A * exp(e.alpha * ((double)nISM + (double)sd - e.tau_dp));
            double nz = orc_noise(seed, (uint32_t)m, room_index, call_offset, (uint32_t)sd, noise_kind);
            o[nISM + sd] = level * sqrt(env) * nz;
        }
        /* optional cross-fade (no reference counterpart; 0 = the reference's hard switch):
           on t in [nISM - X, nISM) the early part fades out linearly, r = (k + 1/2) / X with
           k = t - (nISM - X), and the diffuse model, continued before nISM with noise
           indices nDiff + k, fades in.  The envelope is fitted before the fade. */
        int X = crossfade < nISM ? crossfade : nISM;
        if (nDiff > 0 && X > 0) {
            for (int k = 0; k < X; k++) {
                int t = nISM - X + k;
                double r = ((double)k + 0.5) / (double)X;
                double env = e.A * exp(e.alpha * ((double)t - e.tau_dp));
                double nz = orc_noise(seed, (uint32_t)m, room_index, call_offset, (uint32_t)(nDiff + k), noise_kind);
                o[t] = (1.0 - r) * o[t] + r * level * sqrt(env) * nz;
            }
        }
    }
    free(part);
    free(lut.tab);
    return 0;
}

int orc_simulate_rir(const float room[3], const float beta_f[6], const float* pos_src, int M_src,
                     const float* pos_rcv, int M_rcv, const float* orv_src, const float* orv_rcv,
                     int spkr, int mic, const int nb[3], float Tdiff, float Tmax, float fs, float c,
                     int mode, int noise_kind, uint64_t seed, uint32_t room_index, uint32_t call_offset,
                     double level, int crossfade, double* out,
                     double* env_out /* [M][3]: A, alpha, tau_dp or NULL */) {
    double L[3], b[6];
    for (int i = 0; i < 3; i++) L[i] = room[i];
    for (int i = 0; i < 6; i++) b[i] = beta_f[i];
    double* ps = (double*)malloc(sizeof(double) * 3 * (size_t)M_src);
    double* pr = (double*)malloc(sizeof(double) * 3 * (size_t)M_rcv);
    double* os = (double*)malloc(sizeof(double) * 3 * (size_t)M_src);
    double* orr = (double*)malloc(sizeof(double) * 3 * (size_t)M_rcv);
    for (int i = 0; i < 3 * M_src; i++) { ps[i] = pos_src[i]; os[i] = orv_src[i]; }
    for (int i = 0; i < 3 * M_rcv; i++) { pr[i] = pos_rcv[i]; orr[i] = orv_rcv[i]; }
    int rc = orc_simulate_rir_d(L, b, ps, M_src, pr, M_rcv, os, orr, spkr, mic, nb, Tdiff, Tmax, fs, c, mode,
                                noise_kind, seed, room_index, call_offset, level, crossfade, out, env_out, room, beta_f,
                                pos_src, pos_rcv, orv_src, orv_rcv);
    free(ps); free(pr); free(os); free(orr);
    return rc;
}

/* Table used by the LUT mode (for the tests of the table formula). */
int orc_lut_table(int Tw, double* tab, int cap) {
    orc_lut_t L = orc_lut_make(Tw);
    int n = L.len < cap ? L.len : cap;
    memcpy(tab, L.tab, sizeof(double) * (size_t)n);
    int len = L.len;
    free(L.tab);
    return len;
}

double orc_tap_exact(double x, int Tw) { return orc_tap(x, (double)Tw); }

int orc_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
