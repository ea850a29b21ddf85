// Internal declarations shared by the HIP translation units of libgpurir.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace gpurir {

// ---- error handling: every entry point returns 0 or a negative code and
//      leaves a message retrievable through gpurir_last_error().
void set_error(const std::string& msg);
const char* last_error();

#define GPURIR_HIP(call)                                                                        \
    do {                                                                                        \
        hipError_t _e = (call);                                                                 \
        if (_e != hipSuccess) {                                                                 \
            ::gpurir::set_error(std::string("HIP error: ") + hipGetErrorString(_e) + " at " +  \
                                __FILE__ + ":" + std::to_string(__LINE__) + " (" #call ")");    \
            return -2;                                                                          \
        }                                                                                       \
    } while (0)

#define GPURIR_CHECK(cond, msg)          \
    do {                                 \
        if (!(cond)) {                   \
            ::gpurir::set_error(msg);    \
            return -1;                   \
        }                                \
    } while (0)

// ---- device workspace: one grow-only buffer per (device, slot).
//      Slots keep concurrent users (e.g. an ISM call's partials and its
//      staging copies) apart.
enum WsSlot { WS_PLAN = 0, WS_PARTIAL, WS_STAGE, WS_CONV, WS_GRAD, WS_ADJ_PLAN, WS_COUNT };
int workspace(int slot, size_t bytes, void** ptr);

// Work counters of the persistent accumulation for eager (non-plan) launches: two
// unsigned per device, zero between launches (every launch leaves them zero).
int device_counters(unsigned** counters);

// Host-to-device copy through a pinned staging buffer; the host source may be
// reused as soon as the call returns.  Must not be called during stream capture.
int upload_async(void* dst, const void* src, size_t bytes, hipStream_t stream);

// ---- rounding rules of the reference (src/gpuRIR_cuda.cu:629, :831-835)
int n_samples_rule(float T, float fs);
int tw_rule(float fs);

// Per-room constants, laid out for the device.  One per room of a launch.
struct RoomDev {
    double L[3];
    double fs_over_c;
    double inv_nyz, inv_nz;  // 1/(Ny*Nz), 1/Nz for the image index decode
    float beta[6];
    int nb[3];
    int n_src, n_rcv;
    int src_off, rcv_off;    // first source / receiver of the room in the position arrays
    int nISM, nSamples;
    int n_chunks;            // image chunks per RIR (1: the accumulation writes the output)
    int rir_off;             // global id of the room's first RIR
    long long out_off;       // element offset of the room's [src][rcv][t] block in out
    long long part_off;      // element offset of the room's partial rows
    unsigned room_id;        // noise counter word
    float level;             // diffuse-tail level multiplier
};

// One workgroup of the accumulation kernel.
struct Item {
    int room, rir;           // rir: index inside the room (src * n_rcv + rcv)
    int i0, i1;              // linear image range [i0, i1)
    int t_lo, t_hi;          // time tile [t_lo, t_hi) of the ISM part
    int chunk, pad;
};

// Host-side description of one room of a launch.
struct RoomHost {
    float L[3];
    float beta[6];
    int nb[3];
    int n_src, n_rcv;
    int src_off, rcv_off;
    float Tdiff, Tmax;
    long long out_off;
    unsigned room_id;
    float level;
};

// precision modes (keep in sync with include/gpurir.h)
enum { PREC_FP32 = 0, PREC_LUT = 1, PREC_FP16 = 2, PREC_BF16 = 3 };
enum { METHOD_AUTO = 0, METHOD_TAPS = 1, METHOD_BASIS = 2 };

struct SimOptions {
    int precision;
    int noise;               // 0 logistic, 1 gaussian
    unsigned long long seed;
    unsigned call_offset;
    int spkr, mic;
    float fs, c;
    int chunk_images;        // 0: planner default
    int tile_samples;        // 0: planner default
    int crossfade;           // samples of cross-fade before nISM (0: hard switch)
    int method;              // FP32: METHOD_TAPS or METHOD_BASIS (0 = basis)
};

// Plans and launches a batch of rooms whose positions and output are on the device.
struct SimLaunch;
SimLaunch* sim_new();
void sim_delete(SimLaunch* s);
int sim_prepare(const std::vector<RoomHost>& rooms, const SimOptions& opt, SimLaunch& S);
size_t sim_plan_bytes(const SimLaunch& S);
size_t sim_partial_bytes(const SimLaunch& S);
long long sim_n_rirs(const SimLaunch& S);
const void* sim_blob(const SimLaunch& S);
void sim_set_stream_opts(SimLaunch& S, unsigned long long seed, unsigned call_offset);
// Persistent accumulation: d_counters (two unsigned, zero before the first launch and
// left zero by every launch) switches the launch to the persistent kernel; NULL switches back.
void sim_set_counters(SimLaunch& S, unsigned* d_counters);
// Betas [n_rooms][6] and diffuse levels [n_rooms] read from device memory at every launch
// (either may be null: the values planned from RoomHost).  A captured launch reads the
// arrays' current contents at each replay.
void sim_set_params(SimLaunch& S, const float* d_beta, const float* d_level);
const SimOptions& sim_options_of(const SimLaunch& S);
// true if the launch accumulates by the basis method (no image chunks)
bool sim_is_basis(const SimLaunch& S);
int sim_enqueue(const SimLaunch& S, const void* d_plan, float* d_part, const float* d_pos_src, const float* d_pos_rcv,
                const float* d_orv_src, const float* d_orv_rcv, float* d_out, hipStream_t stream);
int sim_enqueue_stage(const SimLaunch& S, int stage, const void* d_plan, float* d_part, const float* d_pos_src,
                      const float* d_pos_rcv, const float* d_orv_src, const float* d_orv_rcv, float* d_out,
                      hipStream_t stream);
int sim_count_taps(const SimLaunch& S, const void* d_plan, const float* d_pos_src, const float* d_pos_rcv,
                   unsigned long long* d_counter, hipStream_t stream);
// Adjoint of a batch: gradients of L with respect to beta [n_rooms][6], source and
// receiver positions (rows as the position arrays) and level [n_rooms], given the
// forward output d_y and dL/dout d_gy.  Any gradient pointer may be null.
struct AdjLaunch;
AdjLaunch* adj_new();
void adj_delete(AdjLaunch* a);
int adj_prepare(const std::vector<RoomHost>& rooms, const SimOptions& opt, AdjLaunch& A);
size_t adj_plan_bytes(const AdjLaunch& A);
const void* adj_blob(const AdjLaunch& A);
size_t adj_ws_bytes(const AdjLaunch& A);
void adj_set_params(AdjLaunch& A, const float* d_beta, const float* d_level);  // as sim_set_params
int adj_enqueue(const AdjLaunch& A, const void* d_plan, void* d_ws, const float* d_pos_src, const float* d_pos_rcv,
                const float* d_orv_src, const float* d_orv_rcv, const float* d_y, const float* d_gy, float* g_beta,
                float* g_src, float* g_rcv, float* g_level, hipStream_t stream);
int launch_adjoint(const std::vector<RoomHost>& rooms, const float* d_pos_src, const float* d_pos_rcv,
                   const float* d_orv_src, const float* d_orv_rcv, const SimOptions& opt, const float* d_y,
                   const float* d_gy, float* g_beta, float* g_src, float* g_rcv, float* g_level, hipStream_t stream);
int launch_image_list(const RoomHost& room, const float* d_pos_src, const float* d_pos_rcv, const float* d_orv_src,
                      const float* d_orv_rcv, const SimOptions& opt, int* d_n0, float* d_frac, float* d_amp,
                      hipStream_t stream);
int launch_simulation(const std::vector<RoomHost>& rooms, const float* d_pos_src, const float* d_pos_rcv,
                      const float* d_orv_src, const float* d_orv_rcv, const SimOptions& opt, float* d_out,
                      hipStream_t stream);
// The same without allocating device or pinned memory: plan, partial rows and work counters
// in a static per-GPU scratch array of NOALLOC_SCRATCH_BYTES (calls on one GPU share it).
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
int launch_simulation_noalloc(const std::vector<RoomHost>& rooms, const float* d_pos_src, const float* d_pos_rcv,
                              const float* d_orv_src, const float* d_orv_rcv, const SimOptions& opt, float* d_out,
                              hipStream_t stream);
// The same with the source/receiver positions and orientations in host memory ([n][3] each;
// orientations may be null): they travel in the plan's upload, one host-to-device copy.
int launch_simulation_host(const std::vector<RoomHost>& rooms, const float* pos_src, int n_src, const float* pos_rcv,
                           int n_rcv, const float* orv_src, const float* orv_rcv, const SimOptions& opt, float* d_out,
                           hipStream_t stream);

}  // namespace gpurir
