/*
 * rvx.h - C ABI of the MI355X (gfx950) receive chain.
 *
 * This is the drop-in boundary for the GPU half of Revenant's receive chain:
 * what core/gpu (Vulkan context, buffers, pipelines, run_kernel) plus the
 * core/dsp operator contract (POD param structs + twin semantics) give the
 * engine today. Every launch takes the same POD parameter block as the
 * reference kernel's push constants, byte for byte, and a small grid struct
 * for what the reference makes specialization constants.
 *
 * Conventions (after core/plugin/engine_plugin_abi.h:22-60, 129-140):
 *   - int32_t return: RVX_OK (0) or a negative RVX_E* code; the message is in
 *     rvx_last_error(), thread-local, valid until the next call on the thread.
 *   - Device pointers are plain pointers into device memory (from
 *     rvx_buf_alloc, hipMalloc or a torch CUDA tensor); every launch also takes
 *     the element count of each span so the library refuses what would be an
 *     out-of-bounds access, exactly where the reference twin refuses.
 *   - Launches are asynchronous on the context's stream. Host pointers are
 *     read during the call only.
 *   - Nothing throws across this line.
 *
 * Complex samples are interleaved float pairs (Complex32, core/dsp/types.h:45).
 */
#ifndef RVX_H
#define RVX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RVX_OK 0
#define RVX_EINVAL (-1)   /* parameters refused (the twin's validate() rules) */
#define RVX_EDEVICE (-2)  /* a HIP call failed */
#define RVX_ENOMEM (-3)   /* allocation failed */
#define RVX_EAGAIN (-4)   /* not yet: the call needs more stream first */

#define RVX_ABI_VERSION 1

/* ------------------------------------------------------------------------ */
/* Parameter blocks                                                          */
/* ------------------------------------------------------------------------ */

/* core/dsp/pfb.h:32-74 GridParams */
typedef struct rvx_grid {
    uint32_t channels;        /* M, power of two */
    uint32_t taps_per_branch; /* L + 1 */
    uint32_t decimation;      /* D, divides M */
} rvx_grid;

/* core/dsp/convert.h:116-141 ConvertParams */
typedef struct rvx_convert_params {
    uint32_t ring_mask;
    uint32_t dst_offset;
    uint32_t src_offset;
    uint32_t count;
} rvx_convert_params;

/* core/dsp/pfb_branch_reference.h:337-360 PfbBranchParams */
typedef struct rvx_pfb_branch_params {
    uint32_t ring_mask;
    uint32_t base_offset;
    uint32_t block_count;
} rvx_pfb_branch_params;

/* core/dsp/pfb_fft_reference.h:42-82 PfbFftParams */
typedef struct rvx_pfb_fft_params {
    uint32_t channels;
    uint32_t decimation;
    uint32_t stages;
    uint32_t block_base;
    uint32_t block_count;
    uint32_t out_ring_blocks;
    uint32_t out_ring_mask;
} rvx_pfb_fft_params;

/* core/dsp/spectrum_reference.h:55-79 SpectrumParams */
typedef struct rvx_spectrum_params {
    uint32_t channels;
    uint32_t transform;
    uint32_t stages;
    uint32_t chan_blocks;
    uint32_t chan_mask;
    uint32_t in_offset;
} rvx_spectrum_params;

/* core/dsp/spectrum_levels_reference.h SpectrumLevelsParams */
typedef struct rvx_spectrum_levels_params {
    uint32_t bins;
    uint32_t low_permille;
    uint32_t high_permille;
} rvx_spectrum_levels_params;

/* The exponentially averaged power spectrum. No reference counterpart: the
 * reference averages on the host, in double, from dB frames
 * (core/detect/detector.cpp:621-627). Operation order, per channel k:
 *   partial: for each frame f, s = 0; for each of its frame_blocks channel
 *            samples v in ascending order: x = v.re*v.re + v.im*v.im;
 *            s = s + (x - s) * alpha
 *   row:     for each row r of row_frames consecutive frames, S = 0; for its
 *            frames in ascending order: S = S * frame_decay + s_f
 *   combine: P = P * row_decay + S_r for ascending r; row r = P.
 * frame_decay = (1 - alpha)^frame_blocks and row_decay = (1 - alpha)^
 * (frame_blocks * row_frames), each rounded once to float. The row level makes
 * the only serial part one step per row instead of one per frame. */
typedef struct rvx_ema_params {
    uint32_t channels;
    uint32_t chan_blocks;
    uint32_t chan_mask;
    uint32_t in_offset;
    uint32_t frame_blocks;
    uint32_t frames;       /* a multiple of row_frames */
    float alpha;
    float frame_decay;
    uint32_t row_frames;   /* >= 1 */
    float row_decay;
} rvx_ema_params;

/* core/dsp/vrx_reference.h:178-195 VrxFineConfig */
typedef struct rvx_vrx_fine_config {
    uint32_t taps;
    uint32_t phases;
    uint32_t nco_log2;
} rvx_vrx_fine_config;

/* core/dsp/vrx_reference.h:232-271 VrxFineParams (fifteen words) */
typedef struct rvx_vrx_fine_params {
    uint32_t chan_base;
    uint32_t chan_mask;
    uint32_t in_offset;
    uint32_t out_offset;
    uint32_t out_mask;
    uint32_t count;
    uint32_t step_whole;
    uint32_t step_rem;
    uint32_t out_rate;
    uint32_t frac0;
    uint32_t nco_phase_high;
    uint32_t nco_phase_low;
    uint32_t nco_delta_high;
    uint32_t nco_delta_low;
    float inv_out_rate;
} rvx_vrx_fine_params;

/* core/dsp/vrx_reference.h:506-544 VrxDemodConfig */
typedef struct rvx_vrx_demod_config {
    uint32_t mode;
    uint32_t decimation;
    uint32_t audio_taps;
    uint32_t dc_taps;
    uint32_t channels;
    uint32_t pilot_taps;
} rvx_vrx_demod_config;

/* core/dsp/vrx_reference.h:569-584 VrxDemodParams */
typedef struct rvx_vrx_demod_params {
    uint32_t in_mask;
    uint32_t in_offset;
    uint32_t count;
    float gain;
} rvx_vrx_demod_params;

/* Demodulator mode ids, core/dsp/vrx_reference.h:300-331 */
#define RVX_DEMOD_RAW 0
#define RVX_DEMOD_AM 1
#define RVX_DEMOD_NFM 2
#define RVX_DEMOD_WFM 3
#define RVX_DEMOD_USB 4
#define RVX_DEMOD_LSB 5
#define RVX_DEMOD_DSB 6
#define RVX_DEMOD_CW 7
#define RVX_DEMOD_SAM 12

/* Carrier loop kinds, core/dsp/vrx_reference.h:674-676: SAM runs the
 * phase-locked loop, DSB the Costas loop, every other mode none. */
#define RVX_CARRIER_NONE 0
#define RVX_CARRIER_PLL 1
#define RVX_CARRIER_COSTAS 2
#define RVX_CARRIER_FLAG_RESET 1u /* vrx_reference.h:694 */
#define RVX_CARRIER_STATE_WORDS 8 /* vrx_reference.h:723-731 */

/* core/dsp/vrx_reference.h:686-690 VrxCarrierConfig */
typedef struct rvx_vrx_carrier_config {
    uint32_t power;    /* RVX_CARRIER_PLL or RVX_CARRIER_COSTAS */
    uint32_t nco_log2; /* 1..24, the NCO table rvx_build_nco_table makes */
} rvx_vrx_carrier_config;

/* core/dsp/vrx_reference.h:699-718 VrxCarrierParams (fourteen words) */
typedef struct rvx_vrx_carrier_params {
    uint32_t in_mask;   /* fine ring capacity - 1 (a power of two minus one) */
    uint32_t in_offset; /* first fine sample rotated */
    uint32_t count;     /* samples rotated, <= capacity */
    uint32_t flags;     /* RVX_CARRIER_FLAG_RESET */
    float shift;        /* radians per sample added to the integrator first */
    float k1_acquire, k2_acquire, k1_track, k2_track;
    float level_alpha, lock_alpha;
    float clamp_radians; /* integrator bound, (0, pi/2] */
    uint32_t hold_in, hold_out;
} rvx_vrx_carrier_params;

/* Native sample formats, core/dsp/convert.h */
#define RVX_FORMAT_CU8 0
#define RVX_FORMAT_CS8 1
#define RVX_FORMAT_CS16 2
#define RVX_FORMAT_CS24 3
/* Converted samples, cf32 pairs: the fused chain's input when conditioning
 * (impulse excision) cleaned them after conversion (rvx_chain_desc.format). */
#define RVX_FORMAT_CF32 4

/* ------------------------------------------------------------------------ */
/* Context, stream, buffers (core/gpu/context.h:64-143, buffer.h:33-73)      */
/* ------------------------------------------------------------------------ */

typedef struct rvx_ctx rvx_ctx;

/* core/gpu/context.cpp:274 Context::create. Creates a context on HIP device
 * `device` with its own stream (default priority). */
int32_t rvx_ctx_create(int32_t device, rvx_ctx** out);
void rvx_ctx_destroy(rvx_ctx* ctx);
/* Launch on `hip_stream` from now on (e.g. torch's current stream); NULL is
 * the device's null stream. */
int32_t rvx_ctx_set_stream(rvx_ctx* ctx, void* hip_stream);
/* Go back to the context's own stream. */
int32_t rvx_ctx_use_own_stream(rvx_ctx* ctx);
void* rvx_ctx_stream(rvx_ctx* ctx);
/* Wait for everything launched on the context's stream. */
int32_t rvx_stream_sync(rvx_ctx* ctx);
const char* rvx_last_error(void);
int32_t rvx_abi_version(void);

#define RVX_MEM_DEVICE 0   /* MemoryKind::DeviceLocal */
#define RVX_MEM_UPLOAD 1   /* MemoryKind::Upload: pinned host, device-readable */
#define RVX_MEM_READBACK 2 /* MemoryKind::Readback: pinned host */

/* core/gpu/buffer.cpp:24 Buffer::create */
int32_t rvx_buf_alloc(rvx_ctx* ctx, uint64_t bytes, int32_t kind, void** out);
int32_t rvx_buf_free(rvx_ctx* ctx, void* ptr);
/* Copies on the context's stream, synchronous with respect to the host. */
int32_t rvx_buf_write(rvx_ctx* ctx, void* dst, const void* src, uint64_t bytes);
int32_t rvx_buf_read(rvx_ctx* ctx, void* dst, const void* src, uint64_t bytes);
/* core/gpu/kernel.h:126-137 CommandRunner::clear: zero fill by a kernel (not
 * a memset node, so it is safe inside a captured graph). */
int32_t rvx_buf_clear(rvx_ctx* ctx, void* dst, uint64_t bytes);

/* Timing on the context's stream (hipEvents): begin, end, then elapsed ms. */
int32_t rvx_timer_begin(rvx_ctx* ctx);
int32_t rvx_timer_end(rvx_ctx* ctx, float* elapsed_ms);

/* ------------------------------------------------------------------------ */
/* Host-side design (core/dsp/pfb_design.cpp, spectrum_reference.cpp,        */
/* vrx_reference.cpp). FP64, rounded once to float.                          */
/* ------------------------------------------------------------------------ */

/* core/dsp/pfb_design.cpp:381-510 design_prototype. attenuation_db selects
 * Kaiser beta by kaiser_beta (:115); if beta_override >= 0 it is used as the
 * Kaiser shape parameter directly. out receives channels*taps_per_branch
 * floats. stopband_db / group_delay may be NULL. */
int32_t rvx_design_prototype(const rvx_grid* grid, double attenuation_db, double beta_override,
                             float* out, uint64_t out_len, double* stopband_db,
                             uint64_t* group_delay);
/* core/dsp/pfb_design.cpp:515-591 build_twiddles: W_M^j, M cf32. */
int32_t rvx_build_twiddles(uint32_t channels, float* out, uint64_t out_len);
/* core/dsp/spectrum_reference.cpp:195-285 build_spectrum_window: N window
 * taps then N/2 channel-shape correction gains. */
int32_t rvx_build_spectrum_window(uint32_t size, uint32_t taps_per_branch,
                                  double attenuation_db, float* out, uint64_t out_len);
/* core/dsp/vrx_reference.cpp:364 build_nco_table: 2^log2_size cf32. */
int32_t rvx_build_nco_table(uint32_t log2_size, float* out, uint64_t out_len);
/* core/dsp/vrx_reference.cpp:308 turn_fixed64 and :333 nco_delta. */
uint64_t rvx_turn_fixed64(int64_t numerator, int64_t denominator);
int32_t rvx_nco_delta(int64_t hz_numerator, int64_t hz_denominator, int64_t rate,
                      uint64_t* out);

/* ------------------------------------------------------------------------ */
/* Receiver planning (core/engine/vrx_place.cpp, core/dsp/vrx_reference.cpp) */
/* ------------------------------------------------------------------------ */

/* core/engine/vrx.h:213-240 Deemphasis */
#define RVX_DEEMPH_DEFAULT 0
#define RVX_DEEMPH_NONE 1
#define RVX_DEEMPH_US75 2
#define RVX_DEEMPH_EU50 3

/* core/engine/vrx.h VrxParams, the part planning reads. */
typedef struct rvx_vrx_request {
    int64_t center_hz;     /* relative to the source's tuned centre */
    int64_t bandwidth_hz;  /* 0: the mode's default passband */
    int64_t passband_low;  /* both 0: derived from bandwidth_hz */
    int64_t passband_high;
    uint32_t mode;         /* RVX_DEMOD_* */
    uint32_t deemphasis;   /* RVX_DEEMPH_*: DEFAULT takes the mode's own */
    int64_t audio_rate;    /* 0: 48000 */
    int64_t cw_pitch;
    uint32_t mono;         /* nonzero: VrxParams::stereo = false. 0 (the default) decodes
                              FM stereo where resolve_stereo allows (WFM programme audio) */
    uint32_t reserved;
} rvx_vrx_request;

/* core/engine/vrx.h:746-783 VrxPlacement */
typedef struct rvx_vrx_placement {
    uint32_t channel;
    uint32_t bandwidth_clamped;
    int64_t centre_num, centre_den;     /* channel centre, reduced rational Hz */
    int64_t residual_num, residual_den; /* receiver centre minus channel centre */
    int64_t channel_rate;
    int64_t granted_low, granted_high;
} rvx_vrx_placement;

/* core/dsp/vrx_reference.h:1334-1476 VrxPlan, the part the kernels use. */
typedef struct rvx_vrx_plan {
    rvx_vrx_placement placement;
    uint32_t mode;
    uint32_t audio_decimation_taps;
    int64_t channel_rate, audio_rate, demod_rate, output_rate;
    int64_t passband_low, passband_high, bandwidth, deviation;
    int64_t filter_num, filter_den, mix_num, mix_den;
    rvx_vrx_fine_config fine;
    uint32_t deemphasis;  /* the curve resolved for the mode and rate (RVX_DEEMPH_*) */
    uint64_t fine_nco_delta;
    rvx_vrx_demod_config demod;
    float demod_gain;
    double audio_pass_hz, audio_stop_hz;
    rvx_vrx_carrier_config carrier;      /* power RVX_CARRIER_NONE: no loop */
    rvx_vrx_carrier_params carrier_loop; /* design_carrier_loop; per-dispatch words zero */
    uint32_t stereo;  /* resolve_stereo (core/engine/vrx.h:381): L and R, a pilot filter */
} rvx_vrx_plan;

/* core/engine/vrx_place.cpp:187 channel_count_for: the largest power of two M
 * with rate/M >= widest_receiver_hz, at least 2 (also for a width <= 0). The
 * reference narrows the 64-bit answer to 32 bits; this one stops at 2^31. */
uint32_t rvx_channel_count_for(int64_t rate, int64_t widest_receiver_hz);
/* core/dsp/pfb_design.cpp:349 channel_centre: channel k's centre k*rate/M as a
 * reduced rational, channels from M/2 up negative (DC is 0/1). */
int32_t rvx_channel_centre(const rvx_grid* grid, int64_t rate, uint32_t channel,
                           int64_t* numerator, int64_t* denominator);
/* core/dsp/pfb_design.cpp:593 channel_sample_to_input: base + m*D minus the
 * prototype's group delay (rvx_design_prototype's *group_delay, M*L/2),
 * modulo 2^64. */
uint64_t rvx_channel_sample_to_input(const rvx_grid* grid, uint64_t group_delay_samples,
                                     uint64_t base, uint64_t m);
/* core/engine/vrx_place.cpp:216 place: nearest channel round(center*M/rate),
 * exact rational residual, the granted passband (clamp_to_channel; an empty
 * grant is reported, plan_vrx refuses it); an NFM/WFM receiver granted less
 * than its mode's default passband is refused (:398-440). */
int32_t rvx_place(const rvx_grid* grid, int64_t rate, const rvx_vrx_request* request,
                  rvx_vrx_placement* out);
/* The planner's fitting queries (core/dsp/vrx_reference.h, vrx_reference.cpp):
 * resolve_passband (the explicit pair, else the bandwidth shorthand: symmetric
 * and one hertz narrow when odd, [0, B] for USB, [-B, 0] for LSB, else the
 * mode's default), default_passband, max_channel_bandwidth (:1409: the
 * channel rate less twice |residual|, rounded down), clamp_to_channel (:1359:
 * each edge into [-max/2, +max/2]; empty when the channel carries nothing),
 * the shorthand minimum_demod_rate (:1519) and demod_rate_for (:1670, the
 * rate plan_vrx builds, refusing what it refuses). */
int32_t rvx_resolve_passband(const rvx_vrx_request* request, int64_t* low, int64_t* high);
int32_t rvx_default_passband(uint32_t mode, int64_t* low, int64_t* high);
int64_t rvx_max_channel_bandwidth(const rvx_vrx_placement* placement);
int32_t rvx_clamp_to_channel(const rvx_vrx_placement* placement, int64_t* low, int64_t* high);
int64_t rvx_minimum_demod_rate(uint32_t mode, int64_t bandwidth, int64_t cw_pitch);
int32_t rvx_demod_rate_for(const rvx_vrx_request* request, const rvx_vrx_placement* placement,
                           int64_t audio_rate, int64_t* out);
/* core/dsp/vrx_reference.cpp:2568 plan_vrx (AM, NFM, WFM mono, USB, LSB,
 * DSB, CW, SAM, raw; no stereo). The de-emphasis curve (resolve_deemphasis,
 * core/engine/vrx.h:302: 75 us on WFM programme audio by default, none
 * elsewhere) is folded into the audio filter, so demod.audio_taps is the
 * folded length (:2462-2509). SAM and DSB plans carry their carrier loop
 * (:2550-2561). */
int32_t rvx_plan_vrx(const rvx_grid* grid, int64_t rate, const rvx_vrx_request* request,
                     const rvx_vrx_placement* placement, rvx_vrx_plan* out);
/* The plan's tables: fine taps (2*(phases*taps+1) floats, design_fine_taps
 * :1808) and demod weights (audio_taps + dc_taps floats: design_audio_taps
 * :1921 folded with design_deemphasis_taps :2011 by fold_deemphasis :2079,
 * then design_dc_weights :1983). Either pointer may be NULL. */
int32_t rvx_plan_vrx_tables(const rvx_vrx_plan* plan, float* fine_taps, uint64_t n_fine_floats,
                            float* weights, uint64_t n_weights);
/* :2011 design_deemphasis_taps: a resolved curve's taps at an audio rate
 * (out may be NULL to ask the length, returned in *n). */
int32_t rvx_design_deemphasis(uint32_t curve, int64_t audio_rate, float* out, uint32_t max_taps,
                              uint32_t* n);
/* core/dsp/vrx_reference.cpp:2855 fine_block and :2873 demod_block: one
 * dispatch's params from the absolute index of its first output. */
int32_t rvx_fine_block(const rvx_vrx_plan* plan, uint32_t channel_base, uint32_t channel_mask,
                       uint32_t fine_mask, uint64_t first_output, uint32_t count,
                       rvx_vrx_fine_params* out, uint64_t* first_input);
int32_t rvx_demod_block(const rvx_vrx_plan* plan, uint32_t fine_mask, uint64_t first_audio,
                        uint32_t count, rvx_vrx_demod_params* out);
/* core/dsp/vrx_reference.h:555 demod_fine_history. */
uint32_t rvx_demod_history(const rvx_vrx_demod_config* config);

/* ---- The display tap: vrx_reference.h:1639-1768 ---------------------------------
 * A receiver's passband pane: the fine kernel (rvx_vrx_fine) run a second time
 * with a 256-tap, one-phase anti-alias table, reading the channel before any
 * receiver filter, mixing by the fine stage's own frequency and decimating by
 * an integer R from the channel rate's ladder, so the pane is flat across its
 * width and the client draws the receiver's filter over it. */
#define RVX_DISPLAY_TAPS 256

/* vrx_reference.h:1727-1754 VrxDisplayPlan (the tap table is separate) */
typedef struct rvx_vrx_display_plan {
    int64_t channel_rate, display_rate;
    uint32_t decimation, reserved;
    int64_t reach_hz;                 /* the passband's further edge from DC */
    rvx_vrx_fine_config fine;         /* taps 256, phases 1, the plan's NCO */
    uint32_t reserved2;
    uint64_t nco_delta;               /* the fine mix at the display rate */
    double pass_hz, stop_hz, attenuation_db;
    double group_delay_channel_samples;
} rvx_vrx_display_plan;

/* :1721 display_max_decimation (25). */
uint32_t rvx_display_max_decimation(void);
/* :1716 display_ladder: the rungs for a channel rate, ascending from 1 (writes
 * up to max_rungs, returns how many there are). */
uint32_t rvx_display_ladder(int64_t channel_rate, uint32_t* out, uint32_t max_rungs);
/* :1725 display_decimation. */
uint32_t rvx_display_decimation(int64_t channel_rate, int64_t reach_hz);
/* :1758 plan_vrx_display. taps: 2 * (RVX_DISPLAY_TAPS + 1) floats, or NULL. */
int32_t rvx_plan_vrx_display(const rvx_vrx_plan* plan, rvx_vrx_display_plan* out, float* taps,
                             uint64_t n_floats);
/* :1762 display_block: rvx_vrx_fine params for display outputs
 * [first_output, first_output + count). */
int32_t rvx_display_block(const rvx_vrx_display_plan* plan, uint32_t channel_base,
                          uint32_t channel_mask, uint32_t display_mask, uint64_t first_output,
                          uint32_t count, rvx_vrx_fine_params* out, uint64_t* first_input);

/* Carrier loop design, core/dsp/vrx_reference.cpp:918-1185. */
/* :918 carrier_power: RVX_CARRIER_* for a demod mode. */
uint32_t rvx_carrier_power(uint32_t mode);
/* :948 carrier_loop_gains: K1, K2 of the second-order loop for a noise
 * bandwidth and damping at a rate, in double (both zero for a non-positive
 * argument, as the reference returns). */
int32_t rvx_carrier_loop_gains(double noise_bandwidth_hz, double damping, int64_t rate,
                               double* k1, double* k2);
/* :958 design_carrier_loop: the loop for a demodulation rate and a passband
 * reaching reach_hz from the receiver's centre (per-dispatch words zero). */
int32_t rvx_design_carrier_loop(int64_t demod_rate, int64_t reach_hz,
                                rvx_vrx_carrier_params* out);
/* :1187 carrier_retune: a move further than the granted width restarts the
 * loop; a move inside it carries the loop by -2*pi*moved/demod_rate. */
int32_t rvx_carrier_retune(int64_t moved_hz, int64_t granted_width, int64_t demod_rate,
                           int32_t* reset, double* shift);
/* :2911 carrier_block: one dispatch's params from the absolute index of its
 * first fine sample (a reset drops the shift). */
int32_t rvx_carrier_block(const rvx_vrx_plan* plan, uint32_t fine_mask, uint64_t first_fine,
                          uint32_t count, int32_t reset, float shift,
                          rvx_vrx_carrier_params* out);
/* :986 validate(VrxCarrierConfig, VrxCarrierParams). */
int32_t rvx_vrx_carrier_check(const rvx_vrx_carrier_config* config,
                              const rvx_vrx_carrier_params* params);

/* ------------------------------------------------------------------------ */
/* Operators. Each replaces one reference compute shader + its dispatch.     */
/* ------------------------------------------------------------------------ */

/* core/shaders/convert_{cu8,cs8,cs16,cs24}_cf32.comp; twin
 * core/dsp/convert.cpp:153,199,235,271. format = RVX_FORMAT_*. packed is the
 * staging buffer as the device sees it, packed_bytes its size. */
int32_t rvx_convert(rvx_ctx* ctx, int32_t format, const void* packed, uint64_t packed_bytes,
                    const rvx_convert_params* params, void* ring, uint64_t ring_samples);

/* core/shaders/cmul.comp; twin core/dsp/complex_ops.cpp:33. The device
 * denormal probe and plumbing check. */
int32_t rvx_cmul(rvx_ctx* ctx, const void* lhs, const void* rhs, void* out, uint64_t count);

/* core/shaders/pfb_branch.comp; twin core/dsp/pfb_branch_reference.cpp:22.
 * taps: M*(L+1) floats. iq_ring: ring_samples cf32. out: block_count*M cf32,
 * block-major, slot (M - r) & (M - 1). */
int32_t rvx_pfb_branch(rvx_ctx* ctx, const rvx_grid* grid, const float* taps, uint64_t n_taps,
                       const void* iq_ring, uint64_t ring_samples,
                       const rvx_pfb_branch_params* params, void* out, uint64_t out_samples);

/* core/shaders/pfb_fft.comp; twin core/dsp/pfb_fft_reference.cpp:143.
 * branch: block_count*M cf32; twiddles: M cf32; channel ring: M *
 * out_ring_blocks cf32, channel-major. */
int32_t rvx_pfb_fft(rvx_ctx* ctx, const rvx_pfb_fft_params* params, const void* branch,
                    uint64_t branch_samples, const void* twiddles, uint64_t n_twiddles,
                    void* channel_ring, uint64_t ring_samples);

/* Twin core/dsp/pfb_fft_reference.cpp:234 reference_fft_radix2: batch
 * transforms of `size` points, no correction. */
int32_t rvx_fft_radix2(rvx_ctx* ctx, const void* twiddles, uint64_t n_twiddles, const void* in,
                       void* out, uint32_t size, uint32_t batch);

/* core/shaders/spectrum.comp; twin core/dsp/spectrum_reference.cpp:286.
 * window: N + N/2 floats from rvx_build_spectrum_window. frame: M*N/2 dB. */
int32_t rvx_spectrum(rvx_ctx* ctx, const rvx_spectrum_params* params, const void* channel_ring,
                     uint64_t ring_samples, const void* twiddles, uint64_t n_twiddles,
                     const float* window, uint64_t n_window, float* frame, uint64_t n_frame);

/* core/shaders/spectrum_levels.comp; twin
 * core/dsp/spectrum_levels_reference.cpp:38. levels: 2 floats. */
int32_t rvx_spectrum_levels(rvx_ctx* ctx, const rvx_spectrum_levels_params* params,
                            const float* frame, uint64_t n_frame, float* levels);

/* The averaged power spectrum (see rvx_ema_params). partial: frames*channels
 * floats. state: channels floats, read and updated. rows: (frames /
 * row_frames) * channels floats, required. */
int32_t rvx_ema_partial(rvx_ctx* ctx, const rvx_ema_params* params, const void* channel_ring,
                        uint64_t ring_samples, float* partial, uint64_t n_partial);
int32_t rvx_ema_combine(rvx_ctx* ctx, const rvx_ema_params* params, const float* partial,
                        uint64_t n_partial, float* state, float* rows, uint64_t n_rows);

/* core/shaders/vrx_fine.comp; twin core/dsp/vrx_reference.cpp:489. The
 * stateless rational resampler + complex FIR + 0.64 NCO rotation. channel
 * ring: reads chan_base + ((in_offset + n - k) & chan_mask). taps: phases *
 * taps + 1 cf32. nco: 2^nco_log2 cf32. fine ring: out_mask + 1 cf32. */
int32_t rvx_vrx_fine(rvx_ctx* ctx, const rvx_vrx_fine_config* config,
                     const rvx_vrx_fine_params* params, const void* channel_ring,
                     uint64_t ring_samples, const void* taps, uint64_t n_taps, const void* nco,
                     uint64_t n_nco, void* fine_ring, uint64_t fine_samples);

/* core/shaders/vrx_demod.comp; twin core/dsp/vrx_reference.cpp:714. weights:
 * audio_taps + dc_taps (+ 2*pilot_taps + 2*audio_taps for FM stereo) floats.
 * audio: count * channels floats. */
int32_t rvx_vrx_demod(rvx_ctx* ctx, const rvx_vrx_demod_config* config,
                      const rvx_vrx_demod_params* params, const void* fine_ring,
                      uint64_t fine_samples, const float* weights, uint64_t n_weights,
                      float* audio, uint64_t n_audio);

/* core/shaders/vrx_carrier.comp; twin core/dsp/vrx_reference.cpp:1052. The
 * serial carrier loop (PLL for SAM, Costas for DSB): rotates params->count
 * fine samples in place from in_offset onto the recovered carrier and
 * carries RVX_CARRIER_STATE_WORDS words of state (device) between calls. nco:
 * 2^nco_log2 cf32. fine ring: >= in_mask + 1 cf32. */
int32_t rvx_vrx_carrier(rvx_ctx* ctx, const rvx_vrx_carrier_config* config,
                        const rvx_vrx_carrier_params* params, const void* nco, uint64_t n_nco,
                        void* fine_ring, uint64_t fine_samples, uint32_t* state);

/* Many receivers' carrier loops in one launch, one lane per receiver, each
 * bit-identical to rvx_vrx_carrier on its own. No counterpart in the
 * reference, which dispatches one single-invocation kernel per receiver
 * (core/engine/vrx_stage.cpp:1346). Items live in device memory; the host
 * checks each with rvx_vrx_carrier_item_check before uploading it, and the
 * kernel leaves an item whose sizes disagree untouched. Every item uses the
 * one NCO table `nco` (n_nco = 2^config.nco_log2 cf32). Items must not share
 * a fine ring slot or a state block. */
typedef struct rvx_vrx_carrier_item {
    rvx_vrx_carrier_config config;
    rvx_vrx_carrier_params params;
    void* fine_ring;    /* device, fine_samples cf32 */
    uint64_t fine_samples;
    uint32_t* state;    /* device, RVX_CARRIER_STATE_WORDS */
    uint64_t reserved;
} rvx_vrx_carrier_item;
int32_t rvx_vrx_carrier_item_check(const rvx_vrx_carrier_item* item, uint64_t n_nco);
int32_t rvx_vrx_carrier_batch(rvx_ctx* ctx, const rvx_vrx_carrier_item* items, uint32_t n_items,
                              const void* nco, uint64_t n_nco);

/* Batched receivers: many VRXs of any mix of modes and rates in one
 * persistent launch, each computing fine stage -> detector -> audio filter on
 * chip per tile of audio outputs, bit-identical to rvx_vrx_fine followed by
 * rvx_vrx_demod. No counterpart in the reference, which records two or three
 * dispatches per receiver (core/engine/vrx_stage.cpp:1306,1346,1399).
 *
 * One descriptor per receiver and step. `fine` holds the params of the fine
 * block the step needs: fine outputs F0 .. F0 + (count-1)*R + hist, where F0
 * = first_audio * R - hist and hist = rvx_demod_history(&demod); its
 * out_offset/out_mask are unused (fine samples stay on chip). `demod.count`
 * audio outputs are written to `audio` (count * channels floats); its
 * in_offset/in_mask are unused. `tile` is the audio outputs one work item
 * computes (chosen by the tuner). */
typedef struct rvx_vrx_desc {
    rvx_vrx_fine_config fine_config;
    rvx_vrx_demod_config demod_config;
    rvx_vrx_fine_params fine;
    rvx_vrx_demod_params demod;
    uint32_t hist;
    uint32_t tile;
    uint32_t reserved[2];
    const void* channel_ring;  /* device, cf32; fine.chan_base/chan_mask index it */
    const void* taps;          /* device, fine_config.phases*taps + 1 cf32 */
    const float* weights;      /* device, demod weights */
    float* audio;              /* device, demod.count * channels */
} rvx_vrx_desc;

/* Work list: (receiver, tile) pairs in the order the persistent workgroups
 * take them; the host orders them largest-first. */
typedef struct rvx_vrx_work {
    uint32_t vrx;
    uint32_t tile;
} rvx_vrx_work;

/* descs / work / counter are device memory (descs and work written by the
 * caller, e.g. with rvx_buf_write; the counter is one uint32 the launch
 * zeroes itself, so a captured graph replays correctly). nco: 2^16 cf32. */
int32_t rvx_vrx_batch(rvx_ctx* ctx, const rvx_vrx_desc* descs, uint32_t n_desc,
                      const rvx_vrx_work* work, uint32_t n_work, uint32_t* counter,
                      const void* nco, uint64_t n_nco, uint32_t workgroups);

/* No counterpart in the reference: rvx_vrx_batch for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid allocation). Same descriptors, work list, kernels and
 * bits; the queue counter lives in the library's static scratch (256 bytes
 * per device, set up by rvx_ctx_create) instead of a caller's buffer, and
 * nothing is allocated or freed. Every call on one device shares that
 * counter: calls that overlap on one device (two streams, or two threads)
 * race, so the caller serialises them. At most 2^31 - 1 work items. */
int32_t rvx_vrx_batch_noalloc(rvx_ctx* ctx, const rvx_vrx_desc* descs, uint32_t n_desc,
                              const rvx_vrx_work* work, uint32_t n_work, const void* nco,
                              uint64_t n_nco, uint32_t workgroups);

/* Host-side check of one descriptor (sizes are the caller's to guarantee for
 * device pointers; this checks the parameter rules of rvx_vrx_fine/demod and
 * that a tile fits the LDS budget). */
int32_t rvx_vrx_desc_check(const rvx_vrx_desc* desc);

/* Fine samples one batched work item holds on chip: a receiver's tile of n
 * audio outputs needs (n - 1) * R + hist + 1 <= RVX_VRX_TILE_SPAN. */
#define RVX_VRX_TILE_SPAN 3072

/* One receiver's share of a batched launch: audio outputs first_audio ..
 * first_audio + count - 1, written to `audio` (count * channels floats). */
typedef struct rvx_vrx_span {
    const rvx_vrx_plan* plan;  /* host */
    uint32_t channel_base;     /* cf32 index of the receiver's channel in channel_ring */
    uint32_t channel_mask;     /* channel length - 1 */
    const void* channel_ring;  /* device */
    const void* taps;          /* device, the plan's fine taps (phases*taps + 1 cf32) */
    const float* weights;      /* device, the plan's demod weights */
    float* audio;              /* device */
    uint64_t first_audio;
    uint32_t count;
    uint32_t tile;             /* audio outputs per work item; 0: the largest that fits */
} rvx_vrx_span;

/* Host: the descriptors and the work list of one rvx_vrx_batch launch over
 * `spans` (the per-step record of DemodStage::record,
 * core/engine/vrx_stage.cpp:1119-1480, for a whole receiver set). A span is
 * split into several descriptors where the fine block's 32-bit resampler
 * recurrence would overflow; work items are ordered by estimated cost, largest
 * first. With null descs/work only the counts are returned. */
int32_t rvx_vrx_batch_plan(const rvx_vrx_span* spans, uint32_t n_spans, rvx_vrx_desc* descs,
                           uint32_t max_descs, uint32_t* n_descs, rvx_vrx_work* work,
                           uint32_t max_work, uint32_t* n_work);

/* Graph capture of the context's own stream (HIP graphs): begin, end, launch. */
typedef struct rvx_graph rvx_graph;
int32_t rvx_graph_begin(rvx_ctx* ctx);
int32_t rvx_graph_end(rvx_ctx* ctx, rvx_graph** out);
int32_t rvx_graph_launch(rvx_ctx* ctx, rvx_graph* graph);
void rvx_graph_destroy(rvx_graph* graph);

/* Synthetic IQ source (the inputs of tests and benchmarks; after
 * core/dsp/synth, whose scenes are pure functions of the absolute index).
 * Each sample n = start + i is the FP64 sum over carriers of
 *   env * exp(j 2 pi (f n / rate + fm(n)))   with phases reduced exactly
 * plus complex Gaussian noise (Box-Muller over a splitmix64 hash of seed and
 * n), rounded to float and quantised: cs16 (x 32768, saturated), cu8
 * (127.5 + 127.5 x, saturated) or format -1 for cf32. */
#define RVX_CARRIER_CW 0
#define RVX_CARRIER_AM 1
#define RVX_CARRIER_FM 2

typedef struct rvx_carrier {
    int32_t kind;        /* RVX_CARRIER_* */
    int32_t reserved;
    int64_t freq_hz;     /* offset from the tuned centre */
    int64_t tone_hz;     /* modulating tone */
    double amplitude;    /* linear, full scale 1 */
    double index;        /* FM: peak phase deviation in radians (deviation / tone) */
    double depth;        /* AM: modulation depth */
} rvx_carrier;

typedef struct rvx_synth_params {
    int64_t rate;
    uint64_t start;
    uint64_t count;
    uint64_t seed;
    double noise_sigma;  /* per component */
    int32_t format;      /* RVX_FORMAT_CS16, RVX_FORMAT_CU8 or -1 (cf32) */
    int32_t reserved;
} rvx_synth_params;

/* Renders the scene into `out` (count samples) and waits for it. */
int32_t rvx_synth(rvx_ctx* ctx, const rvx_carrier* carriers, uint32_t ncarriers,
                  const rvx_synth_params* params, void* out, uint64_t out_bytes);

/* The fused plain chain: convert -> pfb_branch -> pfb_fft (-> ema_partial) in
 * one launch, bit-identical to running the four operators in sequence
 * (core/engine/graph.cpp:5120-5196 records them as separate dispatches).
 * input: native codes from absolute sample `input_origin` (input_samples of
 * them); block m (absolute, first_block <= m < first_block + block_count) is
 * taken at base m*D and needs samples m*D - M*(L+1) + 1 .. m*D. Writes the
 * channel ring (channel-major, slot m & (out_ring_blocks - 1)) and, when
 * frame_blocks != 0, the EMA partials of block_count / frame_blocks frames.
 * channel_rows (device, M uint32, optional): channel k is stored in ring row
 * channel_rows[k] of a ring of ring_rows rows, and not stored when that is
 * >= ring_rows; null stores every channel in row k (ring_rows is then
 * ignored). A coarse grid that exists only to feed a few wide receivers keeps
 * just their channels; the averaged spectrum still covers every channel.
 * twiddles: rvx_build_twiddles(M). The kernel relies on that table's exact
 * quadrant symmetry (entry k + M/4 is entry k turned by -i, bit for bit) and
 * reads only the entries it needs to form the rest. */
typedef struct rvx_chain_desc {
    int32_t format;           /* RVX_FORMAT_CU8, RVX_FORMAT_CS16 or RVX_FORMAT_CF32 */
    uint32_t channels;
    uint32_t taps_per_branch;
    uint32_t decimation;
    uint64_t input_origin;
    uint64_t input_samples;
    uint64_t first_block;
    uint32_t block_count;
    uint32_t out_ring_blocks;
    uint32_t frame_blocks;
    float alpha;
    uint32_t ring_rows;
    uint32_t taps_scaled;     /* 1: cs16 input and taps from rvx_prescale_taps */
    const uint32_t* channel_rows;
    /* Front-end correction applied to every converted sample before the
     * branch filter, as rvx_iq_correct on the converted ring would (its ring
     * fields are ignored): samples with absolute index >= correction_from take
     * `correction`, earlier ones `correction_before` (NULL: `correction`).
     * NULL `correction`: none. Not with taps_scaled. */
    const struct rvx_iq_correct_params* correction;
    const struct rvx_iq_correct_params* correction_before;
    uint64_t correction_from;
    /* Samples from one channel's ring row to the next: 0 is out_ring_blocks
     * (the reference's channel_ring_index, pfb_fft_reference.h:157-162);
     * otherwise at least out_ring_blocks and a multiple of 4. Row r starts
     * at r * ring_pitch, and ring_samples must cover
     * (rows - 1) * ring_pitch + out_ring_blocks. A power-of-two row length
     * puts every channel's write of one block on the same few HBM channels;
     * padding each row by 32 blocks (256 bytes) spreads them
     * (RVX_RING_PITCH_PAD). */
    uint32_t ring_pitch;
    uint32_t reserved;
} rvx_chain_desc;

/* The row padding (blocks) the benchmark and the engine's padded layout use. */
#define RVX_RING_PITCH_PAD 32

/* taps * 2^-15 exactly (host arrays), for a cs16 chain with taps_scaled = 1:
 * the conversion's scale moves into the taps and the products are unchanged
 * bit for bit. Refuses a tap whose scaled value would not be a normal float
 * (0 < |h| < 2^-111) or is not finite. */
int32_t rvx_prescale_taps(const float* taps, uint64_t n, float* out);

int32_t rvx_chain(rvx_ctx* ctx, const rvx_chain_desc* desc, const void* input,
                  const float* taps, uint64_t n_taps, const void* twiddles, uint64_t n_twiddles,
                  void* channel_ring, uint64_t ring_samples, float* partial, uint64_t n_partial);

/* ---- Front-end correction ---------------------------------------------------
 * A zero-IF front end's DC offset and I/Q imbalance, removed before the
 * channelizer (core/dsp/front_end_correction.h). The device sums five moments
 * per 64-sample chunk of a block; the host averages them into an estimate and
 * turns it into the correction's four numbers; the device applies
 * i' = i - dc_i, q' = cross * i' + scale * (q - dc_q) in place. */
#define RVX_IQ_MOMENTS_CHUNK 64     /* front_end_correction.h:84 */
#define RVX_IQ_MOMENTS_PER_CHUNK 5  /* :88 (I, Q, II, QQ, IQ) */

/* front_end_correction.h:96-107 IqMomentsParams */
typedef struct rvx_iq_moments_params {
    uint32_t ring_mask;
    uint32_t src_offset;
    uint32_t count;
} rvx_iq_moments_params;

/* front_end_correction.h:116-130 IqCorrectParams; the identity is dc 0,
 * cross 0, scale 1 */
struct rvx_iq_correct_params {
    uint32_t ring_mask;
    uint32_t offset;
    uint32_t count;
    float dc_i, dc_q, cross, scale;
};
typedef struct rvx_iq_correct_params rvx_iq_correct_params;

/* core/shaders/iq_moments.comp; twin front_end_correction_reference.cpp:46.
 * sums: ceil(count / 64) * 5 floats (chunk-major). */
int32_t rvx_iq_moments(rvx_ctx* ctx, const rvx_iq_moments_params* params, const void* ring,
                       uint64_t ring_samples, float* sums, uint64_t n_sums);
/* The same chunk sums over native codes (RVX_FORMAT_CU8 / CS16) read from
 * sample `first` of a linear buffer, each converted as rvx_convert converts
 * it: the moments of the ring the reference's convert would have written.
 * No counterpart in the reference, whose moments read the converted ring. */
int32_t rvx_iq_moments_codes(rvx_ctx* ctx, int32_t format, const void* codes,
                             uint64_t code_samples, uint64_t first, uint32_t count, float* sums,
                             uint64_t n_sums);
/* core/shaders/iq_correct.comp; twin front_end_correction_reference.cpp:90.
 * in and out may be the same ring. */
int32_t rvx_iq_correct(rvx_ctx* ctx, const rvx_iq_correct_params* params, const void* in,
                       uint64_t in_samples, void* out, uint64_t out_samples);

/* front_end_correction.h:151-160 IqMoments */
typedef struct rvx_iq_moments_total {
    uint64_t samples;
    double sum_i, sum_q, sum_ii, sum_qq, sum_iq;
} rvx_iq_moments_total;
/* front_end_correction.h:164-189 FrontEndCorrectionConfig */
typedef struct rvx_front_end_config {
    double dc_time_constant_s;   /* 0.1 */
    double iq_time_constant_s;   /* 1.0 */
    double max_gain_error_db;    /* 6 */
    double max_phase_error_deg;  /* 30 */
} rvx_front_end_config;
/* front_end_correction.h:191-219 FrontEndEstimate */
typedef struct rvx_front_end_estimate {
    int32_t measured;
    int32_t iq_plausible;
    uint64_t samples_seen;
    double dc_i, dc_q;
    double gain, sin_phase;
} rvx_front_end_estimate;
typedef struct rvx_front_end rvx_front_end;

/* front_end_correction.cpp:39 total_moments: a block's chunk sums added in
 * chunk order, in double. */
int32_t rvx_total_moments(const float* sums, uint64_t n_sums, uint32_t count,
                          rvx_iq_moments_total* out);
void rvx_front_end_default_config(rvx_front_end_config* out);
/* FrontEndCorrector (front_end_correction.h:222-247); config NULL: defaults. */
int32_t rvx_front_end_create(const rvx_front_end_config* config, rvx_front_end** out);
void rvx_front_end_destroy(rvx_front_end* f);
void rvx_front_end_reset(rvx_front_end* f);
int32_t rvx_front_end_update(rvx_front_end* f, const rvx_iq_moments_total* moments,
                             int64_t rate);
int32_t rvx_front_end_estimate_get(const rvx_front_end* f, rvx_front_end_estimate* out);
/* The correction's numbers (ring fields zero): the identity for a half that
 * is off, not yet measured, or (I/Q) implausible. */
int32_t rvx_front_end_correction(const rvx_front_end* f, int32_t dc, int32_t iq,
                                 rvx_iq_correct_params* out);
/* FrontEndEstimate's readings: dc_dbfs, gain_error_db, phase_error_deg,
 * image_rejection_db (front_end_correction.cpp:20-37); any output may be NULL. */
int32_t rvx_front_end_derived(const rvx_front_end_estimate* e, double* dc_dbfs,
                              double* gain_error_db, double* phase_error_deg,
                              double* image_rejection_db);

/* ---- Impulse excision --------------------------------------------------------
 * core/dsp/impulse_excision.h: wideband impulses cut out of the sample ring
 * before the channelizer (zeroed in place as -0, their values moved to a
 * removed ring so cleaned + removed == raw bit for bit). Four passes per
 * block, planned on the host by the cursor. */
#define RVX_EXCISE_SEGMENT 0
#define RVX_EXCISE_REFERENCE 1
#define RVX_EXCISE_FLAG 2
#define RVX_EXCISE_APPLY 3
#define RVX_EXCISE_COUNTERS 4 /* samples excised, events, runs spared, (unused) */

/* impulse_excision.h:155-168 ExcisionConfig */
typedef struct rvx_excision_config {
    uint32_t pass;
    uint32_t seg_log2;   /* segment of 2^seg_log2 samples, 2..12 */
    uint32_t segments;   /* reference window, 2..64 */
    uint32_t max_width;  /* longest run excised, 1..4096 */
    uint32_t hang, lead; /* guard after and before a flag, <= 16 */
} rvx_excision_config;

/* impulse_excision.h:185-201 ExcisionParams (eight words) */
typedef struct rvx_excision_params {
    uint32_t ring_mask;
    uint32_t first;        /* ring slot of the first sample (or segment's first sample) */
    uint32_t count;        /* samples (flag, apply) or segments (segment, reference) */
    uint32_t seg_mask, flag_mask, removed_mask;
    int32_t history;       /* samples of this run before `first` */
    float threshold;       /* T as a power ratio */
} rvx_excision_params;

/* impulse_excision.h:219-223 ExcisionRings */
typedef struct rvx_excision_rings {
    uint32_t seg_mask, flag_mask, removed_mask;
} rvx_excision_rings;

/* impulse_excision.h:228-243 ExcisionPlan: one block's four dispatches */
typedef struct rvx_excision_plan {
    rvx_excision_params segment, reference, flag, apply;
    uint64_t cleaned_end;  /* one past the last finished sample */
    int32_t restarted;
    uint32_t reserved;
} rvx_excision_plan;

/* impulse_excision_reference.cpp:62 design_excision (max width 20 us, guard
 * 0.1 us (2..16 samples), segments >= a quarter of the width, 32 of them). */
int32_t rvx_design_excision(int64_t source_rate, double max_width_seconds,
                            rvx_excision_config* out);
/* :92 excision_threshold: 10^(dB/10) rounded to float once. */
float rvx_excision_threshold(double threshold_db);
/* :96 validate(ExcisionConfig, ExcisionParams) for config->pass. */
int32_t rvx_excision_check(const rvx_excision_config* config, const rvx_excision_params* params);
/* :160 size_excision_rings. */
int32_t rvx_size_excision_rings(const rvx_excision_config* config, uint64_t max_block,
                                uint32_t frames_in_flight, uint32_t ring_mask,
                                rvx_excision_rings* out);
/* :190-247 ExcisionCursor: which samples each pass covers, block by block. */
typedef struct rvx_excision_cursor rvx_excision_cursor;
int32_t rvx_excision_cursor_create(const rvx_excision_config* config, uint32_t ring_mask,
                                   const rvx_excision_rings* rings, float threshold,
                                   rvx_excision_cursor** out);
void rvx_excision_cursor_destroy(rvx_excision_cursor* cursor);
void rvx_excision_cursor_reset(rvx_excision_cursor* cursor);
int32_t rvx_excision_cursor_plan(rvx_excision_cursor* cursor, uint64_t start, uint32_t count,
                                 rvx_excision_plan* out);
/* core/shaders/impulse_excise.comp, the four passes (the config's pass field
 * is set by each call); twins impulse_excision_reference.cpp:250-387. raw and
 * cleaned may be the same ring. counters: RVX_EXCISE_COUNTERS words the apply
 * pass adds to. */
int32_t rvx_excise_segment(rvx_ctx* ctx, const rvx_excision_config* config,
                           const rvx_excision_params* params, const void* raw,
                           uint64_t raw_samples, float* seg_power, uint64_t n_seg);
int32_t rvx_excise_reference(rvx_ctx* ctx, const rvx_excision_config* config,
                             const rvx_excision_params* params, const float* seg_power,
                             float* seg_level, uint64_t n_seg);
int32_t rvx_excise_flag(rvx_ctx* ctx, const rvx_excision_config* config,
                        const rvx_excision_params* params, const void* raw, uint64_t raw_samples,
                        const float* seg_level, uint64_t n_seg, uint32_t* flags, uint64_t n_flags);
int32_t rvx_excise_apply(rvx_ctx* ctx, const rvx_excision_config* config,
                         const rvx_excision_params* params, const void* raw, uint64_t raw_samples,
                         const uint32_t* flags, uint64_t n_flags, void* cleaned,
                         uint64_t cleaned_samples, void* removed, uint64_t removed_samples,
                         uint32_t* counters);

/* ---- Comb cancellation (device half) ------------------------------------------
 * core/dsp/comb_cancel.h: a harmonic comb of spurs projected onto its lines
 * and subtracted from the sample ring before the channelizer. The lines'
 * frequencies, phases and amplitudes come from the host's comb tracker. */
#define RVX_COMB_MAX_LINES 1024   /* comb_cancel.h:109 */
#define RVX_COMB_TABLE_LOG2 12    /* :117 */
#define RVX_COMB_MAX_SUB_BLOCKS 16 /* :123 */

/* comb_cancel.h:248-254 CombLine (six words) */
typedef struct rvx_comb_line {
    uint32_t phase0;     /* the line's 0.32 phase at the dispatch's first sample */
    uint32_t step;       /* per sample (rvx_comb_step) */
    uint32_t pad0, pad1;
    float amplitude[2];  /* what subtract removes; project ignores it */
} rvx_comb_line;

/* comb_cancel.h:276-285 CombParams (eight words) */
typedef struct rvx_comb_params {
    uint32_t ring_mask;
    uint32_t first;       /* ring slot of the dispatch's first sample */
    uint32_t count;       /* a multiple of sub_blocks */
    uint32_t lines;
    uint32_t sub_blocks;
    uint32_t removed_mask;
    uint32_t pad0, pad1;
} rvx_comb_params;

/* comb_kernels_reference.cpp:44 comb_step: round(f / rate * 2^32) mod 2^32. */
uint32_t rvx_comb_step(double frequency_hz, int64_t rate);
/* :58 comb_table: 2^12 (cos, sin) pairs, 2 * 4096 floats. */
int32_t rvx_comb_table(float* out, uint64_t n_floats);
/* :67 validate(CombParams). */
int32_t rvx_comb_check(const rvx_comb_params* params);
/* core/shaders/comb_project.comp; twin comb_kernels_reference.cpp:100.
 * projections: lines * sub_blocks cf32, line-major. lines, table: device. */
int32_t rvx_comb_project(rvx_ctx* ctx, const rvx_comb_params* params, const void* raw,
                         uint64_t raw_samples, const rvx_comb_line* lines, uint64_t n_lines,
                         const void* table, uint64_t n_table, void* projections,
                         uint64_t n_projections);
/* core/shaders/comb_subtract.comp; twin :143. raw and cleaned may be the same
 * ring; removed is indexed from the same slot under removed_mask. */
int32_t rvx_comb_subtract(rvx_ctx* ctx, const rvx_comb_params* params, const void* raw,
                          uint64_t raw_samples, const rvx_comb_line* lines, uint64_t n_lines,
                          const void* table, uint64_t n_table, void* cleaned,
                          uint64_t cleaned_samples, void* removed, uint64_t removed_samples);

/* ---- Comb cancellation (host half) -------------------------------------------
 * core/dsp/comb_cancel.h: detect_comb/detect_combs (comb_detect.cpp), the
 * CombTracker (comb_tracker.cpp) and the CombBank (comb_bank.cpp). The bank
 * plans each block's lines for rvx_comb_project/subtract (or
 * rvx_engine_set_comb) and reads back that block's projections. */
#define RVX_COMB_MIN_LINES 8            /* comb_cancel.h:112 */
#define RVX_COMB_MAX_COMBS 8            /* :126 */
#define RVX_COMB_MAX_SPURS 16           /* :152 */
#define RVX_COMB_SAME_LINE_HZ 45.0      /* :147 */
#define RVX_COMB_OVERLAP_FRACTION 0.25  /* :141 */

/* comb_cancel.h:310 CombLock */
#define RVX_COMB_IDLE 0
#define RVX_COMB_ACQUIRING 1
#define RVX_COMB_LOCKED 2
#define RVX_COMB_LOST 3
#define RVX_COMB_RECOVERING 4

/* comb_cancel.h:161-199 CombEstimate: line k at offset_hz + k * spacing_hz,
 * k in [k_lo, k_hi]; a spur is one line at offset_hz (k_lo = k_hi = 0). */
typedef struct rvx_comb_estimate {
    double offset_hz;
    double spacing_hz;
    int32_t k_lo, k_hi;
    uint32_t lines_found;
    float median_height_db;
    double ripple_hz;          /* sidebands at multiples of it either side of each line */
    int32_t sideband_orders;
    uint32_t spur;
    double curvature_hz;       /* line k at offset + k spacing + k^2 curvature: 0 from a
                                  detection (straight grids); a pass's own once it has
                                  voted the curved model in */
} rvx_comb_estimate;

/* comb_cancel.h:318-331 CombStatus, over the bank's combs (spurs counted apart). */
typedef struct rvx_comb_status {
    uint32_t lock;                /* RVX_COMB_*: locked if any comb is */
    uint32_t lines_cancelled;
    float suppression_db;         /* mean over cancelled lines, from the projections */
    float spacing_drift_hz_per_s;
    rvx_comb_estimate comb;       /* the comb cancelling the most lines */
    uint32_t combs_active;
    uint32_t spurs_cancelled;
    float residual_rad;           /* its phase fit's weighted RMS residual */
    uint32_t reserved;
} rvx_comb_status;

/* comb_cancel.h:225-231 detect_combs: every comb in a linear power spectrum of
 * `bins` bins in natural order (DC at bins / 2, bin i centred on
 * (i - bins / 2) * rate / bins), at most RVX_COMB_MAX_COMBS, strongest first.
 * `halves` 1 holds each comb as its even and odd halves at twice the spacing
 * (detect_combs); 0 reports a comb whole unless its halves disagree
 * (detect_comb takes the first). `active`: combs already held, left out.
 * Refuses a spectrum of under 64 bins, a rate <= 0 or a bin that is not a
 * finite non-negative power. */
int32_t rvx_comb_detect(const float* power, uint64_t bins, int64_t rate,
                        const rvx_comb_estimate* active, uint32_t n_active, uint32_t halves,
                        rvx_comb_estimate* out, uint32_t max_out, uint32_t* n_out);

typedef struct rvx_comb_bank rvx_comb_bank;
/* comb_cancel.h:601-678 CombBank. Single-threaded: one caller drives it. */
int32_t rvx_comb_bank_create(int64_t rate, rvx_comb_bank** out);
void rvx_comb_bank_destroy(rvx_comb_bank* bank);
int32_t rvx_comb_bank_reset(rvx_comb_bank* bank);
/* New detections (CombBank::detected, comb_bank.cpp): each of at least
 * RVX_COMB_MIN_LINES lines that no running comb tracker holds starts in a
 * free one, strongest first; combs have first claim on the
 * RVX_COMB_MAX_LINES slots (what spurs hold counts as free), each taking its
 * lines if the rest still fit and an even share of what is free otherwise,
 * none under RVX_COMB_MIN_LINES; the newest spurs then give back what a
 * starting comb took, each held off. */
int32_t rvx_comb_bank_detected(rvx_comb_bank* bank, const rvx_comb_estimate* estimates,
                               uint32_t n);
/* Single carriers to cancel (CombBank::spurs, comb_cancel.h:614-621), each in
 * a one-line tracker. A spur missing from a later call is not let go (once
 * cancelled, the wideband detector no longer sees it). A spur tracker ends when
 * it loses its line, when it is withdrawn, or when it is dropped; a spur that
 * went is not started again within RVX_COMB_SAME_LINE_HZ of its detected
 * offset for RVX_COMB_SPUR_HOLD_OFF_S seconds of planned blocks. */
#define RVX_COMB_SPUR_HOLD_OFF_S 3.0
int32_t rvx_comb_bank_spurs(rvx_comb_bank* bank, const rvx_comb_estimate* estimates, uint32_t n);
/* CombBank::drop_spurs (comb_cancel.h:635): every spur goes, each held off. */
int32_t rvx_comb_bank_drop_spurs(rvx_comb_bank* bank);
/* CombBank::withdraw_spurs (comb_cancel.h:621-624): every spur detected within
 * RVX_COMB_SAME_LINE_HZ of one of these offsets goes (the detector stopped
 * calling the carrier there a spur), held off. */
int32_t rvx_comb_bank_withdraw_spurs(rvx_comb_bank* bank, const double* frequencies_hz,
                                     uint32_t n);
/* CombBank::SpurBand and drop_spurs_in (comb_cancel.h:626-634): a baseband
 * range an open receiver listens to, edges included; every spur detected in
 * one goes, held off. */
typedef struct rvx_spur_band {
    double low_hz;
    double high_hz;
} rvx_spur_band;
int32_t rvx_comb_bank_drop_spurs_in(rvx_comb_bank* bank, const rvx_spur_band* bands, uint32_t n);
/* The lines of block [start, start + count) with their phases at `start`:
 * every tracker's in turn; *subtract when any is cancelling (the others'
 * amplitudes are zero); *sub_blocks the count the trackers would like. */
int32_t rvx_comb_bank_plan(rvx_comb_bank* bank, uint64_t start, uint32_t count,
                           rvx_comb_line* lines, uint32_t max_lines, uint32_t* n_lines,
                           uint32_t* sub_blocks, int32_t* subtract);
/* A planned block's projections: lines x sub_blocks cf32, line-major, the
 * lines in plan order (rvx_comb_project's or rvx_engine_comb_projections'
 * layout); sub_blocks >= 2 for the frequency discriminator. */
int32_t rvx_comb_bank_observe(rvx_comb_bank* bank, uint64_t start, uint32_t count,
                              uint32_t sub_blocks, const float* projections, uint64_t n_floats);
int32_t rvx_comb_bank_status(const rvx_comb_bank* bank, rvx_comb_status* out);
/* Every running comb tracker's grid (CombBank::active_estimates). */
int32_t rvx_comb_bank_active(const rvx_comb_bank* bank, rvx_comb_estimate* out, uint32_t max_out,
                             uint32_t* n_out);

/* ---- Adaptive noise profile (host) -------------------------------------------
 * One receiver's learned noise spectrum taken out of its mono audio on the
 * completion thread, after the audio is read back (replaces
 * dsp::NoiseProfile, core/dsp/noise_profile.h:54-118, which
 * core/engine/graph.cpp:2587-2592 runs per receiver slot). A periodic
 * sqrt-Hann STFT at half overlap (1024 points at 32 kHz and up, 512 at 16 kHz,
 * else 256; the window pair reconstructs exactly at unit gain), a Wiener gain
 * per bin against the noise estimate floored at -strength dB, smoothed across
 * frames and across neighbouring bins. The estimate, in order: a profile
 * learned on request (one second of average power, kept until a reset), the
 * average of frames taken while the squelch is shut (auto-learn), or minimum
 * statistics over two seconds (auto-learn, no quiet frame yet). Pure host
 * arithmetic; one object per receiver, single-threaded. */
#define RVX_NOISE_PROFILE_MAX_STRENGTH_DB 20.0
typedef struct rvx_noise_profile rvx_noise_profile;
/* NoiseProfileSettings (noise_profile.h:44-50). A change of learn_count starts
 * a one-second learn; a change of reset_count drops the learned profile and
 * the automatic state; the first settings a profile sees are its baseline. */
typedef struct rvx_noise_profile_settings {
    int32_t enabled;
    int32_t auto_learn;
    double strength_db; /* clamped to 0 .. RVX_NOISE_PROFILE_MAX_STRENGTH_DB */
    uint32_t learn_count;
    uint32_t reset_count;
} rvx_noise_profile_settings;
int32_t rvx_noise_profile_create(rvx_noise_profile** out);
void rvx_noise_profile_destroy(rvx_noise_profile* profile);
/* NoiseProfile::process: `audio` (n floats, mono) in place, `latency` samples
 * late while enabled. `quiet`: the squelch was shut for this block. A change
 * of rate rebuilds the transform and drops every estimate; rate <= 0 leaves
 * the audio and the state untouched, as the reference does. */
int32_t rvx_noise_profile_process(rvx_noise_profile* profile, float* audio, uint64_t n,
                                  double rate, int32_t quiet,
                                  const rvx_noise_profile_settings* settings);
/* NoiseProfile::retune: the automatic estimate is dropped, a learned one kept. */
int32_t rvx_noise_profile_retune(rvx_noise_profile* profile);
/* learning(), learned(), latency() and the number of bins of profile() (0
 * while there is no estimate; size / 2 + 1 otherwise). */
int32_t rvx_noise_profile_state(const rvx_noise_profile* profile, int32_t* learning,
                                int32_t* learned, uint32_t* latency, uint32_t* bins);
/* NoiseProfile::profile: the estimate in use, power per bin (up to max_out). */
int32_t rvx_noise_profile_estimate(const rvx_noise_profile* profile, double* out,
                                   uint32_t max_out, uint32_t* n_out);

/* ---- Per-receiver impulse blanker --------------------------------------------
 * core/dsp/noise_reference.h:59-190: impulses blanked on a receiver's coarse
 * channel at the channel rate, before its fine stage narrows them into
 * ringing bursts. The fine stage then reads the blanked ring (chan_base 0,
 * out_mask as its channel mask). */
#define RVX_BLANK_DETECT 0
#define RVX_BLANK_APPLY 1

/* noise_reference.h:117 BlankerConfig */
typedef struct rvx_blanker_config {
    uint32_t pass;
    uint32_t window;     /* reference window, 16..256 samples */
    uint32_t guard;      /* samples between the window and the tested sample */
    uint32_t hang, lead; /* blanked after the last / before the first flag */
} rvx_blanker_config;

/* noise_reference.h:152 BlankerParams (seven words) */
typedef struct rvx_blanker_params {
    uint32_t chan_base, chan_mask, chan_first; /* the channel ring, as vrx_fine reads it */
    uint32_t out_mask, out_first;              /* the receiver's flag / blanked rings */
    uint32_t count;
    float threshold;                           /* T as a power ratio */
} rvx_blanker_params;

/* noise_reference.cpp:112 design_blanker: 0.5 ms window clamped to [16, 256],
 * hang = lead = 2, guard = lead + 2. */
int32_t rvx_design_blanker(int64_t channel_rate, rvx_blanker_config* out);
/* :123 blanker_threshold: 10^(dB/10) rounded to float once. */
float rvx_blanker_threshold(double threshold_db);
/* :127 validate(BlankerConfig, BlankerParams). */
int32_t rvx_blanker_check(const rvx_blanker_config* config, const rvx_blanker_params* params);
/* core/shaders/noise_blank.comp; twins noise_reference.cpp:158 (detect: flags,
 * out_mask + 1 floats) and :197 (apply: blanked, out_mask + 1 cf32). */
int32_t rvx_blank_detect(rvx_ctx* ctx, const rvx_blanker_config* config,
                         const rvx_blanker_params* params, const void* channel_ring,
                         uint64_t ring_samples, float* flags, uint64_t n_flags);
int32_t rvx_blank_apply(rvx_ctx* ctx, const rvx_blanker_config* config,
                        const rvx_blanker_params* params, const void* channel_ring,
                        uint64_t ring_samples, const float* flags, uint64_t n_flags,
                        void* blanked, uint64_t blanked_samples);

/* ---- The line kernel: manual notch cascade and adaptive line enhancer ----------
 * core/dsp/noise_reference.h:190-420: per receiver, over its mono audio, a
 * cascade of up to four notch sections, then an NLMS predictor (taps split
 * over `lanes` partial sums, summed in lane order) whose error is the output,
 * optionally with a guard section on the subtraction (CW). */
#define RVX_LINE_NOTCH 1u
#define RVX_LINE_ALE 2u
#define RVX_LINE_GUARD 4u
#define RVX_LINE_NOTCH_COUNT_SHIFT 8
#define RVX_LINE_NOTCH_COUNT_MASK (3u << RVX_LINE_NOTCH_COUNT_SHIFT)
#define RVX_MAX_NOTCHES 4
#define RVX_MAX_LINE_STRIDE 8
/* state layout, floats (noise_reference.h:318-333) */
#define RVX_LINE_STATE_POSITION 4
#define RVX_LINE_STATE_GUARD 8
#define RVX_LINE_STATE_MORE 12
#define RVX_LINE_STATE_WEIGHTS 24

/* noise_reference.h:277 LineConfig. Here lanes <= 64 (one wavefront) and
 * taps + history <= 16128 floats (the workgroup's LDS copy). */
typedef struct rvx_line_config {
    uint32_t lanes, taps, history;
} rvx_line_config;

typedef struct rvx_line_section {
    float b0, b1, b2, a1, a2; /* y = b0 x + b1 x1 + b2 x2 - a1 y1 - a2 y2 */
} rvx_line_section;

/* noise_reference.h:346 LineParams (thirty-two words) */
typedef struct rvx_line_params {
    uint32_t count, flags, stride, delay;
    float b0, b1, b2, a1, a2;
    float mu, leak, eps;
    float gb0, gb1, gb2, ga1, ga2;
    rvx_line_section more[RVX_MAX_NOTCHES - 1];
} rvx_line_params;

/* One receiver of a batched launch: its parameters, its audio (count floats,
 * in place) and its state (rvx_line_state_size floats). 144 bytes; the
 * array lives in device memory. */
typedef struct rvx_noise_line_item {
    rvx_line_params params;
    float* audio;
    float* state;
} rvx_noise_line_item;

/* noise_reference.h:395 NotchDesign, as five doubles b0 b1 b2 a1 a2. */
int32_t rvx_design_notch(double audio_hz, double depth_db, double width_hz, int64_t audio_rate,
                         double* out5);
/* :410 notch_section_db: the power response of the float-rounded section. */
double rvx_notch_section_db(const double* design5, double audio_hz, int64_t audio_rate);
uint64_t rvx_line_state_size(const rvx_line_config* config);
/* :390 validate(LineConfig, LineParams), plus this kernel's lane and LDS limits. */
int32_t rvx_line_check(const rvx_line_config* config, const rvx_line_params* params);
/* core/shaders/noise_line.comp; twin noise_reference.cpp:266 reference_line.
 * audio: count floats processed in place; state: rvx_line_state_size floats. */
int32_t rvx_noise_line(rvx_ctx* ctx, const rvx_line_config* config, const rvx_line_params* params,
                       float* audio, uint64_t n_audio, float* state, uint64_t n_state);
/* Many receivers' line kernels in one launch, one workgroup each. The items
 * are in device memory and the launch trusts them: the caller checks each
 * with rvx_noise_line_item_check (its audio and state lengths) before
 * uploading it, as for rvx_vrx_carrier_batch. */
int32_t rvx_noise_line_item_check(const rvx_line_config* config, const rvx_noise_line_item* item,
                                  uint64_t n_audio, uint64_t n_state);
int32_t rvx_noise_line_batch(rvx_ctx* ctx, const rvx_line_config* config,
                             const rvx_noise_line_item* items, uint32_t n_items);

/* ---- Spectral noise reduction -------------------------------------------------
 * noise_reference.h:420-520: 50% overlap-add with a square-root Hann window,
 * a direct DFT, power spectral subtraction with a floor, and Doblinger
 * minimum tracking of the noise floor. */
#define RVX_MIN_SPECTRAL_FRAME 64
#define RVX_MAX_SPECTRAL_FRAME 512

/* noise_reference.h:495 SpectralParams (ten words) */
typedef struct rvx_spectral_params {
    uint32_t count;
    float smoothing, smoothing_complement;
    float gamma, beta, rise;
    float bias, alpha, floor_gain, eps;
} rvx_spectral_params;

typedef struct rvx_noise_spectral_item {
    rvx_spectral_params params;
    uint32_t reserved[2];
    float* audio;
    float* state;
} rvx_noise_spectral_item; /* 64 bytes */

/* :535 spectral_config_for: the frame for an audio rate. */
uint32_t rvx_spectral_frame_for(int64_t audio_rate);
uint64_t rvx_spectral_state_size(uint32_t frame);
/* :540 design_spectral_tables: 4 * frame floats. */
int32_t rvx_design_spectral_tables(uint32_t frame, float* out, uint64_t n_floats);
/* :543 spectral_params(strength), count left at zero. */
int32_t rvx_spectral_params_for(double strength, rvx_spectral_params* out);
int32_t rvx_spectral_check(uint32_t frame, const rvx_spectral_params* params);
/* core/shaders/noise_spectral.comp; twin noise_reference.cpp:568
 * reference_spectral. tables: the designed 4 * frame floats on the device. */
int32_t rvx_noise_spectral(rvx_ctx* ctx, uint32_t frame, const rvx_spectral_params* params,
                           const float* tables, uint64_t n_tables, float* audio, uint64_t n_audio,
                           float* state, uint64_t n_state);
/* Many receivers in one launch, one workgroup each, sharing one frame and
 * its tables; items in device memory, each checked by the caller with
 * rvx_noise_spectral_item_check before uploading. */
int32_t rvx_noise_spectral_item_check(uint32_t frame, const rvx_noise_spectral_item* item,
                                      uint64_t n_audio, uint64_t n_state);
int32_t rvx_noise_spectral_batch(rvx_ctx* ctx, uint32_t frame, const float* tables,
                                 uint64_t n_tables, const rvx_noise_spectral_item* items,
                                 uint32_t n_items);

/* ---- Plain detector ----------------------------------------------------------
 * The decision of core/detect/detector.cpp on an averaged linear spectrum:
 * estimate_noise_floor (:794-876), then find_candidates (:879-1120) through
 * the accepted, grown bands. Noise floor, prefix sums, the summed-bin ladder
 * and the peak search run on the GPU; the strongest-first greedy growth runs
 * on the host over the peaks. DetectorConfig fields (detector.h:579-909). */
typedef struct rvx_detect_config {
    uint32_t bins;
    uint32_t noise_knots;         /* 32 */
    uint32_t noise_window_knots;  /* 8 */
    uint32_t max_sum_bins;        /* 0: bit_floor(bins / 8) */
    uint32_t max_peaks;           /* 0: bins */
    uint32_t split_gap_bins;      /* 8 */
    double bin_width_hz;
    double detection_threshold_db;  /* 6 */
    double noise_percentile;        /* 0.10 */
    double noise_percentile_high;   /* 0.40 */
    double noise_excision_sigma;    /* 4 */
    double edge_floor_fraction;     /* 0 */
    double edge_floor_sigma;        /* 3 */
} rvx_detect_config;

typedef struct rvx_detect_band {
    uint32_t start;      /* first bin */
    uint32_t width;      /* bins */
    double deflection;   /* of the seed window that won */
} rvx_detect_band;

typedef struct rvx_detector rvx_detector;

int32_t rvx_detector_create(rvx_ctx* ctx, const rvx_detect_config* config, rvx_detector** out);
void rvx_detector_destroy(rvx_detector* detector);
/* average: device, bins doubles, an averaged linear power spectrum (values
 * >= 0: the percentile select orders the doubles' bit patterns). allow_prefix
 * (device, bins + 1 int32, optional): the running count of bins a proposal
 * allows; a ladder window is searched only where it holds an allowed bin
 * (null: everywhere). Waits for the result; the three running sums and the
 * greedy growth run on the calling thread. *n_peaks: peaks the ladder found
 * before the budget. */
int32_t rvx_detect(rvx_detector* detector, const double* average, const int32_t* allow_prefix,
                   rvx_detect_band* bands, uint32_t max_bands, uint32_t* n_bands,
                   uint32_t* n_peaks);
/* The last decision's per-bin floor and spread (device, bins doubles each). */
const double* rvx_detector_floor(rvx_detector* detector);
const double* rvx_detector_sigma(rvx_detector* detector);

/* ---- Detection events -------------------------------------------------------
 * core/detect/detector.cpp after the decision (host, as in the reference):
 * accepted bands -> candidates (:1168-1334: occupied band, interior-gap
 * split, the SNR test in 2500 Hz, centre and width in Hz from the bin edges),
 * then per decision reject_residual (:1336) and update_tracks (:1585). A
 * track's first_seen / last_detected / last_seen are the sample indices of
 * the decisions that found it first and last: the event's boundaries. Each
 * decision ends with the emitters' key-up count, the spur rule and the
 * hybrid sideband mark (:1885-1933, mark_spurs :687-780, mark_hd_hybrids
 * :1943-1976). The probes stay outside: the finding the
 * rule reads, a probe's family, is the caller's to record
 * (rvx_tracker_record_probe), and remembers the track's emitter. */
typedef struct rvx_event_config {
    double bin_width_hz;             /* the frame's bin spacing */
    double bin_zero_hz;              /* bin 0's centre, relative to the source centre */
    int64_t source_center_hz;        /* added to every candidate centre */
    int64_t source_rate;             /* samples per second (the hold is in samples) */
    double detection_threshold_db;   /* 6 (detector.h:600) */
    double occupied_power_fraction;  /* 0.99 (:857) */
    double split_gap_db;             /* 1 (:910) */
    uint32_t split_gap_bins;         /* 8 (:909) */
    uint32_t residual_decisions;     /* 4 (:1184); 0: never withheld */
    double average_seconds;          /* 1 (:643) */
    double residual_rate_tolerance;  /* 0.25 (:1183) */
    double association_overlap;      /* 0.3 (:1188) */
    double confidence_rise;          /* 0.35 (:1194) */
    double confidence_half_life_seconds; /* 1 (:1195) */
    double drop_confidence;          /* 0.05 (:1196) */
    double centre_smoothing;         /* 0.4 (:1200) */
    double bootstrap_hold_seconds;   /* 3 (:955) */
    uint32_t birth_hits;             /* 3 (:918) */
    uint32_t max_tracks;             /* 512 (:938) */
    uint32_t grid_channels;          /* the channelizer's M for a track's channel; 0: none (:591) */
    uint32_t reserved;
    double channel_hysteresis;       /* 0.15 of a channel past the midpoint (:1205) */
} rvx_event_config;

/* detector.h:299-360 Candidate (the fields built here) */
typedef struct rvx_candidate {
    uint32_t first_bin, last_bin;
    int64_t center_hz, bandwidth_hz;
    double snr_2500_db, noise_floor_dbfs, peak_dbfs;
    /* characterise::margin_confidence(snr_2500_db, detection_threshold_db):
     * 0 under the threshold, else 1 - exp(-(snr - threshold) / 6 dB) / 2 */
    double margin_confidence;
    /* BandShape (core/detect/shape.h:53-272): measurements of the band's
     * excess (each bin's average over floor, negatives taken as 0), no
     * verdicts: peak over mean; the strongest three adjacent bins' share
     * (1 for bands of three bins or fewer); the lower half's share of the two
     * halves; skirt bins (outside the band, within twice its width and its own
     * region between the neighbouring decision bands, over a tenth of its mean)
     * per band bin, and how many bins the walk had room for. */
    double shape_peak_to_mean, shape_concentration, shape_lower_fraction, shape_skirt_fraction;
    uint32_t shape_skirt_bins_available, shape_measured;
} rvx_candidate;

#define RVX_TRACK_PENDING 0u
#define RVX_TRACK_LIVE 1u
#define RVX_TRACK_HELD 2u
#define RVX_TRACK_MERGED 3u

/* detector.h:362-575 Track (the fields built here) */
typedef struct rvx_track {
    uint64_t id;
    uint32_t state;              /* RVX_TRACK_* */
    uint32_t hits, misses, held_spells;
    int64_t center_hz, bandwidth_hz;
    double snr_2500_db, confidence;
    uint64_t first_seen, last_seen, last_detected;  /* sample indices */
    uint64_t merged_into;        /* parent id while merged */
    uint32_t first_bin, last_bin;  /* the last candidate's bins */
    double margin_confidence;    /* of the smoothed snr_2500_db, as the candidate's */
    /* The band's level in each spectrum frame (rvx_tracker_measure_levels): a
     * running mean and sum of squared deviations (Welford) of 10 log10 of the
     * summed linear power over level_first_bin .. level_last_bin, restarted
     * when the track's bins change; sigma = sqrt(level_m2 / (frames - 1)). */
    double level_mean_db, level_m2;
    uint32_t level_first_bin, level_last_bin, level_frames, level_bins_known;
    /* The grid channel nearest the centre (with grid_channels set): moved
     * only when the centre is more than 0.5 + channel_hysteresis channels
     * from the one held, so a track on a boundary does not flip. */
    int64_t channel_index;       /* unwrapped: round((centre - source centre) M / rate) */
    uint32_t channel;            /* channel_index reduced into [0, M) */
    uint32_t channel_valid;
    /* the band shape of the last candidate taken (not smoothed: it stays what
     * the last detection measured while the track holds) */
    double shape_peak_to_mean, shape_concentration, shape_lower_fraction, shape_skirt_fraction;
    uint32_t shape_skirt_bins_available, shape_measured;
    /* What a probe named the track (RVX_FAMILY_*, with double_sideband for a
     * carrier), as the caller recorded it, and its emitter's key-ups. */
    uint32_t classification, double_sideband, keyups;
    /* An unkeyed carrier the engine may cancel (detector.h:521-535): spur by
     * the rule or held while being cancelled; lo_spur when it stands at the
     * baseband offset of a spur seen under another tuning. */
    uint32_t spur, lo_spur;
    uint32_t reserved_spur;
    /* The emitter it belongs to (refine::EmitterMemory, refine.h:261-295):
     * remembered at the track's first probe, joined by a track born within
     * max(1 kHz, half the narrower width) of it; 0 for none. Coming live after
     * more than a second of the emitter's silence is one of its key-ups (not
     * while the carrier there is being cancelled); an emitter quiet for
     * 1800 s is forgotten, and at most 512 are kept, the longest quiet going. */
    uint64_t emitter;
    /* Hybrid digital sidebands beside wideband FM (detector.h:537-560,
     * mark_hd_hybrids): measured at each decision that detects a live track
     * at least 50 kHz wide; hd_hybrid after 2 passing readings in a row,
     * cleared after 3 failing; the sidebands' level against the analogue
     * signal (dBc), the first passing reading taken whole, then smoothed 0.3. */
    double hd_lower_dbc, hd_upper_dbc;
    uint32_t hd_passes, hd_fails, hd_hybrid, reserved_hd;
} rvx_track;

/* HdSidebandReading (core/detect/hd_hybrid.h): with `measured`, the levels of
 * the digital sidebands (129-198 kHz from the analogue signal's centroid, found
 * within 100 kHz of the given centre) against the analogue excess (dBc), their
 * floor and the 105-120 kHz gap (dB), and each side's inner-over-outer tilt;
 * `hybrid` when each side is 3 dB over its floor and 1.5 dB over its gap,
 * tilted by at most 3 dB, and the sides within 6 dB of each other. */
typedef struct rvx_hd_reading {
    uint32_t measured, hybrid;
    double centre_hz, lower_dbc, upper_dbc, lower_over_floor_db, upper_over_floor_db;
    double lower_over_gap_db, upper_over_gap_db, lower_tilt_db, upper_tilt_db;
} rvx_hd_reading;

#define RVX_FAMILY_UNKNOWN 0u      /* no answer yet */
#define RVX_FAMILY_UNMODULATED 1u  /* a bare carrier */
#define RVX_FAMILY_OTHER 2u        /* any other family a probe names */

/* SpurBook's entry (detector.h:1262-1292): a carrier that passed the spur
 * rule, by its baseband offset and the source centre it was seen under. */
typedef struct rvx_spur_sighting {
    int64_t baseband_hz, center_hz;
    uint64_t last_seen;          /* sample index */
} rvx_spur_sighting;
#define RVX_SPUR_BOOK_CAPACITY 256u

typedef struct rvx_tracker_stats {
    uint64_t decisions, candidates, candidates_residual, residual_parts_degenerate;
    uint64_t tracks_born, tracks_dropped, births_refused, merges, splits;
} rvx_tracker_stats;

void rvx_event_config_defaults(rvx_event_config* config);
/* average, floor: host, bins doubles (the decision's input and its floor);
 * bands: the decision's accepted bands (rvx_detect), ascending and disjoint. */
int32_t rvx_candidates(const rvx_event_config* config, const double* average, const double* floor,
                       uint32_t bins, const rvx_detect_band* bands, uint32_t n_bands,
                       rvx_candidate* out, uint32_t max_out, uint32_t* n_out);
typedef struct rvx_tracker rvx_tracker;
int32_t rvx_tracker_create(const rvx_event_config* config, uint32_t bins, rvx_tracker** out);
void rvx_tracker_destroy(rvx_tracker* tracker);
/* One decision at sample index `now`, elapsed_seconds after the previous one:
 * residual rejection over the candidates, then the tracks' update.
 * *n_kept (may be NULL): candidates left after residual rejection. */
/* Detector::measure_levels (detector.cpp:629-646): one spectrum frame (dB,
 * `bins` floats) into every live track's level statistics. */
int32_t rvx_tracker_measure_levels(rvx_tracker* tracker, const float* power_db, uint32_t bins);
int32_t rvx_tracker_update(rvx_tracker* tracker, const double* average, const double* floor,
                           const rvx_candidate* candidates, uint32_t n, uint64_t now,
                           double elapsed_seconds, uint32_t* n_kept);
int32_t rvx_tracker_tracks(const rvx_tracker* tracker, rvx_track* out, uint32_t max_out,
                           uint32_t* n);
int32_t rvx_tracker_stats_get(const rvx_tracker* tracker, rvx_tracker_stats* out);
/* measure_hd_sidebands (hd_hybrid.cpp): power and floor are a decision's
 * average and floor (linear, `bins` doubles), bin_zero_hz bin 0's centre and
 * centre_hz the track's, both relative to the source centre. */
int32_t rvx_measure_hd_sidebands(const double* power, const double* floor, uint32_t bins,
                                 double bin_zero_hz, double bin_width_hz, double centre_hz,
                                 rvx_hd_reading* out);
/* Detector::record_probe (detector.cpp:465-510), the part the spur rule
 * reads: a probe's family for the track `id`, one the caller lets drive
 * detection; the track's first probe remembers its emitter (at the last
 * decision's sample index). RVX_EINVAL when no track has that id. */
int32_t rvx_tracker_record_probe(rvx_tracker* tracker, uint64_t id, uint32_t family,
                                 uint32_t double_sideband);
/* Detector::withdrawn_spurs (detector.h:1400-1403): absolute centres (Hz) of
 * the spurs the latest decision took back, for the engine to stop cancelling. */
int32_t rvx_tracker_withdrawn_spurs(const rvx_tracker* tracker, int64_t* out, uint32_t max_out,
                                    uint32_t* n);
/* Detector::spur_book / adopt_spur_book (detector.h:1397-1398): the sightings,
 * carried by the caller from the tracker before a retune to the one after.
 * Sightings unseen for 300 s of source time are forgotten at each decision. */
int32_t rvx_tracker_spur_book(const rvx_tracker* tracker, rvx_spur_sighting* out,
                              uint32_t max_out, uint32_t* n);
int32_t rvx_tracker_adopt_spur_book(rvx_tracker* tracker, const rvx_spur_sighting* book,
                                    uint32_t n);

/* ---- Learned proposals ------------------------------------------------------
 * A five-layer 3x3 convolutional network (1 -> 32 -> 32 -> 32 -> 32 -> 1,
 * ReLU) over log-power spectrogram tiles, per-pixel logits of "a signal the
 * plain detector would report is here". bf16 matrix-core convolutions with
 * fp32 accumulation; trained on the GPU with Adam. No counterpart in the
 * reference: it proposes, the plain detector (rvx_detect with an allow mask)
 * confirms. Tiles are device bf16 [batch][height][width], row-major. */
typedef struct rvx_net rvx_net;
int32_t rvx_net_create(rvx_ctx* ctx, uint32_t height, uint32_t width, uint32_t max_batch,
                       uint64_t seed, rvx_net** out);
void rvx_net_destroy(rvx_net* net);
uint32_t rvx_net_param_count(const rvx_net* net);
/* Master fp32 parameters, per layer W[cout][3][3][cin] then b[cout]; write
 * nonzero copies host -> device. */
int32_t rvx_net_params(rvx_net* net, float* host, uint64_t count, int32_t write);
/* The last train step's gradients, same layout (host copy; waits). */
int32_t rvx_net_grads(rvx_net* net, float* host, uint64_t count);
/* logits: device fp32 [batch][height][width]. */
int32_t rvx_net_forward(rvx_net* net, const uint16_t* tiles, uint32_t batch, float* logits);
/* labels: device fp32 0/1. lr 0 computes gradients without updating.
 * *loss (optional): the mean pos_weight-weighted binary cross-entropy; waits. */
int32_t rvx_net_train_step(rvx_net* net, const uint16_t* tiles, const float* labels,
                           uint32_t batch, float lr, float pos_weight, float* loss);
/* fp32 -> bf16, round to nearest even (device arrays). */
int32_t rvx_to_bf16(rvx_ctx* ctx, const float* in, uint64_t count, uint16_t* out);

/* ---- Streaming engine ------------------------------------------------------
 * The per-block orchestration of Graph::on_block (core/engine/graph.cpp:
 * 4770-5480) and the receivers' stream cursors of DemodStage
 * (core/engine/vrx_stage.cpp:1119-1480), over the fused chain and the batched
 * receivers. Samples arrive in pushes of any size; each push channelizes every
 * block whose history is complete (whole spectrum rows when the averaged
 * spectrum is on), then runs every receiver over the audio its channel
 * samples now allow. Every output is a pure function of its absolute index,
 * so results do not depend on how the stream is cut into pushes. */
typedef struct rvx_engine rvx_engine;

typedef struct rvx_engine_config {
    int32_t format;            /* RVX_FORMAT_CU8 or RVX_FORMAT_CS16 */
    uint32_t channels;         /* M, a power of two <= 4096 */
    uint32_t taps_per_branch;
    uint32_t decimation;       /* D, dividing M */
    uint32_t max_push;         /* largest push, samples */
    uint32_t ring_blocks;      /* channel ring depth, a power of two */
    uint32_t frame_blocks;     /* averaged spectrum frame; 0: no averaged spectrum */
    uint32_t row_frames;       /* frames per spectrum row */
    float alpha;
    uint32_t calibration;      /* RVX_CALIBRATE_DC | RVX_CALIBRATE_IQ: front-end correction */
    uint64_t start_index;      /* absolute index of the first sample pushed */
    int64_t sample_rate;       /* Hz; the estimator's time constants (calibration only) */
    uint32_t ring_pitch;       /* channel ring row stride in samples: 0 is ring_blocks,
                                * else >= ring_blocks and a multiple of 4
                                * (rvx_chain_desc.ring_pitch) */
    uint32_t reserved;
} rvx_engine_config;

/* Engine::set_calibration's two halves (core/engine/engine.h, the front-end
 * correction stage): moments of every push are read three pushes later into
 * the estimator, and each push's samples are corrected, inside the
 * channelizer, with the estimate held when they arrived. */
#define RVX_CALIBRATE_DC 1u
#define RVX_CALIBRATE_IQ 2u

typedef struct rvx_engine_stats {
    uint64_t samples_in;       /* samples pushed */
    uint64_t next_block;       /* first block not yet channelized */
    uint64_t blocks;           /* blocks channelized */
    uint64_t spectrum_rows;    /* averaged-spectrum rows folded into the state */
    uint64_t audio_lost;       /* receiver audio outputs skipped because their
                                * channel samples had left the ring */
    uint64_t audio_overrun;    /* audio outputs computed but overwritten in a
                                * receiver's audio ring before rvx_engine_read_audio
                                * copied them (the reader fell more than the ring
                                * behind) */
    uint64_t grid_blocks;      /* blocks channelized on the added grids */
} rvx_engine_stats;

/* The front-end estimate and the correction the newest push's samples took
 * (the identity before the first estimate or with calibration off). */
int32_t rvx_engine_front_end(rvx_engine* engine, rvx_front_end_estimate* estimate,
                             rvx_iq_correct_params* applied);

/* prototype: M * taps_per_branch floats (host), e.g. from rvx_design_prototype. */
int32_t rvx_engine_create(rvx_ctx* ctx, const rvx_engine_config* config, const float* prototype,
                          uint64_t n_prototype, rvx_engine** out);
void rvx_engine_destroy(rvx_engine* engine);
/* A receiver planned on this engine's grid (rvx_plan_vrx); tables are host
 * arrays (fine taps as cf32 pairs). audio_capacity: audio outputs its ring
 * holds until read. tile: 0 = the largest that fits. */
int32_t rvx_engine_add_vrx(rvx_engine* engine, const rvx_vrx_plan* plan, const float* fine_taps,
                           uint64_t n_fine_floats, const float* weights, uint64_t n_weights,
                           uint32_t audio_capacity, uint32_t tile, uint32_t* id);

/* A second grid on the same input stream, for receivers wider than a channel
 * of the engine's own. The reference runs one grid and coarsens it until the
 * widest receiver fits (core/engine/engine.h:160-200, EngineConfig::channels
 * and carry_receiver_hz, engine::channel_count_for); this engine keeps its
 * fine grid for the spectrum and the narrow receivers and adds the coarse one
 * beside it. The grid stores only the channels its receivers sit on: a
 * channel takes one of max_rows ring rows when its first receiver arrives
 * (the row is cleared and its stream starts at the grid's next block) and
 * gives it back when its last receiver leaves. Each push channelizes the
 * grid's complete blocks (the same chain, front-end correction included)
 * after the engine's own. *grid: its id for rvx_engine_add_vrx_on (1, 2, ...;
 * 0 names the engine's own grid). Its history must fit the engine's input
 * window: M * taps_per_branch - 1 + decimation + max_push samples. */
typedef struct rvx_engine_grid {
    uint32_t channels;         /* M, a power of two <= 4096 */
    uint32_t taps_per_branch;
    uint32_t decimation;       /* D, dividing M */
    uint32_t ring_blocks;      /* channel ring depth, a power of two */
    uint32_t max_rows;         /* channels it stores at once, 1..M */
    uint32_t ring_pitch;       /* row stride in samples: 0 is ring_blocks (as
                                * rvx_engine_config.ring_pitch) */
} rvx_engine_grid;
int32_t rvx_engine_add_grid(rvx_engine* engine, const rvx_engine_grid* grid,
                            const float* prototype, uint64_t n_prototype, uint32_t* grid_id);
/* rvx_engine_add_vrx on grid `grid` (the plan made with that grid's
 * parameters: rvx_plan_vrx with its M, D and taps). */
int32_t rvx_engine_add_vrx_on(rvx_engine* engine, uint32_t grid, const rvx_vrx_plan* plan,
                              const float* fine_taps, uint64_t n_fine_floats,
                              const float* weights, uint64_t n_weights, uint32_t audio_capacity,
                              uint32_t tile, uint32_t* id);
/* Impulse excision before the channelizer, as the reference records it
 * between the converted ring and pfb_branch (core/engine/graph.cpp:1490,
 * record_impulse_excision): each push is converted (and corrected, with
 * calibration on, under the correction it arrived with) into a cf32 sample
 * ring, the cursor's four passes clean it in place, and the cleaned samples
 * feed every grid's chain (RVX_FORMAT_CF32). Blocks run up to the excision's
 * cleaned end, which trails the newest sample by max(max_width + lead, one
 * segment). Set before the first push; config from rvx_design_excision,
 * threshold from rvx_excision_threshold. */
int32_t rvx_engine_set_excision(rvx_engine* engine, const rvx_excision_config* config,
                                float threshold);
/* The excision's RVX_EXCISE_COUNTERS words (samples excised, events, runs
 * spared, unused); waits. */
int32_t rvx_engine_excision_counters(rvx_engine* engine, uint32_t* counters);
/* Comb cancellation before the channelizer, in the reference's record order
 * (graph.cpp:1715, record_comb_cancel: after conversion and correction,
 * before excision). Each push projects the samples onto every line (the
 * raw, pre-subtraction samples, one projection per line over the push) and,
 * with subtract, removes the lines' a * tone from them. lines[k].phase0 is
 * line k's 0.32 phase at absolute sample phase_index; each push advances it
 * by step per sample, so results do not depend on how the stream is cut.
 * The lines come from the caller's comb tracker (the reference's comb_bank);
 * the first call, before the first push, opens the stage (the conditioned
 * path, as with excision), later calls replace the lines. */
int32_t rvx_engine_set_comb(rvx_engine* engine, const rvx_comb_line* lines, uint32_t n_lines,
                            uint64_t phase_index, int32_t subtract);
/* The last push's projections over samples [*first_sample, *first_sample +
 * *samples): *lines x *sub_blocks cf32 values, line-major, sub-block j over
 * the j-th equal part of the push (sub_blocks: the largest power of two up
 * to RVX_COMB_MAX_SUB_BLOCKS dividing the push); waits. */
int32_t rvx_engine_comb_projections(rvx_engine* engine, float* out, uint64_t n_floats,
                                    uint64_t* first_sample, uint64_t* samples, uint32_t* lines,
                                    uint32_t* sub_blocks);
/* Grid `grid`'s channel ring (device, row-major: row r's ring_blocks cf32
 * start at r * ring_pitch; max_rows rows; grid 0: the engine's channel ring). */
const void* rvx_engine_grid_ring(rvx_engine* engine, uint32_t grid);
/* The ring row holding `channel` on grid `grid` and the first block written
 * into it (grid 0: row = channel, the engine's first block). */
int32_t rvx_engine_grid_row(rvx_engine* engine, uint32_t grid, uint32_t channel, uint32_t* row,
                            uint64_t* first_block);
int32_t rvx_engine_remove_vrx(rvx_engine* engine, uint32_t id);
/* A mono receiver's audio noise stages (the reference's per-receiver stage,
 * noise_reference.h:190-520, recorded after its demodulator): the line
 * kernel with line_config / line (count ignored) and spectral reduction at
 * spectral_frame with spectral (count ignored), each NULL for off. Every push
 * then runs them in place on the audio it produced, all receivers in batched
 * launches. A call restarts both stages' state (it waits for the stream). */
/* The receiver's display tap (vrx_reference.h:1639-1768, recorded beside its
 * fine stage): every push extends the display stream (the fine kernel with
 * the 256-tap display table) over the samples the fine stage reads, into a
 * display ring of display_ring samples (a power of two, 0 turns it off).
 * rvx_engine_passband transforms the newest `transform` display samples into
 * the pane: transform / 2 dB values, ascending from the display's DC minus a
 * quarter of the display rate (graph.cpp:314-400); *end_output is one past
 * the display output the window ends at. RVX_EAGAIN until enough samples. */
int32_t rvx_engine_set_display(rvx_engine* engine, uint32_t id, uint32_t display_ring,
                               uint32_t transform);
int32_t rvx_engine_passband(rvx_engine* engine, uint32_t id, float* out, uint64_t n_out,
                            uint64_t* end_output, int64_t* display_rate);
/* The receiver's impulse blanker on its coarse channel (noise_reference.h:
 * 59-190, recorded before the fine stage): config (NULL for off; pass is
 * ignored) and the threshold as a power ratio (rvx_blanker_threshold). Each
 * push detects through the newest channel sample and applies through `lead`
 * before it; the receiver's fine stage reads the blanked samples, so its audio
 * runs lead channel samples behind. Enabling restarts the receiver's audio
 * cursor from the blanked stream. */
int32_t rvx_engine_set_blanker(rvx_engine* engine, uint32_t id, const rvx_blanker_config* config,
                               float threshold);
int32_t rvx_engine_set_noise(rvx_engine* engine, uint32_t id, const rvx_line_config* line_config,
                             const rvx_line_params* line, uint32_t spectral_frame,
                             const rvx_spectral_params* spectral);
/* codes: `samples` native samples, host memory (on_device = 0) or device. */
int32_t rvx_engine_push(rvx_engine* engine, const void* codes, uint64_t samples,
                        int32_t on_device);
/* Copies the receiver's unread audio (oldest first, at most max_values floats)
 * to host memory and waits for it; *first_audio is the index of the first
 * output copied, *values the floats copied. It moves the same per-receiver
 * cursors rvx_engine_push moves: calls of the two on one engine must be
 * serialised by the caller (one lock around both), like every call on a
 * context. */
int32_t rvx_engine_read_audio(rvx_engine* engine, uint32_t id, float* out, uint64_t max_values,
                              uint64_t* first_audio, uint64_t* values);
/* The averaged spectrum state (M floats, host) after the last row; waits. */
int32_t rvx_engine_spectrum(rvx_engine* engine, float* out, uint64_t n);
/* The channel ring (device, channel-major: channel k's ring_blocks cf32 start
 * at k * ring_pitch). */
const void* rvx_engine_channel_ring(rvx_engine* engine);
int32_t rvx_engine_stats_get(rvx_engine* engine, rvx_engine_stats* out);

/* The same 4096-channel chain with its transform as two 64 x 64 complex
 * DFT-GEMMs on the matrix cores (configuration 4 of the precision study;
 * no counterpart in the reference). The branch filter stays fp32; the DFT
 * operands are split bf16 (hi + lo, three products) or fp16, fp32 accumulation.
 * Same descriptor, buffers and outputs as rvx_chain (cs16, M = 4096, no
 * channel_rows); not bit-exact: its error is measured against FP64. */
#define RVX_PREC_BF16X3 1
#define RVX_PREC_FP16 2
int32_t rvx_chain_mfma(rvx_ctx* ctx, const rvx_chain_desc* desc, int32_t precision,
                       const void* input, const float* taps, uint64_t n_taps,
                       const void* twiddles, uint64_t n_twiddles, void* channel_ring,
                       uint64_t ring_samples, float* partial, uint64_t n_partial);

/* Display spectrum in FP8 e4m3: out[i] = e4m3(10 log10(max(power[i], 1e-20)) / 4),
 * round to nearest even, saturating (dB / 4 spans -50 .. 112 in e4m3's +-448). */
int32_t rvx_spectrum_fp8(rvx_ctx* ctx, const float* power, uint64_t n, uint8_t* out);

#ifdef __cplusplus
}
#endif

#endif /* RVX_H */
