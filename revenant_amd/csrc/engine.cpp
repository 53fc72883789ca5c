// The streaming engine and the batched receivers' record step.
//
// rvx_vrx_batch_plan: DemodStage::record (core/engine/vrx_stage.cpp:1119-1480)
// for a whole receiver set: per receiver the fine-block and demod-block params
// (rvx_fine_block / rvx_demod_block, vrx_reference.cpp:2786-2909) of the audio
// it can now produce, cut into tiles and ordered largest first.
//
// rvx_engine: the per-block orchestration of Graph::on_block
// (core/engine/graph.cpp:4770-5480) over the fused chain:
//   - samples land in a device window that keeps the filter history
//     (M*L - 1 samples) of the next block, so a push of any size continues the
//     stream exactly (the reference keeps a sample ring; the chain here reads
//     native codes, so the window holds codes);
//   - the first block is the first whose history lies in the stream
//     (graph.cpp:5036-5050), block_count is what the window completes, in
//     whole spectrum rows when the averaged spectrum is on;
//   - each receiver keeps its next audio index (the reference's next_audio_
//     cursor) and produces every output whose channel samples are in the ring;
//     outputs whose samples have left the ring are counted as lost;
//   - receivers with a carrier loop (SAM, DSB) take the reference's three
//     stages (core/engine/vrx_stage.cpp:1306,1346,1399): the fine stage into a
//     per-receiver fine ring, the loops of every such receiver in one launch
//     (the loop runs over the fine stream in order, so it cannot sit inside a
//     tile), then the demodulator from the ring. A receiver's loop carries its
//     state from push to push; a gap in its fine stream (lost audio) restarts
//     it;
//   - with calibration on, the front-end correction stage: each push's
//     moments (read from its codes) reach the estimator three pushes later,
//     as the reference reads a frame's sums when its slot comes round again,
//     and each push's samples take the correction the estimate gave when they
//     arrived. The chain applies it while converting, with one switch point:
//     a new correction waits until no block still to run reads samples from
//     before the previous switch.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <tuple>
#include <vector>

#include "common.h"
#include "plan.h"

namespace {

constexpr uint32_t kUnbounded = 0xFFFFFFFFu;

uint32_t max_tile(const rvx_vrx_plan& p) {
    const uint32_t hist = rvx::demod_fine_history(p.demod);
    if (hist + 1 > RVX_VRX_TILE_SPAN || p.demod.decimation == 0) return 0;
    return (RVX_VRX_TILE_SPAN - hist - 1) / p.demod.decimation + 1;
}

// Multiply count of one tile of n audio outputs, for ordering the work.
uint64_t tile_cost(const rvx_vrx_plan& p, uint32_t n) {
    const auto& d = p.demod;
    const uint64_t span = uint64_t(n - 1) * d.decimation + rvx::demod_fine_history(d) + 1;
    uint64_t cost = span * p.fine.taps * 4;
    if (d.mode == RVX_DEMOD_AM || d.mode == 12) cost += span * d.dc_taps;
    cost += uint64_t(n) * d.audio_taps * (d.channels == 2 && d.pilot_taps ? 3 : 1);
    return cost;
}

// Most audio outputs one descriptor carries: the fine block's recurrence
// frac0 + count * step_rem must stay below 2^32 (vrx_reference.cpp:429-486).
uint64_t max_desc_audio(const rvx_vrx_plan& p) {
    const uint64_t fc = uint64_t(p.channel_rate), fd = uint64_t(p.demod_rate);
    const uint64_t rem = fc % fd;
    const uint64_t fine = rem == 0 ? 0xFFFFFFFFull : (0xFFFFFFFFull - (fd - 1)) / rem;
    const uint64_t hist = rvx::demod_fine_history(p.demod);
    if (fine < hist + 1) return 1;
    return (fine - hist - 1) / p.demod.decimation + 1;
}

// (first_audio, count) of the audio whose every input lies in channel samples
// first .. last: fine output j reads floor(j*Fc/Fd) - (taps-1) .. floor(j*Fc/Fd)
// (vrx_reference.cpp:489-620), audio a reads fine a*R - hist .. a*R.
void audio_range(const rvx_vrx_plan& p, uint64_t first, uint64_t last, uint64_t* a0,
                 uint64_t* n) {
    using u128 = unsigned __int128;
    const u128 fc = uint64_t(p.channel_rate), fd = uint64_t(p.demod_rate);
    const uint64_t R = p.demod.decimation, hist = rvx::demod_fine_history(p.demod);
    const u128 lo_n = u128(first) + p.fine.taps - 1;
    const uint64_t j_lo = uint64_t((lo_n * fd + fc - 1) / fc);
    const uint64_t j_hi = uint64_t(((u128(last) + 1) * fd - 1) / fc);
    const uint64_t a_lo = (j_lo + hist + R - 1) / R;
    const uint64_t a_hi = j_hi / R;
    *a0 = a_lo;
    *n = a_hi >= a_lo ? a_hi - a_lo + 1 : 0;
}

// The 32-bit resampler recurrence's limit on one fine dispatch
// (frac0 + i*step_rem below 2^32).
uint64_t max_fine_outputs(const rvx_vrx_plan& p) {
    const uint64_t fc = uint64_t(p.channel_rate), fd = uint64_t(p.demod_rate);
    const uint64_t rem = fc % fd;
    return rem == 0 ? 0xFFFFFFFFull : (0xFFFFFFFFull - (fd - 1)) / rem + 1;
}

}  // namespace

extern "C" int32_t rvx_vrx_batch_plan(const rvx_vrx_span* spans, uint32_t n_spans,
                                      rvx_vrx_desc* descs, uint32_t max_descs,
                                      uint32_t* n_descs, rvx_vrx_work* work, uint32_t max_work,
                                      uint32_t* n_work) {
    static const char* who = "rvx_vrx_batch_plan";
    if (!n_descs || !n_work || (n_spans && !spans))
        return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    struct Item {
        uint64_t cost;
        uint32_t v, j;
    };
    std::vector<Item> items;
    uint32_t nd = 0;
    for (uint32_t s = 0; s < n_spans; ++s) {
        const rvx_vrx_span& sp = spans[s];
        if (!sp.plan) return rvx::fail(RVX_EINVAL, "%s: span %u has no plan", who, s);
        if (sp.count == 0) continue;
        const rvx_vrx_plan& p = *sp.plan;
        if (p.carrier.power != RVX_CARRIER_NONE)
            return rvx::fail(RVX_EINVAL,
                             "%s: span %u is a mode %u receiver, whose carrier loop runs over "
                             "its fine stream in order: it takes rvx_vrx_fine, "
                             "rvx_vrx_carrier_batch and rvx_vrx_demod, not a batched tile",
                             who, s, p.mode);
        const uint32_t cap = max_tile(p);
        if (cap == 0) return rvx::fail(RVX_EINVAL, "%s: span %u's filters do not fit a tile", who, s);
        uint32_t tile = sp.tile ? std::min(sp.tile, cap) : cap;
        tile = std::min(tile, sp.count);
        uint64_t step = max_desc_audio(p);
        step -= step % tile;
        step = std::max<uint64_t>(step, tile);
        const uint32_t R = p.demod.decimation, hist = rvx::demod_fine_history(p.demod);
        if (sp.first_audio * R < hist)
            return rvx::fail(RVX_EINVAL, "%s: span %u reaches below fine sample 0", who, s);
        for (uint64_t lo = 0; lo < sp.count; lo += step) {
            const uint32_t m = uint32_t(std::min<uint64_t>(step, sp.count - lo));
            if (descs) {
                if (nd >= max_descs) return rvx::fail(RVX_EINVAL, "%s: descriptor array full", who);
                rvx_vrx_desc d{};
                d.fine_config = p.fine;
                d.demod_config = p.demod;
                const uint64_t a0 = sp.first_audio + lo;
                if (int32_t rc = rvx_fine_block(&p, sp.channel_base, sp.channel_mask, kUnbounded,
                                                a0 * R - hist, (m - 1) * R + hist + 1, &d.fine,
                                                nullptr))
                    return rc;
                if (int32_t rc = rvx_demod_block(&p, kUnbounded, a0, m, &d.demod)) return rc;
                d.hist = hist;
                d.tile = tile;
                d.channel_ring = sp.channel_ring;
                d.taps = sp.taps;
                d.weights = sp.weights;
                d.audio = sp.audio + lo * p.demod.channels;
                if (int32_t rc = rvx_vrx_desc_check(&d)) return rc;
                descs[nd] = d;
            }
            const uint32_t tiles = (m + tile - 1) / tile;
            for (uint32_t j = 0; j < tiles; ++j)
                items.push_back({tile_cost(p, std::min(tile, m - j * tile)), nd, j});
            ++nd;
        }
    }
    *n_descs = nd;
    *n_work = uint32_t(items.size());
    if (items.size() > 0xFFFFFFFFull) return rvx::fail(RVX_EINVAL, "%s: too many tiles", who);
    if (!work) return RVX_OK;
    if (items.size() > max_work) return rvx::fail(RVX_EINVAL, "%s: work array full", who);
    std::sort(items.begin(), items.end(), [](const Item& a, const Item& b) {
        return std::make_tuple(b.cost, a.v, a.j) < std::make_tuple(a.cost, b.v, b.j);
    });
    for (size_t i = 0; i < items.size(); ++i) work[i] = {items[i].v, items[i].j};
    return RVX_OK;
}

// ---------------------------------------------------------------------------

namespace {

struct EngineRx {
    uint32_t id = 0;
    uint32_t grid = 0;  // 0: the engine's own grid; k: the k-th added grid
    uint32_t row = 0;   // ring row of the receiver's channel on its grid
    rvx_vrx_plan plan{};
    void* taps = nullptr;
    float* weights = nullptr;
    float* audio = nullptr;
    uint64_t cap = 0;  // audio outputs the ring holds
    uint32_t tile = 0;
    bool started = false;
    uint64_t next_audio = 0;
    uint64_t read_audio = 0;
    uint64_t n_weights = 0;
    // carrier-loop receivers (SAM, DSB) only
    bool carrier = false;
    void* fine = nullptr;        // fine ring, fine_cap cf32
    uint64_t fine_cap = 0;
    uint32_t* cstate = nullptr;  // the loop's RVX_CARRIER_STATE_WORDS
    bool fine_started = false;
    uint64_t next_fine = 0;      // first fine sample not yet produced and rotated
    // audio noise stages (mono receivers): the line kernel, then spectral
    // reduction, in place on the audio each push produces
    bool line_on = false, spectral_on = false;
    rvx_line_config line_cfg{};
    rvx_line_params line{};
    float* line_state = nullptr;
    uint32_t spec_frame = 0;
    rvx_spectral_params spec{};
    float* spec_state = nullptr;
    float* spec_tables = nullptr;
    uint64_t pushed_a0 = 0, pushed_n = 0;  // the audio this push produced
    // the impulse blanker on the receiver's coarse channel: the fine stage
    // then reads the blanked ring (same slots as the channel ring), which
    // runs `lead` channel samples behind the channel
    bool blank_on = false;
    rvx_blanker_config blank_cfg{};
    float blank_threshold = 0.0f;
    float* flags = nullptr;   // ring_blocks floats
    void* blanked = nullptr;  // ring_blocks cf32
    bool blank_started = false;
    uint64_t detect_next = 0, apply_next = 0, apply_first = 0;
    // the display tap: the fine kernel with the display table into a display
    // ring (channel 0 of two; channel 1 stays zero for the pane's transform)
    bool display_on = false;
    rvx_vrx_display_plan display{};
    void* display_taps = nullptr;     // RVX_DISPLAY_TAPS + 1 cf32
    void* display_ring = nullptr;     // 2 * display_cap cf32
    uint64_t display_cap = 0;
    uint32_t transform = 0;
    void* pane_tw = nullptr;          // transform cf32
    float* pane_window = nullptr;     // transform + transform / 2 floats
    float* pane = nullptr;            // transform floats (two channels' central halves)
    bool display_started = false;
    uint64_t display_next = 0;        // first display output not yet produced
    uint64_t display_from = 0;        // first display output of this run
};

void free_rx(EngineRx& r) {
    for (void* q : {r.taps, (void*)r.weights, (void*)r.audio, r.fine, (void*)r.cstate,
                    (void*)r.line_state, (void*)r.spec_state, (void*)r.spec_tables,
                    (void*)r.flags, r.blanked, r.display_taps, r.display_ring, r.pane_tw,
                    (void*)r.pane_window, (void*)r.pane})
        if (q) (void)hipFree(q);
    r.flags = nullptr;
    r.blanked = nullptr;
    r.display_taps = r.display_ring = r.pane_tw = nullptr;
    r.pane_window = r.pane = nullptr;
    r.taps = r.fine = nullptr;
    r.weights = r.audio = nullptr;
    r.cstate = nullptr;
    r.line_state = r.spec_state = r.spec_tables = nullptr;
}

template <class T>
int32_t device_alloc(T** p, uint64_t bytes, const char* who) {
    void* q = nullptr;
    const hipError_t err = hipMalloc(&q, bytes);
    if (err != hipSuccess)
        return rvx::fail(RVX_ENOMEM, "%s: hipMalloc of %llu bytes failed: %s", who,
                         (unsigned long long)bytes, hipGetErrorString(err));
    *p = static_cast<T*>(q);
    return RVX_OK;
}

// A grid added beside the engine's own (rvx_engine_add_grid). It reads the
// same input window and stores only the channels its receivers sit on: each
// gets a ring row when its first receiver arrives (the row is cleared and
// its stream starts at the grid's next block) and gives it back when its
// last receiver leaves.
struct EngineGrid {
    rvx_engine_grid cfg{};
    uint64_t N = 0;  // prototype length M*L
    std::vector<float> proto;  // host copy (re-uploaded unscaled when excision turns on)
    float* taps = nullptr;
    void* tw = nullptr;
    bool taps_scaled = false;
    void* ring = nullptr;          // max_rows rows of ring_blocks cf32, ring_pitch apart
    uint32_t* rows_d = nullptr;    // ring row of each channel (max_rows: not stored)
    uint32_t* rows_h = nullptr;    // pinned staging of rows_d's uploads
    std::vector<uint32_t> row_of;  // host copy of rows_d
    std::vector<uint32_t> users;   // receivers on each row
    std::vector<uint64_t> row_first;  // first block written into each row
    bool rows_dirty = false;       // rows_d is behind row_of
    uint32_t max_blocks = 0;
    uint64_t next_block = 0;

    ~EngineGrid() {
        for (void* q : {(void*)taps, tw, ring, (void*)rows_d})
            if (q) (void)hipFree(q);
        if (rows_h) (void)hipHostFree(rows_h);
    }
};

}  // namespace

struct rvx_engine {
    rvx_ctx* ctx = nullptr;
    rvx_engine_config cfg{};
    uint64_t N = 0;  // prototype length M*L
    uint32_t bps = 0;
    float* taps = nullptr;
    void* tw = nullptr;
    void* nco = nullptr;
    uint32_t nco_entries = 0;
    bool taps_scaled = false;
    uint8_t* win = nullptr;
    uint8_t* bounce = nullptr;
    uint8_t* stage_h = nullptr;  // pinned copy of a host push
    hipEvent_t pushed = nullptr;
    uint64_t win_cap = 0, win_origin = 0, win_fill = 0;
    void* ring = nullptr;
    float* partial = nullptr;
    float* rows = nullptr;
    float* state = nullptr;
    uint32_t max_blocks = 0;
    uint64_t next_block = 0, first_block = 0;
    bool any = false;
    std::vector<EngineRx> rx;
    uint32_t next_id = 1;
    rvx_vrx_desc* descs_h = nullptr;
    rvx_vrx_work* work_h = nullptr;
    void* descs_d = nullptr;
    void* work_d = nullptr;
    uint32_t descs_cap = 0, work_cap = 0;
    uint32_t* counter = nullptr;
    hipEvent_t staged = nullptr;
    rvx_vrx_carrier_item* items_h = nullptr;  // carrier loops of one push
    void* items_d = nullptr;
    uint32_t items_cap = 0;
    uint8_t* noise_h = nullptr;  // line and spectral items of one push
    void* noise_d = nullptr;
    uint64_t noise_cap = 0;      // bytes
    hipEvent_t noise_staged = nullptr;
    // front-end correction (calibration on)
    static constexpr int kInFlight = 3;
    rvx_front_end* fe = nullptr;
    float* sums_d[kInFlight] = {};
    float* sums_h = nullptr;  // pinned, kInFlight slots of sums_per_slot
    uint64_t sums_per_slot = 0;
    uint32_t sums_count[kInFlight] = {};
    hipEvent_t sums_ready[kInFlight] = {};
    uint64_t pushes = 0;
    bool corr_on = false;             // a correction has been applied
    rvx_iq_correct_params corr_before{0, 0, 0, 0.0f, 0.0f, 0.0f, 1.0f};
    rvx_iq_correct_params corr_now{0, 0, 0, 0.0f, 0.0f, 0.0f, 1.0f};
    uint64_t corr_from = 0;           // first sample that takes corr_now
    std::vector<std::unique_ptr<EngineGrid>> grids;  // grid k is grids[k - 1]
    // The added grids' chains run on a stream of their own, forked after the
    // engine's chain and joined before the receivers, so they overlap the
    // averaged spectrum's combine (a latency-bound recurrence over few waves).
    hipStream_t aux = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    std::vector<float> proto;  // host copy of the engine's own prototype
    // The conditioned path (set before the first push by rvx_engine_set_comb
    // or rvx_engine_set_excision): each push is converted (and corrected) into
    // xring, comb cancellation and impulse excision clean it in place, the
    // cleaned samples are copied into cwin (a cf32 window that mirrors win),
    // and the chains read cwin up to cleaned_end.
    bool condition = false;
    uint64_t xring_n = 0;
    bool excise = false;
    rvx_excision_config xcfg{};
    rvx_excision_rings xrings{};
    rvx_excision_cursor* xcursor = nullptr;
    uint32_t xmask = 0;
    void* xring = nullptr;
    float* xseg_power = nullptr;
    float* xseg_level = nullptr;
    uint32_t* xflags = nullptr;
    void* xremoved = nullptr;
    uint32_t* xcounters = nullptr;
    void* cwin = nullptr;
    void* cbounce = nullptr;
    uint64_t cleaned_end = 0;
    // comb cancellation (rvx_engine_set_comb)
    bool comb_on = false, comb_subtract = false;
    std::vector<rvx_comb_line> comb_lines;  // phases at comb_index
    uint64_t comb_index = 0;
    rvx_comb_line* comb_h = nullptr;  // pinned: one push's lines
    void* comb_d = nullptr;
    void* comb_table = nullptr;       // the 4096-entry tone table
    void* comb_removed = nullptr;     // xring_n cf32
    float* comb_proj = nullptr;       // the last push's projections, a cf32 per line
    hipEvent_t comb_staged = nullptr;
    uint64_t proj_start = 0, proj_count = 0;
    uint32_t proj_lines = 0, proj_subs = 0;
    rvx_engine_stats stats{};

    // One past the last sample the chains may read.
    uint64_t readable_end() const { return condition ? cleaned_end : win_origin + win_fill; }

    // The oldest input sample a block still to run on any grid reads.
    uint64_t oldest_needed() const {
        uint64_t need = next_block * cfg.decimation - (N - 1);
        for (const auto& g : grids)
            need = std::min(need, g->next_block * g->cfg.decimation - (g->N - 1));
        return need;
    }

    ~rvx_engine() {
        if (aux) (void)hipStreamSynchronize(aux);
        if (ctx) (void)hipStreamSynchronize(ctx->stream);
        for (auto& r : rx) free_rx(r);
        grids.clear();
        for (void* p : {xring, (void*)xseg_power, (void*)xseg_level, (void*)xflags, xremoved,
                        (void*)xcounters, cwin, cbounce, comb_d, comb_table, comb_removed,
                        (void*)comb_proj})
            if (p) (void)hipFree(p);
        if (comb_h) (void)hipHostFree(comb_h);
        if (comb_staged) (void)hipEventDestroy(comb_staged);
        if (xcursor) rvx_excision_cursor_destroy(xcursor);
        for (void* p : {(void*)taps, tw, nco, (void*)win, (void*)bounce, ring, (void*)partial,
                        (void*)rows, (void*)state, descs_d, work_d, (void*)counter, items_d})
            if (p) (void)hipFree(p);
        if (stage_h) (void)hipHostFree(stage_h);
        if (pushed) (void)hipEventDestroy(pushed);
        if (descs_h) (void)hipHostFree(descs_h);
        if (work_h) (void)hipHostFree(work_h);
        if (items_h) (void)hipHostFree(items_h);
        if (noise_h) (void)hipHostFree(noise_h);
        if (noise_d) (void)hipFree(noise_d);
        if (noise_staged) (void)hipEventDestroy(noise_staged);
        if (fork) (void)hipEventDestroy(fork);
        if (join) (void)hipEventDestroy(join);
        if (aux) (void)hipStreamDestroy(aux);
        if (staged) (void)hipEventDestroy(staged);
        for (int k = 0; k < kInFlight; ++k) {
            if (sums_d[k]) (void)hipFree(sums_d[k]);
            if (sums_ready[k]) (void)hipEventDestroy(sums_ready[k]);
        }
        if (sums_h) (void)hipHostFree(sums_h);
        if (fe) rvx_front_end_destroy(fe);
    }
};

namespace {

// Where a receiver's coarse channel lives: its grid's ring, the first slot of
// its row, the ring depth, the first block of the row's stream, and the
// channel samples the ring holds for it (none when newest < oldest).
struct Source {
    const void* ring;
    uint64_t ring_n;  // cf32 entries
    uint32_t base;
    uint32_t blocks;
    uint32_t max_blocks;  // blocks one push can add
    uint64_t first;
    uint64_t oldest, newest;
    bool any;
};

Source source_of(const rvx_engine& e, const EngineRx& r) {
    Source src{};
    uint64_t next = 0;
    if (r.grid == 0) {
        src.ring = e.ring;
        src.blocks = e.cfg.ring_blocks;
        src.ring_n = uint64_t(e.cfg.channels) * e.cfg.ring_pitch;
        src.base = r.plan.placement.channel * e.cfg.ring_pitch;
        src.max_blocks = e.max_blocks;
        src.first = e.first_block;
        src.any = e.any;
        next = e.next_block;
    } else {
        const EngineGrid& g = *e.grids[r.grid - 1];
        src.ring = g.ring;
        src.blocks = g.cfg.ring_blocks;
        src.ring_n = uint64_t(g.cfg.max_rows) * g.cfg.ring_pitch;
        src.base = r.row * g.cfg.ring_pitch;
        src.max_blocks = g.max_blocks;
        src.first = g.row_first[r.row];
        src.any = g.next_block > src.first;
        next = g.next_block;
    }
    src.newest = next - 1;
    src.oldest = std::max(src.first, next >= src.blocks ? next - src.blocks : 0);
    if (!src.any) src.oldest = src.newest + 1;
    return src;
}

// One carrier-loop receiver's share of a push: audio a0 .. a0+count-1 from
// fine samples first_fine .. the last, after the loop has rotated them.
struct CarrierJob {
    EngineRx* r;
    uint64_t a0, count;
};

// The fine stage of one carrier receiver over its new fine samples, in
// dispatches the resampler recurrence allows, and the loop's item.
int32_t carrier_fine(rvx_engine& e, EngineRx& r, uint64_t a0, uint64_t count,
                     rvx_vrx_carrier_item* item) {
    const uint64_t R = r.plan.demod.decimation, hist = rvx::demod_fine_history(r.plan.demod);
    const uint64_t need = a0 * R - hist;  // audio_range keeps a0*R >= hist
    const uint64_t last = (a0 + count - 1) * R;
    bool reset = false;
    if (!r.fine_started || r.next_fine < need) {  // the first push, or a gap
        r.fine_started = true;
        r.next_fine = need;
        reset = true;
    }
    const uint64_t n_new = last + 1 - r.next_fine;
    const uint32_t mask = uint32_t(r.fine_cap - 1);
    const Source src = source_of(e, r);
    const uint64_t limit = std::min<uint64_t>(max_fine_outputs(r.plan), 0x7FFFFFFFull);
    const uint64_t n_taps = uint64_t(r.plan.fine.phases) * r.plan.fine.taps + 1;
    for (uint64_t at = r.next_fine; at <= last;) {
        const uint32_t n = uint32_t(std::min<uint64_t>(last + 1 - at, limit));
        rvx_vrx_fine_params fp{};
        uint64_t first_input = 0;
        const bool b = r.blank_on;
        if (int32_t rc = rvx_fine_block(&r.plan, b ? 0u : src.base, src.blocks - 1, mask, at, n,
                                        &fp, &first_input))
            return rc;
        if (int32_t rc = rvx_vrx_fine(e.ctx, &r.plan.fine, &fp, b ? r.blanked : src.ring,
                                      b ? uint64_t(src.blocks) : src.ring_n, r.taps, n_taps,
                                      e.nco, e.nco_entries, r.fine, r.fine_cap))
            return rc;
        at += n;
    }
    rvx_vrx_carrier_item it{};
    it.config = r.plan.carrier;
    if (int32_t rc = rvx_carrier_block(&r.plan, mask, r.next_fine, uint32_t(n_new), reset ? 1 : 0,
                                       0.0f, &it.params))
        return rc;
    it.fine_ring = r.fine;
    it.fine_samples = r.fine_cap;
    it.state = r.cstate;
    if (int32_t rc = rvx_vrx_carrier_item_check(&it, e.nco_entries)) return rc;
    *item = it;
    r.next_fine = last + 1;
    return RVX_OK;
}

// The loops of every carrier receiver in one launch, then each receiver's
// demodulator from its rotated fine ring into its audio ring.
int32_t carrier_receivers(rvx_engine& e, const std::vector<CarrierJob>& jobs,
                          const std::vector<rvx_vrx_carrier_item>& items, const char* who) {
    if (jobs.empty()) return RVX_OK;
    const uint32_t n = uint32_t(items.size());
    RVX_TRY_HIP(hipEventSynchronize(e.staged), who);
    if (n > e.items_cap) {
        RVX_TRY_HIP(hipStreamSynchronize(e.ctx->stream), who);
        (void)hipHostFree(e.items_h);
        (void)hipFree(e.items_d);
        e.items_h = nullptr;
        e.items_d = nullptr;
        e.items_cap = std::max(n, 2 * e.items_cap);
        RVX_TRY_HIP(hipHostMalloc((void**)&e.items_h, e.items_cap * sizeof(rvx_vrx_carrier_item),
                                  hipHostMallocDefault),
                    who);
        if (int32_t rc = device_alloc(&e.items_d, e.items_cap * sizeof(rvx_vrx_carrier_item), who))
            return rc;
    }
    std::memcpy(e.items_h, items.data(), n * sizeof(rvx_vrx_carrier_item));
    RVX_TRY_HIP(hipMemcpyAsync(e.items_d, e.items_h, n * sizeof(rvx_vrx_carrier_item),
                               hipMemcpyHostToDevice, e.ctx->stream),
                who);
    RVX_TRY_HIP(hipEventRecord(e.staged, e.ctx->stream), who);
    if (int32_t rc = rvx_vrx_carrier_batch(e.ctx, static_cast<const rvx_vrx_carrier_item*>(e.items_d),
                                           n, e.nco, e.nco_entries))
        return rc;
    for (const CarrierJob& j : jobs) {
        EngineRx& r = *j.r;
        const uint32_t ch = r.plan.demod.channels;
        uint64_t a = j.a0, left = j.count;
        while (left) {  // one dispatch per contiguous stretch of the audio ring
            const uint64_t pos = a % r.cap;
            const uint64_t seg = std::min<uint64_t>({left, r.cap - pos, 0x7FFFFFFFull});
            rvx_vrx_demod_params dp{};
            if (int32_t rc = rvx_demod_block(&r.plan, uint32_t(r.fine_cap - 1), a, uint32_t(seg),
                                             &dp))
                return rc;
            if (int32_t rc = rvx_vrx_demod(e.ctx, &r.plan.demod, &dp, r.fine, r.fine_cap,
                                           r.weights, r.n_weights, r.audio + pos * ch, seg * ch))
                return rc;
            a += seg;
            left -= seg;
        }
    }
    return RVX_OK;
}

// The front-end stage for one push: the moments that have come round feed the
// estimator; this push's samples get the correction the estimate now gives,
// unless the previous switch point is still read by a block to come; then
// this push's moments are summed from its codes and copied out.
int32_t front_end_push(rvx_engine& e, uint64_t push_first, uint64_t samples, const char* who) {
    const int slot = int(e.pushes % rvx_engine::kInFlight);
    if (e.sums_count[slot]) {
        RVX_TRY_HIP(hipEventSynchronize(e.sums_ready[slot]), who);
        rvx_iq_moments_total m{};
        const float* h = e.sums_h + slot * e.sums_per_slot;
        if (int32_t rc = rvx_total_moments(h, e.sums_per_slot, e.sums_count[slot], &m)) return rc;
        if (int32_t rc = rvx_front_end_update(e.fe, &m, e.cfg.sample_rate))
            return rc;
        e.sums_count[slot] = 0;
    }
    rvx_front_end_estimate est{};
    rvx_front_end_estimate_get(e.fe, &est);
    if (est.measured) {
        rvx_iq_correct_params want{};
        rvx_front_end_correction(e.fe, (e.cfg.calibration & RVX_CALIBRATE_DC) != 0,
                                 (e.cfg.calibration & RVX_CALIBRATE_IQ) != 0, &want);
        const bool differs = std::memcmp(&want, &e.corr_now, sizeof(want)) != 0;
        // the oldest sample any block still to run reads
        const uint64_t oldest = e.oldest_needed();
        if (differs && (!e.corr_on || oldest >= e.corr_from)) {
            e.corr_before = e.corr_now;
            e.corr_now = want;
            e.corr_from = push_first;
            e.corr_on = true;
        }
    }
    // this push's moments, from its codes in the window
    const uint32_t n = uint32_t(samples);
    if (int32_t rc = rvx_iq_moments_codes(e.ctx, e.cfg.format, e.win, e.win_fill,
                                          push_first - e.win_origin, n, e.sums_d[slot],
                                          e.sums_per_slot))
        return rc;
    const uint64_t used = (uint64_t(n) + RVX_IQ_MOMENTS_CHUNK - 1) / RVX_IQ_MOMENTS_CHUNK *
                          RVX_IQ_MOMENTS_PER_CHUNK;
    RVX_TRY_HIP(hipMemcpyAsync(e.sums_h + slot * e.sums_per_slot, e.sums_d[slot],
                               used * sizeof(float), hipMemcpyDeviceToHost, e.ctx->stream),
                who);
    RVX_TRY_HIP(hipEventRecord(e.sums_ready[slot], e.ctx->stream), who);
    e.sums_count[slot] = n;
    ++e.pushes;
    return RVX_OK;
}

int32_t batch_receivers(rvx_engine& e, const std::vector<rvx_vrx_span>& spans, const char* who);
int32_t noise_stages(rvx_engine& e, const char* who);

// The display stream over the receiver's source samples lo .. hi (the channel,
// or the blanked channel the fine stage reads): output j reads source samples
// j*R - 255 .. j*R, so outputs up to hi / R are produced, from the first whose
// history is inside [lo, ...] on a new run (vrx_reference.h:1639-1768).
int32_t display_stream(rvx_engine& e, EngineRx& r, uint64_t lo, uint64_t hi, const char* who) {
    const uint64_t R = r.display.decimation;
    const uint64_t hist = RVX_DISPLAY_TAPS - 1;
    const uint64_t j_hi = hi / R;
    if (!r.display_started) {
        r.display_started = true;
        r.display_next = r.display_from = (lo + hist + R - 1) / R;
    }
    if (r.display_next * R < lo + hist) {  // the source moved past the stream: a new run
        r.display_next = r.display_from = (lo + hist + R - 1) / R;
    }
    const bool b = r.blank_on;
    const Source from = source_of(e, r);
    const void* src = b ? r.blanked : from.ring;
    const uint32_t base = b ? 0u : from.base;
    const uint64_t src_n = b ? uint64_t(from.blocks) : from.ring_n;
    const uint64_t n_taps = RVX_DISPLAY_TAPS + 1;
    while (r.display_next <= j_hi) {
        const uint32_t n = uint32_t(std::min<uint64_t>(j_hi + 1 - r.display_next, r.display_cap));
        rvx_vrx_fine_params fp{};
        uint64_t first_input = 0;
        if (int32_t rc = rvx_display_block(&r.display, base, from.blocks - 1,
                                           uint32_t(r.display_cap - 1), r.display_next, n, &fp,
                                           &first_input))
            return rc;
        if (int32_t rc = rvx_vrx_fine(e.ctx, &r.display.fine, &fp, src, src_n, r.display_taps,
                                      n_taps, e.nco, e.nco_entries, r.display_ring,
                                      r.display_cap))
            return rc;
        r.display_next += n;
    }
    (void)who;
    return RVX_OK;
}

// The blanker over a receiver's new channel samples (noise_reference.h:59-190):
// detect through the newest sample, apply through `lead` before it (a sample's
// verdict waits for the flags up to lead after it), in pieces that keep each
// dispatch's reach inside the ring. The first push starts at the stream's first
// block, whose window reads the ring's cleared slots (a zero window never
// flags); a blanker enabled mid-stream starts a window past the oldest sample
// the ring still holds.
int32_t blank_channel(rvx_engine& e, EngineRx& r, uint64_t oldest, uint64_t newest,
                      const char* who) {
    const rvx_blanker_config& c = r.blank_cfg;
    const Source src = source_of(e, r);
    const uint32_t mask = src.blocks - 1;
    const uint64_t below = uint64_t(c.hang) + c.guard + c.window;
    if (!r.blank_started) {
        r.blank_started = true;
        const uint64_t start = oldest == src.first ? src.first : oldest + c.guard + c.window;
        r.detect_next = start;
        r.apply_next = r.apply_first = start;
    }
    const uint64_t piece_max = uint64_t(mask) + 1 - below - c.lead - 1;
    rvx_blanker_params p{};
    p.chan_base = src.base;
    p.chan_mask = mask;
    p.out_mask = mask;
    p.threshold = r.blank_threshold;
    const uint64_t ring_n = src.ring_n;
    while (r.detect_next <= newest) {
        const uint64_t n = std::min(newest + 1 - r.detect_next, piece_max);
        p.chan_first = uint32_t(r.detect_next & mask);
        p.out_first = p.chan_first;
        p.count = uint32_t(n);
        if (int32_t rc = rvx_blank_detect(e.ctx, &c, &p, src.ring, ring_n, r.flags, mask + 1ull))
            return rc;
        r.detect_next += n;
    }
    const uint64_t apply_end = r.detect_next >= c.lead ? r.detect_next - c.lead : 0;
    while (r.apply_next < apply_end) {
        const uint64_t n = std::min(apply_end - r.apply_next, piece_max);
        p.chan_first = uint32_t(r.apply_next & mask);
        p.out_first = p.chan_first;
        p.count = uint32_t(n);
        if (int32_t rc = rvx_blank_apply(e.ctx, &c, &p, src.ring, ring_n, r.flags, mask + 1ull,
                                         r.blanked, mask + 1ull))
            return rc;
        r.apply_next += n;
    }
    (void)who;
    return RVX_OK;
}

// The audio ring holds the last cap outputs: once the writer is more than cap
// ahead of the reader, the oldest unread outputs have been overwritten. They
// are counted, and the reader moves to the oldest output still held.
void reader_follows(rvx_engine& e, EngineRx& r) {
    if (r.next_audio - r.read_audio > r.cap) {
        e.stats.audio_overrun += (r.next_audio - r.cap) - r.read_audio;
        r.read_audio = r.next_audio - r.cap;
    }
}

int32_t engine_receivers(rvx_engine& e, const char* who) {
    std::vector<rvx_vrx_span> spans;
    std::vector<CarrierJob> jobs;
    std::vector<rvx_vrx_carrier_item> items;
    for (auto& r : e.rx) r.pushed_n = 0;
    for (auto& r : e.rx) {
        const Source src = source_of(e, r);
        if (src.newest + 1 <= src.oldest) continue;  // nothing in its row yet
        const uint64_t oldest = src.oldest, newest = src.newest;
        uint64_t a_lo = 0, n = 0;
        if (r.display_on) {
            if (r.blank_on) {
                if (int32_t rc = blank_channel(e, r, oldest, newest, who)) return rc;
                if (r.apply_next > r.apply_first) {
                    if (int32_t rc = display_stream(e, r, std::max(oldest, r.apply_first),
                                                    r.apply_next - 1, who))
                        return rc;
                }
            } else if (int32_t rc = display_stream(e, r, oldest, newest, who)) {
                return rc;
            }
        }
        if (r.blank_on) {
            if (!r.display_on)
                if (int32_t rc = blank_channel(e, r, oldest, newest, who)) return rc;
            if (r.apply_next == r.apply_first) continue;  // nothing blanked yet
            const uint64_t lo = std::max(oldest, r.apply_first);
            audio_range(r.plan, lo, r.apply_next - 1, &a_lo, &n);
        } else {
            audio_range(r.plan, oldest, newest, &a_lo, &n);
        }
        if (n == 0) continue;
        const uint64_t a_hi = a_lo + n - 1;
        if (!r.started) {
            r.started = true;
            r.next_audio = r.read_audio = a_lo;
        }
        if (r.next_audio < a_lo) {
            e.stats.audio_lost += a_lo - r.next_audio;
            r.next_audio = a_lo;
        }
        if (r.next_audio > a_hi) continue;
        uint64_t count = a_hi - r.next_audio + 1;
        if (count > r.cap) {  // only the last cap outputs fit the audio ring
            e.stats.audio_lost += count - r.cap;
            r.next_audio += count - r.cap;
            count = r.cap;
        }
        r.pushed_a0 = r.next_audio;
        r.pushed_n = count;
        if (r.carrier) {
            rvx_vrx_carrier_item it{};
            if (int32_t rc = carrier_fine(e, r, r.next_audio, count, &it)) return rc;
            jobs.push_back({&r, r.next_audio, count});
            items.push_back(it);
            r.next_audio += count;
            reader_follows(e, r);
            continue;
        }
        const uint32_t ch = r.plan.demod.channels;
        uint64_t a = r.next_audio, left = count;
        while (left) {  // one span per contiguous stretch of the audio ring
            const uint64_t pos = a % r.cap;
            const uint64_t seg = std::min<uint64_t>({left, r.cap - pos, 0x7FFFFFFFull});
            rvx_vrx_span s{};
            s.plan = &r.plan;
            s.channel_base = r.blank_on ? 0u : src.base;
            s.channel_mask = src.blocks - 1;
            s.channel_ring = r.blank_on ? r.blanked : src.ring;
            s.taps = r.taps;
            s.weights = r.weights;
            s.audio = r.audio + pos * ch;
            s.first_audio = a;
            s.count = uint32_t(seg);
            s.tile = r.tile;
            spans.push_back(s);
            a += seg;
            left -= seg;
        }
        r.next_audio += count;
        reader_follows(e, r);
    }
    if (int32_t rc = carrier_receivers(e, jobs, items, who)) return rc;
    if (int32_t rc = batch_receivers(e, spans, who)) return rc;
    return noise_stages(e, who);
}

int32_t batch_receivers(rvx_engine& e, const std::vector<rvx_vrx_span>& spans, const char* who) {
    if (spans.empty()) return RVX_OK;
    uint32_t nd = 0, nw = 0;
    if (int32_t rc = rvx_vrx_batch_plan(spans.data(), uint32_t(spans.size()), nullptr, 0, &nd,
                                        nullptr, 0, &nw))
        return rc;
    // The staging buffers are rewritten only after the previous upload left them.
    RVX_TRY_HIP(hipEventSynchronize(e.staged), who);
    if (nd > e.descs_cap || nw > e.work_cap) {
        RVX_TRY_HIP(hipStreamSynchronize(e.ctx->stream), who);
        if (nd > e.descs_cap) {
            (void)hipHostFree(e.descs_h);
            (void)hipFree(e.descs_d);
            e.descs_h = nullptr;
            e.descs_d = nullptr;
            e.descs_cap = std::max(nd, 2 * e.descs_cap);
            RVX_TRY_HIP(hipHostMalloc((void**)&e.descs_h, e.descs_cap * sizeof(rvx_vrx_desc),
                                      hipHostMallocDefault),
                        who);
            if (int32_t rc = device_alloc(&e.descs_d, e.descs_cap * sizeof(rvx_vrx_desc), who))
                return rc;
        }
        if (nw > e.work_cap) {
            (void)hipHostFree(e.work_h);
            (void)hipFree(e.work_d);
            e.work_h = nullptr;
            e.work_d = nullptr;
            e.work_cap = std::max(nw, 2 * e.work_cap);
            RVX_TRY_HIP(hipHostMalloc((void**)&e.work_h, e.work_cap * sizeof(rvx_vrx_work),
                                      hipHostMallocDefault),
                        who);
            if (int32_t rc = device_alloc(&e.work_d, e.work_cap * sizeof(rvx_vrx_work), who))
                return rc;
        }
    }
    if (int32_t rc = rvx_vrx_batch_plan(spans.data(), uint32_t(spans.size()), e.descs_h,
                                        e.descs_cap, &nd, e.work_h, e.work_cap, &nw))
        return rc;
    RVX_TRY_HIP(hipMemcpyAsync(e.descs_d, e.descs_h, nd * sizeof(rvx_vrx_desc),
                               hipMemcpyHostToDevice, e.ctx->stream),
                who);
    RVX_TRY_HIP(hipMemcpyAsync(e.work_d, e.work_h, nw * sizeof(rvx_vrx_work),
                               hipMemcpyHostToDevice, e.ctx->stream),
                who);
    RVX_TRY_HIP(hipEventRecord(e.staged, e.ctx->stream), who);
    return rvx_vrx_batch(e.ctx, static_cast<const rvx_vrx_desc*>(e.descs_d), nd,
                         static_cast<const rvx_vrx_work*>(e.work_d), nw, e.counter, e.nco,
                         e.nco_entries, 0);
}

// The audio noise stages of one push (the reference records them after the
// demodulator, noise_reference.h:190-520): every receiver's line kernel over
// the audio this push produced, then every receiver's spectral reduction, in
// place in the audio ring. Audio that wraps the ring is two pieces of one
// receiver, which must run in order, so the second pieces go in a second
// launch. Spectral items are launched per frame size.
int32_t noise_stages(rvx_engine& e, const char* who) {
    // launch groups: the line kernel per config, spectral reduction per frame
    using LineKey = std::tuple<uint32_t, uint32_t, uint32_t>;
    struct SpecGroup {
        const float* tables = nullptr;
        std::vector<rvx_noise_spectral_item> items;
    };
    std::map<LineKey, std::pair<rvx_line_config, std::vector<rvx_noise_line_item>>> line[2];
    std::map<uint32_t, SpecGroup> spec[2];
    for (auto& r : e.rx) {
        if (!r.pushed_n || (!r.line_on && !r.spectral_on)) continue;
        const uint64_t pos = r.pushed_a0 % r.cap;
        const uint64_t first = std::min(r.pushed_n, r.cap - pos);
        const uint64_t piece[2] = {first, r.pushed_n - first};
        float* at[2] = {r.audio + pos, r.audio};
        for (int k = 0; k < 2; ++k) {
            if (!piece[k]) continue;
            if (r.line_on) {
                rvx_noise_line_item it{};
                it.params = r.line;
                it.params.count = uint32_t(piece[k]);
                it.audio = at[k];
                it.state = r.line_state;
                auto& g = line[k][LineKey{r.line_cfg.lanes, r.line_cfg.taps, r.line_cfg.history}];
                g.first = r.line_cfg;
                g.second.push_back(it);
            }
            if (r.spectral_on) {
                rvx_noise_spectral_item it{};
                it.params = r.spec;
                it.params.count = uint32_t(piece[k]);
                it.audio = at[k];
                it.state = r.spec_state;
                SpecGroup& g = spec[k][r.spec_frame];
                g.tables = r.spec_tables;
                g.items.push_back(it);
            }
        }
    }
    struct Launch {
        const rvx_line_config* line_cfg;  // null: spectral
        uint32_t frame;
        const float* tables;
        uint64_t offset;
        uint32_t n;
    };
    std::vector<Launch> launches;
    uint64_t bytes = 0;
    for (int k = 0; k < 2; ++k) {
        for (auto& g : line[k]) bytes += g.second.second.size() * sizeof(rvx_noise_line_item);
        for (auto& g : spec[k]) bytes += g.second.items.size() * sizeof(rvx_noise_spectral_item);
    }
    if (bytes == 0) return RVX_OK;
    RVX_TRY_HIP(hipEventSynchronize(e.noise_staged), who);
    if (bytes > e.noise_cap) {
        RVX_TRY_HIP(hipStreamSynchronize(e.ctx->stream), who);
        (void)hipHostFree(e.noise_h);
        (void)hipFree(e.noise_d);
        e.noise_h = nullptr;
        e.noise_d = nullptr;
        e.noise_cap = std::max(bytes, 2 * e.noise_cap);
        RVX_TRY_HIP(hipHostMalloc((void**)&e.noise_h, e.noise_cap, hipHostMallocDefault), who);
        if (int32_t rc = device_alloc(&e.noise_d, e.noise_cap, who)) return rc;
    }
    // order: every line piece (first pieces, then the wrapped remainders),
    // then every spectral piece the same way; a receiver's stages and pieces
    // therefore run in stream order
    uint64_t off = 0;
    for (int k = 0; k < 2; ++k)
        for (auto& g : line[k]) {
            const auto& items = g.second.second;
            std::memcpy(e.noise_h + off, items.data(), items.size() * sizeof(rvx_noise_line_item));
            launches.push_back({&g.second.first, 0, nullptr, off, uint32_t(items.size())});
            off += items.size() * sizeof(rvx_noise_line_item);
        }
    for (int k = 0; k < 2; ++k)
        for (auto& g : spec[k]) {
            const auto& items = g.second.items;
            std::memcpy(e.noise_h + off, items.data(),
                        items.size() * sizeof(rvx_noise_spectral_item));
            launches.push_back({nullptr, g.first, g.second.tables, off, uint32_t(items.size())});
            off += items.size() * sizeof(rvx_noise_spectral_item);
        }
    RVX_TRY_HIP(hipMemcpyAsync(e.noise_d, e.noise_h, off, hipMemcpyHostToDevice, e.ctx->stream),
                who);
    RVX_TRY_HIP(hipEventRecord(e.noise_staged, e.ctx->stream), who);
    for (const Launch& l : launches) {
        const uint8_t* base = static_cast<const uint8_t*>(e.noise_d) + l.offset;
        const int32_t rc =
            l.line_cfg
                ? rvx_noise_line_batch(e.ctx, l.line_cfg,
                                       reinterpret_cast<const rvx_noise_line_item*>(base), l.n)
                : rvx_noise_spectral_batch(e.ctx, l.frame, l.tables, 4ull * l.frame,
                                           reinterpret_cast<const rvx_noise_spectral_item*>(base),
                                           l.n);
        if (rc) return rc;
    }
    return RVX_OK;
}

}  // namespace

extern "C" {

int32_t rvx_engine_create(rvx_ctx* ctx, const rvx_engine_config* config, const float* prototype,
                          uint64_t n_prototype, rvx_engine** out) {
    static const char* who = "rvx_engine_create";
    RVX_ENTER(ctx, who);
    if (!config || !prototype || !out) return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    *out = nullptr;
    const rvx_engine_config c = *config;
    const uint32_t M = c.channels, L = c.taps_per_branch, D = c.decimation;
    if (c.format != RVX_FORMAT_CU8 && c.format != RVX_FORMAT_CS16)
        return rvx::fail(RVX_EINVAL, "%s: format %d (the chain reads cu8 or cs16)", who, c.format);
    if (M < 2 || !rvx::pow2(M) || M > 4096 || L == 0 || D == 0 || M % D != 0)
        return rvx::fail(RVX_EINVAL, "%s: invalid grid M=%u L=%u D=%u", who, M, L, D);
    if (n_prototype != uint64_t(M) * L)
        return rvx::fail(RVX_EINVAL, "%s: wants %llu prototype taps", who,
                         (unsigned long long)(uint64_t(M) * L));
    if (c.max_push == 0 || !rvx::pow2(c.ring_blocks))
        return rvx::fail(RVX_EINVAL, "%s: max_push must be > 0 and ring_blocks a power of two",
                         who);
    if (c.ring_pitch && (c.ring_pitch < c.ring_blocks || c.ring_pitch % 4))
        return rvx::fail(RVX_EINVAL, "%s: ring pitch %u (at least %u, a multiple of 4)", who,
                         c.ring_pitch, c.ring_blocks);
    const uint32_t unit = c.frame_blocks ? c.frame_blocks * std::max(1u, c.row_frames) : 1;
    if (c.frame_blocks && c.row_frames == 0)
        return rvx::fail(RVX_EINVAL, "%s: row_frames must be >= 1", who);
    if (c.calibration & ~(RVX_CALIBRATE_DC | RVX_CALIBRATE_IQ))
        return rvx::fail(RVX_EINVAL, "%s: calibration flags 0x%x", who, c.calibration);
    if (c.calibration && c.sample_rate <= 0)
        return rvx::fail(RVX_EINVAL, "%s: calibration needs the sample rate", who);
    auto e = new rvx_engine();
    e->ctx = ctx;
    e->cfg = c;
    e->cfg.row_frames = std::max(1u, c.row_frames);
    e->cfg.ring_pitch = c.ring_pitch ? c.ring_pitch : c.ring_blocks;
    e->N = uint64_t(M) * L;
    e->bps = c.format == RVX_FORMAT_CS16 ? 4 : 2;
    // Blocks one push can complete: its own, one straddling the previous end,
    // and up to a spectrum row's worth waiting from before.
    const uint64_t blocks = uint64_t(c.max_push) / D + 2 + unit;
    e->max_blocks = uint32_t(blocks + (unit - blocks % unit) % unit);
    if (e->max_blocks > c.ring_blocks) {
        delete e;
        return rvx::fail(RVX_EINVAL, "%s: a push can complete %u blocks; the ring holds %u", who,
                         blocks > 0xFFFFFFFF ? 0xFFFFFFFFu : uint32_t(blocks), c.ring_blocks);
    }
    e->win_cap = e->N - 1 + uint64_t(e->max_blocks + 1) * D + c.max_push;
    if (e->win_cap >= (uint64_t(1) << 31)) {
        delete e;
        return rvx::fail(RVX_EINVAL, "%s: a window of %llu samples exceeds 2^31", who,
                         (unsigned long long)e->win_cap);
    }
    e->win_origin = c.start_index;
    e->next_block = (c.start_index + e->N - 1 + D - 1) / D;
    std::vector<float> tw(2 * size_t(M)), nco(2 * size_t(1) << 16);
    int32_t rc = rvx_build_twiddles(M, tw.data(), tw.size());
    if (!rc) rc = rvx_build_nco_table(16, nco.data(), nco.size());
    e->nco_entries = 1u << 16;
    e->proto.assign(prototype, prototype + n_prototype);
    e->cleaned_end = c.start_index;
    const uint64_t frames = c.frame_blocks ? e->max_blocks / c.frame_blocks : 0;
    if (!rc) rc = device_alloc(&e->taps, n_prototype * sizeof(float), who);
    if (!rc) rc = device_alloc(&e->tw, tw.size() * sizeof(float), who);
    if (!rc) rc = device_alloc(&e->nco, nco.size() * sizeof(float), who);
    if (!rc) rc = device_alloc(&e->win, e->win_cap * e->bps, who);
    if (!rc) rc = device_alloc(&e->bounce, (e->N + uint64_t(e->max_blocks + 1) * D) * e->bps, who);
    if (!rc) rc = device_alloc(&e->ring, uint64_t(M) * e->cfg.ring_pitch * 8, who);
    // cleared, so a window reaching below the stream's first block reads zeros
    if (!rc && hipMemset(e->ring, 0, uint64_t(M) * e->cfg.ring_pitch * 8) != hipSuccess)
        rc = rvx::fail(RVX_EDEVICE, "%s: clearing the channel ring failed", who);
    if (!rc) rc = device_alloc(&e->counter, sizeof(uint32_t), who);
    if (!rc && c.calibration) {
        rc = rvx_front_end_create(nullptr, &e->fe);
        e->sums_per_slot = (uint64_t(c.max_push) + RVX_IQ_MOMENTS_CHUNK - 1) /
                           RVX_IQ_MOMENTS_CHUNK * RVX_IQ_MOMENTS_PER_CHUNK;
        for (int k = 0; !rc && k < rvx_engine::kInFlight; ++k)
            rc = device_alloc(&e->sums_d[k], e->sums_per_slot * sizeof(float), who);
        if (!rc && hipHostMalloc((void**)&e->sums_h,
                                 rvx_engine::kInFlight * e->sums_per_slot * sizeof(float),
                                 hipHostMallocDefault) != hipSuccess)
            rc = rvx::fail(RVX_ENOMEM, "%s: pinned moments buffer", who);
        for (int k = 0; !rc && k < rvx_engine::kInFlight; ++k)
            if (hipEventCreateWithFlags(&e->sums_ready[k], hipEventDisableTiming) != hipSuccess)
                rc = rvx::fail(RVX_EDEVICE, "%s: event", who);
    }
    if (!rc && frames) {
        rc = device_alloc(&e->partial, frames * M * sizeof(float), who);
        if (!rc) rc = device_alloc(&e->rows, (frames / e->cfg.row_frames) * M * sizeof(float), who);
        if (!rc) rc = device_alloc(&e->state, uint64_t(M) * sizeof(float), who);
    }
    if (rc) {
        delete e;
        return rc;
    }
    const hipStream_t s = ctx->stream;
    // cs16: taps carrying the conversion's 2^-15 when that is exact (the
    // chain's products are then the same bits with one multiply fewer).
    std::vector<float> scaled(n_prototype);
    const float* upload = prototype;
    // (Not with calibration: the correction needs the converted samples.)
    if (c.format == RVX_FORMAT_CS16 && !c.calibration &&
        rvx_prescale_taps(prototype, n_prototype, scaled.data()) == RVX_OK) {
        upload = scaled.data();
        e->taps_scaled = true;
    }
    hipError_t err = hipMemcpyAsync(e->taps, upload, n_prototype * sizeof(float),
                                    hipMemcpyHostToDevice, s);
    if (err == hipSuccess)
        err = hipMemcpyAsync(e->tw, tw.data(), tw.size() * sizeof(float), hipMemcpyHostToDevice, s);
    if (err == hipSuccess)
        err = hipMemcpyAsync(e->nco, nco.data(), nco.size() * sizeof(float),
                             hipMemcpyHostToDevice, s);
    if (err == hipSuccess) err = hipEventCreateWithFlags(&e->staged, hipEventDisableTiming);
    if (err == hipSuccess) err = hipEventRecord(e->staged, s);
    if (err == hipSuccess) err = hipEventCreateWithFlags(&e->pushed, hipEventDisableTiming);
    if (err == hipSuccess) err = hipEventRecord(e->pushed, s);
    if (err == hipSuccess) err = hipEventCreateWithFlags(&e->noise_staged, hipEventDisableTiming);
    if (err == hipSuccess) err = hipEventRecord(e->noise_staged, s);
    if (err == hipSuccess)
        err = hipHostMalloc((void**)&e->stage_h, uint64_t(c.max_push) * e->bps,
                            hipHostMallocDefault);
    if (err == hipSuccess) err = hipStreamSynchronize(s);
    if (err != hipSuccess) {
        delete e;
        return rvx::fail(RVX_EDEVICE, "%s: upload failed: %s", who, hipGetErrorString(err));
    }
    if (e->state) {
        if ((rc = rvx_buf_clear(ctx, e->state, uint64_t(M) * sizeof(float)))) {
            delete e;
            return rc;
        }
    }
    *out = e;
    return RVX_OK;
}

void rvx_engine_destroy(rvx_engine* engine) { delete engine; }

int32_t rvx_engine_add_vrx_on(rvx_engine* e, uint32_t grid, const rvx_vrx_plan* plan,
                              const float* fine_taps, uint64_t n_fine_floats, const float* weights,
                              uint64_t n_weights, uint32_t audio_capacity, uint32_t tile,
                              uint32_t* id) {
    static const char* who = "rvx_engine_add_vrx";
    if (!e) return rvx::fail(RVX_EINVAL, "%s: null engine", who);
    RVX_ENTER(e->ctx, who);
    if (!plan || !fine_taps || !weights || !id)
        return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    if (grid > e->grids.size())
        return rvx::fail(RVX_EINVAL, "%s: no grid %u (the engine has %zu)", who, grid,
                         e->grids.size() + 1);
    EngineGrid* g = grid ? e->grids[grid - 1].get() : nullptr;
    const uint32_t M = g ? g->cfg.channels : e->cfg.channels;
    const rvx_vrx_plan& p = *plan;
    if (p.placement.channel >= M)
        return rvx::fail(RVX_EINVAL, "%s: channel %u on a %u-channel grid", who,
                         p.placement.channel, M);
    uint32_t row = p.placement.channel;
    if (g) {  // the channel's row, or a free one
        row = g->row_of[p.placement.channel];
        if (row == g->cfg.max_rows) {
            row = uint32_t(std::find(g->users.begin(), g->users.end(), 0u) - g->users.begin());
            if (row == g->cfg.max_rows)
                return rvx::fail(RVX_EINVAL, "%s: grid %u stores %u channels and every one is "
                                 "taken", who, grid, g->cfg.max_rows);
        }
    }
    if (n_fine_floats != 2 * (uint64_t(p.fine.phases) * p.fine.taps + 1))
        return rvx::fail(RVX_EINVAL, "%s: wants %llu fine-tap floats", who,
                         (unsigned long long)(2 * (uint64_t(p.fine.phases) * p.fine.taps + 1)));
    const uint64_t pilot_base = uint64_t(p.demod.audio_taps) + p.demod.dc_taps;
    const uint64_t want_w = p.demod.pilot_taps == 0
                                ? pilot_base
                                : pilot_base + 2ull * p.demod.pilot_taps + 2ull * p.demod.audio_taps;
    if (n_weights != want_w)
        return rvx::fail(RVX_EINVAL, "%s: wants %llu weights", who, (unsigned long long)want_w);
    if (max_tile(p) == 0) return rvx::fail(RVX_EINVAL, "%s: the filters do not fit a tile", who);
    if (p.fine.nco_log2 > 16)
        return rvx::fail(RVX_EINVAL, "%s: NCO table of 2^%u entries (the engine's holds 2^16)",
                         who, p.fine.nco_log2);
    if (audio_capacity == 0) return rvx::fail(RVX_EINVAL, "%s: zero audio capacity", who);
    EngineRx r;
    r.plan = p;
    r.cap = audio_capacity;
    r.tile = tile;
    r.n_weights = n_weights;
    r.carrier = p.carrier.power != RVX_CARRIER_NONE;
    if (r.carrier) {
        if ((uint64_t(1) << p.carrier.nco_log2) != e->nco_entries)
            return rvx::fail(RVX_EINVAL, "%s: carrier NCO table of 2^%u entries (the engine's "
                             "holds %u)", who, p.carrier.nco_log2, e->nco_entries);
        // The ring holds every fine sample one push's audio reads: (cap-1)*R
        // + hist + 1, the history included.
        const uint64_t span = (uint64_t(audio_capacity) - 1) * p.demod.decimation +
                              rvx::demod_fine_history(p.demod) + 1;
        uint64_t cap_f = 2;
        while (cap_f < span) cap_f <<= 1;
        if (cap_f > (uint64_t(1) << 31))
            return rvx::fail(RVX_EINVAL, "%s: a fine ring of %llu samples", who,
                             (unsigned long long)cap_f);
        r.fine_cap = cap_f;
        rvx_vrx_carrier_params probe{};  // the loop's design, checked once here
        if (int32_t rc = rvx_carrier_block(&p, uint32_t(cap_f - 1), 0, 0, 1, 0.0f, &probe))
            return rc;
    }
    int32_t rc = device_alloc(&r.taps, n_fine_floats * sizeof(float), who);
    if (!rc) rc = device_alloc(&r.weights, n_weights * sizeof(float), who);
    if (!rc) rc = device_alloc(&r.audio, uint64_t(audio_capacity) * p.demod.channels * sizeof(float), who);
    if (!rc && r.carrier) rc = device_alloc(&r.fine, r.fine_cap * 2 * sizeof(float), who);
    if (!rc && r.carrier) rc = device_alloc(&r.cstate, RVX_CARRIER_STATE_WORDS * sizeof(uint32_t), who);
    if (!rc && r.carrier) {
        const hipError_t z = hipMemset(r.cstate, 0, RVX_CARRIER_STATE_WORDS * sizeof(uint32_t));
        if (z != hipSuccess) rc = rvx::fail(RVX_EDEVICE, "%s: %s", who, hipGetErrorString(z));
    }
    hipError_t err = hipSuccess;
    if (!rc) {
        err = hipMemcpy(r.taps, fine_taps, n_fine_floats * sizeof(float), hipMemcpyHostToDevice);
        if (err == hipSuccess)
            err = hipMemcpy(r.weights, weights, n_weights * sizeof(float), hipMemcpyHostToDevice);
        if (err != hipSuccess) rc = rvx::fail(RVX_EDEVICE, "%s: %s", who, hipGetErrorString(err));
    }
    if (rc) {
        free_rx(r);
        return rc;
    }
    if (g && g->users[row] == 0) {
        // A row's stream starts at the grid's next block, on cleared slots,
        // so a window reaching below it reads zeros as on a new stream.
        if ((rc = rvx_buf_clear(e->ctx, static_cast<char*>(g->ring) +
                                            uint64_t(row) * g->cfg.ring_pitch * 8,
                                uint64_t(g->cfg.ring_blocks) * 8))) {
            free_rx(r);
            return rc;
        }
        g->row_of[p.placement.channel] = row;
        g->row_first[row] = g->next_block;
        g->rows_dirty = true;
    }
    if (g) ++g->users[row];
    r.grid = grid;
    r.row = row;
    r.id = e->next_id++;
    *id = r.id;
    e->rx.push_back(r);
    return RVX_OK;
}

int32_t rvx_engine_add_vrx(rvx_engine* e, const rvx_vrx_plan* plan, const float* fine_taps,
                           uint64_t n_fine_floats, const float* weights, uint64_t n_weights,
                           uint32_t audio_capacity, uint32_t tile, uint32_t* id) {
    return rvx_engine_add_vrx_on(e, 0, plan, fine_taps, n_fine_floats, weights, n_weights,
                                 audio_capacity, tile, id);
}

int32_t rvx_engine_add_grid(rvx_engine* e, const rvx_engine_grid* config, const float* prototype,
                            uint64_t n_prototype, uint32_t* grid) {
    static const char* who = "rvx_engine_add_grid";
    if (!e) return rvx::fail(RVX_EINVAL, "%s: null engine", who);
    RVX_ENTER(e->ctx, who);
    if (!config || !prototype || !grid) return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    const rvx_engine_grid c = *config;
    const uint32_t M = c.channels, L = c.taps_per_branch, D = c.decimation;
    if (M < 2 || !rvx::pow2(M) || M > 4096 || L == 0 || D == 0 || M % D != 0)
        return rvx::fail(RVX_EINVAL, "%s: invalid grid M=%u L=%u D=%u", who, M, L, D);
    if (n_prototype != uint64_t(M) * L)
        return rvx::fail(RVX_EINVAL, "%s: wants %llu prototype taps", who,
                         (unsigned long long)(uint64_t(M) * L));
    if (!rvx::pow2(c.ring_blocks) || c.max_rows == 0 || c.max_rows > M)
        return rvx::fail(RVX_EINVAL, "%s: ring_blocks must be a power of two and max_rows in "
                         "1..%u", who, M);
    if (c.ring_pitch && (c.ring_pitch < c.ring_blocks || c.ring_pitch % 4))
        return rvx::fail(RVX_EINVAL, "%s: ring pitch %u (at least %u, a multiple of 4)", who,
                         c.ring_pitch, c.ring_blocks);
    auto g = std::make_unique<EngineGrid>();
    g->cfg = c;
    g->cfg.ring_pitch = c.ring_pitch ? c.ring_pitch : c.ring_blocks;
    g->N = uint64_t(M) * L;
    g->proto.assign(prototype, prototype + n_prototype);
    g->max_blocks = uint32_t(std::min<uint64_t>(uint64_t(e->cfg.max_push) / D + 2, 0xFFFFFFFFull));
    if (g->max_blocks > c.ring_blocks)
        return rvx::fail(RVX_EINVAL, "%s: a push can complete %u blocks; the ring holds %u", who,
                         g->max_blocks, c.ring_blocks);
    // After a push the grid keeps at most N - 1 + D samples of history; they
    // and the next push must fit the engine's window and its bounce buffer.
    const uint64_t kept = g->N - 1 + D;
    const uint64_t bounce = e->N + uint64_t(e->max_blocks + 1) * e->cfg.decimation;
    if (kept + e->cfg.max_push > e->win_cap || kept > bounce)
        return rvx::fail(RVX_EINVAL, "%s: the grid keeps %llu samples of history; the engine's "
                         "window holds %llu with a push of %u", who, (unsigned long long)kept,
                         (unsigned long long)e->win_cap, e->cfg.max_push);
    // the first block whose history the window still holds
    g->next_block = (e->win_origin + g->N - 1 + D - 1) / D;
    std::vector<float> tw(2 * size_t(M));
    int32_t rc = rvx_build_twiddles(M, tw.data(), tw.size());
    std::vector<float> scaled(n_prototype);
    const float* upload = prototype;
    if (!rc && e->taps_scaled && rvx_prescale_taps(prototype, n_prototype, scaled.data()) == RVX_OK) {
        upload = scaled.data();
        g->taps_scaled = true;
    }
    const uint64_t ring_bytes = uint64_t(c.max_rows) * g->cfg.ring_pitch * 8;
    if (!rc) rc = device_alloc(&g->taps, n_prototype * sizeof(float), who);
    if (!rc) rc = device_alloc(&g->tw, tw.size() * sizeof(float), who);
    if (!rc) rc = device_alloc(&g->ring, ring_bytes, who);
    if (!rc) rc = device_alloc(&g->rows_d, M * sizeof(uint32_t), who);
    if (!rc && hipHostMalloc((void**)&g->rows_h, M * sizeof(uint32_t), hipHostMallocDefault) !=
                   hipSuccess)
        rc = rvx::fail(RVX_ENOMEM, "%s: pinned row map", who);
    if (rc) return rc;
    g->row_of.assign(M, c.max_rows);
    g->users.assign(c.max_rows, 0);
    g->row_first.assign(c.max_rows, 0);
    hipError_t err = hipMemcpy(g->taps, upload, n_prototype * sizeof(float), hipMemcpyHostToDevice);
    if (err == hipSuccess)
        err = hipMemcpy(g->tw, tw.data(), tw.size() * sizeof(float), hipMemcpyHostToDevice);
    if (err == hipSuccess)
        err = hipMemcpy(g->rows_d, g->row_of.data(), M * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (err == hipSuccess) err = hipMemset(g->ring, 0, ring_bytes);
    if (err != hipSuccess)
        return rvx::fail(RVX_EDEVICE, "%s: upload failed: %s", who, hipGetErrorString(err));
    e->grids.push_back(std::move(g));
    *grid = uint32_t(e->grids.size());
    return RVX_OK;
}

// The conditioned path's buffers, before the first push: the sample ring (at
// least `ring` samples, grown if a stage needs more), the cleaned window and
// its bounce buffer; the chains' taps go back to unscaled, since they now
// read converted samples.
static int32_t condition_setup(rvx_engine& e, uint64_t ring, const char* who) {
    if (e.stats.samples_in)
        return rvx::fail(RVX_EINVAL, "%s: set before the first push (the chains' input changes "
                         "from codes to the conditioned samples)", who);
    const uint64_t bounce = e.N + uint64_t(e.max_blocks + 1) * e.cfg.decimation;
    if (!e.cwin) {
        if (int32_t rc = device_alloc(&e.cwin, e.win_cap * 8, who)) return rc;
        if (int32_t rc = device_alloc(&e.cbounce, bounce * 8, who)) return rc;
    }
    if (e.xring_n < ring) {
        if (e.xring) (void)hipFree(e.xring);
        e.xring = nullptr;
        if (int32_t rc = device_alloc(&e.xring, ring * 8, who)) return rc;
        if (e.comb_removed) {
            (void)hipFree(e.comb_removed);
            e.comb_removed = nullptr;
            if (int32_t rc = device_alloc(&e.comb_removed, ring * 8, who)) return rc;
        }
        e.xring_n = ring;
        e.xmask = uint32_t(ring - 1);
    }
    if (e.taps_scaled) {
        RVX_TRY_HIP(hipMemcpy(e.taps, e.proto.data(), e.proto.size() * 4, hipMemcpyHostToDevice),
                    who);
        e.taps_scaled = false;
    }
    for (auto& g : e.grids)
        if (g->taps_scaled) {
            RVX_TRY_HIP(hipMemcpy(g->taps, g->proto.data(), g->proto.size() * 4,
                                  hipMemcpyHostToDevice),
                        who);
            g->taps_scaled = false;
        }
    e.cleaned_end = e.win_origin;
    e.condition = true;
    return RVX_OK;
}

int32_t rvx_engine_set_excision(rvx_engine* e, const rvx_excision_config* config,
                                float threshold) {
    static const char* who = "rvx_engine_set_excision";
    if (!e || !config) return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    RVX_ENTER(e->ctx, who);
    if (e->excise) return rvx::fail(RVX_EINVAL, "%s: excision is already on", who);
    if (e->stats.samples_in)
        return rvx::fail(RVX_EINVAL, "%s: set before the first push (the chains' input changes "
                         "from codes to the cleaned samples)", who);
    const rvx_excision_config c = *config;
    // the sample ring: a push, the samples still waiting for their scan or
    // their segment, and what the passes' rings need beside them
    const uint64_t segment = uint64_t(1) << std::min(c.seg_log2, 31u);
    const uint64_t lag = std::max<uint64_t>(uint64_t(c.max_width) + c.lead, segment);
    if (lag > e->cfg.max_push)  // the blocks waiting behind the excision must fit the window
        return rvx::fail(RVX_EINVAL, "%s: the excision trails the stream by %llu samples, more "
                         "than a push of %u", who, (unsigned long long)lag, e->cfg.max_push);
    uint64_t ring = 1;
    while (ring < uint64_t(e->cfg.max_push) + lag + segment) ring <<= 1;
    ring = std::max(ring, e->xring_n);
    rvx_excision_rings rings{};
    for (;; ring <<= 1) {
        if (ring > (uint64_t(1) << 31))
            return rvx::fail(RVX_EINVAL, "%s: no sample ring up to 2^31 fits a push of %u", who,
                             e->cfg.max_push);
        if (rvx_size_excision_rings(&c, e->cfg.max_push, 1, uint32_t(ring - 1), &rings) == RVX_OK)
            break;
    }
    for (uint32_t pass : {0u, 2u}) {  // the config and threshold as the passes will check them
        rvx_excision_config probe_c = c;
        probe_c.pass = pass;
        rvx_excision_params probe{};
        probe.ring_mask = uint32_t(ring - 1);
        probe.count = 1;
        probe.seg_mask = rings.seg_mask;
        probe.flag_mask = rings.flag_mask;
        probe.removed_mask = rings.removed_mask;
        probe.threshold = threshold;
        if (int32_t rc = rvx_excision_check(&probe_c, &probe)) return rc;
    }
    if (int32_t rc = condition_setup(*e, ring, who)) return rc;
    int32_t rc = device_alloc(&e->xseg_power, (uint64_t(rings.seg_mask) + 1) * 4, who);
    if (!rc) rc = device_alloc(&e->xseg_level, (uint64_t(rings.seg_mask) + 1) * 4, who);
    if (!rc) rc = device_alloc(&e->xflags, (uint64_t(rings.flag_mask) + 1) * 4, who);
    if (!rc) rc = device_alloc(&e->xremoved, (uint64_t(rings.removed_mask) + 1) * 8, who);
    if (!rc) rc = device_alloc(&e->xcounters, RVX_EXCISE_COUNTERS * 4, who);
    if (!rc) rc = rvx_excision_cursor_create(&c, e->xmask, &rings, threshold, &e->xcursor);
    if (rc) return rc;  // what was allocated is freed with the engine
    RVX_TRY_HIP(hipMemset(e->xcounters, 0, RVX_EXCISE_COUNTERS * 4), who);
    e->xcfg = c;
    e->xrings = rings;
    e->excise = true;
    return RVX_OK;
}

int32_t rvx_engine_set_comb(rvx_engine* e, const rvx_comb_line* lines, uint32_t n_lines,
                            uint64_t phase_index, int32_t subtract) {
    static const char* who = "rvx_engine_set_comb";
    if (!e || (n_lines && !lines)) return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    RVX_ENTER(e->ctx, who);
    if (n_lines > RVX_COMB_MAX_LINES)
        return rvx::fail(RVX_EINVAL, "%s: %u lines is more than the %u the stage cancels", who,
                         n_lines, RVX_COMB_MAX_LINES);
    for (uint32_t k = 0; k < n_lines; ++k)
        if (!std::isfinite(lines[k].amplitude[0]) || !std::isfinite(lines[k].amplitude[1]))
            return rvx::fail(RVX_EINVAL, "%s: line %u's amplitude is not finite", who, k);
    if (!e->comb_on) {  // the first call opens the stage: before the first push
        uint64_t ring = 1;
        while (ring < uint64_t(e->cfg.max_push)) ring <<= 1;
        if (int32_t rc = condition_setup(*e, std::max(ring, e->xring_n), who)) return rc;
        const uint64_t table_n = uint64_t(1) << RVX_COMB_TABLE_LOG2;
        std::vector<float> table(2 * table_n);
        int32_t rc = rvx_comb_table(table.data(), table.size());
        if (!rc) rc = device_alloc(&e->comb_d, RVX_COMB_MAX_LINES * sizeof(rvx_comb_line), who);
        if (!rc) rc = device_alloc(&e->comb_table, table.size() * 4, who);
        if (!rc) rc = device_alloc(&e->comb_removed, e->xring_n * 8, who);
        if (!rc)
            rc = device_alloc(&e->comb_proj,
                              uint64_t(RVX_COMB_MAX_LINES) * RVX_COMB_MAX_SUB_BLOCKS * 8, who);
        if (rc) return rc;
        RVX_TRY_HIP(hipHostMalloc((void**)&e->comb_h, RVX_COMB_MAX_LINES * sizeof(rvx_comb_line),
                                  hipHostMallocDefault),
                    who);
        RVX_TRY_HIP(hipMemcpy(e->comb_table, table.data(), table.size() * 4,
                              hipMemcpyHostToDevice),
                    who);
        RVX_TRY_HIP(hipEventCreateWithFlags(&e->comb_staged, hipEventDisableTiming), who);
        RVX_TRY_HIP(hipEventRecord(e->comb_staged, e->ctx->stream), who);
        e->comb_on = true;
    }
    e->comb_lines.assign(lines, lines + n_lines);
    e->comb_index = phase_index;
    e->comb_subtract = subtract != 0;
    return RVX_OK;
}

int32_t rvx_engine_comb_projections(rvx_engine* e, float* out, uint64_t n_floats,
                                    uint64_t* first_sample, uint64_t* samples, uint32_t* lines,
                                    uint32_t* sub_blocks) {
    static const char* who = "rvx_engine_comb_projections";
    if (!e || !lines || !sub_blocks) return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    RVX_ENTER(e->ctx, who);
    if (!e->comb_on) return rvx::fail(RVX_EINVAL, "%s: the comb stage is off", who);
    *lines = e->proj_lines;
    *sub_blocks = e->proj_subs;
    if (first_sample) *first_sample = e->proj_start;
    if (samples) *samples = e->proj_count;
    const uint64_t values = uint64_t(e->proj_lines) * e->proj_subs;
    if (values == 0) return RVX_OK;
    if (!out || n_floats < 2 * values)
        return rvx::fail(RVX_EINVAL, "%s: wants %llu floats", who,
                         (unsigned long long)(2 * values));
    RVX_TRY_HIP(hipStreamSynchronize(e->ctx->stream), who);
    RVX_TRY_HIP(hipMemcpy(out, e->comb_proj, 2 * values * 4, hipMemcpyDeviceToHost), who);
    return RVX_OK;
}

int32_t rvx_engine_excision_counters(rvx_engine* e, uint32_t* counters) {
    static const char* who = "rvx_engine_excision_counters";
    if (!e || !counters) return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    RVX_ENTER(e->ctx, who);
    if (!e->excise) return rvx::fail(RVX_EINVAL, "%s: excision is off", who);
    RVX_TRY_HIP(hipStreamSynchronize(e->ctx->stream), who);
    RVX_TRY_HIP(hipMemcpy(counters, e->xcounters, RVX_EXCISE_COUNTERS * 4, hipMemcpyDeviceToHost),
                who);
    return RVX_OK;
}

const void* rvx_engine_grid_ring(rvx_engine* e, uint32_t grid) {
    if (!e || grid == 0 || grid > e->grids.size()) return e && grid == 0 ? e->ring : nullptr;
    return e->grids[grid - 1]->ring;
}

int32_t rvx_engine_grid_row(rvx_engine* e, uint32_t grid, uint32_t channel, uint32_t* row,
                            uint64_t* first_block) {
    static const char* who = "rvx_engine_grid_row";
    if (!e || !row) return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    if (grid == 0) {
        if (channel >= e->cfg.channels)
            return rvx::fail(RVX_EINVAL, "%s: channel %u on a %u-channel grid", who, channel,
                             e->cfg.channels);
        *row = channel;
        if (first_block) *first_block = e->first_block;
        return RVX_OK;
    }
    if (grid > e->grids.size()) return rvx::fail(RVX_EINVAL, "%s: no grid %u", who, grid);
    const EngineGrid& g = *e->grids[grid - 1];
    if (channel >= g.cfg.channels || g.row_of[channel] == g.cfg.max_rows)
        return rvx::fail(RVX_EINVAL, "%s: channel %u is not stored on grid %u", who, channel, grid);
    *row = g.row_of[channel];
    if (first_block) *first_block = g.row_first[*row];
    return RVX_OK;
}

int32_t rvx_engine_remove_vrx(rvx_engine* e, uint32_t id) {
    static const char* who = "rvx_engine_remove_vrx";
    if (!e) return rvx::fail(RVX_EINVAL, "%s: null engine", who);
    RVX_ENTER(e->ctx, who);
    auto it = std::find_if(e->rx.begin(), e->rx.end(), [&](const EngineRx& r) { return r.id == id; });
    if (it == e->rx.end()) return rvx::fail(RVX_EINVAL, "%s: no receiver %u", who, id);
    RVX_TRY_HIP(hipStreamSynchronize(e->ctx->stream), who);
    if (it->grid) {
        EngineGrid& g = *e->grids[it->grid - 1];
        if (--g.users[it->row] == 0) {
            g.row_of[it->plan.placement.channel] = g.cfg.max_rows;
            g.rows_dirty = true;
        }
    }
    free_rx(*it);
    e->rx.erase(it);
    return RVX_OK;
}

int32_t rvx_engine_set_display(rvx_engine* e, uint32_t id, uint32_t display_ring,
                               uint32_t transform) {
    static const char* who = "rvx_engine_set_display";
    if (!e) return rvx::fail(RVX_EINVAL, "%s: null engine", who);
    RVX_ENTER(e->ctx, who);
    auto it = std::find_if(e->rx.begin(), e->rx.end(), [&](const EngineRx& r) { return r.id == id; });
    if (it == e->rx.end()) return rvx::fail(RVX_EINVAL, "%s: no receiver %u", who, id);
    EngineRx& r = *it;
    rvx_vrx_display_plan plan{};
    std::vector<float> taps(2 * (RVX_DISPLAY_TAPS + 1));
    if (display_ring) {
        if (!rvx::pow2(display_ring) || !rvx::pow2(transform) || transform < 4 ||
            transform > display_ring)
            return rvx::fail(RVX_EINVAL, "%s: a display ring of %u and a transform of %u (powers "
                             "of two, the ring holding the transform)", who, display_ring,
                             transform);
        if (int32_t rc = rvx_plan_vrx_display(&r.plan, &plan, taps.data(), taps.size()))
            return rc;
        if ((uint64_t(1) << plan.fine.nco_log2) != e->nco_entries)
            return rvx::fail(RVX_EINVAL, "%s: display NCO of 2^%u entries", who,
                             plan.fine.nco_log2);
    }
    RVX_TRY_HIP(hipStreamSynchronize(e->ctx->stream), who);
    for (void** q : {&r.display_taps, &r.display_ring, &r.pane_tw, (void**)&r.pane_window,
                     (void**)&r.pane})
        if (*q) {
            (void)hipFree(*q);
            *q = nullptr;
        }
    r.display_on = false;
    r.display_started = false;
    if (!display_ring) return RVX_OK;
    std::vector<float> tw(2ull * transform), win(transform + transform / 2);
    if (int32_t rc = rvx_build_twiddles(transform, tw.data(), tw.size())) return rc;
    if (int32_t rc = rvx_build_spectrum_window(transform, e->cfg.taps_per_branch, 120.0,
                                               win.data(), win.size()))
        return rc;
    // the pane has no channel seam: unity in place of the droop correction
    for (uint32_t b = 0; b < transform / 2; ++b) win[transform + b] = 1.0f;
    int32_t rc = device_alloc(&r.display_taps, taps.size() * sizeof(float), who);
    if (!rc) rc = device_alloc(&r.display_ring, 2ull * display_ring * 8, who);
    if (!rc) rc = device_alloc(&r.pane_tw, tw.size() * sizeof(float), who);
    if (!rc) rc = device_alloc(&r.pane_window, win.size() * sizeof(float), who);
    if (!rc) rc = device_alloc(&r.pane, uint64_t(transform) * sizeof(float), who);
    if (rc) return rc;
    RVX_TRY_HIP(hipMemset(r.display_ring, 0, 2ull * display_ring * 8), who);
    RVX_TRY_HIP(hipMemcpy(r.display_taps, taps.data(), taps.size() * sizeof(float),
                          hipMemcpyHostToDevice),
                who);
    RVX_TRY_HIP(hipMemcpy(r.pane_tw, tw.data(), tw.size() * sizeof(float), hipMemcpyHostToDevice),
                who);
    RVX_TRY_HIP(hipMemcpy(r.pane_window, win.data(), win.size() * sizeof(float),
                          hipMemcpyHostToDevice),
                who);
    r.display = plan;
    r.display_cap = display_ring;
    r.transform = transform;
    r.display_on = true;
    return RVX_OK;
}

int32_t rvx_engine_passband(rvx_engine* e, uint32_t id, float* out, uint64_t n_out,
                            uint64_t* end_output, int64_t* display_rate) {
    static const char* who = "rvx_engine_passband";
    if (!e) return rvx::fail(RVX_EINVAL, "%s: null engine", who);
    RVX_ENTER(e->ctx, who);
    auto it = std::find_if(e->rx.begin(), e->rx.end(), [&](const EngineRx& r) { return r.id == id; });
    if (it == e->rx.end()) return rvx::fail(RVX_EINVAL, "%s: no receiver %u", who, id);
    EngineRx& r = *it;
    if (!r.display_on) return rvx::fail(RVX_EINVAL, "%s: receiver %u has no display tap", who, id);
    const uint32_t n = r.transform;
    if (!out || n_out < n / 2) return rvx::fail(RVX_EINVAL, "%s: the pane is %u floats", who, n / 2);
    if (!r.display_started || r.display_next - r.display_from < n)
        return rvx::fail(RVX_EAGAIN, "%s: fewer than %u display samples so far", who, n);
    const uint64_t first = r.display_next - n;
    rvx_spectrum_params p{};
    p.channels = 2;
    p.transform = n;
    p.stages = rvx::log2u(n);
    p.chan_blocks = uint32_t(r.display_cap);
    p.chan_mask = uint32_t(r.display_cap - 1);
    p.in_offset = uint32_t(first & (r.display_cap - 1));
    if (int32_t rc = rvx_spectrum(e->ctx, &p, r.display_ring, 2 * r.display_cap, r.pane_tw, n,
                                  r.pane_window, n + n / 2, r.pane, n))
        return rc;
    // channel 0's central half is the frame's upper half
    RVX_TRY_HIP(hipMemcpyAsync(out, r.pane + n / 2, (n / 2) * sizeof(float),
                               hipMemcpyDeviceToHost, e->ctx->stream),
                who);
    RVX_TRY_HIP(hipStreamSynchronize(e->ctx->stream), who);
    if (end_output) *end_output = r.display_next;
    if (display_rate) *display_rate = r.display.display_rate;
    return RVX_OK;
}

int32_t rvx_engine_set_blanker(rvx_engine* e, uint32_t id, const rvx_blanker_config* config,
                               float threshold) {
    static const char* who = "rvx_engine_set_blanker";
    if (!e) return rvx::fail(RVX_EINVAL, "%s: null engine", who);
    RVX_ENTER(e->ctx, who);
    auto it = std::find_if(e->rx.begin(), e->rx.end(), [&](const EngineRx& r) { return r.id == id; });
    if (it == e->rx.end()) return rvx::fail(RVX_EINVAL, "%s: no receiver %u", who, id);
    EngineRx& r = *it;
    const Source src = source_of(*e, r);
    if (config) {
        rvx_blanker_params probe{};
        probe.chan_mask = probe.out_mask = src.blocks - 1;
        probe.count = 1;
        probe.threshold = threshold;
        rvx_blanker_config c = *config;
        c.pass = RVX_BLANK_DETECT;
        if (int32_t rc = rvx_blanker_check(&c, &probe)) return rc;
        // a push's new blocks, the window below them and the lead above must
        // all still be in the ring when the blanker reads them
        const uint64_t reach = uint64_t(c.hang) + c.guard + c.window + c.lead;
        if (uint64_t(src.max_blocks) + reach >= src.blocks)
            return rvx::fail(RVX_EINVAL, "%s: a push of up to %u blocks plus the blanker's reach "
                             "of %llu does not fit a %u-block ring", who, src.max_blocks,
                             (unsigned long long)reach, src.blocks);
    }
    RVX_TRY_HIP(hipStreamSynchronize(e->ctx->stream), who);
    for (void** q : {(void**)&r.flags, &r.blanked})
        if (*q) {
            (void)hipFree(*q);
            *q = nullptr;
        }
    r.blank_on = false;
    r.blank_started = false;
    if (config) {
        const uint64_t n = src.blocks;
        if (int32_t rc = device_alloc(&r.flags, n * sizeof(float), who)) return rc;
        if (int32_t rc = device_alloc(&r.blanked, n * 8, who)) return rc;
        RVX_TRY_HIP(hipMemset(r.flags, 0, n * sizeof(float)), who);  // no flags yet
        RVX_TRY_HIP(hipMemset(r.blanked, 0, n * 8), who);
        r.blank_cfg = *config;
        r.blank_threshold = threshold;
        r.blank_on = true;
        // the fine stage changes source: its audio restarts from what is blanked
        r.started = false;
        r.fine_started = false;
    }
    return RVX_OK;
}

int32_t rvx_engine_set_noise(rvx_engine* e, uint32_t id, const rvx_line_config* line_config,
                             const rvx_line_params* line, uint32_t spectral_frame,
                             const rvx_spectral_params* spectral) {
    static const char* who = "rvx_engine_set_noise";
    if (!e) return rvx::fail(RVX_EINVAL, "%s: null engine", who);
    RVX_ENTER(e->ctx, who);
    auto it = std::find_if(e->rx.begin(), e->rx.end(), [&](const EngineRx& r) { return r.id == id; });
    if (it == e->rx.end()) return rvx::fail(RVX_EINVAL, "%s: no receiver %u", who, id);
    EngineRx& r = *it;
    if ((line || spectral) && r.plan.demod.channels != 1)
        return rvx::fail(RVX_EINVAL, "%s: receiver %u hands out %u channels; the audio stages "
                         "run on one channel of audio", who, id, r.plan.demod.channels);
    if (line) {
        if (!line_config) return rvx::fail(RVX_EINVAL, "%s: a line kernel needs its config", who);
        if (int32_t rc = rvx_line_check(line_config, line)) return rc;
    }
    if (spectral) {
        if (int32_t rc = rvx_spectral_check(spectral_frame, spectral)) return rc;
    }
    // the stages' state restarts: wait for the launches that use the old one
    RVX_TRY_HIP(hipStreamSynchronize(e->ctx->stream), who);
    for (float** q : {&r.line_state, &r.spec_state, &r.spec_tables})
        if (*q) {
            (void)hipFree(*q);
            *q = nullptr;
        }
    r.line_on = r.spectral_on = false;
    if (line) {
        const uint64_t n = rvx_line_state_size(line_config);
        if (int32_t rc = device_alloc(&r.line_state, n * sizeof(float), who)) return rc;
        RVX_TRY_HIP(hipMemset(r.line_state, 0, n * sizeof(float)), who);
        r.line_cfg = *line_config;
        r.line = *line;
        r.line_on = true;
    }
    if (spectral) {
        const uint64_t n = rvx_spectral_state_size(spectral_frame);
        std::vector<float> tables(4ull * spectral_frame);
        if (int32_t rc = rvx_design_spectral_tables(spectral_frame, tables.data(), tables.size()))
            return rc;
        if (int32_t rc = device_alloc(&r.spec_state, n * sizeof(float), who)) return rc;
        if (int32_t rc = device_alloc(&r.spec_tables, tables.size() * sizeof(float), who))
            return rc;
        RVX_TRY_HIP(hipMemset(r.spec_state, 0, n * sizeof(float)), who);
        RVX_TRY_HIP(hipMemcpy(r.spec_tables, tables.data(), tables.size() * sizeof(float),
                              hipMemcpyHostToDevice),
                    who);
        r.spec_frame = spectral_frame;
        r.spec = *spectral;
        r.spectral_on = true;
    }
    return RVX_OK;
}

static int32_t main_push(rvx_engine* e, uint64_t count, hipEvent_t fork, const char* who);

// The conditioned path over one push, in the reference's record order
// (graph.cpp: the converted block, its correction, record_comb_cancel :1715,
// record_impulse_excision :1490): the push's samples converted with the
// correction they arrived under into the sample ring, the comb's lines
// projected and subtracted, the excision cursor's four passes, and the samples
// finished now copied into the cleaned window.
static int32_t condition_push(rvx_engine& e, uint64_t push_first, uint64_t samples,
                              const char* who) {
    const uint64_t ring = e.xring_n;
    if (int32_t rc = rvx::convert_to_ring(e.ctx, e.cfg.format, e.win, e.win_fill,
                                          push_first - e.win_origin, push_first,
                                          uint32_t(samples), e.corr_on ? &e.corr_now : nullptr,
                                          e.xring, e.xmask, who))
        return rc;
    if (e.comb_on && !e.comb_lines.empty()) {
        const uint32_t n = uint32_t(e.comb_lines.size());
        // each line's phase at the push's first sample: phase0 + step * (j0 - index), 0.32
        const uint64_t since = push_first - e.comb_index;
        RVX_TRY_HIP(hipEventSynchronize(e.comb_staged), who);  // the last upload has been read
        for (uint32_t k = 0; k < n; ++k) {
            rvx_comb_line l = e.comb_lines[k];
            l.phase0 = uint32_t(l.phase0 + uint64_t(l.step) * (since & 0xFFFFFFFFull));
            e.comb_h[k] = l;
        }
        RVX_TRY_HIP(hipMemcpyAsync(e.comb_d, e.comb_h, n * sizeof(rvx_comb_line),
                                   hipMemcpyHostToDevice, e.ctx->stream),
                    who);
        RVX_TRY_HIP(hipEventRecord(e.comb_staged, e.ctx->stream), who);
        rvx_comb_params p{};
        p.ring_mask = e.xmask;
        p.first = uint32_t(push_first & e.xmask);
        p.count = uint32_t(samples);
        p.lines = n;
        // sub-blocks: the largest power of two up to 16 dividing the push (a
        // projection's workgroup sums its sub-block in order)
        p.sub_blocks = std::min<uint32_t>(uint32_t(samples & (~samples + 1)),
                                          RVX_COMB_MAX_SUB_BLOCKS);
        p.removed_mask = e.xmask;
        const auto* lines = static_cast<const rvx_comb_line*>(e.comb_d);
        if (int32_t rc = rvx_comb_project(e.ctx, &p, e.xring, ring, lines, n, e.comb_table,
                                          uint64_t(1) << RVX_COMB_TABLE_LOG2, e.comb_proj,
                                          uint64_t(n) * p.sub_blocks))
            return rc;
        e.proj_start = push_first;
        e.proj_count = samples;
        e.proj_lines = n;
        e.proj_subs = p.sub_blocks;
        if (e.comb_subtract)
            if (int32_t rc = rvx_comb_subtract(e.ctx, &p, e.xring, ring, lines, n, e.comb_table,
                                               uint64_t(1) << RVX_COMB_TABLE_LOG2, e.xring, ring,
                                               e.comb_removed, ring))
                return rc;
    }
    uint64_t finished = push_first + samples;
    if (e.excise) {
        rvx_excision_plan plan{};
        if (int32_t rc = rvx_excision_cursor_plan(e.xcursor, push_first, uint32_t(samples), &plan))
            return rc;
        const uint64_t seg = ring >> e.xcfg.seg_log2;
        const uint64_t flags = uint64_t(e.xrings.flag_mask) + 1;
        const uint64_t removed = uint64_t(e.xrings.removed_mask) + 1;
        if (plan.segment.count)
            if (int32_t rc = rvx_excise_segment(e.ctx, &e.xcfg, &plan.segment, e.xring, ring,
                                                e.xseg_power, seg))
                return rc;
        if (plan.reference.count)
            if (int32_t rc = rvx_excise_reference(e.ctx, &e.xcfg, &plan.reference, e.xseg_power,
                                                  e.xseg_level, seg))
                return rc;
        if (plan.flag.count)
            if (int32_t rc = rvx_excise_flag(e.ctx, &e.xcfg, &plan.flag, e.xring, ring,
                                             e.xseg_level, seg, e.xflags, flags))
                return rc;
        if (plan.apply.count)
            if (int32_t rc = rvx_excise_apply(e.ctx, &e.xcfg, &plan.apply, e.xring, ring,
                                              e.xflags, flags, e.xring, ring, e.xremoved,
                                              removed, e.xcounters))
                return rc;
        finished = plan.cleaned_end;
    }
    if (finished > e.cleaned_end) {
        const uint64_t n = finished - e.cleaned_end;
        if (int32_t rc = rvx::ring_to_window(e.ctx, e.xring, e.xmask, e.cleaned_end, uint32_t(n),
                                             e.cwin, e.win_cap, e.cleaned_end - e.win_origin, who))
            return rc;
        e.cleaned_end = finished;
    }
    return RVX_OK;
}

// The complete blocks of an added grid (a push's share), stored into the
// rows its receivers hold; with no row held the blocks are passed over.
static int32_t grid_push(rvx_engine& e, EngineGrid& g, uint64_t stop, bool* ran,
                         const char* who) {
    const uint32_t D = g.cfg.decimation, M = g.cfg.channels;
    if (stop == 0 || (stop - 1) / D < g.next_block) return RVX_OK;
    const uint64_t count = std::min<uint64_t>((stop - 1) / D - g.next_block + 1, g.max_blocks);
    const bool held = std::any_of(g.users.begin(), g.users.end(), [](uint32_t u) { return u; });
    if (held) {
        const hipStream_t s = e.ctx->stream;
        if (g.rows_dirty) {
            RVX_TRY_HIP(hipStreamSynchronize(s), who);  // the last upload has been read
            std::memcpy(g.rows_h, g.row_of.data(), M * sizeof(uint32_t));
            RVX_TRY_HIP(hipMemcpyAsync(g.rows_d, g.rows_h, M * sizeof(uint32_t),
                                       hipMemcpyHostToDevice, s),
                        who);
            g.rows_dirty = false;
        }
        rvx_chain_desc d{};
        d.format = e.condition ? RVX_FORMAT_CF32 : e.cfg.format;
        d.channels = M;
        d.taps_per_branch = g.cfg.taps_per_branch;
        d.decimation = D;
        d.input_origin = e.win_origin;
        d.input_samples = e.readable_end() - e.win_origin;
        d.first_block = g.next_block;
        d.block_count = uint32_t(count);
        d.out_ring_blocks = g.cfg.ring_blocks;
        d.ring_pitch = g.cfg.ring_pitch;
        d.ring_rows = g.cfg.max_rows;
        d.channel_rows = g.rows_d;
        d.taps_scaled = g.taps_scaled ? 1u : 0u;
        if (e.corr_on && !e.condition) {
            d.correction = &e.corr_now;
            d.correction_before = &e.corr_before;
            d.correction_from = e.corr_from;
        }
        if (int32_t rc = rvx_chain(e.ctx, &d, e.condition ? e.cwin : e.win, g.taps, g.N, g.tw, M,
                                   g.ring,
                                   uint64_t(g.cfg.max_rows) * g.cfg.ring_pitch, nullptr, 0))
            return rc;
        e.stats.grid_blocks += count;
        *ran = true;
    }
    g.next_block += count;
    return RVX_OK;
}

int32_t rvx_engine_push(rvx_engine* e, const void* codes, uint64_t samples, int32_t on_device) {
    static const char* who = "rvx_engine_push";
    if (!e) return rvx::fail(RVX_EINVAL, "%s: null engine", who);
    RVX_ENTER(e->ctx, who);
    if (samples == 0) return RVX_OK;
    if (!codes) return rvx::fail(RVX_EINVAL, "%s: null samples", who);
    if (samples > e->cfg.max_push)
        return rvx::fail(RVX_EINVAL, "%s: %llu samples, the engine takes at most %u per push", who,
                         (unsigned long long)samples, e->cfg.max_push);
    const hipStream_t s = e->ctx->stream;
    const uint32_t D = e->cfg.decimation, M = e->cfg.channels;
    const uint64_t end = e->win_origin + e->win_fill;
    if (e->win_fill + samples > e->win_cap) {
        // Keep the history of the next block; move it to the window's front
        // through the bounce buffer (source and destination may overlap).
        const uint64_t need_from = e->oldest_needed();
        const uint64_t keep_from = std::min(std::max(need_from, e->win_origin), end);
        const uint64_t kept = end - keep_from;
        if (kept) {
            const uint64_t off = (keep_from - e->win_origin) * e->bps;
            RVX_TRY_HIP(hipMemcpyAsync(e->bounce, e->win + off, kept * e->bps,
                                       hipMemcpyDeviceToDevice, s),
                        who);
            RVX_TRY_HIP(hipMemcpyAsync(e->win, e->bounce, kept * e->bps, hipMemcpyDeviceToDevice, s),
                        who);
        }
        if (e->condition) {  // the cleaned samples kept move with the codes
            const uint64_t kept_c = e->cleaned_end > keep_from ? e->cleaned_end - keep_from : 0;
            if (kept_c) {
                const uint64_t off = (keep_from - e->win_origin) * 8;
                RVX_TRY_HIP(hipMemcpyAsync(e->cbounce, static_cast<uint8_t*>(e->cwin) + off,
                                           kept_c * 8, hipMemcpyDeviceToDevice, s),
                            who);
                RVX_TRY_HIP(hipMemcpyAsync(e->cwin, e->cbounce, kept_c * 8,
                                           hipMemcpyDeviceToDevice, s),
                            who);
            }
        }
        e->win_origin = keep_from;
        e->win_fill = kept;
    }
    if (on_device) {
        RVX_TRY_HIP(hipMemcpyAsync(e->win + e->win_fill * e->bps, codes, samples * e->bps,
                                   hipMemcpyDeviceToDevice, s),
                    who);
    } else {
        // Host samples go through pinned staging, so the caller's buffer is
        // free on return and the upload is asynchronous.
        RVX_TRY_HIP(hipEventSynchronize(e->pushed), who);
        std::memcpy(e->stage_h, codes, samples * e->bps);
        RVX_TRY_HIP(hipMemcpyAsync(e->win + e->win_fill * e->bps, e->stage_h, samples * e->bps,
                                   hipMemcpyHostToDevice, s),
                    who);
        RVX_TRY_HIP(hipEventRecord(e->pushed, s), who);
    }
    const uint64_t push_first = e->win_origin + e->win_fill;  // absolute index
    e->win_fill += samples;
    e->stats.samples_in += samples;
    if (e->cfg.calibration) {
        if (int32_t rc = front_end_push(*e, push_first, samples, who)) return rc;
    }
    if (e->condition) {
        if (int32_t rc = condition_push(*e, push_first, samples, who)) return rc;
    }

    const uint64_t stop = e->readable_end();  // one past the last sample a chain may read
    bool ran = false;
    uint64_t count = 0;
    if (stop > 0 && (stop - 1) / D >= e->next_block) {
        count = (stop - 1) / D - e->next_block + 1;
        const uint32_t unit = e->cfg.frame_blocks ? e->cfg.frame_blocks * e->cfg.row_frames : 1;
        count -= count % unit;
        count = std::min<uint64_t>(count, e->max_blocks);
    }
    const bool added = !e->grids.empty();
    if (added && !e->aux) {  // all three or none, so a failed push leaves nothing half made
        hipStream_t aux = nullptr;
        hipEvent_t fork = nullptr, join = nullptr;
        hipError_t err = hipStreamCreateWithFlags(&aux, hipStreamNonBlocking);
        if (err == hipSuccess) err = hipEventCreateWithFlags(&fork, hipEventDisableTiming);
        if (err == hipSuccess) err = hipEventCreateWithFlags(&join, hipEventDisableTiming);
        if (err != hipSuccess) {
            if (join) (void)hipEventDestroy(join);
            if (fork) (void)hipEventDestroy(fork);
            if (aux) (void)hipStreamDestroy(aux);
            return rvx::fail(RVX_EDEVICE, "%s: second stream: %s", who, hipGetErrorString(err));
        }
        e->aux = aux;
        e->fork = fork;
        e->join = join;
    }
    if (count) {
        if (int32_t rc = main_push(e, count, added ? e->fork : nullptr, who)) return rc;
        ran = true;
    }
    if (added) {
        if (!count) RVX_TRY_HIP(hipEventRecord(e->fork, s), who);
        RVX_TRY_HIP(hipStreamWaitEvent(e->aux, e->fork, 0), who);
        // The context's calls are serialised by its caller, so it can carry
        // the second stream for the added grids' launches.
        e->ctx->stream = e->aux;
        int32_t rc = RVX_OK;
        for (auto& g : e->grids)
            if ((rc = grid_push(*e, *g, stop, &ran, who))) break;
        e->ctx->stream = s;
        // joined even after an error: later work on the engine's stream stays
        // behind whatever the second stream was given
        RVX_TRY_HIP(hipEventRecord(e->join, e->aux), who);
        RVX_TRY_HIP(hipStreamWaitEvent(s, e->join, 0), who);
        if (rc) return rc;
    }
    if (!ran) return RVX_OK;
    return engine_receivers(*e, who);
}

// The engine's own grid over `count` complete blocks: the chain, the
// averaged spectrum's rows, and the block cursor. `fork` (if any) is recorded
// between the chain and the rows.
static int32_t main_push(rvx_engine* e, uint64_t count, hipEvent_t fork, const char* who) {
    const uint32_t D = e->cfg.decimation, M = e->cfg.channels;

    rvx_chain_desc d{};
    d.format = e->condition ? RVX_FORMAT_CF32 : e->cfg.format;
    d.channels = M;
    d.taps_per_branch = e->cfg.taps_per_branch;
    d.decimation = D;
    d.input_origin = e->win_origin;
    d.input_samples = e->readable_end() - e->win_origin;
    d.first_block = e->next_block;
    d.block_count = uint32_t(count);
    d.out_ring_blocks = e->cfg.ring_blocks;
    d.ring_pitch = e->cfg.ring_pitch;
    d.frame_blocks = e->cfg.frame_blocks;
    d.alpha = e->cfg.alpha;
    d.taps_scaled = e->taps_scaled ? 1u : 0u;
    if (e->corr_on && !e->condition) {
        d.correction = &e->corr_now;
        d.correction_before = &e->corr_before;
        d.correction_from = e->corr_from;
    }
    const uint64_t frames = e->cfg.frame_blocks ? count / e->cfg.frame_blocks : 0;
    if (int32_t rc = rvx_chain(e->ctx, &d, e->condition ? e->cwin : e->win, e->taps, e->N, e->tw, M,
                               e->ring,
                               uint64_t(M) * e->cfg.ring_pitch, e->partial, frames * M))
        return rc;
    if (fork) RVX_TRY_HIP(hipEventRecord(fork, e->ctx->stream), who);
    if (frames) {
        rvx_ema_params p{};
        p.channels = M;
        p.chan_blocks = e->cfg.ring_blocks;
        p.chan_mask = e->cfg.ring_blocks - 1;
        p.in_offset = 0;
        p.frame_blocks = e->cfg.frame_blocks;
        p.frames = uint32_t(frames);
        p.alpha = e->cfg.alpha;
        p.frame_decay = float(std::pow(1.0 - double(e->cfg.alpha), double(e->cfg.frame_blocks)));
        p.row_frames = e->cfg.row_frames;
        p.row_decay = float(std::pow(1.0 - double(e->cfg.alpha),
                                     double(e->cfg.frame_blocks) * e->cfg.row_frames));
        const uint64_t nrows = frames / e->cfg.row_frames;
        if (int32_t rc = rvx_ema_combine(e->ctx, &p, e->partial, frames * M, e->state, e->rows,
                                         nrows * M))
            return rc;
        e->stats.spectrum_rows += nrows;
    }
    if (!e->any) {
        e->any = true;
        e->first_block = e->next_block;
    }
    e->next_block += count;
    e->stats.blocks += count;
    e->stats.next_block = e->next_block;
    return RVX_OK;
}

int32_t rvx_engine_read_audio(rvx_engine* e, uint32_t id, float* out, uint64_t max_values,
                              uint64_t* first_audio, uint64_t* values) {
    static const char* who = "rvx_engine_read_audio";
    if (!e) return rvx::fail(RVX_EINVAL, "%s: null engine", who);
    RVX_ENTER(e->ctx, who);
    if (!first_audio || !values || (max_values && !out))
        return rvx::fail(RVX_EINVAL, "%s: null argument", who);
    auto it = std::find_if(e->rx.begin(), e->rx.end(), [&](const EngineRx& r) { return r.id == id; });
    if (it == e->rx.end()) return rvx::fail(RVX_EINVAL, "%s: no receiver %u", who, id);
    EngineRx& r = *it;
    RVX_TRY_HIP(hipStreamSynchronize(e->ctx->stream), who);
    const uint32_t ch = r.plan.demod.channels;
    const uint64_t n = std::min<uint64_t>(r.next_audio - r.read_audio, max_values / ch);
    *first_audio = r.read_audio;
    uint64_t done = 0;
    while (done < n) {
        const uint64_t pos = (r.read_audio + done) % r.cap;
        const uint64_t seg = std::min(n - done, r.cap - pos);
        RVX_TRY_HIP(hipMemcpy(out + done * ch, r.audio + pos * ch, seg * ch * sizeof(float),
                              hipMemcpyDeviceToHost),
                    who);
        done += seg;
    }
    r.read_audio += n;
    *values = n * ch;
    return RVX_OK;
}

int32_t rvx_engine_spectrum(rvx_engine* e, float* out, uint64_t n) {
    static const char* who = "rvx_engine_spectrum";
    if (!e) return rvx::fail(RVX_EINVAL, "%s: null engine", who);
    RVX_ENTER(e->ctx, who);
    if (!e->state) return rvx::fail(RVX_EINVAL, "%s: the averaged spectrum is off", who);
    if (!out || n < e->cfg.channels)
        return rvx::fail(RVX_EINVAL, "%s: wants %u floats", who, e->cfg.channels);
    RVX_TRY_HIP(hipStreamSynchronize(e->ctx->stream), who);
    RVX_TRY_HIP(hipMemcpy(out, e->state, e->cfg.channels * sizeof(float), hipMemcpyDeviceToHost),
                who);
    return RVX_OK;
}

const void* rvx_engine_channel_ring(rvx_engine* e) { return e ? e->ring : nullptr; }

int32_t rvx_engine_front_end(rvx_engine* e, rvx_front_end_estimate* estimate,
                             rvx_iq_correct_params* applied) {
    static const char* who = "rvx_engine_front_end";
    if (!e) return rvx::fail(RVX_EINVAL, "%s: null engine", who);
    if (estimate) {
        *estimate = rvx_front_end_estimate{};
        estimate->gain = 1.0;
        if (e->fe) rvx_front_end_estimate_get(e->fe, estimate);
    }
    if (applied) *applied = e->corr_now;
    return RVX_OK;
}

int32_t rvx_engine_stats_get(rvx_engine* e, rvx_engine_stats* out) {
    if (!e || !out) return rvx::fail(RVX_EINVAL, "rvx_engine_stats_get: null argument");
    *out = e->stats;
    out->next_block = e->next_block;
    return RVX_OK;
}

}  // extern "C"
