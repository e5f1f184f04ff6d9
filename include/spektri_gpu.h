/*
 * spektri_gpu.h -- C ABI of the MI355X (gfx950) FCFB + Welch hot path.
 *
 * This library replaces the compute seam of spektri's DSP layer:
 *   DspState::init(DspParams) -> (DspState, InputBufferSize)   spektri/src/dsp.rs:76-112
 *   DspState::process_real / process_complex                   spektri/src/dsp.rs:114-160
 *     MultiFft::process_*        (forward FFTs)                spektri/src/dsp/multifft.rs:34-96
 *     SpectrumAccumulator::accumulate (Welch |X|^2 sums)        spektri/src/dsp/spectrum.rs:82-116
 *     Fcfb::add_filter / Fcfb::process / FilterDsp::process     spektri/src/dsp/fcfb.rs:41-100,154-188
 * The caller keeps framing of the raw byte stream (read_exact), the CLI,
 * the dB quantiser, the wire format and ZeroMQ/file output; it hands raw
 * input bytes in and reads channel samples and averaged spectra out.
 *
 * Conventions: every entry point returns an spk_status; nothing throws,
 * panics or aborts across the boundary.  A context is single-threaded
 * (like the ZeroMQ socket it feeds).  The caller owns host buffers; the
 * library owns device memory, its HIP stream (unless one is set) and its
 * kernels.  All arithmetic is FP32.
 */
#ifndef SPEKTRI_GPU_H
#define SPEKTRI_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum spk_status {
  SPK_OK = 0,
  SPK_ERR_INVALID = 1,      /* bad argument (size, pointer, range) */
  SPK_ERR_UNSUPPORTED = 2,  /* valid for the reference, not implemented here */
  SPK_ERR_DEVICE = 3,       /* a HIP call failed */
  SPK_ERR_STATE = 4,        /* call out of order (e.g. output before process) */
  SPK_ERR_CAPACITY = 5      /* more blocks / records than the context was sized for */
} spk_status;

/* DataFormat codes of spektri/src/dsp/data.rs:81-95; every one is an input
 * format (inputformats.rs:12-155).  s16le, f32le, cs16le and cf32le are
 * read in place by the transform kernels; the others pass through one
 * conversion kernel first (into s16, f32 or cf32, with the reference's
 * arithmetic: u8/cu8 subtract 127.4 in f32, big-endian formats swap bytes). */
#define SPK_FMT_S8 0x04
#define SPK_FMT_S16LE 0x08
#define SPK_FMT_S16BE 0x09
#define SPK_FMT_F32LE 0x1C
#define SPK_FMT_F32BE 0x1D
#define SPK_FMT_U8 0x24
#define SPK_FMT_CS8 0x44
#define SPK_FMT_CS16LE 0x48
#define SPK_FMT_CS16BE 0x49
#define SPK_FMT_CF32LE 0x5C
#define SPK_FMT_CF32BE 0x5D
#define SPK_FMT_CU8 0x64

/* Mirrors DspParams (spektri/src/dsp.rs:24-36) minus fs/fc, which only
 * the frequency<->bin helpers below need, and minus the filter list, which
 * is added with spk_add_channel. */
typedef struct spk_config {
  uint32_t fft_size;          /* N: powers of two 1024..524288 (real input) and 1024..262144
                                 (complex input) run in their planned kernels; any other N
                                 from 16 to 2^20 whose transform length (N/2 for even real
                                 N, N otherwise) has no prime factor above 31, or splits
                                 into two factors of at most 480, runs in mixed-radix
                                 passes; other sizes return SPK_ERR_UNSUPPORTED */
  uint32_t input_format;      /* SPK_FMT_*: real or complex input follows the format (main.rs:35) */
  float scaling;              /* input scaling; 0 = input_format_scaling(fmt) (inputformats.rs:47-62) */
  uint32_t ffts_per_buf;      /* F: blocks per spk_process call */
  uint32_t spectrum_averages; /* A: blocks per Welch record */
  uint32_t max_blocks;        /* capacity for spk_process_device; 0 = F */
  uint32_t flags;             /* SPK_FLAG_* */
} spk_config;

/* Keep every block's forward spectrum in device memory (spk_spectra_device).
 * Without it the spectrum is only materialised when a channel kernel needs
 * it; channels of L = 256 are then computed inside the forward kernel. */
#define SPK_FLAG_KEEP_SPECTRA 1u
/* Capture each distinct (input pointer, block count) launch sequence into a
 * HIP graph once and replay it (no counterpart in the reference).  Applies
 * to calls with channel launches after the forward transform (complex
 * input, or channels not run inside the forward kernel); a real-input call
 * whose channels all run inside the forward kernel is two kernels and is
 * launched directly.  Ignored while per-kernel timing is enabled. */
#define SPK_FLAG_GRAPHS 2u
/* SpectrumFormat (spektri/src/dsp/spectrum.rs:29-32): also quantise every
 * Welch record to dB, U8 (0.5 dB per LSB) or U16 (0.05 dB per LSB). */
#define SPK_FLAG_SPECTRUM_U8 4u
#define SPK_FLAG_SPECTRUM_U16 8u

/* InputBufferSize (spektri/src/dsp.rs:38-42), in samples. */
typedef struct spk_buffer_size {
  uint64_t new_samples; /* F * 3N/4 */
  uint64_t overlap;     /* N / 4 */
  uint64_t total;       /* overlap + new */
} spk_buffer_size;

typedef struct spk_ctx spk_ctx;

/* DspState::init (spektri/src/dsp.rs:76-112). */
spk_status spk_create(const spk_config* cfg, spk_ctx** out);
spk_status spk_destroy(spk_ctx* ctx);
spk_status spk_buffer_size_of(const spk_ctx* ctx, spk_buffer_size* out);
/* Message of the last failed call on this context (static string on NULL). */
const char* spk_last_error(const spk_ctx* ctx);

/* Fcfb::add_filter after freq_to_bins_exact (spektri/src/dsp/fcfb.rs:41-64,137-152).
 * first_bin: number of the channel's lowest bin (BinNumbers.first, may be < 0);
 * bins: IFFT size L, a multiple of 4 and <= N (as freq_to_bins makes it,
 * fcfb.rs:225-248): powers of two up to 8192 run in the radix-2^k kernels,
 * other multiples of 4 up to 4096 in a mixed-radix kernel with several
 * channels per workgroup, and every multiple of 4 up to 20480 in a
 * mixed-radix kernel with one channel per workgroup; larger ones return
 * SPK_ERR_UNSUPPORTED.  A channel gives 6 * (L / 8) samples per block
 * (the integer slice buf[L/8 .. L/8*7] of fcfb.rs:176-177: 3L/4 when 8 | L,
 * none for L = 4);
 * weights: L floats in j order (weights[j] multiplies bin first + j), or NULL
 * for the reference's raised-cosine weights (fcfb.rs:197-203). */
spk_status spk_add_channel(spk_ctx* ctx, int64_t first_bin, uint32_t bins, const float* weights,
                           uint32_t* id);
/* Remove a channel (Fcfb::process drops filters that are done,
 * fcfb.rs:96-99; the planned subscription-driven channels need it at run
 * time).  Its id is not reused; the next process call lays the channel set
 * out again. */
spk_status spk_remove_channel(spk_ctx* ctx, uint32_t id);
/* Replace a channel's weights (same L). */
spk_status spk_set_channel_weights(spk_ctx* ctx, uint32_t id, const float* weights);

/* DspState::process_real / process_complex (spektri/src/dsp.rs:114-160) on one
 * buffer: exactly F * 3N/4 new raw samples of the input format from host
 * memory (nbytes = F * 3N/4 * bytes_per_input_sample, inputformats.rs:12-27).
 * The N/4-sample history is kept on the device (zeros at start,
 * main.rs:165).  Synchronous. */
spk_status spk_process(spk_ctx* ctx, const void* raw_new, size_t nbytes);

/* Streaming form of spk_process for a reader loop (main.rs:171-193): copies
 * the buffer into one of two pinned host slots (the caller may reuse its
 * buffer on return), uploads it on a copy stream and enqueues the launch
 * sequence on the context's stream after the upload, then returns without
 * waiting.  The upload of the next call overlaps this call's kernels, and
 * the caller's next read overlaps both.  Outputs are read with the usual
 * readers (which wait for the stream) and belong to the last call: read
 * them before the next call.  No counterpart in the reference, whose
 * process_* calls are synchronous. */
spk_status spk_process_async(spk_ctx* ctx, const void* raw_new, size_t nbytes);
/* The pinned host slot the next spk_process_async call uploads from (waits
 * until the slot's previous upload has left it): a reader can read_exact
 * straight into it (main.rs:176) and pass it back to spk_process_async,
 * which then skips its host copy.  *nbytes = F * 3N/4 raw samples in bytes.
 * Valid until the context is destroyed.  No counterpart in the reference. */
spk_status spk_input_buffer(spk_ctx* ctx, void** host_ptr, size_t* nbytes);

/* Same path on a device-resident stream of nsamples raw samples (4-byte
 * aligned): processes floor(nsamples / (F * 3N/4)) * F blocks, like the
 * reference's main loop over that many bytes (main.rs:171-179), in one
 * launch sequence on the context's stream.  Asynchronous.  For real input
 * the call's Welch records are summed by the next call's forward launch
 * (in the time its early workgroups would idle), or when first asked for:
 * spk_spectrum_records*, spk_spectrum_quantized*, spk_synchronize and
 * spk_flush enqueue the sum. */
spk_status spk_process_device(spk_ctx* ctx, const void* d_raw, uint64_t nsamples);

/* Blocks processed by the last process call. */
spk_status spk_last_blocks(const spk_ctx* ctx, uint32_t* nblocks);

/* FilterDsp::process output of the last call for channel id: nblocks *
 * 6 (L/8) complex samples as interleaved f32 (re, im), i.e. the record body
 * of fcfb.rs:177-187.  cap and *n count floats. */
spk_status spk_channel_output(spk_ctx* ctx, uint32_t id, float* out, size_t cap, size_t* n);
/* Device pointer and float count of the same data (valid until the next call). */
spk_status spk_channel_output_device(spk_ctx* ctx, uint32_t id, const float** dptr, size_t* n);

/* Welch records completed by the last call (spectrum.rs:113-160): each is K
 * floats acc / A (K = N/2 + 1).  *nrec receives the record count; cap counts
 * floats.  The _device form enqueues the records' sum on the context's
 * stream if it is still pending; the pointer is readable after that stream
 * (spk_synchronize). */
spk_status spk_spectrum_records(spk_ctx* ctx, float* out, size_t cap, uint32_t* nrec);
spk_status spk_spectrum_records_device(spk_ctx* ctx, const float** dptr, uint32_t* nrec);

/* Quantised Welch records of the last call (contexts created with
 * SPK_FLAG_SPECTRUM_U8 or SPK_FLAG_SPECTRUM_U16): the record body of
 * spectrum.rs:128-152, K bytes (U8) or K big-endian 16-bit words (U16) per
 * record, computed on the device from the f32 sum with the reference's
 * formula.  cap counts bytes. */
spk_status spk_spectrum_quantized(spk_ctx* ctx, void* out, size_t cap, uint32_t* nrec);
spk_status spk_spectrum_quantized_device(spk_ctx* ctx, const void** dptr, uint32_t* nrec,
                                         uint32_t* bytes_per_bin);

/* The same quantised records as the reference sends them
 * (spectrum.rs:120-153): each is 24 bytes of metadata (data.rs:110-132:
 * u64 seq = first_seq + record index, u64 secs, u32 nanos, u32 0, all
 * little-endian) followed by the body, packed on the device and copied out
 * in one transfer, so the host only writes or sends them.  *nbytes
 * receives the packed size (also when cap is too small); cap counts bytes.
 * No counterpart in the reference, whose host serialises each record. */
spk_status spk_spectrum_records_packed(spk_ctx* ctx, uint64_t first_seq, uint64_t secs, uint32_t nanos, void* out,
                                       size_t cap, uint32_t* nrec, size_t* nbytes);

/* Every live channel's record of the last call, as the reference writes
 * them (fcfb.rs:75-89): in ascending id order, 24 bytes of metadata
 * (data.rs:110-132: u64 seq, u64 secs, u32 nanos, u32 0, little-endian,
 * the same for every channel of a buffer) followed by the channel's
 * nblocks * 6 (L/8) samples as f32 re, im (little-endian), so record i is
 * 24 + 8 * nblocks * 6 (L_i / 8) bytes.  Packed on the device and copied
 * out in one transfer.  *nbytes receives the packed size (also when cap is
 * too small).  No counterpart in the reference, whose host serialises each
 * record. */
spk_status spk_channel_records_packed(spk_ctx* ctx, uint64_t seq, uint64_t secs, uint32_t nanos, void* out,
                                      size_t cap, uint32_t* nrec, size_t* nbytes);

/* Forward spectra of the last call: nblocks rows of K complex f32 with a
 * row stride of *stride complex values (the reference's fft_result_buf).
 * Valid only for contexts created with SPK_FLAG_KEEP_SPECTRA. */
spk_status spk_spectra_device(spk_ctx* ctx, const float** dptr, uint64_t* stride);

/* Clear history, Welch accumulators and counters (as a fresh DspState).
 * Asynchronous: the device state is cleared by a kernel on the context's
 * stream after every call already enqueued there (no memset, so the call
 * can be captured into a graph); the previous call's pending records are
 * dropped. */
spk_status spk_reset(spk_ctx* ctx);
/* Run on the caller's HIP stream (hipStream_t passed as void*); NULL
 * returns to the context's own stream. */
spk_status spk_set_stream(spk_ctx* ctx, void* stream);
/* Sum the last call's pending Welch records (as spk_flush), then wait for
 * the context's stream. */
spk_status spk_synchronize(spk_ctx* ctx);
/* Enqueue the sum of the last call's Welch records if still pending (no
 * counterpart in the reference, whose records are written synchronously);
 * asynchronous. */
spk_status spk_flush(spk_ctx* ctx);

/* Per-kernel timing from HIP events recorded on the context's stream around
 * each launch: average ms per call since spk_enable_timing of ms[0] the
 * forward kernel (FFT, Welch partial sums and the fused channels), ms[1] the
 * channel launches of the channels not fused, ms[2] the Welch finaliser.
 * Enabling adds events only; spk_last_timing synchronizes.  No counterpart
 * in the reference (measurement hook). */
spk_status spk_enable_timing(spk_ctx* ctx, int on);
spk_status spk_last_timing(spk_ctx* ctx, float ms[3]);

/* Launch form of the channel stage for the channels not fused into the
 * forward kernel (no counterpart in the reference, whose filters run one by
 * one on rayon, fcfb.rs:66-100):
 *   SPK_CHAN_PER_SIZE  one launch per distinct IFFT size (the default);
 *   SPK_CHAN_BATCHED   one persistent launch for every channel of every
 *                      size, with XCD-local block chunks; IFFT sizes
 *                      without a power-of-two plan join it up to L = 2048
 *                      (above that they keep their own launch).  */
#define SPK_CHAN_PER_SIZE 0
#define SPK_CHAN_BATCHED 1
spk_status spk_set_channel_mode(spk_ctx* ctx, int mode);

/* On the context's current spectra for nblocks blocks (0 = max_blocks):
 * time every tile shape of every power-of-two IFFT size in its per-size
 * launch and keep the fastest for that form; time the per-size form; for
 * the batched form, with each size's fastest shape and with its fastest
 * shape of at most 16 points per thread (half the LDS per workgroup), search
 * the block chunk, rounds per work item and workgroups per CU; keep the
 * fastest settings and select the batched form if it is at least 2 %
 * faster.  Synchronous; overwrites channel outputs.  No counterpart in the
 * reference. */
typedef struct spk_tune_result {
  float per_size_ms;
  float batched_ms;
  uint32_t chunk_blocks;
  uint32_t rounds_per_item;
  uint32_t wgs_per_cu; /* 0 = as many as fit */
  int mode;            /* SPK_CHAN_* selected */
} spk_tune_result;
spk_status spk_autotune(spk_ctx* ctx, uint32_t nblocks, spk_tune_result* out);

/* Tile shape of the channel transform of power-of-two IFFT size `bins` in
 * one launch form (SPK_CHAN_*): shape s is a radix plan with
 * points_per_thread points per thread, so a (channel, block) tile takes
 * bins / points_per_thread threads.  Sizes 16 to 4096 have two or three
 * shapes.  spk_set_tile_shape fixes one (SPK_ERR_INVALID for a shape the
 * size does not have); the next call lays the channel tables out again.
 * In the batched form a shape of more than 16 points per thread halves the
 * workgroups per CU for every size (twice the LDS each).
 * spk_tile_shape reports the shape in use (the form's default until
 * spk_autotune or spk_set_tile_shape chose one).  No counterpart in the
 * reference: launch geometry for the autotuner and its tests. */
spk_status spk_set_tile_shape(spk_ctx* ctx, uint32_t bins, int form, uint32_t shape);
spk_status spk_tile_shape(spk_ctx* ctx, uint32_t bins, int form, uint32_t* shape, uint32_t* threads_per_tile,
                          uint32_t* points_per_thread);

/* Number of channels whose gather, weighting, IFFT and discard run inside
 * the forward kernel (real input up to N = 65536): every channel when all
 * have IFFT size 256, lie in [0, K), do not overlap and fit the workgroup's
 * 16-thread tiles; otherwise every channel of a power-of-two IFFT size from
 * 8 to 256 when their tiles fit the kernel's LDS and no bin feeds more than
 * two tile positions (overlapping channels, get_bin's mirror at DC and
 * Nyquist); 0 otherwise.  No counterpart in the reference (plan
 * introspection). */
spk_status spk_fused_channels(spk_ctx* ctx, uint32_t* n);

/* Captured launch sequences (SPK_FLAG_GRAPHS): executable graphs held now,
 * graphs instantiated and graphs updated in place since the context was
 * created.  A channel added, removed or reweighted does not drop the
 * graphs: the next call with the same input pointer and block count
 * captures its launch sequence again and updates the executable graph
 * (hipGraphExecUpdate); only a changed launch topology (another kernel
 * count or kernel) instantiates a new one.  No counterpart in the reference
 * (plan introspection).  Any pointer may be NULL. */
spk_status spk_graph_info(spk_ctx* ctx, uint32_t* live, uint32_t* instantiated, uint32_t* updated);

/* ---- Differentiable channel weights (no counterpart in the reference; the
 * reference's weights are fixed raised cosines, fcfb.rs:197-203).  Packed
 * layouts follow channel id order over the channels not removed: weights L floats per channel; channel
 * outputs and their gradients nblocks * 6 (L/8) complex per channel; gathered-
 * bin gradients nblocks * L complex per channel (nblocks = blocks of the
 * last process call).  All pointers are device pointers; the calls are
 * asynchronous on the context's stream unless stated. */

/* Set every channel's weights from a packed device array: a copy on the
 * context's stream (asynchronous, so weights must stay valid until the
 * stream reaches it).  The host copies of the weights are read back only
 * when a channel is added, removed or given host weights, or the channel
 * set is laid out again (a wait on the stream then). */
spk_status spk_set_weights_device(spk_ctx* ctx, const float* weights);
/* Copy the last call's channel outputs into a packed device array of cap floats. */
spk_status spk_channels_output_packed(spk_ctx* ctx, float* dst, size_t cap);
/* Adjoint of weight-multiply, IFFT and discard (FilterDsp::process,
 * fcfb.rs:154-188) for the last process call: ybar holds dLoss/dRe + i
 * dLoss/dIm of every packed output sample; writes wbar (packed weights
 * layout, summed over blocks) and, when xbar is not NULL, the gradient with
 * respect to every gathered bin get_bin(first + j).  Needs a context created
 * with SPK_FLAG_KEEP_SPECTRA (SPK_ERR_STATE otherwise). */
spk_status spk_channels_adjoint(spk_ctx* ctx, const float* ybar, float* wbar, float* xbar);
/* spk_channels_adjoint without any device allocation: no counterpart in the
 * reference.  It is for callers that cannot allocate (code under HIP graph
 * capture, serving loops that forbid hipMalloc).  The per-(channel, block
 * split) weight sums go to one fixed 8 MiB scratch array per GPU, held
 * until the process ends, instead of a context buffer grown on demand;
 * SPK_ERR_CAPACITY when a channel group needs more than 2^21 floats there
 * (nch * S * L, S the group's block splits).  Same kernels and results as
 * spk_channels_adjoint.  Calls that overlap on one GPU (several contexts or
 * streams) share the scratch and race; the caller orders them. */
spk_status spk_channels_adjoint_noalloc(spk_ctx* ctx, const float* ybar, float* wbar, float* xbar);

/* freq_to_bins / bins_to_freq / freq_to_bins_exact (spektri/src/dsp/fcfb.rs:225-285).
 * Return SPK_ERR_INVALID when the reference returns None. */
spk_status spk_freq_to_bins(double fs_in, double fc_in, uint32_t fft_size, double fs_out,
                            double fc_out, uint32_t* bins, int64_t* first);
spk_status spk_bins_to_freq(double fs_in, double fc_in, uint32_t fft_size, uint32_t bins,
                            int64_t first, double* fs_out, double* fc_out);
spk_status spk_freq_to_bins_exact(double fs_in, double fc_in, uint32_t fft_size, double fs_out,
                                  double fc_out, uint32_t* bins, int64_t* first);

#ifdef __cplusplus
}
#endif
#endif /* SPEKTRI_GPU_H */
