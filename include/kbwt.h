/* libkbwt_hip: kolmo's block-sorting path (forward suffix sort with BWT, primary index and sampled indexes, and the
 * inverse BWT) on one MI355X (gfx950), behind kolmo's own C entry points.
 *
 * Drop-in boundary: the three kolmo_sais_* functions below have the signatures, argument meaning, validation and
 * return codes of src/algo/sais/sais.h:13,17,20 (implemented in src/algo/sais/sais.c:12027-12082), so linking this
 * library in place of sais.c leaves every caller (src/codecs/bwt_stream.c:52,73-74,233,294) unchanged and every
 * codec-5 and codec-18 stream byte-identical. The output never depends on `threads`, as sais.h:5 requires.
 *
 * Return codes: 0 ok; -1 bad arguments or indexes out of range (sais.c:12029-12032, 12057-12066); -2 out of memory
 * (host or device), or a block longer than 2^31 - 65 bytes, which the device kernels cannot index; -3 a HIP runtime
 * error (no usable device, a failed launch). A device error is reported on stderr and through kbwt_last_error(), and
 * the library recovers from it (the stream is drained and the cached state reset) before the call returns, so the
 * next call is unaffected. The CPU path is off by default: with kbwt_set_cpu_fallback(1) (or KBWT_CPU_FALLBACK=1) a
 * host-pointer call whose device work fails is computed on the CPU instead, with identical bytes, as kolmo's Metal lane
 * did (src/arch/metal/cubwt_metal.c:4-7), and counted as kolmo counts GPU fallbacks (src/core/hetero.c:62-66).
 *
 * Every function is reentrant; calls from several host threads are serialised per device, except that concurrent
 * kolmo_sais_* calls (kolmo's block workers) are gathered into batches: one caller runs every block queued at that
 * moment as one kbwt batch, and each block still gets exactly its single-block result. When such a batch fails, each
 * of its blocks runs again on its own, so a failure stays with the call it belongs to. */
#ifndef KBWT_H
#define KBWT_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#define KBWT_API __attribute__((visibility("default")))
#else
#define KBWT_API
#endif

#define KBWT_OK 0
#define KBWT_ERR_ARG (-1)
#define KBWT_ERR_NOMEM (-2)
#define KBWT_ERR_DEVICE (-3)

/* ---- drop-in entry points (host pointers) ---- */

/* Replaces kolmo_sais_bwt_aux (sais.h:13; sais.c:12027-12053). BWT of T[0..n) into U (T == U allowed) in libsais'
   convention: U[0] = T[n-1], then T[SA[p]-1] for every suffix-array row p except the one holding suffix 0.
   I[k] = 1 + rank of the suffix starting at k*r, k = 0..(n-1)/r; I[0] is the primary index. r is a power of two
   >= 2. A (n int32 of caller scratch) is not used: its contents on return are unspecified, as in the reference. */
KBWT_API int32_t kolmo_sais_bwt_aux(const uint8_t *T, uint8_t *U, int32_t *A, int32_t n, int32_t r, int32_t *I, int32_t threads);

/* Replaces kolmo_sais_unbwt_aux (sais.h:17; sais.c:12055-12077). Inverse of the above (T == U allowed). Every
   I[t], t = 0..(n-1)/r, must lie in 1..n (else -1); r is n or a power of two >= 2. Only I[0] steers the decode: the
   other samples are a decoding aid that a decoder may ignore (FORMAT.md 22.6), and this one does. A is unused. */
KBWT_API int32_t kolmo_sais_unbwt_aux(const uint8_t *T, uint8_t *U, int32_t *A, int32_t n, int32_t r, const int32_t *I,
                             int32_t threads);

/* Replaces kolmo_sais_unbwt (sais.h:20; sais.c:12079-12082): kolmo_sais_unbwt_aux with r = n and primary index i. */
KBWT_API int32_t kolmo_sais_unbwt(const uint8_t *T, uint8_t *U, int32_t *A, int32_t n, int32_t i, int32_t threads);

/* ---- device-resident forms (no counterpart in the reference: the inputs already sit in HBM) ----
   dT, dU are device pointers of the current device; I is host memory; stream is a hipStream_t (NULL: the library's
   own stream). They return when the result is complete. Same semantics and codes as the drop-in functions;
   dT == dU is allowed. */
KBWT_API int32_t kbwt_dev_bwt_aux(const uint8_t *dT, uint8_t *dU, int32_t n, int32_t r, int32_t *I, void *stream);
KBWT_API int32_t kbwt_dev_unbwt_aux(const uint8_t *dT, uint8_t *dU, int32_t n, int32_t r, const int32_t *I, void *stream);

/* kbwt_dev_unbwt_aux without any allocation. No counterpart in the reference: it is for callers that cannot allocate
   device memory (code under HIP graph capture, tight decode loops, harnesses that forbid hipMalloc). It calls no
   hipMalloc, hipFree or allocating library routine: every buffer of the call lives in a static device array of about
   100 MiB that each device running the library holds until the process ends. n is at most 4194304 (4 MiB; larger
   blocks return -1 with a message in kbwt_last_error). stream NULL is the null stream. Same semantics, codes and
   output as kbwt_dev_unbwt_aux. Calls that overlap on one device share the array and race: serialise them. */
KBWT_API int32_t kbwt_dev_unbwt_aux_noalloc(const uint8_t *dT, uint8_t *dU, int32_t n, int32_t r, const int32_t *I,
                                            void *stream);

/* ---- batches (no counterpart in the reference: kolmo calls the single-block entry points once per block) ----
   count independent blocks in one launch sequence. Block b occupies bytes [offset, offset + n) of T and of U;
   blocks must not overlap and are given in increasing offset order. r: the block's sampling interval, as for the
   single-block functions. I holds each block's (n-1)/r + 1 sampled indexes (1 when n <= 1), concatenated in block
   order. Every block's result is the single-block function's result. count <= 65535; the blocks' total length
   < 2^31. The dev_ forms take device pointers, like kbwt_dev_bwt_aux. */
typedef struct {
    int64_t offset;
    int32_t n;
    int32_t r;
} kbwt_seg;
KBWT_API int32_t kbwt_bwt_aux_batch(const uint8_t *T, uint8_t *U, const kbwt_seg *segs, int32_t count, int32_t *I);
KBWT_API int32_t kbwt_unbwt_aux_batch(const uint8_t *T, uint8_t *U, const kbwt_seg *segs, int32_t count,
                                      const int32_t *I);
KBWT_API int32_t kbwt_dev_bwt_aux_batch(const uint8_t *dT, uint8_t *dU, const kbwt_seg *segs, int32_t count,
                                        int32_t *I, void *stream);
KBWT_API int32_t kbwt_dev_unbwt_aux_batch(const uint8_t *dT, uint8_t *dU, const kbwt_seg *segs, int32_t count,
                                          const int32_t *I, void *stream);

/* ---- learned router (no counterpart in the reference) ----
   Picks the coder and the block size of a level-8 block, the choice src/core/blocks.c:390-411 and :473-507 make today
   by coding a slice of the block both ways: class 0 block sorting of the whole block (codec 18, kolmo_bwt_compress),
   1 KLZMA (codec 24, kolmo_klzma_compress level 8), 2 store the block, 3 block sorting at half the block size (its two
   halves as two codec-18 blocks). The choice is recorded per block, so coders, containers and the decoder are
   unchanged; exhaustive best-per-block (code every way, keep the smallest) stays the fallback for blocks the network
   is unsure of.
   Features, KBWT_ROUTER_FEATURES fp32 per block: sqrt(byte count / n) for the 256 byte values, the order-1 entropy
   H(T[i] | T[i-1]) / 8, and m_r / n for r = 1..5 (the share of suffixes still in rank groups at the start of
   doubling round r). Network: 262 -> 64 -> 64 -> 4, ReLU, softmax; KBWT_ROUTER_PARAMS fp32 parameters, input-major
   (W1[k*64+j], b1, W2[k*64+j], b2, W3[k*4+c], b3). All router arrays below are host memory. */
#define KBWT_ROUTER_FEATURES 262
#define KBWT_ROUTER_CLASSES 4
#define KBWT_ROUTER_PARAMS 21252
#define KBWT_ROUTER_GRAD_MAX 4096   /* samples of one kbwt_router_grad call and of one training batch */

/* kbwt_dev_bwt_aux_batch, and the router's features of every block into features[count * KBWT_ROUTER_FEATURES]. */
KBWT_API int32_t kbwt_dev_bwt_aux_batch_features(const uint8_t *dT, uint8_t *dU, const kbwt_seg *segs, int32_t count,
                                                 int32_t *I, float *features, void *stream);
typedef struct {
    int32_t epochs, batch;       /* passes over the samples; samples per Adam step (<= KBWT_ROUTER_GRAD_MAX) */
    float lr, beta1, beta2, eps; /* Adam; 0 for beta1, beta2, eps takes 0.9, 0.999, 1e-8 */
    uint64_t seed;               /* shuffling */
} kbwt_router_train_cfg;
/* Deterministic He-uniform initialisation from seed. */
KBWT_API int32_t kbwt_router_init(float *params, uint64_t seed);
/* Minibatch Adam on the weighted cross-entropy mean_s w[s] * -log p(y[s] | X[s]) (w NULL: all 1), y in 0..2.
   params: in, trained out. epoch_loss (NULL or cfg->epochs floats): mean loss of each epoch. */
KBWT_API int32_t kbwt_router_train(const float *X, const int32_t *y, const float *w, int32_t ns,
                                   const kbwt_router_train_cfg *cfg, float *params, float *epoch_loss);
/* Gradient of that mean loss at params over ns <= KBWT_ROUTER_GRAD_MAX samples, and the loss. */
KBWT_API int32_t kbwt_router_grad(const float *X, const int32_t *y, const float *w, int32_t ns, const float *params,
                                  float *grad, float *loss);
/* Class probabilities, probs[ns * KBWT_ROUTER_CLASSES]. */
KBWT_API int32_t kbwt_router_predict(const float *X, int32_t ns, const float *params, float *probs);
/* The novelty gate: dist[s] = Euclidean distance from the features of sample s to the nearest of np prototype feature
   vectors (protos[np * KBWT_ROUTER_FEATURES]). A block far from every prototype is unlike the blocks the network was
   trained on, and its decision falls back to exhaustive best-per-block. */
KBWT_API int32_t kbwt_router_gate(const float *X, int32_t ns, const float *protos, int32_t np, float *dist);
/* The router at kolmo's decision site (src/core/blocks.c:390-411, 473-507), for a frame's blocks at once: features
   of each host block (blocks[b], n[b] bytes; one batched forward on the device), the network, the gate, and
   decision[b] = the class of highest probability, or -1 when that probability is below threshold or the block lies
   farther than radius from every prototype (protos NULL: no gate); on -1 the caller codes the block every way and
   keeps the smallest (exhaustive best-per-block). probs (NULL or count * KBWT_ROUTER_CLASSES floats) gets the
   probabilities. params, threshold, protos and radius: a trained model (kolmo_amd/data/router_model.npz). */
KBWT_API int32_t kbwt_router_route(const uint8_t *const *blocks, const int32_t *n, int32_t count, const float *params,
                                   float threshold, const float *protos, int32_t nprotos, float radius,
                                   int32_t *decision, float *probs);

/* ---- autotuner (no counterpart in the reference) ----
   Launch parameters per size class (ceil(log2) of a call's total length): sort_cfg, the radix pass's workgroup and
   tile shape (0: 512 threads x 16 keys with a 4-tile lookback window; 1-10: 256x16, 256x16 with window 1, 512x8,
   256x12, 256x8, 512x16 with window 1, 1024x12, 1024x16, 1024x16 with window 2, 1024x16 with window 8);
   inv_dshift, log2 of the inverse's splitter spacing (2..12); seg_rank, the largest group the doubling rounds rank
   directly instead of radix-sorting (0..1024). Built-in table: 0, 3, 64 up to 1 MiB; 0, 4, 64 up to 2 MiB; 8, 4, 64
   up to 8 MiB; 8, 6, 64 up to 64 MiB; 8, 7, 64 above (a call's class is ceil(log2) of its total length). Digits are
   8 bits wide in every configuration. None of them changes any output. kbwt_autotune times a forward + inverse of
   the given device blocks for each candidate, by coordinate descent from start (NULL: the built-in entry; its ms is
   set to the time at start), installs the fastest for that size class and returns it with its time in best. */
typedef struct {
    int32_t sort_cfg, inv_dshift, seg_rank;
    float ms;
} kbwt_tuning;
KBWT_API int32_t kbwt_autotune(const uint8_t *dT, const kbwt_seg *segs, int32_t count, void *stream, kbwt_tuning *best,
                               kbwt_tuning *start);
/* size_log2 < 0 sets every class; t == NULL puts the class (or every class) back to the built-in table. */
KBWT_API int32_t kbwt_set_tuning(int32_t size_log2, const kbwt_tuning *t);
KBWT_API int32_t kbwt_get_tuning(int32_t size_log2, kbwt_tuning *t);

/* ---- telemetry of the calling thread's last call (a kolmo_sais_* block that ran in another caller's batch leaves
   it with that caller's thread) ---- */
#define KBWT_MAX_ROUNDS 40
typedef struct {
    int32_t n;                         /* block length (a batch: the blocks' total length) */
    int32_t blocks;                    /* blocks of the call */
    int32_t sigma;                     /* distinct byte values */
    int32_t bits_per_symbol;           /* w: bits of the dense symbol code */
    int32_t h0;                        /* symbols in the first sort key */
    int32_t rounds;                    /* doubling rounds after the first sort */
    int32_t passes[KBWT_MAX_ROUNDS];   /* radix passes of round r (round 0 = the first sort) */
    int64_t active[KBWT_MAX_ROUNDS];   /* m_r: suffixes sorted in round r (m_0 = n) */
    int32_t splitters;                 /* inverse: sublists of the list ranking */
    double ms_fwd, ms_inv;             /* device time of the last forward / inverse call (HIP events) */
} kbwt_stats;
KBWT_API int32_t kbwt_last_stats(kbwt_stats *out);

/* ---- measurement (no counterpart in the reference) ----
   kbwt_profile(1) clears and starts per-kernel timing: every launch is bracketed by HIP events on its own stream, and
   kbwt_profile_read fills up to max records (name, summed device ms, launches, summed elements) and returns their
   count. kbwt_profile(0) stops it. kbwt_profile_only(name) clears and times only the launches of that kernel, so
   a timed run carries events around those launches alone. kbwt_dev_copy is a plain 16-byte-vector device copy (16-byte aligned pointers and
   size), the reference workload for scaling hardware byte counters. */
typedef struct {
    char name[32];
    double ms;
    int64_t launches;
    int64_t elems;
} kbwt_kernel_time;
KBWT_API void kbwt_profile(int32_t on);
KBWT_API void kbwt_profile_only(const char *name);
KBWT_API int32_t kbwt_profile_read(kbwt_kernel_time *out, int32_t max);
KBWT_API int32_t kbwt_dev_copy(void *dst, const void *src, size_t bytes, void *stream);

/* ---- error path (no counterpart in the reference's sais.h; it mirrors kolmo's GPU worker, src/core/hetero.c) ----
   kbwt_set_cpu_fallback(on): host-pointer calls (kolmo_sais_*, kbwt_bwt_aux_batch, kbwt_unbwt_aux_batch) whose device
   work fails with -2 or -3 are computed by the library's CPU path instead (src/arch/metal/cubwt_metal.c:4-7). Off by
   default; KBWT_CPU_FALLBACK=1 in the environment turns it on. Returns the previous setting. The device-pointer
   entry points never fall back.
   kbwt_fault_inject(spec): host-side fault injection for tests. spec "site:k" or "site:k:count", site one of alloc,
   launch, sync, copy: occurrences k .. k+count-1 of that fault point, counted from this call, fail as if the HIP
   runtime had returned an error (alloc: -2, the others: -3), before the runtime is called. NULL or "" disarms it.
   KBWT_FAULT in the environment arms it at load. Returns -1 for a spec it cannot parse.
   kbwt_get_counters: process-wide counts since load (kolmo's KOLMO_GPU_STATS, src/core/hetero.c:19-22,38). */
typedef struct {
    int64_t device_errors;     /* calls whose device work failed (-2 or -3 from the device path) */
    int64_t retried_blocks;    /* blocks of a failed combined batch that ran again on their own */
    int64_t cpu_fallbacks;     /* blocks computed by the CPU path after a device error */
    int64_t injected_faults;   /* faults injected by kbwt_fault_inject / KBWT_FAULT */
    int64_t combined_batches;  /* batches of concurrent kolmo_sais_* calls */
    int64_t combined_blocks;   /* blocks those batches held */
    int64_t graph_replays;     /* launch sequences run from an executable graph (the forward's first sort, the inverse) */
} kbwt_counters;
KBWT_API int32_t kbwt_set_cpu_fallback(int32_t on);
KBWT_API int32_t kbwt_fault_inject(const char *spec);
KBWT_API int32_t kbwt_get_counters(kbwt_counters *out);

/* Message of the calling thread's last error ("" when none). */
KBWT_API const char *kbwt_last_error(void);

/* 1 when a gfx950 device is usable, else 0. */
KBWT_API int32_t kbwt_device_ok(void);

/* Releases the cached device workspace of the current device, and its pinned host staging chunks. */
KBWT_API void kbwt_release(void);

#ifdef __cplusplus
}
#endif
#endif
