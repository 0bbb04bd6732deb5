/*
 * oz2_c.h - C ABI of the MI355X (gfx950) Ozaki Scheme II DGEMM/ZGEMM library
 * (liboz2.so).  Plain pointers and sizes only; the C++ drop-in class
 * oz2::Ozaki2 in <oz2/ozaki2.hpp> wraps these entry points.
 *
 * Each entry point replaces one member of the reference's class oz2::Ozaki2
 * (src/ozaki2_gpu.hpp, cited per function).  Matrices are column-major device
 * pointers; complex data are interleaved (re, im) doubles.  All work is
 * enqueued asynchronously on the handle's HIP stream.
 *
 * Return codes (the reference throws the matching std exception types):
 *   OZ2_OK 0, OZ2_INVALID_ARGUMENT 1 (std::invalid_argument),
 *   OZ2_RUNTIME_ERROR 2 (std::runtime_error), OZ2_LOGIC_ERROR 3 (std::logic_error),
 *   OZ2_OVERFLOW_ERROR 4 (std::overflow_error).
 * oz2_last_error() returns the message of the calling thread's last failure.
 */
#ifndef OZ2_C_H
#define OZ2_C_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oz2_engine* oz2_handle;

enum { OZ2_OK = 0, OZ2_INVALID_ARGUMENT = 1, OZ2_RUNTIME_ERROR = 2, OZ2_LOGIC_ERROR = 3, OZ2_OVERFLOW_ERROR = 4 };
enum { OZ2_TRANS_N = 0, OZ2_TRANS_T = 1, OZ2_TRANS_C = 2 }; /* Trans{N,T,C}, hpp:34 */

typedef struct { double scale_ms, split_ms, gemm_ms, crt_ms; } oz2_stage_times_t; /* StageTimes, hpp:36-39 */
typedef struct { int64_t mb, nb, pm, pn, k_chunks; size_t workspace_bytes; } oz2_tile_info_t; /* TileInfo, hpp:47-52 */
/* Tuning, hpp:41-45, plus gemm_tile (-1 auto, else 256 / 128 / 64 / 256128) and overlap_blocks (column blocks
 * of each panel pair whose B split and CRT run on a second stream beside the next block's GEMM: -1 auto,
 * 0 or 1 none, up to 16) */
typedef struct { int batched_gemm, pad_lda, pow2_tiles, gemm_tile, overlap_blocks; } oz2_tuning_t;

/* Ozaki2(sycl::queue&) / ~Ozaki2() (hpp:680-687).  stream: a hipStream_t, NULL = default stream. */
int oz2_create(void* stream, oz2_handle* out);
int oz2_destroy(oz2_handle h);
/* rebinds the handle to another stream (no reference counterpart: a queue is fixed at construction) */
int oz2_set_stream(oz2_handle h, void* stream);

/* Ozaki2::gemm (hpp:739-760): C = alpha*op(A)*op(B) + beta*C with s = num_moduli in [2, 24].
 * deps: ndeps hipEvent_t the stream waits for first (the reference's deps vector).
 * done: if non-NULL, a hipEvent_t recorded after the last kernel (the returned sycl::event). */
int oz2_dgemm(oz2_handle h, int ta, int tb, int64_t m, int64_t n, int64_t k, double alpha, const double* A,
              int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc, int num_moduli,
              void* const* deps, int ndeps, void* done);

/* Ozaki2::zgemm (hpp:764-786): alpha, beta are (re, im) pairs; A, B, C interleaved complex. */
int oz2_zgemm(oz2_handle h, int ta, int tb, int64_t m, int64_t n, int64_t k, const double* alpha, const double* A,
              int64_t lda, const double* B, int64_t ldb, const double* beta, double* C, int64_t ldc, int num_moduli,
              void* const* deps, int ndeps, void* done);

/* oz2_dgemm with nothing allocated (no reference counterpart; for callers that cannot allocate device
 * memory: work under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, semantics and results as oz2_dgemm; every intermediate lives in a fixed device array of
 * OZ2_NOALLOC_SCRATCH_BYTES that each GPU holds from the library's load until the process ends, and C is
 * cut into panels whose intermediates fit it.  A product whose smallest panels (64 x 64 of C, all s
 * residue planes of their k) do not fit returns OZ2_INVALID_ARGUMENT before anything is enqueued.  No
 * stage times are recorded.  All calls on one GPU share its array: calls that overlap on one GPU (two
 * handles on two streams, two threads) race; keep them on one in-order stream. */
#define OZ2_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int oz2_dgemm_noalloc(oz2_handle h, int ta, int tb, int64_t m, int64_t n, int64_t k, double alpha, const double* A,
                      int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc, int num_moduli,
                      void* const* deps, int ndeps, void* done);

/* Ozaki2::precision_bits (hpp:691), precision_bits_complex (hpp:762), moduli_for_bits (hpp:693-700).
 * precision_bits returns tk >= 1, or -1 with the error code in *err (may be NULL). */
int oz2_precision_bits(int64_t k, int num_moduli, int* err);
int oz2_precision_bits_complex(int64_t k, int num_moduli, int* err);
int oz2_moduli_for_bits(int64_t k, int bits);

/* Ozaki2::set_workspace_limit / set_resident_hint / release_workspace (hpp:702-710) */
int oz2_set_workspace_limit(oz2_handle h, size_t bytes);
int oz2_set_resident_hint(oz2_handle h, size_t bytes);
int oz2_release_workspace(oz2_handle h);
/* grows the workspace now (no reference counterpart; lets callers pre-size before graph capture) */
int oz2_reserve_workspace(oz2_handle h, size_t bytes);
/* bytes of workspace a call with these arguments needs (no reference counterpart) */
int oz2_workspace_query(oz2_handle h, int ta, int tb, int64_t m, int64_t n, int64_t k, int num_moduli, int cplx,
                        size_t* bytes);

/* Ozaki2::stage_times / last_tiling / set_tuning / tuning (hpp:712-737) */
int oz2_stage_times(oz2_handle h, oz2_stage_times_t* out);
int oz2_last_tiling(oz2_handle h, oz2_tile_info_t* out);
int oz2_set_tuning(oz2_handle h, const oz2_tuning_t* t);
int oz2_get_tuning(oz2_handle h, oz2_tuning_t* out);
/* stage timing needs a profiling queue in the reference (hpp:681); here it is a switch, default on */
int oz2_set_profiling(oz2_handle h, int on);

/* stage times summed over every call between the two (no reference counterpart; the
 * reference reports only the last call) */
int oz2_profile_begin(oz2_handle h);
int oz2_profile_end(oz2_handle h, oz2_stage_times_t* sum, int64_t* calls);

/* the moduli of a set (coprime_tables.hpp); writes num_moduli values, returns OZ2_INVALID_ARGUMENT if out of range */
int oz2_moduli(int num_moduli, uint16_t* out);

/* ---- Batches of DGEMMs (no reference counterpart: the reference runs one GEMM per call) ----
 * A plan holds a list of DGEMMs, each with the full oz2_dgemm semantics (op, alpha, beta, leading
 * dimensions, num_moduli).  num_moduli > 0 fixes a problem's moduli; num_moduli <= 0 lets the plan
 * choose: the learned selector when OZ2_BATCH_SELECT is set (oz2_selector_choose), else
 * moduli_for_bits(k, bits).  The plan owns one workspace for all problems and
 * a grouped launch sequence (scale, split, one INT8 GEMM per macro-tile shape, CRT: about ten
 * launches for any count); with OZ2_BATCH_GRAPH that sequence is captured in a hipGraph when the
 * plan is created and oz2_batch_run replays it.  Buffers are bound at creation: running the plan
 * again recomputes from their current contents.  The moduli chosen at creation (by the selector or
 * moduli_for_bits) stay fixed: a run over new contents keeps them, even where the new data would need
 * more moduli for the bit target; oz2_batch_reselect re-runs the selector on the current contents.  Each problem's C is bit-identical to oz2_dgemm with
 * the same num_moduli.  tiles: NULL (per-problem heuristic) or count macro-tile codes: 256, 128, 64
 * (square tiles of that many rows and columns), 256128 (256 rows of op(A) x 128 columns of op(B),
 * two workgroups per CU: faster for k up to about 1024) or 0 (heuristic).  oz2_batch_run enqueues on the handle's current stream; oz2_batch_destroy waits for
 * the device first.  Real DGEMMs only. */
typedef struct {
    int ta, tb;
    int64_t m, n, k;
    double alpha;
    const double* A;
    int64_t lda;
    const double* B;
    int64_t ldb;
    double beta;
    double* C;
    int64_t ldc;
    int num_moduli;
} oz2_gemm_desc_t;
typedef struct oz2_batch_s* oz2_batch;
/* GRAPH: capture the launch sequence in a hipGraph; SELECT: learned num_moduli for entries with 0;
 * PERSISTENT: persistent GEMM schedule (a fixed grid drawing tiles from per-XCD work ranges) for
 *   128, 64 and 256128 tiles, else one workgroup per tile (256 tiles always take one per tile);
 * PROBLEM_ORDER: GEMM work in problem order, else largest first */
enum { OZ2_BATCH_GRAPH = 1, OZ2_BATCH_SELECT = 2, OZ2_BATCH_PERSISTENT = 4, OZ2_BATCH_PROBLEM_ORDER = 8 };
int oz2_batch_create(oz2_handle h, int count, const oz2_gemm_desc_t* descs, int bits, const int* tiles, int flags,
                     oz2_batch* out);
int oz2_batch_run(oz2_batch b);
/* For plans made with OZ2_BATCH_SELECT: runs the selector again on the buffers' current contents
 * (synchronizing the handle's stream) and, if any problem's choice changed, rebuilds the plan with the
 * same tiles and flags (new workspace; a new graph with OZ2_BATCH_GRAPH).  *changed (may be NULL):
 * the number of problems whose num_moduli changed; 0 and nothing done without OZ2_BATCH_SELECT. */
int oz2_batch_reselect(oz2_batch b, int* changed);
/* Autotuner: measures whole runs of candidate plans for these problems (tile policies x GEMM
 * schedules) and returns the fastest: tiles_out (count ints, for oz2_batch_create), flags_out (the
 * OZ2_BATCH_PERSISTENT / OZ2_BATCH_PROBLEM_ORDER bits to pass on), best_ms.  Runs the batch about
 * 40 times (C is overwritten); num_moduli as in oz2_batch_create with these flags (SELECT honoured). */
int oz2_batch_autotune(oz2_handle h, int count, const oz2_gemm_desc_t* descs, int bits, int flags, int* tiles_out,
                       int* flags_out, double* best_ms);
/* chosen moduli / macro tiles per problem (count ints each) */
int oz2_batch_moduli(oz2_batch b, int* out);
int oz2_batch_tiles(oz2_batch b, int* out);
/* workspace bytes, kernel launches per run, 1 if replayed from a graph */
int oz2_batch_info(oz2_batch b, size_t* workspace_bytes, int* launches, int* graph);
int oz2_batch_destroy(oz2_batch b);

/* ---- Learned moduli selector (no reference counterpart; moduli_for_bits, hpp:693-700, is its fallback) ----
 * Features: 48 floats per problem from sampled exponent gaps of op(A) rows / op(B) columns plus log2 k
 * and the bit target ("b bits": every entry within 2^(53-b) ulp of the exact product).  The network
 * (48 -> 64 -> 64 -> 23, FP32, OZ2_SEL_PARAMS floats) gives for s = 2..24 the probability that s moduli
 * reach the target; the choice is the smallest admissible s whose probability reaches the trained
 * threshold, else the largest admissible s; moduli_for_bits(k, bits) when no weights are loaded.
 * num_moduli in the descriptors is ignored here.  These calls synchronize
 * the handle's stream; out arrays are host memory. */
enum { OZ2_SEL_FEATURES = 48, OZ2_SEL_OUTPUTS = 23, OZ2_SEL_PARAMS = 8791 };
int oz2_selector_features(oz2_handle h, int count, const oz2_gemm_desc_t* descs, int bits, float* features);
/* probs: count x 23 (zeros when no weights are loaded); s_out: count chosen s; learned: 1 where the
 * network chose, 0 where the fallback did (either may be NULL) */
int oz2_selector_choose(oz2_handle h, int count, const oz2_gemm_desc_t* descs, int bits, float* probs, int* s_out,
                        int* learned);
/* n = OZ2_SEL_PARAMS host floats (n = 0: unload), threshold in (0, 1) */
int oz2_selector_set_weights(oz2_handle h, const float* weights, size_t n, float threshold);
/* copies the loaded weights (returns the count, 0 if none) and the threshold */
int oz2_selector_get_weights(oz2_handle h, float* weights, size_t n, float* threshold, size_t* count);
/* Training kernels on device arrays, enqueued on the handle's stream (layout: W1[64][48] b1 W2[64][64] b2
 * W3[23][64] b3, row-major).  forward: X nb x 48 -> H1, H2 nb x 64, Z logits and P = sigmoid(Z) nb x 23
 * (any output may be NULL).  backward: masked binary cross-entropy summed and scaled by inv_norm;
 * D1, D2 nb x 64 and D3 nb x 23 scratch; G the gradient (OZ2_SEL_PARAMS); loss nb x 23 terms or NULL.
 * adam: one Adam step (step >= 1) with decoupled-free L2 term wd*W added to the gradient. */
int oz2_mlp_forward(oz2_handle h, const float* X, int nb, const float* W, float* H1, float* H2, float* Z, float* P);
int oz2_mlp_backward(oz2_handle h, const float* X, int nb, const float* W, const float* H1, const float* H2,
                     const float* Z, const float* Y, const float* M, float inv_norm, float* D1, float* D2, float* D3,
                     float* G, float* loss);
int oz2_mlp_adam(oz2_handle h, float* W, const float* G, float* M, float* V, int step, float lr, float beta1,
                 float beta2, float eps, float wd);

/* Kernel-level entry points, the counterparts of the reference's detail::launch_split and
 * detail::launch_crt as its unit tests call them (tests/test_ozaki2.cpp:215-221, 306-313).
 * oz2_unit_split: rows of op(X) (row r, column p at X[r*ldx + p] if kcontig, else X[r + p*ldx];
 * complex: interleaved, 3 plane groups Re, +-Im, Re+-Im) -> int8 planes [g][t], each rows_pad x kp
 * bytes blocked as the GEMM reads them: with q = p%64, element (r, p) at
 * ((r/16)*(kp/64) + p/64)*1024 + (r%16)*64 + ((((q/2)%4) ^ ((r/4)%4))*16) + 2*(q/8) + q%2
 * (chunk c of a row holds the k pairs (2c, 2c+1) of each group of 8); exponents ex[r]; kp a
 * multiple of 128 and >= k, rows_pad a multiple of 32 and >= rows.
 * oz2_unit_crt: residue bytes R[t][j][ldr] (ldr a multiple of 4, >= mb) -> C[i + j*ldc] =
 * alpha*CRT(...)*2^-(D[i]+E[j]) (+ beta*C).  Both run on `stream` and do not synchronize. */
int oz2_unit_split(void* stream, const double* X, int64_t ldx, int64_t rows, int64_t k, int kcontig, int cplx, int conj,
                   const int32_t* ex, int num_moduli, int tk, int8_t* planes, int64_t rows_pad, int64_t kp);
int oz2_unit_crt(void* stream, const uint8_t* R, int64_t ldr, int64_t mb, int64_t nb, const int32_t* D,
                 const int32_t* E, int num_moduli, double alpha, double beta, double* C, int64_t ldc);
/* The INT8 GEMM with its residue epilogue as one launch: the counterpart of the reference's oneMKL
 * gemm_batch call (hpp:1029-1033), whose int32 products its tests pin only end to end.  Planes of op(A)
 * (num_moduli x mpad x kp) and of op(B) (num_moduli x npad x kp), blocked as oz2_unit_split writes them
 * -> R[t][j][ldr] = (sum_p op(A)_t[i][p] op(B)_t[j][p]) mod m_t as a byte in [0, m_t), for every
 * i < mpad, j < npad.  tile: 256, 128, 64 or 256128 (rows x columns of the macro tile; mpad and npad
 * multiples of it); kp a multiple of 128 up to 65408 (one k-chunk); ldr >= mpad, a multiple of 16. */
int oz2_unit_gemm(void* stream, const int8_t* A, const int8_t* B, int64_t mpad, int64_t npad, int64_t kp,
                  int num_moduli, int tile, uint8_t* R, int64_t ldr);

/* Ozaki Scheme I with L slices, the reference's comparison baseline of its accuracy benchmark
 * (oz1::Ozaki1::gemm, benchmarks/ozaki1_gpu.hpp:100-151): C = A*B (op N, alpha 1, beta 0), rows of A and
 * columns of B scaled to tk = 7L - 2 bits, L balanced base-128 digits, the L(L+1)/2 digit products with
 * i + j < L as INT8 GEMMs and a double-double Horner over the anti-diagonals.  2 <= L <= 9 and
 * k * 4096 < 2^31 (the reference's limits and "oz1: ..." messages).  Enqueued on the handle's stream,
 * in its workspace. */
int oz2_ozaki1_dgemm(oz2_handle h, int64_t m, int64_t n, int64_t k, const double* A, int64_t lda, const double* B,
                     int64_t ldb, double* C, int64_t ldc, int L);

const char* oz2_last_error(void);
const char* oz2_version(void);

#ifdef __cplusplus
}
#endif
#endif
