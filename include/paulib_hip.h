/*
 * paulib_hip.h -- C ABI of the MI355X PauliSum engine (libpaulib_hip.so).
 *
 * Plain pointers, sizes and integer status codes; no C++ or torch types.
 * Every call returns PLH_OK (0) or a negative PLH_ERR_* code; the text of
 * the last error on the calling thread is plh_last_error().  Nothing throws
 * across this boundary.
 *
 * Layout.  A Pauli sum of AS words per plane (AS = ceil(qubits/64)) is the
 * reference's flat plane-major SoA (include/gpu/GpuPauliSum.h:8-12):
 *   x[k*stride + t], z[k*stride + t]   word k of term t (qubit 64k+b = bit b)
 *   flags[t]                           bit0 = i, bit1 = -1 (i^flags)
 *   val[t]                             FP64 coefficient
 * Host arrays passed to upload/download are packed (stride = length).
 *
 * Ordering.  Every result has a deterministic term order: products are
 * out[i*n + j] = lhs[j] * rhs[i]; commutators list anticommuting pairs in
 * (i over a, j over b) order; collated results are sorted by the
 * reference's key (x words from word 0, z words, flags; unsigned).
 *
 * Streams.  Each context owns one HIP stream; all work of a context is
 * ordered on it.  Calls that return a size synchronise that stream.  A
 * context must not be used from two threads at once.
 */
#ifndef PAULIB_HIP_H
#define PAULIB_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    PLH_OK = 0,
    PLH_ERR_ARG = -1,   /* invalid argument (null, qubit out of range, ...) */
    PLH_ERR_AS = -2,    /* unsupported words-per-plane (supported: 1..16) */
    PLH_ERR_SIZE = -3,  /* size limit exceeded (see each call) */
    PLH_ERR_HIP = -4,   /* HIP runtime error */
    PLH_ERR_OOM = -5,   /* device allocation failed or admission refused */
    PLH_ERR_NOGPU = -6, /* no usable GPU */
    PLH_ERR_STATE = -7  /* object unusable after an earlier failure (see the call) */
};

typedef struct plh_ctx plh_ctx; /* device + stream + scratch */
typedef struct plh_sum plh_sum; /* device-resident PauliSum<N,double> */

/* ------------------------------------------------------------------ misc */
const char* plh_last_error(void);
int plh_version(void);
/* Engine build information (gfx target, flags); never fails. */
const char* plh_build_info(void);

/* ------------------------------------------------------------ contexts */
/* stream == NULL creates a private stream at the default priority. */
int plh_ctx_create(int device, void* stream, plh_ctx** out);
int plh_ctx_destroy(plh_ctx* ctx);
int plh_ctx_synchronize(plh_ctx* ctx);
void* plh_ctx_stream(plh_ctx* ctx);
/* Device memory the context holds beyond its live sums: scratch slots plus
 * the free part of its heap's mapped memory (bytes).  plh_ctx_trim gives
 * the scratch, and every heap chunk without a live sum, back to the device. */
int64_t plh_ctx_scratch_bytes(plh_ctx* ctx);
int plh_ctx_trim(plh_ctx* ctx);
/* Per-device autotuning (no reference counterpart): the radix passes' tile
 * caps (root pass and later passes) from timed sweeps of 32768 x 32768 and
 * 65536 x 65536 products on 64 qubits, and the threshold from which
 * plh_batch_multiply_canonical_each runs a product alone.  With cache_path
 * set, a cached entry for this device (architecture, CU count, memory) is
 * loaded instead of sweeping unless force != 0, and a new sweep's result is
 * written there.  The sweep takes ~10 s, needs ~150 GB of device memory and
 * trims the context. */
int plh_ctx_autotune(plh_ctx* ctx, const char* cache_path, int force);
/* The context's tuned values (0 = built-in default). */
int plh_ctx_tuning(const plh_ctx* ctx, int64_t* tile_root_max, int64_t* tile_pass_max, int64_t* solo_pairs);

/* Kernel timing: while enabled, the engine brackets its kernel launches
 * with HIP events on the context's stream.  profile_read synchronises,
 * aggregates per kernel name (names: 32-byte slots): total milliseconds,
 * launch count and the algorithmic HBM bytes of those launches; returns the
 * number of names in *count and clears the records.  No reference
 * counterpart (the reference times its launcher with cudaEvent,
 * bench/GpuMulBench.cpp:53-65). */
int plh_ctx_profile(plh_ctx* ctx, int enable);
int plh_ctx_profile_read(plh_ctx* ctx, int max_entries, char* names, double* total_ms,
                         int64_t* launches, double* alg_bytes, int* count);

/* --------------------------------------------------------- device sums */
/* GpuPauliSum<N,double> (include/gpu/GpuPauliSum.h:34-148): device
 * buffers owned by the engine.  capacity is a hint. */
int plh_sum_create(plh_ctx* ctx, int array_size, int64_t num_qubits, int64_t capacity,
                   plh_sum** out);
int plh_sum_destroy(plh_sum* s);
/* GpuPauliSum::upload (:75-92); host planes packed x[k*n + t]. */
int plh_sum_upload(plh_sum* s, int64_t n, const uint64_t* x, const uint64_t* z,
                   const uint8_t* flags, const double* val);
/* GpuPauliSum::download (:95-114); arrays sized for plh_sum_length(). */
int plh_sum_download(const plh_sum* s, uint64_t* x, uint64_t* z, uint8_t* flags, double* val);
/* Terms [offset, offset + count) only, planes x[k*count + t]; any output may
 * be null to skip it.  The reference's element accessors
 * (get_pauli_string(i), value[i], include/PauliSumCore.h:20-134) for a
 * device sum too large to download whole. */
int plh_sum_download_range(const plh_sum* s, int64_t offset, int64_t count, uint64_t* x, uint64_t* z,
                           uint8_t* flags, double* val);
int64_t plh_sum_length(const plh_sum* s);
int64_t plh_sum_num_qubits(const plh_sum* s);
int plh_sum_set_num_qubits(plh_sum* s, int64_t num_qubits);
int plh_sum_array_size(const plh_sum* s);
/* Raw device planes (valid until the next call that resizes the sum). */
int plh_sum_device_planes(plh_sum* s, uint64_t** x, uint64_t** z, uint8_t** flags, double** val,
                          int64_t* stride);
int plh_sum_copy(const plh_sum* src, plh_sum* dst);

/* --------------------------------------------------- raw launchers */
/* Replaces launch_gpu_mul_soa (include/gpu/paulib_gpu_api.h:25-33):
 * out[i*n + j] = lhs[j] * rhs[i], caller-allocated device buffers of n*m
 * terms, planes strided by n (lhs), m (rhs), n*m (out).  Asynchronous on
 * `stream`.  array_size 1..16, else PLH_ERR_AS (the reference traps). */
int plh_launch_mul_soa(int array_size, const uint64_t* lhs_x, const uint64_t* lhs_z,
                       const uint8_t* lhs_flags, const double* lhs_val, const uint64_t* rhs_x,
                       const uint64_t* rhs_z, const uint8_t* rhs_flags, const double* rhs_val,
                       uint64_t* out_x, uint64_t* out_z, uint8_t* out_flags, double* out_val,
                       int n, int m, void* stream);

/* ------------------------------------------------------- bulk operations */
/* PauliSum::multiply (include/PauliSumOperations.h:255-306): lhs <- lhs*rhs,
 * uncombined, n*m terms; keeps lhs num_qubits.  n*m <= 2^32. */
int plh_multiply(plh_sum* lhs, const plh_sum* rhs);
/* canonicalize(lhs * rhs) without materialising the uncombined product
 * (the fused product->collation path).  n*m <= 2^32 pairs. */
int plh_multiply_canonical(const plh_sum* lhs, const plh_sum* rhs, plh_sum* out);
/* A batch of independent collated products: for every k < njobs,
 * canonicalize(lhs[k] * rhs[k]) exactly as plh_multiply_canonical gives it,
 * all in one collation (one launch sequence for the whole batch).  out
 * becomes a batched sum holding the jobs one after another; job k is terms
 * offsets[k] .. offsets[k + 1] - 1 (offsets: njobs + 1 host entries; read
 * them with plh_sum_download_range).  Operands share out's context and one
 * array size (1 or 2: up to 128 qubits); the sum of n_k * m_k <= 2^32
 * (PLH_ERR_SIZE otherwise: split the batch).  No reference counterpart:
 * the reference multiplies one pair of sums per call
 * (include/PauliSumOperations.h:255-306). */
int plh_batch_multiply_canonical(int njobs, const plh_sum* const* lhs, const plh_sum* const* rhs,
                                 plh_sum* out, int64_t* offsets);
/* A batch of independent collated products, each into its own output sum:
 * outs[k] = canonicalize(lhs[k] * rhs[k]), exactly as
 * plh_multiply_canonical gives it.  Products of at least solo_pairs pairs
 * (0: 2^22), and those over 128 qubits, run alone, straight into their
 * output; the rest are collated together per array size (up to 2^31 pairs
 * per collation, job plane leading the key) and split into their outputs
 * in one launch.  Operands and outputs share one context; outputs are
 * distinct and alias no operand.  No reference counterpart: the reference
 * multiplies one pair of sums per call (include/PauliSumOperations.h:255-306). */
int plh_batch_multiply_canonical_each(int njobs, const plh_sum* const* lhs, const plh_sum* const* rhs,
                                      plh_sum* const* outs, int64_t solo_pairs);
/* commutator / i_commutator (include/PauliSumOperations.h:348-356), the
 * uncombined terms in (i over a, j over b) order; num_qubits = max. */
int plh_commutator(const plh_sum* a, const plh_sum* b, int multiply_by_i, plh_sum* out);
/* No reference counterpart: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The raw commutator of plh_commutator -- [a, b], or i[a, b]
 * with multiply_by_i, one term per anticommuting pair, (i over a, j over b)
 * order, coefficient (2 a_i) b_j -- from caller-held packed planes (a: stride
 * n, b: stride m) into caller-allocated output planes of stride `capacity`,
 * in the argument order of the reference's launch_gpu_commutator_soa
 * (include/gpu/paulib_gpu_api.h:43-53).  *d_count (device memory) receives
 * the number of anticommuting pairs (it need not be zeroed); terms from
 * index `capacity` on are not written, so a count above capacity means the
 * output was cut.  Stream-ordered on `stream` (the calling thread's current
 * device must be the stream's); returns without synchronising.  Allocates
 * nothing: its scratch is a fixed per-device array of ~12 MiB, so
 * n*m <= 2^32 (PLH_ERR_SIZE beyond); calls that overlap on one device race
 * on that array. */
int plh_launch_commutator_soa_noalloc(int array_size, int multiply_by_i, const uint64_t* a_x,
                                      const uint64_t* a_z, const uint8_t* a_flags, const double* a_val,
                                      const uint64_t* b_x, const uint64_t* b_z, const uint8_t* b_flags,
                                      const double* b_val, uint64_t* out_x, uint64_t* out_z,
                                      uint8_t* out_flags, double* out_val, unsigned long long* d_count,
                                      int64_t n, int64_t m, int64_t capacity, void* stream);
/* canonicalize([a, b]) (or of i[a, b]) fused, without the uncombined list. */
int plh_commutator_canonical(const plh_sum* a, const plh_sum* b, int multiply_by_i, plh_sum* out);
/* PauliSum::sort_and_combine (include/PauliSumOperations.h:368-461). */
int plh_sort_and_combine(plh_sum* s);
/* PauliSum::canonicalize (include/PauliSumUtils.h:314-328). */
int plh_canonicalize(plh_sum* s);
/* operator+= (include/PauliSumCore.h:232-237): append, sort_and_combine. */
int plh_add(plh_sum* lhs, const plh_sum* rhs);
/* operator-= (include/PauliSumCore.h:221-250): append rhs with every
 * coefficient times -1, then sort_and_combine. */
int plh_sub(plh_sum* lhs, const plh_sum* rhs);
/* PauliSum::scale (include/PauliSumUtils.h:395-398): every coefficient times
 * c (operator*= with a scalar, include/PauliSumCore.h:191-207; unary minus
 * is scale(-1), :258-262). */
int plh_scale(plh_sum* s, double c);
/* PauliSum::lookup (include/PauliSumUtils.h:253-268): the coefficient of the
 * pattern (x, z) with phase flags, summed over every term with that
 * pattern (phase relative to flags; terms off by +-i skipped); 0 if absent. */
int plh_lookup(const plh_sum* s, int array_size, const uint64_t* x, const uint64_t* z, uint8_t flags,
               double* out);
/* PauliSum::trace (:231-243): 2^num_qubits * the sum of the identity terms'
 * real coefficients. */
int plh_trace(const plh_sum* s, double* out);
/* PauliSum::frobenius_norm2 (:403-410): sum of squared coefficients (of a
 * canonical sum). */
int plh_frobenius_norm2(const plh_sum* s, double* out);
/* PauliSum::zero_state_expectation (include/PauliSumUtils.h:162-177). */
int plh_zero_state_expectation(const plh_sum* s, double* out);
/* PauliSum::truncate_weight / cutoff (include/PauliSumUtils.h:377-393). */
int plh_truncate_weight(plh_sum* s, int64_t max_weight);
int plh_cutoff(plh_sum* s, double eps);
/* PauliSum::apply_H / apply_S / apply_CNOT (include/PauliSumClifford.h:
 * 43-81, 212-239). */
int plh_apply_H(plh_sum* s, int64_t q);
int plh_apply_S(plh_sum* s, int64_t q);
int plh_apply_CNOT(plh_sum* s, int64_t control, int64_t target);
/* The whole gate set of include/PauliSumClifford.h (:44-498), one
 * table-driven launch per gate.  Gate ids for plh_apply_clifford:
 *   0 I  1 X  2 Y  3 Z  4 H  5 S  6 H_YZ  7 H_XY  8 sqrt_X  9 sqrt_Y
 *  10 CNOT 11 CZ 12 CY 13 XC_X 14 XC_Z 15 XC_Y 16 YC_X 17 YC_Z 18 YC_Y
 *  19 SWAP 20 iSWAP (= SWAP, CZ, S(c), S(t) as :493-498, one launch).
 * q1 is ignored for ids 0-9.  A qubit outside the sum's words, or the same
 * qubit as control and target, is PLH_ERR_ARG (the reference checks
 * neither).  The named entry points below replace PauliSum::apply_<name>
 * one for one (apply_H/S/CNOT are declared above). */
#define PLH_GATE_I 0
#define PLH_GATE_X 1
#define PLH_GATE_Y 2
#define PLH_GATE_Z 3
#define PLH_GATE_H 4
#define PLH_GATE_S 5
#define PLH_GATE_H_YZ 6
#define PLH_GATE_H_XY 7
#define PLH_GATE_SQRT_X 8
#define PLH_GATE_SQRT_Y 9
#define PLH_GATE_CNOT 10
#define PLH_GATE_CZ 11
#define PLH_GATE_CY 12
#define PLH_GATE_XC_X 13
#define PLH_GATE_XC_Z 14
#define PLH_GATE_XC_Y 15
#define PLH_GATE_YC_X 16
#define PLH_GATE_YC_Z 17
#define PLH_GATE_YC_Y 18
#define PLH_GATE_SWAP 19
#define PLH_GATE_ISWAP 20
int plh_apply_clifford(plh_sum* s, int gate, int64_t q0, int64_t q1);
/* apply_I :208-210, apply_X :162-174, apply_Y :177-190, apply_Z :193-205,
 * apply_H_YZ :84-100, apply_H_XY :103-119, apply_sqrt_X :122-138,
 * apply_sqrt_Y :141-159. */
int plh_apply_I(plh_sum* s, int64_t q);
int plh_apply_X(plh_sum* s, int64_t q);
int plh_apply_Y(plh_sum* s, int64_t q);
int plh_apply_Z(plh_sum* s, int64_t q);
int plh_apply_H_YZ(plh_sum* s, int64_t q);
int plh_apply_H_XY(plh_sum* s, int64_t q);
int plh_apply_sqrt_X(plh_sum* s, int64_t q);
int plh_apply_sqrt_Y(plh_sum* s, int64_t q);
/* apply_CZ :242-266, apply_CY :269-294, apply_XC_X :297-321,
 * apply_XC_Z :324-348, apply_XC_Y :351-376, apply_YC_X :379-404,
 * apply_YC_Z :407-432, apply_YC_Y :435-462, apply_SWAP :465-490,
 * apply_iSWAP :493-498 (c = control / first qubit, t = target). */
int plh_apply_CZ(plh_sum* s, int64_t c, int64_t t);
int plh_apply_CY(plh_sum* s, int64_t c, int64_t t);
int plh_apply_XC_X(plh_sum* s, int64_t c, int64_t t);
int plh_apply_XC_Z(plh_sum* s, int64_t c, int64_t t);
int plh_apply_XC_Y(plh_sum* s, int64_t c, int64_t t);
int plh_apply_YC_X(plh_sum* s, int64_t c, int64_t t);
int plh_apply_YC_Z(plh_sum* s, int64_t c, int64_t t);
int plh_apply_YC_Y(plh_sum* s, int64_t c, int64_t t);
int plh_apply_SWAP(plh_sum* s, int64_t c, int64_t t);
int plh_apply_iSWAP(plh_sum* s, int64_t c, int64_t t);
/* Host-only, no GPU needed (no counterpart in the reference): the gate's
 * 16-entry table, new local bits xc | zc<<1 | xt<<2 | zt<<3 as nibble b of
 * *table for input bits b, and the sign flip as bit b of *sign.  For tests
 * and tools that check the gate set without a device. */
int plh_clifford_table(int gate, uint64_t* table, uint32_t* sign);
/* PauliSum::apply_RX / apply_RY / apply_RZ (include/PauliSumOperations.h:
 * 28-104): half-angle cos/sin (host libm), the new terms (iP_q) Q appended
 * in term order, then sort_and_combine; no fast paths, zeros kept. */
int plh_apply_RX(plh_sum* s, int64_t q, double angle);
int plh_apply_RY(plh_sum* s, int64_t q, double angle);
int plh_apply_RZ(plh_sum* s, int64_t q, double angle);
/* PauliSum::elementwise_multiply (include/PauliSumOperations.h:220-250):
 * term i <- term_i * rhs_i, value v_i * rhs.v_i.  rhs must have the same
 * length and word count (PLH_ERR_ARG otherwise; the reference reads rhs[i]
 * unchecked) and the same context; rhs may be s itself. */
int plh_elementwise_multiply(plh_sum* s, const plh_sum* rhs);
/* PauliSum::evolve (include/PauliSumOperations.h:106-183): rotation terms
 * r = 0..nrot-1 in order, Pauli words px/pz[r*as + k], flags pflags[r],
 * angle theta[r]; cos/sin are taken on the host (libm). */
int plh_evolve(plh_sum* s, int64_t nrot, const uint64_t* px, const uint64_t* pz,
               const uint8_t* pflags, const double* theta);
/* apply_truncation (include/PauliPropagation.h:35-47); max_weight < 0,
 * cutoff <= 0, max_terms < 0 mean "not set". */
int plh_apply_truncation(plh_sum* s, int64_t max_weight, double cutoff, int64_t max_terms);
/* apply_truncation(O, opts, keep) (include/PauliPropagation.h:35-47): the
 * trim step is trim(max_terms, keep); keep may be NULL. */
int plh_apply_truncation_keep(plh_sum* s, int64_t max_weight, double cutoff, int64_t max_terms,
                              const plh_sum* keep);
/* PauliSum::trim(max_terms) (include/PauliSumUtils.h:330-343): keep the
 * max_terms terms of largest |c|, in term order.  Among equal magnitudes the
 * reference leaves the choice to std::nth_element; here the lower index
 * is kept. */
int plh_trim(plh_sum* s, int64_t max_terms);
/* PauliSum::trim(max_terms, keep) (include/PauliSumUtils.h:345-375): terms
 * whose (x, z) pattern appears in keep are never dropped; the other terms
 * keep the max(0, max_terms - protected) largest |c| (ties as plh_trim), in
 * term order.  keep must share s's context and array size. */
int plh_trim_keep(plh_sum* s, int64_t max_terms, const plh_sum* keep);
/* propagate (include/PauliPropagation.h:53-63): gates in reverse order,
 * each a one-term evolve then apply_truncation. */
int plh_propagate(plh_sum* s, int64_t nrot, const uint64_t* px, const uint64_t* pz,
                  const uint8_t* pflags, const double* theta, int64_t max_weight, double cutoff,
                  int64_t max_terms);
/* Heisenberg run of a mixed program (application order, processed last to
 * first): kind 0 = H(q0), 1 = S(q0), 2 = CNOT(q0, q1), 3 = rotation by
 * theta about Pauli word px/pz[l*as..], 16 + g = Clifford gate id g of
 * plh_apply_clifford on (q0, q1); each rotation is evolve +
 * apply_truncation as in propagate().  No reference counterpart: it
 * composes the reference's methods the way a brickwork driver would. */
int plh_run_program(plh_sum* s, int64_t nops, const int32_t* kind, const int64_t* q0,
                    const int64_t* q1, const uint64_t* px, const uint64_t* pz,
                    const double* theta, int64_t max_weight, double cutoff, int64_t max_terms);

/* ---- Batched propagation server (continuous batching).
 * No reference counterpart: the reference runs one propagate() per call
 * (include/PauliPropagation.h:53-71).  A server keeps every admitted job's
 * terms in one device-resident sum tagged by job slot, sorted by (job, key).
 * A full plh_server_step advances every running job by its Clifford gates up
 * to and including its next rotation (evolve + apply_truncation with the
 * job's own max_weight and cutoff, exactly as plh_run_program) and collates
 * the batch; a sorted step advances, without a sort, only the jobs whose next
 * op is a rotation, the others waiting (at most 24 steps) for the next full
 * step; either takes a fixed number of launches for all jobs together.  Jobs
 * are admitted at full steps while a slot is free and the batch fits the HBM
 * budget; a finished job's result is its
 * zero_state_expectation (include/PauliSumUtils.h:162-177) and, with
 * keep_terms, its final sum in canonical order.  A job's result does not
 * depend on which other jobs share its steps. */
typedef struct plh_server plh_server;
/* array_size 1 or 2 (up to 128 qubits); max_jobs slots; hbm_budget_bytes <= 0
 * takes half the device's free memory. */
int plh_server_create(plh_ctx* ctx, int array_size, int max_jobs, int64_t hbm_budget_bytes,
                      int keep_terms, plh_server** out);
int plh_server_destroy(plh_server* s);
/* Queue a job: observable (n terms, planes x[k*n+t] for k < ceil(nq/64),
 * flags, values; its (x, z, imag) keys must be distinct, as after
 * canonicalize) and a program as for plh_run_program.  max_weight < 0,
 * cutoff <= 0 and max_terms < 0 disable those truncations (TruncationOptions,
 * include/PauliPropagation.h:26-47: after each rotation the job is
 * canonicalized, filtered by weight and cutoff, then trimmed to its max_terms
 * largest |c| in term order, equal magnitudes lower index first). */
int plh_server_submit(plh_server* s, int64_t n, const uint64_t* x, const uint64_t* z,
                      const uint8_t* flags, const double* values, int64_t num_qubits, int64_t nops,
                      const int32_t* kind, const int64_t* q0, const int64_t* q1,
                      const uint64_t* px, const uint64_t* pz, const double* theta,
                      int64_t max_weight, double cutoff, int64_t max_terms, int64_t* handle);
/* One step: admit, advance every running job, finish jobs whose program is
 * done.  Returns the running and queued job counts.  A job that cannot be
 * admitted (its upload fails) stays queued and the call returns the error;
 * a failure in the step's device work leaves the batch with half-applied
 * gates, so every later step and submit returns PLH_ERR_STATE. */
int plh_server_step(plh_server* s, int* running, int* queued);
/* Finished jobs in completion order: handle, expectation, final term count,
 * latency from submission (seconds). */
int plh_server_poll(plh_server* s, int max_results, int64_t* handles, double* expectation,
                    int64_t* terms, double* latency_s, int* count);
/* Final terms of a finished job (keep_terms servers): planes for the job's
 * own array size, `terms` entries each (as reported by poll). */
int plh_server_take_terms(plh_server* s, int64_t handle, uint64_t* x, uint64_t* z,
                          uint8_t* flags, double* values);
int64_t plh_server_batch_terms(const plh_server* s);
int64_t plh_server_steps(const plh_server* s);

#ifdef __cplusplus
}
#endif
#endif /* PAULIB_HIP_H */
