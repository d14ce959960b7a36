/*
 * qipgpu.h -- C ABI of the MI355X (gfx950) state-vector engine for QIP.
 *
 * The engine is the native layer under a QIP StateType backend (qip/backend.py:14-65). Every
 * entry point below replaces one native function of the reference's CythonBackend stack
 * (qip/backend.py:68-175 -> qip/ext/kronprod.pyx, qip/ext/func_apply.pyx) and cites it.
 *
 * Conventions (the reference's):
 *   - qubit q is state-index bit (n-1-q) (qip/ext/kronprod.pyx:168,187-189);
 *   - inside a group of qubit indices the first index is the most significant bit of the
 *     sub-matrix / register value (kronprod.pyx:186-189, func_apply.pyx:82-85);
 *   - amplitudes are interleaved (re, im): complex128 (QIPGPU_C128) or complex64 (QIPGPU_C64),
 *     the memory layout of numpy.complex128 / numpy.complex64;
 *   - matrices are row-major, 2^k x 2^k, interleaved complex128.
 * All entry points return QIPGPU_OK (0) or a negative status; qipgpu_last_error() then holds a
 * message (per calling thread). Calls on one state handle must come from one thread at a time;
 * each handle owns one HIP stream. Calls that return host data synchronise that stream.
 */
#ifndef QIPGPU_H
#define QIPGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qipgpu_state qipgpu_state;
typedef struct qipgpu_table qipgpu_table;

enum { QIPGPU_C128 = 0, QIPGPU_C64 = 1 };

enum {
    QIPGPU_OK = 0,
    QIPGPU_EINVAL = -1,   /* bad argument: Python raises ValueError */
    QIPGPU_ENOMEM = -2,   /* device allocation failed: MemoryError */
    QIPGPU_EHIP = -3,     /* HIP runtime error: RuntimeError */
    QIPGPU_EZERODIV = -4, /* measured probability is 0: ZeroDivisionError, as 1.0/p raises in
                             kronprod.pyx:428 */
    QIPGPU_EUNSUP = -5    /* operator outside what the engine applies: NotImplementedError */
};

/* Largest number of (non-control) target qubits of one dense or diagonal operator. */
#define QIPGPU_MAX_DENSE 12
/* Largest number of qubits of a state. */
#define QIPGPU_MAX_QUBITS 34

const char *qipgpu_last_error(void);
int qipgpu_device_count(int32_t *count);
const char *qipgpu_version(void);

/* ---- state lifecycle: CythonBackend.make_state (qip/backend.py:73-104) ---------------------- */

/* Allocate a 2^n-amplitude state on the current device. Contents undefined until an init. */
int qipgpu_create(int32_t n, int32_t dtype, qipgpu_state **out);
int qipgpu_destroy(qipgpu_state *s);
int qipgpu_info(const qipgpu_state *s, int32_t *n, int32_t *dtype);

/* Product state: amplitude i = prod over groups g (in order) of feed_g[value of i on group g],
 * zero where any unfed qubit is 1 (qip/backend.py:88-101, qip/util.py:108-124). With
 * n_groups == 0 the state is |0...0> (backend.py:90-91). `feeds` concatenates the 2^len_g
 * complex128 entries of each feed. */
int qipgpu_init_product(qipgpu_state *s, int32_t n_groups, const int32_t *group_len,
                        const int32_t *group_idx, const double *feeds);

/* Copies of amplitude windows [start, end) in the state's own dtype:
 *   get:  get_state / get_relative_range   (qip/backend.py:106-107,168-169)
 *   set:  overwrite_relative_range          (qip/backend.py:171-172)
 *   add:  addto_relative_range              (qip/backend.py:174-175) */
int qipgpu_get_range(qipgpu_state *s, int64_t start, int64_t end, void *host);
int qipgpu_set_range(qipgpu_state *s, int64_t start, int64_t end, const void *host);
int qipgpu_add_range(qipgpu_state *s, int64_t start, int64_t end, const void *host);

/* ---- gate application: cdot_loop (qip/ext/kronprod.pyx:41-200) -------------------------------
 * Applies, in place, the operator that is `mat` on the `targets` qubits, conditioned on every
 * `controls` qubit being 1 (the C_MAT structure diag(I, M), kronprod.pyx:215-225). The engine
 * finds further control qubits and diagonal structure inside `mat` itself, so a caller may pass
 * a whole controlled operator as one dense matrix. k <= QIPGPU_MAX_DENSE. Up to 3 targets are
 * applied from registers, 4..6 through LDS tiles on FP64 MFMA, 7..12 as an FP64 GEMM out of place
 * (a temporary of 2^(n - controls) amplitudes). */
int qipgpu_apply_matrix(qipgpu_state *s, int32_t k, const int32_t *targets, const double *mat,
                        int32_t n_ctrl, const int32_t *controls);

/* Any kronselect_dot descriptor component, by the reference's own formula (cdot_loop,
 * qip/ext/kronprod.pyx:137-197): out[r] = sum over the 2^u columns c that agree with r outside
 * the union (u qubits) of all groups' indices, in increasing order, of
 * prod_g M_g[r|_g, c|_g] * in[c], with M_g's entries by calc_mat_entry (kronprod.pyx:203-235).
 * Group g has group_len[g] indices (concatenated in group_idx, first = MSB); its first levels[g]
 * indices are C_MAT levels (controls); the inner matrix is kinds[g] = 1 (NUMPY: a dense
 * 2^(len - levels) square matrix, row-major complex128, the inner matrices concatenated in `mats`)
 * or 3 (SWAP_MAT of block width params[g]). Groups may share indices. Costs 2^(n+u) entry
 * products and a temporary state (out of place, as the reference); u <= 30. For descriptors
 * no specialised sweep covers. */
int qipgpu_apply_kron(qipgpu_state *s, int32_t n_groups, const int32_t *group_len, const int32_t *group_idx,
                      const int32_t *levels, const int32_t *kinds, const int32_t *params, const double *mats);

/* Product of single-qubit operators, mats[4j..4j+3] (row-major 2x2 complex128) on targets[j], for
 * k <= QIPGPU_MAX_PRODUCT distinct qubits, in one sweep: the uncontrolled single-qubit factors of
 * one kronselect_dot descriptor (e.g. H on a register, qip/operators.py:97-99), which cdot_loop
 * also applies in one sweep. Computed in FP64 on a tile held in LDS, rounded once per amplitude;
 * needs n >= 12. */
#define QIPGPU_MAX_PRODUCT 12
int qipgpu_apply_product(qipgpu_state *s, int32_t k, const int32_t *targets, const double *mats);

/* Diagonal operator diag(d) on `targets`, conditioned on `controls` (k <= QIPGPU_MAX_DENSE).
 * Held-back gates (this entry point and qipgpu_apply_matrix), once controls are found:
 *   - an operator diagonal on at most 2 targets (CZ, controlled phases, T, S, Rz) joins a run of
 *     up to 16 such gates, applied in one sweep over the amplitudes whose bits are 1 on every
 *     control the run shares: each amplitude gets each gate's factor in turn, rounded after each;
 *   - on a state of 12 qubits or more, a dense single-qubit operator (H, X, CNOT, Toffoli, any
 *     2x2 under any controls) with at most 8 of its bits above state bit 3 joins a run of up to 32
 *     such gates whose bits, with state bits 0..3, fit one tile of 2^12 amplitudes: a tile is read
 *     once, the gates are applied to it one after the other in on-chip memory with the arithmetic
 *     of their own sweeps (FP64 from zero, rounded to the state's precision after each gate), and
 *     the tile is written once.
 * A run waits until a gate that does not fit it or any other call on the state (qipgpu_synchronize,
 * qipgpu_fused_flush, any read) applies it first, in order. A run goes gate by gate where separate
 * sweeps would move fewer bytes. The state is bit for bit what one launch per gate gives. */
int qipgpu_apply_diagonal(qipgpu_state *s, int32_t k, const int32_t *targets, const double *diag,
                          int32_t n_ctrl, const int32_t *controls);

/* ---- gate graphs: no counterpart in the reference ----------------------------------------------
 * For callers that apply the same sequence of small gates again and again to states of one size
 * (a circuit run many times on a few hundred thousand amplitudes is bound by kernel launches, not
 * by HBM). Between qipgpu_capture_begin and qipgpu_capture_end the state's gate calls add their
 * kernels to a graph, each after the one before, instead of launching them (the graph is built node
 * by node; no queue is put in a capture mode, so other states and threads are not affected). Only
 * qipgpu_apply_matrix on up to 3 targets, qipgpu_apply_diagonal on up to 6 and qipgpu_apply_product
 * may be called, whose operands travel as kernel arguments; any other call that reads, writes or
 * stages data for the state (wider operators, func_apply, reductions, copies, initialisation,
 * fused blocks) returns QIPGPU_EUNSUP and makes qipgpu_capture_end drop the recording with
 * QIPGPU_EUNSUP. Nothing recorded has run: launch the graph to apply it. A graph updates the
 * state buffer it was captured on; it can be launched on any state that holds that buffer
 * (qipgpu_graph_matches). A destroyed state of up to 256 MiB leaves its buffer and queue to the next
 * state of the same size on the same device (at most 8 such buffers are kept, until the process
 * ends), so a state made again after the first was destroyed can launch the first one's graphs. */
typedef struct qipgpu_graph qipgpu_graph;
int qipgpu_capture_begin(qipgpu_state *s);
int qipgpu_capture_end(qipgpu_state *s, qipgpu_graph **out);
/* 1 when g updates s's buffer (same buffer, size, dtype and device), else 0. */
int qipgpu_graph_matches(const qipgpu_state *s, const qipgpu_graph *g);
/* Enqueue the graph's kernels on s's queue (QIPGPU_EINVAL unless qipgpu_graph_matches). Launches of
 * one graph from several threads at once are not supported. */
int qipgpu_graph_launch(qipgpu_state *s, qipgpu_graph *g);
/* Waits for the graph's last launch to finish, then frees it. */
int qipgpu_graph_destroy(qipgpu_graph *g);

/* ---- classical functions: func_apply (qip/ext/func_apply.pyx:11-112) --------------------------
 * |x>|q>|r> -> |x>|f(x) xor q>|r> with x read MSB-first from reg1 and only the low n_reg2 bits of
 * f(x) used (func_apply.pyx:82-97). `table[x]` = f(x) for x in [0, 2^n_reg1). In place. */
int qipgpu_func_apply(qipgpu_state *s, int32_t n_reg1, const int32_t *reg1, int32_t n_reg2,
                      const int32_t *reg2, const int32_t *table);
/* The same with the f table kept resident on the device across calls. */
int qipgpu_table_create(int32_t n_reg1, const int32_t *table, qipgpu_table **out);
int qipgpu_table_destroy(qipgpu_table *t);
int qipgpu_func_apply_table(qipgpu_state *s, int32_t n_reg1, const int32_t *reg1, int32_t n_reg2,
                            const int32_t *reg2, const qipgpu_table *t);
/* qipgpu_func_apply_table with flags. QIPGPU_FUNC_REFERENCE_BOUND reproduces the reference's loop
 * over the rest qubits (func_apply.pyx:107-111), which sets only the first n_reg2 of them (in qubit
 * order): when more than n_reg2 qubits lie outside both registers, every amplitude with one of the
 * other rest qubits at 1 ends 0, as the reference's zeroed output leaves it (func_apply.pyx:42-44).
 * With n_rest <= n_reg2 the flag changes nothing. Without it the permutation is applied to the
 * whole state. */
#define QIPGPU_FUNC_REFERENCE_BOUND 1
int qipgpu_func_apply_table_ex(qipgpu_state *s, int32_t n_reg1, const int32_t *reg1, int32_t n_reg2,
                               const int32_t *reg2, const qipgpu_table *t, int32_t flags);

/* ---- reductions and measurement (qip/ext/kronprod.pyx:238-490) ------------------------------- */

/* prob_magnitude / total_prob: sum |psi_i|^2 in FP64 (kronprod.pyx:318-326, backend.py:123-124) */
int qipgpu_total_prob(qipgpu_state *s, double *out);

/* measure_probabilities (kronprod.pyx:238-261): out[2^k], qubit indices[j] on bit j (LSB-first) */
int qipgpu_probabilities(qipgpu_state *s, int32_t k, const int32_t *indices, double *out);

/* qipgpu_probabilities without any device allocation: no counterpart in the reference; for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). The reduction's results and partials live in a 16 MiB device array of the library,
 * one per GPU, held until the process ends; k <= 20 (larger k: QIPGPU_EINVAL). The state must have
 * no recorded, unlaunched gates (call qipgpu_fused_flush first; QIPGPU_EINVAL otherwise) and the
 * call's device must be the state's. Calls that overlap on one GPU (several threads or streams)
 * race on the shared array. Same results as qipgpu_probabilities to FP64 rounding (the partial sums
 * may be grouped differently when the fixed array is the limit). */
int qipgpu_probabilities_noalloc(qipgpu_state *s, int32_t k, const int32_t *indices, double *out);

/* Marginals in measurement order: out[m] = sum of |psi_i|^2 over i whose measured bits read m,
 * the smallest measured qubit being the most significant bit of m (entwine_bit, util.pyx:2-23). */
int qipgpu_marginals(qipgpu_state *s, int32_t k, const int32_t *indices, double *out);

/* measure_top_probabilities (kronprod.pyx:264-315): the top_k largest marginals (measurement
 * order), descending, ties by ascending m. Writes *count = min(top_k, 2^k) entries. The marginals
 * stay on the device and are selected there for min(top_k, 2^k) <= 1024; wider selections copy
 * the 2^k marginals to the host. */
int qipgpu_top_probabilities(qipgpu_state *s, int32_t k, const int32_t *indices, int32_t top_k,
                             int64_t *ms, double *ps, int32_t *count);

/* soft_measure (kronprod.pyx:329-383) with the uniform draw u supplied by the caller:
 * r = u * total_prob, then the sequential walk r -= p_m; stop at the first r <= 0. With
 * measured >= 0 only that outcome is evaluated. Returns the raw marginal in *p. */
int qipgpu_soft_measure(qipgpu_state *s, int32_t k, const int32_t *indices, double u,
                        int64_t measured, int64_t *m, double *p);

/* measure (kronprod.pyx:386-446): soft_measure unless both measured >= 0 and measured_prob > 0
 * are given, then collapse in place: psi_i *= sqrt(1/p) where the measured bits read m, else 0. */
int qipgpu_measure(qipgpu_state *s, int32_t k, const int32_t *indices, double u, int64_t measured,
                   double measured_prob, int64_t *m, double *p);

/* reduce_measure (kronprod.pyx:449-490): like measure, but amplitude i of the result is the
 * collapsed amplitude with measured bits m and unmeasured bits the low n-k bits of i, so the
 * reduced state appears tiled 2^k times (the reference keeps its 2^n arena, backend.py:139-142). */
int qipgpu_reduce_measure(qipgpu_state *s, int32_t k, const int32_t *indices, double u,
                          int64_t measured, double measured_prob, int64_t *m, double *p);

/* reduce_measure with the state shrunk to the n-k unmeasured qubits (kronprod.pyx:449-490 with
 * the arena resized, the intended reduce of backend.py:136-151): afterwards the handle holds
 * 2^(n-k) amplitudes, amplitude j being the collapsed amplitude with measured bits m and
 * unmeasured bits j (their order kept). Needs k < n. The old buffer is freed after the slice is
 * gathered (peak memory: the state plus the slice). */
int qipgpu_reduce_measure_shrink(qipgpu_state *s, int32_t k, const int32_t *indices, double u,
                                 int64_t measured, double measured_prob, int64_t *m, double *p);

/* ---- precision form: fused complex64 blocks on MFMA (no reference counterpart: it applies the
 * product of several consecutive cdot_loop operators at once) ---------------------------------- */

/* Apply a dense 2^k x 2^k operator (k <= 6; row-major complex128, targets[0] = MSB) to a
 * QIPGPU_C64 state with split-BF16 MFMA products and FP32 accumulation. The state is first
 * multiplied by the pending FP64 renormalisation factor left by the previous fused call; the
 * norm of the result is accumulated in FP64 and becomes the next pending factor. */
int qipgpu_fused_apply(qipgpu_state *s, int32_t k, const int32_t *targets, const double *mat);
/* Diagnostic, with no counterpart in the reference: load and store every amplitude in place with
 * exactly the tile pattern, grid and double buffering qipgpu_fused_apply uses for these targets, and no
 * arithmetic. The state is unchanged; the launch time is the memory-pattern floor of such a block. */
int qipgpu_fused_pattern_sweep(qipgpu_state *s, int32_t k, const int32_t *targets);
/* Host only (no GPU call): the 2^k x 2^k matrix (row-major complex128) of an ordered list of gates
 * on the k <= 6 distinct qubits `qubits` (first = MSB of the result's index), out = G_last ... G_1.
 * Gate g acts on its gate_k[g] qubits gate_qubits[6 g ..] (all among `qubits`, first = MSB of its
 * own matrix) with the 2^gate_k x 2^gate_k matrix that follows the previous gates' in `mats`: the
 * operator of a fused block, i.e. the product of the consecutive cdot_loop operators
 * (kronprod.pyx:41-200) it replaces. */
int qipgpu_compose(int32_t k, const int32_t *qubits, int32_t n_gates, const int32_t *gate_k,
                   const int32_t *gate_qubits, const double *mats, double *out);
/* Record gates for online fusion: the fused forms' replacement of the gate-by-gate
 * kronselect_dot -> cdot_loop calls (kronprod.pyx:300-345, 41-200), on a complex64 state (the
 * precision form) or a complex128 one (FP64 blocks). Gate g has gate_k[g]
 * (1..width) distinct qubits gate_qubits[6 g ..] (first = MSB of its matrix) and the next
 * 4^gate_k[g] interleaved complex entries of `mats`. Gates are planned into blocks of at most
 * `width` qubits by the light-cone rule (a block grows from the earliest gate whose predecessors are
 * all placed, absorbing every such gate within its qubits and taking the qubits of the one, preferably
 * sharing a qubit with it, that a two-step look-ahead scores best, while the width allows). A block
 * that no later gate can change is composed (qipgpu_compose) and launched when streaming
 * (qipgpu_fused_apply on a complex64 state with n >= 12; the FP64 dense sweep of qipgpu_apply_matrix
 * on a complex128 state, and below 12 qubits); the record is planned again every
 * max(16, pending / 8) gates. For n >= 12 a run
 * of consecutive blocks holding single-qubit gates only is one product of per-qubit 2x2s on the
 * product sweep (qipgpu_apply_product, FP64 arithmetic); a trailing run waits until a final block
 * that is not single-qubit-only follows it, so runs launch whole. Streaming and deferred recording
 * launch the same blocks. Every other entry point that touches the state, and qipgpu_synchronize,
 * launches what is recorded first; qipgpu_init_product drops it. All gates are checked before any is
 * recorded. */
int qipgpu_fused_record(qipgpu_state *s, int32_t n_gates, const int32_t *gate_k, const int32_t *gate_qubits,
                        const double *mats);
/* Fusion width (1..6, default 6) and streaming (1, default) or deferred (0) launches; launches what
 * is recorded first. */
int qipgpu_fused_set_recording(qipgpu_state *s, int32_t width, int32_t streaming);
/* Compose and launch every recorded gate (the blocks back to back), after the held-back diagonal
 * gates (qipgpu_apply_diagonal); on any state, launches what the engine holds back. */
int qipgpu_fused_flush(qipgpu_state *s);
/* Recorded gates not launched yet, and blocks launched since the first recording on this state. */
int qipgpu_fused_recorded(const qipgpu_state *s, int64_t *pending_gates, int64_t *blocks_launched);
/* Host only (no GPU call): the recorder's streaming fusion plan for a gate list on n qubits (<= 64).
 * gate_block[g] = global number of gate g's block (blocks numbered from 0 in launch order);
 * final_after[g] = number of blocks found final, by the plans the streaming recorder makes, once
 * gate g is recorded (launched then, but for a trailing run of single-qubit-only blocks). */
int qipgpu_fuse_online(int32_t n, int32_t width, int32_t n_gates, const int32_t *gate_k, const int32_t *gate_qubits,
                       int64_t *gate_block, int64_t *final_after);
/* Host only (no GPU call): the recorder's launches for a gate list on n qubits recorded with the given
 * width and streaming setting, then flushed. launch_of_block[b] (room for n_gates entries) = index
 * of the launch that applies block b; *n_blocks = number of blocks. Blocks that share a launch are a
 * run of single-qubit-only blocks on the product sweep (n >= 12). */
int qipgpu_fuse_launches(int32_t n, int32_t width, int32_t streaming, int32_t n_gates, const int32_t *gate_k,
                         const int32_t *gate_qubits, int64_t *launch_of_block, int64_t *n_blocks);
/* Launch recorded gates, then apply the pending renormalisation factor (if any) with one scaling
 * sweep. */
int qipgpu_fused_finish(qipgpu_state *s);
/* Norm of the state after the last fused call, before its renormalisation, in FP64. */
int qipgpu_fused_last_norm(qipgpu_state *s, double *out);

/* ---- comparison helpers ------------------------------------------------------------------------ */

/* <a|b> in FP64 (either dtype each, same n) and both squared norms. */
int qipgpu_overlap(qipgpu_state *a, qipgpu_state *b, double *re, double *im, double *norm_a,
                   double *norm_b);
/* max_i |a_i - b_i| in FP64. */
int qipgpu_max_abs_diff(qipgpu_state *a, qipgpu_state *b, double *out);
/* Launch recorded gates, then block until all work queued on the state's stream has finished. */
int qipgpu_synchronize(qipgpu_state *s);
/* The hipStream_t every kernel on this state is launched on, for callers that time or order
 * their own work against it (events, copies). */
int qipgpu_get_stream(qipgpu_state *s, void **stream);

/* ---- batching form: many independent circuits of uneven qubit counts -------------------------
 * One batch holds `count` complex64 states (each in its own slice of one device arena) and a
 * gate list per state; a run applies every state's gates. It replaces `count` independent
 * reference runs, each a CythonBackend.make_state (qip/backend.py:73-104) of |0...0> followed by
 * one kronselect_dot (qip/backend.py:112-116 -> kronprod.pyx:41-200) per gate, as
 * qip/pipeline.py:71-135 issues them; the reference has no batched entry point.
 *
 * Gates are given packed, in program order per state (states may interleave): gate g acts on
 * state gate_state[g], on the gate_k[g] qubits gate_qubits[6 g .. 6 g + k - 1] (first = MSB of
 * the matrix index, qubit q = state bit n-1-q), with the dense 2^k x 2^k row-major complex128
 * matrix that follows the previous gates' matrices in `mats` (controls are part of the dense
 * matrix). 1 <= k <= 6.
 *
 * States with n <= lds_max_qubits run their whole gate list in one workgroup with the state in
 * LDS, gate by gate, in FP64 arithmetic. Larger states have their gates fused greedily into
 * blocks of at most fusion_width qubits (the rule of the precision form), applied with the
 * split-BF16 MFMA tile kernel; block j of every large state runs in wave j, and each block's
 * output is renormalised by the FP64 norm of the previous block's output. Every state that has
 * gates leaves a run renormalised to FP64 norm 1; a state without gates is left as it is. */
typedef struct qipgpu_batch qipgpu_batch;

typedef struct {
    int32_t fusion_width;   /* 1..6: qubits per fused block of the large states (default 6) */
    int32_t lds_max_qubits; /* 11..13: states with n <= this run whole in LDS (default 13) */
    int32_t chunk_tiles;    /* 1..64: 4096-amplitude tiles per workgroup in a block pass (default 8) */
    int32_t use_graph;      /* 1: a run replays one captured hipGraph (default); 0: eager launches */
} qipgpu_batch_tuning;

/* Host-only planning (no GPU call): validate a packed gate list and return the fusion the batch
 * would use. gate_block[g] = index of gate g's block within its state, or -1 if the state runs in
 * LDS; state_blocks[i] = blocks of state i (0 for LDS states). */
int qipgpu_batch_plan(int32_t count, const int32_t *n_qubits, int64_t n_gates, const int32_t *gate_state,
                      const int32_t *gate_k, const int32_t *gate_qubits, int32_t fusion_width,
                      int32_t lds_max_qubits, int32_t *gate_block, int32_t *state_blocks);
/* Allocate `count` states (1 <= n_qubits[i] <= 30), all |0...0>, with no gates. */
int qipgpu_batch_create(int32_t count, const int32_t *n_qubits, qipgpu_batch **out);
int qipgpu_batch_destroy(qipgpu_batch *b);
/* Replace the gate lists (packed as above) and plan them with `tuning` (NULL: defaults). */
int qipgpu_batch_load(qipgpu_batch *b, int64_t n_gates, const int32_t *gate_state, const int32_t *gate_k,
                      const int32_t *gate_qubits, const double *mats, const qipgpu_batch_tuning *tuning);
/* Re-plan the loaded gates with another tuning. */
int qipgpu_batch_set_tuning(qipgpu_batch *b, const qipgpu_batch_tuning *tuning);
int qipgpu_batch_get_tuning(qipgpu_batch *b, qipgpu_batch_tuning *out);
/* Time runs from |0...0> under candidate tunings (chunk_tiles, then lds_max_qubits, then
 * fusion_width, each kept at its best so far; `reps` timed runs per candidate after one warm-up)
 * and keep the fastest. Leaves every state reset to |0...0>. */
int qipgpu_batch_autotune(qipgpu_batch *b, int32_t reps, qipgpu_batch_tuning *chosen, double *best_ms);
/* Set every state to |0...0> (enqueued). */
int qipgpu_batch_reset(qipgpu_batch *b);
/* Apply every state's gates to its current contents (enqueued). */
int qipgpu_batch_run(qipgpu_batch *b);
int qipgpu_batch_synchronize(qipgpu_batch *b);
/* Copy state i (2^n complex64) to / from host memory (synchronous). */
int qipgpu_batch_get_state(qipgpu_batch *b, int32_t i, void *out);
int qipgpu_batch_set_state(qipgpu_batch *b, int32_t i, const void *in);
/* Plan figures: out[0..len) of {states, LDS states, tiled states, gates, blocks, waves, chunks,
 * largest wave (blocks), arena bytes, algorithmic HBM bytes of one run}. */
int qipgpu_batch_info(qipgpu_batch *b, int64_t *out, int32_t len);
int qipgpu_batch_get_stream(qipgpu_batch *b, void **stream);

#ifdef __cplusplus
}
#endif

#endif /* QIPGPU_H */
