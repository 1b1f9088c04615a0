/*
 * neb_hip.h -- C ABI of libnebhip.so, the gfx950 (MI355X) backend for the
 * neural eigenfluid network hot path of baiming-zhang/NEB.
 *
 * The reference has no compiled code: its hot path is CuPy array code whose
 * matmuls/einsums go to cuBLAS and whose fused sine-jet chain is an NVRTC
 * ElementwiseKernel.  Each entry point below replaces one of those operation
 * groups; the reference interface it stands in for is cited (paths relative
 * to the reference root):
 *   HB2P  = experiments/2d/_backend/periodic_core/vortex_merger/src/eigenfluid/
 *           gen/vortex_merger2d/hardcoded_backend.py
 *   VJ2   = experiments/metaball/_backend/flow/cases/metaball/velocity_trace_jet2.py
 *   NCC   = experiments/metaball/_backend/flow/cases/metaball/native_cupy_core.py
 *   VCORE = experiments/metaball/_backend/flow/cases/metaball/velocity/core.py
 *   FUSED = experiments/metaball/_backend/flow/cases/metaball/velocity/fused.py
 *   FAST  = experiments/metaball/_backend/flow/cases/metaball/velocity/fast_path.py
 *   PP    = experiments/metaball/_backend/exports/basis_projection/postprocessing/
 *           paired_projection.py
 *
 * Conventions
 *  - Every buffer is caller-owned device memory passed as a plain pointer;
 *    the library allocates nothing.  Scratch space for split reductions is
 *    a caller-provided workspace sized by the *_workspace_size functions.
 *  - Calls are stream-ordered on `stream` (a hipStream_t, NULL = default
 *    stream) and never synchronize the device.
 *  - Return 0 on success or an NEB_ERR_* code; neb_last_error() returns a
 *    thread-local message.  SHAPE/ARG map to ValueError, HIP to RuntimeError
 *    (the reference's exception classes, HB2P:3299-3305, VCORE:60-61); no C++
 *    exception crosses the ABI.
 *  - Packed jets: a second-order jet of width f over n points is stored
 *    channel-major [C][n][ld] with channel stride n*ld, channels ordered
 *    value, d grads, d(d+1)/2 Hessian entries in the pair order
 *    (0,0),(0,1),(1,1) [2D] or (0,0),(0,1),(0,2),(1,1),(1,2),(2,2) [3D]
 *    (HB2P:154-166, VJ2:9,20-22).  C = 6 in 2D, 10 in 3D.
 *  - prec selects the storage type of jets / staged weights and the MFMA
 *    instruction: NEB_BF16 (bf16 operands, fp32 accumulate; fields, biases
 *    and epilogue math in fp32), NEB_F32 (fp32), NEB_F64 (fp64).
 *  - Leading dimensions are in elements and must keep 16-byte alignment of
 *    every row (multiples of 8 for bf16, 4 for fp32, 2 for fp64).
 *  - Empty inputs: with zero points (n = 0, rows = 0) a call checks its
 *    sizes and touches no point-indexed buffer, which may then be null; the
 *    jet and field calls write nothing, the Gram / cross / column-sum
 *    reductions write beta * G (zeros for beta = 0), like a sum over no
 *    points.
 */
#ifndef NEB_HIP_H
#define NEB_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NEB_ABI_VERSION 16

enum { NEB_OK = 0, NEB_ERR_SHAPE = 1, NEB_ERR_NONFINITE = 2, NEB_ERR_HIP = 3, NEB_ERR_ARG = 4 };
enum { NEB_F32 = 0, NEB_BF16 = 1, NEB_F64 = 2 };             /* prec / dtype codes */
enum { NEB_TANH = 0, NEB_SIN = 1 };                           /* activations */
/* lift codes of the output-layer entry points (their `dim` argument): the 2D
 * velocity + vorticity lift, the 3D curl lift, the 2D velocity + Hessian
 * lift of the causal velocity peel (stiffness rows psi_xx, sqrt2 psi_xy,
 * psi_yy, whose Gram is the Hessian Frobenius form with the pair metric
 * (1, 2, 1) of HB2P:3022), and the 2D streamfunction lift of the Dirichlet
 * Rayleigh peel (mass rows psi, stiffness rows psi_x, psi_y: M = int psi^2,
 * K = int |grad psi|^2, HB2P:3394-3397) */
enum { NEB_LIFT_2D_VORTICITY = 2, NEB_LIFT_3D_CURL = 3, NEB_LIFT_2D_HESSIAN = 4, NEB_LIFT_2D_STREAM = 5 };
/* the streamfunction lift on first-order tapes (value and gradient channels
 * only, the Jet1 stream of HB2P:1716-1800): the fields of
 * NEB_LIFT_2D_STREAM; the output-layer reverse writes a 3-channel adjoint */
enum { NEB_LIFT_2D_STREAM_JET1 = 6 };
/* dim code of the jet-layer entry points (neb_jet_layer_forward / _x3 /
 * _blocks, neb_jet_layer_backward / _blocks) for 2D first-order jets:
 * packed channels (value, d/dx, d/dy), C = 3 */
enum { NEB_DIM_2D_JET1 = 12 };
enum { NEB_CROSS_F32 = 0, NEB_CROSS_BF16 = 1, NEB_CROSS_BF16X3 = 2, NEB_CROSS_F64 = 3 };

int neb_abi_version(void);
const char* neb_last_error(void);

/* Periodic sin/cos feature jets.  Replaces encoding_jet2 periodic branch
 * (HB2P:566-589) / encoding_jet1 (HB2P:418-441) and features1 (NCC:76-82).
 * x: [n][dim] fp32 (F32/BF16) or fp64 (F64).  out: packed [C][n][ldf], the
 * 2*freqs*dim features followed by zero padding up to ldf. */
int neb_feature_jets(int prec, int dim, int64_t n, const void* x, int freqs, const double* lo,
                     const double* hi, void* out, int64_t ldf, void* stream);

/* Hidden layer forward: z = omega*(x W^T + b) on all channels (bias on the
 * value channel only), y = act-jet(z).  Replaces linear_forward_batch_jet2 +
 * omega scaling + sine_forward_jet2 (HB2P:1005-1040, :1836-1837) and
 * VJ2.linear/scaled/sine (VJ2:17,26-34) with the FUSED sine kernel
 * (FUSED:3-8).  x: packed [C][n][ldx] (k valid columns), W: [width][ldw],
 * bias: math precision [width] (fp32, fp64 for F64) or NULL.
 * z, y: packed [C][n][ldy]. */
int neb_jet_layer_forward(int prec, int dim, int act, int64_t n, int64_t k, int64_t width, const void* x,
                          int64_t ldx, const void* w, int64_t ldw, const void* bias, double omega, void* z,
                          void* y, int64_t ldy, void* stream);

/* Output layer, raw jets: out = y W^T (+ bias on the value channel),
 * packed [C][n][ldo] in fp32 (fp64 for F64).  Replaces the last
 * linear_forward_batch_jet2 of forward_stream_jet2 (HB2P:1802-1880) /
 * VJ2.forward (VJ2:65-76). */
int neb_jet_output_jets(int prec, int dim, int64_t n, int64_t k, int64_t nout, const void* y, int64_t ldy,
                        const void* w, int64_t ldw, const void* bias, void* out, int64_t ldo, void* stream);

/* The output jets times a per-point envelope jet e (the Jet2 product rule
 * y = e o, grad y = e grad o + o grad e, hess_ij y = e hess_ij o + o hess_ij e
 * + e_i o_j + e_j o_i) applied in the epilogue of the output-layer GEMM:
 * field = envelope * raw (VJ2:47-50 `product`, NCC:94-95; the periodic
 * configurations use the identity envelope, HB2P:546-547).  env: packed
 * [C][n] jets in the math precision (fp32, fp64 for F64), one per point,
 * shared by all nout outputs.  Other arguments as neb_jet_output_jets.
 * neb_envelope_reverse: its transpose (VJ2:51-55 `product_reverse`): the
 * cotangent adj [C][n][lda] of the enveloped jets (math precision) -> the
 * cotangent of the raw output jets, written [C][n][ldo] in the operand
 * precision (bf16 / fp32 / fp64) of the output-layer reverse
 * (neb_jet_layer_backward with the output weights). */
int neb_jet_output_jets_enveloped(int prec, int dim, int64_t n, int64_t k, int64_t nout, const void* y,
                                  int64_t ldy, const void* w, int64_t ldw, const void* bias, const void* env,
                                  void* out, int64_t ldo, void* stream);
int neb_envelope_reverse(int prec, int dim, int64_t n, int64_t nout, const void* env, const void* adj, int64_t lda,
                         void* out, int64_t ldo, void* stream);

/* Output layer fused with the differential lift, writing field matrices in
 * fp32 (fp64 for F64).  2D: w [K][ldw]; mass = (psi_y, -psi_x) as [2n][ldf]
 * and stiff = omega = -(psi_xx+psi_yy) as [n][ldf] (ELC:35-38 =
 * .../canonical_gauge/scripts/evaluate_large_comparison.py).  3D: w [3K][ldw]
 * with rows 3k+c (vector potential); mass = curl A as [3n][ldf], stiff =
 * grad u as [9n][ldf] rows 3a+j = d_j u_a (NCC:97-99, VJ2:93-101,
 * VCORE:71-77).  dim = NEB_LIFT_2D_HESSIAN: 2D mass as above, stiff = the
 * Hessian rows (psi_xx, sqrt2 psi_xy, psi_yy) as [3n][ldf] (the velocity
 * Rayleigh peel's H1 seminorm, HB2P:2935-3060, metric (1, 2, 1) at
 * HB2P:3022); dim = NEB_LIFT_2D_STREAM: mass = psi as [n][ldf] (without the
 * output bias, which the peel's centering removes), stiff = (psi_x, psi_y)
 * as [2n][ldf] (HB2P:3461-3474); the same lift codes apply to
 * neb_jet_output_fields_split,
 * neb_jet_output_fields_x3 and neb_jet_output_backward.  layout 0 =
 * point-major rows (p*comp+c, the reference's [N*comp, K] layout,
 * VCORE:75-76); 1 = component-major (c*n+p). */
int neb_jet_output_fields(int prec, int dim, int64_t n, int64_t k, int64_t K, const void* y, int64_t ldy,
                          const void* w, int64_t ldw, void* mass, void* stiff, int64_t ldf, int layout,
                          void* stream);

/* The same fields (prec NEB_BF16 only) written as two bf16 planes each:
 * hi = bf16(x) and lo = bf16(x - hi) of the fp32 value x, in the layout of
 * neb_jet_output_fields.  neb_gram_split consumes both planes (the split
 * Gram without re-splitting per tile); the hi plane alone is the bf16
 * operand of neb_jet_output_backward and of the weight-gradient crosses.
 * Replaces the same call sites as neb_jet_output_fields (VCORE:71-77)
 * where the Gram runs in split precision. */
int neb_jet_output_fields_split(int dim, int64_t n, int64_t k, int64_t K, const void* y, int64_t ldy, const void* w,
                                int64_t ldw, void* mass_hi, void* mass_lo, void* stiff_hi, void* stiff_lo,
                                int64_t ldf, int layout, void* stream);

/* Split-precision operands (2D and 3D jets).  A value x is held as two bf16 planes hi = bf16(x), lo = bf16(x - hi); every
 * product of the contraction is W_lo x_hi + W_hi x_lo + W_hi x_hi on bf16
 * MFMA with fp32 accumulation (about 16 significant bits per operand), and
 * the results are stored as hi / lo planes again.  The jets and fields then
 * carry FP32-level accuracy at bf16 MFMA rates; the hi planes are the bf16
 * tape and fields of the reverse pass.
 *
 * neb_split_bf16: hi / lo planes of an fp32 or fp64 matrix (rows x cols,
 * source stride lds, plane stride ldd): the feature jets and master weights
 * enter the split path through it (the `.astype(compute_dtype)` of
 * HB2P:1026-1030, VJ2:68, kept at ~16 bits).
 * neb_jet_layer_forward_x3: neb_jet_layer_forward on split input jets and
 * weights; writes y as hi / lo planes and z as its hi plane, plus its lo
 * plane when z_lo is not null (HB2P:1005-1040, VJ2:20-34).
 * neb_jet_output_fields_x3: neb_jet_output_fields_split on split last-hidden
 * jets and output weights (VCORE:71-77, ELC:35-38; 3D: NCC:97-99,
 * VJ2:93-101, the weight rows 3k+c as in neb_jet_output_fields). */
int neb_split_bf16(int src_dtype, const void* src, int64_t rows, int64_t cols, int64_t lds, void* hi, void* lo,
                   int64_t ldd, void* stream);

/* The inverse of neb_split_bf16: dst = hi + lo (rows x cols, plane stride ld)
 * as NEB_F32 (exact) or NEB_F64.  The split-precision basis enters the FP64
 * full-mass projection loads through it (PP:141-177 reads an fp32 basis). */
/* Column centering x[r][c] -= mean[c] of a rows x cols matrix (stride ld):
 * dtype NEB_F32 / NEB_F64, or NEB_BF16 -- one plane (lo null, the
 * difference rounded to bf16) or split planes (lo not null: hi + lo - mean
 * split again into hi and lo).  mean: FP64 [cols].  The periodic gauge of
 * the Dirichlet streamfunction peel (HB2P:3475-3479, `raw_value -=
 * cp.mean(raw_value, axis=0)`), applied to the field matrix the Grams and
 * the output-layer reverse then read. */
int neb_center_columns(int dtype, void* x, void* lo, int64_t rows, int64_t cols, int64_t ld, const double* mean,
                       void* stream);

int neb_join_bf16(int dst_dtype, const void* hi, const void* lo, int64_t rows, int64_t cols, int64_t ld, void* dst,
                  int64_t ldd, void* stream);
int neb_jet_layer_forward_x3(int dim, int act, int64_t n, int64_t k, int64_t width, const void* x_hi,
                             const void* x_lo, int64_t ldx, const void* w_hi, const void* w_lo, int64_t ldw,
                             const void* bias, double omega, void* z_hi, void* z_lo, void* y_hi, void* y_lo,
                             int64_t ldy, void* stream);
int neb_jet_output_fields_x3(int dim, int64_t n, int64_t k, int64_t K, const void* y_hi, const void* y_lo,
                             int64_t ldy, const void* w_hi, const void* w_lo, int64_t ldw, void* mass_hi,
                             void* mass_lo, void* stiff_hi, void* stiff_lo, int64_t ldf, int layout, void* stream);

/* Output-layer reverse fused with the lift transpose and the transpose of
 * the last hidden activation.  Given field matrices Phi (mass, stiff) and
 * bm = 2w W^T dL/dM, bs = 2w W^T dL/dS ([width][ldb]; in 3D three
 * consecutive blocks, one per component c, each [width][ldb] built from
 * W rows 3k+c), computes the adjoint of the last pre-activation jet:
 * adj = omega * act-jet^T(z, lift^T(2w Phi dL/dG) W).  Replaces
 * VCORE:94-100 (gu = U Cu/n, lift) + VJ2.product_reverse/linear_reverse +
 * FUSED:9-19 (and HB2P:1115-1174 in 2D).  z: tape packed [C][n][ldz];
 * adj: packed [C][n][lda].  fields_dtype: NEB_F32 (or the bf16 hi plane,
 * NEB_BF16) for prec NEB_BF16, NEB_F32 for NEB_F32, NEB_F64 for NEB_F64;
 * fp32 fields meet the bf16 MFMA rounded to bf16, i.e. as the hi plane. */
int neb_jet_output_backward(int prec, int dim, int act, int64_t n, int64_t width, int64_t K, const void* mass,
                            const void* stiff, int64_t ldf, int fields_dtype, int layout, const void* bm,
                            const void* bs, int64_t ldb, const void* z, int64_t ldz, double omega, void* adj,
                            int64_t lda, void* stream);

/* Hidden layer reverse: adj_out = omega_prev * act-jet^T(z_prev, adj_in W).
 * wt = W^T staged as [k][ldw] (k = layer input width).  Replaces
 * linear_backward_batch_jet2's input adjoint (HB2P:1115-1144), VJ2:28-30,
 * then sine_backward_jet2 / FUSED:9-19 for the previous layer. */
int neb_jet_layer_backward(int prec, int dim, int act, int64_t n, int64_t k, int64_t width, const void* adj_in,
                           int64_t lda_in, const void* wt, int64_t ldw, const void* z_prev, int64_t ldz,
                           double omega_prev, void* adj_out, int64_t lda_out, void* stream);

/* Block-diagonal hidden layers (the per-mode networks of HB2P:1590-1880,
 * NCC:111-185: a block of B modes of width h is one layer of width B*h whose
 * weights are zero outside the h x h diagonal blocks).  The same jets,
 * epilogues and precisions as neb_jet_layer_forward / _x3 /
 * neb_jet_layer_backward on a square layer (k = width), except that each
 * tile of output features contracts only the k of the blocks it belongs to
 * (block: h, a multiple of 8 dividing the width); the zeros outside the
 * blocks are not read.  Replaces the per-mode batched matmuls of
 * HB2P:1005-1036 (forward) and HB2P:1115-1144 (input adjoint).
 * neb_cross_blocks: neb_cross restricted to an m x m output's diagonal
 * blocks of `block` (block nesting with the 128-wide tiles: a divisor or a
 * multiple of 128, dividing m): the tiles over those blocks are computed and
 * stored, every other entry of G is left as it is; the per-mode weight
 * gradients (HB2P:1087-1144 per mode).  Workspace:
 * neb_cross_blocks_workspace_size. */
int neb_jet_layer_forward_blocks(int prec, int dim, int act, int64_t n, int64_t width, int64_t block, const void* x,
                                 int64_t ldx, const void* w, int64_t ldw, const void* bias, double omega, void* z,
                                 void* y, int64_t ldy, void* stream);
int neb_jet_layer_forward_x3_blocks(int dim, int act, int64_t n, int64_t width, int64_t block, const void* x_hi,
                                    const void* x_lo, int64_t ldx, const void* w_hi, const void* w_lo, int64_t ldw,
                                    const void* bias, double omega, void* z_hi, void* z_lo, void* y_hi, void* y_lo,
                                    int64_t ldy, void* stream);
int neb_jet_layer_backward_blocks(int prec, int dim, int act, int64_t n, int64_t width, int64_t block,
                                  const void* adj_in, int64_t lda_in, const void* wt, int64_t ldw, const void* z_prev,
                                  int64_t ldz, double omega_prev, void* adj_out, int64_t lda_out, void* stream);
size_t neb_cross_blocks_workspace_size(int64_t m, int64_t block, int64_t rows);
int neb_cross_blocks(int mode, int nseg, const void* const* x, const void* const* y, const int64_t* ldx,
                     const int64_t* ldy, const int64_t* rows, const double* sign, int x_dtype, int y_dtype, int64_t m,
                     int64_t block, double scale, double beta, double* G, int64_t ldg, void* workspace,
                     size_t ws_bytes, void* stream);

/* Mode-diagonal crosses: out[m][j] = beta*out[m][j] + scale * sum_s sign_s
 * sum_r x_s[r][m] y_s[r][m*h + j], m < B, j < h (out: FP64 [B][h], row
 * stride h) -- the diagonal 1 x h blocks of neb_cross's B x B*h result,
 * i.e. each mode's output-weight products (a block's field-cotangent column
 * m against its own h last-hidden features), without the products against
 * the block's other modes.  x_s: [rows_s][ldx_s] (F32 / F64, ldx_s >= B);
 * y_s: [rows_s][ldy_s] (BF16 / F32 / F64, 16-byte aligned rows, ldy_s >=
 * B*h); h a multiple of 8; up to 16 segments.  FP64 accumulation in a fixed
 * order (deterministic).  Replaces the per-mode `cp.matmul(gu.T, y)` output
 * weight gradients of HB2P:2087-2139.  Workspace:
 * neb_cross_modes_workspace_size(B, h, max rows). */
size_t neb_cross_modes_workspace_size(int64_t B, int64_t h, int64_t rows);
int neb_cross_modes(int nseg, const void* const* x, const void* const* y, const int64_t* ldx, const int64_t* ldy,
                    const int64_t* rows, const double* sign, int x_dtype, int y_dtype, int64_t B, int64_t h,
                    double scale, double beta, double* out, void* workspace, size_t ws_bytes, void* stream);

/* The output layer + lift of a block of B per-mode networks (hidden width h)
 * on split operands: the fields of neb_jet_output_fields_x3 for the block's
 * B modes, contracting each mode's ncomp output rows over its own h hidden
 * features only.  y: the block's last hidden jets ([C][n][ldy], width B*h,
 * hi / lo); w: the block-diagonal output weights ([B*ncomp][ldw], row
 * m*ncomp + c holds mode m's component c at columns m*h .. m*h+h-1, hi /
 * lo); the fields go to columns 0 .. B-1 of the split planes (point-major
 * rows; offset the pointers to the block's first column).  lift: 2, 3 or
 * NEB_LIFT_2D_HESSIAN; h a multiple of 32, B a multiple of 8.  Replaces the
 * per-mode output layer and lift of HB2P:1802-1880 / VJ2:93-101. */
int neb_mode_output_fields_x3(int lift, int64_t n, int64_t h, int64_t B, const void* y_hi, const void* y_lo,
                              int64_t ldy, const void* w_hi, const void* w_lo, int64_t ldw, void* mass_hi,
                              void* mass_lo, void* stiff_hi, void* stiff_lo, int64_t ldf, void* stream);

/* The output-layer reverse of a block of per-mode networks: the result of
 * neb_jet_output_backward with K = width / block for weights with one
 * nonzero per row, output column f = m*block + j (of each component)
 * reading only field column m = f / block -- mode m's output weights
 * against mode m's field cotangents.  bm / bs hold just those nonzeros, as
 * [ncomp][ldb] vectors (ldb >= width; entry f of component c: mode f /
 * block's weight on hidden feature f % block), in the operand precision.
 * One thread per point and 8 or 4 features: the lift's transpose and the
 * activation jet chain on the tape, no contraction over the block's other
 * modes.  block: the hidden width h, a multiple of 8 dividing the width.
 * Replaces the per-mode output-layer reverse of HB2P:2087-2139 (`gu`, lift
 * transpose and `linear_backward_batch_jet2` of the last layer, per mode). */
int neb_jet_output_backward_blocks(int prec, int dim, int act, int64_t n, int64_t width, int64_t block,
                                   const void* mass, const void* stiff, int64_t ldf, int fields_dtype, int layout,
                                   const void* bm, const void* bs, int64_t ldb, const void* z, int64_t ldz,
                                   double omega, void* adj, int64_t lda, void* stream);

/* Factored output-layer reverse (split-precision tapes).  The fields are
 * linear in the output weights, Phi_a = sum_t sg_t Y_{c_t} W_{s_t}^T over the
 * lift's terms t (channel c_t of the last hidden jets Y, output-weight slice
 * s_t: rows 3k+s in 3D, all rows in 2D), so the field adjoint
 * 2w Phi dL/dG^T through the lift transpose and the output layer is
 *   adj Y_c = sum_c' Y_c' R_c'c,   R_c'c = 2w sum sg sg' W_s'^T dL/dG^T W_s
 * with width x width blocks R formed in FP64 from the master weights: a
 * GEMM whose inner length is the hidden width instead of K, and whose
 * K x K contraction (where an ill-conditioned mass makes dL/dM large and
 * the sum cancel) runs in FP64.  It replaces the same operations as
 * neb_jet_output_backward (VCORE:94-100 gu = U Cu/n and the lift /
 * product / linear transposes of VJ2:77-82, NCC:97-109), in three steps:
 *
 * neb_stack_channels: channels ch[0..nch-1] of a packed bf16 jet tensor
 * ([C][n][ld], hi plane and optionally lo plane) side by side per point:
 * dst[p][i*width + j] = src[ch[i]][p][j] ([n][ldd]; width a multiple of 8,
 * nch <= 16).  The stacked last hidden jets are the operand of the GEMM
 * below and of the channel Grams Y_c'^T Y_c (neb_gram_split) that give the
 * output-weight gradient 2w sum sg sg' dL/dG W_s' Y_c'^T Y_c.
 * neb_gemm_split: out[p][f] = sum_k x[p][k] w[f][k] on split operands (hi /
 * lo bf16 planes, the three products of the split layers), fp32 result
 * [n][ldo]: the stacked adjoint Y_stack R, with w = R^T staged as planes.
 * neb_jet_act_reverse: the transpose of the last hidden activation on its
 * own: adj = omega * act-jet^T(z, a), where a holds the adjoints of jet
 * channels ch[i] as fp32 columns [i*width, (i+1)*width) of [n][lda] (other
 * channels zero) and z / adj are the packed tape and result ([C][n][ld*],
 * prec NEB_BF16 or NEB_F32); FUSED:9-19, HB2P:1147-1174. */
int neb_stack_channels(const void* src_hi, const void* src_lo, int64_t n, int64_t ld, int64_t width, int nch,
                       const int* ch, void* dst_hi, void* dst_lo, int64_t ldd, void* stream);
int neb_gemm_split(int64_t n, int64_t k, int64_t nout, const void* x_hi, const void* x_lo, int64_t ldx,
                   const void* w_hi, const void* w_lo, int64_t ldw, float* out, int64_t ldo, void* stream);
int neb_jet_act_reverse(int prec, int dim, int act, int64_t n, int64_t width, const float* adj, int64_t lda, int nch,
                        const int* ch, const void* z, int64_t ldz, double omega, void* out, int64_t ldo,
                        void* stream);

/* Split-R cross product  G = beta*G + scale * sum_s sign_s X_s^T diag(roww) Y_s
 * over rows, FP64 result [m][ldg].  Segment s: X_s [rows_s][ldx_s] with m
 * used columns, Y_s [rows_s][ldy_s] with n used columns; x_dtype / y_dtype
 * are NEB_F32 / NEB_BF16 / NEB_F64 (NEB_CROSS_F64: fp64 operands, or fp32
 * X widened exactly to fp64 against fp64 or fp32 Y).  symmetric != 0 requires X == Y and
 * computes upper tiles only.  Replaces the FP64 Gram blocks of VCORE:40-46,
 * experiments/metaball/run.py:99-107, experiments/periodic/paper_rollout.py:
 * 178-179, the weight-gradient einsums HB2P:1087-1144 / VJ2:29 / FAST:27-29
 * and the projection loads PP:148-154.  roww (fp64, per row) is allowed with
 * a single segment only. */
size_t neb_cross_workspace_size(int64_t m, int64_t n, int64_t rows, int symmetric);
int neb_cross(int mode, int nseg, const void* const* x, const void* const* y, const int64_t* ldx,
              const int64_t* ldy, const int64_t* rows, const double* sign, int x_dtype, int y_dtype, int64_t m,
              int64_t n, const double* roww, int symmetric, double scale, double beta, double* G, int64_t ldg,
              void* workspace, size_t ws_bytes, void* stream);

/* Quadrature-weighted Gram  G = beta*G + w * Phi^T diag(roww) Phi  (FP64,
 * full symmetric K x K).  Phi: [rows][ld], K columns, dtype phi_dtype.
 * Replaces VCORE:40-46 `grams` and run.py:99-116 `fullmass`. */
int neb_gram(int mode, const void* phi, int64_t rows, int64_t K, int64_t ld, int phi_dtype, double w,
             const double* roww, double beta, double* G, int64_t ldg, void* workspace, size_t ws_bytes,
             void* stream);

/* Split-precision Gram  G = beta*G + w * Phi^T Phi  (FP64, full symmetric)
 * from pre-split planes (neb_jet_output_fields_split): three bf16 MFMA
 * products hi.hi + hi.lo + lo.hi per pair, fp32 accumulation flushed into
 * fp32 partial sums every 256 rows, FP64 split reduction: the products of
 * neb_gram(NEB_CROSS_BF16X3) on the fp32 matrix the planes came from, equal
 * up to the rounding of differently split partial sums.  ld < 2^25.
 * Workspace: neb_cross_workspace_size(K, K, rows, 1).  Replaces VCORE:40-46
 * `grams` on the split-precision path. */
int neb_gram_split(const void* phi_hi, const void* phi_lo, int64_t rows, int64_t K, int64_t ld, double w, double beta,
                   double* G, int64_t ldg, void* workspace, size_t ws_bytes, void* stream);

/* neb_gram_split without a workspace argument: the same kernels and the same
 * result, bit for bit, with the split-R partial tiles held in a file-scope
 * device array of NEB_NOALLOC_SCRATCH_BYTES (32 MiB) that each GPU the
 * library runs on keeps until the process ends.  The call allocates nothing
 * (no hipMalloc, hipFree or allocating library routine); its kernels write
 * every partial tile before reading it, so nothing has to start at zero.
 * Limit: K up to 3968 (31 column tiles of 128) for any row count; larger K is
 * rejected with NEB_ERR_ARG.  Every call on one GPU shares that GPU's array:
 * calls that overlap in time on one GPU (other streams, other threads) race
 * and must be ordered by the caller.  No counterpart in the reference: it is
 * for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). */
#define NEB_NOALLOC_SCRATCH_BYTES ((size_t)32 << 20)
int neb_gram_split_noalloc(const void* phi_hi, const void* phi_lo, int64_t rows, int64_t K, int64_t ld, double w,
                           double beta, double* G, int64_t ldg, void* stream);

/* Caps the workgroups that each of this thread's later bf16 cross launches
 * (neb_cross on bf16 operands, neb_gram_split) keeps resident: 0 = as many
 * as fit (two per CU), otherwise at most max_workgroups (the row splits
 * shrink to fit; results change only by the rounding of differently split
 * partial sums).  No counterpart in the reference: a scheduling knob for
 * products queued on another stream beside latency-bound launches (the FP64
 * SPD sweeps), which then find whole CUs free.  NEB_ERR_ARG if negative. */
int neb_set_cross_occupancy(int max_workgroups);

/* Column sums in FP64: out[c] = beta*out[c] + sum_r x[r][c] (bias
 * gradients, HB2P:1101 `grad_bias`, VJ2:30).  Workspace from
 * neb_colsum_workspace_size. */
size_t neb_colsum_workspace_size(int64_t rows, int64_t cols);
int neb_colsum(int dtype, const void* x, int64_t rows, int64_t cols, int64_t ld, double beta, double* out,
               void* workspace, size_t ws_bytes, void* stream);

/* Sum of squares in FP64 into device *out (gradient norm of the Adam clip,
 * NCC:157).  Workspace: 256 doubles. */
int neb_sumsq(int dtype, const void* x, int64_t n, double* out, void* workspace, size_t ws_bytes, void* stream);

/* Adam step with bias correction on flat fp32/fp64 arrays, gradients scaled
 * by `scale` (the clip factor min(1, clip/|g|)).  Replaces Model.adam
 * (NCC:156-171) / adam_step_jet1_device (HB2P:2184-2278). */
int neb_adam(int dtype, void* p, const void* g, void* m, void* v, int64_t n, double lr, double scale, double b1,
             double b2, double eps, int64_t t, void* stream);

/* Adam step as neb_adam with the step's gates evaluated on the device, so a
 * training step needs no host synchronisation before the update is queued:
 * the clip factor min(1, clip/|g|) is formed from the device squared norm
 * *norm2 (neb_sumsq), and p, m, v are left untouched when |g| is zero or
 * nonfinite or when *gate (device int, nullable: an earlier gate of the same
 * step, e.g. the mass positive-definiteness flag) is nonzero; the host reads
 * *norm2 and *gate afterwards to raise.  Replaces Model.adam (NCC:156-171:
 * clip_grad_norm_ + the FloatingPointError gate + Adam). */
int neb_adam_gated(int dtype, void* p, const void* g, void* m, void* v, int64_t n, double lr, double clip,
                   double b1, double b2, double eps, int64_t t, const double* norm2, const int* gate, void* stream);

/* B = (A + A^T)/2 + shift I for K x K FP64 matrices (B may alias A): the
 * ridge-shifted symmetric mass M_r of the spectral objective
 * (MB/velocity_trace_jet2.py:107-109, `0.5 * (M + M.T)` + ridge). */
int neb_sym_shift(int64_t K, const double* A, int64_t lda, double shift, double* B, int64_t ldb, void* stream);

/* Terms of the spectral objective L = <X, sym(S)> + lam ||M - I||_F^2 with
 * X = M_r^-1 (MB/velocity_trace_jet2.py:107-117; the trace tr(M_r^-1 S) as
 * the Frobenius product): symmetrises X in place, writes Ss = sym(S), and
 * writes the device scalars *loss and *gate.  *gate = *info when info is
 * non-null and nonzero (first nonpositive pivot of neb_spd_inverse), else -1
 * when the condition estimate ||M_r||_F ||X||_F (M_r = sym(M) + ridge I) is
 * nonfinite or above cond_limit, else 0: the gates of
 * MB/velocity_trace_jet2.py:109-111 as a device flag for neb_adam_gated.
 * Deterministic.  Workspace: neb_spectral_workspace_size(K). */
size_t neb_spectral_workspace_size(int64_t K);
int neb_spectral_terms(int64_t K, const double* M, int64_t ldm, const double* S, int64_t lds, double* X,
                       int64_t ldx, const int* info, double lam, double ridge, double cond_limit, double* Ss,
                       int64_t ldss, double* loss, int* gate, void* workspace, size_t ws_bytes, void* stream);

/* Mass adjoint of the spectral objective, dL/dM = -sym(C) + 2 lam (M - I)
 * with C = X Ss X (MB/velocity_trace_jet2.py:113-117).  dM may alias C, not M. */
int neb_spectral_adjoint(int64_t K, const double* C, int64_t ldc, const double* M, int64_t ldm, double lam, double* dM,
                         int64_t ldd, void* stream);

/* FP64 inverse of a symmetric positive definite K x K matrix (blocked
 * Gauss-Jordan sweeps, 64-column pivot blocks).  X must not alias A.  *info
 * (device int) is set to 0, or to the 1-based index of the first
 * nonpositive pivot, i.e. the matrix is not positive definite (the
 * FloatingPointError gate of velocity_trace_jet2.py:109-111); X is then
 * meaningless.  Replaces cholesky + solve / solve_triangular of
 * velocity_trace_jet2.py:112 and velocity/core.py:52-56.  Workspace from
 * neb_spd_inverse_workspace_size. */
size_t neb_spd_inverse_workspace_size(int64_t K);
int neb_spd_inverse(int64_t K, const double* A, int64_t lda, double* X, int64_t ldx, int* info, void* workspace,
                    size_t ws_bytes, void* stream);

/* neb_spd_inverse that also writes T (K x K, row stride ldt) = U^-1, the
 * inverse of the upper Cholesky factor of A (A = U^T U, T^T A T = I, T upper
 * triangular with a positive diagonal, zeros below it), from the same
 * sweeps: the factor the coefficient peel orthonormalises the sorted mass
 * with.  Replaces cholesky + solve_triangular(L^T, I) of velocity/core.py:
 * 53-56 (both passes).  A, X and T must not alias; *info as for
 * neb_spd_inverse (T is then meaningless); same workspace size. */
int neb_spd_inverse_chol(int64_t K, const double* A, int64_t lda, double* X, int64_t ldx, double* T, int64_t ldt,
                         int* info, void* workspace, size_t ws_bytes, void* stream);

/* Reduced-order right-hand side of the 2D periodic rollout on the network
 * basis (experiments/periodic/paper_rollout.py:37-41 `rhs`): with the
 * orthonormalised basis fields gu, gv (velocity) and gw (vorticity) as
 * [n][ld] matrices (dtype NEB_F32 or NEB_F64, K <= 8192 columns) and the
 * coefficients a [K] (FP64):
 *   u = gu a, v = gv a, omega = gw a (per point),
 *   rhs = gu^T (w omega v) - gv^T (w omega u)   (FP64, [K]).
 * Workspace: neb_galerkin_workspace_size(n, K) bytes. */
size_t neb_galerkin_workspace_size(int64_t n, int64_t K);
int neb_galerkin_rhs2d(int dtype, const void* gu, const void* gv, const void* gw, int64_t n, int64_t K, int64_t ld,
                       const double* a, double w, double* rhs, void* workspace, size_t ws_bytes, void* stream);

/* One block of B per-mode networks as a single layer of a grouped GEMM
 * (the reference's mode blocks, HB2P:1716-1880 / NCC:84-89 `cp.matmul` over
 * [B, out, in] stacks): from the mode-stacked master weights [K][out][in]
 * (fp32 or fp64), modes m0 .. m0+B-1, writes the [B*out][ld] operand --
 * diag = 0: the modes' weights stacked by rows (a first layer that reads
 * shared input features); diag = 1: block-diagonal (each mode reads its own
 * block of B*in inputs) -- in dst_dtype (bf16 / fp32 / fp64), columns
 * beyond the width zero; lo: the split lo plane (bf16 only) or NULL; wt: the
 * transpose [width][ldt] (for the reverse) or NULL.  diag = 2: the diagonal
 * blocks of diag = 1 only (B*out*in entries of hi / lo and of wt), every
 * other entry left as it is -- for buffers the caller keeps zero off the
 * diagonal blocks from one update to the next. */
int neb_stage_mode_block(int master_dtype, const void* master, int64_t K, int64_t out, int64_t in, int64_t m0,
                         int64_t B, int diag, int dst_dtype, void* hi, void* lo, int64_t ld, void* wt, int64_t ldt,
                         void* stream);

/* Every layer's bf16 operand forms in one launch, from the flat master
 * parameters (fp32 or fp64): layer l's weight W_l is the row-major
 * [rows[l]][cols[l]] block at element offset off[l]; hi[l] = bf16(W_l) as
 * [rows][ld[l]] (columns beyond cols zero), lo[l] = bf16(W_l - hi[l]) in the
 * same layout when lo[l] is not NULL (the rounding of neb_split_bf16), and
 * wt[l] = bf16(W_l^T) as [cols][ldt[l]] (columns beyond rows zero) when wt[l]
 * is not NULL.  At most 16 layers.  Replaces the per-update
 * `.astype(compute_dtype)` copies of the weight stacks after Adam
 * (HB2P:1026-1030, HB2P:2184-2278): one launch instead of one per form and
 * layer. */
int neb_stage_net(int master_dtype, const void* master, int nl, const int64_t* off, const int64_t* rows,
                  const int64_t* cols, void* const* hi, void* const* lo, const int64_t* ld, void* const* wt,
                  const int64_t* ldt, void* stream);

/* Weight staging: dst[r][c] = scale * (transpose ? src[c][r] : src[r][c])
 * converted from src_dtype to dst_dtype, zero outside the source extent
 * (dst is drows x dcols).  Replaces the per-call `.astype(compute_dtype)`
 * of the master weights (HB2P:1026-1030, VJ2:68). */
int neb_stage(int src_dtype, int dst_dtype, const void* src, int64_t rows, int64_t cols, int64_t lds, void* dst,
              int64_t ldd, int64_t drows, int64_t dcols, int transpose, double scale, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* NEB_HIP_H */
